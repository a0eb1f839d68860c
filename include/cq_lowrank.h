/*
 * cq_lowrank.h -- FP64 low-rank decomposition of the viscoacoustic step's
 * mixed-domain operator (libcq_host.so).
 *
 * Counterpart of the reference's acsglowrank2d / checka_error / consist_Vt
 * (sgpstd-qmodeling/sglowrank.c:63-480), which approximate a mixed-domain
 * symbol W(x, k) by  W ~ sum_j sum_i U(x, j) A(j, i) V(i, k)  from a few of its
 * columns (wavenumbers) and rows (cells), and whose apply kernels
 * (sgpstd-qmodeling/acps4.cu:1337-1459) form sum_j U(x, j) sum_i A(j, i)
 * Re F^-1[V_i F f] / N.  In the reference that path is dead (not linked, its
 * pseudo-inverse 4GINV.C missing, srand(time) seeding); here it is seeded,
 * has its own pseudo-inverse (one-sided Jacobi SVD) and feeds the engine's
 * rank-r apply (cq_job_desc::lr_rank).
 *
 * The symbol is the operator the viscoacoustic step applies to the
 * divergence of the particle velocity before the p update (sum_total,
 * sgpstd-qmodeling/acsglw.cu:1213-1245), with per-cell gam, vwd and
 * cc = dt^2 v^2 lam / 12 (transfermodel, func.c:912-929) and k2 = kx^2 + kz^2
 * on the reference's float wavenumber grid (cal_k2d, func.c:550-583):
 *   CQH_LR_TAYLOR4 : (1 - 32 gam) + 32 gam vwd k2^(1/32)
 *                    - cc ((1 - 64 gam) k2 + 64 gam vwd k2^(33/32)),
 *                    exactly the reference's live rank-4 operator;
 *   CQH_LR_FRACLAP : a^(2 gam) - cc k2 a^(4 gam),  a = vwd^16 |k| = |k| v / (2 pi fd),
 *                    the fractional-Laplacian dispersion operator that the
 *                    Taylor form expands to first order in gam.
 * Wavenumbers are sampled on the quarter plane kx = 0..nxx/2, kz = 0..nzz/2
 * (Nq = (nxx/2+1)(nzz/2+1) points, kz fastest), mirrored to negative ones
 * (consist_Vt, sglowrank.c:104-145).
 */
#ifndef CQ_LOWRANK_H
#define CQ_LOWRANK_H

#ifdef __cplusplus
extern "C" {
#endif

enum { CQH_LR_TAYLOR4 = 0, CQH_LR_FRACLAP = 1 };

/* W(x, k) of one cell at k2 = kx^2 + kz^2 (see above). */
double cqh_lr_symbol(int symbol, double k2, double gam, double vwd, double cc);

/* Quarter-plane wavenumber squares k2[ikx * nzh + ikz] (Nq of them) on the
 * reference's float grid (dk = 2 pi / (n d), pi = 3.1415926). */
void cqh_lr_kgrid(int nxx, int nzz, float dx, float dz, double *k2);

/* Decomposition of W over the window (N1 = nxx nzz cells, x-major) at rank
 * r (1..16) with the seeded randomised column / row selection of
 * sglowrank.c:148-480: r wavenumber columns and r cell rows chosen by
 * pivoted Gram-Schmidt on random samples, A = pinv(U_s) W_s pinv(V_s) on a
 * random sub-grid.  Outputs U[N1 * r] (U[x * r + j]), A[r * r]
 * (A[j * r + i]), V[r * Nq] (V[i * Nq + k]), all FP64.  Returns 0, or -1 if a
 * pseudo-inverse fails or an argument is bad. */
int cqh_lowrank(int nxx, int nzz, float dx, float dz, const float *gam, const float *vwd, const float *cc,
                int symbol, int rank, unsigned long long seed, double *U, double *A, double *V);

/* The engine's inputs from a decomposition: per-cell weights
 * wts[i * N1 + x] = sum_j U[x][j] A[j][i] (computed in FP64, stored as
 * float) and symbols sym[i * Nq + k] = V[i][k] / (nxx nzz) as float (the
 * inverse transform's 1/N folded in, as multv_kernel does). */
void cqh_lowrank_pack(int nxx, int nzz, int rank, const double *U, const double *A, const double *V,
                      float *wts, float *sym);

/* checka_error (sglowrank.c:63-101) over `ncell` cells: the largest
 * |W - sum U A V| over all wavenumbers of those cells, divided by the
 * largest |W| there. */
double cqh_lowrank_error(int nxx, int nzz, float dx, float dz, const float *gam, const float *vwd,
                         const float *cc, int symbol, int rank, const double *U, const double *A,
                         const double *V, int ncell, const int *cells);

#ifdef __cplusplus
}
#endif
#endif
