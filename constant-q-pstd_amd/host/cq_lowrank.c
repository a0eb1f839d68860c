/*
 * cq_lowrank.c -- FP64 low-rank decomposition of the viscoacoustic step's
 * mixed-domain operator (include/cq_lowrank.h).
 *
 * The method is the reference's acsglowrank2d (sgpstd-qmodeling/sglowrank.c:
 * 148-480), a randomised skeleton decomposition: W(x, k) ~ U A V with U the
 * values of W at r selected wavenumbers (all cells), V its values at r
 * selected cells (all wavenumbers) and A fitted on a random sub-grid.
 * Columns and rows are chosen by pivoted Gram-Schmidt (largest remaining
 * norm first, the others orthogonalised against it) on random samples of
 * r ln N rows or columns.  What differs, by design: every step is FP64, the
 * sampling is a seeded SplitMix64 stream (the reference seeds rand() with
 * the time of day), and the pseudo-inverses are computed here by one-sided
 * Jacobi SVD (the reference's 4GINV.C is not in its repository).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cq_lowrank.h"

#define LR_PI 3.1415926   /* the reference's constant (funca.c:6, func.c:6) */

double cqh_lr_symbol(int symbol, double k2, double gam, double vwd, double cc)
{
    if (symbol == CQH_LR_TAYLOR4) {
        /* sum_total, acsglw.cu:1232-1240, with A, K, B the |k|^(1/16), |k|^2,
         * |k|^(33/16) multiples of div */
        const double s = pow(k2, 1.0 / 32.0);
        return (1.0 - 32.0 * gam) + 32.0 * gam * vwd * s - cc * ((1.0 - 64.0 * gam) * k2 + 64.0 * gam * vwd * k2 * s);
    }
    /* a^(2 gam) - cc k2 a^(4 gam), a = vwd^16 |k| */
    if (k2 <= 0.0) return gam > 0.0 ? 0.0 : 1.0;
    const double a = pow(vwd, 16.0) * sqrt(k2);
    return pow(a, 2.0 * gam) - cc * k2 * pow(a, 4.0 * gam);
}

void cqh_lr_kgrid(int nxx, int nzz, float dx, float dz, double *k2)
{
    const int nxh = nxx / 2 + 1, nzh = nzz / 2 + 1;
    const float dkx = 2 * LR_PI / (nxx * dx), dkz = 2 * LR_PI / (nzz * dz);   /* cal_k2d, func.c:555-556 */
    for (int i = 0; i < nxh; i++)
        for (int j = 0; j < nzh; j++) {
            const float kx = i * dkx, kz = j * dkz;
            const float s = kx * kx + kz * kz;   /* float, as k2d1 */
            k2[(size_t)i * nzh + j] = s;
        }
}

/* ---- seeded sampling ---------------------------------------------------- */
static unsigned long long sm_next(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* m distinct indices of 0..n-1 (random_select, sglowrank.c:5-23) */
static void random_select(int n, int m, unsigned long long *seed, int *out)
{
    char *used = (char *)calloc((size_t)n, 1);
    for (int i = 0; i < m;) {
        const int r = (int)(sm_next(seed) % (unsigned long long)n);
        if (used[r]) continue;
        used[r] = 1;
        out[i++] = r;
    }
    free(used);
}

/* ---- pseudo-inverse by one-sided Jacobi SVD ------------------------------ */
/* P (n x m, row-major) = pinv(B) for B (m x n, row-major) with m >= n. */
static int pinv_tall(const double *B, int m, int n, double *P)
{
    double *W = (double *)malloc(sizeof(double) * (size_t)m * n);   /* columns of B: W[c * m + i] */
    double *V = (double *)calloc((size_t)n * n, sizeof(double));    /* V[c * n + r] */
    if (!W || !V) {
        free(W);
        free(V);
        return -1;
    }
    for (int i = 0; i < m; i++)
        for (int c = 0; c < n; c++) W[(size_t)c * m + i] = B[(size_t)i * n + c];
    for (int c = 0; c < n; c++) V[(size_t)c * n + c] = 1.0;
    for (int sweep = 0; sweep < 80; sweep++) {
        double off = 0.0;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double al = 0, be = 0, ga = 0;
                double *wp = W + (size_t)p * m, *wq = W + (size_t)q * m;
                for (int i = 0; i < m; i++) {
                    al += wp[i] * wp[i];
                    be += wq[i] * wq[i];
                    ga += wp[i] * wq[i];
                }
                if (al == 0.0 || be == 0.0) continue;
                const double c0 = fabs(ga) / sqrt(al * be);
                if (c0 > off) off = c0;
                if (c0 < 1e-15) continue;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                for (int i = 0; i < m; i++) {
                    const double a = wp[i], b = wq[i];
                    wp[i] = cs * a - sn * b;
                    wq[i] = sn * a + cs * b;
                }
                double *vp = V + (size_t)p * n, *vq = V + (size_t)q * n;
                for (int i = 0; i < n; i++) {
                    const double a = vp[i], b = vq[i];
                    vp[i] = cs * a - sn * b;
                    vq[i] = sn * a + cs * b;
                }
            }
        if (off < 1e-15) break;
    }
    /* singular values are the column norms; pinv = sum_c v_c u_c^T / s_c */
    double smax = 0.0;
    double *sv = (double *)malloc(sizeof(double) * (size_t)n);
    for (int c = 0; c < n; c++) {
        double s = 0;
        for (int i = 0; i < m; i++) s += W[(size_t)c * m + i] * W[(size_t)c * m + i];
        sv[c] = sqrt(s);
        if (sv[c] > smax) smax = sv[c];
    }
    memset(P, 0, sizeof(double) * (size_t)n * m);
    int kept = 0;
    for (int c = 0; c < n; c++) {
        /* rank cut at 1e-8 of the largest singular value (the reference
         * calls ginv with eps = 1e-8, sglowrank.c:438): smaller ones are the
         * sampling noise of an operator of lower numerical rank */
        if (!(sv[c] > 1e-8 * smax)) continue;
        kept++;
        const double inv2 = 1.0 / (sv[c] * sv[c]);   /* u_c / s_c = w_c / s_c^2 */
        for (int r = 0; r < n; r++) {
            const double vr = V[(size_t)c * n + r] * inv2;
            for (int i = 0; i < m; i++) P[(size_t)r * m + i] += vr * W[(size_t)c * m + i];
        }
    }
    free(sv);
    free(W);
    free(V);
    return kept > 0 ? 0 : -1;
}

/* P (n x m) = pinv(B) for B (m x n), any shape. */
static int pinv(const double *B, int m, int n, double *P)
{
    if (m >= n) return pinv_tall(B, m, n, P);
    double *Bt = (double *)calloc((size_t)m * n, sizeof(double)), *Pt = (double *)calloc((size_t)m * n, sizeof(double));
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) Bt[(size_t)j * m + i] = B[(size_t)i * n + j];
    const int rc = pinv_tall(Bt, n, m, Pt);   /* Pt: m x n = pinv(B^T) */
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) P[(size_t)j * m + i] = Pt[(size_t)i * n + j];
    free(Bt);
    free(Pt);
    return rc;
}

/* ---- decomposition -------------------------------------------------------- */
static double cell_symbol(int symbol, const double *k2, int k, const float *gam, const float *vwd, const float *cc,
                          size_t x)
{
    return cqh_lr_symbol(symbol, k2[k], gam[x], vwd[x], cc[x]);
}

/* Pivoted Gram-Schmidt over the columns of S (rows x cols, row-major):
 * picks `pick` columns of largest remaining norm, orthogonalising the rest
 * against each pick (sglowrank.c:192-246 for columns, :300-357 for rows,
 * with the matrix transposed).  sel[] receives the chosen column indices. */
static void pivot_columns(double *S, int rows, int cols, int pick, int *sel)
{
    char *free_ = (char *)malloc((size_t)cols);
    double *piv = (double *)malloc(sizeof(double) * (size_t)rows);
    memset(free_, 1, (size_t)cols);
    for (int n = 0; n < pick; n++) {
        double mx = -1.0;
        int best = -1;
        for (int j = 0; j < cols; j++) {
            if (!free_[j]) continue;
            double s = 0;
            for (int i = 0; i < rows; i++) s += S[(size_t)i * cols + j] * S[(size_t)i * cols + j];
            if (s > mx) {
                mx = s;
                best = j;
            }
        }
        sel[n] = best;
        free_[best] = 0;
        double pp = 0;
        for (int i = 0; i < rows; i++) {
            piv[i] = S[(size_t)i * cols + best];
            pp += piv[i] * piv[i];
        }
        if (n == pick - 1 || pp == 0.0) continue;
        for (int j = 0; j < cols; j++) {
            if (!free_[j]) continue;
            double d = 0;
            for (int i = 0; i < rows; i++) d += piv[i] * S[(size_t)i * cols + j];
            const double f = d / pp;
            for (int i = 0; i < rows; i++) S[(size_t)i * cols + j] -= f * piv[i];
        }
    }
    free(free_);
    free(piv);
}

int cqh_lowrank(int nxx, int nzz, float dx, float dz, const float *gam, const float *vwd, const float *cc,
                int symbol, int rank, unsigned long long seed, double *U, double *A, double *V)
{
    if (rank < 1 || rank > 16 || nxx < 2 || nzz < 2 || !gam || !vwd || !cc || !U || !A || !V) return -1;
    const int N1 = nxx * nzz, Nq = (nxx / 2 + 1) * (nzz / 2 + 1);
    const int col = rank, row = rank;
    if (rank > Nq || rank > N1) return -1;
    double *k2 = (double *)malloc(sizeof(double) * (size_t)Nq);
    cqh_lr_kgrid(nxx, nzz, dx, dz, k2);
    unsigned long long st = seed;

    /* columns (wavenumbers) from mm random cells (sglowrank.c:157-246) */
    int mm = (int)(col * log((double)N1));
    if (mm < col) mm = col;
    if (mm > N1) mm = N1;
    int *r1 = (int *)malloc(sizeof(int) * (size_t)mm);
    random_select(N1, mm, &st, r1);
    double *Sr = (double *)malloc(sizeof(double) * (size_t)mm * Nq);
    for (int i = 0; i < mm; i++)
        for (int k = 0; k < Nq; k++) Sr[(size_t)i * Nq + k] = cell_symbol(symbol, k2, k, gam, vwd, cc, (size_t)r1[i]);
    int *jrr = (int *)malloc(sizeof(int) * (size_t)col);
    pivot_columns(Sr, mm, Nq, col, jrr);
    free(Sr);
    for (size_t x = 0; x < (size_t)N1; x++)
        for (int j = 0; j < col; j++) U[x * col + j] = cell_symbol(symbol, k2, jrr[j], gam, vwd, cc, x);

    /* rows (cells) from nn random wavenumbers (sglowrank.c:267-371) */
    int nn = (int)(row * log((double)Nq));
    if (nn < row) nn = row;
    if (nn > Nq) nn = Nq;
    int *r2 = (int *)malloc(sizeof(int) * (size_t)nn);
    random_select(Nq, nn, &st, r2);
    /* S^T: nn x N1, so that rows of W become columns */
    double *Sc = (double *)malloc(sizeof(double) * (size_t)nn * N1);
    for (int j = 0; j < nn; j++)
        for (size_t x = 0; x < (size_t)N1; x++) Sc[(size_t)j * N1 + x] = cell_symbol(symbol, k2, r2[j], gam, vwd, cc, x);
    int *irr = (int *)malloc(sizeof(int) * (size_t)row);
    pivot_columns(Sc, nn, N1, row, irr);
    free(Sc);
    for (int i = 0; i < row; i++)
        for (int k = 0; k < Nq; k++) V[(size_t)i * Nq + k] = cell_symbol(symbol, k2, k, gam, vwd, cc, (size_t)irr[i]);

    /* A = pinv(Us) M pinv(Vs) on np random cells and wavenumbers (:374-463) */
    const int np = mm < nn ? mm : nn;
    int *r3 = (int *)malloc(sizeof(int) * (size_t)np), *r4 = (int *)malloc(sizeof(int) * (size_t)np);
    random_select(N1, np, &st, r3);
    random_select(Nq, np, &st, r4);
    double *Us = (double *)malloc(sizeof(double) * (size_t)np * col);
    double *Vs = (double *)malloc(sizeof(double) * (size_t)row * np);
    double *M = (double *)malloc(sizeof(double) * (size_t)np * np);
    for (int i = 0; i < np; i++)
        for (int j = 0; j < col; j++) Us[(size_t)i * col + j] = U[(size_t)r3[i] * col + j];
    for (int i = 0; i < row; i++)
        for (int j = 0; j < np; j++) Vs[(size_t)i * np + j] = V[(size_t)i * Nq + r4[j]];
    for (int i = 0; i < np; i++)
        for (int j = 0; j < np; j++) M[(size_t)i * np + j] = cell_symbol(symbol, k2, r4[j], gam, vwd, cc, (size_t)r3[i]);
    double *Pu = (double *)malloc(sizeof(double) * (size_t)col * np);
    double *Pv = (double *)malloc(sizeof(double) * (size_t)np * row);
    int rc = pinv(Us, np, col, Pu);
    if (rc == 0) rc = pinv(Vs, row, np, Pv);
    if (rc == 0) {
        double *T = (double *)calloc((size_t)col * np, sizeof(double));
        for (int a = 0; a < col; a++)
            for (int b = 0; b < np; b++) {
                double s = 0;
                for (int c = 0; c < np; c++) s += Pu[(size_t)a * np + c] * M[(size_t)c * np + b];
                T[(size_t)a * np + b] = s;
            }
        for (int a = 0; a < col; a++)
            for (int b = 0; b < row; b++) {
                double s = 0;
                for (int c = 0; c < np; c++) s += T[(size_t)a * np + c] * Pv[(size_t)c * row + b];
                A[(size_t)a * row + b] = s;
            }
        free(T);
    }
    free(k2); free(r1); free(r2); free(r3); free(r4); free(jrr); free(irr);
    free(Us); free(Vs); free(M); free(Pu); free(Pv);
    return rc;
}

void cqh_lowrank_pack(int nxx, int nzz, int rank, const double *U, const double *A, const double *V,
                      float *wts, float *sym)
{
    const size_t N1 = (size_t)nxx * nzz, Nq = (size_t)(nxx / 2 + 1) * (nzz / 2 + 1);
    for (size_t x = 0; x < N1; x++)
        for (int i = 0; i < rank; i++) {
            double s = 0;
            for (int j = 0; j < rank; j++) s += U[x * rank + j] * A[(size_t)j * rank + i];
            wts[(size_t)i * N1 + x] = (float)s;
        }
    const double inv = 1.0 / ((double)nxx * nzz);
    for (int i = 0; i < rank; i++)
        for (size_t k = 0; k < Nq; k++) sym[(size_t)i * Nq + k] = (float)(V[(size_t)i * Nq + k] * inv);
}

double cqh_lowrank_error(int nxx, int nzz, float dx, float dz, const float *gam, const float *vwd,
                         const float *cc, int symbol, int rank, const double *U, const double *A,
                         const double *V, int ncell, const int *cells)
{
    const int Nq = (nxx / 2 + 1) * (nzz / 2 + 1);
    double *k2 = (double *)malloc(sizeof(double) * (size_t)Nq);
    double *tp = (double *)malloc(sizeof(double) * (size_t)rank);
    cqh_lr_kgrid(nxx, nzz, dx, dz, k2);
    double emax = 0.0, wmax = 0.0;
    for (int c = 0; c < ncell; c++) {
        const size_t x = (size_t)cells[c];
        for (int i = 0; i < rank; i++) {
            double s = 0;
            for (int j = 0; j < rank; j++) s += U[x * rank + j] * A[(size_t)j * rank + i];
            tp[i] = s;
        }
        for (int k = 0; k < Nq; k++) {
            const double we = cell_symbol(symbol, k2, k, gam, vwd, cc, x);
            double wa = 0;
            for (int i = 0; i < rank; i++) wa += tp[i] * V[(size_t)i * Nq + k];
            if (fabs(we - wa) > emax) emax = fabs(we - wa);
            if (fabs(we) > wmax) wmax = fabs(we);
        }
    }
    free(k2);
    free(tp);
    return wmax > 0 ? emax / wmax : emax;
}
