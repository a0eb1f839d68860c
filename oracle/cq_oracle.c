/*
 * cq_oracle.c -- CPU restatement of the reference's per-time-step PSTD kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this file:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and only as the checker (or, for the CPU baseline, as the timed reference-
 * faithful CPU path).
 *
 * The same source is compiled twice:
 *   -DORC_REAL=double  -> liboracle64.so : the FP64 oracle used for parity;
 *   -DORC_REAL=float   -> liboracle32.so : the FP32 "reference-faithful" CPU
 *                                          path timed as the CPU baseline.
 *
 * It follows the reference step for step: full 2-D complex-to-complex FFTs of
 * the padded nxx x nzz grid (z fastest), the same multiplier kernels, the same
 * real-part extraction and the same update order.
 *   acoustic step   : sgpstd-qmodeling/acps4.cu:508-645 (kernels :1001-1730)
 *   viscoacoustic   : sgpstd-qmodeling/acsglw.cu:511-683 (kernels :1213-1245,
 *                     :1500-1537)
 *   RTM / Q-RTM     : sgpstd-qrtm/acps4.cu:447-945, sgpstd-qrtm/acsglw.cu:470-1079
 * The FFT is this file's own mixed-radix transform (no FFTW in the image; prime
 * factors above 64 by Bluestein's chirp-z algorithm over a power-of-two
 * length): the unnormalised DFT with exp(-i) forward and exp(+i) inverse,
 * which is what the reference's cufftExecC2C computes (third-party cuFFT,
 * version unpinned).
 *
 * Coefficient arrays (CPML profiles, shift multipliers, |k| powers, windowed
 * models) are inputs: the tests pass the FP32 arrays the reference's own host
 * builders produce, so the oracle runs exactly the arithmetic the reference's
 * kernels would run on them.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
typedef struct { real re, im; } cpx;

#define ORC_MAXFAC 40
#define ORC_MAXRADIX 4096

/* ------------------------------------------------------------------------ */
/* Mixed-radix FFT                                                           */
/* ------------------------------------------------------------------------ */

typedef struct fplan_s {
    int n;
    int nf;
    int fac[ORC_MAXFAC];
    cpx *tw; /* tw[j] = exp(-2*pi*i*j/n), computed in double */
    /* lengths with a prime factor above ORC_BLUE_PRIME: Bluestein's chirp-z
     * transform over M = the smallest power of two >= 2n - 1 points */
    int M;
    struct fplan_s *pm;   /* plan of the M-point transforms */
    cpx *chirp;           /* exp(-i pi j^2 / n), j < n */
    cpx *bspec;           /* FFT_M of the conjugate chirp wrapped cyclically, / M */
} fplan;

#define ORC_BLUE_PRIME 64

static void fft_line(const struct fplan_s *pl, const cpx *in, int xs, cpx *out, int inv);

static void plan_init(fplan *p, int n)
{
    int m = n, f;
    p->n = n;
    p->nf = 0;
    p->M = 0;
    p->pm = NULL;
    p->chirp = p->bspec = NULL;
    /* radix 4 first, then 2, 3, 5, then odd factors in increasing order */
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (f = 3; f * f <= m; f += 2)
        while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
    if (m > 1) p->fac[p->nf++] = m;
    p->tw = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    for (int j = 0; j < n; j++) {
        double a = -2.0 * M_PI * (double)j / (double)n;
        p->tw[j].re = (real)cos(a);
        p->tw[j].im = (real)sin(a);
    }
    if (p->fac[p->nf - 1] > ORC_BLUE_PRIME) {
        /* X[k] = c[k] sum_j (x[j] c[j]) conj c[k - j], c[j] = exp(-i pi j^2 / n):
         * a linear convolution of n-point sequences, done cyclically on M points */
        int M = 1;
        while (M < 2 * n - 1) M *= 2;
        p->M = M;
        p->pm = (fplan *)malloc(sizeof(fplan));
        plan_init(p->pm, M);
        p->chirp = (cpx *)malloc(sizeof(cpx) * (size_t)n);
        p->bspec = (cpx *)malloc(sizeof(cpx) * (size_t)M);
        cpx *b = (cpx *)calloc((size_t)M, sizeof(cpx));
        for (int j = 0; j < n; j++) {
            const long long r = (long long)j * j % (2LL * n);   /* exact phase reduction */
            const double a = -M_PI * (double)r / (double)n;
            p->chirp[j].re = (real)cos(a);
            p->chirp[j].im = (real)sin(a);
            b[j].re = (real)cos(a);
            b[j].im = (real)-sin(a);
            if (j) b[M - j] = b[j];
        }
        fft_line(p->pm, b, 1, p->bspec, 0);
        for (int k = 0; k < M; k++) {
            p->bspec[k].re /= (real)M;
            p->bspec[k].im /= (real)M;
        }
        free(b);
    }
}

static void plan_free(fplan *p)
{
    free(p->tw);
    p->tw = NULL;
    if (p->pm) {
        plan_free(p->pm);
        free(p->pm);
        free(p->chirp);
        free(p->bspec);
        p->pm = NULL;
    }
}

/* Recursive decimation in time: x has stride xs, X is contiguous.  At a level
 * of length n the twiddle stride into the top-level table is tws = N/n; the
 * children (length n/p) use tws*p. */
static void fft_rec_top(const cpx *x, int xs, cpx *X, int n, const int *fac,
                        const cpx *tw, int tws, int inv, cpx *tmp)
{
    if (n == 1) { X[0] = x[0]; return; }
    int p = fac[0], m = n / p;
    for (int r = 0; r < p; r++)
        fft_rec_top(x + (size_t)r * xs, xs * p, X + (size_t)r * m, m, fac + 1, tw, tws * p, inv, tmp);
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            cpx y = X[(size_t)r * m + k];
            cpx w = tw[(long)r * k * tws];
            if (inv) w.im = -w.im;
            tmp[r].re = y.re * w.re - y.im * w.im;
            tmp[r].im = y.re * w.im + y.im * w.re;
        }
        for (int q = 0; q < p; q++) {
            real sr = 0, si = 0;
            for (int r = 0; r < p; r++) {
                cpx w = tw[(long)((r * q) % p) * tws * m];
                if (inv) w.im = -w.im;
                sr += tmp[r].re * w.re - tmp[r].im * w.im;
                si += tmp[r].re * w.im + tmp[r].im * w.re;
            }
            X[(size_t)q * m + k].re = sr;
            X[(size_t)q * m + k].im = si;
        }
    }
}

/* One line: in (stride xs) -> out (contiguous).  An inverse Bluestein
 * transform is the conjugate of the forward transform of the conjugate. */
static void fft_line(const fplan *pl, const cpx *in, int xs, cpx *out, int inv)
{
    if (pl->pm) {
        const int n = pl->n, M = pl->M;
        cpx *a = (cpx *)calloc((size_t)M * 2, sizeof(cpx)), *A = a + M;
        for (int j = 0; j < M; j++) {
            if (j < n) {
                cpx x = in[(size_t)j * xs], c = pl->chirp[j];
                if (inv) x.im = -x.im;
                a[j].re = x.re * c.re - x.im * c.im;
                a[j].im = x.re * c.im + x.im * c.re;
            } else {
                a[j].re = a[j].im = 0;
            }
        }
        fft_line(pl->pm, a, 1, A, 0);
        for (int k = 0; k < M; k++) {
            const cpx u = A[k], w = pl->bspec[k];
            A[k].re = u.re * w.re - u.im * w.im;
            A[k].im = u.re * w.im + u.im * w.re;
        }
        fft_line(pl->pm, A, 1, a, 1);
        for (int k = 0; k < n; k++) {
            const cpx r = a[k], c = pl->chirp[k];
            out[k].re = r.re * c.re - r.im * c.im;
            out[k].im = r.re * c.im + r.im * c.re;
            if (inv) out[k].im = -out[k].im;
        }
        free(a);
        return;
    }
    cpx tmp[ORC_MAXRADIX > 64 ? 64 : ORC_MAXRADIX];
    cpx *t = tmp;
    int big = 0;
    for (int i = 0; i < pl->nf; i++) if (pl->fac[i] > 64) big = 1;
    if (big) t = (cpx *)malloc(sizeof(cpx) * (size_t)pl->n);
    fft_rec_top(in, xs, out, pl->n, pl->fac, pl->tw, 1, inv, t);
    if (big) free(t);
}

/* 2-D C2C, layout a[ix*nzz + iz], unnormalised; inv=1 uses exp(+i). */
static void fft2d(const fplan *px, const fplan *pz, cpx *a, int nxx, int nzz, int inv)
{
    #pragma omp parallel
    {
        cpx *lo = (cpx *)malloc(sizeof(cpx) * (size_t)(nxx > nzz ? nxx : nzz));
        #pragma omp for schedule(static)
        for (int ix = 0; ix < nxx; ix++) {
            fft_line(pz, a + (size_t)ix * nzz, 1, lo, inv);
            memcpy(a + (size_t)ix * nzz, lo, sizeof(cpx) * (size_t)nzz);
        }
        #pragma omp for schedule(static)
        for (int iz = 0; iz < nzz; iz++) {
            fft_line(px, a + iz, nzz, lo, inv);
            for (int ix = 0; ix < nxx; ix++) a[(size_t)ix * nzz + iz] = lo[ix];
        }
        free(lo);
    }
}

/* Exported 1-D / 2-D transforms for the FFT tests (interleaved re,im). */
void orc_fft1d(int n, const real *in, real *out, int inv)
{
    fplan p;
    plan_init(&p, n);
    fft_line(&p, (const cpx *)in, 1, (cpx *)out, inv);
    plan_free(&p);
}

void orc_fft2d(int nxx, int nzz, real *data, int inv)
{
    fplan px, pz;
    plan_init(&px, nxx);
    plan_init(&pz, nzz);
    fft2d(&px, &pz, (cpx *)data, nxx, nzz, inv);
    plan_free(&px);
    plan_free(&pz);
}

int orc_real_bytes(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------------ */
/* Per-step kernels (restated from sgpstd-qmodeling/acps4.cu)                */
/* ------------------------------------------------------------------------ */

/* Staggered derivative multiply, acps4.cu:1047-1108 (v->p, kx "sm", kz "sc")
 * and :1210-1271 (p->v, kx "sc", kz "sm"):
 *   out = +i (kr + i ki) in / NXZ   for index <= n/2,
 *   out = -i (kr + i ki) in / NXZ   for index >  n/2. */
static inline cpx shift_mul(cpx in, real kr, real ki, int neg, real inv_n)
{
    cpx o;
    if (!neg) {
        o.re = -kr * in.im - ki * in.re;
        o.im = kr * in.re - ki * in.im;
    } else {
        o.re = kr * in.im + ki * in.re;
        o.im = -kr * in.re + ki * in.im;
    }
    o.re *= inv_n;
    o.im *= inv_n;
    return o;
}

typedef struct {
    int nxx, nzz;
    real dt;
    /* window fields, nxx*nzz */
    const real *vp, *rho, *bbx, *bbz;
    const real *gam, *lam, *vwd, *tao;        /* viscoacoustic only */
    /* 1-D CPML coefficients */
    const real *iax, *ibx, *iaz, *ibz, *hax, *hbx, *haz, *hbz;
    /* 1-D shift multipliers */
    const real *kxsmr, *kxsmi, *kxscr, *kxsci, *kzsmr, *kzsmi, *kzscr, *kzsci;
    /* 2-D |k| powers, nxx*nzz */
    const real *k2d1, *k2d2, *k2d3, *k2da;
} orc_model;

typedef struct {
    real *p, *vx, *vz, *fipx, *fipz, *fivxx, *fivzz, *aux;
} orc_state;

typedef struct {
    fplan px, pz;
    cpx *d0, *d1, *d2, *d3, *d4;
    real *total;
} orc_work;

static void work_init(orc_work *w, int nxx, int nzz)
{
    size_t n = (size_t)nxx * nzz;
    plan_init(&w->px, nxx);
    plan_init(&w->pz, nzz);
    w->d0 = (cpx *)calloc(n, sizeof(cpx));
    w->d1 = (cpx *)calloc(n, sizeof(cpx));
    w->d2 = (cpx *)calloc(n, sizeof(cpx));
    w->d3 = (cpx *)calloc(n, sizeof(cpx));
    w->d4 = (cpx *)calloc(n, sizeof(cpx));
    w->total = (real *)calloc(n, sizeof(real));
}

static void work_free(orc_work *w)
{
    plan_free(&w->px); plan_free(&w->pz);
    free(w->d0); free(w->d1); free(w->d2); free(w->d3); free(w->d4); free(w->total);
}

static void state_alloc(orc_state *s, size_t n)
{
    s->p = (real *)calloc(n, sizeof(real));
    s->vx = (real *)calloc(n, sizeof(real));
    s->vz = (real *)calloc(n, sizeof(real));
    s->fipx = (real *)calloc(n, sizeof(real));
    s->fipz = (real *)calloc(n, sizeof(real));
    s->fivxx = (real *)calloc(n, sizeof(real));
    s->fivzz = (real *)calloc(n, sizeof(real));
    s->aux = (real *)calloc(n, sizeof(real));
}

static void state_free(orc_state *s)
{
    free(s->p); free(s->vx); free(s->vz); free(s->fipx); free(s->fipz);
    free(s->fivxx); free(s->fivzz); free(s->aux);
}

/* v -> div: fuzhi_kernel, 2 FFT, shift_vxz_kernel, 2 IFFT, update_vxx_vzz,
 * sumpxpz (acps4.cu:517-547).  Leaves div (complex) in w->d0. */
static void step_div(const orc_model *m, orc_state *s, orc_work *w)
{
    const int nxx = m->nxx, nzz = m->nzz;
    const size_t N = (size_t)nxx * nzz;
    const real inv_n = (real)1 / (real)(double)N;
    for (size_t i = 0; i < N; i++) {
        w->d0[i].re = s->vx[i]; w->d0[i].im = 0;
        w->d1[i].re = s->vz[i]; w->d1[i].im = 0;
    }
    fft2d(&w->px, &w->pz, w->d0, nxx, nzz, 0);
    fft2d(&w->px, &w->pz, w->d1, nxx, nzz, 0);
    #pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nxx; ix++)
        for (int iz = 0; iz < nzz; iz++) {
            size_t a = (size_t)ix * nzz + iz;
            w->d2[a] = shift_mul(w->d0[a], m->kxsmr[ix], m->kxsmi[ix], ix > nxx / 2, inv_n);
            w->d3[a] = shift_mul(w->d1[a], m->kzscr[iz], m->kzsci[iz], iz > nzz / 2, inv_n);
        }
    fft2d(&w->px, &w->pz, w->d2, nxx, nzz, 1);
    fft2d(&w->px, &w->pz, w->d3, nxx, nzz, 1);
    /* update_vxx_vzz (acps4.cu:1111-1144) and sumpxpz (:1146-1172) */
    for (int ix = 0; ix < nxx; ix++)
        for (int iz = 0; iz < nzz; iz++) {
            size_t a = (size_t)ix * nzz + iz;
            s->fivxx[a] = m->ibx[ix] * s->fivxx[a] + m->iax[ix] * w->d2[a].re;
            w->d2[a].re = w->d2[a].re + s->fivxx[a];
            s->fivzz[a] = m->ibz[iz] * s->fivzz[a] + m->iaz[iz] * w->d3[a].re;
            w->d3[a].re = w->d3[a].re + s->fivzz[a];
            w->d0[a].re = w->d2[a].re + w->d3[a].re;
            w->d0[a].im = w->d2[a].im + w->d3[a].im;
        }
}

/* dst = IFFT( FFT(src) * k / NXZ ) with k a 2-D real array (multtt_kernel,
 * acps4.cu:1709-1730).  fsrc holds FFT(src) already. */
static void mult_k_ifft(const orc_model *m, orc_work *w, const cpx *fsrc, const real *k, cpx *dst)
{
    const int nxx = m->nxx, nzz = m->nzz;
    const size_t N = (size_t)nxx * nzz;
    const real inv_n = (real)1 / (real)(double)N;
    for (size_t i = 0; i < N; i++) {
        dst[i].re = fsrc[i].re * k[i] * inv_n;
        dst[i].im = fsrc[i].im * k[i] * inv_n;
    }
    fft2d(&w->px, &w->pz, dst, nxx, nzz, 1);
}

/* p -> v: fuzhi_kernel_p, FFT, shift_pxz_kernel, 2 IFFT, vx_cal, vz_cal
 * (acps4.cu:571-607).  Leaves FFT(p) in w->d0 for the viscoacoustic aux. */
static void step_vel(const orc_model *m, orc_state *s, orc_work *w)
{
    const int nxx = m->nxx, nzz = m->nzz;
    const size_t N = (size_t)nxx * nzz;
    const real inv_n = (real)1 / (real)(double)N;
    const real dt = m->dt;
    for (size_t i = 0; i < N; i++) { w->d0[i].re = s->p[i]; w->d0[i].im = 0; }
    fft2d(&w->px, &w->pz, w->d0, nxx, nzz, 0);
    #pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nxx; ix++)
        for (int iz = 0; iz < nzz; iz++) {
            size_t a = (size_t)ix * nzz + iz;
            w->d1[a] = shift_mul(w->d0[a], m->kxscr[ix], m->kxsci[ix], ix > nxx / 2, inv_n);
            w->d2[a] = shift_mul(w->d0[a], m->kzsmr[iz], m->kzsmi[iz], iz > nzz / 2, inv_n);
        }
    fft2d(&w->px, &w->pz, w->d1, nxx, nzz, 1);
    fft2d(&w->px, &w->pz, w->d2, nxx, nzz, 1);
    for (int ix = 0; ix < nxx; ix++)
        for (int iz = 0; iz < nzz; iz++) {
            size_t a = (size_t)ix * nzz + iz;
            /* vx_cal_kernel, acps4.cu:1274-1303 */
            s->fipx[a] = m->hbx[ix] * s->fipx[a] + m->hax[ix] * w->d1[a].re;
            s->vx[a] = s->vx[a] + dt * m->bbx[a] * (w->d1[a].re + s->fipx[a]);
            /* vz_cal_kernel, acps4.cu:1305-1333 */
            s->fipz[a] = m->hbz[iz] * s->fipz[a] + m->haz[iz] * w->d2[a].re;
            s->vz[a] = s->vz[a] + dt * m->bbz[a] * (w->d2[a].re + s->fipz[a]);
        }
}

/* Acoustic p update (addsa_kernel, acps4.cu:1462-1498), d0 = div, d2 = L. */
static void p_update_ac(const orc_model *m, orc_state *s, orc_work *w)
{
    const size_t N = (size_t)m->nxx * m->nzz;
    const real dt = m->dt;
    for (size_t a = 0; a < N; a++) {
        real v = m->vp[a];
        s->p[a] = s->p[a] + m->rho[a] * v * v * dt *
                  (w->d0[a].re - (real)(1.0 / 12.0) * v * dt * v * dt * w->d2[a].re);
    }
}

/* The rank-r low-rank operator (set by orc_set_lowrank; rank 0: none).  U is
 * N1 x r (U[x * r + j]), A is r x r (A[j * r + i]), V is r x Nq on the
 * quarter plane (V[i * Nq + ikx * nzh + ikz]), as cqh_lowrank returns them. */
static struct {
    int rank;
    const real *U, *A, *V;
} g_lr;

void orc_set_lowrank(int rank, const real *U, const real *A, const real *V)
{
    g_lr.rank = rank;
    g_lr.U = U;
    g_lr.A = A;
    g_lr.V = V;
}

/* total = sum_j U(x, j) sum_i A(j, i) Re IFFT(V_i FFT div / N): the apply
 * kernels of the reference's low-rank path (sgpstd-qmodeling/acps4.cu:
 * 1337-1459): multv_kernel (V mirrored to the full grid as consist_Vt,
 * sglowrank.c:104-145), the inverse transform, fuzhi_r_kernel + sumlw_kernel
 * (accumulators in double) and sumlw2_kernel.  fdiv holds FFT(div). */
static void lowrank_total(const orc_model *m, orc_work *w, const cpx *fdiv)
{
    const int nxx = m->nxx, nzz = m->nzz, r = g_lr.rank, nzh = nzz / 2 + 1;
    const size_t N = (size_t)nxx * nzz, Nq = (size_t)(nxx / 2 + 1) * nzh;
    const real inv_n = (real)1 / (real)(double)N;
    double *pst = (double *)calloc(N * (size_t)r, sizeof(double));
    for (int i = 0; i < r; i++) {
        for (int ix = 0; ix < nxx; ix++)
            for (int iz = 0; iz < nzz; iz++) {
                const size_t a = (size_t)ix * nzz + iz;
                const int ikx = ix <= nxx / 2 ? ix : nxx - ix, ikz = iz <= nzz / 2 ? iz : nzz - iz;
                const real v = g_lr.V[(size_t)i * Nq + (size_t)ikx * nzh + ikz];
                w->d2[a].re = fdiv[a].re * v * inv_n;
                w->d2[a].im = fdiv[a].im * v * inv_n;
            }
        fft2d(&w->px, &w->pz, w->d2, nxx, nzz, 1);
        for (int j = 0; j < r; j++) {
            const double aji = (double)g_lr.A[(size_t)j * r + i];
            for (size_t a = 0; a < N; a++) pst[(size_t)j * N + a] += aji * (double)w->d2[a].re;
        }
    }
    for (size_t a = 0; a < N; a++) {
        double t = 0.0;
        for (int j = 0; j < r; j++) t += (double)g_lr.U[a * r + j] * pst[(size_t)j * N + a];
        w->total[a] = (real)t;
    }
    free(pst);
}

/* Viscoacoustic: sum_total (acsglw.cu:1213-1245), FFT(total)*k2da
 * (:578-588) and the p update (addsa_kernel, :1500-1537).  sgn = +1 for
 * modeling; -1 flips the tao terms for Q-compensated back-propagation
 * (sgpstd-qrtm/acsglw.cu:1769-1806). */
static void p_update_q(const orc_model *m, orc_state *s, orc_work *w, real sgn)
{
    const int nxx = m->nxx, nzz = m->nzz;
    const size_t N = (size_t)nxx * nzz;
    const real dt = m->dt;
    /* d1 = FFT(div) */
    memcpy(w->d1, w->d0, sizeof(cpx) * N);
    fft2d(&w->px, &w->pz, w->d1, nxx, nzz, 0);
    if (g_lr.rank > 0) {
        lowrank_total(m, w, w->d1);
    } else {
    mult_k_ifft(m, w, w->d1, m->k2d2, w->d2);
    mult_k_ifft(m, w, w->d1, m->k2d1, w->d3);
    mult_k_ifft(m, w, w->d1, m->k2d3, w->d4);
    for (size_t a = 0; a < N; a++) {
        real g = m->gam[a];
        real y1 = (real)1.0 - (real)32.0 * g;
        real y2 = (real)32.0 * g * m->vwd[a];
        real y3 = (real)1.0 - (real)64.0 * g;
        real y4 = (real)64.0 * g * m->vwd[a];
        real vpf = m->vp[a] * m->vp[a];
        real dtf = dt * dt;
        w->total[a] = y1 * w->d0[a].re + y2 * w->d2[a].re -
                      (real)(1.0 / 12.0) * dtf * vpf * m->lam[a] * (y3 * w->d3[a].re + y4 * w->d4[a].re);
    }
    }
    for (size_t a = 0; a < N; a++) { w->d1[a].re = w->total[a]; w->d1[a].im = 0; }
    fft2d(&w->px, &w->pz, w->d1, nxx, nzz, 0);
    mult_k_ifft(m, w, w->d1, m->k2da, w->d0);
    for (size_t a = 0; a < N; a++) {
        real v = m->vp[a];
        s->p[a] = s->p[a] + m->rho[a] * v * v * dt * m->lam[a] *
                  (w->total[a] - sgn * (real)0.5 * dt * m->tao[a] * w->d0[a].re) -
                  sgn * dt * m->tao[a] * s->aux[a];
    }
}

/* aux = Re IFFT(|k|^1.01 FFT(p))/NXZ with FFT(p) in d0 (acsglw.cu:623-631);
 * for Q-RTM the multiplier is the stabilised one (sgpstd-qrtm/acsglw.cu:1895-1924). */
static void aux_update(const orc_model *m, orc_state *s, orc_work *w, int stab,
                       real eta, int nt, real taoc)
{
    const int nxx = m->nxx, nzz = m->nzz;
    const size_t N = (size_t)nxx * nzz;
    const real inv_n = (real)1 / (real)(double)N;
    const real dt = m->dt;
    for (size_t a = 0; a < N; a++) {
        real k = m->k2da[a];
        real f = k;
        if (stab) {
            real tt = (real)nt * dt;
            f = k - (real)2.0 * eta * k * ((real)1.0 + taoc * dt * k) / (real)1000000.0 /
                        (exp(-taoc * tt * k) + eta * exp(taoc * dt * k) / (real)1000000.0);
        }
        w->d3[a].re = w->d0[a].re * f * inv_n;
        w->d3[a].im = w->d0[a].im * f * inv_n;
    }
    fft2d(&w->px, &w->pz, w->d3, nxx, nzz, 1);
    for (size_t a = 0; a < N; a++) s->aux[a] = w->d3[a].re;
}

/* ------------------------------------------------------------------------ */
/* Modeling drivers                                                          */
/* ------------------------------------------------------------------------ */

static void model_from_args(orc_model *m, int nxx, int nzz, real dt,
                            const real *const *f, const real *const *c1d,
                            const real *const *k2)
{
    memset(m, 0, sizeof(*m));
    m->nxx = nxx; m->nzz = nzz; m->dt = dt;
    m->vp = f[0]; m->rho = f[1]; m->bbx = f[2]; m->bbz = f[3];
    m->gam = f[4]; m->lam = f[5]; m->vwd = f[6]; m->tao = f[7];
    m->iax = c1d[0]; m->ibx = c1d[1]; m->iaz = c1d[2]; m->ibz = c1d[3];
    m->hax = c1d[4]; m->hbx = c1d[5]; m->haz = c1d[6]; m->hbz = c1d[7];
    m->kxsmr = c1d[8]; m->kxsmi = c1d[9]; m->kxscr = c1d[10]; m->kxsci = c1d[11];
    m->kzsmr = c1d[12]; m->kzsmi = c1d[13]; m->kzscr = c1d[14]; m->kzsci = c1d[15];
    m->k2d1 = k2[0]; m->k2d2 = k2[1]; m->k2d3 = k2[2]; m->k2da = k2[3];
}

/*
 * One-shot forward modeling.  kind 0 = acoustic (acps4.cu:508-645), kind 1 =
 * viscoacoustic (acsglw.cu:511-683).
 *   fields[8] : vp, rho, bbx, bbz, gam, lam, vwd, tao (window, nxx*nzz; the
 *               last four only for kind 1)
 *   c1d[16]   : iax, ibx, iaz, ibz, hax, hbx, haz, hbz, kxsmr, kxsmi, kxscr,
 *               kxsci, kzsmr, kzsmi, kzscr, kzsci
 *   k2[4]     : k2d1, k2d2, k2d3, k2da (nxx*nzz; k2d2/k2d3/k2da kind 1 only)
 *   rec       : nt*nr gather, time-major (acps4.cu:611-621)
 *   p_out     : optional final p (nxx*nzz)
 * Returns 0.
 */
int orc_forward_ic(int kind, int nxx, int nzz, int nt, real dt,
                   const real *const *fields, const real *const *c1d, const real *const *k2,
                   const real *rik, int xsg, int zsg,
                   int nr, const int *gxg, const int *gzg, real *rec, real *p_out,
                   real *vx_out, real *vz_out, const real *p0, const real *vx0, const real *vz0)
{
    orc_model m;
    orc_state s;
    orc_work w;
    const size_t N = (size_t)nxx * nzz;
    model_from_args(&m, nxx, nzz, dt, fields, c1d, k2);
    state_alloc(&s, N);
    work_init(&w, nxx, nzz);
    /* optional initial wavefield (the reference always starts from rest,
     * initialize_kernel; an initial state only serves the closed-form tests) */
    if (p0) memcpy(s.p, p0, sizeof(real) * N);
    if (vx0) memcpy(s.vx, vx0, sizeof(real) * N);
    if (vz0) memcpy(s.vz, vz0, sizeof(real) * N);
    for (int t = 0; t < nt; t++) {
        step_div(&m, &s, &w);
        if (kind == 0) {
            memcpy(w.d1, w.d0, sizeof(cpx) * N);
            fft2d(&w.px, &w.pz, w.d1, nxx, nzz, 0);
            mult_k_ifft(&m, &w, w.d1, m.k2d1, w.d2);
            p_update_ac(&m, &s, &w);
        } else {
            p_update_q(&m, &s, &w, (real)1);
        }
        s.p[(size_t)xsg * nzz + zsg] += rik[t] * dt * (real)1000.0;
        for (int i = 0; i < nr; i++)
            rec[(size_t)t * nr + i] = s.p[(size_t)gxg[i] * nzz + gzg[i]];
        step_vel(&m, &s, &w);
        if (kind == 1) aux_update(&m, &s, &w, 0, 0, 0, 0);
    }
    if (p_out) memcpy(p_out, s.p, sizeof(real) * N);
    if (vx_out) memcpy(vx_out, s.vx, sizeof(real) * N);
    if (vz_out) memcpy(vz_out, s.vz, sizeof(real) * N);
    state_free(&s);
    work_free(&w);
    return 0;
}

int orc_forward(int kind, int nxx, int nzz, int nt, real dt,
                const real *const *fields, const real *const *c1d, const real *const *k2,
                const real *rik, int xsg, int zsg,
                int nr, const int *gxg, const int *gzg, real *rec, real *p_out,
                real *vx_out, real *vz_out)
{
    return orc_forward_ic(kind, nxx, nzz, nt, dt, fields, c1d, k2, rik, xsg, zsg, nr, gxg, gzg, rec, p_out,
                          vx_out, vz_out, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* RTM / Q-RTM driver (sgpstd-qrtm/acps4.cu:447-945, acsglw.cu:470-1079)     */
/* ------------------------------------------------------------------------ */

static void state_copy(orc_state *dst, const orc_state *src, size_t n)
{
    memcpy(dst->p, src->p, sizeof(real) * n);
    memcpy(dst->vx, src->vx, sizeof(real) * n);
    memcpy(dst->vz, src->vz, sizeof(real) * n);
    memcpy(dst->fipx, src->fipx, sizeof(real) * n);
    memcpy(dst->fipz, src->fipz, sizeof(real) * n);
    memcpy(dst->fivxx, src->fivxx, sizeof(real) * n);
    memcpy(dst->fivzz, src->fivzz, sizeof(real) * n);
    memcpy(dst->aux, src->aux, sizeof(real) * n);
}

/* One time step of the RTM source (rcv == 0) or receiver (rcv == 1)
 * wavefield.  q: viscoacoustic with compensation (sign -1) and the stabilised
 * aux multiplier (eta, it = its time index).  Source: p[src] += rik[t] dt 1000.
 * Receiver: p = obs[row] at the receiver cells (addsa_b_kernel). */
static void rtm_step(const orc_model *m, orc_state *s, orc_work *w, int q, int rcv, int t,
                     const real *rik, size_t src_cell, const int *flagrec, const int *rindex,
                     const real *obs, int nr, int row, real eta, real taoc)
{
    const size_t N = (size_t)m->nxx * m->nzz;
    step_div(m, s, w);
    if (!q) {
        memcpy(w->d1, w->d0, sizeof(cpx) * N);
        fft2d(&w->px, &w->pz, w->d1, m->nxx, m->nzz, 0);
        mult_k_ifft(m, w, w->d1, m->k2d1, w->d2);
        p_update_ac(m, s, w);
    } else {
        p_update_q(m, s, w, (real)-1);
    }
    if (!rcv) {
        s->p[src_cell] += rik[t] * m->dt * (real)1000.0;
    } else {
        for (size_t a = 0; a < N; a++)
            if (flagrec[a]) s->p[a] = obs[(size_t)row * nr + rindex[a]];
    }
    step_vel(m, s, w);
    if (q) aux_update(m, s, w, 1, eta, t, taoc);
}

/*
 * RTM of one shot following a revolve action trace (act[i], capo[i],
 * check[i]; actions numbered as revolve.h: 0 advance, 1 takeshot, 2 restore,
 * 3 firsturn, 4 youturn, 5 terminate).  kind 2: acoustic RTM; 3: Q-RTM.
 * obs: nt*nr (direct wave removed).  img, ssg: raw interior image and
 * illumination, nx*nz each (nx = nxx - 2 pml), zeroed here.
 */
int orc_rtm(int kind, int nxx, int nzz, int pml, int nt, real dt,
            const real *const *fields, const real *const *c1d, const real *const *k2,
            const real *rik, int xsg, int zsg, int nr, const int *gxg, const int *gzg,
            const real *obs, const int *act, const int *capo, const int *check, int nact, int snaps,
            real etas, real etar, real *img, real *ssg)
{
    orc_model m;
    orc_state src, rcv, *slot;
    orc_work w;
    const int q = (kind == 3);
    const size_t N = (size_t)nxx * nzz;
    const int nx = nxx - 2 * pml, nz = nzz - 2 * pml;
    model_from_args(&m, nxx, nzz, dt, fields, c1d, k2);
    state_alloc(&src, N);
    state_alloc(&rcv, N);
    work_init(&w, nxx, nzz);
    slot = (orc_state *)malloc(sizeof(orc_state) * (size_t)snaps);
    for (int i = 0; i < snaps; i++) state_alloc(&slot[i], N);
    int *flagrec = (int *)calloc(N, sizeof(int));
    int *rindex = (int *)calloc(N, sizeof(int));
    for (int i = 0; i < nr; i++) {      /* acps4.cu:416-420: last receiver wins */
        size_t a = (size_t)gxg[i] * nzz + gzg[i];
        flagrec[a] = 1;
        rindex[a] = i;
    }
    real taoc = 0;
    if (q)   /* get_max_tao over the padded window (acsglw.cu:354-356) */
        for (size_t a = 0; a < N; a++)
            if (m.tao[a] >= taoc) taoc = m.tao[a];
    memset(img, 0, sizeof(real) * (size_t)nx * nz);
    memset(ssg, 0, sizeof(real) * (size_t)nx * nz);
    const size_t src_cell = (size_t)xsg * nzz + zsg;
    int ts = 0, restart = 0;
    for (int i = 0; i < nact; i++) {
        const int a = act[i];
        if (a == 1) {                    /* takeshot (+ addsource when capo == 0) */
            state_copy(&slot[check[i]], &src, N);
            if (capo[i] == 0) slot[check[i]].p[src_cell] += rik[0] * (real)1000.0 * dt;
        } else if (a == 0) {             /* advance t = ts..capo */
            for (int t = ts; t <= capo[i]; t++) {
                rtm_step(&m, &src, &w, q, 0, t, rik, src_cell, NULL, NULL, NULL, 0, 0, etas, taoc);
                if (restart == 0)        /* illumination, first sweep only */
                    for (int ix = 0; ix < nx; ix++)
                        for (int iz = 0; iz < nz; iz++) {
                            real p = src.p[(size_t)(ix + pml) * nzz + iz + pml];
                            ssg[(size_t)ix * nz + iz] += p * p;
                        }
            }
            ts = capo[i] + 1;
        } else if (a == 3 || a == 4) {   /* firsturn / youturn: receiver step + image */
            if (a == 3) restart = 1;
            const int row = (a == 3) ? nt - 1 : capo[i];
            rtm_step(&m, &rcv, &w, q, 1, row, rik, src_cell, flagrec, rindex, obs, nr, row, etar, taoc);
            for (int ix = 0; ix < nx; ix++)
                for (int iz = 0; iz < nz; iz++) {
                    size_t c = (size_t)(ix + pml) * nzz + iz + pml;
                    img[(size_t)ix * nz + iz] += src.p[c] * rcv.p[c];
                }
        } else if (a == 2) {             /* restore */
            state_copy(&src, &slot[check[i]], N);
            ts = capo[i] + 1;
        } else if (a == 5) {
            break;
        } else {
            break;
        }
    }
    for (int i = 0; i < snaps; i++) state_free(&slot[i]);
    free(slot);
    free(flagrec);
    free(rindex);
    state_free(&src);
    state_free(&rcv);
    work_free(&w);
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
