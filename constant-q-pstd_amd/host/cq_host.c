/*
 * cq_host.c -- parameter files, binary formats, coefficient builders and shot
 * windowing for the constant-Q PSTD engine (libcq_host.so).
 *
 * The builders keep the reference's arithmetic exactly (float storage, double
 * libm, the constant pi = 3.1415926 of sgpstd-qmodeling/funca.c:6) because the
 * engine's results are checked against the reference's kernels run on these
 * very arrays; each function cites the reference routine it replaces.  Unlike
 * the reference, nothing here calls exit(): errors come back as return codes.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cq_host.h"

#define PI CQH_PI

/* ------------------------------------------------------------------------ */
/* parameter files                                                           */
/* ------------------------------------------------------------------------ */

/* One label line followed by one value (the reference reads the label with
 * fgets and the value with fscanf, funca.c:21-111). */
static int rd_label(FILE *fp)
{
    char buf[256];
    if (fscanf(fp, " ") < 0) return -1;
    if (!fgets(buf, sizeof buf, fp)) return -1;
    return 0;
}
static int rd_int(FILE *fp, int *v) { return (rd_label(fp) || fscanf(fp, " %d", v) != 1) ? -1 : 0; }
static int rd_flt(FILE *fp, float *v) { return (rd_label(fp) || fscanf(fp, " %f", v) != 1) ? -1 : 0; }
static int rd_str(FILE *fp, char *s, int cap)
{
    char buf[256];
    if (rd_label(fp) || fscanf(fp, " %255s", buf) != 1) return -1;
    if ((int)strlen(buf) >= cap) return -2;
    strcpy(s, buf);
    return 0;
}

int cqh_read_params(const char *path, int fmt, cqh_params *o)
{
    FILE *fp = fopen(path, "r");
    int e = 0;
    if (!fp) return -1;
    memset(o, 0, sizeof(*o));
    o->pml = 10;           /* acousticsgps4.c:4 #define pml 10 */
    o->snapnum = 20;       /* acousticsgps4.c:131 snapnum = 20 (modeling) */
    o->fo = 0.f;
    const int q = (fmt == CQH_FMT_VAPS || fmt == CQH_FMT_QRTM);
    const int rtm = (fmt == CQH_FMT_ARTM || fmt == CQH_FMT_QRTM);
    e |= rd_int(fp, &o->shotnum);
    e |= rd_int(fp, &o->tmax);
    e |= rd_flt(fp, &o->dt_ms);
    e |= rd_flt(fp, &o->fd);
    if (q) e |= rd_flt(fp, &o->fo);
    e |= rd_int(fp, &o->nx);
    e |= rd_int(fp, &o->nz);
    e |= rd_flt(fp, &o->dx);
    e |= rd_flt(fp, &o->dz);
    e |= rd_flt(fp, &o->disx);
    if (rtm) e |= rd_flt(fp, &o->v0);
    if (fmt == CQH_FMT_QRTM) {
        e |= rd_flt(fp, &o->etas);
        e |= rd_flt(fp, &o->etar);
    }
    e |= rd_str(fp, o->fname1, 40);
    e |= rd_str(fp, o->fname2, 40);
    e |= rd_str(fp, o->fname3, 40);
    e |= rd_str(fp, o->fname4, 40);
    e |= rd_str(fp, o->fname5, 40);
    e |= rd_str(fp, o->vpname, 80);
    e |= rd_str(fp, o->rhoname, 80);
    if (q) e |= rd_str(fp, o->qname, 80);
    if (rtm) e |= rd_str(fp, o->recname, 60);
    else e |= rd_int(fp, &o->flagout);
    e |= rd_int(fp, &o->flagsource);
    if (rtm) e |= rd_int(fp, &o->snapnum);
    if (e) { fclose(fp); return -2; }
    /* optional trailing "label / value" pairs: "pml" and "low-rank" */
    {
        char lab[256];
        while (fscanf(fp, " ") >= 0 && fgets(lab, sizeof lab, fp)) {
            int v;
            if (fscanf(fp, " %d", &v) != 1) break;
            if (strstr(lab, "low-rank") || strstr(lab, "lowrank")) o->lr_rank = v;
            else if (strstr(lab, "pml") || strstr(lab, "PML")) o->pml = v;
        }
    }
    fclose(fp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* binary files                                                              */
/* ------------------------------------------------------------------------ */

static int rd_bin(const char *path, void *dst, size_t sz, long count)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    size_t got = fread(dst, sz, (size_t)count, fp);
    fclose(fp);
    return got == (size_t)count ? 0 : -2;
}

int cqh_read_f32(const char *path, float *dst, long count) { return rd_bin(path, dst, 4, count); }
int cqh_read_i32(const char *path, int *dst, long count) { return rd_bin(path, dst, 4, count); }

int cqh_write_f32(const char *path, const float *src, long count)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    size_t put = fwrite(src, 4, (size_t)count, fp);
    fclose(fp);
    return put == (size_t)count ? 0 : -2;
}

int cqh_read_geophones(const char *prefix, int shotid, int nr, float *dst)
{
    char name[512];
    snprintf(name, sizeof name, "%s%d.bin", prefix, shotid);
    return cqh_read_f32(name, dst, nr);
}

/* ------------------------------------------------------------------------ */
/* coefficient builders                                                      */
/* ------------------------------------------------------------------------ */

void cqh_get_max_min(int n, const float *v, float *vmax, float *vmin)
{
    float mn = 10000000.0f, mx = -10000000.0f;
    for (int i = 0; i < n; i++) {
        if (v[i] - mn < 1.0e-4) mn = v[i];
        if (v[i] - mx > 1.0e-4) mx = v[i];
    }
    *vmax = mx;
    *vmin = mn;
}

float cqh_check_stability(float vmax, float dt, float dx, float dz)
{
    float d = PI * vmax * dt / 2.0 * sqrt(1.0 / dx / dx + 1.0 / dz / dz);
    return d;
}

void cqh_bbx_bbz(int nxx, int nzz, const float *dens, float *bbx, float *bbz)
{
    for (int ix = 0; ix < nxx - 1; ix++)
        for (int iz = 0; iz < nzz; iz++) {
            long a = (long)ix * nzz + iz;
            bbx[a] = 2.0 / (dens[a] + dens[a + nzz]);
        }
    for (int iz = 0; iz < nzz; iz++) {
        long a = (long)(nxx - 1) * nzz + iz;
        bbx[a] = 1.0 / dens[a];
    }
    for (int ix = 0; ix < nxx; ix++) {
        for (int iz = 0; iz < nzz - 1; iz++) {
            long a = (long)ix * nzz + iz;
            bbz[a] = 2.0 / (dens[a] + dens[a + 1]);
        }
        long a = (long)ix * nzz + nzz - 1;
        bbz[a] = 1.0 / dens[a];
    }
}

void cqh_ricker(int tmax, float dt, float fd, float *rik, int flag)
{
    float t0 = 1.0 / fd;
    for (int i = 0; i < tmax; i++) {
        float u;
        if (flag == 2) {
            u = pow(PI * fd * (i * dt - t0), 2.0);
            rik[i] = (1.0 - 2.0 * u) * exp(-u);
        } else {
            u = pow(2.0 * PI * fd * (i * dt - t0), 2.0);
            rik[i] = (i * dt - t0) * exp(-u);
        }
    }
}

/* One axis of the CPML profile (integer and half grid), funca.c:223-340:
 * d(i) = dmax ((pml-1-i)/pml)^3, alpha(i) = pi fm i/(pml-1),
 * b = exp(-dt (d+alpha)), a = d/(d+alpha) (b-1). */
static void pml_axis(int pml, float alpha_max, float dmax, int n, float dt,
                     float *ia, float *ib, float *ha, float *hb)
{
    const int n123 = 3, n3 = 1;
    float *idd = (float *)malloc(sizeof(float) * n);
    float *ial = (float *)malloc(sizeof(float) * n);
    float *hd = (float *)malloc(sizeof(float) * n);
    float *hal = (float *)malloc(sizeof(float) * n);
    for (int i = 0; i < pml; i++) {
        float w1 = pow(1.0 * (pml - 1 - i) / pml, n123);
        float w3 = pow(1.0 * i / (pml - 1), n3);
        idd[i] = dmax * w1;
        ial[i] = alpha_max * w3;
        idd[n - 1 - i] = idd[i];
        ial[n - 1 - i] = ial[i];
    }
    for (int i = pml; i < n - pml; i++) {
        idd[i] = 0.0f;
        ial[i] = ial[pml - 1];
    }
    for (int i = 0; i < pml - 1; i++) {
        float w2 = pow(1.0 * (pml - 1.5 - i) / pml, n123);
        float w4 = pow(1.0 * (i + 0.5) / (pml - 1), n3);
        hd[i] = dmax * w2;
        hd[n - 2 - i] = hd[i];
        hal[i] = alpha_max * w4;
        hal[n - 2 - i] = hal[i];
    }
    for (int i = pml - 1; i < n - pml; i++) {
        hd[i] = 0.0f;
        hal[i] = hal[pml - 2];
    }
    hd[n - 1] = 0.0f;
    hal[n - 1] = hal[n - 2];
    for (int i = 0; i < n; i++) {
        float s = idd[i] + ial[i];
        ib[i] = exp(-dt * s);
        ia[i] = idd[i] / s * (ib[i] - 1.0);
        s = hd[i] + hal[i];
        hb[i] = exp(-dt * s);
        ha[i] = hd[i] / s * (hb[i] - 1.0);
    }
    free(idd); free(ial); free(hd); free(hal);
}

void cqh_pml_coeffs(int pml, float fm,
                    float *i_ax, float *i_bx, float *i_az, float *i_bz,
                    float *h_ax, float *h_bx, float *h_az, float *h_bz,
                    int nxx, int nzz, float dx, float dz, float dt)
{
    const float Re = 1e-5f, vmax = 5000.0f;
    float alpha_max = PI * fm;
    float wx = pml * dx, wz = pml * dz;
    float dmaxx = (1 + 2 + 1) * vmax * log(1.0 / Re) / (2.0 * wx);
    float dmaxz = (1 + 2 + 1) * vmax * log(1.0 / Re) / (2.0 * wz);
    pml_axis(pml, alpha_max, dmaxx, nxx, dt, i_ax, i_bx, h_ax, h_bx);
    pml_axis(pml, alpha_max, dmaxz, nzz, dt, i_az, i_bz, h_az, h_bz);
}

/* |k| on the FFT grid, mirrored by kk[i] = kk[n-i] above n/2 (funca.c:363-385) */
static void kgrid(int n, float d, float *kk)
{
    float dk = 2 * PI / (n * d);
    int nh = n / 2 + 1;
    for (int i = 0; i < nh; i++) kk[i] = i * dk;
    for (int i = nh; i < n; i++) kk[i] = kk[n - i];
}

/* Half-cell phase exp(i pi k/n), with the reference's mirror index
 * n/2*2 - i for the negative frequencies (funca.c:387-396). */
static void half_shift(int n, float *sr, float *si)
{
    int nh = n / 2 + 1;
    for (int i = 0; i < nh; i++) {
        float a = i * PI / n;
        sr[i] = cos(a);
        si[i] = sin(a);
    }
    for (int i = nh; i < n; i++) {
        sr[i] = sr[n / 2 * 2 - i];
        si[i] = -si[n / 2 * 2 - i];
    }
}

void cqh_shift_coeffs(float *kxsmr, float *kxscr, float *kzsmr, float *kzscr,
                      float *kxsmi, float *kxsci, float *kzsmi, float *kzsci,
                      float dx, float dz, int nxx, int nzz)
{
    float *kk = (float *)malloc(sizeof(float) * (nxx > nzz ? nxx : nzz));
    float *sr = (float *)malloc(sizeof(float) * (nxx > nzz ? nxx : nzz));
    float *si = (float *)malloc(sizeof(float) * (nxx > nzz ? nxx : nzz));
    kgrid(nxx, dx, kk);
    half_shift(nxx, sr, si);
    for (int i = 0; i < nxx; i++) {
        kxsmr[i] = kk[i] * sr[i];
        kxsmi[i] = -kk[i] * si[i];
        kxscr[i] = kk[i] * sr[i];
        kxsci[i] = kk[i] * si[i];
    }
    kgrid(nzz, dz, kk);
    half_shift(nzz, sr, si);
    for (int j = 0; j < nzz; j++) {
        kzsmr[j] = kk[j] * sr[j];
        kzsmi[j] = kk[j] * si[j];
        kzscr[j] = kk[j] * sr[j];
        kzsci[j] = -kk[j] * si[j];
    }
    free(kk); free(sr); free(si);
}

void cqh_k2d(float dx, float dz, int nxx, int nzz,
             float *k2d1, float *k2d2, float *k2d3, float *k2da)
{
    float *kx = (float *)malloc(sizeof(float) * nxx);
    float *kz = (float *)malloc(sizeof(float) * nzz);
    kgrid(nxx, dx, kx);
    kgrid(nzz, dz, kz);
    for (int i = 0; i < nxx; i++)
        for (int j = 0; j < nzz; j++) {
            long a = (long)i * nzz + j;
            float s = kx[i] * kx[i] + kz[j] * kz[j];
            if (k2da) k2da[a] = pow(kx[i] * kx[i] + kz[j] * kz[j], 0.505);
            k2d1[a] = s;
            if (k2d2) k2d2[a] = pow(k2d1[a], 1.0 / 32.0);
            if (k2d3) k2d3[a] = pow(k2d1[a], 33.0 / 32.0);
        }
    free(kx); free(kz);
}

void cqh_transfermodel(int n, float fd, float fo, const float *vp, const float *qp,
                       float *gam, float *lam, float *vwd, float *tao)
{
    float wd = 2 * PI * fd;
    for (int i = 0; i < n; i++) {
        gam[i] = 1.0 / PI / qp[i];
        lam[i] = pow(fd / fo, 2.0 * gam[i]);
        vwd[i] = pow(vp[i] / wd, 1.0 / 16.0);
        tao[i] = lam[i] * vp[i] / qp[i];
    }
}

float cqh_get_max_tao(int n, const float *tao)
{
    float mx = 0.0f;
    for (int i = 0; i < n; i++)
        if (tao[i] >= mx) mx = tao[i];
    return mx;
}

/* ------------------------------------------------------------------------ */
/* windowing                                                                 */
/* ------------------------------------------------------------------------ */

void cqh_select_part_model(int pml, float mlx, float dx, int nxt, int nzt, int nx, int nz,
                           int nxx, int nzz, const float *full, float *win)
{
    (void)nz; (void)nxx;
    if (mlx >= 0.0) {
        /* window starts inside the model: first column truncated */
        int c0 = (int)(mlx / dx);
        int c1 = c0 + nx - 1;
        int last = c1 < nxt ? c1 : nxt - 1;
        for (int i = c0; i <= last; i++)
            for (int j = 0; j < nzt; j++)
                win[(long)(i - c0 + pml) * nzz + j + pml] = full[(long)i * nzt + j];
        /* past the right model edge: replicate the last model column
         * (interior depths only; expand_pml fills the rest) */
        for (int i = nxt; i <= c1; i++)
            for (int j = pml; j < nzz - pml; j++)
                win[(long)(i - c0 + pml) * nzz + j] = win[(long)(nxt - 1 - c0 + pml) * nzz + j];
    } else {
        /* window starts left of the model: offset rounded */
        int off = (int)(fabs(mlx) / dx + 0.5);
        int c1 = nx - off;
        int ncopy = c1 <= nxt ? c1 : nxt;
        for (int i = 0; i < ncopy; i++)
            for (int j = 0; j < nzt; j++)
                win[(long)(i + off + pml) * nzz + j + pml] = full[(long)i * nzt + j];
        if (c1 > nxt)
            for (int i = nxt; i < c1; i++)
                for (int j = pml; j < nzz - pml; j++)
                    win[(long)(i + off + pml) * nzz + j] = win[(long)(nxt - 1 + off + pml) * nzz + j];
        for (int i = 0; i < off; i++)
            for (int j = pml; j < nzz - pml; j++)
                win[(long)(i + pml) * nzz + j] = win[(long)(off + pml) * nzz + j];
    }
}

void cqh_expand_pml(int nx, int nz, int nxx, int nzz, int pml, float *win)
{
    (void)nx; (void)nz;
    for (int i = pml; i < nxx - pml; i++) {
        float top = win[(long)i * nzz + pml];
        float bot = win[(long)i * nzz + nzz - pml - 1];
        for (int j = 0; j < pml; j++) win[(long)i * nzz + j] = top;
        for (int j = nzz - pml; j < nzz; j++) win[(long)i * nzz + j] = bot;
    }
    for (int j = 0; j < nzz; j++) {
        float l = win[(long)pml * nzz + j];
        float r = win[(long)(nxx - pml - 1) * nzz + j];
        for (int i = 0; i < pml; i++) win[(long)i * nzz + j] = l;
        for (int i = nxx - pml; i < nxx; i++) win[(long)i * nzz + j] = r;
    }
}

void cqh_projectomodel(int pml, float mlx, float dx, int nxt, int nzt, int nx, int nz,
                       int nxx, int nzz, float *full, const float *win)
{
    (void)pml; (void)nxx; (void)nzz;
    if (mlx >= 0.0) {
        int c0 = (int)(mlx / dx);
        int c1 = c0 + nx - 1;
        int last = c1 < nxt ? c1 : nxt - 1;
        for (int i = c0; i <= last; i++)
            for (int j = 0; j < nzt; j++)
                full[(long)i * nzt + j] = win[(long)(i - c0) * nz + j];
    } else {
        int off = (int)(fabs(mlx) / dx + 0.5);
        int c1 = nx - off;
        int ncopy = c1 <= nxt ? c1 : nxt;
        for (int i = 0; i < ncopy; i++)
            for (int j = 0; j < nzt; j++)
                full[(long)i * nzt + j] = win[(long)(i + off) * nz + j];
    }
}

void cqh_removedirect(int tmax, int nr, float dt, float fd, float xs, float zs,
                      const float *gxp, const float *gzp, float v0, float *obs)
{
    for (int i = 0; i < nr; i++) {
        float dis = sqrt(pow((xs - gxp[i]), 2.0) + pow((zs - gzp[i]), 2.0));
        int ts = (int)(1.0 * dis / v0 / dt + 0.5);
        int te = ts + (int)(2.0 / fd / dt + 0.5);
        if (te >= tmax) te = tmax - 1;
        for (int j = ts; j <= te; j++) obs[(long)j * nr + i] = 0.0f;
    }
}

/* ------------------------------------------------------------------------ */
/* shot geometry                                                             */
/* ------------------------------------------------------------------------ */

float cqh_window_sroffmx(int s0, int s1, const int *tracenum, const float *shotxp,
                         const char *gxprefix, int *nx_out, float disx, float dx)
{
    float sroffmx = -1.0f;
    int nrmx = 0;
    for (int i = s0; i < s1; i++) if (tracenum[i] > nrmx) nrmx = tracenum[i];
    float *gx = (float *)malloc(sizeof(float) * (nrmx > 0 ? nrmx : 1));
    for (int i = s0; i < s1; i++) {
        float xr, xl;
        if (cqh_read_geophones(gxprefix, i, tracenum[i], gx)) { free(gx); return -1.0f; }
        cqh_get_max_min(tracenum[i], gx, &xr, &xl);
        float a = fabs(xr - shotxp[i]);
        float b = fabs(shotxp[i] - xl);
        float off = a > b ? a : b;
        if (off > sroffmx) sroffmx = off;
    }
    free(gx);
    float simxlen = 2.0 * (sroffmx + disx);
    if (nx_out) *nx_out = (int)(simxlen / dx + 1.5);
    return sroffmx;
}

int cqh_shot_indices(int pml, int nxx, int nzz, float dx, float dz,
                     float sroffmx, float disx, float xs, float zs,
                     int nr, const float *gxp, const float *gzp,
                     int *gxg, int *gzg, cqh_shot_geom *g)
{
    g->mlx = xs - sroffmx - disx;
    g->zsg = (int)(zs / dz + 0.5) + pml;
    g->xsg = (int)((sroffmx + disx) / dx + 0.5) + pml;
    g->nr = nr;
    g->lgest = 10000;
    g->rgest = -1;
    g->nclamped = 0;
    if (g->xsg < 0 || g->xsg >= nxx || g->zsg < 0 || g->zsg >= nzz) return -1;
    for (int i = 0; i < nr; i++) {
        gzg[i] = (int)(gzp[i] / dz + 0.5) + pml;
        if (gxp[i] >= xs) gxg[i] = g->xsg + (int)((gxp[i] - xs) / dx + 0.5);
        else gxg[i] = g->xsg - (int)((xs - gxp[i]) / dx + 0.5);
        if (gxg[i] < pml) { gxg[i] = pml; g->nclamped++; }
        if (gxg[i] >= nxx - pml) { gxg[i] = nxx - pml - 1; g->nclamped++; }
        if (gxg[i] > g->rgest) g->rgest = gxg[i];
        if (gxg[i] < g->lgest) g->lgest = gxg[i];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* RTM image post-processing                                                 */
/* ------------------------------------------------------------------------ */

void cqh_filter_image(int nx, int nz, float dx, float dz, const float *img, float *imgf)
{
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            long a = (long)ix * nz + iz;
            if (ix == 0 || iz == 0 || ix == nx - 1 || iz == nz - 1) {
                imgf[a] = 0.0f;
                continue;
            }
            imgf[a] = (img[a + 1] + img[a - 1] - 2.0 * img[a]) / dz / dz +
                      (img[a + nz] + img[a - nz] - 2.0 * img[a]) / dx / dx;
        }
}

void cqh_mute_image(int nx, int nz, int pml, int lgest, int rgest, float *img)
{
    int l = lgest - pml, r = rgest - pml;
    for (int iz = 0; iz < nz; iz++) {
        for (int ix = 0; ix < l && ix < nx; ix++) img[(long)ix * nz + iz] = 0.0f;
        for (int ix = r < 0 ? 0 : r; ix < nx; ix++) img[(long)ix * nz + iz] = 0.0f;
    }
}
