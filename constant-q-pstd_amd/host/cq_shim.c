/*
 * cq_shim.c -- cq_shim_aps: the reference's aps acforward on the engine
 * (declared in include/cq_shim.h).  Per shot it does what
 * sgpstd-qmodeling/acps4.cu:330-426 does on the host (model windows, CPML
 * expansion, source and receiver indices), then runs every shot in one
 * batch and writes the reference's output files (:482-494, :610-640,
 * :690-699).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cq.h"
#include "cq_host.h"
#include "cq_shim.h"

#define SHIM_EFILE (-100)
#define SHIM_ESOURCE (-101)

typedef struct {
    float *vp, *rho, *bbx, *bbz;
    int *gxg, *gzg;
    int nr;
} shim_shot;

static float *shim_window(const float *full, int pml, float mlx, float dx, int nxt, int nzt, int nx, int nz, int nxx,
                          int nzz)
{
    float *w = (float *)calloc((size_t)nxx * nzz, sizeof(float));
    if (!w) return NULL;
    cqh_select_part_model(pml, mlx, dx, nxt, nzt, nx, nz, nxx, nzz, full, w);
    cqh_expand_pml(nx, nz, nxx, nzz, pml, w);
    return w;
}

static void shim_free(shim_shot *s)
{
    free(s->vp);
    free(s->rho);
    free(s->bbx);
    free(s->bbz);
    free(s->gxg);
    free(s->gzg);
}

int cq_shim_aps(int tmax, float dt, const float *rik, int nshots, const int *shotid, const float *sxp,
                const float *szp, int nx, int nz, int nxx, int nzz, int pml, float dx, float dz, float sroffmx,
                float disx, const float *vpt, const float *rhot, const float *bbxt, const float *bbzt, int nxt,
                int nzt, const int *tracenum, int flagout, const float *i_ax, const float *i_bx,
                const float *i_az, const float *i_bz, const float *h_ax, const float *h_bx, const float *h_az,
                const float *h_bz, const float *kxsmr, const float *kxscr, const float *kzsmr,
                const float *kzscr, const float *kxsmi, const float *kxsci, const float *kzsmi,
                const float *kzsci, const char *fname4, const char *fname5, const char *outdir)
{
    if (nshots < 1) return 0;
    int rc = 0;
    char name[1024];
    shim_shot *sh = (shim_shot *)calloc((size_t)nshots, sizeof(shim_shot));
    cq_job_desc *jd = (cq_job_desc *)calloc((size_t)nshots, sizeof(cq_job_desc));
    float *gxp = NULL, *gzp = NULL, *buf = NULL;
    cq_ctx ctx = NULL;
    cq_batch bt = NULL;
    if (!sh || !jd) {
        rc = CQ_ENOMEM;
        goto out;
    }
    for (int k = 0; k < nshots; k++) {
        const int sid = shotid[k], nr = tracenum[sid];
        shim_shot *s = &sh[k];
        const float xs = sxp[k], zs = szp[k], mlx = xs - sroffmx - disx;
        s->nr = nr;
        s->vp = shim_window(vpt, pml, mlx, dx, nxt, nzt, nx, nz, nxx, nzz);
        s->rho = shim_window(rhot, pml, mlx, dx, nxt, nzt, nx, nz, nxx, nzz);
        s->bbx = shim_window(bbxt, pml, mlx, dx, nxt, nzt, nx, nz, nxx, nzz);
        s->bbz = shim_window(bbzt, pml, mlx, dx, nxt, nzt, nx, nz, nxx, nzz);
        s->gxg = (int *)malloc(sizeof(int) * (nr > 0 ? nr : 1));
        s->gzg = (int *)malloc(sizeof(int) * (nr > 0 ? nr : 1));
        free(gxp);
        free(gzp);
        gxp = (float *)malloc(sizeof(float) * (nr > 0 ? nr : 1));
        gzp = (float *)malloc(sizeof(float) * (nr > 0 ? nr : 1));
        if (!s->vp || !s->rho || !s->bbx || !s->bbz || !s->gxg || !s->gzg || !gxp || !gzp) {
            rc = CQ_ENOMEM;
            goto out;
        }
        if (cqh_read_geophones(fname4, sid, nr, gxp) || cqh_read_geophones(fname5, sid, nr, gzp)) {
            fprintf(stderr, "cq_shim_aps: cannot read the receiver positions of shot %d\n", sid);
            rc = SHIM_EFILE;
            goto out;
        }
        cqh_shot_geom g;
        if (cqh_shot_indices(pml, nxx, nzz, dx, dz, sroffmx, disx, xs, zs, nr, gxp, gzp, s->gxg, s->gzg, &g)) {
            fprintf(stderr, "%d shot grid index exceeds model!\n", sid);
            rc = SHIM_ESOURCE;
            goto out;
        }
        cq_job_desc *j = &jd[k];
        j->kind = CQ_KIND_ACOUSTIC;
        j->nxx = nxx; j->nzz = nzz; j->pml = pml; j->nt = tmax; j->dt = dt;
        j->vp = s->vp; j->rho = s->rho; j->bbx = s->bbx; j->bbz = s->bbz;
        j->iax = i_ax; j->ibx = i_bx; j->hax = h_ax; j->hbx = h_bx;
        j->iaz = i_az; j->ibz = i_bz; j->haz = h_az; j->hbz = h_bz;
        j->kxsmr = kxsmr; j->kxsmi = kxsmi; j->kxscr = kxscr; j->kxsci = kxsci;
        j->kzsmr = kzsmr; j->kzsmi = kzsmi; j->kzscr = kzscr; j->kzsci = kzsci;
        j->dx = dx; j->dz = dz; j->rik = rik;
        j->xsg = g.xsg; j->zsg = g.zsg;
        j->nr = nr; j->gxg = s->gxg; j->gzg = s->gzg;
        printf("shotid=%d nxx=%d nzz=%d xs=%f zs=%f\n", sid, nxx, nzz, xs, zs);
    }
    if ((rc = cq_open(0, &ctx)) != 0) goto out;
    if ((rc = cq_batch_create(ctx, jd, nshots, &bt)) != 0) goto out;
    buf = (float *)malloc(sizeof(float) * (size_t)nxx * nzz);
    if (!buf) {
        rc = CQ_ENOMEM;
        goto out;
    }
    if (flagout == 1) {
        /* the interior of p after steps t = 0, 10, 20, ..., appended x-major
         * to <outdir>/snapa<shot>_<nz>_<nx>.bin, each file truncated first */
        for (int k = 0; k < nshots; k++) {
            snprintf(name, sizeof name, "%s/snapa%d_%d_%d.bin", outdir, shotid[k], nz, nx);
            FILE *fs = fopen(name, "wb");
            if (!fs) {
                rc = SHIM_EFILE;
                goto out;
            }
            fclose(fs);
        }
        for (int t = 0; t < tmax; t++) {
            if ((rc = cq_batch_step(bt, 1)) != 0) goto out;
            if (t % 10) continue;
            for (int k = 0; k < nshots; k++) {
                if ((rc = cq_batch_read_p(bt, k, buf)) != 0) goto out;
                snprintf(name, sizeof name, "%s/snapa%d_%d_%d.bin", outdir, shotid[k], nz, nx);
                FILE *fs = fopen(name, "ab");
                if (!fs) {
                    rc = SHIM_EFILE;
                    goto out;
                }
                for (int ix = pml; ix < nxx - pml; ix++)
                    if (fwrite(buf + (size_t)ix * nzz + pml, sizeof(float), (size_t)nz, fs) != (size_t)nz) rc = SHIM_EFILE;
                fclose(fs);
                if (rc) goto out;
            }
        }
    }
    if ((rc = cq_batch_run(bt)) != 0) goto out;
    for (int k = 0; k < nshots; k++) {
        const int nr = sh[k].nr;
        float *rec = (float *)malloc(sizeof(float) * (size_t)tmax * (nr > 0 ? nr : 1));
        if (!rec) {
            rc = CQ_ENOMEM;
            goto out;
        }
        rc = cq_batch_read_rec(bt, k, rec);
        if (rc == 0) {
            snprintf(name, sizeof name, "%s/reca%d_%d.bin", outdir, shotid[k], nr);
            if (cqh_write_f32(name, rec, (long)tmax * nr)) rc = SHIM_EFILE;
        }
        free(rec);
        if (rc) goto out;
    }
out:
    if (bt) cq_batch_destroy(bt);
    if (ctx) cq_close(ctx);
    if (sh)
        for (int k = 0; k < nshots; k++) shim_free(&sh[k]);
    free(sh);
    free(jd);
    free(gxp);
    free(gzp);
    free(buf);
    return rc;
}
