/*
 * cq_shim.c -- the bodies of the reference's four acforward entry points on
 * the engine (declared in include/cq_shim.h).  Per shot they do what the
 * reference does on the host (sgpstd-qmodeling/acps4.cu:330-426 and
 * acsglw.cu:330-430: model windows, CPML expansion, source and receiver
 * indices; sgpstd-qrtm/acps4.cu:330-440: the observed data and removedirect),
 * then run every shot of the call in one batch and write the reference's
 * outputs: gathers and snapshots (sgpstd-qmodeling/acps4.cu:482-494,
 * 610-640, 690-699) or the filtered, muted and projected image and
 * illumination summed into img_sp / ssg_sp (sgpstd-qrtm/acps4.cu:909-945).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cq.h"
#include "cq_host.h"
#include "cq_shim.h"

#define SHIM_EFILE (-100)
#define SHIM_ESOURCE (-101)

/* the inputs every entry point shares, in acforward's terms */
typedef struct {
    int kind;   /* cq_kind */
    int tmax;
    float dt;
    const float *rik;
    int nshots;
    const int *shotid;
    const float *sxp, *szp;
    int nx, nz, nxx, nzz, pml;
    float dx, dz, sroffmx, disx;
    const float *vpt, *rhot, *bbxt, *bbzt;
    const float *vwdt, *taot, *lamt, *gamt;   /* viscoacoustic kinds */
    int nxt, nzt;
    const int *tracenum;
    int flagout;
    const float *c[16];   /* i_ax i_bx i_az i_bz h_ax h_bx h_az h_bz kxsmr kxscr kzsmr kzscr kxsmi kxsci kzsmi kzsci */
    const char *fname4, *fname5;
    /* RTM kinds */
    int snaps;
    float *img_sp, *ssg_sp;
    float fobs, v0, etas, etar;
    const char *recname;
    const char *outdir;
} shim_args;

typedef struct {
    float *f[8];   /* vp rho bbx bbz, then gam lam vwd tao */
    float *obs;
    int *gxg, *gzg;
    cqh_shot_geom geo;
} shim_shot;

static float *shim_window(const shim_args *a, const float *full, float mlx)
{
    float *w = (float *)calloc((size_t)a->nxx * a->nzz, sizeof(float));
    if (!w) return NULL;
    cqh_select_part_model(a->pml, mlx, a->dx, a->nxt, a->nzt, a->nx, a->nz, a->nxx, a->nzz, full, w);
    cqh_expand_pml(a->nx, a->nz, a->nxx, a->nzz, a->pml, w);
    return w;
}

static void shim_free(shim_shot *s)
{
    for (int i = 0; i < 8; i++) free(s->f[i]);
    free(s->obs);
    free(s->gxg);
    free(s->gzg);
}

/* <outdir>/<stem><shot>_<nz>_<nx>.bin: the interior of p after steps t = 0,
 * 10, 20, ..., x-major, each file truncated first */
static int shim_snapshots(const shim_args *a, cq_batch bt, const char *stem)
{
    char name[1024];
    float *buf = (float *)malloc(sizeof(float) * (size_t)a->nxx * a->nzz);
    if (!buf) return CQ_ENOMEM;
    int rc = 0;
    for (int k = 0; k < a->nshots && !rc; k++) {
        snprintf(name, sizeof name, "%s/%s%d_%d_%d.bin", a->outdir, stem, a->shotid[k], a->nz, a->nx);
        FILE *fs = fopen(name, "wb");
        if (!fs) rc = SHIM_EFILE;
        else fclose(fs);
    }
    for (int t = 0; t < a->tmax && !rc; t++) {
        if ((rc = cq_batch_step(bt, 1)) != 0) break;
        if (t % 10) continue;
        for (int k = 0; k < a->nshots && !rc; k++) {
            if ((rc = cq_batch_read_p(bt, k, buf)) != 0) break;
            snprintf(name, sizeof name, "%s/%s%d_%d_%d.bin", a->outdir, stem, a->shotid[k], a->nz, a->nx);
            FILE *fs = fopen(name, "ab");
            if (!fs) {
                rc = SHIM_EFILE;
                break;
            }
            for (int ix = a->pml; ix < a->nxx - a->pml; ix++)
                if (fwrite(buf + (size_t)ix * a->nzz + a->pml, sizeof(float), (size_t)a->nz, fs) != (size_t)a->nz)
                    rc = SHIM_EFILE;
            fclose(fs);
        }
    }
    free(buf);
    return rc;
}

static int shim_run(const shim_args *a)
{
    if (a->nshots < 1) return 0;
    const int q = a->kind == CQ_KIND_VISCO || a->kind == CQ_KIND_QRTM;
    const int rtm = a->kind == CQ_KIND_RTM || a->kind == CQ_KIND_QRTM;
    int rc = 0;
    char name[1024];
    shim_shot *sh = (shim_shot *)calloc((size_t)a->nshots, sizeof(shim_shot));
    cq_job_desc *jd = (cq_job_desc *)calloc((size_t)a->nshots, sizeof(cq_job_desc));
    float *gxp = NULL, *gzp = NULL;
    cq_ctx ctx = NULL;
    cq_batch bt = NULL;
    if (!sh || !jd) {
        rc = CQ_ENOMEM;
        goto out;
    }
    for (int k = 0; k < a->nshots; k++) {
        const int sid = a->shotid[k], nr = a->tracenum[sid];
        shim_shot *s = &sh[k];
        const float xs = a->sxp[k], zs = a->szp[k], mlx = xs - a->sroffmx - a->disx;
        const float *full[8] = {a->vpt, a->rhot, a->bbxt, a->bbzt, a->gamt, a->lamt, a->vwdt, a->taot};
        for (int i = 0; i < (q ? 8 : 4); i++)
            if (!(s->f[i] = shim_window(a, full[i], mlx))) {
                rc = CQ_ENOMEM;
                goto out;
            }
        const size_t nrb = (size_t)(nr > 0 ? nr : 1);
        s->gxg = (int *)malloc(sizeof(int) * nrb);
        s->gzg = (int *)malloc(sizeof(int) * nrb);
        free(gxp);
        free(gzp);
        gxp = (float *)malloc(sizeof(float) * nrb);
        gzp = (float *)malloc(sizeof(float) * nrb);
        if (!s->gxg || !s->gzg || !gxp || !gzp) {
            rc = CQ_ENOMEM;
            goto out;
        }
        if (cqh_read_geophones(a->fname4, sid, nr, gxp) || cqh_read_geophones(a->fname5, sid, nr, gzp)) {
            fprintf(stderr, "cq_shim: cannot read the receiver positions of shot %d\n", sid);
            rc = SHIM_EFILE;
            goto out;
        }
        if (cqh_shot_indices(a->pml, a->nxx, a->nzz, a->dx, a->dz, a->sroffmx, a->disx, xs, zs, nr, gxp, gzp, s->gxg,
                             s->gzg, &s->geo)) {
            fprintf(stderr, "%d shot grid index exceeds model!\n", sid);
            rc = SHIM_ESOURCE;
            goto out;
        }
        if (rtm) {
            /* observed data <recname><shot>_<nr>.bin with the direct wave removed */
            s->obs = (float *)malloc(sizeof(float) * (size_t)a->tmax * nrb);
            snprintf(name, sizeof name, "%s%d_%d.bin", a->recname, sid, nr);
            if (!s->obs || cqh_read_f32(name, s->obs, (long)a->tmax * nr)) {
                fprintf(stderr, "cq_shim: cannot read the observed data %s\n", name);
                rc = SHIM_EFILE;
                goto out;
            }
            cqh_removedirect(a->tmax, nr, a->dt, a->fobs, xs, zs, gxp, gzp, a->v0, s->obs);
        }
        cq_job_desc *j = &jd[k];
        j->kind = a->kind;
        j->nxx = a->nxx; j->nzz = a->nzz; j->pml = a->pml; j->nt = a->tmax; j->dt = a->dt;
        j->vp = s->f[0]; j->rho = s->f[1]; j->bbx = s->f[2]; j->bbz = s->f[3];
        if (q) {
            j->gam = s->f[4]; j->lam = s->f[5]; j->vwd = s->f[6]; j->tao = s->f[7];
        }
        j->iax = a->c[0]; j->ibx = a->c[1]; j->iaz = a->c[2]; j->ibz = a->c[3];
        j->hax = a->c[4]; j->hbx = a->c[5]; j->haz = a->c[6]; j->hbz = a->c[7];
        j->kxsmr = a->c[8]; j->kxscr = a->c[9]; j->kzsmr = a->c[10]; j->kzscr = a->c[11];
        j->kxsmi = a->c[12]; j->kxsci = a->c[13]; j->kzsmi = a->c[14]; j->kzsci = a->c[15];
        j->dx = a->dx; j->dz = a->dz; j->rik = a->rik;
        j->xsg = s->geo.xsg; j->zsg = s->geo.zsg;
        j->nr = nr; j->gxg = s->gxg; j->gzg = s->gzg;
        j->obs = s->obs; j->snaps = a->snaps; j->etas = a->etas; j->etar = a->etar;
        printf("shotid=%d nxx=%d nzz=%d xs=%f zs=%f\n", sid, a->nxx, a->nzz, xs, zs);
    }
    if ((rc = cq_open(0, &ctx)) != 0) goto out;
    if ((rc = cq_batch_create(ctx, jd, a->nshots, &bt)) != 0) goto out;
    if (!rtm && a->flagout == 1 && (rc = shim_snapshots(a, bt, q ? "snapva" : "snapa")) != 0) goto out;
    if ((rc = cq_batch_run(bt)) != 0) goto out;
    for (int k = 0; k < a->nshots && !rc; k++) {
        const shim_shot *s = &sh[k];
        const int nr = a->tracenum[a->shotid[k]];
        if (!rtm) {
            float *rec = (float *)malloc(sizeof(float) * (size_t)a->tmax * (nr > 0 ? nr : 1));
            if (!rec) {
                rc = CQ_ENOMEM;
                break;
            }
            rc = cq_batch_read_rec(bt, k, rec);
            if (rc == 0) {
                snprintf(name, sizeof name, "%s/%s%d_%d.bin", a->outdir, q ? "recva" : "reca", a->shotid[k], nr);
                if (cqh_write_f32(name, rec, (long)a->tmax * nr)) rc = SHIM_EFILE;
            }
            free(rec);
        } else {
            const size_t ni = (size_t)a->nx * a->nz, nm = (size_t)a->nxt * a->nzt;
            float *img = (float *)malloc(sizeof(float) * ni), *imgf = (float *)malloc(sizeof(float) * ni);
            float *ssg = (float *)malloc(sizeof(float) * ni), *pj = (float *)calloc(nm, sizeof(float));
            if (!img || !imgf || !ssg || !pj) rc = CQ_ENOMEM;
            if (rc == 0) rc = cq_batch_read_image(bt, k, img, ssg);
            if (rc == 0) {
                cqh_filter_image(a->nx, a->nz, a->dx, a->dz, img, imgf);
                cqh_mute_image(a->nx, a->nz, a->pml, s->geo.lgest, s->geo.rgest, imgf);
                cqh_projectomodel(a->pml, s->geo.mlx, a->dx, a->nxt, a->nzt, a->nx, a->nz, a->nxx, a->nzz, pj, imgf);
                for (size_t i = 0; i < nm; i++) a->img_sp[i] += pj[i];
                memset(pj, 0, sizeof(float) * nm);
                cqh_projectomodel(a->pml, s->geo.mlx, a->dx, a->nxt, a->nzt, a->nx, a->nz, a->nxx, a->nzz, pj, ssg);
                for (size_t i = 0; i < nm; i++) a->ssg_sp[i] += pj[i];
            }
            free(img);
            free(imgf);
            free(ssg);
            free(pj);
        }
    }
out:
    if (bt) cq_batch_destroy(bt);
    if (ctx) cq_close(ctx);
    if (sh)
        for (int k = 0; k < a->nshots; k++) shim_free(&sh[k]);
    free(sh);
    free(jd);
    free(gxp);
    free(gzp);
    return rc;
}

#define SHIM_COMMON(KIND)                                                                                     \
    shim_args a;                                                                                              \
    memset(&a, 0, sizeof a);                                                                                  \
    a.kind = KIND; a.tmax = tmax; a.dt = dt; a.rik = rik; a.nshots = nshots; a.shotid = shotid;             \
    a.sxp = sxp; a.szp = szp; a.nx = nx; a.nz = nz; a.nxx = nxx; a.nzz = nzz; a.pml = pml; a.dx = dx;        \
    a.dz = dz; a.sroffmx = sroffmx; a.disx = disx; a.vpt = vpt; a.rhot = rhot; a.bbxt = bbxt;                \
    a.bbzt = bbzt; a.nxt = nxt; a.nzt = nzt; a.tracenum = tracenum; a.fname4 = fname4; a.fname5 = fname5;  \
    {                                                                                                         \
        const float *c_[16] = {i_ax, i_bx, i_az, i_bz, h_ax, h_bx, h_az, h_bz,                              \
                               kxsmr, kxscr, kzsmr, kzscr, kxsmi, kxsci, kzsmi, kzsci};                     \
        memcpy(a.c, c_, sizeof c_);                                                                           \
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
    SHIM_COMMON(CQ_KIND_ACOUSTIC)
    a.flagout = flagout;
    a.outdir = outdir;
    return shim_run(&a);
}

int cq_shim_vaps(int tmax, float dt, const float *rik, int nshots, const int *shotid, const float *sxp,
                 const float *szp, int nx, int nz, int nxx, int nzz, int pml, float dx, float dz, float sroffmx,
                 float disx, const float *vpt, const float *vwdt, const float *taot, const float *lamt,
                 const float *gamt, const float *rhot, const float *bbxt, const float *bbzt, int nxt, int nzt,
                 const int *tracenum, int flagout, const float *i_ax, const float *i_bx, const float *i_az,
                 const float *i_bz, const float *h_ax, const float *h_bx, const float *h_az, const float *h_bz,
                 const float *kxsmr, const float *kxscr, const float *kzsmr, const float *kzscr,
                 const float *kxsmi, const float *kxsci, const float *kzsmi, const float *kzsci,
                 const char *fname4, const char *fname5, const char *outdir)
{
    SHIM_COMMON(CQ_KIND_VISCO)
    a.vwdt = vwdt; a.taot = taot; a.lamt = lamt; a.gamt = gamt;
    a.flagout = flagout;
    a.outdir = outdir;
    return shim_run(&a);
}

int cq_shim_artm(int snapnum, int tmax, float dt, const float *rik, int nshots, const int *shotid,
                 const float *sxp, const float *szp, int nx, int nz, int nxx, int nzz, int pml, float dx,
                 float dz, float sroffmx, float disx, const float *vpt, const float *rhot, const float *bbxt,
                 const float *bbzt, int nxt, int nzt, const int *tracenum, const float *i_ax,
                 const float *i_bx, const float *i_az, const float *i_bz, const float *h_ax, const float *h_bx,
                 const float *h_az, const float *h_bz, const float *kxsmr, const float *kxscr,
                 const float *kzsmr, const float *kzscr, const float *kxsmi, const float *kxsci,
                 const float *kzsmi, const float *kzsci, float *img_sp, float *ssg_sp, float fobs, float v0,
                 const char *fname4, const char *fname5, const char *recname)
{
    SHIM_COMMON(CQ_KIND_RTM)
    a.snaps = snapnum; a.img_sp = img_sp; a.ssg_sp = ssg_sp; a.fobs = fobs; a.v0 = v0; a.recname = recname;
    return shim_run(&a);
}

int cq_shim_qrtm(int snapnum, int tmax, float dt, const float *rik, int nshots, const int *shotid,
                 const float *sxp, const float *szp, int nx, int nz, int nxx, int nzz, int pml, float dx,
                 float dz, float sroffmx, float disx, const float *vpt, const float *vwdt, const float *taot,
                 const float *lamt, const float *gamt, const float *rhot, const float *bbxt, const float *bbzt,
                 int nxt, int nzt, const int *tracenum, const float *i_ax, const float *i_bx,
                 const float *i_az, const float *i_bz, const float *h_ax, const float *h_bx, const float *h_az,
                 const float *h_bz, const float *kxsmr, const float *kxscr, const float *kzsmr,
                 const float *kzscr, const float *kxsmi, const float *kxsci, const float *kzsmi,
                 const float *kzsci, float *img_sp, float *ssg_sp, float fobs, float v0, float etas,
                 float etar, const char *fname4, const char *fname5, const char *recname)
{
    SHIM_COMMON(CQ_KIND_QRTM)
    a.vwdt = vwdt; a.taot = taot; a.lamt = lamt; a.gamt = gamt;
    a.snaps = snapnum; a.img_sp = img_sp; a.ssg_sp = ssg_sp; a.fobs = fobs; a.v0 = v0;
    a.etas = etas; a.etar = etar; a.recname = recname;
    return shim_run(&a);
}
