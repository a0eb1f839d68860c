/*
 * cq_driver.c -- the aps / vaps / artm / qrtm drivers on the MI355X engine.
 *
 * Replaces the reference's MPI drivers sgpstd-qmodeling/acousticsgps4.c (aps),
 * sgpstd-qmodeling/visacousticsgps4.c (vaps), sgpstd-qrtm/main.c (artm) and
 * sgpstd-qrtm/acrtmain.c (qrtm): the same parameter files, geometry and model
 * files and the same outputs (reca<shot>_<nr>.bin / recva<shot>_<nr>.bin gathers,
 * RTM.BIN / QRTM.BIN = [image | illumination]).  The MPI shot loop becomes
 * on-GPU batching: the shots of a chunk of `batch` shots advance in the same
 * launches (RTM shots each on their own revolve schedule).
 *
 * usage: <aps|vaps|artm|qrtm> [parameter file] [gather directory] [batch]
 *   parameter file   default ../input/<reference name>.txt
 *   gather directory default ../obsdata (modeling output, as acps4.cu:694)
 *   batch            shots per batch, default 64 (halved while a batch does
 *                    not fit the device)
 * The mode is taken from the program name (argv[0]) or CQ_MODE.
 */
#include <libgen.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "cq.h"
#include "cq_host.h"
#include "cq_lowrank.h"

#define DIE(...)                                   \
    do {                                           \
        fprintf(stderr, __VA_ARGS__);              \
        fputc('\n', stderr);                       \
        return 1;                                  \
    } while (0)

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

typedef struct {
    float *vp, *rho, *bbx, *bbz, *gam, *lam, *vwd, *tao;
    float *lrw, *lrs;          /* low-rank weights and symbols (vaps / qrtm with "low-rank") */
    int *gxg, *gzg;
    float *obs;
    cqh_shot_geom geo;
    int shotid, nr;
} shot_buf;

static float *win_field(const float *full, int nxt, int nzt, int nx, int nz, int pml, float mlx, float dx)
{
    const int nxx = nx + 2 * pml, nzz = nz + 2 * pml;
    float *w = (float *)calloc((size_t)nxx * nzz, sizeof(float));
    cqh_select_part_model(pml, mlx, dx, nxt, nzt, nx, nz, nxx, nzz, full, w);
    cqh_expand_pml(nx, nz, nxx, nzz, pml, w);
    return w;
}

static void free_shot(shot_buf *s)
{
    free(s->vp); free(s->rho); free(s->bbx); free(s->bbz);
    free(s->gam); free(s->lam); free(s->vwd); free(s->tao);
    free(s->gxg); free(s->gzg); free(s->obs);
    free(s->lrw); free(s->lrs);
    memset(s, 0, sizeof(*s));
}

int main(int argc, char **argv)
{
    const char *mode = getenv("CQ_MODE");
    if (!mode) mode = basename(argv[0]);
    int fmt;
    const char *defpar;
    if (!strcmp(mode, "aps")) { fmt = CQH_FMT_APS; defpar = "../input/ac2drealmodeling.txt"; }
    else if (!strcmp(mode, "vaps")) { fmt = CQH_FMT_VAPS; defpar = "../input/visac2drealmodeling.txt"; }
    else if (!strcmp(mode, "artm")) { fmt = CQH_FMT_ARTM; defpar = "../input/ac2drtm.txt"; }
    else if (!strcmp(mode, "qrtm")) { fmt = CQH_FMT_QRTM; defpar = "../input/visac2drtm.txt"; }
    else DIE("unknown mode %s (aps, vaps, artm, qrtm)", mode);
    const int q = (fmt == CQH_FMT_VAPS || fmt == CQH_FMT_QRTM);
    const int rtm = (fmt == CQH_FMT_ARTM || fmt == CQH_FMT_QRTM);
    const char *parfile = argc > 1 ? argv[1] : defpar;
    const char *outdir = argc > 2 ? argv[2] : "../obsdata";
    const int batch = argc > 3 ? atoi(argv[3]) : 64;

    cqh_params P;
    if (cqh_read_params(parfile, fmt, &P)) DIE("cannot read parameter file %s", parfile);
    const int shotnum = P.shotnum, tmax = P.tmax, nxt = P.nx, nzt = P.nz, pml = P.pml;
    const float dt = P.dt_ms / 1000.0;     /* acousticsgps4.c:182 */
    const float dx = P.dx, dz = P.dz, fd = P.fd;
    const size_t nzxt = (size_t)nxt * nzt;

    int *tracenum = (int *)malloc(sizeof(int) * shotnum);
    float *shotxp = (float *)malloc(sizeof(float) * shotnum), *shotzp = (float *)malloc(sizeof(float) * shotnum);
    float *vpt = (float *)malloc(sizeof(float) * nzxt), *rhot = (float *)malloc(sizeof(float) * nzxt);
    float *bbxt = (float *)malloc(sizeof(float) * nzxt), *bbzt = (float *)malloc(sizeof(float) * nzxt);
    float *qt = NULL, *gamt = NULL, *lamt = NULL, *vwdt = NULL, *taot = NULL;
    if (cqh_read_i32(P.fname1, tracenum, shotnum)) DIE("Cannot open the tracenumber file %s", P.fname1);
    if (cqh_read_f32(P.fname2, shotxp, shotnum)) DIE("Cannot open the shot x position file %s", P.fname2);
    if (cqh_read_f32(P.fname3, shotzp, shotnum)) DIE("Cannot open the shot z position file %s", P.fname3);
    if (cqh_read_f32(P.vpname, vpt, nzxt)) DIE("Cannot open the velocity file %s", P.vpname);
    if (cqh_read_f32(P.rhoname, rhot, nzxt)) DIE("Cannot open the density file %s", P.rhoname);
    float vmax, vmin;
    cqh_get_max_min((int)nzxt, vpt, &vmax, &vmin);
    printf(" CFL number= %f %f %f\n", cqh_check_stability(vmax, dt, dx, dz), vmax, vmin);
    if (q) {
        qt = (float *)malloc(sizeof(float) * nzxt);
        gamt = (float *)malloc(sizeof(float) * nzxt); lamt = (float *)malloc(sizeof(float) * nzxt);
        vwdt = (float *)malloc(sizeof(float) * nzxt); taot = (float *)malloc(sizeof(float) * nzxt);
        if (cqh_read_f32(P.qname, qt, nzxt)) DIE("Cannot open the Q file %s", P.qname);
        cqh_transfermodel((int)nzxt, fd, P.fo, vpt, qt, gamt, lamt, vwdt, taot);
    }
    cqh_bbx_bbz(nxt, nzt, rhot, bbxt, bbzt);
    float *rik = (float *)malloc(sizeof(float) * tmax);
    cqh_ricker(tmax, dt, fd, rik, P.flagsource);

    const float xmax = (nxt - 1) * dx, zmax = (nzt - 1) * dz;
    for (int s = 0; s < shotnum; s++)
        if (shotxp[s] < 0 || shotxp[s] > xmax || shotzp[s] < 0 || shotzp[s] > zmax)
            DIE("%d shot exceeds model!", s);
    cq_ctx ctx;
    if (cq_open(0, &ctx)) DIE("cq_open: %s", cq_last_error());
    int nrmx = 0;
    for (int s = 0; s < shotnum; s++) if (tracenum[s] > nrmx) nrmx = tracenum[s];
    float *gxp = (float *)malloc(sizeof(float) * nrmx), *gzp = (float *)malloc(sizeof(float) * nrmx);
    float *img_sp = NULL, *ssg_sp = NULL;
    if (rtm) {
        img_sp = (float *)calloc(nzxt, sizeof(float));
        ssg_sp = (float *)calloc(nzxt, sizeof(float));
    }
    /* The reference splits the shots over numprocs MPI ranks of GPU_N GPUs each
     * (acousticsgps4.c:255-273): rank r runs shots offsp .. offsp + nshot GPU_N - 1
     * with a window sized over its own shots, and the last shotnum % (numprocs
     * GPU_N) % GPU_N shots run nowhere.  CQ_SHOT_SPLIT="numprocs,GPU_N" reproduces
     * that split (window per rank, dropped shots); by default every shot runs
     * with one window over all shots. */
    int np = 1, gpn = 1, split = 0;
    const char *sp = getenv("CQ_SHOT_SPLIT");
    if (sp && sscanf(sp, "%d,%d", &np, &gpn) == 2 && np > 0 && gpn > 0) split = 1;
    else np = gpn = 1;
    const double t0 = now_s();
    for (int rank = 0; rank < np; rank++) {
        int g0 = 0, g1 = shotnum;
        if (split) {
            const int rid = shotnum % (np * gpn), ave = (shotnum - rid) / (np * gpn), ridg = rid / gpn;
            const int nshot = rank < ridg ? ave + 1 : ave;
            g0 = rank < ridg ? rank * (ave + 1) * gpn : ridg * (ave + 1) * gpn + (rank - ridg) * ave * gpn;
            g1 = g0 + nshot * gpn;
            if (rank == np - 1) printf("Run total shot number is %d\n", g1);
            if (g1 <= g0) continue;
        }
        int nx = 0;
        const float sroffmx = cqh_window_sroffmx(g0, g1, tracenum, shotxp, P.fname4, &nx, P.disx, dx);
        if (sroffmx < 0) DIE("Cannot open receiver positions file %s<shot>.bin", P.fname4);
        const int nz = nzt, nxx = nx + 2 * pml, nzz = nz + 2 * pml;
        printf(" window %d x %d (padded %d x %d), %d shots, tmax %d\n", nx, nz, nxx, nzz, g1 - g0, tmax);

        float *iax = (float *)malloc(4 * nxx), *ibx = (float *)malloc(4 * nxx), *hax = (float *)malloc(4 * nxx),
              *hbx = (float *)malloc(4 * nxx), *iaz = (float *)malloc(4 * nzz), *ibz = (float *)malloc(4 * nzz),
              *haz = (float *)malloc(4 * nzz), *hbz = (float *)malloc(4 * nzz);
        cqh_pml_coeffs(pml, fd, iax, ibx, iaz, ibz, hax, hbx, haz, hbz, nxx, nzz, dx, dz, dt);
        float *kxsmr = (float *)malloc(4 * nxx), *kxscr = (float *)malloc(4 * nxx), *kxsmi = (float *)malloc(4 * nxx),
              *kxsci = (float *)malloc(4 * nxx), *kzsmr = (float *)malloc(4 * nzz), *kzscr = (float *)malloc(4 * nzz),
              *kzsmi = (float *)malloc(4 * nzz), *kzsci = (float *)malloc(4 * nzz);
        cqh_shift_coeffs(kxsmr, kxscr, kzsmr, kzscr, kxsmi, kxsci, kzsmi, kzsci, dx, dz, nxx, nzz);

        /* shots run in batches of `chunk`; a batch that does not fit the
         * device (an RTM shot holds snaps checkpoints of its state) is retried
         * with half as many shots.  RTM images are stacked on the host per
         * shot, so nothing needs all shots in one batch. */
        int chunk = batch > 0 ? batch : 64;
        for (int s0 = g0; s0 < g1;) {
            const int ns = (s0 + chunk <= g1) ? chunk : g1 - s0;
            shot_buf *sb = (shot_buf *)calloc(ns, sizeof(shot_buf));
            cq_job_desc *jd = (cq_job_desc *)calloc(ns, sizeof(cq_job_desc));
            for (int k = 0; k < ns; k++) {
                const int sid = s0 + k;
                shot_buf *b = &sb[k];
                b->shotid = sid;
                b->nr = tracenum[sid];
                const float xs = shotxp[sid], zs = shotzp[sid];
                const float mlx = xs - sroffmx - P.disx;
                b->vp = win_field(vpt, nxt, nzt, nx, nz, pml, mlx, dx);
                b->rho = win_field(rhot, nxt, nzt, nx, nz, pml, mlx, dx);
                b->bbx = win_field(bbxt, nxt, nzt, nx, nz, pml, mlx, dx);
                b->bbz = win_field(bbzt, nxt, nzt, nx, nz, pml, mlx, dx);
                if (q) {
                    b->gam = win_field(gamt, nxt, nzt, nx, nz, pml, mlx, dx);
                    b->lam = win_field(lamt, nxt, nzt, nx, nz, pml, mlx, dx);
                    b->vwd = win_field(vwdt, nxt, nzt, nx, nz, pml, mlx, dx);
                    b->tao = win_field(taot, nxt, nzt, nx, nz, pml, mlx, dx);
                }
                if (cqh_read_geophones(P.fname4, sid, b->nr, gxp) || cqh_read_geophones(P.fname5, sid, b->nr, gzp))
                    DIE("Cannot open receiver positions file for shot %d", sid);
                b->gxg = (int *)malloc(sizeof(int) * (b->nr > 0 ? b->nr : 1));
                b->gzg = (int *)malloc(sizeof(int) * (b->nr > 0 ? b->nr : 1));
                if (cqh_shot_indices(pml, nxx, nzz, dx, dz, sroffmx, P.disx, xs, zs, b->nr, gxp, gzp, b->gxg, b->gzg,
                                     &b->geo))
                    DIE("%d shot grid index exceeds model!", sid);
                if (b->geo.nclamped) printf("warning: %d receivers of shot %d clamped out of the x pml\n", b->geo.nclamped, sid);
                if (rtm) {
                    char name[512];
                    b->obs = (float *)malloc(sizeof(float) * (size_t)tmax * b->nr);
                    snprintf(name, sizeof name, "%s%d_%d.bin", P.recname, sid, b->nr);
                    if (cqh_read_f32(name, b->obs, (long)tmax * b->nr)) DIE("Cannot open observed data %s", name);
                    cqh_removedirect(tmax, b->nr, dt, fd, xs, zs, gxp, gzp, P.v0, b->obs);
                }
                cq_job_desc *j = &jd[k];
                j->kind = rtm ? (q ? CQ_KIND_QRTM : CQ_KIND_RTM) : (q ? CQ_KIND_VISCO : CQ_KIND_ACOUSTIC);
                j->nxx = nxx; j->nzz = nzz; j->pml = pml; j->nt = tmax; j->dt = dt;
                j->vp = b->vp; j->rho = b->rho; j->bbx = b->bbx; j->bbz = b->bbz;
                j->gam = b->gam; j->lam = b->lam; j->vwd = b->vwd; j->tao = b->tao;
                j->iax = iax; j->ibx = ibx; j->hax = hax; j->hbx = hbx;
                j->iaz = iaz; j->ibz = ibz; j->haz = haz; j->hbz = hbz;
                j->kxsmr = kxsmr; j->kxsmi = kxsmi; j->kxscr = kxscr; j->kxsci = kxsci;
                j->kzsmr = kzsmr; j->kzsmi = kzsmi; j->kzscr = kzscr; j->kzsci = kzsci;
                j->dx = dx; j->dz = dz; j->rik = rik;
                j->xsg = b->geo.xsg; j->zsg = b->geo.zsg;
                j->nr = b->nr; j->gxg = b->gxg; j->gzg = b->gzg;
                j->obs = b->obs; j->snaps = P.snapnum; j->etas = P.etas; j->etar = P.etar;
                if (q && P.lr_rank > 0) {
                    /* the viscoacoustic operator in low-rank form: FP64
                     * decomposition of this shot's window, seeded by its id */
                    const int r = P.lr_rank;
                    const size_t N1 = (size_t)nxx * nzz, Nq = (size_t)(nxx / 2 + 1) * (nzz / 2 + 1);
                    float *cc = (float *)malloc(sizeof(float) * N1);
                    const float dtf = dt * dt, c12 = (float)(1.0 / 12.0);
                    for (size_t a = 0; a < N1; a++) cc[a] = c12 * dtf * (b->vp[a] * b->vp[a]) * b->lam[a];
                    double *U = (double *)malloc(sizeof(double) * N1 * r), *A = (double *)malloc(sizeof(double) * r * r),
                           *V = (double *)malloc(sizeof(double) * Nq * r);
                    if (cqh_lowrank(nxx, nzz, dx, dz, b->gam, b->vwd, cc, CQH_LR_FRACLAP, r,
                                    1ULL + (unsigned long long)sid, U, A, V))
                        DIE("low-rank decomposition of shot %d failed", sid);
                    b->lrw = (float *)malloc(sizeof(float) * N1 * r);
                    b->lrs = (float *)malloc(sizeof(float) * Nq * r);
                    cqh_lowrank_pack(nxx, nzz, r, U, A, V, b->lrw, b->lrs);
                    free(U); free(A); free(V); free(cc);
                    j->lr_rank = r; j->lr_w = b->lrw; j->lr_sym = b->lrs;
                }
                printf("shotid=%d nxx=%d nzz=%d xs=%f zs=%f\n", sid, nxx, nzz, xs, zs);
            }
            cq_batch bt;
            const int rc = cq_batch_create(ctx, jd, ns, &bt);
            if (rc == CQ_ENOMEM && ns > 1) {
                printf(" %d shots do not fit the device (%s); retrying with %d\n", ns, cq_last_error(), (ns + 1) / 2);
                for (int k = 0; k < ns; k++) free_shot(&sb[k]);
                free(sb);
                free(jd);
                chunk = (ns + 1) / 2;
                continue;
            }
            if (rc) DIE("cq_batch_create: %s", cq_last_error());
            if (!rtm && P.flagout == 1) {
                /* snapshots (sgpstd-qmodeling/acps4.cu:482-494, 624-640): the
                 * interior of p after every step t with t % 10 == 0, appended
                 * x-major to ../obsdata/snapa<shot>_<nz>_<nx>.bin (snapva for the
                 * viscoacoustic binary), each file truncated first */
                float *pf = (float *)malloc(sizeof(float) * (size_t)nxx * nzz);
                char name[512];
                for (int k = 0; k < ns; k++) {
                    snprintf(name, sizeof name, "%s/%s%d_%d_%d.bin", outdir, q ? "snapva" : "snapa", sb[k].shotid, nz, nx);
                    FILE *fs = fopen(name, "wb");
                    if (!fs) DIE("cannot write %s", name);
                    fclose(fs);
                }
                for (int t = 0; t < tmax; t++) {
                    if (cq_batch_step(bt, 1)) DIE("cq_batch_step: %s", cq_last_error());
                    if (t % 10) continue;
                    for (int k = 0; k < ns; k++) {
                        if (cq_batch_read_p(bt, k, pf)) DIE("cq_batch_read_p: %s", cq_last_error());
                        snprintf(name, sizeof name, "%s/%s%d_%d_%d.bin", outdir, q ? "snapva" : "snapa", sb[k].shotid, nz, nx);
                        FILE *fs = fopen(name, "ab");
                        if (!fs) DIE("cannot write %s", name);
                        for (int ix = pml; ix < nxx - pml; ix++)
                            if (fwrite(pf + (size_t)ix * nzz + pml, sizeof(float), (size_t)nz, fs) != (size_t)nz)
                                DIE("short write of %s", name);
                        fclose(fs);
                    }
                }
                free(pf);
            }
            if (cq_batch_run(bt)) DIE("cq_batch_run: %s", cq_last_error());
            for (int k = 0; k < ns; k++) {
                shot_buf *b = &sb[k];
                if (!rtm) {
                    char name[512];
                    float *rec = (float *)malloc(sizeof(float) * (size_t)tmax * (b->nr > 0 ? b->nr : 1));
                    if (cq_batch_read_rec(bt, k, rec)) DIE("shot %d: %s", b->shotid, cq_last_error());
                    snprintf(name, sizeof name, "%s/%s%d_%d.bin", outdir, q ? "recva" : "reca", b->shotid, b->nr);
                    if (cqh_write_f32(name, rec, (long)tmax * b->nr)) DIE("cannot write %s", name);
                    free(rec);
                } else {
                    /* post-processing of one shot (sgpstd-qrtm/acps4.cu:909-945) */
                    float *img = (float *)malloc(sizeof(float) * (size_t)nx * nz);
                    float *imgf = (float *)malloc(sizeof(float) * (size_t)nx * nz);
                    float *ssg = (float *)malloc(sizeof(float) * (size_t)nx * nz);
                    float *pj = (float *)calloc(nzxt, sizeof(float));
                    if (cq_batch_read_image(bt, k, img, ssg)) DIE("shot %d: %s", b->shotid, cq_last_error());
                    cqh_filter_image(nx, nz, dx, dz, img, imgf);
                    cqh_mute_image(nx, nz, pml, b->geo.lgest, b->geo.rgest, imgf);
                    cqh_projectomodel(pml, b->geo.mlx, dx, nxt, nzt, nx, nz, nxx, nzz, pj, imgf);
                    for (size_t i = 0; i < nzxt; i++) img_sp[i] += pj[i];
                    memset(pj, 0, sizeof(float) * nzxt);
                    cqh_projectomodel(pml, b->geo.mlx, dx, nxt, nzt, nx, nz, nxx, nzz, pj, ssg);
                    for (size_t i = 0; i < nzxt; i++) ssg_sp[i] += pj[i];
                    free(img); free(imgf); free(ssg); free(pj);
                }
                free_shot(b);
            }
            cq_batch_destroy(bt);
            free(sb);
            free(jd);
            s0 += ns;
        }
        free(iax); free(ibx); free(hax); free(hbx); free(iaz); free(ibz); free(haz); free(hbz);
        free(kxsmr); free(kxscr); free(kxsmi); free(kxsci); free(kzsmr); free(kzscr); free(kzsmi); free(kzsci);
    }
    const double el = now_s() - t0;
    printf("The cost of the run time is %f seconds\n", el);
    if (rtm) {
        const char *out = q ? "QRTM.BIN" : "RTM.BIN";
        FILE *fp = fopen(out, "wb");
        if (!fp) DIE("cannot write %s", out);
        size_t w1 = fwrite(img_sp, sizeof(float), nzxt, fp);
        size_t w2 = fwrite(ssg_sp, sizeof(float), nzxt, fp);
        fclose(fp);
        if (w1 != nzxt || w2 != nzxt) DIE("short write of %s", out);
    }
    cq_close(ctx);
    return 0;
}
