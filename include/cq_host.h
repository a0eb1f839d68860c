/*
 * cq_host.h -- host-side formats, parameter files and coefficient builders of
 * the constant-Q PSTD engine (plain C, libcq_host.so).
 *
 * Everything here runs on the CPU and produces the FP32 arrays the engine
 * consumes.  Each builder states the reference function it replaces; the
 * arithmetic (float storage, double libm calls, pi = 3.1415926) is the
 * reference's, so the arrays are bit-identical to what the reference's host
 * code produces on the same inputs.
 */
#ifndef CQ_HOST_H
#define CQ_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

#define CQH_PI 3.1415926 /* sgpstd-qmodeling/funca.c:6 */

/* ---- parameter files (label line, value line; optional trailing keys) ---- */

/* Superset of the four reference parameter files:
 *   ac2drealmodeling.txt  (sgpstd-qmodeling/funca.c:9-114)
 *   visac2drealmodeling.txt (sgpstd-qmodeling/func.c:9-126)
 *   ac2drtm.txt          (sgpstd-qrtm/funca.c:9-129)
 *   visac2drtm.txt       (sgpstd-qrtm/func.c:9-146)
 * Optional trailing label/value pairs understood after the reference's last
 * value: "pml" (CPML width, default 10 = the reference's #define) and
 * "low-rank" (vaps / qrtm: rank of the low-rank form of the viscoacoustic
 * operator, cq_lowrank.h; default 0 = the reference's Taylor form). */
typedef struct {
    int shotnum, tmax;
    float dt_ms, fd, fo;
    int nx, nz;
    float dx, dz, disx;
    float v0, etas, etar;         /* RTM: direct-wave velocity, stabilisers */
    char fname1[40], fname2[40], fname3[40], fname4[40], fname5[40];
    char vpname[80], rhoname[80], qname[80];
    char recname[60];
    int flagout, flagsource, snapnum;
    int pml;                       /* optional trailing key, default 10 */
    int lr_rank;                   /* optional trailing key, default 0 */
} cqh_params;

enum { CQH_FMT_APS = 0, CQH_FMT_VAPS = 1, CQH_FMT_ARTM = 2, CQH_FMT_QRTM = 3 };

/* Returns 0 on success, negative on a missing file or a short read. */
int cqh_read_params(const char *path, int fmt, cqh_params *out);

/* ---- binary formats (float32 / int32, native endianness) ---- */
int cqh_read_f32(const char *path, float *dst, long count);
int cqh_read_i32(const char *path, int *dst, long count);
int cqh_write_f32(const char *path, const float *src, long count);
/* "<prefix><shotid>.bin" receiver coordinates (funca.c:587-600) */
int cqh_read_geophones(const char *prefix, int shotid, int nr, float *dst);

/* ---- coefficient builders ---- */

/* funca.c:504-518 get_max_min (with its 1e-4 hysteresis) */
void cqh_get_max_min(int n, const float *v, float *vmax, float *vmin);
/* funca.c:533-538 */
float cqh_check_stability(float vmax, float dt, float dx, float dz);
/* funca.c:169-197 */
void cqh_bbx_bbz(int nxx, int nzz, const float *dens, float *bbx, float *bbz);
/* funca.c:201-220 (flag 2: Ricker, else derivative of Gaussian) */
void cqh_ricker(int tmax, float dt, float fd, float *rik, int flag);
/* funca.c:223-340 */
void cqh_pml_coeffs(int pml, float fm,
                    float *i_ax, float *i_bx, float *i_az, float *i_bz,
                    float *h_ax, float *h_bx, float *h_az, float *h_bz,
                    int nxx, int nzz, float dx, float dz, float dt);
/* funca.c:342-427 */
void cqh_shift_coeffs(float *kxsmr, float *kxscr, float *kzsmr, float *kzscr,
                      float *kxsmi, float *kxsci, float *kzsmi, float *kzsci,
                      float dx, float dz, int nxx, int nzz);
/* funca.c:540-573 (k2d1 only; k2d2/k2d3/k2da may be NULL) and
 * func.c:550-583 (all four). */
void cqh_k2d(float dx, float dz, int nxx, int nzz,
             float *k2d1, float *k2d2, float *k2d3, float *k2da);
/* func.c:912-929 */
void cqh_transfermodel(int n, float fd, float fo, const float *vp, const float *qp,
                       float *gam, float *lam, float *vwd, float *tao);
/* func.c:933-943 */
float cqh_get_max_tao(int n, const float *tao);

/* ---- windowing ---- */

/* funca.c:623-733: copy the shot window of the full model into the padded
 * window (interior part; PML filled by cqh_expand_pml). */
void cqh_select_part_model(int pml, float mlx, float dx, int nxt, int nzt, int nx, int nz,
                           int nxx, int nzz, const float *full, float *win);
/* funca.c:735-770 */
void cqh_expand_pml(int nx, int nz, int nxx, int nzz, int pml, float *win);
/* funca.c:803-876: window interior (nx*nz) -> full model (nxt*nzt) */
void cqh_projectomodel(int pml, float mlx, float dx, int nxt, int nzt, int nx, int nz,
                       int nxx, int nzz, float *full, const float *win);
/* funca.c:878-899 */
void cqh_removedirect(int tmax, int nr, float dt, float fd, float xs, float zs,
                      const float *gxp, const float *gzp, float v0, float *obs);

/* ---- shot geometry (acousticsgps4.c:307-336, acps4.cu:391-426) ---- */

/* Window half-width scan over shots [s0, s1): returns sroffmx in metres, or a
 * negative value if a receiver file cannot be read; *nx_out receives
 * (int)(2*(sroffmx+disx)/dx + 1.5). */
float cqh_window_sroffmx(int s0, int s1, const int *tracenum, const float *shotxp,
                         const char *gxprefix, int *nx_out, float disx, float dx);

typedef struct {
    int xsg, zsg;        /* source grid index in the padded window */
    int nr;
    int lgest, rgest;    /* leftmost / rightmost receiver column (padded) */
    int nclamped;        /* receivers clamped out of the x PML */
    float mlx;           /* window left edge in metres */
} cqh_shot_geom;

/* Receiver and source indices, bit-exact with acps4.cu:391-426.  gxg/gzg get
 * nr entries.  Returns 0, or -1 if the source falls outside the window. */
int cqh_shot_indices(int pml, int nxx, int nzz, float dx, float dz,
                     float sroffmx, float disx, float xs, float zs,
                     int nr, const float *gxp, const float *gzp,
                     int *gxg, int *gzg, cqh_shot_geom *geo);

/* ---- RTM image post-processing (sgpstd-qrtm/acps4.cu:909-945) ---- */

/* filter_image (acps4.cu:1723-1745): 5-point Laplacian of the raw image over
 * the interior nx*nz, zero on the border. */
void cqh_filter_image(int nx, int nz, float dx, float dz, const float *img, float *imgf);
/* Zero image columns ix < lgest - pml and ix >= rgest - pml (acps4.cu:926-935;
 * lgest/rgest are padded-grid receiver columns, so the rightmost receiver
 * column is zeroed too). */
void cqh_mute_image(int nx, int nz, int pml, int lgest, int rgest, float *img);

#ifdef __cplusplus
}
#endif
#endif
