/*
 * cq_shim.h -- the bodies of the reference's four acforward entry points on
 * the engine (libcq_shim.so), for maintainers who keep the reference's own
 * drivers.
 *
 * Each reference binary links one acforward (sgpstd-qmodeling/acps4.cu:184-198
 * for aps, acsglw.cu:185-200 for vaps, sgpstd-qrtm/acps4.cu:183-198 for artm,
 * sgpstd-qrtm/acsglw.cu:189-204 for qrtm) that takes one MPI rank's shots
 * through `struct MultiGPU singpu[]` (statement.h:15), windows the full
 * models per shot (select_part_model / expand_pml_areas), computes the
 * source and receiver grid indices (acps4.cu:391-426), steps each shot with
 * cuFFT and writes its outputs: ../obsdata/reca<shot>_<nr>.bin (recva for
 * vaps, time-major) and, with flagout = 1, the interior of p after every
 * 10th step to ../obsdata/snapa<shot>_<nz>_<nx>.bin (snapva); for artm/qrtm
 * it reads <recname><shot>_<nr>.bin, removes the direct wave, migrates, and
 * adds the Laplacian-filtered, muted, projected image and the projected
 * illumination of every shot to img_sp / ssg_sp (sgpstd-qrtm/acps4.cu:909-945).
 * The functions below do the same for shots given as plain arrays (the shot
 * ids and source positions the caller reads out of
 * singpu[ip].shotid/sxp/szp[ns]), advancing all of them together in one
 * batch.  The remaining arguments are acforward's own, in its order; k2d1,
 * k2da, k2d2, k2d3, myid, numprocs, nrmx and (modeling) snapnum are not
 * needed.  `outdir` replaces the hard-coded "../obsdata".
 *
 * Each returns 0, or a negative cq_status of the engine (message in
 * cq_last_error()), or -100 for a file that cannot be read or written and
 * -101 for a shot whose source lies outside the window (the reference
 * prints "shot grid index exceeds model!" and exits; the message is printed
 * here too, nothing exits).
 */
#ifndef CQ_SHIM_H
#define CQ_SHIM_H

#ifdef __cplusplus
extern "C" {
#endif

int cq_shim_aps(int tmax, float dt, const float *rik, int nshots, const int *shotid, const float *sxp,
                const float *szp, int nx, int nz, int nxx, int nzz, int pml, float dx, float dz, float sroffmx,
                float disx, const float *vpt, const float *rhot, const float *bbxt, const float *bbzt, int nxt,
                int nzt, const int *tracenum, int flagout, const float *i_ax, const float *i_bx,
                const float *i_az, const float *i_bz, const float *h_ax, const float *h_bx, const float *h_az,
                const float *h_bz, const float *kxsmr, const float *kxscr, const float *kzsmr,
                const float *kzscr, const float *kxsmi, const float *kxsci, const float *kzsmi,
                const float *kzsci, const char *fname4, const char *fname5, const char *outdir);

int cq_shim_vaps(int tmax, float dt, const float *rik, int nshots, const int *shotid, const float *sxp,
                 const float *szp, int nx, int nz, int nxx, int nzz, int pml, float dx, float dz, float sroffmx,
                 float disx, const float *vpt, const float *vwdt, const float *taot, const float *lamt,
                 const float *gamt, const float *rhot, const float *bbxt, const float *bbzt, int nxt, int nzt,
                 const int *tracenum, int flagout, const float *i_ax, const float *i_bx, const float *i_az,
                 const float *i_bz, const float *h_ax, const float *h_bx, const float *h_az, const float *h_bz,
                 const float *kxsmr, const float *kxscr, const float *kzsmr, const float *kzscr,
                 const float *kxsmi, const float *kxsci, const float *kzsmi, const float *kzsci,
                 const char *fname4, const char *fname5, const char *outdir);

int cq_shim_artm(int snapnum, int tmax, float dt, const float *rik, int nshots, const int *shotid,
                 const float *sxp, const float *szp, int nx, int nz, int nxx, int nzz, int pml, float dx,
                 float dz, float sroffmx, float disx, const float *vpt, const float *rhot, const float *bbxt,
                 const float *bbzt, int nxt, int nzt, const int *tracenum, const float *i_ax,
                 const float *i_bx, const float *i_az, const float *i_bz, const float *h_ax, const float *h_bx,
                 const float *h_az, const float *h_bz, const float *kxsmr, const float *kxscr,
                 const float *kzsmr, const float *kzscr, const float *kxsmi, const float *kxsci,
                 const float *kzsmi, const float *kzsci, float *img_sp, float *ssg_sp, float fobs, float v0,
                 const char *fname4, const char *fname5, const char *recname);

int cq_shim_qrtm(int snapnum, int tmax, float dt, const float *rik, int nshots, const int *shotid,
                 const float *sxp, const float *szp, int nx, int nz, int nxx, int nzz, int pml, float dx,
                 float dz, float sroffmx, float disx, const float *vpt, const float *vwdt, const float *taot,
                 const float *lamt, const float *gamt, const float *rhot, const float *bbxt, const float *bbzt,
                 int nxt, int nzt, const int *tracenum, const float *i_ax, const float *i_bx,
                 const float *i_az, const float *i_bz, const float *h_ax, const float *h_bx, const float *h_az,
                 const float *h_bz, const float *kxsmr, const float *kxscr, const float *kzsmr,
                 const float *kzscr, const float *kxsmi, const float *kxsci, const float *kzsmi,
                 const float *kzsci, float *img_sp, float *ssg_sp, float fobs, float v0, float etas,
                 float etar, const char *fname4, const char *fname5, const char *recname);

#ifdef __cplusplus
}
#endif
#endif
