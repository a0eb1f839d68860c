/*
 * cq.h -- C ABI of the MI355X constant-Q PSTD engine (libcq_hip.so).
 *
 * This is the drop-in boundary for the reference's per-binary GPU entry point
 *     extern "C" void acforward(...)
 * of sgpstd-qmodeling/acps4.cu:184-198 (aps), sgpstd-qmodeling/acsglw.cu:185-200
 * (vaps), sgpstd-qrtm/acps4.cu:183-198 (artm) and sgpstd-qrtm/acsglw.cu:189-204
 * (qrtm).  The reference passes one MultiGPU struct per device
 * (sgpstd-qmodeling/statement.h:15-195) holding host arrays the driver filled
 * and device pointers acforward allocates; here the driver describes each shot
 * with a cq_job_desc of plain host pointers and sizes, and the engine owns all
 * device memory.  No torch types, no exit(): every call returns a cq_status and
 * cq_last_error() describes the last failure of the calling thread.
 *
 * Layout of every 2-D array: x-major with z fastest, a[ix * nzz + iz], the
 * padded window of the reference (nxx = nx + 2 pml, nzz = nz + 2 pml).
 * Gathers are time-major, rec[t * nr + i] (acps4.cu:620).
 */
#ifndef CQ_H
#define CQ_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CQ_OK = 0,
    CQ_EINVAL = -1,        /* bad argument */
    CQ_EDEVICE = -2,       /* HIP runtime error */
    CQ_ENOMEM = -3,        /* device allocation failed */
    CQ_EUNSUPPORTED = -4,  /* e.g. a window whose transforms exceed one workgroup's LDS */
    CQ_ESTATE = -5,        /* call out of order (e.g. reading results before running) */
    CQ_ENUMERIC = -6       /* non-finite wavefield detected */
} cq_status;

typedef enum {
    CQ_KIND_ACOUSTIC = 0,  /* aps : acps4.cu (modeling)               */
    CQ_KIND_VISCO = 1,     /* vaps: acsglw.cu (constant-Q modeling)   */
    CQ_KIND_RTM = 2,       /* artm: sgpstd-qrtm/acps4.cu              */
    CQ_KIND_QRTM = 3       /* qrtm: sgpstd-qrtm/acsglw.cu             */
} cq_kind;

typedef struct cq_ctx_s *cq_ctx;
typedef struct cq_batch_s *cq_batch;

/* One shot.  All pointers are host memory owned by the caller and only read
 * during cq_batch_create (outputs are written by cq_batch_read_*). */
typedef struct {
    int kind;                 /* cq_kind */
    int nxx, nzz, pml;        /* padded window and CPML width */
    int nt;                   /* time steps (tmax) */
    float dt;                 /* seconds (the reference divides the ms value by 1000) */
    /* window fields, nxx*nzz (select_part_model + expand_pml_areas output) */
    const float *vp, *rho, *bbx, *bbz;
    const float *gam, *lam, *vwd, *tao;   /* kind VISCO / QRTM (transfermodel) */
    /* CPML profiles (get_pml_parameters): x arrays nxx, z arrays nzz */
    const float *iax, *ibx, *hax, *hbx;
    const float *iaz, *ibz, *haz, *hbz;
    /* staggered shift multipliers (cal_shift_parameters) */
    const float *kxsmr, *kxsmi, *kxscr, *kxsci;   /* nxx */
    const float *kzsmr, *kzsmi, *kzscr, *kzsci;   /* nzz */
    float dx, dz;             /* grid spacing (for the |k| tables) */
    const float *rik;         /* source wavelet, nt (rik_cal) */
    int xsg, zsg;             /* source grid index */
    int nr;                   /* receivers */
    const int *gxg, *gzg;     /* receiver grid indices, nr each */
    /* RTM / Q-RTM (kind 2, 3) */
    const float *obs;         /* observed gather nt*nr, time-major, direct wave
                               * already removed (removedirect, funca.c:878-899) */
    int snaps;                /* revolve checkpoints, 1..64 (input file "snapnum") */
    float etas, etar;         /* Q-RTM stabilisers of the source / receiver
                               * wavefields, in units of 1e-6 (visac2drtm.txt) */
    /* Kinds VISCO / QRTM: the operator applied to the velocity divergence,
     * total = sum_i w_i(x) F^-1[s_i(k) F div], as a rank-r low-rank form
     * (cq_lowrank.h: cqh_lowrank + cqh_lowrank_pack, the FP64 host
     * decomposition of sgpstd-qmodeling/sglowrank.c:148-480 applied like
     * acps4.cu:1337-1459).  lr_rank 0 keeps the reference's live Taylor form
     * (sum_total, acsglw.cu:1213-1245). */
    int lr_rank;              /* 0, or 1..8 */
    const float *lr_w;        /* lr_rank * nxx*nzz per-cell weights, x-major */
    const float *lr_sym;      /* lr_rank * (nxx/2+1)*(nzz/2+1) symbols / (nxx nzz) on the
                               * quarter plane |kx|, |kz| (kz fastest) */
} cq_job_desc;

/* Opens device `device` (HIP ordinal) and creates the engine's stream. */
int cq_open(int device, cq_ctx *out);
int cq_close(cq_ctx ctx);
/* Text of the last error on this thread ("" if none). */
const char *cq_last_error(void);
/* Library version string. */
const char *cq_version(void);

/* Uploads njobs shots (models, coefficients, receiver maps) and zeroes their
 * state.  Shots of different sizes and kinds (modeling, RTM, Q-RTM) may be
 * mixed; every launch advances all of them.  Replaces the per-shot
 * acforward calls of sgpstd-qmodeling/acps4.cu:184-198, acsglw.cu:185-200,
 * sgpstd-qrtm/acps4.cu:183-198 and acsglw.cu:189-204. */
int cq_batch_create(cq_ctx ctx, const cq_job_desc *jobs, int njobs, cq_batch *out);
/* Enqueues nsteps global steps (asynchronous).  A modeling shot takes one
 * time step per global step until its nt; an RTM shot takes the next step of
 * its own revolve schedule (src.c:314-470: forward source steps, adjoint
 * receiver steps and the checkpoint copies between them), which has
 * numforw + nt steps (cqh_numforw).  The shots are dealt into four launch
 * groups that run on the context's stream and three streams of the batch;
 * they are joined back into the context's stream before this returns. */
int cq_batch_step(cq_batch b, int nsteps);
/* Waits for the batch's queued work. */
int cq_batch_sync(cq_batch b);
/* Runs every shot to its end and waits. */
int cq_batch_run(cq_batch b);
/* Copies modeling shot `job`'s gather (nt*nr floats) to host memory.
 * Returns CQ_ENUMERIC (after copying) if the gather holds a non-finite value,
 * i.e. the shot blew up (for instance a time step beyond the CFL limit; the
 * reference's own check is disabled, acousticsgps4.c:221-224). */
int cq_batch_read_rec(cq_batch b, int job, float *rec);
/* Copies shot `job`'s pressure field (nxx*nzz floats) to host memory. */
int cq_batch_read_p(cq_batch b, int job, float *p);
/* RTM shots: copies shot `job`'s raw cross-correlation image and source
 * illumination, nx*nz floats each over the interior (x-major), before the
 * Laplacian filter, mute and projection of sgpstd-qrtm/acps4.cu:909-945.
 * Returns CQ_ENUMERIC (after copying) if either holds a non-finite value. */
int cq_batch_read_image(cq_batch b, int job, float *img, float *ssg);
/* Global steps taken so far by the batch. */
int cq_batch_time(cq_batch b);
/* Device bytes held by the batch. */
long long cq_batch_bytes(cq_batch b);
int cq_batch_destroy(cq_batch b);

/* Serving (continuous batching; no single counterpart in the reference, whose
 * drivers loop over shots with MPI, sgpstd-qmodeling/acousticsgps4.c:255-452).
 * A server keeps a device-resident batch.  Submitted jobs wait in a queue
 * until the device memory budget and the job limit admit them, then join
 * the launches of the jobs already running at the next global step.
 * Finished jobs leave the batch, their results stay on the host until
 * released.  Workgroup tiles are chosen for the largest window; a job whose
 * transforms those do not fit within the LDS budget (e.g. a length with a
 * large prime factor) runs with tiles of its own in launches of its own. */
typedef struct cq_server_s *cq_server;

typedef struct {
    int max_nxx, max_nzz;   /* largest padded window a job may have */
    long long hbm_budget;   /* device bytes the resident jobs may hold; 0: 90% of free memory */
    int max_jobs;           /* jobs resident at once; 0: no limit */
    /* Work cells advanced per global step (0: every resident job advances at
     * every step, the throughput-first default).  Otherwise the resident jobs
     * with the least remaining work (steps left x cells, viscoacoustic cells
     * counted twice) advance, as many as fit this many work cells (at least
     * one); the others wait in HBM and resume where they stopped, chosen
     * again whenever a job finishes or is admitted and every 256 global
     * steps.  A job's results do not depend on when it runs. */
    long long run_cells;
} cq_server_desc;

/* CQ_JOB_FAILED: the job could not be admitted (e.g. its device allocation
 * failed) or its wavefield became non-finite (CQ_ENUMERIC, checked every few
 * hundred global steps and at completion); cq_server_result then returns that
 * status and cq_last_error() its reason.  The other jobs are not affected. */
typedef enum { CQ_JOB_QUEUED = 0, CQ_JOB_RUNNING = 1, CQ_JOB_DONE = 2, CQ_JOB_FAILED = 3 } cq_job_state;

int cq_server_open(cq_ctx ctx, const cq_server_desc *desc, cq_server *out);
/* Validates and copies the job (all arrays); returns its id. */
int cq_server_submit(cq_server s, const cq_job_desc *job, int *id);
/* Admits queued jobs, enqueues up to nsteps global steps and retires the
 * jobs that finish (their results are copied to the host); returns when the
 * steps ran or no job is left.  Optional outputs: jobs running and queued. */
int cq_server_advance(cq_server s, int nsteps, int *running, int *queued);
int cq_server_poll(cq_server s, int id, int *state);
/* Global steps at which the job was submitted, started and finished. */
int cq_server_job_steps(cq_server s, int id, int *submitted, int *started, int *finished);
/* A finished job's result: the gather (nt*nr) for modeling jobs, the raw
 * image and illumination (nx*nz each, as cq_batch_read_image) for RTM jobs
 * (ssg may be NULL for modeling jobs). */
int cq_server_result(cq_server s, int id, float *out, float *ssg);
int cq_server_release(cq_server s, int id);
int cq_server_close(cq_server s);

/* The plain path: one modeling shot (kind ACOUSTIC or VISCO; RTM kinds are
 * rejected with CQ_EINVAL), one launch sequence per time step, the gather
 * row copied to the host and the stream synchronised after every step (the
 * reference's per-step structure, acps4.cu:508-621).  rec receives nt*nr.
 * Stops with CQ_ENUMERIC at the first gather row holding a non-finite value. */
int cq_plain_forward(cq_ctx ctx, const cq_job_desc *job, float *rec);

/* Size of the static device scratch of cq_forward_static (64 MiB per GPU,
 * held from the library's load until the process ends). */
#define CQ_STATIC_BYTES (64LL << 20)

/* Allocation-free variant of cq_plain_forward.  No counterpart in the
 * reference (whose acforward allocates per call, sgpstd-qmodeling/acps4.cu:797-928):
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It runs
 * one modeling shot (kind ACOUSTIC or VISCO) to its end and copies the gather
 * (nt*nr floats) to rec; the results equal cq_plain_forward's bit for bit
 * (the same passes and tile shapes).  Every device buffer -- wavefield
 * state, coefficients, FFT twiddles, gather, work lists -- lives in one
 * static device array of CQ_STATIC_BYTES per GPU, and the call makes no
 * hipMalloc, hipFree, hipHostMalloc or stream creation; it synchronises the
 * context's stream before returning.  A shot whose device footprint exceeds
 * 90% of CQ_STATIC_BYTES (about a 1200 x 680 acoustic window at nt = 3000)
 * is refused with CQ_EINVAL.  The array is shared by every call on that
 * GPU and keeps the last call's contents (each call clears what must start
 * at zero): calls that overlap on one GPU, from several threads or
 * contexts, race and must be serialised by the caller. */
int cq_forward_static(cq_ctx ctx, const cq_job_desc *job, float *rec);

/* Tile shapes and FFT plans of a batch (no counterpart in the reference, whose
 * kernels take fixed CUDA block shapes and cuFFT plans, sgpstd-qmodeling/
 * acps4.cu:871-872).  Every pass runs one workgroup per tile:
 *   tz   depths per row tile of k_row_ac and k_q_r3: 4, 8, 12 or 16
 *   cw   columns per tile of k_col_ac: 2, 4, 8, 12, 16, 24 or 32
 *   cwq  columns per tile of the viscoacoustic column passes (k_q_c1, k_q_c2,
 *        k_q_c3, k_lr_c2): the same values
 *   tk   kz rows per tile of k_q_r1 and k_lr_r1: 1, 2, 4 or 8
 *   tk2  kz rows per tile of k_q_r2: 1, 2, 3, 4, 6, 8, 12 or 16
 *   plan FFT factorisation: 0 the fewest stages (powers of two as 16, odd
 *        parts in the fewest of 3..17 and the composites 15, 21, 25, 27),
 *        1 radices of at most 8 (no 16, 9 or composites)
 * A field of 0 (plan: -1) leaves that choice to the engine. */
typedef struct {
    int tz, cw, cwq, tk, tk2, plan;
} cq_tiles;

/* cq_batch_create with the given tiles (NULL: the engine's choice).  Fails
 * with CQ_EUNSUPPORTED if a window's transforms do not fit the tiles. */
int cq_batch_create_tiled(cq_ctx ctx, const cq_job_desc *jobs, int njobs, const cq_tiles *tiles, cq_batch *out);
/* The tiles a batch runs with. */
int cq_batch_get_tiles(cq_batch b, cq_tiles *out);

/* Autotuner: the tiles and FFT plan that run a batch of these jobs fastest on
 * this device.  Looks the windows (kind class and padded size of every job)
 * up in the text file `cache` (NULL: no cache); on a miss it times a few
 * steps of trial batches of the same windows -- every candidate of each
 * tile field and each plan, per-pass device times, then the whole step of
 * the best combination against the engine's own choice -- and appends the
 * winner to the cache.  The jobs themselves are not advanced (the trials use
 * copies of their descriptions with at most 8 steps).  *tuned_ms receives
 * the winner's ms per global step (optional); *from_cache 1 if the cache
 * answered (optional). */
int cq_autotune(cq_ctx ctx, const cq_job_desc *jobs, int njobs, const char *cache, cq_tiles *out,
                double *tuned_ms, int *from_cache);

/* Graph-captured steps (no counterpart in the reference, which launches
 * every kernel of every step from the host, sgpstd-qmodeling/acps4.cu:508-645).
 * With steps_per_graph = k > 0, cq_batch_step / cq_batch_run run spans of k
 * global steps as one HIP graph (the four launch groups' pass chains and
 * their step counters, captured once per combination of active RTM roles and
 * replayed), wherever the span holds no checkpoint copy and the same roles
 * throughout; other steps, and all steps while per-pass timing is on, are
 * launched directly.  Results are identical to direct launches bit for bit.
 * The graph of the roles active at the batch's current step is captured by
 * this call; other combinations when first met.  0 switches graphs off (the
 * default).  k at most 64.  CQ_EDEVICE if the capture failed (graphs are then
 * off). */
int cq_batch_use_graphs(cq_batch b, int steps_per_graph);

/* Per-pass device timing with HIP events on the engine's stream.  While
 * enabled, every launch of cq_batch_step is bracketed by events; passes are
 * numbered 0..CQ_NPASS-1 (cq_pass_name gives the kernel name): acoustic row,
 * acoustic column, then the viscoacoustic passes C1 R1 LR1 C2 LR2 R2 C3 R3
 * (LR1 / LR2 replace R1 / C2 for low-rank jobs). */
#define CQ_NPASS 10
int cq_batch_enable_timing(cq_batch b, int on);
/* Average ms per launch of each pass (arrays of CQ_NPASS) and launch counts. */
int cq_batch_kernel_times(cq_batch b, double *ms, int *calls);
const char *cq_pass_name(int pass);
/* Acoustic row / column averages (passes 0 and 1). */
int cq_batch_pass_times(cq_batch b, double *row_ms, double *col_ms, int *count);

/* Measurement helper (no counterpart in the reference): times `reps` runs of
 * a streaming device copy of `bytes` bytes (read + write) to calibrate HBM
 * byte counters and achievable bandwidth. */
int cq_copy_calibrate(cq_ctx ctx, long long bytes, int reps, double *ms_per_copy);

/* The FFT plan the engine takes for an n-point transform of which one
 * workgroup holds `points` points per stage (lines x n), in plan variant 0 or
 * 1 (cq_tiles::plan), for inspection and tests; needs no GPU (no counterpart
 * in the reference, whose cuFFT plans are opaque).  Writes the stages'
 * radices, first to last, to radices[0 .. cap) and returns their count;
 * *inner is 0 for a direct transform, or, for a length with a prime factor
 * above 17, the length M of the Bluestein transform's inner transforms (the
 * radices are then M's).  CQ_EUNSUPPORTED when no plan holds `points`. */
int cq_fft_plan(int n, int points, int variant, int *radices, int cap, int *inner);

#ifdef __cplusplus
}
#endif
#endif
