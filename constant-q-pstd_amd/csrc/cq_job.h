// cq_job.h -- device-side description of one wavefield of one shot (shared by
// host and kernels).  A modeling shot has one; an RTM shot has two (the
// source wavefield, role 0, and the receiver wavefield, role 1) that share the
// model coefficients and the per-step scratch.
#pragma once
#include "cq_fft.h"

// Layout of every real field: depth quads major.  Cell (ix, iz) is float
// (iz & 3) of the 16-byte quad (ix, iz >> 2) at cq_q4(nxx, ix, iz >> 2):
// the quads of one depth band are contiguous along x (a row tile reads one
// contiguous run per band) and a column tile reads CW consecutive quads per
// band.  ldz (nzz rounded up to whole row tiles and 32) depths are stored;
// the padding depths are never read as data.
__host__ __device__ inline size_t cq_q4(int nxx, int ix, int zq) { return ((size_t)zq * nxx + ix) * 4; }
// Half spectra in kz (viscoacoustic path): float2, kz in [0, nzh), columns in
// blocks of 8: (ix, kz) at cq_hs(ldk, ix, kz).  A kz-row tile reads one
// contiguous run of TK x 8 values per block; a column tile reads, per kz,
// one 64-byte run per block it covers.  nxx is rounded up to whole blocks
// (cq_hs_cols); the extra columns are never read as data.
__host__ __device__ inline size_t cq_hs(int ldk, int ix, int kz) { return ((size_t)(ix >> 3) * ldk + kz) * 8 + (ix & 7); }
__host__ __device__ inline int cq_hs_cols(int nxx) { return (nxx + 7) & ~7; }
//
// The fields of a job record.  G qualifies the device pointers: empty in the
// host's JobD, the global address space in the kernels' JobG view of the same
// bytes, so that kernel accesses compile to global rather than flat
// instructions.
#define CQ_JOB_FIELDS(G)                                                                          \
    int kind, nxx, nzz, ldz, pml, nt, xsg, zsg, nr;                                               \
    int nzh, ldk;                                                                                 \
    int role;          /* 0: source / modeling wavefield, 1: receiver wavefield (RTM) */          \
    int nx, nz;        /* interior size (image and illumination) */                               \
    int gather;        /* record rec[t][i] (modeling) */                                          \
    float qsign;       /* +1 attenuating (modeling), -1 compensating (Q-RTM) */                   \
    int stab;          /* stabilised |k|^1.01 multiplier for aux (Q-RTM) */                       \
    int r1_fused;      /* k_q_r1 holds 3 tk lines: symbols inverted two at a time (else one) */   \
    float eta, taoc, dt; /* stabiliser (sgpstd-qrtm/acsglw.cu:1895-1924) */                       \
    float src_scale;   /* dt * 1000 (acps4.cu:1491) */                                            \
    float dkx, dkz, invN; /* wavenumber steps (funca.c:555-556) and 1/(nxx nzz) */                \
    /* steps: the wavefield takes part in global steps t0 .. t0 + nsched - 1; sched[s] = */       \
    /* {role, t, t_aux, flags} of its step s (RTM), or none: t = s, t_aux = s - 1 (modeling) */   \
    int t0, nsched;                                                                               \
    const int4 G *sched;                                                                          \
    /* the launch group's global step counter, read when a pass is launched with step < 0 */      \
    /* (graph-captured steps) */                                                                  \
    const int G *dstep;                                                                           \
    CqPlan px, pz;                                                                                \
    const float2 G *twx;                                                                          \
    const float2 G *twz;                                                                          \
    /* LDS geometry of the transforms: twiddle entries (the transform length, M for a */          \
    /* Bluestein plan) and line strides along x and z */                                          \
    int twlx, twlz, lsx, lsz;                                                                     \
    /* Hermitian-symmetrised 1-D multipliers / n: v->p (m), p->v (p), d2 (2) */                   \
    const float2 G *mxm; const float2 G *mxp; const float2 G *mx2;                                \
    const float2 G *mzm; const float2 G *mzp; const float2 G *mz2;                                \
    /* wavefield state */                                                                         \
    float G *p; float G *vx; float G *vz; float G *div;                                           \
    float G *lz; float G *dvx; float G *dvz; float G *tot;                                        \
    float G *scr;      /* per-cell scratch of the viscoacoustic column pass C3 */                 \
    /* per-cell coefficients; acoustic: rho v^2 dt, rho v^4 dt^3/12 */                            \
    const float G *c1; const float G *c2; const float G *dtbx; const float G *dtbz;               \
    /* viscoacoustic (see build_job) */                                                           \
    const float G *qa; const float G *qb;                                                         \
    const float G *qg; const float G *qy2; const float G *qcc;                                    \
    /* CPML memory variables, strips only */                                                      \
    float G *psix; float G *phix; float G *psiz; float G *phiz;                                   \
    const float G *iax; const float G *ibx; const float G *hax; const float G *hbx;               \
    const float G *iaz; const float G *ibz; const float G *haz; const float G *hbz;               \
    /* half-spectrum buffers (viscoacoustic): s2 is state (F_z p), s1/s3/s4 scratch */            \
    float2 G *s1; float2 G *s2; float2 G *s3; float2 G *s4;                                       \
    const float G *rik;                                                                           \
    float G *rec;                                                                                 \
    /* gather, by x tile */                                                                       \
    const int G *rcv_x; const int G *rcv_z; const int G *rcv_i; const int G *rcv_off;             \
    /* RTM: image and illumination over the interior (nx*nz), source p (receiver role), */        \
    /* nt*nr observed data, injection cells by row or x tile */                                   \
    float G *img; float G *ssg;                                                                   \
    const float G *psrc;                                                                          \
    const float G *obs;                                                                           \
    const int G *inj_x; const int G *inj_z; const int G *inj_i; const int G *inj_off;             \
    /* rank-r low-rank operator (lr_rank > 0): weights lrw[i] (field layout, nxx ldz each), */    \
    /* symbols / N lrv[i] on the quarter plane (nxh x nzh, kz fastest), and the rank's    */    \
    /* half spectra lrs[i] (half-spectrum layout, cq_hs_cols(nxx) ldk each)                */    \
    int lr_rank, nxh;                                                                             \
    const float G *lrw; const float G *lrv; float2 G *lrs;

struct JobD {
    CQ_JOB_FIELDS()
};

struct JobG {
    CQ_JOB_FIELDS(CQ_GLOBAL)
};
static_assert(sizeof(JobD) == sizeof(JobG), "JobD and JobG share one layout");

// per-launch flags
#define CQ_F_SSG 1   // accumulate the illumination ssg += p^2 (first forward sweep)
