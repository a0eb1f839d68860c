// cq_job.h -- device-side description of one wavefield of one shot (shared by
// host and kernels).  A modeling shot has one; an RTM shot has two (the
// source wavefield, role 0, and the receiver wavefield, role 1) that share the
// model coefficients and the per-step scratch.
#pragma once
#include "cq_fft.h"

// Layout of every real field: x-major, z fastest, leading dimension ldz
// (nzz rounded up to 32 floats = 128 B, padding rows kept at zero).
// Half spectra in kz (viscoacoustic path): float2 [nxx][ldk], kz in [0, nzh).
struct JobD {
    int kind, nxx, nzz, ldz, pml, nt, xsg, zsg, nr;
    int nzh, ldk;
    int role;                  // 0: source / modeling wavefield, 1: receiver wavefield (RTM)
    int nx, nz;                // interior size (image and illumination)
    int gather;                // record rec[t][i] (modeling)
    float qsign;               // +1 attenuating (modeling), -1 compensating (Q-RTM)
    int stab;                  // stabilised |k|^1.01 multiplier for aux (Q-RTM)
    float eta, taoc, dt;       // stabiliser (sgpstd-qrtm/acsglw.cu:1895-1924)
    float src_scale;           // dt * 1000 (acps4.cu:1491)
    float dkx, dkz, invN;      // wavenumber steps (funca.c:555-556) and 1/(nxx nzz)
    CqPlan px, pz;
    const float2 *twx, *twz;
    // Hermitian-symmetrised 1-D multipliers / n: v->p (m), p->v (p), d2 (2)
    const float2 *mxm, *mxp, *mx2, *mzm, *mzp, *mz2;
    // wavefield state
    float *p, *vx, *vz, *div, *lz, *dvx, *dvz, *tot;
    // per-cell coefficients
    const float *c1, *c2, *dtbx, *dtbz;                 // acoustic: rho v^2 dt, rho v^4 dt^3/12
    const float *qa, *qb, *qe, *qg, *qy2, *qcc;         // viscoacoustic (see build_job)
    // CPML memory variables, strips only
    float *psix, *phix, *psiz, *phiz;
    const float *iax, *ibx, *hax, *hbx, *iaz, *ibz, *haz, *hbz;
    // half-spectrum buffers (viscoacoustic): s2 is state (F_z p), s1/s3/s4 scratch
    float2 *s1, *s2, *s3, *s4;
    const float *rik;
    float *rec;
    const int *rcv_x, *rcv_z, *rcv_i, *rcv_off;          // gather, by x tile
    // RTM
    float *img, *ssg;                                    // nx*nz interior
    const float *psrc;                                   // source p (receiver role)
    const float *obs;                                    // nt*nr observed data
    const int *inj_x, *inj_z, *inj_i, *inj_off;          // injection cells, by row or x tile
};

// per-launch flags
#define CQ_F_SSG 1   // accumulate the illumination ssg += p^2 (first forward sweep)
