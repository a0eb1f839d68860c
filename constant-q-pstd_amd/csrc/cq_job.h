// cq_job.h -- device-side description of one shot (shared by host and kernels).
#pragma once
#include "cq_fft.h"

// Layout of every real field: x-major, z fastest, leading dimension ldz
// (nzz rounded up to 32 floats = 128 B, padding rows kept at zero).
// Half spectra in kz (viscoacoustic path): float2 [nxx][ldk], kz in [0, nzh).
struct JobD {
    int kind, nxx, nzz, ldz, pml, nt, xsg, zsg, nr;
    int nzh, ldk;
    float src_scale;           // dt * 1000 (acps4.cu:1491)
    float dkx, dkz, invN;      // wavenumber steps (funca.c:555-556) and 1/(nxx nzz)
    CqPlan px, pz;
    const float2 *twx, *twz;
    // Hermitian-symmetrised 1-D multipliers / n: v->p (m), p->v (p), d2 (2)
    const float2 *mxm, *mxp, *mx2, *mzm, *mzp, *mz2;
    // wavefield state
    float *p, *vx, *vz, *div, *lz, *dvx, *tot;
    // per-cell coefficients
    const float *c1, *c2, *dtbx, *dtbz;                 // acoustic: rho v^2 dt, rho v^4 dt^3/12
    const float *qa, *qb, *qe, *qg, *qy2, *qcc;         // viscoacoustic (see build_job)
    // CPML memory variables, strips only
    float *psix, *phix, *psiz, *phiz;
    const float *iax, *ibx, *hax, *hbx, *iaz, *ibz, *haz, *hbz;
    // half-spectrum work buffers (viscoacoustic)
    float2 *s1, *s2, *s3, *s4;
    const float *rik;
    float *rec;
    const int *rcv_x, *rcv_z, *rcv_i, *rcv_off;
};
