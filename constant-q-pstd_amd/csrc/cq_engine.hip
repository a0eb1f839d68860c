// cq_engine.hip -- MI355X (gfx950) engine for staggered-grid pseudo-spectral
// time stepping: acoustic modeling with CPML (the reference's acps4.cu path).
//
// Design (vs the reference's 9 full 2-D C2C FFTs and 10 elementwise kernels
// per step, sgpstd-qmodeling/acps4.cu:514-607):
//   * Every acoustic operator is separable: the staggered derivatives depend on
//     kx or kz alone and k^2 = kx^2 + kz^2, so the 2-D transforms reduce
//     exactly to 1-D transforms along x or z.  The reference keeps only the
//     real part of each complex result; for a real input that equals a 1-D
//     transform with the Hermitian-symmetrised multiplier
//     m_h(k) = (m(k) + conj(m(-k)))/2, which also reproduces the reference's
//     odd-n half-cell phase quirk (funca.c:394-395) exactly.
//   * With Hermitian multipliers two real lines ride in one complex FFT
//     (real and imaginary parts), halving the transform work.
//   * One time step is two fused passes:
//       row pass (x lines, z tiles):  L = Lx + Lz, p update + source,
//                                     dp/dx, x-CPML, vx update, dvx/dx + x-CPML
//       col pass (z lines, x tiles):  receiver gather, dp/dz, z-CPML, vz update,
//                                     dvz/dz + z-CPML, div, Lz = d2/dz2 div
//     reading 11 and writing 6 float words per cell (68 B/cell-step plus the
//     CPML strips), all FFT work staying in LDS.
//   * CPML memory variables are zero outside the absorbing strips (a = 0
//     there), so they are stored for the strips only.
//   * Many shots of different sizes advance in the same launches: the grid is a
//     list of (shot, tile) work items over a device table of shots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cq.h"
#include "cq_fft.h"

#define CQ_VERSION "cq-pstd-mi355x 0.1"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int set_err(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int set_err(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_err(CQ_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// device-side shot table
// ---------------------------------------------------------------------------
struct JobD {
    int kind, nxx, nzz, ldz, pml, nt, xsg, zsg, nr;
    float src_scale;  // dt * 1000 (acps4.cu:1491)
    CqPlan px, pz;
    const float2 *twx, *twz;
    const float2 *mxm, *mxp, *mx2, *mzm, *mzp, *mz2;
    float *p, *vx, *vz, *div, *lz, *dvx;
    const float *c1, *c2, *dtbx, *dtbz;
    float *psix, *phix, *psiz, *phiz;
    const float *iax, *ibx, *hax, *hbx, *iaz, *ibz, *haz, *hbz;
    const float *rik;
    float *rec;
    const int *rcv_x, *rcv_z, *rcv_i, *rcv_off;
};

__device__ __forceinline__ float4 ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ void st4(float *p, float4 v) { *reinterpret_cast<float4 *>(p) = v; }
__device__ __forceinline__ float f4get(const float4 &v, int k)
{
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}
__device__ __forceinline__ void f4set(float4 &v, int k, float s)
{
    if (k == 0) v.x = s;
    else if (k == 1) v.y = s;
    else if (k == 2) v.z = s;
    else v.w = s;
}

// ---------------------------------------------------------------------------
// acoustic row pass (x lines).  Work item = (shot, z0); the tile is TZ
// consecutive depths (TZ/2 packed complex lines of length nxx).
// ---------------------------------------------------------------------------
template <int TZ>
__global__ __launch_bounds__(CQ_NT) void k_row_ac(const JobD *__restrict__ jobs,
                                                  const int2 *__restrict__ items, int step)
{
    constexpr int L = TZ / 2;
    constexpr int Q = TZ / 4;
    const int2 it = items[blockIdx.x];
    const JobD &J = jobs[it.x];
    const int t = step;
    if (t >= J.nt) return;
    const int nxx = J.nxx, ldz = J.ldz, z0 = it.y, pml = J.pml;
    extern __shared__ float2 lds[];
    float2 *tw = lds;
    float2 *buf = lds + nxx;
    float4 *sav = reinterpret_cast<float4 *>(buf + L * nxx);
    for (int i = threadIdx.x; i < nxx; i += CQ_NT) tw[i] = J.twx[i];
    const int ne = nxx * Q;

    // phase 1: div -> LDS (kept for the p update)
    for (int e = threadIdx.x; e < ne; e += CQ_NT) {
        const int ix = e / Q, q = e % Q;
        const float4 d = ld4(J.div + (size_t)ix * ldz + z0 + 4 * q);
        sav[e] = d;
        buf[(2 * q) * nxx + ix] = make_float2(d.x, d.y);
        buf[(2 * q + 1) * nxx + ix] = make_float2(d.z, d.w);
    }
    __syncthreads();
    cq_fft_fwd(J.px, buf, nxx, L, tw, J.mx2);
    cq_fft_inv(J.px, buf, nxx, L, tw);

    // phase 2: p += c1 div - c2 (Lx + Lz); source (addsa_kernel, acps4.cu:1462-1498)
    const float src = J.rik[t] * J.src_scale;
    for (int e = threadIdx.x; e < ne; e += CQ_NT) {
        const int ix = e / Q, q = e % Q;
        const size_t g = (size_t)ix * ldz + z0 + 4 * q;
        const float2 a = buf[(2 * q) * nxx + ix], b = buf[(2 * q + 1) * nxx + ix];
        const float4 p = ld4(J.p + g), c1 = ld4(J.c1 + g), c2 = ld4(J.c2 + g),
                     lz = ld4(J.lz + g), d = sav[e];
        float4 pn;
        pn.x = p.x + c1.x * d.x - c2.x * (a.x + lz.x);
        pn.y = p.y + c1.y * d.y - c2.y * (a.y + lz.y);
        pn.z = p.z + c1.z * d.z - c2.z * (b.x + lz.z);
        pn.w = p.w + c1.w * d.w - c2.w * (b.y + lz.w);
        if (ix == J.xsg) {
            const int dz = J.zsg - (z0 + 4 * q);
            if (dz >= 0 && dz < 4) f4set(pn, dz, f4get(pn, dz) + src);
        }
        st4(J.p + g, pn);
        buf[(2 * q) * nxx + ix] = make_float2(pn.x, pn.y);
        buf[(2 * q + 1) * nxx + ix] = make_float2(pn.z, pn.w);
    }
    __syncthreads();
    cq_fft_fwd(J.px, buf, nxx, L, tw, J.mxp);
    cq_fft_inv(J.px, buf, nxx, L, tw);

    // phase 3: x-CPML on dp/dx and vx update (vx_cal_kernel, acps4.cu:1274-1303)
    for (int e = threadIdx.x; e < ne; e += CQ_NT) {
        const int ix = e / Q, q = e % Q;
        const size_t g = (size_t)ix * ldz + z0 + 4 * q;
        const float2 a = buf[(2 * q) * nxx + ix], b = buf[(2 * q + 1) * nxx + ix];
        float4 dp = make_float4(a.x, a.y, b.x, b.y);
        const float4 vx = ld4(J.vx + g), bx = ld4(J.dtbx + g);
        if (ix < pml || ix >= nxx - pml) {
            const int sx = ix < pml ? ix : ix - (nxx - 2 * pml);
            float *php = J.phix + (size_t)sx * ldz + z0 + 4 * q;
            const float hb = J.hbx[ix], ha = J.hax[ix];
            float4 ph = ld4(php);
            ph.x = hb * ph.x + ha * dp.x;
            ph.y = hb * ph.y + ha * dp.y;
            ph.z = hb * ph.z + ha * dp.z;
            ph.w = hb * ph.w + ha * dp.w;
            st4(php, ph);
            dp.x += ph.x; dp.y += ph.y; dp.z += ph.z; dp.w += ph.w;
        }
        float4 vn;
        vn.x = vx.x + bx.x * dp.x;
        vn.y = vx.y + bx.y * dp.y;
        vn.z = vx.z + bx.z * dp.z;
        vn.w = vx.w + bx.w * dp.w;
        st4(J.vx + g, vn);
        buf[(2 * q) * nxx + ix] = make_float2(vn.x, vn.y);
        buf[(2 * q + 1) * nxx + ix] = make_float2(vn.z, vn.w);
    }
    __syncthreads();
    cq_fft_fwd(J.px, buf, nxx, L, tw, J.mxm);
    cq_fft_inv(J.px, buf, nxx, L, tw);

    // phase 4: dvx/dx for the next step with its x-CPML (update_vxx_vzz, :1111-1144)
    for (int e = threadIdx.x; e < ne; e += CQ_NT) {
        const int ix = e / Q, q = e % Q;
        const size_t g = (size_t)ix * ldz + z0 + 4 * q;
        const float2 a = buf[(2 * q) * nxx + ix], b = buf[(2 * q + 1) * nxx + ix];
        float4 dv = make_float4(a.x, a.y, b.x, b.y);
        if (ix < pml || ix >= nxx - pml) {
            const int sx = ix < pml ? ix : ix - (nxx - 2 * pml);
            float *psp = J.psix + (size_t)sx * ldz + z0 + 4 * q;
            const float ib = J.ibx[ix], ia = J.iax[ix];
            float4 ps = ld4(psp);
            ps.x = ib * ps.x + ia * dv.x;
            ps.y = ib * ps.y + ia * dv.y;
            ps.z = ib * ps.z + ia * dv.z;
            ps.w = ib * ps.w + ia * dv.w;
            st4(psp, ps);
            dv.x += ps.x; dv.y += ps.y; dv.z += ps.z; dv.w += ps.w;
        }
        st4(J.dvx + g, dv);
    }
}

// ---------------------------------------------------------------------------
// acoustic column pass (z lines).  Work item = (shot, x0); CW columns packed
// in CW/2 complex lines of length nzz.
// ---------------------------------------------------------------------------
template <int CW>
__global__ __launch_bounds__(CQ_NT) void k_col_ac(const JobD *__restrict__ jobs,
                                                  const int2 *__restrict__ items, int step)
{
    constexpr int L = CW / 2;
    const int2 it = items[blockIdx.x];
    const JobD &J = jobs[it.x];
    const int t = step;
    if (t >= J.nt) return;
    const int nxx = J.nxx, nzz = J.nzz, ldz = J.ldz, x0 = it.y, pml = J.pml;
    const int nq = (nzz + 3) >> 2;
    extern __shared__ float2 lds[];
    float2 *tw = lds;
    float2 *buf = lds + nzz;
    float *bufs = reinterpret_cast<float *>(buf);
    for (int i = threadIdx.x; i < nzz; i += CQ_NT) tw[i] = J.twz[i];
    const int ne = CW * nq;

    // phase 1: p -> LDS
    for (int e = threadIdx.x; e < ne; e += CQ_NT) {
        const int c = e / nq, zq = e - c * nq, ix = x0 + c;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ix < nxx) v = ld4(J.p + (size_t)ix * ldz + 4 * zq);
        float *line = bufs + 2 * (c >> 1) * nzz + (c & 1);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (4 * zq + k < nzz) line[2 * (4 * zq + k)] = f4get(v, k);
    }
    __syncthreads();
    // receiver gather of the updated p (acps4.cu:611-621)
    {
        const int tile = x0 / CW;
        const int r0 = J.rcv_off[tile], r1 = J.rcv_off[tile + 1];
        for (int r = r0 + threadIdx.x; r < r1; r += CQ_NT) {
            const int c = J.rcv_x[r] - x0, z = J.rcv_z[r];
            J.rec[(size_t)t * J.nr + J.rcv_i[r]] = bufs[2 * ((c >> 1) * nzz + z) + (c & 1)];
        }
    }
    cq_fft_fwd(J.pz, buf, nzz, L, tw, J.mzp);
    cq_fft_inv(J.pz, buf, nzz, L, tw);

    // phase 2: z-CPML on dp/dz and vz update (vz_cal_kernel, acps4.cu:1305-1333)
    for (int e = threadIdx.x; e < ne; e += CQ_NT) {
        const int c = e / nq, zq = e - c * nq, ix = x0 + c;
        if (ix >= nxx) continue;
        const size_t g = (size_t)ix * ldz + 4 * zq;
        float *line = bufs + 2 * (c >> 1) * nzz + (c & 1);
        const float4 vz = ld4(J.vz + g), bz = ld4(J.dtbz + g);
        float4 vn = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int iz = 4 * zq + k;
            if (iz >= nzz) continue;
            float dp = line[2 * iz];
            if (iz < pml || iz >= nzz - pml) {
                const int sz = iz < pml ? iz : iz - (nzz - 2 * pml);
                float *php = J.phiz + (size_t)ix * (2 * pml) + sz;
                const float ph = J.hbz[iz] * (*php) + J.haz[iz] * dp;
                *php = ph;
                dp += ph;
            }
            const float v = f4get(vz, k) + f4get(bz, k) * dp;
            f4set(vn, k, v);
            line[2 * iz] = v;
        }
        st4(J.vz + g, vn);
    }
    __syncthreads();
    cq_fft_fwd(J.pz, buf, nzz, L, tw, J.mzm);
    cq_fft_inv(J.pz, buf, nzz, L, tw);

    // phase 3: dvz/dz with z-CPML; div = dvx + dvz (update_vxx_vzz + sumpxpz)
    for (int e = threadIdx.x; e < ne; e += CQ_NT) {
        const int c = e / nq, zq = e - c * nq, ix = x0 + c;
        float *line = bufs + 2 * (c >> 1) * nzz + (c & 1);
        if (ix >= nxx) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (4 * zq + k < nzz) line[2 * (4 * zq + k)] = 0.f;
            continue;
        }
        const size_t g = (size_t)ix * ldz + 4 * zq;
        const float4 dvx = ld4(J.dvx + g);
        float4 dv = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int iz = 4 * zq + k;
            if (iz >= nzz) continue;
            float d = line[2 * iz];
            if (iz < pml || iz >= nzz - pml) {
                const int sz = iz < pml ? iz : iz - (nzz - 2 * pml);
                float *psp = J.psiz + (size_t)ix * (2 * pml) + sz;
                const float ps = J.ibz[iz] * (*psp) + J.iaz[iz] * d;
                *psp = ps;
                d += ps;
            }
            const float s = f4get(dvx, k) + d;
            f4set(dv, k, s);
            line[2 * iz] = s;
        }
        st4(J.div + g, dv);
    }
    __syncthreads();
    cq_fft_fwd(J.pz, buf, nzz, L, tw, J.mz2);
    cq_fft_inv(J.pz, buf, nzz, L, tw);

    // phase 4: Lz = d2/dz2 div
    for (int e = threadIdx.x; e < ne; e += CQ_NT) {
        const int c = e / nq, zq = e - c * nq, ix = x0 + c;
        if (ix >= nxx) continue;
        const float *line = bufs + 2 * (c >> 1) * nzz + (c & 1);
        float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int iz = 4 * zq + k;
            if (iz < nzz) f4set(o, k, line[2 * iz]);
        }
        st4(J.lz + (size_t)ix * ldz + 4 * zq, o);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct cq_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<int, std::pair<CqPlan, float2 *>> plans;  // n -> plan, device twiddles
    ~cq_ctx_s()
    {
        for (auto &kv : plans) (void)hipFree(kv.second.second);
        if (stream) (void)hipStreamDestroy(stream);
    }
};


// Factorisation into supported radices: powers of two as 16/8/4/2, powers of
// three as 9/3, then 5, 7, 11, 13, 17.
static bool make_plan(int n, CqPlan &p)
{
    memset(&p, 0, sizeof p);
    p.n = n;
    int m = n;
    while (m % 16 == 0) { p.rad[p.ns++] = 16; m /= 16; }
    if (m % 8 == 0) { p.rad[p.ns++] = 8; m /= 8; }
    if (m % 4 == 0) { p.rad[p.ns++] = 4; m /= 4; }
    if (m % 2 == 0) { p.rad[p.ns++] = 2; m /= 2; }
    while (m % 9 == 0) { p.rad[p.ns++] = 9; m /= 9; }
    while (m % 3 == 0) { p.rad[p.ns++] = 3; m /= 3; }
    for (int r : {5, 7, 11, 13, 17})
        while (m % r == 0) {
            if (p.ns >= CQ_MAXST) return false;
            p.rad[p.ns++] = r;
            m /= r;
        }
    return m == 1 && p.ns <= CQ_MAXST;
}

static int plan_capacity(const CqPlan &p)
{
    int cap = 1 << 30;
    for (int s = 0; s < p.ns; s++) cap = std::min(cap, cq_stage_capacity(p.rad[s]));
    return cap;
}

static int get_plan(cq_ctx ctx, int n, CqPlan &p, const float2 *&tw)
{
    auto itp = ctx->plans.find(n);
    if (itp != ctx->plans.end()) {
        p = itp->second.first;
        tw = itp->second.second;
        return CQ_OK;
    }
    if (!make_plan(n, p))
        return set_err(CQ_EUNSUPPORTED,
                       "FFT length %d has a prime factor other than 2, 3, 5, 7, 11, 13, 17", n);
    std::vector<float2> h(n);
    for (int j = 0; j < n; j++) {
        const double a = -2.0 * M_PI * (double)j / (double)n;
        h[j] = make_float2((float)cos(a), (float)sin(a));
    }
    float2 *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(float2) * n));
    HIPCHK(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    ctx->plans[n] = {p, d};
    tw = d;
    return CQ_OK;
}

// Symmetrised staggered-derivative multiplier / n for one axis:
// m(k) = +-i (kr + i ki), sign + for k <= n/2 (shift_vxz_kernel, acps4.cu:1069-1105)
static void deriv_table(int n, const float *kr, const float *ki, std::vector<float2> &out)
{
    out.resize(n);
    auto m = [&](int k) {
        const double s = (k <= n / 2) ? 1.0 : -1.0;
        // s * i * (kr + i ki) = s * (-ki + i kr)
        return std::make_pair(-s * (double)ki[k], s * (double)kr[k]);
    };
    for (int k = 0; k < n; k++) {
        auto a = m(k);
        auto b = m((n - k) % n);
        const double re = 0.5 * (a.first + b.first) / n;
        const double im = 0.5 * (a.second - b.second) / n;
        out[k] = make_float2((float)re, (float)im);
    }
}

// |k|^2 / n with the reference's float wavenumber grid (funca.c:552-567)
static void lap_table(int n, float d, std::vector<float2> &out)
{
    out.resize(n);
    const float dk = 2 * 3.1415926 / (n * d);
    std::vector<float> kk(n);
    const int nh = n / 2 + 1;
    for (int i = 0; i < nh; i++) kk[i] = i * dk;
    for (int i = nh; i < n; i++) kk[i] = kk[n - i];
    for (int i = 0; i < n; i++) out[i] = make_float2((float)((double)(kk[i] * kk[i]) / n), 0.f);
}

struct Arena {
    size_t off = 0;
    template <typename T>
    size_t take(size_t count)
    {
        size_t o = (off + 255) & ~(size_t)255;
        off = o + count * sizeof(T);
        return o;
    }
};

struct JobH {
    JobD d;
    void *mem = nullptr;
    size_t bytes = 0;
    int nt = 0, nr = 0, nxx = 0, nzz = 0, ldz = 0;
};

struct cq_batch_s {
    cq_ctx ctx = nullptr;
    std::vector<JobH> jobs;
    JobD *djobs = nullptr;
    int2 *ditems_row = nullptr, *ditems_col = nullptr;
    int nrow = 0, ncol = 0;
    int tz = 16, cw = 16;
    size_t lds_row = 0, lds_col = 0;
    int t = 0, tmax = 0;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    ~cq_batch_s()
    {
        for (auto &j : jobs) (void)hipFree(j.mem);
        (void)hipFree(djobs);
        (void)hipFree(ditems_row);
        (void)hipFree(ditems_col);
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

extern "C" const char *cq_last_error(void) { return g_err.c_str(); }
extern "C" const char *cq_version(void) { return CQ_VERSION; }

extern "C" int cq_open(int device, cq_ctx *out)
{
    if (!out) return set_err(CQ_EINVAL, "cq_open: null output");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_err(CQ_EINVAL, "cq_open: device %d of %d", device, n);
    HIPCHK(hipSetDevice(device));
    auto *c = new cq_ctx_s();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return set_err(CQ_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = c;
    g_err.clear();
    return CQ_OK;
}

extern "C" int cq_close(cq_ctx ctx)
{
    if (!ctx) return CQ_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return CQ_OK;
}

static int validate(const cq_job_desc &j, int idx)
{
    if (j.kind != CQ_KIND_ACOUSTIC)
        return set_err(CQ_EUNSUPPORTED, "job %d: kind %d not supported by this build", idx, j.kind);
    if (j.nxx < 8 || j.nzz < 8 || j.pml < 2 || 2 * j.pml > j.nxx || 2 * j.pml > j.nzz)
        return set_err(CQ_EINVAL, "job %d: bad window %dx%d pml %d", idx, j.nxx, j.nzz, j.pml);
    if (j.nt < 1 || !(j.dt > 0.f)) return set_err(CQ_EINVAL, "job %d: bad nt/dt", idx);
    if (!j.vp || !j.rho || !j.bbx || !j.bbz || !j.rik || !j.iax || !j.ibx || !j.hax || !j.hbx ||
        !j.iaz || !j.ibz || !j.haz || !j.hbz || !j.kxsmr || !j.kxsmi || !j.kxscr || !j.kxsci ||
        !j.kzsmr || !j.kzsmi || !j.kzscr || !j.kzsci)
        return set_err(CQ_EINVAL, "job %d: missing input array", idx);
    if (j.xsg < 0 || j.xsg >= j.nxx || j.zsg < 0 || j.zsg >= j.nzz)
        return set_err(CQ_EINVAL, "job %d: source index outside the window", idx);
    if (j.nr < 0 || (j.nr > 0 && (!j.gxg || !j.gzg)))
        return set_err(CQ_EINVAL, "job %d: bad receivers", idx);
    for (int i = 0; i < j.nr; i++)
        if (j.gxg[i] < 0 || j.gxg[i] >= j.nxx || j.gzg[i] < 0 || j.gzg[i] >= j.nzz)
            return set_err(CQ_EINVAL, "job %d: receiver %d outside the window", idx, i);
    if (!(j.dx > 0.f) || !(j.dz > 0.f)) return set_err(CQ_EINVAL, "job %d: bad dx/dz", idx);
    return CQ_OK;
}

static int build_job(cq_batch b, const cq_job_desc &j, int cw, JobH &h)
{
    cq_ctx ctx = b->ctx;
    const int nxx = j.nxx, nzz = j.nzz, pml = j.pml;
    const int ldz = (nzz + 31) & ~31;
    const size_t N = (size_t)nxx * ldz;
    JobD &d = h.d;
    memset(&d, 0, sizeof d);
    d.kind = j.kind; d.nxx = nxx; d.nzz = nzz; d.ldz = ldz; d.pml = pml; d.nt = j.nt;
    d.xsg = j.xsg; d.zsg = j.zsg; d.nr = j.nr;
    d.src_scale = j.dt * 1000.0f;
    int rc;
    if ((rc = get_plan(ctx, nxx, d.px, d.twx)) != CQ_OK) return rc;
    if ((rc = get_plan(ctx, nzz, d.pz, d.twz)) != CQ_OK) return rc;

    const int ntile_x = (nxx + cw - 1) / cw;
    Arena A;
    const size_t o_p = A.take<float>(N), o_vx = A.take<float>(N), o_vz = A.take<float>(N);
    const size_t o_div = A.take<float>(N), o_lz = A.take<float>(N), o_dvx = A.take<float>(N);
    const size_t o_c1 = A.take<float>(N), o_c2 = A.take<float>(N);
    const size_t o_bx = A.take<float>(N), o_bz = A.take<float>(N);
    const size_t o_psix = A.take<float>((size_t)2 * pml * ldz), o_phix = A.take<float>((size_t)2 * pml * ldz);
    const size_t o_psiz = A.take<float>((size_t)2 * pml * nxx), o_phiz = A.take<float>((size_t)2 * pml * nxx);
    const size_t o_1x = A.take<float>((size_t)4 * nxx), o_1z = A.take<float>((size_t)4 * nzz);
    const size_t o_mx = A.take<float2>((size_t)3 * nxx), o_mz = A.take<float2>((size_t)3 * nzz);
    const size_t o_rik = A.take<float>(j.nt);
    const size_t o_rec = A.take<float>((size_t)j.nt * std::max(j.nr, 1));
    const size_t o_rcv = A.take<int>((size_t)3 * std::max(j.nr, 1));
    const size_t o_off = A.take<int>((size_t)ntile_x + 1);
    const size_t total = A.take<char>(0);
    char *mem = nullptr;
    hipError_t e = hipMalloc(&mem, total);
    if (e != hipSuccess) return set_err(CQ_ENOMEM, "hipMalloc(%zu): %s", total, hipGetErrorString(e));
    h.mem = mem;
    h.bytes = total;
    h.nt = j.nt; h.nr = j.nr; h.nxx = nxx; h.nzz = nzz; h.ldz = ldz;
    HIPCHK(hipMemsetAsync(mem, 0, total, ctx->stream));

    d.p = (float *)(mem + o_p); d.vx = (float *)(mem + o_vx); d.vz = (float *)(mem + o_vz);
    d.div = (float *)(mem + o_div); d.lz = (float *)(mem + o_lz); d.dvx = (float *)(mem + o_dvx);
    d.c1 = (float *)(mem + o_c1); d.c2 = (float *)(mem + o_c2);
    d.dtbx = (float *)(mem + o_bx); d.dtbz = (float *)(mem + o_bz);
    d.psix = (float *)(mem + o_psix); d.phix = (float *)(mem + o_phix);
    d.psiz = (float *)(mem + o_psiz); d.phiz = (float *)(mem + o_phiz);
    float *x1 = (float *)(mem + o_1x), *z1 = (float *)(mem + o_1z);
    d.iax = x1; d.ibx = x1 + nxx; d.hax = x1 + 2 * nxx; d.hbx = x1 + 3 * nxx;
    d.iaz = z1; d.ibz = z1 + nzz; d.haz = z1 + 2 * nzz; d.hbz = z1 + 3 * nzz;
    float2 *mx = (float2 *)(mem + o_mx), *mz = (float2 *)(mem + o_mz);
    d.mxm = mx; d.mxp = mx + nxx; d.mx2 = mx + 2 * nxx;
    d.mzm = mz; d.mzp = mz + nzz; d.mz2 = mz + 2 * nzz;
    d.rik = (float *)(mem + o_rik);
    d.rec = (float *)(mem + o_rec);
    int *rcv = (int *)(mem + o_rcv);
    d.rcv_x = rcv; d.rcv_z = rcv + j.nr; d.rcv_i = rcv + 2 * j.nr;
    d.rcv_off = (int *)(mem + o_off);

    // per-cell coefficients (ldz-strided, zero padding):
    //   c1 = rho v^2 dt, c2 = rho v^2 dt * v^2 dt^2 / 12 (addsa_kernel, acps4.cu:1488)
    std::vector<float> hc1(N, 0.f), hc2(N, 0.f), hbx(N, 0.f), hbz(N, 0.f);
    const float dt = j.dt;
    for (int ix = 0; ix < nxx; ix++)
        for (int iz = 0; iz < nzz; iz++) {
            const size_t a = (size_t)ix * nzz + iz, g = (size_t)ix * ldz + iz;
            const float v = j.vp[a];
            const float r = j.rho[a] * v * v * dt;
            hc1[g] = r;
            hc2[g] = (float)((double)r * (1.0 / 12.0) * v * dt * v * dt);
            hbx[g] = dt * j.bbx[a];
            hbz[g] = dt * j.bbz[a];
        }
    std::vector<float> h1x(4 * nxx), h1z(4 * nzz);
    memcpy(&h1x[0], j.iax, 4 * nxx); memcpy(&h1x[nxx], j.ibx, 4 * nxx);
    memcpy(&h1x[2 * nxx], j.hax, 4 * nxx); memcpy(&h1x[3 * nxx], j.hbx, 4 * nxx);
    memcpy(&h1z[0], j.iaz, 4 * nzz); memcpy(&h1z[nzz], j.ibz, 4 * nzz);
    memcpy(&h1z[2 * nzz], j.haz, 4 * nzz); memcpy(&h1z[3 * nzz], j.hbz, 4 * nzz);
    std::vector<float2> t1, t2, t3, hmx(3 * nxx), hmz(3 * nzz);
    deriv_table(nxx, j.kxsmr, j.kxsmi, t1);   // v -> p, x: "sm" (acps4.cu:526)
    deriv_table(nxx, j.kxscr, j.kxsci, t2);   // p -> v, x: "sc" (acps4.cu:594)
    lap_table(nxx, j.dx, t3);
    std::copy(t1.begin(), t1.end(), hmx.begin());
    std::copy(t2.begin(), t2.end(), hmx.begin() + nxx);
    std::copy(t3.begin(), t3.end(), hmx.begin() + 2 * nxx);
    deriv_table(nzz, j.kzscr, j.kzsci, t1);   // v -> p, z: "sc" (acps4.cu:527)
    deriv_table(nzz, j.kzsmr, j.kzsmi, t2);   // p -> v, z: "sm" (acps4.cu:595)
    lap_table(nzz, j.dz, t3);
    std::copy(t1.begin(), t1.end(), hmz.begin());
    std::copy(t2.begin(), t2.end(), hmz.begin() + nzz);
    std::copy(t3.begin(), t3.end(), hmz.begin() + 2 * nzz);
    // receivers sorted by column, with per-x-tile offsets
    std::vector<int> order(j.nr);
    for (int i = 0; i < j.nr; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return j.gxg[a] < j.gxg[c]; });
    std::vector<int> hr(3 * std::max(j.nr, 1), 0), hoff(ntile_x + 1, 0);
    for (int k = 0; k < j.nr; k++) {
        hr[k] = j.gxg[order[k]];
        hr[j.nr + k] = j.gzg[order[k]];
        hr[2 * j.nr + k] = order[k];
    }
    for (int tl = 0, k = 0; tl <= ntile_x; tl++) {
        while (k < j.nr && hr[k] < tl * cw) k++;
        hoff[tl] = k;
    }
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync((void *)d.c1, hc1.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void *)d.c2, hc2.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void *)d.dtbx, hbx.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void *)d.dtbz, hbz.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(x1, h1x.data(), 4 * h1x.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(z1, h1z.data(), 4 * h1z.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(mx, hmx.data(), 8 * hmx.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(mz, hmz.data(), 8 * hmz.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void *)d.rik, j.rik, 4 * (size_t)j.nt, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(rcv, hr.data(), 4 * hr.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void *)d.rcv_off, hoff.data(), 4 * hoff.size(), hipMemcpyHostToDevice, s));
    // the host vectors above must outlive the copies
    HIPCHK(hipStreamSynchronize(s));
    return CQ_OK;
}

extern "C" int cq_batch_create(cq_ctx ctx, const cq_job_desc *jobs, int njobs, cq_batch *out)
{
    if (!ctx || !jobs || njobs < 1 || !out) return set_err(CQ_EINVAL, "cq_batch_create: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    for (int i = 0; i < njobs; i++) {
        int rc = validate(jobs[i], i);
        if (rc) return rc;
    }
    std::unique_ptr<cq_batch_s> b(new cq_batch_s());
    b->ctx = ctx;
    // tile shapes: the largest tiles whose packed lines fit one workgroup's
    // per-stage capacity for every shot of the batch
    int tz = 16, cw = 32;
    for (int i = 0; i < njobs; i++) {
        CqPlan px, pz;
        if (!make_plan(jobs[i].nxx, px))
            return set_err(CQ_EUNSUPPORTED, "job %d: FFT length nxx=%d has an unsupported prime factor", i, jobs[i].nxx);
        if (!make_plan(jobs[i].nzz, pz))
            return set_err(CQ_EUNSUPPORTED, "job %d: FFT length nzz=%d has an unsupported prime factor", i, jobs[i].nzz);
        while (tz > 8 && (tz / 2) * jobs[i].nxx > plan_capacity(px)) tz /= 2;
        while (cw > 2 && (cw / 2) * jobs[i].nzz > plan_capacity(pz)) cw /= 2;
        if ((tz / 2) * jobs[i].nxx > plan_capacity(px) || (cw / 2) * jobs[i].nzz > plan_capacity(pz))
            return set_err(CQ_EUNSUPPORTED, "job %d: window %dx%d too large for one workgroup", i, jobs[i].nxx, jobs[i].nzz);
    }
    b->tz = tz;
    b->cw = cw;
    b->jobs.resize(njobs);
    int maxnxx = 0, maxnzz = 0;
    std::vector<int2> irow, icol;
    for (int i = 0; i < njobs; i++) {
        int rc = build_job(b.get(), jobs[i], cw, b->jobs[i]);
        if (rc) return rc;
        const JobH &h = b->jobs[i];
        maxnxx = std::max(maxnxx, h.nxx);
        maxnzz = std::max(maxnzz, h.nzz);
        b->tmax = std::max(b->tmax, h.nt);
        for (int z0 = 0; z0 < h.nzz; z0 += tz) irow.push_back(make_int2(i, z0));
        for (int x0 = 0; x0 < h.nxx; x0 += cw) icol.push_back(make_int2(i, x0));
    }
    b->lds_row = (size_t)maxnxx * 8 + (size_t)(tz / 2) * maxnxx * 8 + (size_t)tz * maxnxx * 4;
    b->lds_col = (size_t)maxnzz * 8 + (size_t)(cw / 2) * maxnzz * 8;
    if (b->lds_row > 160 * 1024 || b->lds_col > 160 * 1024)
        return set_err(CQ_EUNSUPPORTED, "window too large for LDS (row %zu B, col %zu B)", b->lds_row, b->lds_col);
    std::vector<JobD> hd(njobs);
    for (int i = 0; i < njobs; i++) hd[i] = b->jobs[i].d;
    HIPCHK(hipMalloc(&b->djobs, sizeof(JobD) * njobs));
    HIPCHK(hipMalloc(&b->ditems_row, sizeof(int2) * irow.size()));
    HIPCHK(hipMalloc(&b->ditems_col, sizeof(int2) * icol.size()));
    HIPCHK(hipMemcpy(b->djobs, hd.data(), sizeof(JobD) * njobs, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->ditems_row, irow.data(), sizeof(int2) * irow.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->ditems_col, icol.data(), sizeof(int2) * icol.size(), hipMemcpyHostToDevice));
    b->nrow = (int)irow.size();
    b->ncol = (int)icol.size();
    *out = b.release();
    return CQ_OK;
}

template <typename K>
static hipError_t set_lds(K kernel, size_t bytes)
{
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

static int launch_row(cq_batch b, int step)
{
    hipStream_t s = b->ctx->stream;
    if (b->tz == 16) {
        HIPCHK(set_lds(k_row_ac<16>, b->lds_row));
        hipLaunchKernelGGL(k_row_ac<16>, dim3(b->nrow), dim3(CQ_NT), b->lds_row, s, b->djobs, b->ditems_row, step);
    } else {
        HIPCHK(set_lds(k_row_ac<8>, b->lds_row));
        hipLaunchKernelGGL(k_row_ac<8>, dim3(b->nrow), dim3(CQ_NT), b->lds_row, s, b->djobs, b->ditems_row, step);
    }
    HIPCHK(hipGetLastError());
    return CQ_OK;
}

static int launch_col(cq_batch b, int step)
{
    hipStream_t s = b->ctx->stream;
    switch (b->cw) {
    case 32:
        HIPCHK(set_lds(k_col_ac<32>, b->lds_col));
        hipLaunchKernelGGL(k_col_ac<32>, dim3(b->ncol), dim3(CQ_NT), b->lds_col, s, b->djobs, b->ditems_col, step);
        break;
    case 16:
        HIPCHK(set_lds(k_col_ac<16>, b->lds_col));
        hipLaunchKernelGGL(k_col_ac<16>, dim3(b->ncol), dim3(CQ_NT), b->lds_col, s, b->djobs, b->ditems_col, step);
        break;
    case 8:
        HIPCHK(set_lds(k_col_ac<8>, b->lds_col));
        hipLaunchKernelGGL(k_col_ac<8>, dim3(b->ncol), dim3(CQ_NT), b->lds_col, s, b->djobs, b->ditems_col, step);
        break;
    case 4:
        HIPCHK(set_lds(k_col_ac<4>, b->lds_col));
        hipLaunchKernelGGL(k_col_ac<4>, dim3(b->ncol), dim3(CQ_NT), b->lds_col, s, b->djobs, b->ditems_col, step);
        break;
    default:
        HIPCHK(set_lds(k_col_ac<2>, b->lds_col));
        hipLaunchKernelGGL(k_col_ac<2>, dim3(b->ncol), dim3(CQ_NT), b->lds_col, s, b->djobs, b->ditems_col, step);
        break;
    }
    HIPCHK(hipGetLastError());
    return CQ_OK;
}

extern "C" int cq_batch_step(cq_batch b, int nsteps)
{
    if (!b || nsteps < 0) return set_err(CQ_EINVAL, "cq_batch_step: bad arguments");
    HIPCHK(hipSetDevice(b->ctx->device));
    for (int s = 0; s < nsteps && b->t < b->tmax; s++) {
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
        if (b->timing) {
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            HIPCHK(hipEventCreate(&e2));
            HIPCHK(hipEventRecord(e0, b->ctx->stream));
        }
        int rc = launch_row(b, b->t);
        if (rc) return rc;
        if (b->timing) HIPCHK(hipEventRecord(e1, b->ctx->stream));
        rc = launch_col(b, b->t);
        if (rc) return rc;
        if (b->timing) {
            HIPCHK(hipEventRecord(e2, b->ctx->stream));
            b->ev.push_back(e0);
            b->ev.push_back(e1);
            b->ev.push_back(e2);
        }
        b->t++;
    }
    return CQ_OK;
}

extern "C" int cq_batch_sync(cq_batch b)
{
    if (!b) return set_err(CQ_EINVAL, "cq_batch_sync: null batch");
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return CQ_OK;
}

extern "C" int cq_batch_run(cq_batch b)
{
    if (!b) return set_err(CQ_EINVAL, "cq_batch_run: null batch");
    int rc = cq_batch_step(b, b->tmax - b->t);
    if (rc) return rc;
    return cq_batch_sync(b);
}

extern "C" int cq_batch_read_rec(cq_batch b, int job, float *rec)
{
    if (!b || job < 0 || job >= (int)b->jobs.size() || !rec) return set_err(CQ_EINVAL, "cq_batch_read_rec: bad arguments");
    const JobH &h = b->jobs[job];
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    if (h.nr > 0) HIPCHK(hipMemcpy(rec, h.d.rec, sizeof(float) * (size_t)h.nt * h.nr, hipMemcpyDeviceToHost));
    return CQ_OK;
}

extern "C" int cq_batch_read_p(cq_batch b, int job, float *p)
{
    if (!b || job < 0 || job >= (int)b->jobs.size() || !p) return set_err(CQ_EINVAL, "cq_batch_read_p: bad arguments");
    const JobH &h = b->jobs[job];
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    HIPCHK(hipMemcpy2D(p, sizeof(float) * h.nzz, h.d.p, sizeof(float) * h.ldz, sizeof(float) * h.nzz, h.nxx,
                       hipMemcpyDeviceToHost));
    return CQ_OK;
}

extern "C" int cq_batch_time(cq_batch b) { return b ? b->t : -1; }

extern "C" long long cq_batch_bytes(cq_batch b)
{
    if (!b) return -1;
    long long s = 0;
    for (auto &j : b->jobs) s += (long long)j.bytes;
    return s;
}

extern "C" int cq_batch_destroy(cq_batch b)
{
    if (!b) return CQ_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return CQ_OK;
}

extern "C" int cq_batch_enable_timing(cq_batch b, int on)
{
    if (!b) return set_err(CQ_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    for (auto e : b->ev) (void)hipEventDestroy(e);
    b->ev.clear();
    b->timing = on != 0;
    return CQ_OK;
}

extern "C" int cq_batch_pass_times(cq_batch b, double *row_ms, double *col_ms, int *count)
{
    if (!b || !row_ms || !col_ms || !count) return set_err(CQ_EINVAL, "bad arguments");
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    double sr = 0, sc = 0;
    int n = 0;
    for (size_t i = 0; i + 2 < b->ev.size(); i += 3) {
        float a = 0, c = 0;
        HIPCHK(hipEventElapsedTime(&a, b->ev[i], b->ev[i + 1]));
        HIPCHK(hipEventElapsedTime(&c, b->ev[i + 1], b->ev[i + 2]));
        sr += a;
        sc += c;
        n++;
    }
    *row_ms = n ? sr / n : 0.0;
    *col_ms = n ? sc / n : 0.0;
    *count = n;
    return CQ_OK;
}

extern "C" int cq_plain_forward(cq_ctx ctx, const cq_job_desc *job, float *rec)
{
    if (!ctx || !job || !rec) return set_err(CQ_EINVAL, "cq_plain_forward: bad arguments");
    cq_batch b = nullptr;
    int rc = cq_batch_create(ctx, job, 1, &b);
    if (rc) return rc;
    const JobH &h = b->jobs[0];
    float *pinned = nullptr;
    if (h.nr > 0 && hipHostMalloc(&pinned, sizeof(float) * h.nr, hipHostMallocDefault) != hipSuccess) {
        cq_batch_destroy(b);
        return set_err(CQ_ENOMEM, "hipHostMalloc failed");
    }
    for (int t = 0; t < h.nt && rc == CQ_OK; t++) {
        rc = cq_batch_step(b, 1);
        if (rc == CQ_OK && h.nr > 0) {
            hipError_t e = hipMemcpyAsync(pinned, h.d.rec + (size_t)t * h.nr, sizeof(float) * h.nr,
                                          hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) rc = set_err(CQ_EDEVICE, "plain gather copy: %s", hipGetErrorString(e));
            else memcpy(rec + (size_t)t * h.nr, pinned, sizeof(float) * h.nr);
        }
    }
    if (pinned) (void)hipHostFree(pinned);
    cq_batch_destroy(b);
    return rc;
}
