// cq_engine.hip -- MI355X (gfx950) engine for staggered-grid pseudo-spectral
// time stepping: acoustic (acps4.cu) and fractional-Laplacian constant-Q
// (acsglw.cu) modeling with CPML, behind the C ABI of include/cq.h.
//
// Design (vs the reference's 9 (acoustic) / 14 (viscoacoustic) full 2-D C2C
// FFTs and 10-17 elementwise kernels per step, sgpstd-qmodeling/acps4.cu:514-607,
// acsglw.cu:517-645):
//   * The staggered derivatives depend on kx or kz alone and k^2 = kx^2 + kz^2,
//     so those 2-D transforms reduce exactly to 1-D transforms along x or z.
//     The reference keeps only the real part of each complex result; for a real
//     input that equals a 1-D transform with the Hermitian-symmetrised
//     multiplier m_h(k) = (m(k) + conj(m(-k)))/2, which also reproduces the
//     reference's odd-n half-cell phase quirk (funca.c:394-395) exactly.
//   * Hermitian multipliers let two real lines ride in one complex FFT.
//   * Acoustic step = 2 fused passes (row: p update, vx; column: vz, div, Lz),
//     68 B/cell-step of HBM traffic plus the CPML strips.
//   * Viscoacoustic step = 6 fused passes over real fields and kz half spectra
//     (see cq_passes.h), about 152 B/cell-step.
//   * CPML memory variables are zero outside the absorbing strips (a = 0
//     there), so they are stored for the strips only.
//   * Shots of different sizes and kinds advance in the same launches: the
//     grid is a list of (shot, tile) work items over a device table of shots,
//     ordered so neighbouring tiles run on the same XCD.
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
#include "cq_passes.h"

#define CQ_VERSION "cq-pstd-mi355x 0.2"

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

// Streaming copy used to calibrate HBM byte counters and bandwidth.
__global__ __launch_bounds__(256) void k_copy_f4(const float4 *__restrict__ a, float4 *__restrict__ b, long long n4)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) b[i] = a[i];
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
    while (m % 16 == 0 && p.ns < CQ_MAXST) { p.rad[p.ns++] = 16; m /= 16; }
    if (m % 8 == 0) { p.rad[p.ns++] = 8; m /= 8; }
    if (m % 4 == 0) { p.rad[p.ns++] = 4; m /= 4; }
    if (m % 2 == 0) { p.rad[p.ns++] = 2; m /= 2; }
    while (m % 9 == 0 && p.ns < CQ_MAXST) { p.rad[p.ns++] = 9; m /= 9; }
    while (m % 3 == 0 && p.ns < CQ_MAXST) { p.rad[p.ns++] = 3; m /= 3; }
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
static void deriv_table(int n, const float *kr, const float *ki, float2 *out)
{
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

// Wavenumber step of the reference's float grid (funca.c:363)
static float k_step(int n, float d) { return (float)(2 * 3.1415926 / (n * d)); }

// |k|^2 / n with the reference's float wavenumber grid (funca.c:552-567)
static void lap_table(int n, float d, float2 *out)
{
    const float dk = k_step(n, d);
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
    int nt = 0, nr = 0, nxx = 0, nzz = 0, ldz = 0, kind = 0;
};

// One launch of the per-step sequence.
enum PassId { P_AROW, P_ACOL, P_QC1, P_QR1, P_QC2, P_QR2, P_QC3, P_QR3, P_COUNT };
static const char *kPassName[P_COUNT] = {"k_row_ac", "k_col_ac", "k_q_c1", "k_q_r1",
                                         "k_q_c2",   "k_q_r2",   "k_q_c3", "k_q_r3"};

struct cq_batch_s {
    cq_ctx ctx = nullptr;
    std::vector<JobH> jobs;
    JobD *djobs = nullptr;
    int2 *ditems[P_COUNT] = {};
    int nitems[P_COUNT] = {};
    size_t lds[P_COUNT] = {};
    int tz = 16, cw = 16, tk = 8;
    int t = 0, tmax = 0;
    bool timing = false;
    std::vector<hipEvent_t> ev;        // per step: start + one per launch
    std::vector<int> ev_pass;          // pass id of each event after the start
    std::vector<int> step_nev;         // launches recorded per step
    ~cq_batch_s()
    {
        for (auto &j : jobs) (void)hipFree(j.mem);
        (void)hipFree(djobs);
        for (int i = 0; i < P_COUNT; i++) (void)hipFree(ditems[i]);
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
    if (j.kind != CQ_KIND_ACOUSTIC && j.kind != CQ_KIND_VISCO)
        return set_err(CQ_EUNSUPPORTED, "job %d: kind %d is not a modeling kind", idx, j.kind);
    if (j.nxx < 8 || j.nzz < 8 || j.pml < 2 || 2 * j.pml > j.nxx || 2 * j.pml > j.nzz)
        return set_err(CQ_EINVAL, "job %d: bad window %dx%d pml %d", idx, j.nxx, j.nzz, j.pml);
    if (j.nt < 1 || !(j.dt > 0.f)) return set_err(CQ_EINVAL, "job %d: bad nt/dt", idx);
    if (!j.vp || !j.rho || !j.bbx || !j.bbz || !j.rik || !j.iax || !j.ibx || !j.hax || !j.hbx ||
        !j.iaz || !j.ibz || !j.haz || !j.hbz || !j.kxsmr || !j.kxsmi || !j.kxscr || !j.kxsci ||
        !j.kzsmr || !j.kzsmi || !j.kzscr || !j.kzsci)
        return set_err(CQ_EINVAL, "job %d: missing input array", idx);
    if (j.kind == CQ_KIND_VISCO && (!j.gam || !j.lam || !j.vwd || !j.tao))
        return set_err(CQ_EINVAL, "job %d: viscoacoustic job without gam/lam/vwd/tao", idx);
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

static int build_job(cq_batch b, const cq_job_desc &j, JobH &h)
{
    cq_ctx ctx = b->ctx;
    const int cw = b->cw, tk = b->tk;
    const bool q = j.kind == CQ_KIND_VISCO;
    const int nxx = j.nxx, nzz = j.nzz, pml = j.pml;
    const int ldz = (nzz + 31) & ~31;
    const int nzh = nzz / 2 + 1;
    const int ldk = (nzh + tk - 1) / tk * tk;
    const size_t N = (size_t)nxx * ldz;
    JobD &d = h.d;
    memset(&d, 0, sizeof d);
    d.kind = j.kind; d.nxx = nxx; d.nzz = nzz; d.ldz = ldz; d.pml = pml; d.nt = j.nt;
    d.xsg = j.xsg; d.zsg = j.zsg; d.nr = j.nr; d.nzh = nzh; d.ldk = ldk;
    d.src_scale = j.dt * 1000.0f;
    d.dkx = k_step(nxx, j.dx);
    d.dkz = k_step(nzz, j.dz);
    d.invN = (float)(1.0 / ((double)nxx * nzz));
    int rc;
    if ((rc = get_plan(ctx, nxx, d.px, d.twx)) != CQ_OK) return rc;
    if ((rc = get_plan(ctx, nzz, d.pz, d.twz)) != CQ_OK) return rc;

    const int ntile_x = (nxx + cw - 1) / cw;
    Arena A;
    const size_t o_p = A.take<float>(N), o_vx = A.take<float>(N), o_vz = A.take<float>(N);
    const size_t o_div = A.take<float>(N), o_dvx = A.take<float>(N);
    const size_t o_bx = A.take<float>(N), o_bz = A.take<float>(N);
    // acoustic: lz, c1, c2 ; viscoacoustic: tot, qa, qb, qe, qg, qy2, qcc, 4 half spectra
    const size_t o_f1 = A.take<float>(N), o_f2 = A.take<float>(N), o_f3 = A.take<float>(N);
    size_t o_q[4] = {0, 0, 0, 0}, o_s[4] = {0, 0, 0, 0};
    if (q) {
        for (int i = 0; i < 4; i++) o_q[i] = A.take<float>(N);
        for (int i = 0; i < 4; i++) o_s[i] = A.take<float2>((size_t)nxx * ldk);
    }
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
    h.nt = j.nt; h.nr = j.nr; h.nxx = nxx; h.nzz = nzz; h.ldz = ldz; h.kind = j.kind;
    HIPCHK(hipMemsetAsync(mem, 0, total, ctx->stream));

    d.p = (float *)(mem + o_p); d.vx = (float *)(mem + o_vx); d.vz = (float *)(mem + o_vz);
    d.div = (float *)(mem + o_div); d.dvx = (float *)(mem + o_dvx);
    d.dtbx = (float *)(mem + o_bx); d.dtbz = (float *)(mem + o_bz);
    float *f1 = (float *)(mem + o_f1), *f2 = (float *)(mem + o_f2), *f3 = (float *)(mem + o_f3);
    if (!q) {
        d.lz = f1; d.c1 = f2; d.c2 = f3;
    } else {
        d.tot = f1; d.qa = f2; d.qb = f3;
        d.qe = (float *)(mem + o_q[0]); d.qg = (float *)(mem + o_q[1]);
        d.qy2 = (float *)(mem + o_q[2]); d.qcc = (float *)(mem + o_q[3]);
        d.s1 = (float2 *)(mem + o_s[0]); d.s2 = (float2 *)(mem + o_s[1]);
        d.s3 = (float2 *)(mem + o_s[2]); d.s4 = (float2 *)(mem + o_s[3]);
    }
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

    // per-cell coefficients (ldz-strided, zero padding)
    const float dt = j.dt;
    std::vector<float> hbx(N, 0.f), hbz(N, 0.f), hf1(N, 0.f), hf2(N, 0.f), hf3(N, 0.f);
    std::vector<float> hq[4];
    if (q)
        for (int i = 0; i < 4; i++) hq[i].assign(N, 0.f);
    for (int ix = 0; ix < nxx; ix++)
        for (int iz = 0; iz < nzz; iz++) {
            const size_t a = (size_t)ix * nzz + iz, g = (size_t)ix * ldz + iz;
            const float v = j.vp[a];
            hbx[g] = dt * j.bbx[a];
            hbz[g] = dt * j.bbz[a];
            if (!q) {
                // c1 = rho v^2 dt, c2 = rho v^2 dt * v^2 dt^2 / 12 (addsa_kernel, acps4.cu:1488)
                const float r = j.rho[a] * v * v * dt;
                hf2[g] = r;
                hf3[g] = (float)((double)r * (1.0 / 12.0) * v * dt * v * dt);
            } else {
                // acsglw.cu:1528-1529: p += (rho v^2 dt lam)(total - (dt tao/2) T) - (dt tao) aux
                const float gm = j.gam[a], lam = j.lam[a], tao = j.tao[a], vwd = j.vwd[a];
                hf2[g] = j.rho[a] * v * v * dt * lam;
                hf3[g] = (float)(0.5 * dt * tao);
                hq[0][g] = dt * tao;
                // sum_total (acsglw.cu:1232-1240): y2 = 32 gam vwd, c = dt^2 v^2 lam / 12
                hq[1][g] = gm;
                hq[2][g] = (float)(32.0 * gm * vwd);
                const float dtf = dt * dt, vpf = v * v;
                hq[3][g] = (float)(1.0 / 12.0 * dtf * vpf * lam);
            }
        }
    std::vector<float> h1x(4 * nxx), h1z(4 * nzz);
    memcpy(&h1x[0], j.iax, 4 * nxx); memcpy(&h1x[nxx], j.ibx, 4 * nxx);
    memcpy(&h1x[2 * nxx], j.hax, 4 * nxx); memcpy(&h1x[3 * nxx], j.hbx, 4 * nxx);
    memcpy(&h1z[0], j.iaz, 4 * nzz); memcpy(&h1z[nzz], j.ibz, 4 * nzz);
    memcpy(&h1z[2 * nzz], j.haz, 4 * nzz); memcpy(&h1z[3 * nzz], j.hbz, 4 * nzz);
    std::vector<float2> hmx(3 * nxx), hmz(3 * nzz);
    deriv_table(nxx, j.kxsmr, j.kxsmi, &hmx[0]);         // v -> p, x: "sm" (acps4.cu:526)
    deriv_table(nxx, j.kxscr, j.kxsci, &hmx[nxx]);       // p -> v, x: "sc" (acps4.cu:594)
    lap_table(nxx, j.dx, &hmx[2 * nxx]);
    deriv_table(nzz, j.kzscr, j.kzsci, &hmz[0]);         // v -> p, z: "sc" (acps4.cu:527)
    deriv_table(nzz, j.kzsmr, j.kzsmi, &hmz[nzz]);       // p -> v, z: "sm" (acps4.cu:595)
    lap_table(nzz, j.dz, &hmz[2 * nzz]);
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
    HIPCHK(hipMemcpyAsync((void *)d.dtbx, hbx.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void *)d.dtbz, hbz.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(f2, hf2.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(f3, hf3.data(), 4 * N, hipMemcpyHostToDevice, s));
    if (q)
        for (int i = 0; i < 4; i++)
            HIPCHK(hipMemcpyAsync(mem + o_q[i], hq[i].data(), 4 * N, hipMemcpyHostToDevice, s));
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

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8),
// each XCD with its own L2.  Re-order the work items so that XCD x runs the
// contiguous chunk [x*chunk, (x+1)*chunk) of the natural order: neighbouring
// tiles (which share 128-B cache lines along z) then run on the same XCD at
// about the same time.  Padding items carry job -1 and exit at once.
static const int kNumXcd = 8;

static std::vector<int2> xcd_order(const std::vector<int2> &items)
{
    const size_t n = items.size();
    const size_t chunk = (n + kNumXcd - 1) / kNumXcd;
    std::vector<int2> out(chunk * kNumXcd, make_int2(-1, 0));
    for (size_t b = 0; b < out.size(); b++) {
        const size_t src = (b % kNumXcd) * chunk + b / kNumXcd;
        if (src < n) out[b] = items[src];
    }
    return out;
}

static const size_t kLdsMax = 160 * 1024;

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
    // Tile shapes: the largest tiles whose lines fit one workgroup's per-stage
    // capacity and the LDS for every shot of the batch.
    int tz = 16, cw = 32, tk = 8;
    int maxnxx = 0, maxnzz = 0;
    for (int i = 0; i < njobs; i++) {
        const cq_job_desc &j = jobs[i];
        CqPlan px, pz;
        if (!make_plan(j.nxx, px))
            return set_err(CQ_EUNSUPPORTED, "job %d: FFT length nxx=%d has a prime factor other than 2, 3, 5, 7, 11, 13, 17", i, j.nxx);
        if (!make_plan(j.nzz, pz))
            return set_err(CQ_EUNSUPPORTED, "job %d: FFT length nzz=%d has a prime factor other than 2, 3, 5, 7, 11, 13, 17", i, j.nzz);
        const bool q = j.kind == CQ_KIND_VISCO;

        while (tz > 8 && ((tz / 2) * j.nxx > plan_capacity(px) || (size_t)(1 + tz / 2) * j.nxx * 8 > kLdsMax)) tz /= 2;
        while (cw > 2 && ((cw / 2) * j.nzz > plan_capacity(pz) ||
                          (size_t)(1 + cw / 2) * j.nzz * 8 + (q ? (size_t)cw * j.nzz * 4 : 0) > kLdsMax))
            cw /= 2;
        if (q)
            while (tk > 2 && (tk * j.nxx > plan_capacity(px) || (size_t)(1 + 2 * tk) * j.nxx * 8 > kLdsMax)) tk /= 2;
        if ((tz / 2) * j.nxx > plan_capacity(px) || (size_t)(1 + tz / 2) * j.nxx * 8 > kLdsMax ||
            (cw / 2) * j.nzz > plan_capacity(pz) || (q && (tk * j.nxx > plan_capacity(px) ||
                                                           (size_t)(1 + 2 * tk) * j.nxx * 8 > kLdsMax)))
            return set_err(CQ_EUNSUPPORTED, "job %d: window %dx%d too large for one workgroup", i, j.nxx, j.nzz);
        maxnxx = std::max(maxnxx, j.nxx);
        maxnzz = std::max(maxnzz, j.nzz);
    }
    b->tz = tz;
    b->cw = cw;
    b->tk = tk;
    b->jobs.resize(njobs);
    std::vector<int2> items[P_COUNT];
    for (int i = 0; i < njobs; i++) {
        int rc = build_job(b.get(), jobs[i], b->jobs[i]);
        if (rc) return rc;
        const JobH &h = b->jobs[i];
        b->tmax = std::max(b->tmax, h.nt);
        if (h.kind == CQ_KIND_ACOUSTIC) {
            for (int z0 = 0; z0 < h.nzz; z0 += tz) items[P_AROW].push_back(make_int2(i, z0));
            for (int x0 = 0; x0 < h.nxx; x0 += cw) items[P_ACOL].push_back(make_int2(i, x0));
        } else {
            for (int x0 = 0; x0 < h.nxx; x0 += cw) {
                items[P_QC1].push_back(make_int2(i, x0));
                items[P_QC2].push_back(make_int2(i, x0));
                items[P_QC3].push_back(make_int2(i, x0));
            }
            for (int k0 = 0; k0 < h.d.nzh; k0 += tk) {
                items[P_QR1].push_back(make_int2(i, k0));
                items[P_QR2].push_back(make_int2(i, k0));
            }
            for (int z0 = 0; z0 < h.nzz; z0 += tz) items[P_QR3].push_back(make_int2(i, z0));
        }
    }
    const size_t lrow = (size_t)(1 + tz / 2) * maxnxx * 8;
    const size_t lcol = (size_t)(1 + cw / 2) * maxnzz * 8;
    b->lds[P_AROW] = lrow;
    b->lds[P_QR3] = lrow;
    b->lds[P_ACOL] = lcol;
    b->lds[P_QC1] = lcol;
    b->lds[P_QC2] = lcol + (size_t)cw * maxnzz * 4;
    b->lds[P_QC3] = lcol + (size_t)cw * maxnzz * 4;
    b->lds[P_QR1] = (size_t)(1 + 2 * tk) * maxnxx * 8;
    b->lds[P_QR2] = (size_t)(1 + tk) * maxnxx * 8;
    for (int p = 0; p < P_COUNT; p++)
        if (!items[p].empty() && b->lds[p] > kLdsMax)
            return set_err(CQ_EUNSUPPORTED, "pass %s needs %zu B of LDS", kPassName[p], b->lds[p]);
    std::vector<JobD> hd(njobs);
    for (int i = 0; i < njobs; i++) hd[i] = b->jobs[i].d;
    HIPCHK(hipMalloc(&b->djobs, sizeof(JobD) * njobs));
    HIPCHK(hipMemcpy(b->djobs, hd.data(), sizeof(JobD) * njobs, hipMemcpyHostToDevice));
    for (int p = 0; p < P_COUNT; p++) {
        if (items[p].empty()) continue;
        std::vector<int2> o = xcd_order(items[p]);
        HIPCHK(hipMalloc(&b->ditems[p], sizeof(int2) * o.size()));
        HIPCHK(hipMemcpy(b->ditems[p], o.data(), sizeof(int2) * o.size(), hipMemcpyHostToDevice));
        b->nitems[p] = (int)o.size();
    }
    *out = b.release();
    return CQ_OK;
}

template <typename K>
static int launch(cq_batch b, K kernel, int pass, int step)
{
    const size_t lds = b->lds[pass];
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kernel, dim3(b->nitems[pass]), dim3(CQ_NT), lds, b->ctx->stream, b->djobs, b->ditems[pass],
                       step);
    HIPCHK(hipGetLastError());
    return CQ_OK;
}

#define CQ_DISPATCH_TZ(K, pass)                                                   \
    (b->tz == 16 ? launch(b, K<16>, pass, step) : launch(b, K<8>, pass, step))
#define CQ_DISPATCH_CW(K, pass)                                                   \
    (b->cw == 32   ? launch(b, K<32>, pass, step)                                 \
     : b->cw == 16 ? launch(b, K<16>, pass, step)                                 \
     : b->cw == 8  ? launch(b, K<8>, pass, step)                                  \
     : b->cw == 4  ? launch(b, K<4>, pass, step)                                  \
                   : launch(b, K<2>, pass, step))
#define CQ_DISPATCH_TK(K, pass)                                                   \
    (b->tk == 8 ? launch(b, K<8>, pass, step)                                     \
     : b->tk == 4 ? launch(b, K<4>, pass, step) : launch(b, K<2>, pass, step))

static int launch_pass(cq_batch b, int pass, int step)
{
    switch (pass) {
    case P_AROW: return CQ_DISPATCH_TZ(k_row_ac, pass);
    case P_ACOL: return CQ_DISPATCH_CW(k_col_ac, pass);
    case P_QC1: return CQ_DISPATCH_CW(k_q_c1, pass);
    case P_QR1: return CQ_DISPATCH_TK(k_q_r1, pass);
    case P_QC2: return CQ_DISPATCH_CW(k_q_c2, pass);
    case P_QR2: return CQ_DISPATCH_TK(k_q_r2, pass);
    case P_QC3: return CQ_DISPATCH_CW(k_q_c3, pass);
    case P_QR3: return CQ_DISPATCH_TZ(k_q_r3, pass);
    }
    return set_err(CQ_EINVAL, "unknown pass %d", pass);
}

extern "C" int cq_batch_step(cq_batch b, int nsteps)
{
    if (!b || nsteps < 0) return set_err(CQ_EINVAL, "cq_batch_step: bad arguments");
    HIPCHK(hipSetDevice(b->ctx->device));
    for (int s = 0; s < nsteps && b->t < b->tmax; s++) {
        int nev = 0;
        if (b->timing) {
            hipEvent_t e0;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventRecord(e0, b->ctx->stream));
            b->ev.push_back(e0);
            b->ev_pass.push_back(-1);
        }
        for (int p = 0; p < P_COUNT; p++) {
            if (b->nitems[p] == 0) continue;
            int rc = launch_pass(b, p, b->t);
            if (rc) return rc;
            if (b->timing) {
                hipEvent_t e1;
                HIPCHK(hipEventCreate(&e1));
                HIPCHK(hipEventRecord(e1, b->ctx->stream));
                b->ev.push_back(e1);
                b->ev_pass.push_back(p);
                nev++;
            }
        }
        if (b->timing) b->step_nev.push_back(nev);
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
    b->ev_pass.clear();
    b->step_nev.clear();
    b->timing = on != 0;
    return CQ_OK;
}

extern "C" int cq_batch_kernel_times(cq_batch b, double *ms, int *calls)
{
    if (!b || !ms || !calls) return set_err(CQ_EINVAL, "bad arguments");
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    for (int p = 0; p < P_COUNT; p++) {
        ms[p] = 0.0;
        calls[p] = 0;
    }
    for (size_t i = 1; i < b->ev.size(); i++) {
        const int p = b->ev_pass[i];
        if (p < 0) continue;
        float a = 0.f;
        HIPCHK(hipEventElapsedTime(&a, b->ev[i - 1], b->ev[i]));
        ms[p] += a;
        calls[p]++;
    }
    for (int p = 0; p < P_COUNT; p++)
        if (calls[p]) ms[p] /= calls[p];
    return CQ_OK;
}

extern "C" const char *cq_pass_name(int pass) { return (pass >= 0 && pass < P_COUNT) ? kPassName[pass] : ""; }

extern "C" int cq_batch_pass_times(cq_batch b, double *row_ms, double *col_ms, int *count)
{
    if (!b || !row_ms || !col_ms || !count) return set_err(CQ_EINVAL, "bad arguments");
    double ms[P_COUNT];
    int calls[P_COUNT];
    int rc = cq_batch_kernel_times(b, ms, calls);
    if (rc) return rc;
    *row_ms = ms[P_AROW];
    *col_ms = ms[P_ACOL];
    *count = calls[P_AROW];
    return CQ_OK;
}

extern "C" int cq_copy_calibrate(cq_ctx ctx, long long bytes, int reps, double *ms_per_copy)
{
    if (!ctx || bytes < 1024 || reps < 1 || !ms_per_copy) return set_err(CQ_EINVAL, "cq_copy_calibrate: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    const long long n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n4 * 16));
    if (hipMalloc(&b, n4 * 16) != hipSuccess) {
        (void)hipFree(a);
        return set_err(CQ_ENOMEM, "cq_copy_calibrate: hipMalloc failed");
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemsetAsync(a, 0, n4 * 16, ctx->stream));
    hipLaunchKernelGGL(k_copy_f4, dim3(4096), dim3(256), 0, ctx->stream, a, b, n4);
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_f4, dim3(4096), dim3(256), 0, ctx->stream, a, b, n4);
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_copy = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
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
