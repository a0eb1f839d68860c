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
//     ordered so neighbouring tiles run on the same XCD.  The shots are dealt
//     into four launch groups (two per kind) with their own streams, so the
//     kernels of different groups overlap on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <set>
#include <tuple>
#include <string>
#include <vector>

#include "cq.h"
#include "cq_passes.h"
#include "cq_revolve.h"

#define CQ_VERSION "cq-pstd-mi355x 0.3"

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

// Streaming copy used to calibrate HBM byte counters and to measure the
// bandwidth a streaming kernel attains: one block per 16 KB, each lane's
// four 16-byte loads issued before its stores, nontemporal both ways (as the
// passes stream their fields).
typedef float cq_cv4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_f4(const float4 *__restrict__ a, float4 *__restrict__ b, long long n4)
{
    const long long base = (long long)blockIdx.x * 1024 + threadIdx.x;
    const cq_cv4 *src = reinterpret_cast<const cq_cv4 *>(a);
    cq_cv4 *dst = reinterpret_cast<cq_cv4 *>(b);
    cq_cv4 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const long long i = base + u * 256;
        if (i < n4) v[u] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const long long i = base + u * 256;
        if (i < n4) __builtin_nontemporal_store(v[u], dst + i);
    }
}

// ---------------------------------------------------------------------------
// static scratch of the allocation-free entry point (cq_forward_static)
// ---------------------------------------------------------------------------
// Every GPU the library runs on holds one copy of this array until the
// process ends.  Host code never passes `scratch` itself to a kernel or a
// copy (in host code the name is a host-side stand-in, not a device
// address): cq_open looks up each device's own address once with
// hipGetSymbolAddress and the entry point uses that.
static __device__ __align__(256) char scratch[CQ_STATIC_BYTES];

// Device address of `scratch` on the current device (looked up once per
// device), or null with the error set.
static char *scratch_address(int device)
{
    static std::mutex mu;
    static std::map<int, char *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) return nullptr;
    addr[device] = static_cast<char *>(ptr);
    return static_cast<char *>(ptr);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct cq_ctx_s {
    int device = 0;
    char *scratch = nullptr;   // this device's address of `scratch`
    hipStream_t stream = nullptr;
    std::map<std::vector<int>, float2 *> twiddles;  // radix sequence -> device twiddle table
    // (n, M) -> device chirp and spectrum tables of an n-point Bluestein transform
    std::map<std::pair<int, int>, std::pair<float2 *, float2 *>> bluestein;
    // open batches and servers; cq_close destroys what is left of them
    std::set<cq_batch_s *> batches;
    std::set<cq_server_s *> servers;
    ~cq_ctx_s()
    {
        for (auto &kv : twiddles) (void)hipFree(kv.second);
        for (auto &kv : bluestein) {
            (void)hipFree(kv.second.first);
            (void)hipFree(kv.second.second);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Fewest odd radices (each of per-stage capacity >= need) whose product is
// o, ties broken by the smaller radix sum (fewer registers per butterfly);
// radices are tried largest first, so `cur` stays non-increasing.
static void odd_split(int o, int need, int from, std::vector<int> &cur, std::vector<int> &best)
{
    static const int kOdd[] = {27, 25, 21, 17, 15, 13, 11, 9, 7, 5, 3};
    auto sum = [](const std::vector<int> &v) {
        int t = 0;
        for (int r : v) t += r;
        return t;
    };
    if (o == 1) {
        if (best.empty() || cur.size() < best.size() || (cur.size() == best.size() && sum(cur) < sum(best)))
            best = cur;
        return;
    }
    if (!best.empty() && cur.size() >= best.size()) return;   // needs one more radix: no better
    for (int i = from; i < (int)(sizeof kOdd / sizeof kOdd[0]); i++) {
        const int r = kOdd[i];
        if (o % r != 0 || cq_stage_capacity(r) < need) continue;
        cur.push_back(r);
        odd_split(o / r, need, i, cur, best);
        cur.pop_back();
    }
}

// Factorisation of n into supported radices for a workgroup that transforms
// `need` points at once.  Variant 0: the fewest stages: powers of two as
// 16/8/4/2, the odd part in the fewest odd radices (odd_split: 3 ... 17 and
// the composites 15, 21, 25, 27).  Order: a length with a pad period
// (cq_pad_period) starts with its largest power-of-two radix (the padded
// first-stage layout), then come the composites (largest first), the other
// powers of two, and the odd primes and 9 (smallest first), so that the last
// stage, which may carry a multiplier or a compressed twiddle table, is never
// a composite (those are compiled without either); a length whose odd radices
// are all composites and that has no power of two to end on takes the plain
// greedy factorisation below.  Variant 1: radices of at most 8 (powers of two
// as 8/4/2, powers of three as 3), more stages of fewer registers each.
// Only radices whose per-stage capacity (cq_stage_capacity) holds `need`
// points are taken.  False if n has another prime factor or a prime factor's
// radix is too small.
static bool make_direct_plan(int n, int need, CqPlan &p, int variant = 0)
{
    memset(&p, 0, sizeof p);
    p.n = n;
    int m = n;
    auto push = [&](int r) {
        if (p.ns >= CQ_MAXST) return false;
        p.rad[p.ns++] = r;
        m /= r;
        return true;
    };
    if (variant != 1) {
        std::vector<int> pow2;
        for (int r : {16, 8, 4, 2}) {
            if (cq_stage_capacity(r) < need) continue;
            while (m % r == 0) {
                pow2.push_back(r);
                m /= r;
            }
        }
        std::vector<int> cur, odd;
        odd_split(m, need, 0, cur, odd);
        if (m != 1 && odd.empty()) return false;
        m = n;
        auto composite = [](int r) { return r == 15 || r == 21 || r == 25 || r == 27; };
        size_t i2 = 0;
        if (cq_pad_period(n) && !pow2.empty() && pow2[0] == cq_pad_period(n) && !push(pow2[i2++])) return false;
        for (int r : odd)
            if (composite(r) && !push(r)) return false;
        for (; i2 < pow2.size(); i2++)
            if (!push(pow2[i2])) return false;
        for (auto it = odd.rbegin(); it != odd.rend(); ++it)
            if (!composite(*it) && !push(*it)) return false;
        if (m == 1 && (p.ns == 0 || !composite(p.rad[p.ns - 1]))) return true;
        memset(&p, 0, sizeof p);
        p.n = n;
        m = n;
    }
    static const int kOrder[2][11] = {{16, 8, 4, 2, 9, 3, 5, 7, 11, 13, 17}, {8, 4, 2, 3, 5, 7, 11, 13, 17, 0, 0}};
    for (int r : kOrder[variant == 1 ? 1 : 0]) {
        if (r == 0) break;
        if (cq_stage_capacity(r) < need) continue;
        while (m % r == 0)
            if (!push(r)) return false;
    }
    return m == 1;
}

// True if every prime factor of n is in `primes`.
static bool smooth(int n, std::initializer_list<int> primes)
{
    if (n < 1) return false;
    for (int q : primes)
        while (n % q == 0) n /= q;
    return n == 1;
}

// Length of the inner transforms of a Bluestein transform of n points: the
// smallest 7-smooth M >= 2 n - 1 (the cyclic convolution then equals the
// linear one of two n-point sequences).
static int bluestein_len(int n)
{
    for (int m = 2 * n - 1;; m++)
        if (smooth(m, {2, 3, 5, 7})) return m;
}

// Points per line a transform of n points occupies: n when its prime factors
// are radices of the stages (2, 3, 5, 7, 11, 13, 17), else the Bluestein
// length.
static int fft_len(int n) { return smooth(n, {2, 3, 5, 7, 11, 13, 17}) ? n : bluestein_len(n); }

// LDS budget per workgroup that the tile choice aims for: 80 KB lets two
// workgroups share a CU.
#ifndef CQ_LDS_TILE
#define CQ_LDS_TILE (80 * 1024)
#endif

// Plan of an n-point transform for `need` points (need / n lines) per
// workgroup: direct, or Bluestein over fft_len(n) points per line.
// Transforms longer than kTcMin keep a compressed last-stage twiddle table
// (CqPlan::tc): for them the full table would take LDS from the lines.  So
// does a shorter one whose lines and full table would exceed the tile budget
// CQ_LDS_TILE while its lines and compressed table fit (1080 points, 8 lines:
// 86.4 KB with the full table of 1080 entries, 79.7 KB with 239).
static const int kTcMin = 2048;

static bool make_plan(int n, int need, CqPlan &p, int variant = 0)
{
    if (n < 1) return false;
    const int M = fft_len(n);
    const int lines = (need + n - 1) / n;
    if (M == n) {
        if (!make_direct_plan(n, need, p, variant)) return false;
    } else {
        if (!make_direct_plan(M, lines * M, p, variant)) return false;
        p.bn = n;
    }
    p.tc = p.n > kTcMin && p.ns > 1 ? 1 : 0;
    if (!p.tc && p.ns > 1) {
        const size_t ls = (size_t)lines * cq_line_stride(p.n);
        const size_t full = (p.n + ls) * 8, comp = (2 * (p.n / p.rad[p.ns - 1]) - 1 + ls) * 8;
        if (full > (size_t)CQ_LDS_TILE && comp <= (size_t)CQ_LDS_TILE) p.tc = 1;
    }
    // the device compiles composite stages only as non-last stages (no
    // multiplier, no compressed table); make_direct_plan never ends on one
    const int last = p.ns > 0 ? p.rad[p.ns - 1] : 0;
    return !(last == 15 || last == 21 || last == 25 || last == 27);
}

// Entries of a plan's twiddle table: n (n - 1 used), or 2 n / R - 1 with a
// compressed last stage of radix R (its Ns = n / R first powers after the
// n / R - 1 entries of the earlier stages).
static int tw_len(const CqPlan &p)
{
    return p.tc ? 2 * (p.n / p.rad[p.ns - 1]) - 1 : p.n;
}

static bool has_plan(int n, int need, int variant = 0)
{
    CqPlan p;
    return make_plan(n, need, p, variant);
}

// In-place FFT of a 7-smooth length in double precision (host), by
// decimation in time over the smallest prime factor; sign -1: exp(-i).
static void host_fft(std::vector<std::complex<double>> &x, int sign)
{
    const int n = (int)x.size();
    if (n == 1) return;
    int p = 2;
    while (n % p) p++;
    const int m = n / p;
    std::vector<std::vector<std::complex<double>>> sub(p, std::vector<std::complex<double>>(m));
    for (int r = 0; r < p; r++)
        for (int j = 0; j < m; j++) sub[r][j] = x[(size_t)j * p + r];
    for (int r = 0; r < p; r++) host_fft(sub[r], sign);
    for (int k = 0; k < n; k++) {
        std::complex<double> acc = 0;
        for (int r = 0; r < p; r++)
            acc += sub[r][k % m] * std::polar(1.0, sign * 2.0 * M_PI * (double)((long long)r * k % n) / n);
        x[k] = acc;
    }
}

// Bluestein tables of an n-point transform with M-point inner transforms:
// chirp c[j] = exp(-i pi j^2 / n) (j^2 reduced mod 2n in integers), and
// B = FFT_M(b) / M of b[m] = conj c[|m|] for |m| < n, wrapped cyclically.
static void bluestein_tables(int n, int M, std::vector<float2> &c, std::vector<float2> &B)
{
    std::vector<std::complex<double>> cd(n), b(M, 0.0);
    c.resize(n);
    for (int j = 0; j < n; j++) {
        const long long r = (long long)j * j % (2LL * n);
        cd[j] = std::polar(1.0, -M_PI * (double)r / n);
        c[j] = make_float2((float)cd[j].real(), (float)cd[j].imag());
    }
    b[0] = std::conj(cd[0]);
    for (int m = 1; m < n; m++) b[m] = b[M - m] = std::conj(cd[m]);
    host_fft(b, -1);
    B.resize(M);
    for (int k = 0; k < M; k++) B[k] = make_float2((float)(b[k].real() / M), (float)(b[k].imag() / M));
}

static int pool_take(cq_batch_s *b, size_t bytes, void **out);

// Uploads `count` float2 host values: into the pool of a pooled batch, else
// into a new device buffer that the context caches.
static int upload_table(cq_ctx ctx, cq_batch_s *pb, const std::vector<float2> &h, float2 **out)
{
    const size_t bytes = sizeof(float2) * h.size();
    if (pb) {
        const int rc = pool_take(pb, bytes, (void **)out);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(*out, h.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return CQ_OK;
    }
    HIPCHK(hipMalloc(out, bytes));
    HIPCHK(hipMemcpy(*out, h.data(), bytes, hipMemcpyHostToDevice));
    return CQ_OK;
}

// Plan and twiddle table of an n-point transform for `need` points per
// workgroup (and a Bluestein plan's chirp tables).  The tables are cached in
// the context, or, for a pooled batch, placed in its pool.
static int get_plan(cq_ctx ctx, cq_batch_s *pb, int n, int need, int variant, CqPlan &p, const float2 *&tw)
{
    if (!make_plan(n, need, p, variant))
        return set_err(CQ_EUNSUPPORTED,
                       "FFT length %d (transform length %d): no factorisation into radices 2..17 for %d points "
                       "per workgroup", n, fft_len(n), need / n * fft_len(n));
    if (p.bn) {
        const std::pair<int, int> bk(p.bn, p.n);
        auto itb = ctx->bluestein.find(bk);
        if (!pb && itb != ctx->bluestein.end()) {
            p.chirp = itb->second.first;
            p.bspec = itb->second.second;
        } else {
            std::vector<float2> c, B;
            bluestein_tables(p.bn, p.n, c, B);
            float2 *dc = nullptr, *dB = nullptr;
            int rc = upload_table(ctx, pb, c, &dc);
            if (rc == CQ_OK) rc = upload_table(ctx, pb, B, &dB);
            if (rc) return rc;
            if (!pb) ctx->bluestein[bk] = std::make_pair(dc, dB);
            p.chirp = dc;
            p.bspec = dB;
        }
    }
    std::vector<int> key(p.rad, p.rad + p.ns);
    if (p.tc) key.push_back(-1);
    auto itp = ctx->twiddles.find(key);
    if (!pb && itp != ctx->twiddles.end()) {
        tw = itp->second;
        return CQ_OK;
    }
    // per-stage tables (cq_fft.h): stage s holds w_{Ns R}^(r k) at
    // (r-1) Ns + k, r = 1..R-1, k < Ns; n - 1 entries in all (one spare);
    // a compressed last stage holds r = 1 only
    std::vector<float2> h(tw_len(p), make_float2(1.f, 0.f));
    size_t o = 0;
    for (int s = 0, Ns = 1; s < p.ns; Ns *= p.rad[s], s++) {
        const int R = p.rad[s];
        const int rmax = (p.tc && s == p.ns - 1) ? 2 : R;
        for (int r = 1; r < rmax; r++)
            for (int k = 0; k < Ns; k++) {
                const double a = -2.0 * M_PI * (double)r * (double)k / ((double)Ns * R);
                h[o + (size_t)(r - 1) * Ns + k] = make_float2((float)cos(a), (float)sin(a));
            }
        o += (size_t)(R - 1) * Ns;
    }
    float2 *d = nullptr;
    const int rc = upload_table(ctx, pb, h, &d);
    if (rc) return rc;
    if (!pb) ctx->twiddles[key] = d;
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

// Per-shot device memory.  The state of a wavefield (what a revolve
// checkpoint must hold) is one contiguous block, so a takeshot or restore is a
// single device copy:  p, vx, vz, div, X, Y, psix, phix, psiz, phiz with
// X = lz (acoustic) or dvz (viscoacoustic) and Y = dvx (acoustic) or the kz
// half spectrum of p (viscoacoustic).
struct StateLayout {
    size_t p, vx, vz, div, x, y, psix, phix, psiz, phiz, bytes;
};

static StateLayout state_layout(int nxx, int ldz, int ldk, int pml, bool q)
{
    Arena A;
    StateLayout s;
    const size_t N = (size_t)nxx * ldz;
    s.p = A.take<float>(N);
    s.vx = A.take<float>(N);
    s.vz = A.take<float>(N);
    s.div = A.take<float>(N);
    s.x = A.take<float>(N);
    s.y = q ? A.take<float2>((size_t)cq_hs_cols(nxx) * ldk) : A.take<float>(N);
    s.psix = A.take<float>((size_t)2 * pml * ldz);
    s.phix = A.take<float>((size_t)2 * pml * ldz);
    s.psiz = A.take<float>((size_t)2 * pml * nxx);
    s.phiz = A.take<float>((size_t)2 * pml * nxx);
    s.bytes = (A.take<char>(0) + 255) & ~(size_t)255;
    return s;
}

static void bind_state(JobD &d, char *base, const StateLayout &s, bool q)
{
    d.p = (float *)(base + s.p);
    d.vx = (float *)(base + s.vx);
    d.vz = (float *)(base + s.vz);
    d.div = (float *)(base + s.div);
    if (q) {
        d.dvz = (float *)(base + s.x);
        d.s2 = (float2 *)(base + s.y);
    } else {
        d.lz = (float *)(base + s.x);
        d.dvx = (float *)(base + s.y);
    }
    d.psix = (float *)(base + s.psix);
    d.phix = (float *)(base + s.phix);
    d.psiz = (float *)(base + s.psiz);
    d.phiz = (float *)(base + s.phiz);
}

// A checkpoint copy of an RTM shot, issued before its local step `at`:
// dir 1 state -> slot (takeshot), 0 slot -> state (restore); src0 adds the
// wavelet's first sample at the source (taken at capo 0).
struct CkptOp {
    int at, dir, slot, src0;
};

// The revolve schedule of an RTM shot of nt steps with `snaps` checkpoints
// (sgpstd-qrtm/acps4.cu:447-905, sgpstd-qrtm/acsglw.cu:470-1040) as one step
// record {role, t, t_aux, flags} per executed step plus the checkpoint
// copies between steps:
//   advance  : source steps t = ts .. capo, t_aux = t - 1, ssg during the first sweep
//   firsturn : receiver step t = nt - 1, t_aux = nt
//   youturn  : receiver step t = capo, t_aux = capo + 1
//   takeshot / restore : copies; a restore resumes the source at capo + 1
static int rtm_schedule(int nt, int snaps, std::vector<int4> &steps, std::vector<CkptOp> &ops)
{
    cqh_revolve rv;
    cqh_revolve_init(&rv, nt, snaps);
    int ts = 0, restart = 0;
    steps.clear();
    ops.clear();
    for (;;) {
        const int a = cqh_revolve_next(&rv);
        const int at = (int)steps.size();
        if (a == CQH_RV_TAKESHOT) {
            ops.push_back({at, 1, rv.check, rv.capo == 0 ? 1 : 0});
        } else if (a == CQH_RV_ADVANCE) {
            for (int t = ts; t <= rv.capo; t++) steps.push_back(make_int4(0, t, t - 1, restart ? 0 : CQ_F_SSG));
            ts = rv.capo + 1;
        } else if (a == CQH_RV_FIRSTURN) {
            restart = 1;
            steps.push_back(make_int4(1, nt - 1, nt, 0));
        } else if (a == CQH_RV_YOUTURN) {
            steps.push_back(make_int4(1, rv.capo, rv.capo + 1, 0));
        } else if (a == CQH_RV_RESTORE) {
            ops.push_back({at, 0, rv.check, 0});
            ts = rv.capo + 1;
        } else if (a == CQH_RV_TERMINATE) {
            return CQ_OK;
        } else {
            return set_err(CQ_EINVAL, "revolve error, info %d", rv.info);
        }
    }
}

// Workgroup tiles of one job's passes: tz z rows per row tile, cw columns per
// column tile (of the job's kind), tk / tk2 kz rows per k_q_r1 / k_q_r2 tile.
struct JobTiles {
    int tz = 0, cw = 0, tk = 0, tk2 = 0;
};

struct JobH {
    JobD d[2];              // role 0: source / modeling wavefield; role 1: receiver wavefield (RTM)
    JobTiles tl;            // the tiles its data layout and plans are built for
    size_t lds[CQ_NPASS] = {};   // LDS bytes of each pass with those tiles
    int nroles = 1;
    void *mem = nullptr;
    size_t bytes = 0;
    char *state[2] = {nullptr, nullptr};
    char *ckpt = nullptr;   // snaps state blocks
    size_t state_bytes = 0;
    int snaps = 0;
    int nt = 0, nr = 0, nxx = 0, nzz = 0, ldz = 0, kind = 0, nx = 0, nz = 0;
    size_t src_cell = 0;    // field index of (xsg, zsg)
    float src0 = 0.f;       // rik[0] * dt * 1000 (checkpoint-0 quirk)
    int t0 = 0, nsched = 0; // global steps t0 .. t0 + nsched - 1
    std::vector<CkptOp> ops;  // RTM checkpoint copies, by local step
    std::vector<int4> hsched; // RTM step records (host copy of d[r].sched)
    size_t op_next = 0;
    int grp = 0;            // launch group (rebuild_tables)
};

// One launch of the per-step sequence.
static const size_t kLdsMax = 160 * 1024;
// LDS budget per workgroup that the tile choice aims for (CQ_LDS_TILE);
// tiles fall back to kLdsMax when a window needs it.
static const size_t kLdsTile = CQ_LDS_TILE;

// In launch order: a job's passes run in this order within a step (low-rank
// jobs take LR1 / LR2 where the others take QR1 / QC2).
enum PassId { P_AROW, P_ACOL, P_QC1, P_QR1, P_LR1, P_QC2, P_LR2, P_QR2, P_QC3, P_QR3, P_COUNT };
static_assert(P_COUNT == CQ_NPASS, "pass count of the C ABI");
static const int kGroups = 4;   // launch groups of a batch (see cq_batch_s::segs)
static const char *kPassName[P_COUNT] = {"k_row_ac", "k_col_ac", "k_q_c1", "k_q_r1", "k_lr_r1",
                                         "k_q_c2",   "k_lr_c2",  "k_q_r2", "k_q_c3", "k_q_r3"};

// A set of shots advanced together: every launch of a pass covers the tiles
// of every wavefield in the set, and each wavefield decides from its step
// record whether it advances at the launch's global step.
struct cq_batch_s {
    cq_ctx ctx = nullptr;
    std::vector<JobH> jobs;
    bool rtm = false;
    // job table: entry i = role 0 of jobs[i], entry cap + i = role 1
    JobD *djobs = nullptr;
    size_t djobs_cap = 0;                  // entries per role the table holds
    int2 *ditems[2][P_COUNT] = {};         // per role: (table entry, tile origin), XCD-ordered
    size_t ditems_cap[2][P_COUNT] = {};
    int nitems[2][P_COUNT] = {};
    size_t lds[P_COUNT] = {};
    int tz = 16, cw = 16, cwq = 16, tk = 8;   // row, acoustic column, Q column, kz-row tiles
    int tk2 = 8;                              // kz-row tile of k_q_r2 (tk, 2 tk or 3 tk by default)
    int plan = 0;                             // FFT factorisation variant (make_direct_plan)
    int maxnxx = 0, maxnzz = 0;
    int t = 0, tmax = 0;                   // next global step; end of the last wavefield
    bool timing = false;
    std::vector<hipEvent_t> ev;            // start of step, then one per launch
    std::vector<int> ev_pass;
    // Item groups: the jobs of a batch are split into kGroups groups (two of
    // viscoacoustic jobs, two of acoustic jobs), each with its own chain of
    // launches on its own stream (group 0 on the context's stream), so that
    // kernels of different groups overlap on the device.  A pass's items
    // come in launch segments, one per tile width of the pass among the jobs
    // (one unless a server job has tiles of its own, tiles_for); group g's
    // items of segment `sg` are ditems[r][p][sg.off[g] .. sg.off[g + 1]).
    // `split` is set while the side streams have forked from the context
    // stream.
    struct Seg {
        int val = 0;       // the pass's tile width (tz, cw / cwq, tk or tk2)
        size_t lds = 0;    // dynamic LDS of its launches
        int off[kGroups + 1] = {};
    };
    std::vector<Seg> segs[2][P_COUNT];
    // own_tiles (a server's batch): a job whose window the batch's tiles do
    // not fit within the LDS budget takes tiles of its own (tiles_for)
    bool own_tiles = false;
    hipStream_t side[kGroups] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kGroups] = {};
    bool split = false;
    // A pooled batch (cq_forward_static) allocates nothing: its arenas,
    // twiddle tables, job table and work items are carved out of `pool`
    // (the device's static scratch), and it runs on the context's stream
    // alone (no side streams or events).
    char *pool = nullptr;
    size_t pool_cap = 0, pool_off = 0;
    // Graph-captured steps (cq_batch_use_graphs): graph_k global steps per
    // graph, one executable graph per combination of active roles; each
    // launch group's kernels read the global step from its counter dstep[g]
    // (launched with step -1), which the graph advances.  dstep_t: the step
    // the counters hold (-1: unknown).
    int graph_k = 0;
    bool dry = false;   // launch() only raises the LDS limits (before a capture)
    int *dstep = nullptr;
    int dstep_t = -1;
    std::map<int, hipGraphExec_t> graphs;
    void drop_graphs()
    {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    ~cq_batch_s()
    {
        drop_graphs();
        if (!pool && dstep) (void)hipFree(dstep);
        if (!pool) {
            for (auto &j : jobs) (void)hipFree(j.mem);
            (void)hipFree(djobs);
            for (int r = 0; r < 2; r++)
                for (int i = 0; i < P_COUNT; i++) (void)hipFree(ditems[r][i]);
        }
        for (auto e : ev) (void)hipEventDestroy(e);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        for (int g = 1; g < kGroups; g++) {
            if (join_ev[g]) (void)hipEventDestroy(join_ev[g]);
            if (side[g]) (void)hipStreamDestroy(side[g]);
        }
    }
};

// `bytes` of a pooled batch's pool (256-byte aligned), or CQ_ENOMEM when
// the static scratch is exhausted.
static int pool_take(cq_batch_s *b, size_t bytes, void **out)
{
    const size_t o = (b->pool_off + 255) & ~(size_t)255;
    if (o + bytes > b->pool_cap)
        return set_err(CQ_ENOMEM, "the shot needs more than the %zu-byte static scratch of cq_forward_static",
                       b->pool_cap);
    *out = b->pool + o;
    b->pool_off = o + bytes;
    return CQ_OK;
}

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
    c->scratch = scratch_address(device);
    if (!c->scratch) {
        delete c;
        return set_err(CQ_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d", device);
    }
    *out = c;
    g_err.clear();
    return CQ_OK;
}

static void destroy_children(cq_ctx ctx);

extern "C" int cq_close(cq_ctx ctx)
{
    if (!ctx) return CQ_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    destroy_children(ctx);
    delete ctx;
    return CQ_OK;
}

static bool is_q(int kind) { return kind == CQ_KIND_VISCO || kind == CQ_KIND_QRTM; }
static bool is_rtm(int kind) { return kind == CQ_KIND_RTM || kind == CQ_KIND_QRTM; }

static int validate(const cq_job_desc &j, int idx)
{
    if (j.kind < CQ_KIND_ACOUSTIC || j.kind > CQ_KIND_QRTM)
        return set_err(CQ_EINVAL, "job %d: unknown kind %d", idx, j.kind);
    if (j.nxx < 8 || j.nzz < 8 || j.pml < 2 || 2 * j.pml >= j.nxx || 2 * j.pml >= j.nzz)
        return set_err(CQ_EINVAL, "job %d: bad window %dx%d pml %d", idx, j.nxx, j.nzz, j.pml);
    if (j.nt < 1 || !(j.dt > 0.f)) return set_err(CQ_EINVAL, "job %d: bad nt/dt", idx);
    if (!j.vp || !j.rho || !j.bbx || !j.bbz || !j.rik || !j.iax || !j.ibx || !j.hax || !j.hbx ||
        !j.iaz || !j.ibz || !j.haz || !j.hbz || !j.kxsmr || !j.kxsmi || !j.kxscr || !j.kxsci ||
        !j.kzsmr || !j.kzsmi || !j.kzscr || !j.kzsci)
        return set_err(CQ_EINVAL, "job %d: missing input array", idx);
    if (is_q(j.kind) && (!j.gam || !j.lam || !j.vwd || !j.tao))
        return set_err(CQ_EINVAL, "job %d: viscoacoustic job without gam/lam/vwd/tao", idx);
    if (j.xsg < 0 || j.xsg >= j.nxx || j.zsg < 0 || j.zsg >= j.nzz)
        return set_err(CQ_EINVAL, "job %d: source index outside the window", idx);
    if (j.nr < 0 || (j.nr > 0 && (!j.gxg || !j.gzg)))
        return set_err(CQ_EINVAL, "job %d: bad receivers", idx);
    for (int i = 0; i < j.nr; i++)
        if (j.gxg[i] < 0 || j.gxg[i] >= j.nxx || j.gzg[i] < 0 || j.gzg[i] >= j.nzz)
            return set_err(CQ_EINVAL, "job %d: receiver %d outside the window", idx, i);
    if (!(j.dx > 0.f) || !(j.dz > 0.f)) return set_err(CQ_EINVAL, "job %d: bad dx/dz", idx);
    if (j.lr_rank < 0 || j.lr_rank > 8) return set_err(CQ_EINVAL, "job %d: lr_rank %d outside 0..8", idx, j.lr_rank);
    if (j.lr_rank > 0 && (!is_q(j.kind) || !j.lr_w || !j.lr_sym))
        return set_err(CQ_EINVAL, "job %d: a low-rank operator needs a viscoacoustic kind, lr_w and lr_sym", idx);
    if (is_rtm(j.kind)) {
        // revolve limits checkup = 64 (src.c:176) and snaps >= 1
        if (j.snaps < 1 || j.snaps > 64) return set_err(CQ_EINVAL, "job %d: snaps %d outside 1..64", idx, j.snaps);
        if (j.nr > 0 && !j.obs) return set_err(CQ_EINVAL, "job %d: RTM job without observed data", idx);
    }
    return CQ_OK;
}

// CSR of items sorted by key into tiles of width `tw` (ntile + 1 offsets)
static void tile_csr(const std::vector<int> &key, int tw, int ntile, std::vector<int> &order, std::vector<int> &off)
{
    const int n = (int)key.size();
    order.resize(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return key[a] < key[c]; });
    off.assign(ntile + 1, 0);
    for (int tl = 0, k = 0; tl <= ntile; tl++) {
        while (k < n && key[order[k]] < tl * tw) k++;
        off[tl] = k;
    }
}

static void job_lds(const JobTiles &tl, const JobD &d, bool q, size_t *out);
static int tiles_for(cq_batch b, const cq_job_desc &j, JobTiles &out);
static void job_needs(const cq_job_desc &j, const JobTiles &t, int variant, bool &fused, int &need_x, int &need_z);
static JobTiles batch_tiles(cq_batch b, bool q);

// Allocates and uploads one shot for its tile shapes (tiles_for).  With
// bytes_only != null only its device footprint is computed.
static int build_job(cq_batch b, const cq_job_desc &j, JobH &h, size_t *bytes_only = nullptr)
{
    cq_ctx ctx = b->ctx;
    int rc = tiles_for(b, j, h.tl);
    if (rc) return rc;
    const int cw = h.tl.cw, tk = h.tl.tk, tk2 = h.tl.tk2, tz = h.tl.tz;
    const bool q = is_q(j.kind), rtm = is_rtm(j.kind);
    const int nxx = j.nxx, nzz = j.nzz, pml = j.pml;
    const int ldz = ((nzz + tz - 1) / tz * tz + 31) & ~31;   // whole row tiles, 128-B rows
    const int nzh = nzz / 2 + 1;
    const int kt = std::lcm(tk, tk2);
    const int ldk = (nzh + kt - 1) / kt * kt;   // whole tiles of both kz-row passes
    const int nx = nxx - 2 * pml, nz = nzz - 2 * pml;
    const size_t N = (size_t)nxx * ldz;
    const int ntile_x = (nxx + cw - 1) / cw, ntile_z = (nzz + tz - 1) / tz;
    const int nrole = rtm ? 2 : 1;
    const int snaps = rtm ? j.snaps : 0;

    JobD base;
    memset(&base, 0, sizeof base);
    base.kind = j.kind; base.nxx = nxx; base.nzz = nzz; base.ldz = ldz; base.pml = pml; base.nt = j.nt;
    base.xsg = j.xsg; base.zsg = j.zsg; base.nr = j.nr; base.nzh = nzh; base.ldk = ldk;
    base.nx = nx; base.nz = nz;
    base.src_scale = j.dt * 1000.0f;
    base.dt = j.dt;
    base.dkx = k_step(nxx, j.dx);
    base.dkz = k_step(nzz, j.dz);
    base.invN = (float)(1.0 / ((double)nxx * nzz));
    base.qsign = (j.kind == CQ_KIND_QRTM) ? -1.0f : 1.0f;
    base.stab = (j.kind == CQ_KIND_QRTM) ? 1 : 0;
    base.gather = rtm ? 0 : 1;
    bool fused;
    int need_x, need_z;
    job_needs(j, h.tl, b->plan, fused, need_x, need_z);
    base.r1_fused = fused ? 1 : 0;
    if (bytes_only) {
        // footprint only: the plans must exist, nothing is uploaded
        if (!make_plan(nxx, need_x, base.px, b->plan) || !make_plan(nzz, need_z, base.pz, b->plan))
            return set_err(CQ_EUNSUPPORTED, "window %dx%d: no FFT plan for the batch's tiles", nxx, nzz);
    } else {
        cq_batch_s *pb = b->pool ? b : nullptr;
        if ((rc = get_plan(ctx, pb, nxx, need_x, b->plan, base.px, base.twx)) != CQ_OK) return rc;
        if ((rc = get_plan(ctx, pb, nzz, need_z, b->plan, base.pz, base.twz)) != CQ_OK) return rc;
    }
    // lines hold the transform length (M points for a Bluestein plan)
    base.twlx = tw_len(base.px);
    base.twlz = tw_len(base.pz);
    base.lsx = cq_line_stride(base.px.n);
    base.lsz = cq_line_stride(base.pz.n);
    {
        // the passes' LDS for this window with its tiles; a Bluestein
        // length may need more than the batch's widest window
        size_t need_lds[P_COUNT];
        job_lds(h.tl, base, q, need_lds);
        for (int pp = 0; pp < P_COUNT; pp++)
            if (need_lds[pp] > kLdsMax)
                return set_err(CQ_EUNSUPPORTED, "window %dx%d: pass %s needs %zu bytes of LDS (at most %zu)", nxx, nzz,
                               kPassName[pp], need_lds[pp], kLdsMax);
        for (int pp = 0; pp < P_COUNT; pp++) h.lds[pp] = need_lds[pp];
        // a job with the batch's tiles raises the batch's launches; one with
        // tiles of its own sizes its own (rebuild_tables)
        const JobTiles bt = batch_tiles(b, q);
        const bool own = bt.tz != tz || bt.cw != cw || bt.tk != tk || bt.tk2 != tk2;
        if (!bytes_only && !own)
            for (int pp = 0; pp < P_COUNT; pp++) b->lds[pp] = std::max(b->lds[pp], need_lds[pp]);
    }

    const StateLayout sl = state_layout(nxx, ldz, ldk, pml, q);
    Arena A;
    size_t o_state[2] = {0, 0};
    for (int r = 0; r < nrole; r++) o_state[r] = A.take<char>(sl.bytes);
    const size_t o_ckpt = A.take<char>(sl.bytes * (size_t)snaps);
    const size_t o_bx = A.take<float>(N), o_bz = A.take<float>(N);
    const size_t o_f2 = A.take<float>(N), o_f3 = A.take<float>(N);
    size_t o_q[3] = {0, 0, 0}, o_tot = 0, o_scr = 0, o_s1 = 0, o_s3 = 0, o_s4 = 0;
    if (q) {
        for (int i = 0; i < 3; i++) o_q[i] = A.take<float>(N);
        o_tot = A.take<float>(N);
        o_scr = A.take<float>(N);
        o_s1 = A.take<float2>((size_t)cq_hs_cols(nxx) * ldk);
        o_s3 = A.take<float2>((size_t)cq_hs_cols(nxx) * ldk);
        o_s4 = A.take<float2>((size_t)cq_hs_cols(nxx) * ldk);
    }
    // low-rank operator: weights (field layout), quarter-plane symbols, half spectra
    const int lr = q ? j.lr_rank : 0, nxh = nxx / 2 + 1;
    size_t o_lrw = 0, o_lrv = 0, o_lrs = 0;
    if (lr) {
        o_lrw = A.take<float>(N * lr);
        o_lrv = A.take<float>((size_t)nxh * nzh * lr);
        o_lrs = A.take<float2>((size_t)cq_hs_cols(nxx) * ldk * lr);
    }
    const size_t o_1x = A.take<float>((size_t)4 * nxx), o_1z = A.take<float>((size_t)4 * nzz);
    const size_t o_mx = A.take<float2>((size_t)3 * nxx), o_mz = A.take<float2>((size_t)3 * nzz);
    const size_t o_rik = A.take<float>(j.nt);
    const size_t o_rec = A.take<float>(rtm ? 1 : (size_t)j.nt * std::max(j.nr, 1));
    const size_t o_rcv = A.take<int>((size_t)3 * std::max(j.nr, 1));
    const size_t o_off = A.take<int>((size_t)ntile_x + 1);
    size_t o_obs = 0, o_img = 0, o_ssg = 0, o_inj = 0, o_ioff = 0;
    const int ntile_inj = q ? ntile_x : ntile_z;
    if (rtm) {
        o_obs = A.take<float>((size_t)j.nt * std::max(j.nr, 1));
        o_img = A.take<float>(N);   // field layout, interior cells used
        o_ssg = A.take<float>(N);
        o_inj = A.take<int>((size_t)3 * std::max(j.nr, 1));
        o_ioff = A.take<int>((size_t)ntile_inj + 1);
    }
    std::vector<int4> hsched;
    size_t o_sched = 0;
    if (rtm) {
        int rc0 = rtm_schedule(j.nt, j.snaps, hsched, h.ops);
        if (rc0) return rc0;
        o_sched = A.take<int4>(hsched.size());
    }
    h.nsched = rtm ? (int)hsched.size() : j.nt;
    h.hsched = hsched;
    h.t0 = 0;
    h.op_next = 0;
    const size_t total = A.take<char>(0);
    if (bytes_only) {
        *bytes_only = total;
        return CQ_OK;
    }
    char *mem = nullptr;
    if (b->pool) {
        rc = pool_take(b, total, (void **)&mem);
        if (rc) return rc;
    } else {
        hipError_t e = hipMalloc(&mem, total);
        if (e != hipSuccess) return set_err(CQ_ENOMEM, "hipMalloc(%zu): %s", total, hipGetErrorString(e));
        h.mem = mem;
    }
    h.bytes = total;
    h.nroles = nrole;
    h.state_bytes = sl.bytes;
    h.snaps = snaps;
    h.ckpt = mem + o_ckpt;
    h.nt = j.nt; h.nr = j.nr; h.nxx = nxx; h.nzz = nzz; h.ldz = ldz; h.kind = j.kind; h.nx = nx; h.nz = nz;
    h.src_cell = cq_q4(nxx, j.xsg, j.zsg >> 2) + (j.zsg & 3);
    h.src0 = j.rik[0] * j.dt * 1000.0f;
    HIPCHK(hipMemsetAsync(mem, 0, total, ctx->stream));

    base.dtbx = (float *)(mem + o_bx); base.dtbz = (float *)(mem + o_bz);
    float *f2 = (float *)(mem + o_f2), *f3 = (float *)(mem + o_f3);
    if (!q) {
        base.c1 = f2; base.c2 = f3;
    } else {
        base.qa = f2; base.qb = f3;
        base.qg = (float *)(mem + o_q[0]);
        base.qy2 = (float *)(mem + o_q[1]); base.qcc = (float *)(mem + o_q[2]);
        base.tot = (float *)(mem + o_tot);
        base.scr = (float *)(mem + o_scr);
        base.s1 =(float2 *)(mem + o_s1); base.s3 = (float2 *)(mem + o_s3); base.s4 = (float2 *)(mem + o_s4);
    }
    base.lr_rank = lr;
    base.nxh = nxh;
    if (lr) {
        base.lrw = (float *)(mem + o_lrw);
        base.lrv = (float *)(mem + o_lrv);
        base.lrs = (float2 *)(mem + o_lrs);
    }
    float *x1 = (float *)(mem + o_1x), *z1 = (float *)(mem + o_1z);
    base.iax = x1; base.ibx = x1 + nxx; base.hax = x1 + 2 * nxx; base.hbx = x1 + 3 * nxx;
    base.iaz = z1; base.ibz = z1 + nzz; base.haz = z1 + 2 * nzz; base.hbz = z1 + 3 * nzz;
    float2 *mx = (float2 *)(mem + o_mx), *mz = (float2 *)(mem + o_mz);
    base.mxm = mx; base.mxp = mx + nxx; base.mx2 = mx + 2 * nxx;
    base.mzm = mz; base.mzp = mz + nzz; base.mz2 = mz + 2 * nzz;
    base.rik = (float *)(mem + o_rik);
    base.rec = (float *)(mem + o_rec);
    int *rcv = (int *)(mem + o_rcv);
    base.rcv_x = rcv; base.rcv_z = rcv + j.nr; base.rcv_i = rcv + 2 * j.nr;
    base.rcv_off = (int *)(mem + o_off);
    if (rtm) {
        base.obs = (float *)(mem + o_obs);
        base.img = (float *)(mem + o_img);
        base.ssg = (float *)(mem + o_ssg);
        int *inj = (int *)(mem + o_inj);
        base.inj_x = inj; base.inj_z = inj + j.nr; base.inj_i = inj + 2 * j.nr;
        base.inj_off = (int *)(mem + o_ioff);
        // stabiliser: tao_c = max tao over the padded window (sgpstd-qrtm/acsglw.cu:354-356)
        if (q) {
            float mx_ = 0.0f;
            for (size_t a = 0; a < (size_t)nxx * nzz; a++)
                if (j.tao[a] >= mx_) mx_ = j.tao[a];
            base.taoc = mx_;
        }
    }
    base.t0 = 0;
    base.nsched = h.nsched;
    base.sched = rtm ? (const int4 *)(mem + o_sched) : nullptr;
    for (int r = 0; r < nrole; r++) {
        h.d[r] = base;
        h.d[r].role = r;
        h.state[r] = mem + o_state[r];
        bind_state(h.d[r], h.state[r], sl, q);
        if (r == 1) {
            h.d[r].psrc = h.d[0].p;
            h.d[r].eta = j.etar;
        } else {
            h.d[r].eta = j.etas;
        }
    }

    // per-cell coefficients (field layout, zero padding)
    const float dt = j.dt;
    std::vector<float> hbx(N, 0.f), hbz(N, 0.f), hf2(N, 0.f), hf3(N, 0.f);
    std::vector<float> hq[3];
    if (q)
        for (int i = 0; i < 3; i++) hq[i].assign(N, 0.f);
    for (int ix = 0; ix < nxx; ix++)
        for (int iz = 0; iz < nzz; iz++) {
            const size_t a = (size_t)ix * nzz + iz, g = cq_q4(nxx, ix, iz >> 2) + (iz & 3);
            const float v = j.vp[a];
            hbx[g] = dt * j.bbx[a];
            hbz[g] = dt * j.bbz[a];
            if (!q) {
                // c1 = rho v^2 dt, c2 = rho v^2 dt * v^2 dt^2 / 12 (addsa_kernel, acps4.cu:1488)
                const float r = j.rho[a] * v * v * dt;
                hf2[g] = r;
                hf3[g] = (float)((double)r * (1.0 / 12.0) * v * dt * v * dt);
            } else {
                // acsglw.cu:1528-1529: p += (rho v^2 dt lam)(total - (dt tao/2) T) - (dt tao) aux.
                // dt tao is not stored: the float dt * tao is exactly twice
                // (float)(0.5 dt tao) (the double product of two floats is
                // exact, and halving commutes with rounding), so the pass
                // forms it as 2 qb bit for bit
                const float gm = j.gam[a], lam = j.lam[a], tao = j.tao[a], vwd = j.vwd[a];
                hf2[g] = j.rho[a] * v * v * dt * lam;
                hf3[g] = (float)(0.5 * dt * tao);
                // sum_total (acsglw.cu:1232-1240): y2 = 32 gam vwd, c = dt^2 v^2 lam / 12
                hq[0][g] = gm;
                hq[1][g] = (float)(32.0 * gm * vwd);
                const float dtf = dt * dt, vpf = v * v;
                hq[2][g] = (float)(1.0 / 12.0 * dtf * vpf * lam);
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
    // receivers sorted by column, with per-x-tile offsets (gather)
    std::vector<int> gx(j.gxg, j.gxg + j.nr), order, hoff;
    tile_csr(gx, cw, ntile_x, order, hoff);
    std::vector<int> hr(3 * std::max(j.nr, 1), 0);
    for (int k = 0; k < j.nr; k++) {
        hr[k] = j.gxg[order[k]];
        hr[j.nr + k] = j.gzg[order[k]];
        hr[2 * j.nr + k] = order[k];
    }
    // RTM data injection: one entry per distinct receiver cell, holding the
    // last receiver mapped to it (rindex, acps4.cu:418-420), by row tile
    // (acoustic) or column tile (viscoacoustic)
    std::vector<int> hinj(3 * std::max(j.nr, 1), 0), hioff;
    if (rtm) {
        std::map<std::pair<int, int>, int> cell;
        for (int i = 0; i < j.nr; i++) cell[{j.gxg[i], j.gzg[i]}] = i;
        std::vector<int> cx, cz, ci, key;
        for (auto &kv : cell) {
            cx.push_back(kv.first.first);
            cz.push_back(kv.first.second);
            ci.push_back(kv.second);
            key.push_back(q ? kv.first.first : kv.first.second);
        }
        std::vector<int> ord;
        tile_csr(key, q ? cw : tz, ntile_inj, ord, hioff);
        const int nc = (int)cx.size();
        for (int k = 0; k < nc; k++) {
            hinj[k] = cx[ord[k]];
            hinj[j.nr + k] = cz[ord[k]];
            hinj[2 * j.nr + k] = ci[ord[k]];
        }
    }
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync((void *)base.dtbx, hbx.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void *)base.dtbz, hbz.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(f2, hf2.data(), 4 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(f3, hf3.data(), 4 * N, hipMemcpyHostToDevice, s));
    if (q)
        for (int i = 0; i < 3; i++)
            HIPCHK(hipMemcpyAsync(mem + o_q[i], hq[i].data(), 4 * N, hipMemcpyHostToDevice, s));
    std::vector<float> hlw;
    if (lr) {
        // weights to the field layout; symbols as given (quarter plane)
        hlw.assign(N * lr, 0.f);
        for (int i = 0; i < lr; i++)
            for (int ix = 0; ix < nxx; ix++)
                for (int iz = 0; iz < nzz; iz++)
                    hlw[N * i + cq_q4(nxx, ix, iz >> 2) + (iz & 3)] = j.lr_w[(size_t)i * nxx * nzz + (size_t)ix * nzz + iz];
        HIPCHK(hipMemcpyAsync((void *)base.lrw, hlw.data(), 4 * hlw.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync((void *)base.lrv, j.lr_sym, 4 * (size_t)nxh * nzh * lr, hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipMemcpyAsync(x1, h1x.data(), 4 * h1x.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(z1, h1z.data(), 4 * h1z.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(mx, hmx.data(), 8 * hmx.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(mz, hmz.data(), 8 * hmz.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void *)base.rik, j.rik, 4 * (size_t)j.nt, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(rcv, hr.data(), 4 * hr.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void *)base.rcv_off, hoff.data(), 4 * hoff.size(), hipMemcpyHostToDevice, s));
    if (rtm) {
        if (j.nr > 0)
            HIPCHK(hipMemcpyAsync((void *)base.obs, j.obs, 4 * (size_t)j.nt * j.nr, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync((void *)base.inj_x, hinj.data(), 4 * hinj.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync((void *)base.inj_off, hioff.data(), 4 * hioff.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync((void *)base.sched, hsched.data(), sizeof(int4) * hsched.size(), hipMemcpyHostToDevice, s));
    }
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


// LDS bytes of a workgroup holding the twiddle table of an n-point transform
// and `lines` lines of it (cq_line_stride slots apart), for the transform
// length fft_len(n) (M points for a Bluestein length) and the plan of that
// many lines (its table length); no plan: more than any workgroup has.
static size_t lds_lines(int n, int lines)
{
    if (n < 1) return 0;   // no window of this kind
    CqPlan p;
    if (!make_plan(n, lines * n, p)) return (size_t)1 << 40;
    return (size_t)(tw_len(p) + lines * cq_line_stride(p.n)) * 8;
}
// The same for a job's own plan (twl table entries, line stride ls).
static size_t lds_job(int twl, int ls, int lines) { return (size_t)(twl + lines * ls) * 8; }
// Candidate tile widths, largest first: z rows per row tile (two per packed
// line), columns per column tile (two per packed line), kz rows per kz-row
// tile.  Widths that are not powers of two fill a workgroup's 512 threads
// better when a power of two does not fit the LDS budget.
static const int kTzCand[] = {16, 12, 8, 4};
static const int kCwCand[] = {32, 24, 16, 12, 8, 4, 2};
static const int kTkCand[] = {8, 4, 2, 1};

// k_q_r1 holds 3 tk lines and transforms 2 tk at a time (the div and p
// spectra, then two symbols at a time) when the LDS budget allows it, else
// inverts one symbol at a time (2 tk lines: spectrum + work, tk at a time).
static bool r1_fused_for(int nxx, int tk, int variant)
{
    return has_plan(nxx, 2 * tk * nxx, variant) && lds_lines(nxx, 3 * tk) <= kLdsTile;
}
static size_t r1_lds(int nxx, int tk, int variant)
{
    return lds_lines(nxx, r1_fused_for(nxx, tk, variant) ? 3 * tk : 2 * tk);
}
static bool r1_ok(int nxx, int tk, size_t lim, int variant = 0)
{
    return (has_plan(nxx, 2 * tk * nxx, variant) && lds_lines(nxx, 3 * tk) <= lim) ||
           (has_plan(nxx, tk * nxx, variant) && lds_lines(nxx, 2 * tk) <= lim);
}

// LDS of each pass for one job (its plans' table lengths and line strides)
// with tiles tl.
static void job_lds(const JobTiles &tl, const JobD &d, bool q, size_t *out)
{
    for (int p = 0; p < P_COUNT; p++) out[p] = 0;
    auto xl = [&](int lines) { return lds_job(d.twlx, d.lsx, lines); };
    auto zl = [&](int lines) { return lds_job(d.twlz, d.lsz, lines); };
    if (!q) {
        out[P_AROW] = xl(tl.tz / 2);
        out[P_ACOL] = zl(tl.cw / 2);
    } else {
        out[P_QC1] = out[P_QC2] = out[P_QC3] = out[P_LR2] = zl(tl.cw / 2);
        out[P_QR1] = xl(d.r1_fused ? 3 * tl.tk : 2 * tl.tk);
        out[P_LR1] = xl(2 * tl.tk);
        out[P_QR2] = xl(tl.tk2);
        out[P_QR3] = xl(tl.tz / 2);
    }
}

// LDS of each pass for windows up to ax x az (acoustic) and qx x qz
// (viscoacoustic); k_q_r1 takes the largest need of the lengths in qxs.
static void set_lds(cq_batch b, int ax, int az, int qx, int qz, const std::vector<int> &qxs)
{
    b->lds[P_AROW] = lds_lines(ax, b->tz / 2);
    b->lds[P_ACOL] = lds_lines(az, b->cw / 2);
    b->lds[P_QC1] = lds_lines(qz, b->cwq / 2);
    b->lds[P_QR1] = 0;
    for (int n : qxs) b->lds[P_QR1] = std::max(b->lds[P_QR1], r1_lds(n, b->tk, b->plan));
    b->lds[P_QC2] = lds_lines(qz, b->cwq / 2);
    b->lds[P_QR2] = lds_lines(qx, b->tk2);
    b->lds[P_QC3] = lds_lines(qz, b->cwq / 2);
    b->lds[P_QR3] = lds_lines(qx, b->tz / 2);
}

// Tile shapes for the windows a batch will hold: the largest tiles whose
// transforms fit one workgroup's per-stage capacity (a plan exists) and the
// LDS budget kLdsTile, else kLdsMax.  Row tiles are shared by both kinds,
// column tiles are chosen per kind (cw acoustic, cwq viscoacoustic).
struct TileReq {
    int nxx, nzz;
    bool q;
};

struct TileSet {
    int tz = 16, cw = 16, cwq = 16, tk = 8, tk2 = 8;
};

static int pick_tiles(const std::vector<TileReq> &req, TileSet &t)
{
    int iz = 0, ic = 0, icq = 0, ik = 0;   // indices into the candidate lists
    std::vector<int> qxs;
    for (size_t i = 0; i < req.size(); i++) {
        const TileReq &j = req[i];
        if (j.q) qxs.push_back(j.nxx);
        if (!has_plan(j.nxx, j.nxx))
            return set_err(CQ_EUNSUPPORTED, "job %d: no FFT plan for nxx=%d (transform length %d)", (int)i, j.nxx,
                           fft_len(j.nxx));
        if (!has_plan(j.nzz, j.nzz))
            return set_err(CQ_EUNSUPPORTED, "job %d: no FFT plan for nzz=%d (transform length %d)", (int)i, j.nzz,
                           fft_len(j.nzz));
        const bool q = j.q;
        auto row_bad = [&](int t, size_t lim) {
            return !has_plan(j.nxx, (t / 2) * j.nxx) || lds_lines(j.nxx, t / 2) > lim;
        };
        auto col_bad = [&](int c, size_t lim) {
            return !has_plan(j.nzz, (c / 2) * j.nzz) || lds_lines(j.nzz, c / 2) > lim;
        };
        auto krow_bad = [&](int k, size_t lim) { return !r1_ok(j.nxx, k, lim); };
        int &ick = q ? icq : ic;
        while (iz < 3 && row_bad(kTzCand[iz], kLdsTile)) iz++;
        while (ick < 6 && col_bad(kCwCand[ick], kLdsTile)) ick++;
        if (q)
            while (ik < 3 && krow_bad(kTkCand[ik], kLdsTile)) ik++;
        if (row_bad(kTzCand[iz], kLdsMax) || col_bad(kCwCand[ick], kLdsMax) || (q && krow_bad(kTkCand[ik], kLdsMax)))
            return set_err(CQ_EUNSUPPORTED, "job %d: window %dx%d too large for one workgroup", (int)i, j.nxx, j.nzz);
    }
    t.tz = kTzCand[iz];
    t.cw = kCwCand[ic];
    t.cwq = kCwCand[icq];
    t.tk = kTkCand[ik];
    // k_q_r2 takes three or two kz-row tiles at once where every length
    // allows it (more lines per transform keep more of the workgroup busy)
    t.tk2 = t.tk;
    for (int m : {3, 2}) {
        const int t2 = m * t.tk;
        if (t2 > 16) continue;
        bool ok = true;
        for (int n : qxs) ok = ok && has_plan(n, t2 * n) && lds_lines(n, t2) <= kLdsTile;
        if (ok) {
            t.tk2 = t2;
            break;
        }
    }
    return CQ_OK;
}

static int choose_tiles(cq_batch b, const std::vector<TileReq> &req)
{
    TileSet t;
    const int rc = pick_tiles(req, t);
    if (rc) return rc;
    b->tz = t.tz;
    b->cw = t.cw;
    b->cwq = t.cwq;
    b->tk = t.tk;
    b->tk2 = t.tk2;
    int ax = 0, az = 0, qx = 0, qz = 0;
    std::vector<int> qxs;
    for (const TileReq &j : req) {
        int &mx = j.q ? qx : ax, &mz = j.q ? qz : az;
        mx = std::max(mx, j.nxx);
        mz = std::max(mz, j.nzz);
        if (j.q) qxs.push_back(j.nxx);
    }
    b->maxnxx = std::max(ax, qx);
    b->maxnzz = std::max(az, qz);
    set_lds(b, ax, az, qx, qz, qxs);
    return CQ_OK;
}

// The batch's tiles for a job of its kind.
static JobTiles batch_tiles(cq_batch b, bool q) { return JobTiles{b->tz, q ? b->cwq : b->cw, b->tk, b->tk2}; }

// Points per workgroup transform of a job's x and z transforms with tiles t:
// row tiles tz/2 lines, kz-row tiles 2 tk (fused k_q_r1) or tk lines and tk2
// lines (k_q_r2) of nxx; column tiles cw/2 lines of nzz.
static void job_needs(const cq_job_desc &j, const JobTiles &t, int variant, bool &fused, int &need_x, int &need_z)
{
    const bool q = is_q(j.kind);
    fused = q && r1_fused_for(j.nxx, t.tk, variant);
    need_x = std::max((t.tz / 2) * j.nxx, q ? std::max((fused ? 2 : 1) * t.tk, t.tk2) * j.nxx : 0);
    need_z = (t.cw / 2) * j.nzz;
}

// True if the job's transforms have plans with the batch's tiles t and no
// pass needs more LDS than the batch's launches of it have or the tile
// budget kLdsTile allows, whichever is more.
static bool tiles_fit(cq_batch b, const cq_job_desc &j, const JobTiles &t)
{
    const int variant = b->plan;
    bool fused;
    int need_x, need_z;
    job_needs(j, t, variant, fused, need_x, need_z);
    CqPlan px, pz;
    if (!make_plan(j.nxx, need_x, px, variant) || !make_plan(j.nzz, need_z, pz, variant)) return false;
    JobD d;
    memset(&d, 0, sizeof d);
    d.twlx = tw_len(px);
    d.twlz = tw_len(pz);
    d.lsx = cq_line_stride(px.n);
    d.lsz = cq_line_stride(pz.n);
    d.r1_fused = fused ? 1 : 0;
    size_t l[P_COUNT];
    job_lds(t, d, is_q(j.kind), l);
    for (int p = 0; p < P_COUNT; p++)
        if (l[p] > std::max(kLdsTile, b->lds[p])) return false;
    return true;
}

// A job's tiles: the batch's, or, for a server job whose window those do not
// fit (tiles_fit: e.g. a Bluestein length, whose inner transform is about
// twice as long), the tiles a batch of that job alone gets, so that one such
// job neither is refused nor makes every job's launches take more LDS.
// Either way a function of the job and the server's tiles: a job's results
// do not depend on the other jobs.
static int tiles_for(cq_batch b, const cq_job_desc &j, JobTiles &out)
{
    const bool q = is_q(j.kind);
    out = batch_tiles(b, q);
    if (!b->own_tiles || tiles_fit(b, j, out)) return CQ_OK;
    TileSet t;
    const int rc = pick_tiles({TileReq{j.nxx, j.nzz, q}}, t);
    if (rc) return rc;
    out = JobTiles{t.tz, q ? t.cwq : t.cw, t.tk, t.tk2};
    return CQ_OK;
}

// The tile width a pass is launched with: z rows (row passes), columns
// (column passes) or kz rows (k_q_r1 / k_lr_r1: tk, k_q_r2: tk2).
static int pass_tile(const JobTiles &t, int p)
{
    switch (p) {
    case P_AROW:
    case P_QR3: return t.tz;
    case P_QR1:
    case P_LR1: return t.tk;
    case P_QR2: return t.tk2;
    default: return t.cw;
    }
}

// Device job table and per-pass work items for the batch's current jobs.
// Table entry i is role 0 of jobs[i], entry djobs_cap + i its role 1; items
// cover every tile of every wavefield.  Nothing may be in flight on the
// stream (the old table and items are overwritten).
static int rebuild_tables(cq_batch b)
{
    hipStream_t s = b->ctx->stream;
    const size_t n = b->jobs.size();
    if (n > b->djobs_cap || !b->djobs) {
        if (b->pool) {
            // a pooled batch builds its tables once
            b->djobs_cap = std::max(n, (size_t)1);
            const int rc = pool_take(b, sizeof(JobD) * 2 * b->djobs_cap, (void **)&b->djobs);
            if (rc) return rc;
        } else {
            (void)hipFree(b->djobs);
            b->djobs = nullptr;
            b->djobs_cap = std::max(n, (size_t)2 * b->djobs_cap);
            if (b->djobs_cap == 0) b->djobs_cap = 1;
            HIPCHK(hipMalloc(&b->djobs, sizeof(JobD) * 2 * b->djobs_cap));
        }
    }
    b->drop_graphs();
    b->dstep_t = -1;   // groups may have changed: the counters are set again before a graph runs
    if (!b->dstep) {
        if (b->pool) {
            const int rc = pool_take(b, sizeof(int) * kGroups, (void **)&b->dstep);
            if (rc) return rc;
        } else {
            HIPCHK(hipMalloc(&b->dstep, sizeof(int) * kGroups));
        }
        HIPCHK(hipMemsetAsync(b->dstep, 0, sizeof(int) * kGroups, s));
        b->dstep_t = -1;
    }
    const int cap = (int)b->djobs_cap;
    std::vector<JobD> hd(2 * (size_t)cap);
    memset(hd.data(), 0, sizeof(JobD) * hd.size());
    // per role and pass: tile width -> each group's items, and the width's
    // largest LDS among jobs with tiles of their own
    std::map<int, std::array<std::vector<int2>, kGroups>> items[2][P_COUNT];
    std::map<int, size_t> seg_lds[2][P_COUNT];
    b->tmax = 0;
    int nq = 0, na = 0;   // jobs of each kind so far (round-robin over the kind's two groups)
    for (size_t i = 0; i < n; i++) {
        JobH &h = b->jobs[i];
        hd[i] = h.d[0];
        hd[cap + i] = h.nroles > 1 ? h.d[1] : h.d[0];
        b->tmax = std::max(b->tmax, h.t0 + h.nsched);
        const int g = is_q(h.kind) ? (nq++ & 1) : 2 + (na++ & 1);
        h.grp = g;
        hd[i].dstep = hd[cap + i].dstep = b->dstep + g;
        const JobTiles &tl = h.tl;
        const JobTiles bt = batch_tiles(b, is_q(h.kind));
        const bool own = bt.tz != tl.tz || bt.cw != tl.cw || bt.tk != tl.tk || bt.tk2 != tl.tk2;
        for (int r = 0; r < h.nroles; r++) {
            const int e = r ? cap + (int)i : (int)i;
            // tiles of width w from 0 to n of pass p (origin = tile start)
            auto add = [&](int p, int w, int n) {
                std::vector<int2> &v = items[r][p][w][g];
                for (int o = 0; o < n; o += w) v.push_back(make_int2(e, o));
                size_t &l = seg_lds[r][p][w];
                if (own) l = std::max(l, h.lds[p]);
            };
            if (!is_q(h.kind)) {
                add(P_AROW, tl.tz, h.nzz);
                add(P_ACOL, tl.cw, h.nxx);
            } else {
                const bool lr = h.d[0].lr_rank > 0;
                add(P_QC1, tl.cw, h.nxx);
                add(lr ? P_LR2 : P_QC2, tl.cw, h.nxx);
                add(P_QC3, tl.cw, h.nxx);
                add(lr ? P_LR1 : P_QR1, tl.tk, h.d[0].nzh);
                add(P_QR2, tl.tk2, h.d[0].nzh);
                add(P_QR3, tl.tz, h.nzz);
            }
        }
    }
    std::vector<int2> ord[2][P_COUNT];
    HIPCHK(hipMemcpyAsync(b->djobs, hd.data(), sizeof(JobD) * hd.size(), hipMemcpyHostToDevice, s));
    for (int r = 0; r < 2; r++)
        for (int p = 0; p < P_COUNT; p++) {
            b->nitems[r][p] = 0;
            b->segs[r][p].clear();
            // segments one after another; in each the groups one after
            // another, each in XCD order.  A segment of the batch's width
            // launches with the batch's LDS (raised by its jobs, build_job),
            // or more if a job with tiles of its own shares the width
            for (auto &kv : items[r][p]) {
                cq_batch_s::Seg sg;
                sg.val = kv.first;
                sg.lds = seg_lds[r][p][kv.first];
                if (kv.first == pass_tile(batch_tiles(b, p != P_AROW && p != P_ACOL), p))
                    sg.lds = std::max(sg.lds, b->lds[p]);
                for (int g = 0; g < kGroups; g++) {
                    sg.off[g] = (int)ord[r][p].size();
                    if (kv.second[g].empty()) continue;
                    const std::vector<int2> o = xcd_order(kv.second[g]);
                    ord[r][p].insert(ord[r][p].end(), o.begin(), o.end());
                }
                sg.off[kGroups] = (int)ord[r][p].size();
                b->segs[r][p].push_back(sg);
            }
            if (ord[r][p].empty()) continue;
            if (ord[r][p].size() > b->ditems_cap[r][p]) {
                if (b->pool) {
                    b->ditems_cap[r][p] = ord[r][p].size();
                    const int rc = pool_take(b, sizeof(int2) * b->ditems_cap[r][p], (void **)&b->ditems[r][p]);
                    if (rc) return rc;
                } else {
                    (void)hipFree(b->ditems[r][p]);
                    b->ditems[r][p] = nullptr;
                    b->ditems_cap[r][p] = std::max(ord[r][p].size(), 2 * b->ditems_cap[r][p]);
                    HIPCHK(hipMalloc(&b->ditems[r][p], sizeof(int2) * b->ditems_cap[r][p]));
                }
            }
            HIPCHK(hipMemcpyAsync(b->ditems[r][p], ord[r][p].data(), sizeof(int2) * ord[r][p].size(),
                                  hipMemcpyHostToDevice, s));
            b->nitems[r][p] = (int)ord[r][p].size();
        }
    // the host vectors must outlive the copies
    HIPCHK(hipStreamSynchronize(s));
    return CQ_OK;
}

// Candidate k_q_r2 tiles (instantiated widths).
static const int kTk2Cand[] = {16, 12, 8, 6, 4, 3, 2, 1};

template <size_t N>
static bool in_set(int v, const int (&set)[N])
{
    return std::find(set, set + N, v) != set + N;
}

// The caller's tile fields over the engine's choice (choose_tiles has run):
// each value must be an instantiated width and every window's transforms
// must fit it (a plan of the batch's variant, at most kLdsMax of LDS).
static int apply_tiles(cq_batch b, const std::vector<TileReq> &req, const cq_tiles *t)
{
    if (!t) return CQ_OK;
    if ((t->tz && !in_set(t->tz, kTzCand)) || (t->cw && !in_set(t->cw, kCwCand)) ||
        (t->cwq && !in_set(t->cwq, kCwCand)) || (t->tk && !in_set(t->tk, kTkCand)) ||
        (t->tk2 && !in_set(t->tk2, kTk2Cand)) || t->plan < -1 || t->plan > 1)
        return set_err(CQ_EINVAL, "tiles tz=%d cw=%d cwq=%d tk=%d tk2=%d plan=%d: a value outside its set", t->tz,
                       t->cw, t->cwq, t->tk, t->tk2, t->plan);
    if (t->tz) b->tz = t->tz;
    if (t->cw) b->cw = t->cw;
    if (t->cwq) b->cwq = t->cwq;
    if (t->tk) b->tk = t->tk;
    if (t->tk2) b->tk2 = t->tk2;
    if (t->plan >= 0) b->plan = t->plan;
    const int v = b->plan;
    auto fits = [&](int n, int lines) { return has_plan(n, lines * n, v) && lds_lines(n, lines) <= kLdsMax; };
    int ax = 0, az = 0, qx = 0, qz = 0;
    std::vector<int> qxs;
    for (size_t i = 0; i < req.size(); i++) {
        const TileReq &j = req[i];
        bool ok = fits(j.nxx, b->tz / 2) && fits(j.nzz, (j.q ? b->cwq : b->cw) / 2);
        if (j.q) ok = ok && r1_ok(j.nxx, b->tk, kLdsMax, v) && fits(j.nxx, b->tk2);
        if (!ok)
            return set_err(CQ_EUNSUPPORTED, "job %d: window %dx%d does not fit tiles tz=%d cw=%d cwq=%d tk=%d tk2=%d "
                           "plan=%d", (int)i, j.nxx, j.nzz, b->tz, b->cw, b->cwq, b->tk, b->tk2, v);
        int &mx = j.q ? qx : ax, &mz = j.q ? qz : az;
        mx = std::max(mx, j.nxx);
        mz = std::max(mz, j.nzz);
        if (j.q) qxs.push_back(j.nxx);
    }
    set_lds(b, ax, az, qx, qz, qxs);
    return CQ_OK;
}

extern "C" int cq_batch_create(cq_ctx ctx, const cq_job_desc *jobs, int njobs, cq_batch *out)
{
    return cq_batch_create_tiled(ctx, jobs, njobs, nullptr, out);
}

extern "C" int cq_batch_get_tiles(cq_batch b, cq_tiles *out)
{
    if (!b || !out) return set_err(CQ_EINVAL, "cq_batch_get_tiles: bad arguments");
    out->tz = b->tz;
    out->cw = b->cw;
    out->cwq = b->cwq;
    out->tk = b->tk;
    out->tk2 = b->tk2;
    out->plan = b->plan;
    return CQ_OK;
}

extern "C" int cq_batch_create_tiled(cq_ctx ctx, const cq_job_desc *jobs, int njobs, const cq_tiles *tiles,
                                     cq_batch *out)
{
    if (!ctx || !jobs || njobs < 1 || !out) return set_err(CQ_EINVAL, "cq_batch_create: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    for (int i = 0; i < njobs; i++) {
        int rc = validate(jobs[i], i);
        if (rc) return rc;
    }
    std::unique_ptr<cq_batch_s> b(new cq_batch_s());
    b->ctx = ctx;
    std::vector<TileReq> req;
    for (int i = 0; i < njobs; i++) req.push_back({jobs[i].nxx, jobs[i].nzz, is_q(jobs[i].kind)});
    int rc = choose_tiles(b.get(), req);
    if (rc == CQ_OK) rc = apply_tiles(b.get(), req, tiles);
    if (rc) return rc;
    b->jobs.resize(njobs);
    for (int i = 0; i < njobs; i++) {
        rc = build_job(b.get(), jobs[i], b->jobs[i]);
        if (rc) return rc;
    }
    rc = rebuild_tables(b.get());
    if (rc) return rc;
    *out = b.release();
    ctx->batches.insert(*out);
    return CQ_OK;
}

// The dynamic-LDS limit is a per-kernel, per-device attribute shared by every
// context of the process: raise it when a launch needs more (one host call
// per kernel and size instead of one per launch); it never shrinks.
static hipError_t raise_lds_limit(int device, const void *kernel, size_t lds)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, size_t> limit;
    std::lock_guard<std::mutex> lock(mu);
    size_t &cur = limit[std::make_pair(device, kernel)];
    if (lds <= cur) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) cur = lds;
    return e;
}

template <typename K>
static int launch(cq_batch b, K kernel, const cq_batch_s::Seg &sg, int role, int pass, int step)
{
    const size_t lds = sg.lds;
    HIPCHK(raise_lds_limit(b->ctx->device, reinterpret_cast<const void *>(kernel), lds));
    if (b->dry) return CQ_OK;
    // one launch per group on its stream, or, unsplit, one for all groups
    for (int g = 0; g < kGroups; g++) {
        const int o = b->split ? sg.off[g] : sg.off[0];
        const int cnt = (b->split ? sg.off[g + 1] : sg.off[kGroups]) - o;
        if (cnt > 0) {
            hipStream_t st = b->split && g > 0 ? b->side[g] : b->ctx->stream;
            hipLaunchKernelGGL(kernel, dim3(cnt), dim3(CQ_NT), lds, st, b->djobs, b->ditems[role][pass] + o, step);
            HIPCHK(hipGetLastError());
        }
        if (!b->split) break;
    }
    return CQ_OK;
}

#define CQ_ARGS sg, role, pass, step
#define CQ_DISPATCH_TZ(K, T)                                                       \
    (T == 16   ? launch(b, K<16>, CQ_ARGS)                                         \
     : T == 12 ? launch(b, K<12>, CQ_ARGS)                                         \
     : T == 8  ? launch(b, K<8>, CQ_ARGS) : launch(b, K<4>, CQ_ARGS))
#define CQ_DISPATCH_CW(K, W)                                                       \
    (W == 32   ? launch(b, K<32>, CQ_ARGS)                                         \
     : W == 24 ? launch(b, K<24>, CQ_ARGS)                                         \
     : W == 16 ? launch(b, K<16>, CQ_ARGS)                                         \
     : W == 12 ? launch(b, K<12>, CQ_ARGS)                                         \
     : W == 8  ? launch(b, K<8>, CQ_ARGS)                                          \
     : W == 4  ? launch(b, K<4>, CQ_ARGS)                                          \
               : launch(b, K<2>, CQ_ARGS))
#define CQ_DISPATCH_TK(K, T)                                                       \
    (T == 16   ? launch(b, K<16>, CQ_ARGS)                                         \
     : T == 12 ? launch(b, K<12>, CQ_ARGS)                                         \
     : T == 8  ? launch(b, K<8>, CQ_ARGS)                                          \
     : T == 6  ? launch(b, K<6>, CQ_ARGS)                                          \
     : T == 4  ? launch(b, K<4>, CQ_ARGS)                                          \
     : T == 3  ? launch(b, K<3>, CQ_ARGS)                                          \
     : T == 2  ? launch(b, K<2>, CQ_ARGS) : launch(b, K<1>, CQ_ARGS))

static int launch_seg(cq_batch b, const cq_batch_s::Seg &sg, int pass, int role, int step)
{
    const int v = sg.val;
    switch (pass) {
    case P_AROW: return CQ_DISPATCH_TZ(k_row_ac, v);
    case P_ACOL: return CQ_DISPATCH_CW(k_col_ac, v);
    case P_QC1: return CQ_DISPATCH_CW(k_q_c1, v);
    case P_QR1:
        return v == 8   ? launch(b, k_q_r1<8>, CQ_ARGS)
               : v == 4 ? launch(b, k_q_r1<4>, CQ_ARGS)
               : v == 2 ? launch(b, k_q_r1<2>, CQ_ARGS) : launch(b, k_q_r1<1>, CQ_ARGS);
    case P_LR1:
        return v == 8   ? launch(b, k_lr_r1<8>, CQ_ARGS)
               : v == 4 ? launch(b, k_lr_r1<4>, CQ_ARGS)
               : v == 2 ? launch(b, k_lr_r1<2>, CQ_ARGS) : launch(b, k_lr_r1<1>, CQ_ARGS);
    case P_QC2: return CQ_DISPATCH_CW(k_q_c2, v);
    case P_LR2: return CQ_DISPATCH_CW(k_lr_c2, v);
    case P_QR2: return CQ_DISPATCH_TK(k_q_r2, v);
    case P_QC3: return CQ_DISPATCH_CW(k_q_c3, v);
    case P_QR3: return CQ_DISPATCH_TZ(k_q_r3, v);
    }
    return set_err(CQ_EINVAL, "unknown pass %d", pass);
}

// Every launch segment of a pass (one per tile width among the jobs).
static int launch_pass(cq_batch b, int pass, int role, int step)
{
    for (const cq_batch_s::Seg &sg : b->segs[role][pass]) {
        const int rc = launch_seg(b, sg, pass, role, step);
        if (rc) return rc;
    }
    return CQ_OK;
}

// Global step `step` of every wavefield of the batch: one launch per pass
// and role, for the roles that advance at this step (a modeling shot's one
// wavefield while it runs; an RTM shot's source or receiver wavefield as its
// schedule says).
// Bit r of the result: some wavefield of role r advances at global step t.
static int active_roles(cq_batch b, int t)
{
    int m = 0;
    for (const JobH &h : b->jobs) {
        const int s = t - h.t0;
        if (s < 0 || s >= h.nsched) continue;
        m |= 1 << (h.hsched.empty() ? 0 : h.hsched[s].x);
    }
    return m;
}

// The launches of one global step for the roles in `mask`; `step` is passed
// to the kernels (-1 in a captured graph: each group's counter).
static int step_launches(cq_batch b, int step, int mask)
{
    for (int r = 0; r < 2; r++) {
        if (!(mask >> r & 1)) continue;
        for (int p = 0; p < P_COUNT; p++) {
            if (b->nitems[r][p] == 0) continue;
            int rc = launch_pass(b, p, r, step);
            if (rc) return rc;
        }
    }
    return CQ_OK;
}

static int one_step(cq_batch b, int step)
{
    const int mask = active_roles(b, step);
    const bool active[2] = {(mask & 1) != 0, (mask & 2) != 0};
    if (!b->timing) return step_launches(b, step, mask);
    if (b->timing) {
        hipEvent_t e0;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventRecord(e0, b->ctx->stream));
        b->ev.push_back(e0);
        b->ev_pass.push_back(-1);
    }
    for (int r = 0; r < 2; r++) {
        if (!active[r]) continue;
        for (int p = 0; p < P_COUNT; p++) {
            if (b->nitems[r][p] == 0) continue;
            int rc = launch_pass(b, p, r, step);
            if (rc) return rc;
            if (b->timing) {
                hipEvent_t e1;
                HIPCHK(hipEventCreate(&e1));
                HIPCHK(hipEventRecord(e1, b->ctx->stream));
                b->ev.push_back(e1);
                b->ev_pass.push_back(p);
            }
        }
    }
    return CQ_OK;
}

__global__ void k_add_point(float *p, float v) { *p += v; }

// One checkpoint copy of an RTM shot: dir 1 state -> slot, 0 slot -> state
// (storecheckpoints, sgpstd-qrtm/acps4.cu:1797-1844).  A slot taken at capo 0
// also gets the reference's wavelet sample rik[0] dt 1000 added at the
// source (addsource, :1846-1863).
static int ckpt_apply(cq_batch b, const JobH &h, const CkptOp &op)
{
    // on the stream of the job's launch group: ordered with its own passes,
    // independent of the other groups
    hipStream_t s = b->split && h.grp > 0 ? b->side[h.grp] : b->ctx->stream;
    char *ck = h.ckpt + (size_t)op.slot * h.state_bytes;
    if (op.dir) {
        HIPCHK(hipMemcpyAsync(ck, h.state[0], h.state_bytes, hipMemcpyDeviceToDevice, s));
        if (op.src0) {
            const size_t off = (size_t)((char *)h.d[0].p - h.state[0]);
            hipLaunchKernelGGL(k_add_point, dim3(1), dim3(1), 0, s,
                               reinterpret_cast<float *>(ck + off) + h.src_cell, h.src0);
            HIPCHK(hipGetLastError());
        }
    } else {
        HIPCHK(hipMemcpyAsync(h.state[0], ck, h.state_bytes, hipMemcpyDeviceToDevice, s));
    }
    return CQ_OK;
}

// Run global steps b->t .. end - 1; before each, the checkpoint copies that
// the RTM shots' schedules place there.
// Start (fork) or end (join) the side streams of launch groups 1 ..
// kGroups - 1.  Per-launch timing keeps one stream so that its events
// bracket single passes.
static int make_side(cq_batch b)
{
    if (!b->fork_ev) {
        HIPCHK(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
        for (int g = 1; g < kGroups; g++) {
            HIPCHK(hipStreamCreate(&b->side[g]));
            HIPCHK(hipEventCreateWithFlags(&b->join_ev[g], hipEventDisableTiming));
        }
    }
    return CQ_OK;
}
static int fork_side(cq_batch b)
{
    if (b->timing || b->pool) return CQ_OK;
    const int rc = make_side(b);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b->fork_ev, b->ctx->stream));
    for (int g = 1; g < kGroups; g++) HIPCHK(hipStreamWaitEvent(b->side[g], b->fork_ev, 0));
    b->split = true;
    return CQ_OK;
}
static int join_side(cq_batch b)
{
    if (!b->split) return CQ_OK;
    b->split = false;
    for (int g = 1; g < kGroups; g++) {
        HIPCHK(hipEventRecord(b->join_ev[g], b->side[g]));
        HIPCHK(hipStreamWaitEvent(b->ctx->stream, b->join_ev[g], 0));
    }
    return CQ_OK;
}

__global__ void k_step_inc(int *c) { c[0] += 1; }
__global__ void k_set_steps(int *c, int v)
{
    if (threadIdx.x < kGroups) c[threadIdx.x] = v;
}

// The executable graph of graph_k global steps for the roles in `mask`
// (captured on first use): fork to the group streams, per step each group's
// passes (step -1: the group's counter) and its counter increment, join.
// Null (and graphs switched off, with the reason in the last error) if the
// runtime refuses the capture.
static hipGraphExec_t step_graph(cq_batch b, int mask)
{
    auto it = b->graphs.find(mask);
    if (it != b->graphs.end()) return it->second;
    hipStream_t s = b->ctx->stream;
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    // streams, events and the passes' LDS limits are set up outside the capture
    int rc = make_side(b);
    if (rc == CQ_OK) {
        b->dry = true;
        rc = step_launches(b, -1, mask);
        b->dry = false;
    }
    if (rc) return nullptr;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
        set_err(CQ_EDEVICE, "hipStreamBeginCapture failed");
        return nullptr;
    }
    bool used[kGroups] = {};
    for (const JobH &h : b->jobs) used[h.grp] = true;
    rc = fork_side(b);
    for (int k = 0; rc == CQ_OK && k < b->graph_k; k++) {
        rc = step_launches(b, -1, mask);
        for (int gi = 0; rc == CQ_OK && gi < kGroups; gi++) {
            if (!used[gi]) continue;   // no job reads this counter
            hipStream_t st = gi > 0 ? b->side[gi] : s;
            hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(1), 0, st, b->dstep + gi);
            if (hipGetLastError() != hipSuccess) rc = set_err(CQ_EDEVICE, "step counter launch failed");
        }
    }
    if (rc == CQ_OK) rc = join_side(b);
    b->split = false;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc == CQ_OK && e == hipSuccess && g) {
        if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
    }
    if (g) (void)hipGraphDestroy(g);
    if (!ex) {
        if (rc == CQ_OK) set_err(CQ_EDEVICE, "graph capture of the step launches failed");
        b->graph_k = 0;
        return nullptr;
    }
    b->graphs[mask] = ex;
    return ex;
}

// Whether the global steps t .. t + k - 1 can run as one graph: no
// checkpoint copy inside them (after their first step) and the same roles
// active throughout.
static bool graph_span(cq_batch b, int t, int k, int &mask)
{
    mask = active_roles(b, t);
    if (!mask) return false;
    for (int u = t + 1; u < t + k; u++)
        if (active_roles(b, u) != mask) return false;
    for (const JobH &h : b->jobs)
        for (size_t o = h.op_next; o < h.ops.size(); o++) {
            const int at = h.ops[o].at + h.t0;
            if (at >= t + k) break;
            if (at > t) return false;
        }
    return true;
}

static int advance_steps(cq_batch b, int end)
{
    while (b->t < end) {
        for (auto &h : b->jobs) {
            const int s = b->t - h.t0;
            while (h.op_next < h.ops.size() && h.ops[h.op_next].at <= s) {
                if (h.ops[h.op_next].at == s) {
                    int rc = ckpt_apply(b, h, h.ops[h.op_next]);
                    if (rc) return rc;
                }
                h.op_next++;
            }
        }
        int mask = 0;
        if (b->graph_k > 0 && !b->timing && !b->pool && end - b->t >= b->graph_k &&
            graph_span(b, b->t, b->graph_k, mask)) {
            // the graph orders after the context stream only: the groups'
            // direct launches join it first (and before any capture)
            int rc = join_side(b);
            if (rc) return rc;
            hipGraphExec_t ex = step_graph(b, mask);
            if (ex) {
                if (b->dstep_t != b->t) {
                    hipLaunchKernelGGL(k_set_steps, dim3(1), dim3(64), 0, b->ctx->stream, b->dstep, b->t);
                    HIPCHK(hipGetLastError());
                }
                HIPCHK(hipGraphLaunch(ex, b->ctx->stream));
                b->t += b->graph_k;
                b->dstep_t = b->t;
                continue;
            }
        }
        if (!b->split) {
            int rc = fork_side(b);
            if (rc) return rc;
        }
        int rc = one_step(b, b->t);
        if (rc) return rc;
        b->t++;
    }
    return CQ_OK;
}

// Global steps b->t .. end - 1; on return all work is ordered on the
// context stream.
static int advance_to(cq_batch b, int end)
{
    const int rc = advance_steps(b, end);
    const int rj = join_side(b);
    return rc ? rc : rj;
}

extern "C" int cq_batch_step(cq_batch b, int nsteps)
{
    if (!b || nsteps < 0) return set_err(CQ_EINVAL, "cq_batch_step: bad arguments");
    HIPCHK(hipSetDevice(b->ctx->device));
    return advance_to(b, std::min(b->tmax, b->t + nsteps));
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
    HIPCHK(hipSetDevice(b->ctx->device));
    int rc = advance_to(b, b->tmax);
    if (rc) return rc;
    return cq_batch_sync(b);
}

// Copies a field (device field layout) to host memory as f[ix * nzz + iz]
// over ix in [x0, x0 + nx), iz in [z0, z0 + nz) (x-major, z fastest).
static int read_field(const JobH &h, const float *dev, int x0, int nx, int z0, int nz, float *out)
{
    std::vector<float> f((size_t)h.nxx * h.ldz);
    HIPCHK(hipMemcpy(f.data(), dev, sizeof(float) * f.size(), hipMemcpyDeviceToHost));
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            const int gx = x0 + ix, gz = z0 + iz;
            out[(size_t)ix * nz + iz] = f[cq_q4(h.nxx, gx, gz >> 2) + (gz & 3)];
        }
    return CQ_OK;
}

// Index of the first non-finite value of a[0 .. n), or -1.
static long first_nonfinite(const float *a, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(a[i])) return (long)i;
    return -1;
}

static int read_image(cq_ctx ctx, const JobH &h, float *img, float *ssg)
{
    if (!is_rtm(h.kind)) return set_err(CQ_ESTATE, "not an RTM shot");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int pml = (h.nxx - h.nx) / 2;
    int rc = read_field(h, h.d[0].img, pml, h.nx, pml, h.nz, img);
    if (rc == CQ_OK) rc = read_field(h, h.d[0].ssg, pml, h.nx, pml, h.nz, ssg);
    if (rc == CQ_OK) {
        const size_t n = (size_t)h.nx * h.nz;
        if (first_nonfinite(img, n) >= 0 || first_nonfinite(ssg, n) >= 0)
            return set_err(CQ_ENUMERIC, "RTM shot: non-finite image or illumination (the wavefield blew up)");
    }
    return rc;
}

static int read_rec(cq_ctx ctx, const JobH &h, float *rec)
{
    if (is_rtm(h.kind)) return set_err(CQ_ESTATE, "RTM shots have no gathers");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (h.nr > 0) {
        HIPCHK(hipMemcpy(rec, h.d[0].rec, sizeof(float) * (size_t)h.nt * h.nr, hipMemcpyDeviceToHost));
        const long i = first_nonfinite(rec, (size_t)h.nt * h.nr);
        if (i >= 0)
            return set_err(CQ_ENUMERIC, "non-finite gather value at step %ld, receiver %ld (the wavefield blew up)",
                           i / h.nr, i % h.nr);
    }
    return CQ_OK;
}

extern "C" int cq_batch_read_image(cq_batch b, int job, float *img, float *ssg)
{
    if (!b || job < 0 || job >= (int)b->jobs.size() || !img || !ssg)
        return set_err(CQ_EINVAL, "cq_batch_read_image: bad arguments");
    return read_image(b->ctx, b->jobs[job], img, ssg);
}

extern "C" int cq_batch_read_rec(cq_batch b, int job, float *rec)
{
    if (!b || job < 0 || job >= (int)b->jobs.size() || !rec) return set_err(CQ_EINVAL, "cq_batch_read_rec: bad arguments");
    return read_rec(b->ctx, b->jobs[job], rec);
}

extern "C" int cq_batch_read_p(cq_batch b, int job, float *p)
{
    if (!b || job < 0 || job >= (int)b->jobs.size() || !p) return set_err(CQ_EINVAL, "cq_batch_read_p: bad arguments");
    const JobH &h = b->jobs[job];
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return read_field(h, h.d[0].p, 0, h.nxx, 0, h.nzz, p);
}

extern "C" int cq_batch_time(cq_batch b) { return b ? b->t : -1; }

extern "C" long long cq_batch_bytes(cq_batch b)
{
    if (!b) return -1;
    long long s = 0;
    for (auto &j : b->jobs) s += (long long)j.bytes;
    return s;
}

// ---------------------------------------------------------------------------
// serving: continuous batching over one device-resident batch
// ---------------------------------------------------------------------------
// A submitted job owns copies of all its host arrays until it is admitted.
struct ServerJob {
    int id = 0, state = CQ_JOB_QUEUED;
    cq_job_desc d;
    std::vector<float> f[32];
    std::vector<int> gx, gz;
    int t_submit = 0, t_start = -1, t_end = -1;
    size_t bytes = 0;
    std::vector<float> out, ssg;   // results (gather, or image + illumination)
    int err = CQ_OK;               // status of a failed job
    std::string msg;               // and its reason
};

struct cq_server_s {
    cq_batch_s b;
    cq_server_desc desc;
    size_t budget = 0, resident = 0;
    int next_id = 1;
    std::map<int, std::unique_ptr<ServerJob>> jobs;
    std::vector<int> queue;         // ids waiting for admission, FIFO
    std::vector<int> running;       // ids of b.jobs, same order
    std::vector<char> blown;        // per running job: non-finite wavefield found
    int last_check = 0;             // global step of the last non-finite check
    int *dflag = nullptr;           // device flags of that check, one per wavefield
    size_t dflag_cap = 0;
    std::set<int> fail_admit;       // CQ_FAULT_ADMIT: ids whose admission is made to fail
    // run_cells > 0: resident jobs outside the run set, their ids and the
    // steps each had taken when it stopped (its device memory stays)
    std::vector<JobH> paused;
    std::vector<int> paused_ids, paused_done;
    ~cq_server_s()
    {
        (void)hipFree(dflag);
        for (auto &h : paused) (void)hipFree(h.mem);
    }
};

// Global steps between two checks of the running jobs' wavefields for
// non-finite values (one read of each pressure field: 4 B per cell every
// kCheckEvery steps).
static const int kCheckEvery = 256;

// Tile shapes fixed for every window up to max_nxx x max_nzz: the largest
// tiles whose LDS fits the budget and whose point count fits the smallest
// per-stage capacity of the power-of-two radices.
static int server_tiles(cq_batch b, int mx, int mz)
{
    int iz = 0, ic = 0, ik = 0;
    // tiles of `lines` lines of the largest window's length n: bad if over
    // the LDS budget or if no plan holds that many points (a job with a
    // shorter length whose plan does not fit is refused when submitted)
    auto lim = [&](int n, int lines, size_t budget) { return lds_lines(n, lines) > budget || !has_plan(n, lines * n); };
    auto r1_bad = [&](int k, size_t budget) { return lim(mx, 3 * k, budget) && lim(mx, 2 * k, budget); };
    while (iz < 3 && lim(mx, kTzCand[iz] / 2, kLdsTile)) iz++;
    while (ic < 6 && lim(mz, kCwCand[ic] / 2, kLdsTile)) ic++;
    while (ik < 3 && r1_bad(kTkCand[ik], kLdsTile)) ik++;
    b->tz = kTzCand[iz];
    b->cw = kCwCand[ic];
    b->cwq = kCwCand[ic];
    b->tk = kTkCand[ik];
    if (lim(mx, b->tz / 2, kLdsMax) || lim(mz, b->cw / 2, kLdsMax) || r1_bad(b->tk, kLdsMax))
        return set_err(CQ_EUNSUPPORTED, "server windows up to %dx%d are too large for one workgroup", mx, mz);
    b->tk2 = b->tk;
    for (int m : {3, 2})
        if (m * b->tk <= 16 && !lim(mx, m * b->tk, kLdsTile)) {
            b->tk2 = m * b->tk;
            break;
        }
    b->maxnxx = mx;
    b->maxnzz = mz;
    b->own_tiles = true;
    set_lds(b, mx, mz, mx, mz, {});
    // any job up to mx: fused k_q_r1 within the tile budget, or split
    b->lds[P_QR1] = std::max(std::min(kLdsTile, lds_lines(mx, 3 * b->tk)), lds_lines(mx, 2 * b->tk));
    return CQ_OK;
}

extern "C" int cq_server_open(cq_ctx ctx, const cq_server_desc *desc, cq_server *out)
{
    if (!ctx || !desc || !out || desc->max_nxx < 8 || desc->max_nzz < 8 || desc->hbm_budget < 0 || desc->max_jobs < 0)
        return set_err(CQ_EINVAL, "cq_server_open: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    std::unique_ptr<cq_server_s> s(new cq_server_s());
    s->b.ctx = ctx;
    s->desc = *desc;
    int rc = server_tiles(&s->b, desc->max_nxx, desc->max_nzz);
    if (rc) return rc;
    if (desc->hbm_budget > 0) {
        s->budget = (size_t)desc->hbm_budget;
    } else {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        s->budget = fr / 10 * 9;
    }
    // fault injection for tests of the admission-failure path: a comma list
    // of job ids (in submission order, from 1) whose admission fails as if
    // their device allocation had
    if (const char *fa = getenv("CQ_FAULT_ADMIT")) {
        for (const char *c = fa; *c;) {
            char *e = nullptr;
            const long v = strtol(c, &e, 10);
            if (e == c) break;
            s->fail_admit.insert((int)v);
            c = (*e == ',') ? e + 1 : e;
        }
    }
    *out = s.release();
    ctx->servers.insert(*out);
    return CQ_OK;
}

// Deep copy of a job description (arrays the engine reads).
static void copy_job(ServerJob &sj, const cq_job_desc &j)
{
    sj.d = j;
    const size_t nc = (size_t)j.nxx * j.nzz;
    int k = 0;
    auto own = [&](const float *&p, size_t n) {
        if (!p) return;
        sj.f[k].assign(p, p + n);
        p = sj.f[k].data();
        k++;
    };
    own(sj.d.vp, nc); own(sj.d.rho, nc); own(sj.d.bbx, nc); own(sj.d.bbz, nc);
    own(sj.d.gam, nc); own(sj.d.lam, nc); own(sj.d.vwd, nc); own(sj.d.tao, nc);
    own(sj.d.iax, j.nxx); own(sj.d.ibx, j.nxx); own(sj.d.hax, j.nxx); own(sj.d.hbx, j.nxx);
    own(sj.d.iaz, j.nzz); own(sj.d.ibz, j.nzz); own(sj.d.haz, j.nzz); own(sj.d.hbz, j.nzz);
    own(sj.d.kxsmr, j.nxx); own(sj.d.kxsmi, j.nxx); own(sj.d.kxscr, j.nxx); own(sj.d.kxsci, j.nxx);
    own(sj.d.kzsmr, j.nzz); own(sj.d.kzsmi, j.nzz); own(sj.d.kzscr, j.nzz); own(sj.d.kzsci, j.nzz);
    own(sj.d.rik, (size_t)j.nt);
    if (is_q(j.kind) && j.lr_rank > 0) {
        own(sj.d.lr_w, (size_t)j.lr_rank * nc);
        own(sj.d.lr_sym, (size_t)j.lr_rank * (j.nxx / 2 + 1) * (j.nzz / 2 + 1));
    }
    if (is_rtm(j.kind) && j.nr > 0) own(sj.d.obs, (size_t)j.nt * j.nr);
    else sj.d.obs = nullptr;
    if (j.nr > 0) {
        sj.gx.assign(j.gxg, j.gxg + j.nr);
        sj.gz.assign(j.gzg, j.gzg + j.nr);
        sj.d.gxg = sj.gx.data();
        sj.d.gzg = sj.gz.data();
    }
}

extern "C" int cq_server_submit(cq_server s, const cq_job_desc *job, int *id)
{
    if (!s || !job || !id) return set_err(CQ_EINVAL, "cq_server_submit: bad arguments");
    int rc = validate(*job, 0);
    if (rc) return rc;
    if (job->nxx > s->desc.max_nxx || job->nzz > s->desc.max_nzz)
        return set_err(CQ_EUNSUPPORTED, "window %dx%d exceeds the server's %dx%d", job->nxx, job->nzz,
                       s->desc.max_nxx, s->desc.max_nzz);
    HIPCHK(hipSetDevice(s->b.ctx->device));
    std::unique_ptr<ServerJob> sj(new ServerJob());
    copy_job(*sj, *job);
    JobH probe;
    rc = build_job(&s->b, sj->d, probe, &sj->bytes);   // plans exist, footprint
    if (rc) return rc;
    if (sj->bytes > s->budget)
        return set_err(CQ_ENOMEM, "job needs %zu device bytes, the server's budget is %zu", sj->bytes, s->budget);
    sj->id = s->next_id++;
    sj->t_submit = s->b.t;
    *id = sj->id;
    s->queue.push_back(sj->id);
    s->jobs[sj->id] = std::move(sj);
    return CQ_OK;
}

__global__ __launch_bounds__(256) void k_nonfinite(const float *__restrict__ f, long long n, int *__restrict__ flag)
{
    int bad = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        bad |= !isfinite(f[i]);
    if (__syncthreads_or(bad) && threadIdx.x == 0) *flag = 1;
}

// Marks the running jobs whose pressure field (either wavefield of an RTM
// shot) holds a non-finite value; server_turnover retires them as failed.
static int check_nonfinite(cq_server s)
{
    cq_batch b = &s->b;
    s->last_check = b->t;
    const size_t n = 2 * b->jobs.size();
    if (n == 0) return CQ_OK;
    if (n > s->dflag_cap) {
        (void)hipFree(s->dflag);
        s->dflag = nullptr;
        s->dflag_cap = std::max(n, 2 * s->dflag_cap);
        HIPCHK(hipMalloc(&s->dflag, sizeof(int) * s->dflag_cap));
    }
    hipStream_t st = b->ctx->stream;
    HIPCHK(hipMemsetAsync(s->dflag, 0, sizeof(int) * n, st));
    for (size_t i = 0; i < b->jobs.size(); i++) {
        const JobH &h = b->jobs[i];
        for (int r = 0; r < h.nroles; r++) {
            hipLaunchKernelGGL(k_nonfinite, dim3(64), dim3(256), 0, st, (const float *)h.d[r].p,
                               (long long)h.nxx * h.ldz, s->dflag + 2 * i + r);
            HIPCHK(hipGetLastError());
        }
    }
    std::vector<int> f(n);
    HIPCHK(hipMemcpyAsync(f.data(), s->dflag, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    s->blown.assign(b->jobs.size(), 0);
    for (size_t i = 0; i < b->jobs.size(); i++) s->blown[i] = (char)(f[2 * i] | f[2 * i + 1]);
    return CQ_OK;
}

// Work cells of a resident job (cq_server_desc::run_cells): cells per step,
// viscoacoustic cells counted twice (their step runs six passes to two).
static double work_cells(const JobH &h) { return (double)h.nxx * h.nzz * (is_q(h.kind) ? 2.0 : 1.0); }

// run_cells > 0: the resident jobs with the least remaining work advance, as
// many as fit run_cells work cells (at least one); the others are paused
// (taken out of the batch, their state kept on the device) and resume with
// their start step moved so that their local step continues where it
// stopped.  Sets `changed` when the batch's jobs changed; the caller
// rebuilds the tables.  Nothing may be in flight on the stream, and the
// running jobs' non-finite flags have been acted on (server_turnover
// retires blown jobs before it chooses).
static void server_reschedule(cq_server s, bool &changed)
{
    if (s->desc.run_cells <= 0) return;
    cq_batch b = &s->b;
    struct Cand {
        bool paused;
        size_t idx;
        int id;
        double work, cells;
    };
    std::vector<Cand> c;
    for (size_t i = 0; i < b->jobs.size(); i++) {
        const JobH &h = b->jobs[i];
        c.push_back({false, i, s->running[i], (double)(h.nsched - (b->t - h.t0)) * work_cells(h), work_cells(h)});
    }
    for (size_t i = 0; i < s->paused.size(); i++) {
        const JobH &h = s->paused[i];
        c.push_back({true, i, s->paused_ids[i], (double)(h.nsched - s->paused_done[i]) * work_cells(h),
                     work_cells(h)});
    }
    std::sort(c.begin(), c.end(), [](const Cand &a, const Cand &z) {
        return a.work != z.work ? a.work < z.work : a.id < z.id;
    });
    std::vector<char> run(c.size(), 0);
    double sum = 0;
    for (size_t k = 0; k < c.size(); k++)
        if (k == 0 || sum < (double)s->desc.run_cells) {
            run[k] = 1;
            sum += c[k].cells;
        }
    bool any = false;
    for (size_t k = 0; k < c.size(); k++) any = any || (run[k] != 0) == c[k].paused;
    if (!any) return;
    // new running and paused sets, in the order chosen
    std::vector<JobH> nj, np;
    std::vector<int> nr, npi, npd;
    std::vector<char> nb;
    for (size_t k = 0; k < c.size(); k++) {
        if (run[k]) {
            if (c[k].paused) {
                JobH &h = s->paused[c[k].idx];
                h.t0 = b->t - s->paused_done[c[k].idx];
                for (int r = 0; r < 2; r++) h.d[r].t0 = h.t0;
                nj.push_back(std::move(h));
                nb.push_back(0);
            } else {
                nj.push_back(std::move(b->jobs[c[k].idx]));
                nb.push_back(s->blown.size() > c[k].idx ? s->blown[c[k].idx] : 0);
            }
            nr.push_back(c[k].id);
        } else {
            if (c[k].paused) {
                npd.push_back(s->paused_done[c[k].idx]);
                np.push_back(std::move(s->paused[c[k].idx]));
            } else {
                JobH &h = b->jobs[c[k].idx];
                npd.push_back(b->t - h.t0);
                np.push_back(std::move(h));
            }
            npi.push_back(c[k].id);
        }
    }
    b->jobs = std::move(nj);
    s->running = std::move(nr);
    s->blown = std::move(nb);
    s->paused = std::move(np);
    s->paused_ids = std::move(npi);
    s->paused_done = std::move(npd);
    changed = true;
}

static void fail_job(ServerJob &sj, int err, int t)
{
    sj.state = CQ_JOB_FAILED;
    sj.err = err;
    sj.msg = g_err;
    sj.t_end = t;
    sj.out.clear();
    sj.ssg.clear();
    for (auto &v : sj.f) std::vector<float>().swap(v);
}

// Retires finished jobs (results copied to the host) and jobs whose
// wavefield blew up (failed, CQ_ENUMERIC), freeing their device memory; then
// admits queued jobs (FIFO) that fit the budget and the job limit.  A job
// whose admission fails is marked failed and the next one is tried.  The
// stream is idle on return when anything changed, and the tables are rebuilt.
// Only HIP runtime errors are returned.
static int server_turnover(cq_server s)
{
    cq_batch b = &s->b;
    bool changed = false, synced = false;
    s->blown.resize(b->jobs.size(), 0);
    for (size_t i = 0; i < b->jobs.size();) {
        JobH &h = b->jobs[i];
        const bool blown = s->blown[i] != 0;
        if (!blown && b->t < h.t0 + h.nsched) {
            i++;
            continue;
        }
        ServerJob &sj = *s->jobs[s->running[i]];
        int rc = CQ_OK;
        if (blown) {
            rc = set_err(CQ_ENUMERIC, "job %d: non-finite wavefield by global step %d (the wavefield blew up)", sj.id,
                         b->t);
        } else if (is_rtm(h.kind)) {
            sj.out.resize((size_t)h.nx * h.nz);
            sj.ssg.resize((size_t)h.nx * h.nz);
            rc = read_image(b->ctx, h, sj.out.data(), sj.ssg.data());
        } else {
            sj.out.resize((size_t)h.nt * h.nr);
            rc = read_rec(b->ctx, h, sj.out.data());
        }
        if (rc == CQ_ENUMERIC) fail_job(sj, rc, b->t);
        else if (rc) return rc;
        else {
            sj.state = CQ_JOB_DONE;
            sj.t_end = b->t;
        }
        if (!synced) HIPCHK(hipStreamSynchronize(b->ctx->stream));
        synced = true;
        HIPCHK(hipFree(h.mem));
        s->resident -= sj.bytes;
        b->jobs.erase(b->jobs.begin() + i);
        s->running.erase(s->running.begin() + i);
        s->blown.erase(s->blown.begin() + i);
        changed = true;
    }
    while (!s->queue.empty()) {
        ServerJob &sj = *s->jobs[s->queue.front()];
        if (s->resident + sj.bytes > s->budget) break;
        if (s->desc.max_jobs > 0 && (int)b->jobs.size() >= s->desc.max_jobs) break;
        if (!synced) HIPCHK(hipStreamSynchronize(b->ctx->stream));
        synced = true;
        b->jobs.emplace_back();
        JobH &h = b->jobs.back();
        int rc = s->fail_admit.count(sj.id)
                     ? set_err(CQ_ENOMEM, "job %d: device allocation failed (injected by CQ_FAULT_ADMIT)", sj.id)
                     : build_job(b, sj.d, h);
        s->queue.erase(s->queue.begin());
        if (rc) {
            // the job leaves the queue as failed; what build_job had
            // allocated is returned, and the next job is tried
            (void)hipStreamSynchronize(b->ctx->stream);
            if (h.mem) (void)hipFree(h.mem);
            b->jobs.pop_back();
            fail_job(sj, rc, b->t);
            continue;
        }
        h.t0 = b->t;
        for (int r = 0; r < h.nroles; r++) h.d[r].t0 = b->t;
        s->resident += sj.bytes;
        s->running.push_back(sj.id);
        s->blown.push_back(0);
        sj.state = CQ_JOB_RUNNING;
        sj.t_start = b->t;
        for (auto &v : sj.f) std::vector<float>().swap(v);   // inputs now live on the device
        changed = true;
    }
    // run_cells > 0: the run set is chosen again at every turnover (a job
    // finished or was admitted, a non-finite check, the end of an advance)
    if (s->desc.run_cells > 0 && (changed || !s->paused.empty())) {
        if (!synced) HIPCHK(hipStreamSynchronize(b->ctx->stream));
        server_reschedule(s, changed);
    }
    if (changed) return rebuild_tables(b);
    return CQ_OK;
}

extern "C" int cq_server_advance(cq_server s, int nsteps, int *running, int *queued)
{
    if (!s || nsteps < 0) return set_err(CQ_EINVAL, "cq_server_advance: bad arguments");
    cq_batch b = &s->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    const int end = b->t + nsteps;
    int rc = server_turnover(s);
    while (rc == CQ_OK && b->t < end && !b->jobs.empty()) {
        // run to the next finishing job, non-finite check or the end, then
        // turn over
        int next = std::min(end, s->last_check + kCheckEvery);
        for (auto &h : b->jobs) next = std::min(next, h.t0 + h.nsched);
        rc = advance_to(b, std::max(next, b->t + 1));
        if (rc == CQ_OK && b->t >= s->last_check + kCheckEvery) rc = check_nonfinite(s);
        if (rc == CQ_OK && b->t >= next) {
            HIPCHK(hipStreamSynchronize(b->ctx->stream));
            rc = server_turnover(s);
        }
    }
    if (rc) return rc;
    if (running) *running = (int)(b->jobs.size() + s->paused.size());
    if (queued) *queued = (int)s->queue.size();
    return CQ_OK;
}

static ServerJob *server_job(cq_server s, int id)
{
    auto it = s->jobs.find(id);
    return it == s->jobs.end() ? nullptr : it->second.get();
}

extern "C" int cq_server_poll(cq_server s, int id, int *state)
{
    ServerJob *sj = s ? server_job(s, id) : nullptr;
    if (!sj || !state) return set_err(CQ_EINVAL, "cq_server_poll: unknown job %d", id);
    *state = sj->state;
    return CQ_OK;
}

extern "C" int cq_server_job_steps(cq_server s, int id, int *submitted, int *started, int *finished)
{
    ServerJob *sj = s ? server_job(s, id) : nullptr;
    if (!sj) return set_err(CQ_EINVAL, "cq_server_job_steps: unknown job %d", id);
    if (submitted) *submitted = sj->t_submit;
    if (started) *started = sj->t_start;
    if (finished) *finished = sj->t_end;
    return CQ_OK;
}

extern "C" int cq_server_result(cq_server s, int id, float *out, float *ssg)
{
    ServerJob *sj = s ? server_job(s, id) : nullptr;
    if (!sj || !out) return set_err(CQ_EINVAL, "cq_server_result: bad arguments");
    if (sj->state == CQ_JOB_FAILED) return set_err(sj->err, "%s", sj->msg.c_str());
    if (sj->state != CQ_JOB_DONE) return set_err(CQ_ESTATE, "job %d has not finished", id);
    memcpy(out, sj->out.data(), sizeof(float) * sj->out.size());
    if (!sj->ssg.empty()) {
        if (!ssg) return set_err(CQ_EINVAL, "cq_server_result: RTM job %d needs ssg", id);
        memcpy(ssg, sj->ssg.data(), sizeof(float) * sj->ssg.size());
    }
    return CQ_OK;
}

extern "C" int cq_server_release(cq_server s, int id)
{
    ServerJob *sj = s ? server_job(s, id) : nullptr;
    if (!sj) return set_err(CQ_EINVAL, "cq_server_release: unknown job %d", id);
    if (sj->state != CQ_JOB_DONE && sj->state != CQ_JOB_FAILED) return set_err(CQ_ESTATE, "job %d has not finished", id);
    s->jobs.erase(id);
    return CQ_OK;
}

extern "C" int cq_server_close(cq_server s)
{
    if (!s) return CQ_OK;
    (void)hipSetDevice(s->b.ctx->device);
    (void)hipStreamSynchronize(s->b.ctx->stream);
    s->b.ctx->servers.erase(s);
    delete s;
    return CQ_OK;
}

extern "C" int cq_batch_destroy(cq_batch b)
{
    if (!b) return CQ_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    b->ctx->batches.erase(b);
    delete b;
    return CQ_OK;
}

static void destroy_children(cq_ctx ctx)
{
    while (!ctx->servers.empty()) cq_server_close(*ctx->servers.begin());
    while (!ctx->batches.empty()) cq_batch_destroy(*ctx->batches.begin());
}

extern "C" int cq_batch_use_graphs(cq_batch b, int steps_per_graph)
{
    if (!b || steps_per_graph < 0 || steps_per_graph > 64)
        return set_err(CQ_EINVAL, "cq_batch_use_graphs: steps_per_graph %d outside 0..64", steps_per_graph);
    if (b->pool) return set_err(CQ_EINVAL, "cq_batch_use_graphs: not for the static scratch batch");
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    b->drop_graphs();
    b->graph_k = steps_per_graph;
    // capture the graph of the roles active now (capture and instantiation
    // take milliseconds: done here, the first steps that replay it do not
    // pay for them); other role combinations are captured when first met
    if (b->graph_k > 0 && !b->jobs.empty()) {
        const int mask = active_roles(b, b->t);
        if (mask && !step_graph(b, mask) && b->graph_k == 0) return CQ_EDEVICE;
    }
    return CQ_OK;
}

extern "C" int cq_batch_enable_timing(cq_batch b, int on)
{
    if (!b) return set_err(CQ_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    for (auto e : b->ev) (void)hipEventDestroy(e);
    b->ev.clear();
    b->ev_pass.clear();
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
    const unsigned grid = (unsigned)((n4 + 1023) / 1024);
    hipLaunchKernelGGL(k_copy_f4, dim3(grid), dim3(256), 0, ctx->stream, a, b, n4);
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_f4, dim3(grid), dim3(256), 0, ctx->stream, a, b, n4);
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

extern "C" int cq_fft_plan(int n, int points, int variant, int *radices, int cap, int *inner)
{
    if (n < 1 || points < n || variant < 0 || variant > 1 || (cap > 0 && !radices) || cap < 0)
        return set_err(CQ_EINVAL, "cq_fft_plan: bad arguments");
    CqPlan p;
    if (!make_plan(n, points, p, variant))
        return set_err(CQ_EUNSUPPORTED, "no FFT plan for %d points (%d per workgroup, variant %d)", n, points,
                       variant);
    for (int s = 0; s < p.ns && s < cap; s++) radices[s] = p.rad[s];
    if (inner) *inner = p.bn ? p.n : 0;
    return p.ns;
}

extern "C" int cq_plain_forward(cq_ctx ctx, const cq_job_desc *job, float *rec)
{
    if (!ctx || !job || !rec) return set_err(CQ_EINVAL, "cq_plain_forward: bad arguments");
    if (is_rtm(job->kind))
        return set_err(CQ_EINVAL, "cq_plain_forward: modeling shots only (kind %d is RTM)", job->kind);
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
            hipError_t e = hipMemcpyAsync(pinned, h.d[0].rec + (size_t)t * h.nr, sizeof(float) * h.nr,
                                          hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) rc = set_err(CQ_EDEVICE, "plain gather copy: %s", hipGetErrorString(e));
            else memcpy(rec + (size_t)t * h.nr, pinned, sizeof(float) * h.nr);
            const long i = rc == CQ_OK ? first_nonfinite(pinned, (size_t)h.nr) : -1;
            if (i >= 0)
                rc = set_err(CQ_ENUMERIC, "non-finite gather value at step %d, receiver %ld (the wavefield blew up)",
                             t, i);
        }
    }
    if (pinned) (void)hipHostFree(pinned);
    cq_batch_destroy(b);
    return rc;
}

// The allocation-free variant of cq_plain_forward: the same passes with the
// same tile shapes (so the same results, bit for bit), every device buffer
// carved out of this device's static scratch.
extern "C" int cq_forward_static(cq_ctx ctx, const cq_job_desc *job, float *rec)
{
    if (!ctx || !job || !rec) return set_err(CQ_EINVAL, "cq_forward_static: bad arguments");
    if (is_rtm(job->kind))
        return set_err(CQ_EINVAL, "cq_forward_static: modeling shots only (kind %d is RTM)", job->kind);
    int rc = validate(*job, 0);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->scratch) return set_err(CQ_EDEVICE, "cq_forward_static: no static scratch on device %d", ctx->device);
    cq_batch_s b;
    b.ctx = ctx;
    rc = choose_tiles(&b, {TileReq{job->nxx, job->nzz, is_q(job->kind)}});
    if (rc) return rc;
    // the arena's size is known before anything is placed: refuse a shot
    // that cannot fit, with its size
    JobH probe;
    size_t need = 0;
    rc = build_job(&b, *job, probe, &need);
    if (rc) return rc;
    if (need > (size_t)CQ_STATIC_BYTES / 10 * 9)
        return set_err(CQ_EINVAL, "cq_forward_static: the shot needs %zu device bytes; the static scratch holds %zu",
                       need, (size_t)CQ_STATIC_BYTES);
    b.pool = ctx->scratch;
    b.pool_cap = CQ_STATIC_BYTES;
    // the previous call's contents are still in the scratch: build_job clears
    // the arena (state, checkpoints, gathers) on the stream before any pass
    b.jobs.resize(1);
    rc = build_job(&b, *job, b.jobs[0]);
    if (rc == CQ_OK) rc = rebuild_tables(&b);
    if (rc == CQ_OK) rc = advance_to(&b, b.tmax);
    if (rc == CQ_OK) rc = read_rec(ctx, b.jobs[0], rec);
    (void)hipStreamSynchronize(ctx->stream);
    return rc;
}

// ---------------------------------------------------------------------------
// autotuner (cq_autotune)
// ---------------------------------------------------------------------------
// Cache key of a job set: library version, device architecture and CU count,
// then every window class (a: acoustic or RTM, q: viscoacoustic or Q-RTM,
// l<r>: low-rank of rank r) with its padded size and count, sorted.
static std::string tune_key(cq_ctx ctx, const cq_job_desc *jobs, int njobs)
{
    std::map<std::string, int> win;
    char buf[64];
    for (int i = 0; i < njobs; i++) {
        const cq_job_desc &j = jobs[i];
        if (!is_q(j.kind)) snprintf(buf, sizeof buf, "a%dx%d", j.nxx, j.nzz);
        else if (j.lr_rank > 0) snprintf(buf, sizeof buf, "l%d:%dx%d", j.lr_rank, j.nxx, j.nzz);
        else snprintf(buf, sizeof buf, "q%dx%d", j.nxx, j.nzz);
        win[buf]++;
    }
    std::string arch = "unknown";
    int cus = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) {
        arch = prop.gcnArchName;
        cus = prop.multiProcessorCount;
    }
    std::string key = std::string(CQ_VERSION) + "|" + arch + "|" + std::to_string(cus) + "|";
    for (auto &kv : win) key += kv.first + "*" + std::to_string(kv.second) + ",";
    for (char &c : key)
        if (c == ' ' || c == '\t' || c == '\n' || c == '\r') c = '_';
    return key;
}

// Cache file: one line per tuned job set, "key<TAB>tz cw cwq tk tk2 plan ms";
// '#' lines are comments; the last line of a key wins.
static bool tune_lookup(const char *path, const std::string &key, cq_tiles &t, double &ms)
{
    FILE *f = fopen(path, "r");
    if (!f) return false;
    bool found = false;
    std::string line;
    int c;
    for (;;) {
        line.clear();
        while ((c = fgetc(f)) != EOF && c != '\n') line.push_back((char)c);
        if (!line.empty() && line[0] != '#') {
            const size_t tab = line.find('\t');
            cq_tiles x;
            double m;
            if (tab != std::string::npos && line.compare(0, tab, key) == 0 && tab == key.size() &&
                sscanf(line.c_str() + tab + 1, "%d %d %d %d %d %d %lf", &x.tz, &x.cw, &x.cwq, &x.tk, &x.tk2, &x.plan,
                       &m) == 7) {
                t = x;
                ms = m;
                found = true;
            }
        }
        if (c == EOF) break;
    }
    fclose(f);
    return found;
}

static int tune_store(const char *path, const std::string &key, const cq_tiles &t, double ms)
{
    FILE *f = fopen(path, "a");
    if (!f) return set_err(CQ_EINVAL, "cq_autotune: cannot write the cache %s", path);
    if (ftell(f) == 0) fprintf(f, "# cq_autotune cache: key<TAB>tz cw cwq tk tk2 plan ms_per_step\n");
    fprintf(f, "%s\t%d %d %d %d %d %d %.6f\n", key.c_str(), t.tz, t.cw, t.cwq, t.tk, t.tk2, t.plan, ms);
    fclose(f);
    return CQ_OK;
}

static const int kTuneJobs = 24;   // trial batch: at most this many jobs
static const int kTuneWarm = 2, kTuneMeas = 6;

// The trial batch's jobs: every job up to kTuneJobs, else one of each window
// class and an even spread of the rest; RTM kinds run as their modeling
// kinds (the same passes), every job for kTuneWarm + kTuneMeas steps.
static std::vector<cq_job_desc> tune_jobs(const cq_job_desc *jobs, int njobs)
{
    std::vector<int> pick;
    if (njobs <= kTuneJobs) {
        for (int i = 0; i < njobs; i++) pick.push_back(i);
    } else {
        std::set<std::tuple<bool, int, int, int>> seen;
        std::vector<bool> used(njobs, false);
        for (int i = 0; i < njobs; i++) {
            const auto k = std::make_tuple(is_q(jobs[i].kind), jobs[i].lr_rank, jobs[i].nxx, jobs[i].nzz);
            if (seen.insert(k).second) {
                pick.push_back(i);
                used[i] = true;
            }
        }
        const int room = kTuneJobs - (int)pick.size();
        for (int r = 0; r < room; r++) {
            const int i = (int)((long long)r * njobs / std::max(room, 1));
            if (!used[i]) {
                pick.push_back(i);
                used[i] = true;
            }
        }
    }
    std::vector<cq_job_desc> out;
    for (int i : pick) {
        cq_job_desc d = jobs[i];
        if (d.kind == CQ_KIND_RTM) d.kind = CQ_KIND_ACOUSTIC;
        if (d.kind == CQ_KIND_QRTM) d.kind = CQ_KIND_VISCO;
        d.nt = std::min(d.nt, kTuneWarm + kTuneMeas);
        out.push_back(d);
    }
    return out;
}

// One trial: a batch of the trial jobs with tiles t, kTuneWarm steps, then
// kTuneMeas steps timed per pass (pm != null, per-pass events on one stream)
// or as whole steps (step_ms, the launch groups overlapping).
static int tune_trial(cq_ctx ctx, const std::vector<cq_job_desc> &tj, const cq_tiles &t, double *pm, double *step_ms)
{
    cq_batch b = nullptr;
    int rc = cq_batch_create_tiled(ctx, tj.data(), (int)tj.size(), &t, &b);
    if (rc) return rc;
    rc = cq_batch_step(b, kTuneWarm);
    if (rc == CQ_OK) rc = cq_batch_sync(b);
    if (rc == CQ_OK && pm) {
        int calls[P_COUNT];
        rc = cq_batch_enable_timing(b, 1);
        if (rc == CQ_OK) rc = cq_batch_step(b, kTuneMeas);
        if (rc == CQ_OK) rc = cq_batch_kernel_times(b, pm, calls);
    } else if (rc == CQ_OK) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            rc = set_err(CQ_EDEVICE, "cq_autotune: hipEventCreate failed");
        } else {
            float ms = 0.f;
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = set_err(CQ_EDEVICE, "cq_autotune: event record");
            if (rc == CQ_OK) rc = cq_batch_step(b, kTuneMeas);
            if (rc == CQ_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = set_err(CQ_EDEVICE, "cq_autotune: event timing failed");
            *step_ms = ms / kTuneMeas;
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    cq_batch_destroy(b);
    return rc;
}

extern "C" int cq_autotune(cq_ctx ctx, const cq_job_desc *jobs, int njobs, const char *cache, cq_tiles *out,
                           double *tuned_ms, int *from_cache)
{
    if (!ctx || !jobs || njobs < 1 || !out) return set_err(CQ_EINVAL, "cq_autotune: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    for (int i = 0; i < njobs; i++) {
        const int rc = validate(jobs[i], i);
        if (rc) return rc;
    }
    const std::string key = tune_key(ctx, jobs, njobs);
    double ms = 0.0;
    if (from_cache) *from_cache = 0;
    if (cache && tune_lookup(cache, key, *out, ms)) {
        if (tuned_ms) *tuned_ms = ms;
        if (from_cache) *from_cache = 1;
        return CQ_OK;
    }
    const std::vector<cq_job_desc> tj = tune_jobs(jobs, njobs);
    bool has_a = false, has_q = false;
    for (const cq_job_desc &d : tj) (is_q(d.kind) ? has_q : has_a) = true;
    // the engine's own choice, per-pass times as the reference point
    cq_tiles base;
    {
        cq_batch b = nullptr;
        int rc = cq_batch_create_tiled(ctx, tj.data(), (int)tj.size(), nullptr, &b);
        if (rc) return rc;
        cq_batch_get_tiles(b, &base);
        cq_batch_destroy(b);
    }
    double bpm[P_COUNT];
    int rc = tune_trial(ctx, tj, base, bpm, nullptr);
    if (rc) return rc;
    // one field at a time from the engine's choice; a candidate replaces the
    // current value when the passes it shapes run at least 1% faster
    struct Field {
        int cq_tiles::*m;
        const int *cand;
        int ncand;
        std::vector<int> passes;
        bool used;
    };
    const Field fields[] = {
        {&cq_tiles::tz, kTzCand, (int)(sizeof kTzCand / sizeof(int)), {P_AROW, P_QR3}, true},
        {&cq_tiles::cw, kCwCand, (int)(sizeof kCwCand / sizeof(int)), {P_ACOL}, has_a},
        {&cq_tiles::cwq, kCwCand, (int)(sizeof kCwCand / sizeof(int)), {P_QC1, P_QC2, P_LR2, P_QC3}, has_q},
        {&cq_tiles::tk, kTkCand, (int)(sizeof kTkCand / sizeof(int)), {P_QR1, P_LR1}, has_q},
        {&cq_tiles::tk2, kTk2Cand, (int)(sizeof kTk2Cand / sizeof(int)), {P_QR2}, has_q},
    };
    cq_tiles best = base;
    for (const Field &f : fields) {
        if (!f.used) continue;
        double best_t = 0.0;
        for (int p : f.passes) best_t += bpm[p];
        for (int c = 0; c < f.ncand; c++) {
            if (f.cand[c] == base.*f.m) continue;
            cq_tiles t = base;
            t.*f.m = f.cand[c];
            double pm[P_COUNT];
            if (tune_trial(ctx, tj, t, pm, nullptr) != CQ_OK) continue;   // does not fit: skipped
            double tt = 0.0;
            for (int p : f.passes) tt += pm[p];
            if (tt < 0.99 * best_t) {
                best_t = tt;
                best.*f.m = f.cand[c];
            }
        }
    }
    // whole steps: the engine's choice, the per-field winners, and the
    // winners with the other FFT factorisation; the fastest that runs wins
    cq_tiles cands[3] = {base, best, best};
    cands[2].plan = 1 - best.plan;
    cq_tiles win = base;
    double win_ms = 0.0;
    bool have = false;
    for (const cq_tiles &t : cands) {
        double sm = 0.0;
        if (tune_trial(ctx, tj, t, nullptr, &sm) != CQ_OK) continue;
        if (!have || sm < win_ms) {
            win = t;
            win_ms = sm;
            have = true;
        }
    }
    if (!have) return set_err(CQ_EDEVICE, "cq_autotune: no trial ran");
    *out = win;
    if (tuned_ms) *tuned_ms = win_ms;
    if (cache) {
        rc = tune_store(cache, key, win, win_ms);
        if (rc) return rc;
    }
    g_err.clear();
    return CQ_OK;
}
