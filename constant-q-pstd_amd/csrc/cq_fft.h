// cq_fft.h -- LDS-resident mixed-radix Stockham FFT for CDNA4 (gfx950).
//
// A workgroup transforms L lines of length n held in LDS (line stride LS
// float2).  Each stage of radix R reads R strided inputs per butterfly into
// registers, applies the stage twiddles from an LDS table of n roots
// (w_n^j = exp(-2 pi i j / n), generated in double on the host), runs an
// in-register DFT_R with compile-time constants, then writes the autosorted
// outputs back to the same buffer after a barrier.  Supported radices:
// 2, 3, 4, 5, 7, 8, 9, 11, 13, 16, 17 and the composite odd radices 15, 21,
// 25, 27 (one Cooley-Tukey split in registers), so any length whose prime
// factors are at most 17; the composites take the 3/5/7-rich lengths in fewer
// stages, i.e. fewer passes through LDS.
// A complex multiplier can be fused into the last forward stage's store.
#pragma once
#include <hip/hip_runtime.h>

// Global address space qualifier for device pointers (global_load/store
// instead of flat accesses).  It only exists in the device compilation pass;
// the host pass, which never dereferences these pointers, sees plain ones.
#if defined(__HIP_DEVICE_COMPILE__)
#define CQ_GLOBAL __attribute__((address_space(1)))
#else
#define CQ_GLOBAL
#endif


// Workgroups of 512 threads.  A thread holds the butterflies it computes in
// one stage in registers (cq_nb(R) of them: 16-22 complex values, 32 for
// radices 7/8 whose butterflies need few temporaries), which keeps every pass
// at <= 128 VGPRs, i.e. 4 waves per SIMD when two workgroups share a CU.
// Two radix-11 butterflies per thread let 880-point columns (16 5 11) be
// transformed eight lines at a time; two radix-13 or radix-17 ones spill.
#ifndef CQ_NT
#define CQ_NT 512
#endif
#define CQ_MAXST 12      // max stages per plan
// Occupancy every pass is compiled for: two 512-thread workgroups per CU,
// i.e. 4 waves per SIMD, which caps a kernel at 128 VGPRs.
#define CQ_WAVES_PER_SIMD 4

// butterflies per thread per stage for radix R
__host__ __device__ constexpr int cq_nb(int R)
{
    return R == 2 ? 12 : R == 3 ? 6 : R == 4 ? 5 : R == 5 ? 4 : R == 7 ? 3 : R == 8 ? 3 : R == 9 ? 2 : R == 11 ? 2 : R == 15 ? 2 : 1;
}

// A transform plan: stages of radices rad[0 .. ns) for n points.  A length
// with a prime factor above 17 is done by Bluestein's chirp-z algorithm:
// then n is the smooth length M >= 2 bn - 1 of the inner transforms, bn the
// true length, chirp[j] = exp(-i pi j^2 / bn) (j < bn) and bspec the M-point
// spectrum of the conjugate chirp, wrapped cyclically, divided by M.
// tc: the last stage's twiddles are stored as their first powers only
// (w^k, k < Ns) and raised to the powers r = 2 .. R-1 in registers, which
// shrinks the table from n - 1 to 2 n / R - 1 entries (long lines, whose
// tables would otherwise take LDS from the lines).
struct CqPlan {
    int n;
    int ns;
    int rad[CQ_MAXST];
    int bn;   // 0: direct transform of n points
    int tc;
    const float2 *chirp;   // device pointers (global memory)
    const float2 *bspec;
};

// LDS line layout: a line holds its n values in order (slot i = value i) and
// lines are cq_line_stride(n) slots apart (odd, so that consecutive lines
// start on different banks).  One exception: between the first and second
// stage of a transform whose first radix equals the length's pad period P
// (16 when 16 | n, else 8 when 8 | n), value i sits at slot i + i / P.  The
// first stage's outputs, R apart in value order, then land R + 1 apart (no
// bank conflicts), and both stages still address their R values at one
// constant stride.  Lines longer than CQ_PAD_MAX points are not padded: there
// the n / P extra slots per line (4725 instead of 4201 slots at n = 4200)
// would push two-line tiles past 80 KB and halve the workgroups per CU, which
// costs more than the first stage's bank conflicts.
#define CQ_PAD_MAX 2048
__host__ __device__ constexpr int cq_pad_period(int n)
{
    return n > CQ_PAD_MAX ? 0 : n % 16 == 0 ? 16 : n % 8 == 0 ? 8 : 0;
}
__host__ __device__ constexpr int cq_line_stride(int n)
{
    return (n + (cq_pad_period(n) ? n / cq_pad_period(n) : 0)) | 1;
}

// ---------------------------------------------------------------------------
// compile-time trigonometry for the butterfly constants
// ---------------------------------------------------------------------------
constexpr double cq_pi = 3.14159265358979323846264338327950288;

constexpr double cq_csin(double x)
{
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 24; i++) {
        term *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cq_ccos(double x)
{
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 24; i++) {
        term *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
// angle 2 pi k / R reduced to [-pi, pi]
constexpr double cq_ang(int k, int R)
{
    int r = k % R;
    if (2 * r > R) r -= R;
    return 2.0 * cq_pi * (double)r / (double)R;
}

template <int R>
struct CqRoots {
    float c[R];
    float s[R];
    constexpr CqRoots() : c(), s()
    {
        for (int k = 0; k < R; k++) {
            c[k] = (float)cq_ccos(cq_ang(k, R));
            s[k] = (float)cq_csin(cq_ang(k, R));
        }
    }
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)  // a * conj(b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 cmul_mi(float2 a)
{
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// in-register DFTs
// ---------------------------------------------------------------------------
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
    __device__ __forceinline__ static void run(float2 *v)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV>
struct Dft<4, INV> {
    __device__ __forceinline__ static void run(float2 *v)
    {
        float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]), t3 = cmul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

// power-of-two radices from two half-size DFTs
template <int R, bool INV>
struct DftPow2 {
    __device__ __forceinline__ static void run(float2 *v)
    {
        constexpr int H = R / 2;
        constexpr CqRoots<R> T{};
        float2 e[H], o[H];
#pragma unroll
        for (int i = 0; i < H; i++) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        Dft<H, INV>::run(e);
        Dft<H, INV>::run(o);
#pragma unroll
        for (int k = 0; k < H; k++) {
            float2 t;
            if (k == 0) {
                t = o[k];
            } else if (4 * k == R) {
                t = cmul_mi<INV>(o[k]);
            } else {
                float2 w = make_float2(T.c[k], INV ? T.s[k] : -T.s[k]);
                t = cmul(o[k], w);
            }
            v[k] = cadd(e[k], t);
            v[k + H] = csub(e[k], t);
        }
    }
};
template <bool INV>
struct Dft<8, INV> : DftPow2<8, INV> {};
template <bool INV>
struct Dft<16, INV> : DftPow2<16, INV> {};

// odd radices: pairwise symmetric form
//   y_q = x0 + sum_j (x_j + x_{R-j}) cos(2 pi j q/R) -/+ i sum_j (x_j - x_{R-j}) sin(2 pi j q/R)
template <int R, bool INV>
struct DftOdd {
    __device__ __forceinline__ static void run(float2 *v)
    {
        constexpr int H = (R - 1) / 2;
        constexpr CqRoots<R> T{};
        float2 s[H], d[H];
        float2 y0 = v[0];
#pragma unroll
        for (int j = 1; j <= H; j++) {
            s[j - 1] = cadd(v[j], v[R - j]);
            d[j - 1] = csub(v[j], v[R - j]);
            y0 = cadd(y0, s[j - 1]);
        }
        float2 out[R];
        out[0] = y0;
#pragma unroll
        for (int q = 1; q <= H; q++) {
            float2 A = v[0];
            float2 B = make_float2(0.f, 0.f);
#pragma unroll
            for (int j = 1; j <= H; j++) {
                const int idx = (j * q) % R;
                A.x += s[j - 1].x * T.c[idx];
                A.y += s[j - 1].y * T.c[idx];
                B.x += d[j - 1].x * T.s[idx];
                B.y += d[j - 1].y * T.s[idx];
            }
            if (!INV) {
                out[q] = make_float2(A.x + B.y, A.y - B.x);
                out[R - q] = make_float2(A.x - B.y, A.y + B.x);
            } else {
                out[q] = make_float2(A.x - B.y, A.y + B.x);
                out[R - q] = make_float2(A.x + B.y, A.y - B.x);
            }
        }
#pragma unroll
        for (int q = 0; q < R; q++) v[q] = out[q];
    }
};
template <bool INV> struct Dft<3, INV> : DftOdd<3, INV> {};
template <bool INV> struct Dft<5, INV> : DftOdd<5, INV> {};
template <bool INV> struct Dft<7, INV> : DftOdd<7, INV> {};
template <bool INV> struct Dft<9, INV> : DftOdd<9, INV> {};
template <bool INV> struct Dft<11, INV> : DftOdd<11, INV> {};
template <bool INV> struct Dft<13, INV> : DftOdd<13, INV> {};
template <bool INV> struct Dft<17, INV> : DftOdd<17, INV> {};

// Composite radix R = R1 R2 by one Cooley-Tukey split in registers: input
// x[R2 n1 + n2]; DFT_R1 over n1 for each n2, times w_R^(n2 k1), DFT_R2 over
// n2 for each k1 gives X[k1 + R1 k2].  Each input is dead once its inner DFT
// has read it and each inner result once its outer DFT has, so about R
// complex values are live at a time.
template <int R1, int R2, bool INV>
struct DftCT {
    __device__ __forceinline__ static void run(float2 *v)
    {
        constexpr int R = R1 * R2;
        constexpr CqRoots<R> T{};
        float2 y[R2][R1];
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) {
            float2 t[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; n1++) t[n1] = v[R2 * n1 + n2];
            Dft<R1, INV>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < R1; k1++) {
                const int e = (n2 * k1) % R;
                y[n2][k1] = e == 0 ? t[k1] : cmul(t[k1], make_float2(T.c[e], INV ? T.s[e] : -T.s[e]));
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) {
            float2 t[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; n2++) t[n2] = y[n2][k1];
            Dft<R2, INV>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = t[k2];
        }
    }
};
template <bool INV> struct Dft<15, INV> : DftCT<3, 5, INV> {};
template <bool INV> struct Dft<21, INV> : DftCT<3, 7, INV> {};
template <bool INV> struct Dft<25, INV> : DftCT<5, 5, INV> {};
template <bool INV> struct Dft<27, INV> : DftCT<3, 9, INV> {};

// Dynamic LDS of every kernel.  The out-of-line stage functions address it
// through this symbol (by float2 offsets) so that their accesses compile to
// LDS instructions rather than generic (flat) ones.
extern __shared__ float2 cq_lds[];

// ---------------------------------------------------------------------------
// one Stockham stage over L lines in LDS
// ---------------------------------------------------------------------------
// bo   : offset of the first line in cq_lds; L lines LS slots apart
// to   : offset of this stage's twiddles, tw[(r-1) Ns + k] = w_{Ns R}^(r k)
//        (r = 1..R-1, k < Ns), so lanes with consecutive k read consecutive
//        slots (no bank conflicts)
// Ns   : product of the radices of the previous stages
// mh   : optional multiplier (global, length n) applied to the outputs
// MODE : 0 values in order on both sides; 1 first stage of several, R = pad
//        period, outputs in the padded layout; 2 second stage after a MODE 1
//        stage (Ns = pad period), inputs in the padded layout
// TC  : compressed table (CqPlan::tc): tw[k] = w_{Ns R}^k only
template <int R, bool INV, int MODE, int CN = 0, int CNS = 0, int CLSM = 1, bool MH = true, bool TC = false>
__device__ __noinline__ void cq_stage(int bo, int LS, int n, int L, int Ns, int to,
                                      const float2 CQ_GLOBAL *__restrict__ mh)
{
    // the arguments are uniform but arrive in vector registers: move them to
    // scalar registers so that the index and stride arithmetic runs there
    bo = __builtin_amdgcn_readfirstlane(bo);
    LS = __builtin_amdgcn_readfirstlane(LS);
    n = __builtin_amdgcn_readfirstlane(n);
    L = __builtin_amdgcn_readfirstlane(L);
    Ns = __builtin_amdgcn_readfirstlane(Ns);
    to = __builtin_amdgcn_readfirstlane(to);
    // a stage compiled for one length (CN, lines CLSM cq_line_stride(CN)
    // apart, CNS = Ns): strides and offsets become constants, so the LDS
    // accesses take immediate offsets instead of one address add each
    if constexpr (CN > 0) {
        n = CN;
        Ns = CNS;
        LS = CLSM * cq_line_stride(CN);
    }
    float2 *const buf = cq_lds + bo;
    const float2 *const tw = cq_lds + to;
    constexpr int NB = cq_nb(R);
    const int m = n / R;
    const int total = L * m;
    // exact quotients by the hardware reciprocal (1 ulp): operands are below
    // 2^14, so (i + 0.5) / d lies at least 0.5 / d from an integer while the
    // product is off by less than 2^-8 / d.  Every index product below has
    // operands under 2^23: a full-rate 24-bit multiply (__mul24), not the
    // quarter-rate 32-bit one
    const float inv_m = __builtin_amdgcn_rcpf((float)m), inv_ns = __builtin_amdgcn_rcpf((float)Ns);
    // slots between a butterfly's inputs: m, or m + m / Ns in the padded
    // layout (the period Ns divides m)
    const int rs = MODE == 2 ? m + m / Ns : m;
    const int psh = MODE == 2 ? 31 - __builtin_clz(Ns) : 0;
    // the first stage (Ns = 1) has only unit twiddles; later stages multiply
    // every input, the k = 0 entries of the table being exactly (1, 0)
    const bool twiddle = MODE != 1 && Ns > 1;
    float2 v[NB][R];
    int base[NB], kk[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int idx = threadIdx.x + b * CQ_NT;
        base[b] = -1;
        kk[b] = 0;
        if (idx < total) {
            const int l = (int)(((float)idx + 0.5f) * inv_m);
            const int j = idx - __mul24(l, m);
            const int k = j - __mul24((int)(((float)j + 0.5f) * inv_ns), Ns);
            const float2 *src = buf + __mul24(l, LS) + (MODE == 2 ? j + (j >> psh) : j);
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = src[r * rs];
            if (twiddle) {
                const float2 *twk = tw + k;
                if constexpr (TC) {
                    // w^r = w^(r/2) w^(r - r/2): at most log2(R) products deep
                    float2 wp[R];
                    wp[1] = twk[0];
#pragma unroll
                    for (int r = 2; r < R; r++) wp[r] = cmul(wp[r >> 1], wp[r - (r >> 1)]);
#pragma unroll
                    for (int r = 1; r < R; r++) v[b][r] = INV ? cmulc(v[b][r], wp[r]) : cmul(v[b][r], wp[r]);
                } else {
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        const float2 w = twk[(r - 1) * Ns];
                        v[b][r] = INV ? cmulc(v[b][r], w) : cmul(v[b][r], w);
                    }
                }
            }
            Dft<R, INV>::run(v[b]);
            // output q is value kk + q Ns: slot kk + q Ns, or, padded (Ns = 1,
            // k = 0), j (R + 1) + q
            kk[b] = __mul24(j - k, R) + k;
            base[b] = __mul24(l, LS) + (MODE == 1 ? __mul24(j, R + 1) : kk[b]);
            if (MH && mh != nullptr) {
#pragma unroll
                for (int q = 0; q < R; q++) v[b][q] = cmul(v[b][q], mh[kk[b] + q * Ns]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; b++) {
        if (base[b] >= 0) {
            float2 *dst = buf + base[b];
#pragma unroll
            for (int q = 0; q < R; q++) dst[q * Ns] = v[b][q];
        }
    }
    __syncthreads();
}

// Composite radices hold R values and their sub-DFTs' temporaries, so they
// are compiled without the multiplier code and the compressed-table form:
// plans never make a composite the last stage (make_direct_plan), which is
// the only stage that carries a multiplier or a compressed table.  With
// those forms they would spill to scratch, and a kernel with any callee
// that uses scratch sets up scratch for every wave it launches.
#define CQ_STAGE_COMPOSITE(INV, MODE)                                                          \
    case 15: cq_stage<15, INV, MODE, 0, 0, 1, false>(bo, LS, n, L, Ns, to, nullptr); break;    \
    case 21: cq_stage<21, INV, MODE, 0, 0, 1, false>(bo, LS, n, L, Ns, to, nullptr); break;    \
    case 25: cq_stage<25, INV, MODE, 0, 0, 1, false>(bo, LS, n, L, Ns, to, nullptr); break;    \
    case 27: cq_stage<27, INV, MODE, 0, 0, 1, false>(bo, LS, n, L, Ns, to, nullptr); break;
#define CQ_STAGE_CASES(INV, MODE, TC)                                                           \
    case 2: cq_stage<2, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;         \
    case 3: cq_stage<3, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;         \
    case 4: cq_stage<4, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;         \
    case 5: cq_stage<5, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;         \
    case 7: cq_stage<7, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;         \
    case 8: cq_stage<8, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;         \
    case 9: cq_stage<9, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;         \
    case 11: cq_stage<11, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;       \
    case 13: cq_stage<13, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;       \
    case 16: cq_stage<16, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;       \
    case 17: cq_stage<17, INV, MODE, 0, 0, 1, true, TC>(bo, LS, n, L, Ns, to, mh); break;       \
    default: break;

// tc: this is the last stage of a plan with a compressed table
template <bool INV>
__device__ __forceinline__ void cq_stage_rt(int R, int mode, int bo, int LS, int n, int L, int Ns, int to,
                                            const float2 CQ_GLOBAL *__restrict__ mh, bool tc)
{
    if (mode == 1) {
        if (R == 16) cq_stage<16, INV, 1>(bo, LS, n, L, Ns, to, mh);
        else cq_stage<8, INV, 1>(bo, LS, n, L, Ns, to, mh);
    } else if (tc) {
        if (mode == 2) {
            switch (R) { CQ_STAGE_CASES(INV, 2, true) }
        } else {
            switch (R) { CQ_STAGE_CASES(INV, 0, true) }
        }
    } else if (mode == 2) {
        switch (R) { CQ_STAGE_COMPOSITE(INV, 2) CQ_STAGE_CASES(INV, 2, false) }
    } else {
        switch (R) { CQ_STAGE_COMPOSITE(INV, 0) CQ_STAGE_CASES(INV, 0, false) }
    }
}

// Layout mode of a plan's first stage (the second follows: 1 -> 2, 0 -> 0).
__device__ __forceinline__ int cq_first_mode(const CqPlan &p)
{
    return p.ns > 1 && p.rad[0] == cq_pad_period(p.n) ? 1 : 0;
}

// Twiddle table of a plan (n - 1 entries, built on the host by the same rule):
// the stages' tables one after another, stage s holding (R_s - 1) Ns entries
// tw[(r-1) Ns + k] = w_{Ns R_s}^(r k) = exp(-2 pi i r k / (Ns R_s)).

// Forward transform of L lines at buf, the plan's twiddle table at tw (both
// inside cq_lds), optional multiplier fused into the last stage.
// The stages R, RS... of a plan of length N compiled for that length, lines
// LSM cq_line_stride(N) apart; MODE, NS and TO are the current stage's layout
// mode, Ns and twiddle offset (the generic loop's run-time values).
// TCL: the last stage reads a compressed table (CqPlan::tc).
template <bool INV, int N, int LSM, int MODE, int NS, int TO, bool TCL, int R, int... RS>
__device__ __forceinline__ void cq_fft_stages(int bo, int L, int to, const float2 CQ_GLOBAL *__restrict__ mh)
{
    constexpr int LS = LSM * cq_line_stride(N);
    if constexpr (sizeof...(RS) > 0) {
        cq_stage<R, INV, MODE, N, NS, LSM>(bo, LS, N, L, NS, to + TO, nullptr);
        cq_fft_stages<INV, N, LSM, (MODE == 1 ? 2 : 0), NS * R, TO + (R - 1) * NS, TCL, RS...>(bo, L, to, mh);
    } else if constexpr (TCL) {
        cq_stage<R, INV, MODE, N, NS, LSM, true, true>(bo, LS, N, L, NS, to + TO, mh);
    } else if constexpr (!INV && R == 17) {
        // compiled for the length with its multiplier code, a forward
        // radix-17 last stage needs 120 VGPRs and spills: with a multiplier
        // it keeps the generic form, without one it is compiled without
        // that code
        if (mh != nullptr) cq_stage<R, INV, MODE>(bo, LS, N, L, NS, to + TO, mh);
        else cq_stage<R, INV, MODE, N, NS, LSM, false>(bo, LS, N, L, NS, to + TO, nullptr);
    } else {
        cq_stage<R, INV, MODE, N, NS, LSM>(bo, LS, N, L, NS, to + TO, mh);
    }
}

// A plan R0 RS... of length N as compiled stages, if p is that plan and its
// lines are LSM line strides apart; the first stage's mode follows
// cq_first_mode.  TCOK: also compiled with a compressed last-stage table.
template <bool INV, bool TCOK, int N, int LSM, int R0, int... RS>
__device__ __forceinline__ bool cq_fft_plan(const CqPlan &p, int LS, int bo, int L, int to,
                                            const float2 CQ_GLOBAL *__restrict__ mh)
{
    constexpr int rad[] = {R0, RS...};
    constexpr int ns = 1 + sizeof...(RS);
    static_assert((R0 * ... * RS) == N, "plan");
    if (p.n != N || p.ns != ns || LS != LSM * cq_line_stride(N)) return false;
    for (int s = 0; s < ns; s++)
        if (p.rad[s] != rad[s]) return false;
    constexpr int mode0 = ns > 1 && R0 == cq_pad_period(N) ? 1 : 0;
    if (p.tc) {
        if constexpr (TCOK) {
            cq_fft_stages<INV, N, LSM, mode0, 1, 0, true, R0, RS...>(bo, L, to, mh);
            return true;
        }
        return false;
    }
    cq_fft_stages<INV, N, LSM, mode0, 1, 0, false, R0, RS...>(bo, L, to, mh);
    return true;
}

// Lengths with compiled plans: the windows of the batched workload (680 deep,
// 1296 wide) and of the Q-RTM workload (880 deep, 2100 = 25 21 4 wide); lines one or,
// for k_q_c2's interleaved lines, two line strides apart.  True if the
// transform ran.
template <bool INV>
__device__ __forceinline__ bool cq_fft_fixed(const CqPlan &p, int LS, int bo, int L, int to,
                                             const float2 CQ_GLOBAL *__restrict__ mh)
{
    return cq_fft_plan<INV, false, 680, 1, 8, 5, 17>(p, LS, bo, L, to, mh) ||
           cq_fft_plan<INV, false, 680, 2, 8, 5, 17>(p, LS, bo, L, to, mh) ||
           cq_fft_plan<INV, false, 1296, 1, 16, 9, 9>(p, LS, bo, L, to, mh) ||
           cq_fft_plan<INV, false, 880, 1, 16, 5, 11>(p, LS, bo, L, to, mh) ||
           cq_fft_plan<INV, false, 880, 2, 16, 5, 11>(p, LS, bo, L, to, mh) ||
           cq_fft_plan<INV, true, 2100, 1, 25, 21, 4>(p, LS, bo, L, to, mh);
}

// The stages of plan p as run-time stages (any plan), multiplier on the last.
// Inlined: out of line it would need a stack frame (the loop's state lives
// across the stage calls), i.e. scratch for every wave of every pass, and the
// measured row passes ran 12% slower that way.
template <bool INV>
__device__ __forceinline__ void cq_fft_generic(const CqPlan &p, int bo, int LS, int L, int to,
                                               const float2 CQ_GLOBAL *__restrict__ mh)
{
    int Ns = 1, mode = cq_first_mode(p);
    for (int s = 0; s < p.ns; s++) {
        const int R = p.rad[s];
        const bool last = s == p.ns - 1;
        cq_stage_rt<INV>(R, mode, bo, LS, p.n, L, Ns, to, last ? mh : nullptr, last && p.tc);
        to += (R - 1) * Ns;
        Ns *= R;
        mode = mode == 1 ? 2 : 0;
    }
}

// Bluestein, first step: a[j] = x[j] chirp[j] (the forward transform of
// conj x for an inverse one) for j < n, zero up to M, in place in each line.
template <bool INV>
__device__ __noinline__ void cq_chirp_in(int bo, int LS, int n, int M, int L, const float2 CQ_GLOBAL *__restrict__ c)
{
    bo = __builtin_amdgcn_readfirstlane(bo);
    LS = __builtin_amdgcn_readfirstlane(LS);
    n = __builtin_amdgcn_readfirstlane(n);
    M = __builtin_amdgcn_readfirstlane(M);
    float2 *const buf = cq_lds + bo;
    // exact quotient by the hardware reciprocal, as in cq_stage (e < 2^15)
    const float inv_m = __builtin_amdgcn_rcpf((float)M);
    for (int e = threadIdx.x; e < L * M; e += CQ_NT) {
        const int l = (int)(((float)e + 0.5f) * inv_m), j = e - l * M;
        float2 *line = buf + l * LS;
        float2 v = make_float2(0.f, 0.f);
        if (j < n) {
            float2 x = line[j];
            if (INV) x.y = -x.y;
            v = cmul(x, c[j]);
        }
        line[j] = v;
    }
    __syncthreads();
}

// Bluestein, last step: X[k] = chirp[k] r[k] (conjugated for an inverse
// transform), times the optional multiplier, for k < n.
template <bool INV>
__device__ __noinline__ void cq_chirp_out(int bo, int LS, int n, int L, const float2 CQ_GLOBAL *__restrict__ c,
                                          const float2 CQ_GLOBAL *__restrict__ mh)
{
    bo = __builtin_amdgcn_readfirstlane(bo);
    LS = __builtin_amdgcn_readfirstlane(LS);
    n = __builtin_amdgcn_readfirstlane(n);
    float2 *const buf = cq_lds + bo;
    const float inv_n = __builtin_amdgcn_rcpf((float)n);
    for (int e = threadIdx.x; e < L * n; e += CQ_NT) {
        const int l = (int)(((float)e + 0.5f) * inv_n), k = e - l * n;
        float2 *line = buf + l * LS;
        float2 y = cmul(line[k], c[k]);
        if (INV) y.y = -y.y;
        if (mh != nullptr) y = cmul(y, mh[k]);
        line[k] = y;
    }
    __syncthreads();
}

// Transform of length p.bn by Bluestein's algorithm: with c[j] = e^{-i pi j^2/n},
// X[k] = c[k] sum_j (x[j] c[j]) conj c[k - j], a linear convolution done as a
// cyclic one of M >= 2n - 1 points (forward M-point transform, times bspec,
// inverse); an inverse transform is the conjugate of the forward transform
// of the conjugate.  The lines hold M points (LS >= the M-point line
// stride); the n results replace the first n values.
template <bool INV>
__device__ __forceinline__ void cq_bluestein(const CqPlan &p, int bo, int LS, int L, int to,
                                             const float2 CQ_GLOBAL *__restrict__ mh)
{
    const float2 CQ_GLOBAL *c = (const float2 CQ_GLOBAL *)p.chirp;
    cq_chirp_in<INV>(bo, LS, p.bn, p.n, L, c);
    cq_fft_generic<false>(p, bo, LS, L, to, (const float2 CQ_GLOBAL *)p.bspec);
    cq_fft_generic<true>(p, bo, LS, L, to, nullptr);
    cq_chirp_out<INV>(bo, LS, p.bn, L, c, mh);
}

__device__ __forceinline__ void cq_fft_fwd(const CqPlan &p, float2 *buf, int LS, int L,
                                           const float2 *tw, const float2 CQ_GLOBAL *__restrict__ mh)
{
    const int bo = (int)(buf - cq_lds);
    const int to = (int)(tw - cq_lds);
    if (p.bn) {
        cq_bluestein<false>(p, bo, LS, L, to, mh);
        return;
    }
    if (cq_fft_fixed<false>(p, LS, bo, L, to, mh)) return;
    cq_fft_generic<false>(p, bo, LS, L, to, mh);
}

__device__ __forceinline__ void cq_fft_inv(const CqPlan &p, float2 *buf, int LS, int L,
                                           const float2 *tw)
{
    const int bo = (int)(buf - cq_lds);
    const int to = (int)(tw - cq_lds);
    if (p.bn) {
        cq_bluestein<true>(p, bo, LS, L, to, nullptr);
        return;
    }
    if (cq_fft_fixed<true>(p, LS, bo, L, to, nullptr)) return;
    cq_fft_generic<true>(p, bo, LS, L, to, nullptr);
}

// Largest number of complex points one workgroup can transform per stage for
// a given radix (host and device agree on this).
__host__ __device__ constexpr int cq_stage_capacity(int R)
{
    return CQ_NT * cq_nb(R) * R;
}
