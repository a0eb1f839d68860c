// cq_passes.h -- fused time-step passes (HIP, gfx950).
//
// Work items are (shot, tile) pairs; one workgroup of CQ_NT threads owns one
// tile and runs every transform of the pass in LDS.
//   row tiles  : TZ consecutive depths x all nxx columns, TZ/2 packed complex
//                lines (real rows 2q and 2q+1 ride as real/imag parts)
//   col tiles  : CW consecutive columns x all nzz depths, CW/2 packed lines
//   krow tiles : TK consecutive kz rows of a half spectrum x all nxx columns,
//                TK complex lines
// A packed pair of real lines is only ever multiplied by Hermitian
// multipliers, so the real and imaginary parts never mix.
#pragma once
#include "cq_job.h"

// Field and spectrum traffic is streamed: every pass touches each cell once
// and the batch's working set is far beyond L2 and the Infinity Cache, so the
// quads are loaded and stored with the nontemporal hint (2-4% off the column
// and update passes).  Tables that every workgroup re-reads (twiddles,
// multipliers, CPML profiles) keep ordinary loads.
typedef float cq_v4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld4(const float CQ_GLOBAL *p)
{
    const cq_v4 w = __builtin_nontemporal_load(reinterpret_cast<const cq_v4 CQ_GLOBAL *>(p));
    return make_float4(w.x, w.y, w.z, w.w);
}
__device__ __forceinline__ void st4(float CQ_GLOBAL *p, float4 v)
{
    const cq_v4 w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<cq_v4 CQ_GLOBAL *>(p));
}
// ordinary (cached) accesses, for quads a workgroup reads twice or writes and
// reads back within one pass, so that the second access can hit in L2
__device__ __forceinline__ float4 ld4c(const float CQ_GLOBAL *p) { return *reinterpret_cast<const float4 CQ_GLOBAL *>(p); }
__device__ __forceinline__ void st4c(float CQ_GLOBAL *p, float4 v) { *reinterpret_cast<float4 CQ_GLOBAL *>(p) = v; }
// two complex half-spectrum values
__device__ __forceinline__ void st4v(float2 CQ_GLOBAL *p, float4 v) { st4(reinterpret_cast<float CQ_GLOBAL *>(p), v); }
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

// The loop e = threadIdx.x, + CQ_NT, ... < total in groups of U iterations:
// pre(e) of all U first (the global loads; it returns what it loaded), then
// post(e, loaded) of all U, so that U iterations' loads are in flight
// together rather than one iteration's at a time.  CLAMP: iterations past
// the end load element total - 1 again (the same address as another lane's
// load of the batch, or one already read) instead of
// branching round their loads, which lets the compiler keep all U in flight
// where the branches made it wait early (the column passes); their post is
// skipped as before.
template <int U, bool CLAMP = false, typename Pre, typename Post>
__device__ __forceinline__ void cq_loop(int total, Pre pre, Post post)
{
    using T = decltype(pre(0));
#pragma unroll 1
    for (int e0 = threadIdx.x; e0 < total; e0 += U * CQ_NT) {
        T d[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (CLAMP) d[u] = pre(min(e0 + u * CQ_NT, total - 1));
            else if (e0 + u * CQ_NT < total) d[u] = pre(e0 + u * CQ_NT);
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (e0 + u * CQ_NT < total) post(e0 + u * CQ_NT, d[u]);
    }
}

// e / n for 0 <= e < 2^21 by the hardware reciprocal of n (inv = 1 / n as
// __builtin_amdgcn_rcpf gives it): (e + 0.5) / n lies at least 0.5 / n from
// an integer, the product is off by less than (e + 0.5) / n * 2^-22.
__device__ __forceinline__ int cq_div(int e, float inv) { return (int)(((float)e + 0.5f) * inv); }

struct F4x2 {
    float4 a, b;
};
struct F4x3 {
    float4 a, b, c;
};
struct F4x5 {
    float4 a, b, c, d, e;
};

__device__ __forceinline__ void load_tw(float2 *tw, const float2 CQ_GLOBAL *__restrict__ g, int n)
{
    cq_loop<4>(n, [&](int i) { return g[i]; }, [&](int i, float2 v) { tw[i] = v; });
}

// ===========================================================================
// row helpers (real rows, packed)
// ===========================================================================
// CACHED: the pass reads src again later (ordinary loads, see ld4c)
template <int Q, bool CACHED = false>
__device__ __forceinline__ void row_fill(float2 *buf, int nxx, int LS, const float CQ_GLOBAL *__restrict__ src, int z0)
{
    const float inv = __builtin_amdgcn_rcpf((float)nxx);
    cq_loop<8>(
        nxx * Q,
        [&](int e) {
            const float CQ_GLOBAL *a = src + (size_t)(z0 >> 2) * nxx * 4 + (size_t)e * 4;
            return CACHED ? ld4c(a) : ld4(a);
        },
        [&](int e, float4 d) {
            const int q = cq_div(e, inv), ix = e - q * nxx;
            buf[(2 * q) * LS + ix] = make_float2(d.x, d.y);
            buf[(2 * q + 1) * LS + ix] = make_float2(d.z, d.w);
        });
}

// x-CPML on dp/dx (in LDS) and the vx update; leaves vx_new in LDS
// (vx_cal_kernel, sgpstd-qmodeling/acps4.cu:1274-1303)
template <int Q>
__device__ __forceinline__ void row_vx_phase(const JobG &J, float2 *buf, int z0)
{
    const int nxx = J.nxx, pml = J.pml, LS = J.lsx;
    const float inv = __builtin_amdgcn_rcpf((float)nxx);
    const size_t g0 = (size_t)(z0 >> 2) * nxx * 4;
    cq_loop<4>(
        nxx * Q,
        [&](int e) {
            const int q = cq_div(e, inv), ix = e - q * nxx;
            F4x3 r;
            r.a = ld4(J.vx + g0 + (size_t)e * 4);
            r.b = ld4(J.dtbx + g0 + (size_t)e * 4);
            if (ix < pml || ix >= nxx - pml) {
                const int sx = ix < pml ? ix : ix - (nxx - 2 * pml);
                r.c = ld4(J.phix + cq_q4(2 * pml, sx, (z0 >> 2) + q));
            }
            return r;
        },
        [&](int e, const F4x3 &r) {
            const int q = cq_div(e, inv), ix = e - q * nxx;
            const float2 a = buf[(2 * q) * LS + ix], b = buf[(2 * q + 1) * LS + ix];
            float4 dp = make_float4(a.x, a.y, b.x, b.y);
            const float4 vx = r.a, bx = r.b;
            if (ix < pml || ix >= nxx - pml) {
                const int sx = ix < pml ? ix : ix - (nxx - 2 * pml);
                const float hb = J.hbx[ix], ha = J.hax[ix];
                float4 ph = r.c;
                ph.x = hb * ph.x + ha * dp.x;
                ph.y = hb * ph.y + ha * dp.y;
                ph.z = hb * ph.z + ha * dp.z;
                ph.w = hb * ph.w + ha * dp.w;
                st4(J.phix + cq_q4(2 * pml, sx, (z0 >> 2) + q), ph);
                dp.x += ph.x; dp.y += ph.y; dp.z += ph.z; dp.w += ph.w;
            }
            float4 vn;
            vn.x = vx.x + bx.x * dp.x;
            vn.y = vx.y + bx.y * dp.y;
            vn.z = vx.z + bx.z * dp.z;
            vn.w = vx.w + bx.w * dp.w;
            st4(J.vx + g0 + (size_t)e * 4, vn);
            buf[(2 * q) * LS + ix] = make_float2(vn.x, vn.y);
            buf[(2 * q + 1) * LS + ix] = make_float2(vn.z, vn.w);
        });
}

// dvx/dx (in LDS) with its x-CPML memory variable -> dvx for the next step
// (update_vxx_vzz, acps4.cu:1111-1144).  TO_DIV: add dvz (already computed
// by the column pass) and store div instead (sumpxpz, acps4.cu:1146-1172).
template <int Q, bool TO_DIV = false>
__device__ __forceinline__ void row_dvx_phase(const JobG &J, const float2 *buf, int z0)
{
    const int nxx = J.nxx, pml = J.pml, LS = J.lsx;
    const float inv = __builtin_amdgcn_rcpf((float)nxx);
    const size_t g0 = (size_t)(z0 >> 2) * nxx * 4;
    cq_loop<4>(
        nxx * Q,
        [&](int e) {
            const int q = cq_div(e, inv), ix = e - q * nxx;
            F4x2 r;
            if (TO_DIV) r.b = ld4(J.dvz + g0 + (size_t)e * 4);
            if (ix < pml || ix >= nxx - pml) {
                const int sx = ix < pml ? ix : ix - (nxx - 2 * pml);
                r.a = ld4(J.psix + cq_q4(2 * pml, sx, (z0 >> 2) + q));
            }
            return r;
        },
        [&](int e, const F4x2 &r) {
            const int q = cq_div(e, inv), ix = e - q * nxx;
            const float2 a = buf[(2 * q) * LS + ix], b = buf[(2 * q + 1) * LS + ix];
            float4 dv = make_float4(a.x, a.y, b.x, b.y);
            if (ix < pml || ix >= nxx - pml) {
                const int sx = ix < pml ? ix : ix - (nxx - 2 * pml);
                const float ib = J.ibx[ix], ia = J.iax[ix];
                float4 ps = r.a;
                ps.x = ib * ps.x + ia * dv.x;
                ps.y = ib * ps.y + ia * dv.y;
                ps.z = ib * ps.z + ia * dv.z;
                ps.w = ib * ps.w + ia * dv.w;
                st4(J.psix + cq_q4(2 * pml, sx, (z0 >> 2) + q), ps);
                dv.x += ps.x; dv.y += ps.y; dv.z += ps.z; dv.w += ps.w;
            }
            if (TO_DIV) {
                const float4 dz = r.b;
                dv.x += dz.x; dv.y += dz.y; dv.z += dz.z; dv.w += dz.w;
                st4(J.div + g0 + (size_t)e * 4, dv);
            } else {
                st4(J.dvx + g0 + (size_t)e * 4, dv);
            }
        });
}

// ===========================================================================
// column helpers (real columns, packed; bufs = float view of the lines)
// ===========================================================================
__device__ __forceinline__ float *col_line(float *bufs, int LS, int c)
{
    return bufs + 2 * (c >> 1) * LS + (c & 1);
}

template <int CW>
__device__ __forceinline__ void col_fill(float *bufs, const JobG &J, const float CQ_GLOBAL *__restrict__ src, int x0)
{
    const int nzz = J.nzz, nq = (nzz + 3) >> 2;
    cq_loop<8, true>(
        CW * nq,
        [&](int e) {
            const int zq = e / CW, c = e - zq * CW;
            return ld4(src + cq_q4(J.nxx, min(x0 + c, J.nxx - 1), zq));
        },
        [&](int e, float4 v) {
            const int zq = e / CW, c = e - zq * CW;
            float *line = col_line(bufs, J.lsz, c);
            if (x0 + c >= J.nxx) v = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (4 * zq + k < nzz) line[2 * (4 * zq + k)] = f4get(v, k);
        });
}

// gather rec[t][i] from a column tile held as plain floats (src(c, iz))
template <int CW, typename F>
__device__ __forceinline__ void col_gather(const JobG &J, int x0, int t, F src)
{
    const int tile = x0 / CW;
    const int r0 = J.rcv_off[tile], r1 = J.rcv_off[tile + 1];
    for (int r = r0 + threadIdx.x; r < r1; r += CQ_NT)
        J.rec[(size_t)t * J.nr + J.rcv_i[r]] = src(J.rcv_x[r] - x0, J.rcv_z[r]);
}

// z-CPML on dp/dz and the vz update; leaves vz_new in LDS
// (vz_cal_kernel, acps4.cu:1305-1333)
template <int CW>
__device__ __forceinline__ void col_vz_phase(const JobG &J, float *bufs, int x0)
{
    const int nzz = J.nzz, nq = (nzz + 3) >> 2, pml = J.pml;
    cq_loop<4, true>(
        CW * nq,
        [&](int e) {
            // clamped column, masked in the second lambda (col_insert)
            const int zq = e / CW, c = e - zq * CW, ix = min(x0 + c, J.nxx - 1);
            F4x3 r;
            const size_t g = cq_q4(J.nxx, ix, zq);
            r.a = ld4(J.vz + g);
            r.b = ld4(J.dtbz + g);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int iz = 4 * zq + k;
                if (iz < nzz && (iz < pml || iz >= nzz - pml)) {
                    const int sz = iz < pml ? iz : iz - (nzz - 2 * pml);
                    f4set(r.c, k, J.phiz[(size_t)sz * J.nxx + ix]);
                }
            }
            return r;
        },
        [&](int e, const F4x3 &r) {
            const int zq = e / CW, c = e - zq * CW, ix = x0 + c;
            if (ix >= J.nxx) return;
            const size_t g = cq_q4(J.nxx, ix, zq);
            float *line = col_line(bufs, J.lsz, c);
            float4 vn = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int iz = 4 * zq + k;
                if (iz >= nzz) continue;
                float dp = line[2 * iz];
                if (iz < pml || iz >= nzz - pml) {
                    const int sz = iz < pml ? iz : iz - (nzz - 2 * pml);
                    const float ph = J.hbz[iz] * f4get(r.c, k) + J.haz[iz] * dp;
                    J.phiz[(size_t)sz * J.nxx + ix] = ph;
                    dp += ph;
                }
                const float v = f4get(r.a, k) + f4get(r.b, k) * dp;
                f4set(vn, k, v);
                line[2 * iz] = v;
            }
            st4(J.vz + g, vn);
        });
}

// dvz/dz with z-CPML, div = dvx + dvz (update_vxx_vzz + sumpxpz,
// acps4.cu:1111-1172); stores div and leaves it in LDS.  TO_DVZ: store dvz
// alone (the row pass adds dvx).
template <int CW, bool TO_DVZ = false>
__device__ __forceinline__ void col_div_phase(const JobG &J, float *bufs, int x0)
{
    const int nzz = J.nzz, nq = (nzz + 3) >> 2, pml = J.pml;
    cq_loop<4, true>(
        CW * nq,
        [&](int e) {
            // clamped column, masked in the second lambda (col_insert)
            const int zq = e / CW, c = e - zq * CW, ix = min(x0 + c, J.nxx - 1);
            F4x2 r;
            r.a = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!TO_DVZ) r.a = ld4(J.dvx + cq_q4(J.nxx, ix, zq));
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int iz = 4 * zq + k;
                if (iz < nzz && (iz < pml || iz >= nzz - pml)) {
                    const int sz = iz < pml ? iz : iz - (nzz - 2 * pml);
                    f4set(r.b, k, J.psiz[(size_t)sz * J.nxx + ix]);
                }
            }
            return r;
        },
        [&](int e, const F4x2 &r) {
            const int zq = e / CW, c = e - zq * CW, ix = x0 + c;
            float *line = col_line(bufs, J.lsz, c);
            if (ix >= J.nxx) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (4 * zq + k < nzz) line[2 * (4 * zq + k)] = 0.f;
                return;
            }
            float4 dv = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int iz = 4 * zq + k;
                if (iz >= nzz) continue;
                float d = line[2 * iz];
                if (iz < pml || iz >= nzz - pml) {
                    const int sz = iz < pml ? iz : iz - (nzz - 2 * pml);
                    const float ps = J.ibz[iz] * f4get(r.b, k) + J.iaz[iz] * d;
                    J.psiz[(size_t)sz * J.nxx + ix] = ps;
                    d += ps;
                }
                const float sm = f4get(r.a, k) + d;
                f4set(dv, k, sm);
                line[2 * iz] = sm;
            }
            st4((TO_DVZ ? J.dvz : J.div) + cq_q4(J.nxx, ix, zq), dv);
        });
}

// After a forward z-FFT of packed columns (c0 + i c1): split into the two
// half spectra G0 = (Z[k] + conj Z[-k])/2, G1 = (Z[k] - conj Z[-k])/(2i),
// k in [0, nzh), and store them.  With mh != null also multiply Z by mh in
// place (each thread owns the pair {k, -k}).
template <int CW>
__device__ __forceinline__ void col_extract(const JobG &J, float2 *buf, int x0, float2 CQ_GLOBAL *__restrict__ S,
                                            const float2 CQ_GLOBAL *__restrict__ mh)
{
    constexpr int L = CW / 2;
    const int nzz = J.nzz, nzh = J.nzh, ldk = J.ldk;
    // lanes run over the column pairs of one kz, then the next kz
    cq_loop<4, true>(
        L * nzh,
        [&](int e) {
            float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
            if (mh != nullptr) {
                const int k = e / L;
                const float2 a = mh[k], b = mh[k ? nzz - k : 0];
                m = make_float4(a.x, a.y, b.x, b.y);
            }
            return m;
        },
        [&](int e, float4 m) {
            const int k = e / L, l = e - k * L;
            const int k2 = k ? nzz - k : 0;
            float2 *line = buf + l * J.lsz;
            const float2 zk = line[k], zm = line[k2];
            const float2 g0 = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            const float2 g1 = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
            const int ix0 = x0 + 2 * l;
            // columns ix0, ix0 + 1 share a block (ix0 even); with odd nxx the
            // last pair's second column is a block's spare column
            if (ix0 < J.nxx)
                st4v(S + cq_hs(ldk, ix0, k), make_float4(g0.x, g0.y, g1.x, g1.y));
            if (mh != nullptr) {
                line[k] = cmul(zk, make_float2(m.x, m.y));
                if (k2 != k) line[k2] = cmul(zm, make_float2(m.z, m.w));
            }
        });
}

// Inverse of col_extract: rebuild the packed full spectrum W = H0 + i H1 from
// the two half spectra (W[k] = H0[k] + i H1[k], W[-k] = conj H0[k] + i conj
// H1[k]; the imaginary parts at k = 0 and k = nzz/2 are dropped, as a real
// column's spectrum has none).
template <int CW>
__device__ __forceinline__ void col_insert(const JobG &J, float2 *buf, int x0, const float2 CQ_GLOBAL *__restrict__ S)
{
    constexpr int L = CW / 2;
    const int nzz = J.nzz, nzh = J.nzh, ldk = J.ldk;
    cq_loop<8, true>(
        L * nzh,
        [&](int e) {
            const int k = e / L, l = e - k * L;
            const int ix0 = min(x0 + 2 * l, (J.nxx - 1) & ~1);
            return *reinterpret_cast<const float4 CQ_GLOBAL *>(S + cq_hs(ldk, ix0, k));
        },
        [&](int e, float4 h) {
            const int k = e / L, l = e - k * L;
            const int k2 = k ? nzz - k : 0;
            // columns past nxx are masked here rather than branched round at
            // the load (the loads are issued unconditionally from a clamped
            // column, so that all of cq_loop's batch stays in flight)
            if (x0 + 2 * l >= J.nxx) h = make_float4(0.f, 0.f, 0.f, 0.f);
            if (x0 + 2 * l + 1 >= J.nxx) h.z = h.w = 0.f;
            float2 h0 = make_float2(h.x, h.y), h1 = make_float2(h.z, h.w);
            float2 *line = buf + l * J.lsz;
            if (k2 == k) {
                h0.y = 0.f;
                h1.y = 0.f;
            }
            line[k] = make_float2(h0.x - h1.y, h0.y + h1.x);
            if (k2 != k) line[k2] = make_float2(h0.x + h1.y, h1.x - h0.y);
        });
}

// Packed full spectrum of two real columns from their half spectra h0, h1
// at kz k (and -k = k2): W[k] = h0 + i h1, W[-k] = conj h0 + i conj h1; the
// imaginary parts at k = 0 and k = nzz/2 are dropped, as a real column's
// spectrum has none.
__device__ __forceinline__ void pack_pair(float2 *line, int k, int k2, float2 h0, float2 h1)
{
    if (k2 == k) {
        h0.y = 0.f;
        h1.y = 0.f;
    }
    line[k] = make_float2(h0.x - h1.y, h0.y + h1.x);
    if (k2 != k) line[k2] = make_float2(h0.x + h1.y, h1.x - h0.y);
}

// NC columns x0 .. x0 + NC - 1, one per line (line j at buf + j LS): the
// packed spectrum of the two real signals Sa and Sb of column x0 + j
// (inverse-transformed, Sa lands in the real and Sb in the imaginary part).
template <int NC>
__device__ __forceinline__ void col2_insert(const JobG &J, float2 *buf, int LS, int x0,
                                            const float2 CQ_GLOBAL *__restrict__ Sa,
                                            const float2 CQ_GLOBAL *__restrict__ Sb)
{
    const int nzz = J.nzz, nzh = J.nzh, ldk = J.ldk;
    const float inv = __builtin_amdgcn_rcpf((float)__builtin_amdgcn_readfirstlane(NC));   // opaque: see k_q_c2
    cq_loop<8, true>(
        NC * nzh,
        [&](int e) {
            const int k = cq_div(e, inv), ix = min(x0 + (e - k * NC), J.nxx - 1);   // as in col_insert
            const float2 a = Sa[cq_hs(ldk, ix, k)], b = Sb[cq_hs(ldk, ix, k)];
            return make_float4(a.x, a.y, b.x, b.y);
        },
        [&](int e, float4 h) {
            const int k = cq_div(e, inv), j = e - k * NC;
            if (x0 + j >= J.nxx) h = make_float4(0.f, 0.f, 0.f, 0.f);
            pack_pair(buf + j * LS, k, k ? nzz - k : 0, make_float2(h.x, h.y), make_float2(h.z, h.w));
        });
}

// NP column pairs x0 + 2 l, + 1 (line l at buf + l LS), two columns per
// line: col_insert / col_extract on explicit lines.
template <int NP>
__device__ __forceinline__ void colpair_insert(const JobG &J, float2 *buf, int LS, int x0,
                                               const float2 CQ_GLOBAL *__restrict__ S)
{
    const int nzz = J.nzz, nzh = J.nzh, ldk = J.ldk;
    const float inv = __builtin_amdgcn_rcpf((float)__builtin_amdgcn_readfirstlane(NP));
    cq_loop<8, true>(
        NP * nzh,
        [&](int e) {
            const int k = cq_div(e, inv), ix0 = min(x0 + 2 * (e - k * NP), (J.nxx - 1) & ~1);
            return *reinterpret_cast<const float4 CQ_GLOBAL *>(S + cq_hs(ldk, ix0, k));
        },
        [&](int e, float4 h) {
            const int k = cq_div(e, inv), l = e - k * NP;
            if (x0 + 2 * l >= J.nxx) h = make_float4(0.f, 0.f, 0.f, 0.f);   // as in col_insert
            if (x0 + 2 * l + 1 >= J.nxx) h.z = h.w = 0.f;
            pack_pair(buf + l * LS, k, k ? nzz - k : 0, make_float2(h.x, h.y), make_float2(h.z, h.w));
        });
}
template <int NP>
__device__ __forceinline__ void colpair_extract(const JobG &J, const float2 *buf, int LS, int x0,
                                                float2 CQ_GLOBAL *__restrict__ S)
{
    const int nzz = J.nzz, nzh = J.nzh, ldk = J.ldk;
    const float inv = __builtin_amdgcn_rcpf((float)__builtin_amdgcn_readfirstlane(NP));
    for (int e = threadIdx.x; e < NP * nzh; e += CQ_NT) {
        const int k = cq_div(e, inv), l = e - k * NP;
        const int k2 = k ? nzz - k : 0;
        const float2 *line = buf + l * LS;
        const float2 zk = line[k], zm = line[k2];
        const int ix0 = x0 + 2 * l;
        if (ix0 < J.nxx)
            st4v(S + cq_hs(ldk, ix0, k), make_float4(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y), 0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x)));
    }
}

// ===========================================================================
// kz-row helpers (half spectra, TK complex lines of length nxx)
// ===========================================================================
// kz rows k0 .. k0 + TK - 1 of a half spectrum: per block of 8 columns one
// contiguous run of TK x 8 values, read or written as 4 TK float4 pieces
// (piece w: row w / 4, columns 2 (w % 4), + 1 of the block).
template <int TK>
__device__ __forceinline__ void krow_fill(float2 *buf, const JobG &J, const float2 CQ_GLOBAL *__restrict__ S, int k0)
{
    constexpr int P = 4 * TK;
    const int nxx = J.nxx, LS = J.lsx;
    cq_loop<4>(
        ((nxx + 7) >> 3) * P,
        [&](int e) {
            const int b = e / P, w = e - b * P;
            return *reinterpret_cast<const float4 CQ_GLOBAL *>(S + ((size_t)b * J.ldk + k0 + (w >> 2)) * 8 +
                                                                2 * (w & 3));
        },
        [&](int e, float4 v) {
            const int b = e / P, w = e - b * P;
            const int ix = 8 * b + 2 * (w & 3);
            float2 *line = buf + (w >> 2) * LS;
            if (ix < nxx) line[ix] = make_float2(v.x, v.y);
            if (ix + 1 < nxx) line[ix + 1] = make_float2(v.z, v.w);
        });
}

template <int TK>
__device__ __forceinline__ void krow_store(const float2 *buf, const JobG &J, float2 CQ_GLOBAL *__restrict__ S, int k0)
{
    constexpr int P = 4 * TK;
    const int nxx = J.nxx, LS = J.lsx;
    const int total = ((nxx + 7) >> 3) * P;
    for (int e = threadIdx.x; e < total; e += CQ_NT) {
        const int b = e / P, w = e - b * P;
        const int ix = 8 * b + 2 * (w & 3);
        const float2 *line = buf + (w >> 2) * LS;
        const float2 a = ix < nxx ? line[ix] : make_float2(0.f, 0.f);
        const float2 c = ix + 1 < nxx ? line[ix + 1] : make_float2(0.f, 0.f);
        st4v(S + ((size_t)b * J.ldk + k0 + (w >> 2)) * 8 + 2 * (w & 3), make_float4(a.x, a.y, c.x, c.y));
    }
}

// Fractional-Laplacian symbols of the viscoacoustic step on the reference's
// float wavenumber grid (cal_k2d, sgpstd-qmodeling/func.c:550-583):
//   0: |k|^(1/16) = (kx^2+kz^2)^(1/32)   (k2d2)
//   1: |k|^2                              (k2d1)
//   2: |k|^(33/16)                        (k2d3)
//   3: |k|^1.01 = (kx^2+kz^2)^0.505       (k2da)
// as exp2(a log2 k2) on the hardware log2 / exp2 (v_log_f32, v_exp_f32, about
// 1 ulp): k2 is 0 (log2 -inf, symbol 0) or at least the smallest nonzero
// wavenumber squared, never denormal, so the library wrappers' denormal
// scaling is not needed.
__device__ __forceinline__ float frac_symbol(int which, float k2)
{
    if (which == 1) return k2;
    const float l = __builtin_amdgcn_logf(k2);
    if (which == 3) return __builtin_amdgcn_exp2f(0.505f * l);
    const float r = __builtin_amdgcn_exp2f(l * (1.0f / 32.0f));
    return which == 0 ? r : k2 * r;
}

// |k|^2 on the float wavenumber grid at kx (FFT order) and kz = kzi.
__device__ __forceinline__ float krow_k2(const JobG &J, int kx, int kzi)
{
    const int ikx = kx <= J.nxx / 2 ? kx : J.nxx - kx;
    const float fkx = (float)ikx * J.dkx;
    const float fkz = (float)kzi * J.dkz;
    return __fadd_rn(__fmul_rn(fkx, fkx), __fmul_rn(fkz, fkz));
}

// The multiplier of the previous step's p: symbol 3 (|k|^1.01), or, which = 4,
// the stabilised compensation multiplier of Q-RTM
//   S = k - 2 e k (1 + tc dt k) / (exp(-tc tt k) + e exp(tc dt k)),  k = |k|^1.01,
// e = eta * 1e-6, tc = max tao, tt = t_aux dt (multt_kernel,
// sgpstd-qrtm/acsglw.cu:1895-1924).
__device__ __forceinline__ float aux_symbol_of(const JobG &J, int which, float k2, float tt)
{
    if (which != 4) return frac_symbol(which, k2);
    const float k = frac_symbol(3, k2);
    const float es = J.eta * 1.0e-6f;
    const float a = J.taoc * J.dt * k;
    return k - 2.0f * es * k * (1.0f + a) / (expf(-J.taoc * tt * k) + es * expf(a));
}

// Fused R1, first round.  Lines 0.. hold the div spectrum, lines TK.. the p
// spectrum (TK lines each): p becomes aux = p x aux symbol / N in place and
// lines 2 TK.. receive K = div x |k|^2 / N.
template <int TK>
__device__ __forceinline__ void krow_round1(const JobG &J, float2 *buf, int k0, int which, float tt)
{
    const int nxx = J.nxx, LS = J.lsx;
    const float inv = __builtin_amdgcn_rcpf((float)nxx);
    for (int e = threadIdx.x; e < nxx * TK; e += CQ_NT) {
        const int l = cq_div(e, inv), kx = e - l * nxx, pe = l * LS + kx;
        const float k2 = krow_k2(J, kx, k0 + l);
        const float ms = aux_symbol_of(J, which, k2, tt) * J.invN, mk = k2 * J.invN;
        const float2 v = buf[pe], w = buf[TK * LS + pe];
        buf[TK * LS + pe] = make_float2(w.x * ms, w.y * ms);
        buf[2 * TK * LS + pe] = make_float2(v.x * mk, v.y * mk);
    }
}

// Fused R1, second round: the div spectrum (lines 0..) becomes A = div x
// |k|^(1/16) / N in place and lines TK.. receive B = div x |k|^(33/16) / N.
template <int TK>
__device__ __forceinline__ void krow_round2(const JobG &J, float2 *buf, int k0)
{
    const int nxx = J.nxx, LS = J.lsx;
    const float inv = __builtin_amdgcn_rcpf((float)nxx);
    for (int e = threadIdx.x; e < nxx * TK; e += CQ_NT) {
        const int l = cq_div(e, inv), kx = e - l * nxx, pe = l * LS + kx;
        const float k2 = krow_k2(J, kx, k0 + l);
        const float r = frac_symbol(0, k2);
        const float ma = r * J.invN, mb = (k2 * r) * J.invN;
        const float2 v = buf[pe];
        buf[pe] = make_float2(v.x * ma, v.y * ma);
        buf[TK * LS + pe] = make_float2(v.x * mb, v.y * mb);
    }
}

// Line dst + l (of TK) = spectrum line l x frac_symbol(which) / N.
template <int TK>
__device__ __forceinline__ void krow_mult_to(const JobG &J, float2 *buf, int dst, int k0, int which)
{
    const int nxx = J.nxx, LS = J.lsx;
    const float inv = __builtin_amdgcn_rcpf((float)nxx);
    for (int e = threadIdx.x; e < nxx * TK; e += CQ_NT) {
        const int l = cq_div(e, inv), kx = e - l * nxx, pe = l * LS + kx;
        const float m = frac_symbol(which, krow_k2(J, kx, k0 + l)) * J.invN;
        const float2 v = buf[pe];
        buf[dst * LS + pe] = make_float2(v.x * m, v.y * m);
    }
}

// buf[l][kx] *= aux_symbol_of(which) (kx, k0 + l) / N.
template <int TK>
__device__ __forceinline__ void krow_mult(const JobG &J, float2 *buf, int k0, int which, float tt = 0.f)
{
    const int nxx = J.nxx, LS = J.lsx;
    const float inv = __builtin_amdgcn_rcpf((float)nxx);
    for (int e = threadIdx.x; e < nxx * TK; e += CQ_NT) {
        const int l = cq_div(e, inv), kx = e - l * nxx, pe = l * LS + kx;
        const float m = aux_symbol_of(J, which, krow_k2(J, kx, k0 + l), tt) * J.invN;
        const float2 v = buf[pe];
        buf[pe] = make_float2(v.x * m, v.y * m);
    }
}

// ===========================================================================
// acoustic passes (sgpstd-qmodeling/acps4.cu:508-645)
// ===========================================================================
// Kernel arguments: the job table, the work items (table entry, tile origin)
// and the global step.  Each wavefield finds its own step record: t = time
// index of this step (wavelet sample for the source wavefield, data sample
// for the receiver wavefield), t_aux = time index the stabilised aux
// multiplier uses, flags = CQ_F_* bits.  It returns false when the wavefield
// does not advance at this global step (not started, finished, or the other
// role's turn in an RTM schedule).
__device__ __forceinline__ bool cq_step_of(const JobG &J, int step, int &t, int &t_aux, int &flags)
{
    const int s = step - J.t0;
    if (s < 0 || s >= J.nsched) return false;
    if (J.sched == nullptr) {
        t = s;
        t_aux = s - 1;
        flags = 0;
        return true;
    }
    const int4 r = J.sched[s];
    if (r.x != J.role) return false;
    t = r.y;
    t_aux = r.z;
    flags = r.w;
    return true;
}

// The job record is copied by value into registers (J): the kernel's global
// stores could alias it as far as the compiler knows, so a reference would
// make it reload the record's fields (scalar loads behind a wait for every
// outstanding store) in each loop iteration after a store.  The FFT plans,
// indexed by stage at run time, are read through the reference (Jr).
// A negative step argument (a launch recorded in a graph that is replayed
// for many steps) means: the global step is the launch group's counter.
#define CQ_ITEM_PROLOGUE                                             \
    const int2 it = items[blockIdx.x];                               \
    if (it.x < 0) return;                                            \
    const JobG &Jr = reinterpret_cast<const JobG *>(jobs)[it.x];     \
    const JobG J = Jr;                                               \
    int t, t_aux, flags;                                             \
    if (step < 0) step = *J.dstep;                                   \
    if (!cq_step_of(J, step, t, t_aux, flags)) return;               \
    (void)t_aux;                                                     \
    (void)flags;

// dst[interior(ix, iz0 + k)] += a.k * b.k for the 4 depths of a float4
// (image: rtm_image, sgpstd-qrtm/acps4.cu:1694-1721; illumination: addsa_kernel,
// :1656)
__device__ __forceinline__ void interior_accum(const JobG &J, float CQ_GLOBAL *dst, int ix, int iz0, float4 a, float4 b)
{
    const int xi = ix - J.pml;
    if (xi < 0 || xi >= J.nx) return;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int zi = iz0 + k - J.pml;
        if (zi >= 0 && zi < J.nz) dst[cq_q4(J.nxx, ix, iz0 >> 2) + k] += f4get(a, k) * f4get(b, k);
    }
}

// row pass: L = Lx + Lz, p update + source, dp/dx -> vx, dvx/dx -> dvx
template <int TZ>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_row_ac(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                  int step)
{
    constexpr int L = TZ / 2, Q = TZ / 4;
    CQ_ITEM_PROLOGUE
    const int nxx = J.nxx, z0 = it.y, LS = J.lsx;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlx;
    load_tw(tw, J.twx, J.twlx);
    row_fill<Q, true>(buf, nxx, J.lsx, J.div, z0);
    __syncthreads();
    cq_fft_fwd(Jr.px, buf, LS, L, tw, J.mx2);
    cq_fft_inv(Jr.px, buf, LS, L, tw);

    // p += c1 div - c2 (Lx + Lz); source (addsa_kernel, acps4.cu:1462-1498);
    // illumination ssg += p^2 during the first forward sweep of RTM
    // (sgpstd-qrtm/acps4.cu:1623-1659)
    const bool rcv = J.role == 1;
    const float src = rcv ? 0.f : J.rik[t] * J.src_scale;
    const float inv = __builtin_amdgcn_rcpf((float)nxx);
    const size_t g0 = (size_t)(z0 >> 2) * nxx * 4;
    cq_loop<2>(
        nxx * Q,
        [&](int e) {
            const size_t g = g0 + (size_t)e * 4;
            F4x5 r;
            r.a = ld4(J.p + g);
            r.b = ld4(J.c1 + g);
            r.c = ld4(J.c2 + g);
            r.d = ld4(J.lz + g);
            r.e = ld4c(J.div + g);
            return r;
        },
        [&](int e, const F4x5 &r) {
        const int q = cq_div(e, inv), ix = e - q * nxx;
        const size_t g = g0 + (size_t)e * 4;
        const float2 a = buf[(2 * q) * LS + ix], b = buf[(2 * q + 1) * LS + ix];
        const float4 p = r.a, c1 = r.b, c2 = r.c, lz = r.d, d = r.e;
        float4 pn;
        pn.x = p.x + c1.x * d.x - c2.x * (a.x + lz.x);
        pn.y = p.y + c1.y * d.y - c2.y * (a.y + lz.y);
        pn.z = p.z + c1.z * d.z - c2.z * (b.x + lz.z);
        pn.w = p.w + c1.w * d.w - c2.w * (b.y + lz.w);
        if (!rcv) {
            if (ix == J.xsg) {
                const int dz = J.zsg - (z0 + 4 * q);
                if (dz >= 0 && dz < 4) f4set(pn, dz, f4get(pn, dz) + src);
            }
            st4(J.p + g, pn);
            if (flags & CQ_F_SSG) interior_accum(J, J.ssg, ix, z0 + 4 * q, pn, pn);
        }
        buf[(2 * q) * LS + ix] = make_float2(pn.x, pn.y);
        buf[(2 * q + 1) * LS + ix] = make_float2(pn.z, pn.w);
        });
    __syncthreads();
    if (rcv) {
        // receiver wavefield: p = obs at the receiver cells (addsa_b_kernel,
        // sgpstd-qrtm/acps4.cu:1662-1692), then img += p_src p (rtm_image)
        const int tile = z0 / TZ;
        for (int r = J.inj_off[tile] + threadIdx.x; r < J.inj_off[tile + 1]; r += CQ_NT) {
            const int dz = J.inj_z[r] - z0;
            float *f = reinterpret_cast<float *>(buf + (dz >> 1) * LS + J.inj_x[r]) + (dz & 1);
            *f = J.obs[(size_t)t * J.nr + J.inj_i[r]];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nxx * Q; e += CQ_NT) {
            const int q = e / nxx, ix = e - q * nxx;
            const size_t g = cq_q4(nxx, ix, (z0 >> 2) + q);
            const float2 a = buf[(2 * q) * LS + ix], b = buf[(2 * q + 1) * LS + ix];
            const float4 pn = make_float4(a.x, a.y, b.x, b.y);
            st4(J.p + g, pn);
            interior_accum(J, J.img, ix, z0 + 4 * q, pn, ld4(J.psrc + g));
        }
        __syncthreads();
    }
    cq_fft_fwd(Jr.px, buf, LS, L, tw, J.mxp);
    cq_fft_inv(Jr.px, buf, LS, L, tw);
    row_vx_phase<Q>(J, buf, z0);
    __syncthreads();
    cq_fft_fwd(Jr.px, buf, LS, L, tw, J.mxm);
    cq_fft_inv(Jr.px, buf, LS, L, tw);
    row_dvx_phase<Q>(J, buf, z0);
}

// column pass: gather, dp/dz -> vz, dvz/dz -> div, Lz = d2/dz2 div
template <int CW>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_col_ac(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                  int step)
{
    constexpr int L = CW / 2;
    CQ_ITEM_PROLOGUE
    const int nzz = J.nzz, x0 = it.y, nq = (nzz + 3) >> 2;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlz;
    float *bufs = reinterpret_cast<float *>(buf);
    load_tw(tw, J.twz, J.twlz);
    col_fill<CW>(bufs, J, J.p, x0);
    __syncthreads();
    // receiver gather of the updated p (acps4.cu:611-621)
    if (J.gather) col_gather<CW>(J, x0, t, [&](int c, int z) { return col_line(bufs, J.lsz, c)[2 * z]; });
    cq_fft_fwd(Jr.pz, buf, J.lsz, L, tw, J.mzp);
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    col_vz_phase<CW>(J, bufs, x0);
    __syncthreads();
    cq_fft_fwd(Jr.pz, buf, J.lsz, L, tw, J.mzm);
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    col_div_phase<CW>(J, bufs, x0);
    __syncthreads();
    cq_fft_fwd(Jr.pz, buf, J.lsz, L, tw, J.mz2);
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    for (int e = threadIdx.x; e < CW * nq; e += CQ_NT) {
        const int zq = e / CW, c = e - zq * CW, ix = x0 + c;
        if (ix >= J.nxx) continue;
        const float *line = col_line(bufs, J.lsz, c);
        float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (4 * zq + k < nzz) f4set(o, k, line[2 * (4 * zq + k)]);
        st4(J.lz + cq_q4(J.nxx, ix, zq), o);
    }
}

// ===========================================================================
// viscoacoustic passes (sgpstd-qmodeling/acsglw.cu:511-683)
// One step = C1 R1 C2 R2 C3 R3; between steps the state is the reference's
// state after the step (p, vx, vz, CPML) plus quantities derived from it:
// div and the dvx/dvz CPML memory of the next step, and F_z p (s2).
//   C1 (cols) : F_z div -> s1
//   R1 (kz)   : F_x, x3 symbols -> A (s1), K (s3), B (s4); F_x p x |k|^1.01 -> aux (s2)
//   C2 (cols) : F_z^-1 A, K, B; total (sum_total, acsglw.cu:1213-1245); F_z total (s1)
//   R2 (kz)   : F_x total x |k|^1.01 -> T (s1)
//   C3 (cols) : F_z^-1 T, aux; p update + source (acsglw.cu:1500-1537); gather;
//               F_z p -> s2; dp/dz -> vz; dvz/dz -> dvz
//   R3 (rows) : dp/dx -> vx; dvx/dx; div = dvx + dvz
// ===========================================================================
template <int CW>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_q_c1(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                int step)
{
    constexpr int L = CW / 2;
    CQ_ITEM_PROLOGUE
    const int x0 = it.y;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlz;
    float *bufs = reinterpret_cast<float *>(buf);
    load_tw(tw, J.twz, J.twlz);
    col_fill<CW>(bufs, J, J.div, x0);
    __syncthreads();
    cq_fft_fwd(Jr.pz, buf, J.lsz, L, tw, nullptr);
    col_extract<CW>(J, buf, x0, J.s1, nullptr);
}

template <int TK>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_q_r1(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                int step)
{
    CQ_ITEM_PROLOGUE
    const int k0 = it.y, LS = J.lsx;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlx;
    load_tw(tw, J.twx, J.twlx);
    // A = |k|^(1/16) div, K = |k|^2 div, B = |k|^(33/16) div (acsglw.cu:553-566);
    // aux = |k|^1.01 p of the previous step (acsglw.cu:623-631)
    const int aux_symbol = J.stab ? 4 : 3;
    const float tt = (float)t_aux * J.dt;
    if (J.r1_fused) {
        // 3 TK lines: the div and p spectra transformed together, then two
        // rounds of two symbols each, inverse-transformed together
        krow_fill<TK>(buf, J, J.s1, k0);
        krow_fill<TK>(buf + TK * LS, J, J.s2, k0);
        __syncthreads();
        cq_fft_fwd(Jr.px, buf, LS, 2 * TK, tw, nullptr);
        krow_round1<TK>(J, buf, k0, aux_symbol, tt);   // aux (lines TK..), K (lines 2 TK..)
        __syncthreads();
        cq_fft_inv(Jr.px, buf + TK * LS, LS, 2 * TK, tw);
        krow_store<TK>(buf + TK * LS, J, J.s2, k0);
        krow_store<TK>(buf + 2 * TK * LS, J, J.s3, k0);
        __syncthreads();
        krow_round2<TK>(J, buf, k0);                   // A (lines 0..), B (lines TK..)
        __syncthreads();
        cq_fft_inv(Jr.px, buf, LS, 2 * TK, tw);
        krow_store<TK>(buf, J, J.s1, k0);
        krow_store<TK>(buf + TK * LS, J, J.s4, k0);
        return;
    }
    krow_fill<TK>(buf, J, J.s1, k0);
    __syncthreads();
    cq_fft_fwd(Jr.px, buf, LS, TK, tw, nullptr);
    {
        // one symbol at a time (long lines: 2 TK lines of LDS)
        float2 CQ_GLOBAL *outs[3] = {J.s1, J.s3, J.s4};
        for (int w = 0; w < 3; w++) {
            krow_mult_to<TK>(J, buf, TK, k0, w);
            __syncthreads();
            cq_fft_inv(Jr.px, buf + TK * LS, LS, TK, tw);
            krow_store<TK>(buf + TK * LS, J, outs[w], k0);
            __syncthreads();
        }
    }
    __syncthreads();
    krow_fill<TK>(buf, J, J.s2, k0);
    __syncthreads();
    cq_fft_fwd(Jr.px, buf, LS, TK, tw, nullptr);
    krow_mult<TK>(J, buf, k0, aux_symbol, tt);
    __syncthreads();
    cq_fft_inv(Jr.px, buf, LS, TK, tw);
    krow_store<TK>(buf, J, J.s2, k0);
}

// ===========================================================================
// low-rank form of the viscoacoustic operator (jobs with lr_rank > 0):
//   total = sum_i w_i(x) F^-1[s_i(k) F div]
// in place of sum_total (the apply of sgpstd-qmodeling/acps4.cu:1337-1459:
// multv_kernel, then sumlw / sumlw2 with the A factor folded into w_i).
//   LR1 (kz rows) : aux as R1; F_x div once, then per rank x s_i, F_x^-1 -> lrs[i]
//   LR2 (columns) : per rank F_z^-1 lrs[i], total += w_i x; F_z total -> s1
// ===========================================================================
// Line dst + l (of TK) = spectrum line l x s_i(|kx|, kz) (symbols already
// divided by N); padding kz rows (kz >= nzh) get zero.
template <int TK>
__device__ __forceinline__ void krow_lr_mult(const JobG &J, float2 *buf, int dst, int k0, int i)
{
    const int nxx = J.nxx, LS = J.lsx;
    const float inv = __builtin_amdgcn_rcpf((float)nxx);
    const float CQ_GLOBAL *v = J.lrv + (size_t)i * J.nxh * J.nzh;
    for (int e = threadIdx.x; e < nxx * TK; e += CQ_NT) {
        const int l = cq_div(e, inv), kx = e - l * nxx, pe = l * LS + kx, kz = k0 + l;
        const int ikx = kx <= nxx / 2 ? kx : nxx - kx;
        const float m = kz < J.nzh ? v[(size_t)ikx * J.nzh + kz] : 0.f;
        const float2 w = buf[pe];
        buf[dst * LS + pe] = make_float2(w.x * m, w.y * m);
    }
}

template <int TK>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_lr_r1(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                 int step)
{
    CQ_ITEM_PROLOGUE
    const int k0 = it.y, LS = J.lsx;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlx;
    load_tw(tw, J.twx, J.twlx);
    // aux = |k|^1.01 p of the previous step (stabilised for Q-RTM), as k_q_r1
    krow_fill<TK>(buf, J, J.s2, k0);
    __syncthreads();
    cq_fft_fwd(Jr.px, buf, LS, TK, tw, nullptr);
    krow_mult<TK>(J, buf, k0, J.stab ? 4 : 3, (float)t_aux * J.dt);
    __syncthreads();
    cq_fft_inv(Jr.px, buf, LS, TK, tw);
    krow_store<TK>(buf, J, J.s2, k0);
    __syncthreads();
    // the div spectrum stays in lines 0..TK; each rank's product goes
    // through lines TK..2 TK
    krow_fill<TK>(buf, J, J.s1, k0);
    __syncthreads();
    cq_fft_fwd(Jr.px, buf, LS, TK, tw, nullptr);
    const size_t hsz = (size_t)cq_hs_cols(J.nxx) * J.ldk;
    for (int i = 0; i < J.lr_rank; i++) {
        krow_lr_mult<TK>(J, buf, TK, k0, i);
        __syncthreads();
        cq_fft_inv(Jr.px, buf + TK * LS, LS, TK, tw);
        krow_store<TK>(buf + TK * LS, J, J.lrs + hsz * i, k0);
        __syncthreads();
    }
}

template <int CW>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_lr_c2(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                 int step)
{
    constexpr int L = CW / 2;
    CQ_ITEM_PROLOGUE
    const int nzz = J.nzz, x0 = it.y, nq = (nzz + 3) >> 2;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlz;
    float *bufs = reinterpret_cast<float *>(buf);
    load_tw(tw, J.twz, J.twlz);
    const size_t hsz = (size_t)cq_hs_cols(J.nxx) * J.ldk, fsz = (size_t)J.nxx * J.ldz;
    const int r = J.lr_rank;
    for (int i = 0; i < r; i++) {
        col_insert<CW>(J, buf, x0, J.lrs + hsz * i);
        __syncthreads();
        cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
        // total accumulates in the field `tot`; each thread re-reads only the
        // quads it wrote; the last rank leaves total in the lines as well
        const bool first = i == 0, last = i == r - 1;
        cq_loop<2, true>(
            CW * nq,
            [&](int e) {
                const int zq = e / CW, c = e - zq * CW;
                const size_t g = cq_q4(J.nxx, min(x0 + c, J.nxx - 1), zq);
                F4x2 v;
                v.a = ld4(J.lrw + fsz * i + g);
                v.b = first ? make_float4(0.f, 0.f, 0.f, 0.f) : ld4c(J.tot + g);
                return v;
            },
            [&](int e, const F4x2 &v) {
                const int zq = e / CW, c = e - zq * CW, ix = x0 + c;
                float *line = col_line(bufs, J.lsz, c);
                if (ix >= J.nxx) {
                    if (last)
#pragma unroll
                        for (int k = 0; k < 4; k++)
                            if (4 * zq + k < nzz) line[2 * (4 * zq + k)] = 0.f;
                    return;
                }
                float4 acc = v.b;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int iz = 4 * zq + k;
                    if (iz >= nzz) continue;
                    const float tv = f4get(acc, k) + f4get(v.a, k) * line[2 * iz];
                    f4set(acc, k, tv);
                    if (last) line[2 * iz] = tv;
                }
                st4c(J.tot + cq_q4(J.nxx, ix, zq), acc);
            });
        __syncthreads();
    }
    cq_fft_fwd(Jr.pz, buf, J.lsz, L, tw, nullptr);
    col_extract<CW>(J, buf, x0, J.s1, nullptr);
}

template <int CW>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_q_c2(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                int step)
{
    constexpr int L = CW / 2;
    CQ_ITEM_PROLOGUE
    const int nzz = J.nzz, x0 = it.y, nq = (nzz + 3) >> 2;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlz;
    float *bufs = reinterpret_cast<float *>(buf);
    load_tw(tw, J.twz, J.twlz);
    // total = y1 div + y2 A - (dt^2 v^2 lam/12)(y3 K + y4 B), with
    // y1 = 1-32 gam, y2 = 32 gam vwd, y3 = 1-64 gam, y4 = 64 gam vwd = 2 y2.
    if constexpr (CW >= 24) {
        // Two halves of NC columns (narrower tiles keep the one-symbol
        // schedule below: their halves would transform too few lines).  K and B of one column share a line
        // (real and imaginary part), so the partial sum
        //   P = cc (y3 K + y4 B) - y1 div
        // of a cell is formed in LDS; P of column pairs is then packed into
        // the even lines, A goes into the odd lines, total = y2 A - P.
        constexpr int NC = CW / 2, NP = NC / 2;
        const int LS = J.lsz;
        // item -> (quad, column) by the reciprocal of an opaque NC: with a
        // compile-time power of two the compiler hoists per-thread offsets
        // of every loop out of the half loop and runs out of registers
        const float inv_nc = __builtin_amdgcn_rcpf((float)__builtin_amdgcn_readfirstlane(NC));
#pragma unroll 1
        for (int h = 0; h < 2; h++) {
            const int xh = x0 + h * NC;
            if (xh >= J.nxx) break;
            col2_insert<NC>(J, buf, LS, xh, J.s3, J.s4);   // K + i B
            __syncthreads();
            cq_fft_inv(Jr.pz, buf, LS, NC, tw);
            cq_loop<2, true>(
                NC * nq,
                [&](int e) {
                    // clamped column, masked in the second lambda (col_insert)
                    const int zq = cq_div(e, inv_nc), ix = min(xh + (e - zq * NC), J.nxx - 1);
                    const size_t g = cq_q4(J.nxx, ix, zq);
                    F4x5 r;
                    r.a = ld4c(J.qg + g);
                    r.b = ld4c(J.qy2 + g);
                    r.c = ld4c(J.qcc + g);
                    r.d = ld4c(J.div + g);
                    return r;
                },
                [&](int e, const F4x5 &r) {
                    const int zq = cq_div(e, inv_nc), j = e - zq * NC;
                    if (xh + j >= J.nxx) return;
                    float2 *line = buf + j * LS;
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int iz = 4 * zq + k;
                        if (iz >= nzz) continue;
                        const float gm = f4get(r.a, k), y2 = f4get(r.b, k);
                        const float2 kb = line[iz];
                        line[iz].x = f4get(r.c, k) * ((1.0f - 64.0f * gm) * kb.x + 2.0f * y2 * kb.y) -
                                     (1.0f - 32.0f * gm) * f4get(r.d, k);
                    }
                });
            __syncthreads();
            const float inv_z = __builtin_amdgcn_rcpf((float)nzz);
            for (int e = threadIdx.x; e < NP * nzz; e += CQ_NT) {
                const int pr = cq_div(e, inv_z), z = e - pr * nzz;
                buf[2 * pr * LS + z].y = buf[(2 * pr + 1) * LS + z].x;
            }
            __syncthreads();
            colpair_insert<NP>(J, buf + LS, 2 * LS, xh, J.s1);   // A
            __syncthreads();
            cq_fft_inv(Jr.pz, buf + LS, 2 * LS, NP, tw);
            cq_loop<4, true>(
                NC * nq,
                [&](int e) {
                    const int zq = cq_div(e, inv_nc), ix = min(xh + (e - zq * NC), J.nxx - 1);
                    return ld4c(J.qy2 + cq_q4(J.nxx, ix, zq));
                },
                [&](int e, float4 y2) {
                    const int zq = cq_div(e, inv_nc), j = e - zq * NC, ix = xh + j;
                    float *la = reinterpret_cast<float *>(buf + (2 * (j >> 1) + 1) * LS) + (j & 1);
                    const float *lp = reinterpret_cast<const float *>(buf + 2 * (j >> 1) * LS) + (j & 1);
                    if (ix >= J.nxx) {
#pragma unroll
                        for (int k = 0; k < 4; k++)
                            if (4 * zq + k < nzz) la[2 * (4 * zq + k)] = 0.f;
                        return;
                    }
                    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int iz = 4 * zq + k;
                        if (iz >= nzz) continue;
                        const float tv = f4get(y2, k) * la[2 * iz] - lp[2 * iz];
                        f4set(o, k, tv);
                        la[2 * iz] = tv;
                    }
                    st4(J.tot + cq_q4(J.nxx, ix, zq), o);
                });
            __syncthreads();
            cq_fft_fwd(Jr.pz, buf + LS, 2 * LS, NP, tw, nullptr);
            colpair_extract<NP>(J, buf + LS, 2 * LS, xh, J.s1);
            __syncthreads();
        }
        return;
    }
    // Two columns per line, one symbol at a time: the partial sums live in
    // `tot` itself; each thread re-reads only the elements it wrote.
    col_insert<CW>(J, buf, x0, J.s3);   // K
    __syncthreads();
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    // column c of the tile: ix = x0 + c; quads beyond nxx are skipped
    auto cell = [&](int e, int &zq, int &c) {
        zq = e / CW;
        c = e - zq * CW;
        return x0 + c < J.nxx;
    };
    cq_loop<4, true>(
        CW * nq,
        [&](int e) {
            int zq, c;
            float4 gm = make_float4(0.f, 0.f, 0.f, 0.f);
            if (cell(e, zq, c)) gm = ld4c(J.qg + cq_q4(J.nxx, x0 + c, zq));
            return gm;
        },
        [&](int e, float4 gm) {
            int zq, c;
            if (!cell(e, zq, c)) return;
            const float *line = col_line(bufs, J.lsz, c);
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int iz = 4 * zq + k;
                if (iz < nzz) f4set(a, k, (1.0f - 64.0f * f4get(gm, k)) * line[2 * iz]);
            }
            st4(J.tot + cq_q4(J.nxx, x0 + c, zq), a);
        });
    __syncthreads();
    col_insert<CW>(J, buf, x0, J.s4);   // B
    __syncthreads();
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    cq_loop<4, true>(
        CW * nq,
        [&](int e) {
            int zq, c;
            F4x3 r;
            if (cell(e, zq, c)) {
                const size_t g = cq_q4(J.nxx, x0 + c, zq);
                r.a = ld4c(J.qy2 + g);
                r.b = ld4c(J.qcc + g);
                r.c = ld4c(J.tot + g);
            }
            return r;
        },
        [&](int e, const F4x3 &r) {
            int zq, c;
            if (!cell(e, zq, c)) return;
            const float *line = col_line(bufs, J.lsz, c);
            const float4 y2 = r.a, cc = r.b;
            float4 a = r.c;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int iz = 4 * zq + k;
                if (iz < nzz) f4set(a, k, f4get(cc, k) * (f4get(a, k) + 2.0f * f4get(y2, k) * line[2 * iz]));
            }
            st4(J.tot + cq_q4(J.nxx, x0 + c, zq), a);
        });
    __syncthreads();
    col_insert<CW>(J, buf, x0, J.s1);   // A
    __syncthreads();
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    cq_loop<3, true>(
        CW * nq,
        [&](int e) {
            int zq, c;
            F4x5 r;
            if (cell(e, zq, c)) {
                const size_t g = cq_q4(J.nxx, x0 + c, zq);
                r.a = ld4c(J.qg + g);
                r.b = ld4c(J.qy2 + g);
                r.c = ld4c(J.div + g);
                r.d = ld4c(J.tot + g);
            }
            return r;
        },
        [&](int e, const F4x5 &r) {
            int zq, c;
            const bool in = cell(e, zq, c);
            float *line = col_line(bufs, J.lsz, c);
            if (!in) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (4 * zq + k < nzz) line[2 * (4 * zq + k)] = 0.f;
                return;
            }
            const float4 gm = r.a, y2 = r.b, d = r.c, a = r.d;
            float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int iz = 4 * zq + k;
                if (iz >= nzz) continue;
                const float tv = (1.0f - 32.0f * f4get(gm, k)) * f4get(d, k) + f4get(y2, k) * line[2 * iz] -
                                 f4get(a, k);
                f4set(o, k, tv);
                line[2 * iz] = tv;
            }
            st4(J.tot + cq_q4(J.nxx, x0 + c, zq), o);
        });
    __syncthreads();
    cq_fft_fwd(Jr.pz, buf, J.lsz, L, tw, nullptr);
    col_extract<CW>(J, buf, x0, J.s1, nullptr);
}

template <int TK>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_q_r2(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                int step)
{
    CQ_ITEM_PROLOGUE
    const int k0 = it.y;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlx;
    load_tw(tw, J.twx, J.twlx);
    krow_fill<TK>(buf, J, J.s1, k0);
    __syncthreads();
    cq_fft_fwd(Jr.px, buf, J.lsx, TK, tw, nullptr);
    krow_mult<TK>(J, buf, k0, 3);   // T = |k|^1.01 total (acsglw.cu:578-588)
    __syncthreads();
    cq_fft_inv(Jr.px, buf, J.lsx, TK, tw);
    krow_store<TK>(buf, J, J.s1, k0);
}

template <int CW>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_q_c3(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                int step)
{
    constexpr int L = CW / 2;
    CQ_ITEM_PROLOGUE
    const int nzz = J.nzz, x0 = it.y, nq = (nzz + 3) >> 2;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlz;
    float *bufs = reinterpret_cast<float *>(buf);
    load_tw(tw, J.twz, J.twlz);
    col_insert<CW>(J, buf, x0, J.s1);   // T
    __syncthreads();
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    // T waits in the scratch field (each thread re-reads only what it wrote)
    for (int e = threadIdx.x; e < CW * nq; e += CQ_NT) {
        const int zq = e / CW, c = e - zq * CW, ix = x0 + c;
        if (ix >= J.nxx) continue;
        const float *line = col_line(bufs, J.lsz, c);
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int iz = 4 * zq + k;
            if (iz < nzz) f4set(a, k, line[2 * iz]);
        }
        st4c(J.scr + cq_q4(J.nxx, ix, zq), a);
    }
    __syncthreads();
    col_insert<CW>(J, buf, x0, J.s2);   // aux of the previous step
    __syncthreads();
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    // p += rho v^2 dt lam (total - s dt tao T / 2) - s dt tao aux; source.
    // s = +1 attenuates (acsglw.cu:1500-1537); s = -1 compensates (Q-RTM,
    // sgpstd-qrtm/acsglw.cu:1769-1841).  The new p replaces aux in the
    // packed columns (each thread overwrites the values it read).
    const bool rcv = J.role == 1;
    const float src = rcv ? 0.f : J.rik[t] * J.src_scale;
    const float s = J.qsign;
    struct C3In {
        float4 p, tot, qa, qb, tt;
    };
    cq_loop<2, true>(
        CW * nq,
        [&](int e) {
            // clamped column, masked in the second lambda (col_insert)
            const int zq = e / CW, c = e - zq * CW;
            const size_t g = cq_q4(J.nxx, min(x0 + c, J.nxx - 1), zq);
            C3In r;
            r.p = ld4(J.p + g);
            r.tot = ld4(J.tot + g);
            r.qa = ld4(J.qa + g);
            r.qb = ld4(J.qb + g);
            r.tt = ld4c(J.scr + g);
            return r;
        },
        [&](int e, const C3In &r) {
            const int zq = e / CW, c = e - zq * CW, ix = x0 + c;
            float *line = col_line(bufs, J.lsz, c);
            if (ix >= J.nxx) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (4 * zq + k < nzz) line[2 * (4 * zq + k)] = 0.f;
                return;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int iz = 4 * zq + k;
                if (iz >= nzz) continue;
                // dt tao = 2 qb exactly (see build_job)
                const float qb = f4get(r.qb, k);
                float v = f4get(r.p, k) + f4get(r.qa, k) * (f4get(r.tot, k) - s * qb * f4get(r.tt, k)) -
                          s * (2.0f * qb) * line[2 * iz];
                if (!rcv && ix == J.xsg && iz == J.zsg) v += src;
                line[2 * iz] = v;
            }
        });
    __syncthreads();
    if (rcv) {
        // receiver wavefield: p = obs at the receiver cells (addsa_b_kernel,
        // sgpstd-qrtm/acsglw.cu:1809-1841)
        const int tile = x0 / CW;
        for (int r = J.inj_off[tile] + threadIdx.x; r < J.inj_off[tile + 1]; r += CQ_NT)
            col_line(bufs, J.lsz, J.inj_x[r] - x0)[2 * J.inj_z[r]] = J.obs[(size_t)t * J.nr + J.inj_i[r]];
        __syncthreads();
    }
    for (int e = threadIdx.x; e < CW * nq; e += CQ_NT) {
        const int zq = e / CW, c = e - zq * CW, ix = x0 + c;
        if (ix >= J.nxx) continue;
        const float *line = col_line(bufs, J.lsz, c);
        const size_t g = cq_q4(J.nxx, ix, zq);
        float4 pn = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (4 * zq + k < nzz) f4set(pn, k, line[2 * (4 * zq + k)]);
        st4(J.p + g, pn);
        if (rcv) interior_accum(J, J.img, ix, 4 * zq, pn, ld4(J.psrc + g));
        else if (flags & CQ_F_SSG) interior_accum(J, J.ssg, ix, 4 * zq, pn, pn);
    }
    if (J.gather) col_gather<CW>(J, x0, t, [&](int c, int z) { return col_line(bufs, J.lsz, c)[2 * z]; });
    __syncthreads();
    cq_fft_fwd(Jr.pz, buf, J.lsz, L, tw, nullptr);
    col_extract<CW>(J, buf, x0, J.s2, J.mzp);   // F_z p for the next aux; x dp/dz multiplier
    __syncthreads();
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    col_vz_phase<CW>(J, bufs, x0);
    __syncthreads();
    cq_fft_fwd(Jr.pz, buf, J.lsz, L, tw, J.mzm);
    cq_fft_inv(Jr.pz, buf, J.lsz, L, tw);
    col_div_phase<CW, true>(J, bufs, x0);
}

template <int TZ>
__global__ __launch_bounds__(CQ_NT, CQ_WAVES_PER_SIMD) void k_q_r3(const JobD *__restrict__ jobs, const int2 *__restrict__ items,
                                                int step)
{
    constexpr int L = TZ / 2, Q = TZ / 4;
    CQ_ITEM_PROLOGUE
    const int nxx = J.nxx, z0 = it.y;
    float2 *const lds = cq_lds;
    float2 *tw = lds, *buf = lds + J.twlx;
    load_tw(tw, J.twx, J.twlx);
    row_fill<Q>(buf, nxx, J.lsx, J.p, z0);
    __syncthreads();
    cq_fft_fwd(Jr.px, buf, J.lsx, L, tw, J.mxp);
    cq_fft_inv(Jr.px, buf, J.lsx, L, tw);
    row_vx_phase<Q>(J, buf, z0);
    __syncthreads();
    cq_fft_fwd(Jr.px, buf, J.lsx, L, tw, J.mxm);
    cq_fft_inv(Jr.px, buf, J.lsx, L, tw);
    row_dvx_phase<Q, true>(J, buf, z0);
}
