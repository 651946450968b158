// Tile bodies shared by the evaluation kernels (one launch per call) and the persistent batched
// kernel: a "C tile" keeps NC candidates per lane in VGPRs and streams zeros through LDS (every
// lane reads the same zero record, an LDS broadcast); a "Z tile" keeps CZ candidates in LDS and
// gives each lane its own zeros, then reduces the per-lane sums with wavefront shuffles.  C tiles
// serve long candidate lists, Z tiles the short requests of a batch, where a C tile would leave
// most lanes idle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "k2_device.h"

namespace k2dev {

#define K2_WG 256   // lanes per workgroup = zeros per LDS tile

// workgroups are dealt round-robin to the 8 XCDs; this renumbers them so that consecutive
// linear indices run on one XCD (and share its L2)
__device__ __forceinline__ int xcd_linear_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    return (nb % 8 == 0) ? ((b & 7) * (nb >> 3) + (b >> 3)) : b;
}

// The split, candidate tile and zero range [z0, z1) of this workgroup of a C-tile grid over
// (split, candidate tile).  Uniform splits (n_long = 0) of zps zeros from z_begin, in
// xcd_linear_block's order.  Guided splits (the host's guided_splits: the first n_long of
// zps_long zeros, then zps each, up to z_end; the grid a multiple of 8 splits): XCD x, which gets
// every 8th block, walks splits x, x + 8, ... in order, each over all candidate tiles, so that
// every XCD runs its share of the long splits first and of the short ones last.  z0 >= z_end
// marks a padding split past the end, which writes no row.
struct BlockSplit {
    int split, ct, z0, z1;
};
__device__ __forceinline__ BlockSplit block_split(int n_ctiles, int z_begin, int z_end, int zps,
                                                  int n_long, int zps_long) {
    BlockSplit r;
    if (n_long > 0) {
        const int b = blockIdx.x, k = b >> 3, j = k / n_ctiles;
        r.ct = k - j * n_ctiles;
        r.split = 8 * j + (b & 7);
    } else {
        const int L = xcd_linear_block();
        r.split = L / n_ctiles;
        r.ct = L - r.split * n_ctiles;
    }
    const bool lng = r.split < n_long;
    r.z0 = z_begin + (lng ? r.split * zps_long : n_long * zps_long + (r.split - n_long) * zps);
    r.z1 = min(z_end, r.z0 + (lng ? zps_long : zps));
    return r;
}

// one k's weighted zero arrays (padded with zero weights past the table end)
struct ZeroView {
    const uint4* g;
    const double2* a;   // 2 e^{-k g^2} w
    const double2* b;   // g a
    const double2* e;   // g^2 a
    const float2* af;
    const float2* bf;
    const float2* ef;
    const float4* gf;   // g0..g3 times 2 pi / 2^44 as floats (the FP32 window's column 3)
};

// ---- accumulators -------------------------------------------------------------------------
template <bool F32> struct Acc;

template <> struct Acc<false> {
    double s0, s1, s2;
    __device__ __forceinline__ void zero() { s0 = s1 = s2 = 0.0; }
    __device__ __forceinline__ double get(int q) const { return q == 0 ? s0 : (q == 1 ? s1 : s2); }
};

// FP32 sums with Kahan compensation (comp holds the negated lost low part)
__device__ __forceinline__ void kahan_add(float& sum, float& comp, float x) {
    const float y = x - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

template <> struct Acc<true> {
    float s0, c0, s1, c1, s2, c2;
    __device__ __forceinline__ void zero() { s0 = c0 = s1 = c1 = s2 = c2 = 0.0f; }
    __device__ __forceinline__ double get(int q) const {
        return q == 0 ? (double)s0 - (double)c0
                      : (q == 1 ? (double)s1 - (double)c1 : (double)s2 - (double)c2);
    }
};

// FP64 sums with error-free accumulation (enclosure mode): h + l carries the sum, TwoSum
// (Knuth) moves each addition's rounding error into l exactly
__device__ __forceinline__ void two_sum_acc(double& h, double& l, double p) {
    const double s = h + p;
    const double bp = s - h;
    const double e = (h - (s - bp)) + (p - bp);
    h = s;
    l += e;
}

struct AccDD {
    double h0, l0, h1, l1;
    __device__ __forceinline__ void zero() { h0 = l0 = h1 = l1 = 0.0; }
};

// one term: K += Re(a e^{i th}), K' += Re(i b e^{i th}), K'' -= Re(e e^{i th})
template <int FL>
__device__ __forceinline__ void accumulate(Acc<false>& A, double S, double C, double2 a,
                                           double2 b, double2 e) {
    A.s0 = fma(a.x, C, A.s0);
    A.s0 = fma(-a.y, S, A.s0);
    if (FL & K2_D1) {
        A.s1 = fma(-b.x, S, A.s1);
        A.s1 = fma(-b.y, C, A.s1);
    }
    if (FL & K2_D2) {
        A.s2 = fma(-e.x, C, A.s2);
        A.s2 = fma(e.y, S, A.s2);
    }
}

template <int FL>
__device__ __forceinline__ void accumulate(AccDD& A, double S, double C, double2 a, double2 b,
                                           double2 e) {
    two_sum_acc(A.h0, A.l0, fma(a.x, C, -a.y * S));
    if (FL & K2_D1) two_sum_acc(A.h1, A.l1, -fma(b.x, S, b.y * C));
}

// Grouped error-free sums (enclosure mode): the products of up to K2_ENC_GROUP consecutive zeros
// go into a plain FP64 group sum by FMAs (2 K2_ENC_GROUP roundings per group, bounded in the
// radius), and each group sum enters hi + lo by TwoSum (exact).  flush() ends a group.
#define K2_ENC_GROUP 8
struct AccGD {
    double g0, g1, h0, l0, h1, l1;
    __device__ __forceinline__ void zero() { g0 = g1 = h0 = l0 = h1 = l1 = 0.0; }
    template <int FL>
    __device__ __forceinline__ void flush() {
        two_sum_acc(h0, l0, g0);
        g0 = 0.0;
        if (FL & K2_D1) {
            two_sum_acc(h1, l1, g1);
            g1 = 0.0;
        }
    }
};

template <int FL>
__device__ __forceinline__ void accumulate(AccGD& A, double S, double C, double2 a, double2 b,
                                           double2 e) {
    A.g0 = fma(a.x, C, A.g0);
    A.g0 = fma(-a.y, S, A.g0);
    if (FL & K2_D1) {
        A.g1 = fma(-b.x, S, A.g1);
        A.g1 = fma(-b.y, C, A.g1);
    }
}

// zeros per group of an accumulator (0: no groups)
template <class ACC> struct AccGroup { static constexpr int value = 0; };
template <> struct AccGroup<AccGD> { static constexpr int value = K2_ENC_GROUP; };

// one zero's FP32 terms (K, K', K'') added into a group sum with two FMAs each
template <int FL>
__device__ __forceinline__ void add_terms(float3& p, float S, float C, float2 a, float2 b,
                                          float2 e) {
    p.x = fmaf(-a.y, S, fmaf(a.x, C, p.x));
    if (FL & K2_D1) p.y = fmaf(-b.y, C, fmaf(-b.x, S, p.y));
    if (FL & K2_D2) p.z = fmaf(-e.x, C, fmaf(e.y, S, p.z));
}

// a group sum into the compensated per-lane sums
template <int FL>
__device__ __forceinline__ void add_group(Acc<true>& A, const float3& p) {
    kahan_add(A.s0, A.c0, p.x);
    if (FL & K2_D1) kahan_add(A.s1, A.c1, p.y);
    if (FL & K2_D2) kahan_add(A.s2, A.c2, p.z);
}

// ---- C tile ---------------------------------------------------------------------------------
// LDS: two buffers of K2_WG zero records holding only what the requested sums read
// (FP64: g 16 B, a 16 B, b 16 B with K2', e 16 B with K2''; FP32: g, a|b 16 B, e 8 B).
template <int FL>
struct CTileSmem64 {
    static constexpr int NB = (FL & K2_D1) ? 1 : 0, NE = (FL & K2_D2) ? 1 : 0;
    // one buffer's records together, so that a zero's fields sit at fixed offsets from one
    // address (one address add per zero in the loop)
    struct Buf {
        uint4 g[K2_WG];
        double2 w[(1 + NB + NE) * K2_WG];   // a, then b, then e
    };
    Buf bf[2];
    __device__ __forceinline__ uint4& gz(int buf, int z) { return bf[buf].g[z]; }
    __device__ __forceinline__ double2& a(int buf, int z) { return bf[buf].w[z]; }
    __device__ __forceinline__ double2& b(int buf, int z) { return bf[buf].w[K2_WG + z]; }
    __device__ __forceinline__ double2& e(int buf, int z) {
        return bf[buf].w[(1 + NB) * K2_WG + z];
    }
};

template <int FL>
struct CTileSmem32 {
    static constexpr int NE = (FL & K2_D2) ? 1 : 0;
    uint4 g[2][K2_WG];
    float4 gf[2][K2_WG];
    float4 ab[2][K2_WG];
    // no entries without K2'' (a zero-length array): the records then take exactly 24 KiB, so
    // that with the 16 KiB table four workgroups share a CU's 160 KiB
    float2 ew[2][NE * K2_WG];
    __device__ __forceinline__ float2& e(int buf, int z) { return ew[buf][z]; }
};

// raw LDS large enough for either record layout (the batched kernel runs both)
template <int FL>
constexpr size_t ctile_smem_bytes() {
    return sizeof(CTileSmem64<FL>) > sizeof(CTileSmem32<FL>) ? sizeof(CTileSmem64<FL>)
                                                               : sizeof(CTileSmem32<FL>);
}

// Sums zeros [z0, z1) into acc for the NC candidates of this lane (ACC: Acc<false>, or AccGD
// in the enclosure mode: its group sums are flushed every K2_ENC_GROUP zeros and at tile ends).
// Enters and leaves with sm free (the last action is a barrier).
// tab: the (cos, sin) table in LDS of the matching width.
// NT threads per workgroup (256, or 512: two threads per zero slot share the staging, the first
// half g and a, the second half b and e, and twice the candidates read each staged record).
template <int NC, int FL, class ACC, int NT = K2_WG>
__device__ __forceinline__ void ctile_f64(const ZeroView& zv, int z0, int z1, const Limbs4 (&v)[NC],
                                          ACC (&acc)[NC], CTileSmem64<FL>& sm,
                                          const Tab64* tab) {
    static_assert(NT == K2_WG || NT == 2 * K2_WG, "256 or 512 threads");
    constexpr bool D1 = (FL & K2_D1) != 0;
    constexpr bool D2 = (FL & K2_D2) != 0;
    const int tid = threadIdx.x;
    const int zl = tid & (K2_WG - 1);
    const bool lo_half = NT == K2_WG || tid < K2_WG;   // stages g, a
    const bool hi_half = NT == K2_WG || tid >= K2_WG;  // stages b, e
    const int ntiles = z1 > z0 ? (z1 - z0 + K2_WG - 1) / K2_WG : 0;
    // the next tile's record of this lane, held in registers while the current tile is summed.
    // Returned by value: assigning captured locals by reference under these conditions made
    // the compiler keep them in a private (scratch) array, stored and reloaded every tile.
    struct Rec { uint4 g; double2 a, b, e; };
    auto fetch = [&](int t) -> Rec {
        const int j = z0 + t * K2_WG + zl;
        Rec r;
        r.g = make_uint4(0, 0, 0, 0);
        r.a = r.b = r.e = make_double2(0.0, 0.0);
        if (j < z1) {
            if (lo_half) { r.g = zv.g[j]; r.a = zv.a[j]; }
            if (hi_half) {
                if (D1) r.b = zv.b[j];
                if (D2) r.e = zv.e[j];
            }
        }
        return r;
    };
    auto stage = [&](int buf, const Rec& r) {
        if (lo_half) { sm.gz(buf, zl) = r.g; sm.a(buf, zl) = r.a; }
        if (hi_half) {
            if (D1) sm.b(buf, zl) = r.b;
            if (D2) sm.e(buf, zl) = r.e;
        }
    };
    Rec pf;
    if (ntiles > 0) { pf = fetch(0); stage(0, pf); }
    __syncthreads();
    const SinCosCoef kc = sincos_coef();
    for (int t = 0; t < ntiles; ++t) {
        const int buf = t & 1;
        if (t + 1 < ntiles) pf = fetch(t + 1);
        const int zn = min(K2_WG, z1 - z0 - t * K2_WG);   // zeros of this tile (short splits)
#pragma unroll(NT == K2_WG ? 2 : 1)
        for (int z = 0; z < zn; ++z) {
            const uint4 gg = sm.gz(buf, z);
            const double2 aa = sm.a(buf, z);
            double2 bb = make_double2(0.0, 0.0), ee = bb;
            if (D1) bb = sm.b(buf, z);
            if (D2) ee = sm.e(buf, z);
            const Limbs4 g{gg.x, gg.y, gg.z, gg.w};
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                double S, C;
                sincos_tab64_at(phase_rem(g, v[i]), tab, S, C, kc);
                accumulate<FL>(acc[i], S, C, aa, bb, ee);
            }
            if constexpr (AccGroup<ACC>::value > 0) {
                if (z % AccGroup<ACC>::value == AccGroup<ACC>::value - 1) {   // uniform: z is
#pragma unroll
                    for (int i = 0; i < NC; ++i) acc[i].template flush<FL>();
                }
            }
        }
        if constexpr (AccGroup<ACC>::value > 0) {   // a group never spans two tiles
#pragma unroll
            for (int i = 0; i < NC; ++i) acc[i].template flush<FL>();
        }
        if (t + 1 < ntiles) stage(buf ^ 1, pf);
        __syncthreads();
    }
}

// ---- C tile of the one-launch FP64 kernel ---------------------------------------------------
// As ctile_f64, with four changes that take instructions and LDS traffic out of the loop:
//  * the zero records come straight from global memory into scalar registers: the record index
//    is the same in every lane, so the loads are scalar (s_load, through the scalar cache) and
//    the VALU instructions take the record as their one scalar operand -- no LDS staging, no
//    barriers, no per-zero address arithmetic, and LDS holds only the table;
//  * the 2,048-point split sin/cos table (sincos_tabw_at: one FMA fewer per evaluation);
//  * column 5 of the phase as one packed dot product: each zero's (g2, g3) low halves are packed
//    on the scalar unit (s_pack_ll_b32_b16), the candidate's once per call (phase_columns_dp);
//  * one carry-case test for the lane's NC candidates: the largest of their masked bits
//    [98,111) equals the mask exactly when one of them is all ones (NC - 1 full-rate maxima and
//    one compare instead of NC compares), and only then does any lane take the fix-up branch
//    (phase_carry_case_w: the words drop W bit 0, so bit 110 joins the test and halves the rate).

// p[j] read through the constant address space: with j the same in every lane the load is a
// scalar one (s_load) whatever else the kernel writes (a kernel that stores or uses atomics
// through other pointers would otherwise read the records with vector loads into VGPRs).  The
// data must not change while the kernel runs.
__device__ __forceinline__ uint4 ld_const(const uint4* p, int j) {
    typedef __attribute__((address_space(4))) const uint32_t CU;
    const CU* q = (const CU*)(p + j);   // a C cast: reinterpret_cast cannot change the space
    return make_uint4(q[0], q[1], q[2], q[3]);
}

__device__ __forceinline__ double2 ld_const(const double2* p, int j) {
    typedef __attribute__((address_space(4))) const double CD;
    const CD* q = (const CD*)(p + j);
    return make_double2(q[0], q[1]);
}

__device__ __forceinline__ double ld_const(const double* p, int j) {
    typedef __attribute__((address_space(4))) const double CD;
    return ((const CD*)p)[j];
}

// sin/cos of one zero's phase for the lane's NC candidates.  SG: the zero's limbs and gp sit in
// scalar registers.
template <int NC, bool SG = false>
__device__ __forceinline__ void ctw_sincos(const uint4 gg, uint32_t gp, const Limbs4 (&v)[NC],
                                           const uint32_t (&vp)[NC], const double* tab,
                                           const SinCosCoefW& kc, double (&S)[NC],
                                           double (&C)[NC]) {
    const Limbs4 g{gg.x, gg.y, gg.z, gg.w};
    uint32_t c3[NC], c4[NC], c5[NC], ct[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i)
        phase_columns_dp<SG>(g, gp, v[i], vp[i], c3[i], c4[i], c5[i], ct[i]);
    uint32_t m = c3[0] & 0x7FFCu;
#pragma unroll
    for (int i = 1; i < NC; ++i) m = max(m, c3[i] & 0x7FFCu);
    PhaseRem r[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) r[i] = phase_rem_w_cols(c3[i], c4[i], c5[i]);
    if (__builtin_expect(m == 0x7FFCu, 0)) {   // some candidate's carry case
#pragma unroll
        for (int i = 0; i < NC; ++i)
            if (phase_carry_case_w(c3[i]))
                r[i] = phase_rem_w_of(phase_word_fixed(g, v[i], c3[i], c4[i], c5[i], ct[i]));
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) sincos_tabw_at(r[i], tab, S[i], C[i], kc);
}

// One zero's terms for the lane's NC candidates.  SG: the record sits in scalar registers.
template <int NC, int FL, bool SG = false, class ACC = Acc<false>>
__device__ __forceinline__ void ctw_zero(const uint4 gg, uint32_t gp, const double2 aa,
                                         const double2 bb, const double2 ee,
                                         const Limbs4 (&v)[NC], const uint32_t (&vp)[NC],
                                         ACC (&acc)[NC], const double* tab,
                                         const SinCosCoefW& kc) {
    double S[NC], C[NC];
    ctw_sincos<NC, SG>(gg, gp, v, vp, tab, kc, S, C);
#pragma unroll
    for (int i = 0; i < NC; ++i) accumulate<FL>(acc[i], S[i], C[i], aa, bb, ee);
}

// Sums zeros [z0, z1) (uniform over the workgroup) into acc for the NC candidates of this lane
// (ACC: Acc<false>, or AccGD in the enclosure mode: its group sums are flushed every
// K2_ENC_GROUP zeros from z0 and at the end); tab: the split table in LDS, staged and made
// visible by the caller.  The next zero's record is loaded while the current one is summed.
template <int NC, int FL, class ACC = Acc<false>>
__device__ __forceinline__ void ctile_f64s(const uint4* __restrict__ zg,
                                           const double2* __restrict__ za,
                                           const double2* __restrict__ zb,
                                           const double2* __restrict__ ze, int z0, int z1,
                                           const Limbs4 (&v)[NC], ACC (&acc)[NC],
                                           const double* tab) {
    constexpr bool D1 = (FL & K2_D1) != 0;
    constexpr bool D2 = (FL & K2_D2) != 0;
    if (z1 <= z0) return;
    uint32_t vp[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) vp[i] = pack_lo16(v[i].x3, v[i].x2);
    const SinCosCoefW kc = sincos_coef_w();
    struct Rec { uint4 g; double2 a, b, e; };
    auto load = [&](int j) -> Rec {
        Rec r;
        r.g = ld_const(zg, j);
        r.a = ld_const(za, j);
        r.b = r.e = make_double2(0.0, 0.0);
        if (D1) r.b = ld_const(zb, j);
        if (D2) r.e = ld_const(ze, j);
        return r;
    };
    constexpr int GR = AccGroup<ACC>::value;
    Rec cur = load(z0);
    for (int j = z0; j < z1; ++j) {
        const Rec nxt = load(min(j + 1, z1 - 1));   // in flight while zero j is summed
        ctw_zero<NC, FL, true>(cur.g, pack_lo16_s(cur.g.z, cur.g.w), cur.a, cur.b, cur.e, v, vp,
                               acc, tab, kc);
        if constexpr (GR > 0) {
            if ((j - z0) % GR == GR - 1) {   // uniform
#pragma unroll
                for (int i = 0; i < NC; ++i) acc[i].template flush<FL>();
            }
        }
        cur = nxt;
    }
    if constexpr (GR > 0) {
#pragma unroll
        for (int i = 0; i < NC; ++i) acc[i].template flush<FL>();
    }
}

#define K2_F32_GROUP 16   // zeros per compensated add in the FP32 C tile (divides K2_WG)

// FP32 form of ctile_f64: the phase window, a 2^TB-point FP32 table (tab), FP32 weights and
// compensated FP32 sums
template <int NC, int FL, int NT = K2_WG, int TB = K2_TAB32_BITS>
__device__ __forceinline__ void ctile_f32(const ZeroView& zv, int z0, int z1, const Limbs4 (&v)[NC],
                                          Acc<true> (&acc)[NC], CTileSmem32<FL>& sm,
                                          const Tab32* tab) {
    static_assert(NT == K2_WG || NT == 2 * K2_WG, "256 or 512 threads");
    constexpr bool D1 = (FL & K2_D1) != 0;
    constexpr bool D2 = (FL & K2_D2) != 0;
    const int tid = threadIdx.x;
    const int zl = tid & (K2_WG - 1);
    const bool lo_half = NT == K2_WG || tid < K2_WG;   // stages g, gf
    const bool hi_half = NT == K2_WG || tid >= K2_WG;  // stages a|b, e
    const int ntiles = z1 > z0 ? (z1 - z0 + K2_WG - 1) / K2_WG : 0;
    float4 vf[NC];
    uint32_t vp[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        vf[i] = window_vf(v[i]);
        vp[i] = pack_lo16(v[i].x3, v[i].x2);
    }
    // prefetch registers: r0, r1 = g, gf for the first half of the lanes and a|b, e for the
    // second (with 512 lanes each half stages one kind, so the two share registers)
    uint4 r0 = make_uint4(0, 0, 0, 0), r1 = r0, r2 = r0, r3 = r0;
    auto fetch = [&](int t) {
        const int j = z0 + t * K2_WG + zl;
        uint4 x0 = make_uint4(0, 0, 0, 0), x1 = x0, x2 = x0, x3 = x0;
        if (j < z1) {
            if (lo_half) {
                x0 = zv.g[j];
                x0.x = pack_lo16(x0.z, x0.w);   // column 5's packed limbs in place of g0
                const float4 f = zv.gf[j];
                x1 = make_uint4(__float_as_uint(f.x), __float_as_uint(f.y), __float_as_uint(f.z),
                                __float_as_uint(f.w));
            }
            if (hi_half) {
                const float2 a = zv.af[j];
                float2 b = make_float2(0.f, 0.f), e = b;
                if (D1) b = zv.bf[j];
                if (D2) e = zv.ef[j];
                x2 = make_uint4(__float_as_uint(a.x), __float_as_uint(a.y), __float_as_uint(b.x),
                                __float_as_uint(b.y));
                x3 = make_uint4(__float_as_uint(e.x), __float_as_uint(e.y), 0u, 0u);
            }
        }
        if constexpr (NT == 2 * K2_WG) {
            r0 = lo_half ? x0 : x2;
            r1 = lo_half ? x1 : x3;
        } else {
            r0 = x0; r1 = x1; r2 = x2; r3 = x3;
        }
    };
    auto stage = [&](int buf) {
        const uint4 g = r0, f = r1;
        const uint4 ab = NT == 2 * K2_WG ? r0 : r2, e = NT == 2 * K2_WG ? r1 : r3;
        if (lo_half) {
            sm.g[buf][zl] = g;
            sm.gf[buf][zl] = make_float4(__uint_as_float(f.x), __uint_as_float(f.y),
                                         __uint_as_float(f.z), __uint_as_float(f.w));
        }
        if (hi_half) {
            sm.ab[buf][zl] = make_float4(__uint_as_float(ab.x), __uint_as_float(ab.y),
                                         __uint_as_float(ab.z), __uint_as_float(ab.w));
            if (D2) sm.e(buf, zl) = make_float2(__uint_as_float(e.x), __uint_as_float(e.y));
        }
    };
    if (ntiles > 0) { fetch(0); stage(0); }
    __syncthreads();
    for (int t = 0; t < ntiles; ++t) {
        const int buf = t & 1;
        if (t + 1 < ntiles) fetch(t + 1);
        // zeros of this tile rounded up to whole groups (the staged records past z1 are zeros)
        const int zn = min(K2_WG, (z1 - z0 - t * K2_WG + K2_F32_GROUP - 1) / K2_F32_GROUP *
                                      K2_F32_GROUP);
        // K2_F32_GROUP zeros per step: their terms are accumulated in FP32 by FMAs, then the
        // group sum enters the compensated sum (blocked compensated summation: about two
        // roundings per term inside a group, against four adds per term to compensate every
        // term)
        for (int z = 0; z < zn; z += K2_F32_GROUP) {
            float3 tp[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) tp[i] = make_float3(0.f, 0.f, 0.f);
#pragma unroll
            for (int h = 0; h < K2_F32_GROUP; ++h) {
                const uint4 gg = sm.g[buf][z + h];
                const float4 gf = sm.gf[buf][z + h];
                const float4 ab = sm.ab[buf][z + h];
                float2 ee = make_float2(0.f, 0.f);
                if (D2) ee = sm.e(buf, z + h);
                const Limbs4 g{gg.x, gg.y, gg.z, gg.w};
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    uint32_t hi, lo;
                    float t3;
                    phase_window_f<TB>(g, gg.x, gf, v[i], vp[i], vf[i], hi, lo, t3);
                    float S, C;
                    sincos_window_f<TB>(hi, lo, t3, tab, S, C);
                    add_terms<FL>(tp[i], S, C, make_float2(ab.x, ab.y), make_float2(ab.z, ab.w),
                                  ee);
                }
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) add_group<FL>(acc[i], tp[i]);
        }
        if (t + 1 < ntiles) stage(buf ^ 1);
        __syncthreads();
    }
}

// ---- Z tile ---------------------------------------------------------------------------------
#define K2_ZG 4   // candidates a Z-tile lane evaluates per pass over its zeros
#define K2_ZF32_GROUP 4   // zeros per lane per compensated add in the FP32 Z tile

// a value the same in every lane, as an SGPR
__device__ __forceinline__ uint32_t uniform(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

__device__ __forceinline__ float uniform_f(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

template <int CZ>
struct ZTileSmem {
    uint4 v[CZ];
    double red[K2_WG / 64][3 * CZ];
};

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Sums zeros [z0, z1) for candidates cand[c_first .. c_first + c_count) (c_count <= CZ); lanes
// stride over the zeros, ZG candidates at a time (the lane's zero records are re-read from
// cache for each group, which bounds the register footprint).  Writes
// out[q * CZ + c] (q = value, K', K'') after a fixed-order reduction: butterfly within each
// wave, then the four wave sums in order.
template <int CZ, int FL, bool F32, int TB = K2_TAB32_BITS>
__device__ __forceinline__ void ztile(const ZeroView& zv, int z0, int z1, const uint4* cand,
                                      int c_first, int c_count, double* out, ZTileSmem<CZ>& sm,
                                      const Tab64* tab64, const Tab32* tab32) {
    constexpr bool D1 = (FL & K2_D1) != 0;
    constexpr bool D2 = (FL & K2_D2) != 0;
    constexpr int NQ = D2 ? 3 : (D1 ? 2 : 1);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    // candidates per pass: the FP32 path keeps 8 in registers (each zero record loaded once
    // per 8 candidates), the FP64 path 4
    constexpr int ZG = (F32 && CZ % 8 == 0) ? 8 : K2_ZG;
    if (tid < CZ) sm.v[tid] = tid < c_count ? cand[c_first + tid] : make_uint4(0, 0, 0, 0);
    __syncthreads();
    // passes with no live candidate are skipped (their outputs are written as zeros below)
    const int cp_end = min(CZ, (c_count + ZG - 1) / ZG * ZG);
    const SinCosCoef kc = sincos_coef();
    for (int cp = 0; cp < cp_end; cp += ZG) {
        Acc<F32> acc[ZG];
        Limbs4 v[ZG];
        uint32_t vp[F32 ? ZG : 1];
        float4 vf[F32 ? ZG : 1];
#pragma unroll
        for (int c = 0; c < ZG; ++c) {
            acc[c].zero();
            // the same candidates in every lane: their limbs go to SGPRs
            const uint4 cv = sm.v[cp + c];
            v[c] = Limbs4{uniform(cv.x), uniform(cv.y), uniform(cv.z), uniform(cv.w)};
            if constexpr (F32) {
                vp[c] = uniform(pack_lo16(cv.w, cv.z));
                const float4 f = window_vf(v[c]);
                vf[c] = make_float4(uniform_f(f.x), uniform_f(f.y), uniform_f(f.z), uniform_f(f.w));
            }
        }
        if constexpr (F32) {
            // K2_ZF32_GROUP zeros per lane per compensated add, as in ctile_f32
            for (int jb = z0 + tid; jb < z1; jb += K2_WG * K2_ZF32_GROUP) {
                float3 tp[ZG];
#pragma unroll
                for (int c = 0; c < ZG; ++c) tp[c] = make_float3(0.f, 0.f, 0.f);
#pragma unroll
                for (int h = 0; h < K2_ZF32_GROUP; ++h) {
                    const int j = jb + h * K2_WG;
                    if (j < z1) {
                        const uint4 gg = zv.g[j];
                        const Limbs4 g{gg.x, gg.y, gg.z, gg.w};
                        const uint32_t gp = pack_lo16(gg.z, gg.w);
                        const float4 gf = zv.gf[j];
                        const float2 a = zv.af[j];
                        const float2 b = D1 ? zv.bf[j] : make_float2(0.f, 0.f);
                        const float2 e = D2 ? zv.ef[j] : make_float2(0.f, 0.f);
#pragma unroll
                        for (int c = 0; c < ZG; ++c) {
                            uint32_t hi, lo;
                            float t3;
                            phase_window_f<TB, true>(g, gp, gf, v[c], vp[c], vf[c], hi, lo,
                                                      t3);
                            float S, C;
                            sincos_window_f<TB>(hi, lo, t3, tab32, S, C);
                            add_terms<FL>(tp[c], S, C, a, b, e);
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < ZG; ++c) add_group<FL>(acc[c], tp[c]);
            }
        } else {
            for (int j = z0 + tid; j < z1; j += K2_WG) {
                const uint4 gg = zv.g[j];
                const Limbs4 g{gg.x, gg.y, gg.z, gg.w};
                const double2 a = zv.a[j];
                const double2 b = D1 ? zv.b[j] : make_double2(0.0, 0.0);
                const double2 e = D2 ? zv.e[j] : make_double2(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < ZG; ++c) {
                    double S, C;
                    sincos_tab64_at(phase_rem<true>(g, v[c]), tab64, S, C, kc);
                    accumulate<FL>(acc[c], S, C, a, b, e);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < ZG; ++c) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const double x = wave_sum(acc[c].get(q));
                if (lane == 0) sm.red[wid][q * CZ + cp + c] = x;
            }
        }
    }
    __syncthreads();
    if (tid < 3 * CZ) {
        double s = 0.0;
        if (tid < NQ * CZ && tid % CZ < cp_end) {
#pragma unroll
            for (int w = 0; w < K2_WG / 64; ++w) s += sm.red[w][tid];
        }
        out[tid] = s;
    }
    __syncthreads();
}

}  // namespace k2dev
