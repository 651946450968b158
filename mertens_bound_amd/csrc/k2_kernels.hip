// HIP kernels of the K2 evaluator for gfx950 (MI355X, CDNA4, wave64).
//
// K2(v) = 2 sum_j e^{-k g_j^2} Re(w_j e^{i g_j v}) and its first two v-derivatives, for a list of
// candidates v against a zero table.  This is a streamed trig sum (no contraction), so it runs on
// the VALU: exact integer phase (k2_device.h), FP64 (or FP32) sin/cos, FP64 accumulation.
//
// Work decomposition of the evaluation kernels:
//   * a workgroup (256 lanes = 4 waves) owns a tile of 256*NC candidates, NC per lane in VGPRs;
//   * it walks one contiguous zero range ("split") in tiles of 256 zeros staged in LDS
//     (double-buffered, one barrier per tile); every lane reads the same zero record, an LDS
//     broadcast, so one staged zero feeds 256*NC candidate evaluations;
//   * each workgroup writes its per-candidate partial sums to partial[split][3][n_cand_pad];
//     k2_reduce_kernel adds the splits in a fixed order, so results are deterministic;
//   * blockIdx -> (split, candidate tile) is XCD-aware: workgroups are dealt round-robin to the 8
//     XCDs, and consecutive workgroups of one XCD get the same split, so a zero range is pulled
//     into one XCD's L2 once and re-read from there.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "k2_device.h"
#include "k2_internal.h"
#include "k2_tiles.h"

using namespace k2dev;

#define K2_ZT 256   // zeros per LDS tile = threads per workgroup

// ---- per-k weighted zero arrays ---------------------------------------------------------------
// a = 2 e^{-k g^2} w, b = g a, e = g^2 a (FP64) and the same in FP32; padded entries are zero.
__global__ __launch_bounds__(256) void k2_weights_kernel(
    const double* __restrict__ gamma, const double2* __restrict__ w, int n, int n_pad, double k,
    double2* __restrict__ a, double2* __restrict__ b, double2* __restrict__ e,
    float2* __restrict__ af, float2* __restrict__ bf, float2* __restrict__ ef) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_pad) return;
    double2 va = make_double2(0.0, 0.0), vb = va, ve = va;
    if (j < n) {
        const double g = gamma[j];
        const double f = 2.0 * exp(-k * g * g);
        const double2 ww = w[j];
        va = make_double2(f * ww.x, f * ww.y);
        vb = make_double2(g * va.x, g * va.y);
        ve = make_double2(g * vb.x, g * vb.y);
    }
    a[j] = va; b[j] = vb; e[j] = ve;
    af[j] = make_float2((float)va.x, (float)va.y);
    bf[j] = make_float2((float)vb.x, (float)vb.y);
    ef[j] = make_float2((float)ve.x, (float)ve.y);
}

// sum |a_j| per K2_ABS_BLOCK zeros (fixed order): the magnitudes that size the mixed form's head
__global__ __launch_bounds__(256) void k2_abs_blocks_kernel(const double2* __restrict__ a,
                                                            double* __restrict__ blk) {
    __shared__ double red[4];
    const size_t base = (size_t)blockIdx.x * K2_ABS_BLOCK;
    double s = 0.0;
    for (int i = threadIdx.x; i < K2_ABS_BLOCK; i += 256) {
        const double2 v = a[base + i];
        s += sqrt(v.x * v.x + v.y * v.y);
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// ---- one-launch evaluation: C tiles over (split, candidate tile) workgroups -----------------------
// Splits s < n_long cover zeros [z_begin + s*zps_long, ...) (zps_long each), the later ones
// zps each after those, up to z_end; split s writes partial row row_base + s.
// F32 selects the precision form (phase window, FP32 sin/cos and weights, compensated FP32
// sums); the mixed form launches an FP64 head and an FP32 tail side by side.
// FP64: the zero records come into scalar registers (ctile_f64s) and LDS holds only the
// 32 KiB table, so VGPRs set the occupancy: about 100 at NC = 4 (four waves per SIMD).
// FP32: NC <= 2 keeps to 128 VGPRs so that four workgroups share a CU; NT = 512 puts twice the
// waves on one staged tile and one sin/cos table (LDS is what limits the workgroups per CU),
// walks the zeros without unrolling and stays within 80 VGPRs (six waves per SIMD; 96 and five
// waves with K2'', whose third sum would spill at 80)
template <int NC, int FL, bool F32, int NT>
__global__ __launch_bounds__(NT, (NT == 512 ? ((FL & K2_D2) ? 5 : (F32 && NC == 2 ? 8 : 6))
                                            : (NC <= 2 ? 4 : 2))) void k2_eval_kernel(
    ZeroView zv, int z_begin, int z_end, int zps, int n_long, int zps_long,
    const uint4* __restrict__ cand, int n_cand, int n_ctiles, double* __restrict__ partial,
    int row_base, int n_cand_pad, const Tab64* __restrict__ tab64,
    const Tab32* __restrict__ tab32) {
    // FP32: the C tile's records and the 2,048-point FP32 table; FP64: only the split
    // 2,048-point table (after the K2_TAB interleaved entries of the FP64 table buffer) -- the
    // FP64 C tile reads its zero records into scalar registers
    __shared__ typename std::conditional<F32, CTileSmem32<FL>, char>::type sm;
    constexpr int NTAB = F32 ? K2_TAB32 : K2_TABW_DOUBLES;
    __shared__ typename std::conditional<F32, Tab32, double>::type s_tab[NTAB];
    const BlockSplit bs = block_split(n_ctiles, z_begin, z_end, zps, n_long, zps_long);
    if (bs.z0 >= z_end) return;   // a padding split: no row
    const int split = bs.split, ct = bs.ct, z0 = bs.z0, z1 = bs.z1;
    const int tid = threadIdx.x;
    for (int i = tid; i < NTAB; i += NT) {
        if constexpr (F32) s_tab[i] = tab32[i];
        else s_tab[i] = reinterpret_cast<const double*>(tab64 + K2_TAB)[i];
    }
    Limbs4 v[NC];
    Acc<F32> acc[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = ct * (NT * NC) + i * NT + tid;
        uint4 cv = make_uint4(0, 0, 0, 0);
        if (c < n_cand) cv = cand[c];
        v[i] = Limbs4{cv.x, cv.y, cv.z, cv.w};
        acc[i].zero();
    }
    if constexpr (F32) ctile_f32<NC, FL, NT>(zv, z0, z1, v, acc, sm, s_tab);
    else {
        (void)sm;
        __syncthreads();   // the table
        ctile_f64s<NC, FL>(zv.g, zv.a, zv.b, zv.e, z0, z1, v, acc, s_tab);
    }
    double* prow = partial + (size_t)(row_base + split) * 3 * n_cand_pad;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = ct * (NT * NC) + i * NT + tid;
        prow[c] = acc[i].get(0);
        if (FL & K2_D1) prow[n_cand_pad + c] = acc[i].get(1);
        if (FL & K2_D2) prow[2 * n_cand_pad + c] = acc[i].get(2);
    }
}

// ---- mixed form in one launch (small problems, where a second stream costs more than it saves):
// splits [0, sp_head) cover the FP64 head [0, head_end), the rest the FP32 tail [head_end, z_end)
template <int NC, int FL>
__global__ __launch_bounds__(256, 4) void k2_eval_mixed_kernel(
    ZeroView zv, int head_end, int zps_head, int sp_head, int z_end, int zps_tail,
    const uint4* __restrict__ cand, int n_cand, int n_ctiles, double* __restrict__ partial,
    int n_cand_pad, const Tab64* __restrict__ tab64, const Tab32* __restrict__ tab32) {
    constexpr size_t TILE = ctile_smem_bytes<FL>();
    __shared__ __align__(16) unsigned char sm_raw[TILE];
    __shared__ Tab64 s_tab64[K2_TAB];   // a tail block's FP32 table takes the same 16 KiB
    const int L = xcd_linear_block();
    const int split = L / n_ctiles, ct = L - split * n_ctiles;
    const int tid = threadIdx.x;
    const bool head = split < sp_head;
    static_assert(sizeof(Tab32) * K2_TAB32 <= sizeof(Tab64) * K2_TAB, "FP32 table fits");
    Tab32* s_tab32 = reinterpret_cast<Tab32*>(s_tab64);
    if (head) {
        for (int i = tid; i < K2_TAB; i += K2_ZT) s_tab64[i] = tab64[i];
    } else {
        for (int i = tid; i < K2_TAB32; i += K2_ZT) s_tab32[i] = tab32[i];
    }
    Limbs4 v[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = ct * (K2_ZT * NC) + i * K2_ZT + tid;
        uint4 cv = make_uint4(0, 0, 0, 0);
        if (c < n_cand) cv = cand[c];
        v[i] = Limbs4{cv.x, cv.y, cv.z, cv.w};
    }
    double* prow = partial + (size_t)split * 3 * n_cand_pad;
    if (head) {
        const int z0 = split * zps_head, z1 = min(head_end, z0 + zps_head);
        Acc<false> acc[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) acc[i].zero();
        ctile_f64<NC, FL>(zv, z0, z1, v, acc, *reinterpret_cast<CTileSmem64<FL>*>(sm_raw),
                          s_tab64);
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const int c = ct * (K2_ZT * NC) + i * K2_ZT + tid;
            prow[c] = acc[i].get(0);
            if (FL & K2_D1) prow[n_cand_pad + c] = acc[i].get(1);
            if (FL & K2_D2) prow[2 * n_cand_pad + c] = acc[i].get(2);
        }
    } else {
        const int z0 = head_end + (split - sp_head) * zps_tail, z1 = min(z_end, z0 + zps_tail);
        Acc<true> acc[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) acc[i].zero();
        ctile_f32<NC, FL>(zv, z0, z1, v, acc, *reinterpret_cast<CTileSmem32<FL>*>(sm_raw),
                          s_tab32);
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const int c = ct * (K2_ZT * NC) + i * K2_ZT + tid;
            prow[c] = acc[i].get(0);
            if (FL & K2_D1) prow[n_cand_pad + c] = acc[i].get(1);
            if (FL & K2_D2) prow[2 * n_cand_pad + c] = acc[i].get(2);
        }
    }
}

// ---- one-launch evaluation for few candidates: Z tiles over (chunk, candidate group) ------------
template <int FL, bool F32>
__global__ __launch_bounds__(256) void k2_eval_z_kernel(
    ZeroView zv, int z_begin, int z_end, int zpc, const uint4* __restrict__ cand, int n_cand,
    int n_groups, double* __restrict__ partial, int row_base, const Tab64* __restrict__ tab64,
    const Tab32* __restrict__ tab32) {
    __shared__ ZTileSmem<K2_Z_CZ> sm;
    __shared__ Tab64 s64[F32 ? 1 : K2_TAB];
    __shared__ Tab32 s32[F32 ? K2_TAB32 : 1];
    for (int i = threadIdx.x; i < (F32 ? K2_TAB32 : K2_TAB); i += K2_ZT) {
        if constexpr (F32) s32[i] = tab32[i];
        else s64[i] = tab64[i];
    }
    // ztile's first barrier orders the table stores before any lookup
    const int L = xcd_linear_block();
    const int chunk = L / n_groups, g = L - chunk * n_groups;
    const int z0 = z_begin + chunk * zpc;
    const int z1 = min(z_end, z0 + zpc);
    const int c_first = g * K2_Z_CZ;
    double* out = partial + ((size_t)(row_base + chunk) * n_groups + g) * 3 * K2_Z_CZ;
    ztile<K2_Z_CZ, FL, F32>(zv, z0, z1, cand, c_first, min(K2_Z_CZ, n_cand - c_first), out, sm,
                            s64, s32);
}

// one wavefront per candidate: lanes stride over the rows, then a fixed butterfly
__global__ __launch_bounds__(256) void k2_reduce_z_kernel(
    const double* __restrict__ partial, int n_rows, int n_groups, int n_cand, unsigned flags,
    double* __restrict__ K, double* __restrict__ dK, double* __restrict__ d2K) {
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= n_cand) return;   // whole wavefronts leave together
    const int g = c / K2_Z_CZ, pos = c - g * K2_Z_CZ;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int r = lane; r < n_rows; r += 64) {
        const double* row = partial + ((size_t)r * n_groups + g) * 3 * K2_Z_CZ;
        s0 += row[pos];
        s1 += row[K2_Z_CZ + pos];
        s2 += row[2 * K2_Z_CZ + pos];
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        if (K) K[c] = s0;
        if (dK) dK[c] = (flags & K2_D1) ? s1 : 0.0;
        if (d2K) d2K[c] = (flags & K2_D2) ? s2 : 0.0;
    }
}

// ---- deterministic split reduction ---------------------------------------------------------------
// A workgroup covers 256/SG candidates; its SG thread rows sum the splits s = row (mod SG) in order
// (consecutive candidates read consecutive doubles of a partial row), then row sums are added in
// row order.  SG grows with the split count so that few candidates over many splits still spread.
template <int SG>
__global__ __launch_bounds__(256) void k2_reduce_kernel(
    const double* __restrict__ partial, int n_splits, int n_cand, int n_cand_pad, unsigned flags,
    double* __restrict__ K, double* __restrict__ dK, double* __restrict__ d2K) {
    constexpr int CB = 256 / SG;
    __shared__ double red[3][SG][CB];
    const int cl = threadIdx.x % CB, sg = threadIdx.x / CB;
    const int c = blockIdx.x * CB + cl;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (c < n_cand) {
        // four rows' loads in flight per step, added in row order (the same sums as one row at
        // a time): the loop is bound by load latency, not bandwidth
        const bool d1 = (flags & K2_D1) != 0, d2 = (flags & K2_D2) != 0;
        const size_t rs = (size_t)3 * n_cand_pad;
        const double* p0 = partial + c;
        int s = sg;
        for (; s + 3 * SG < n_splits; s += 4 * SG) {
            double a[4], b[4], e[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double* row = p0 + (size_t)(s + u * SG) * rs;
                a[u] = row[0];
                b[u] = d1 ? row[n_cand_pad] : 0.0;
                e[u] = d2 ? row[2 * n_cand_pad] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                s0 += a[u];
                s1 += b[u];
                s2 += e[u];
            }
        }
        for (; s < n_splits; s += SG) {
            const double* row = p0 + (size_t)s * rs;
            s0 += row[0];
            if (d1) s1 += row[n_cand_pad];
            if (d2) s2 += row[2 * n_cand_pad];
        }
    }
    if constexpr (SG > 1) {
        red[0][sg][cl] = s0; red[1][sg][cl] = s1; red[2][sg][cl] = s2;
        __syncthreads();
        if (sg != 0) return;
        for (int g = 1; g < SG; ++g) {
            s0 += red[0][g][cl]; s1 += red[1][g][cl]; s2 += red[2][g][cl];
        }
    }
    if (c >= n_cand) return;
    if (K) K[c] = s0;
    if (dK) dK[c] = (flags & K2_D1) ? s1 : 0.0;
    if (d2K) d2K[c] = (flags & K2_D2) ? s2 : 0.0;
}

// ---- debug: phase words --------------------------------------------------------------------------
// mode 0: fast W (skip + fix-up), 1: schoolbook W, 2: the exact-column phase window, bits
// [128,174) of the partial product (the phase floored to 2^-46 turn, low by at most 4 units),
// 3: the phase the FP32 kernels use, in the same units: the window of phase_window_f without
// its rounding offset, plus its FP32 column 3 rounded to a unit
__global__ __launch_bounds__(256) void k2_phase_words_kernel(
    const uint4* __restrict__ zg, const float4* __restrict__ zgf, int z0, int nz,
    const uint4* __restrict__ cand, int n_cand, int mode, uint64_t* __restrict__ out) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n_cand * nz) return;
    const int c = (int)(idx / nz), j = (int)(idx - (long long)c * nz);
    const uint4 gg = zg[z0 + j], cv = cand[c];
    const Limbs4 g{gg.x, gg.y, gg.z, gg.w}, v{cv.x, cv.y, cv.z, cv.w};
    uint64_t W;
    if (mode == 0) W = phase_word(g, v);
    else if (mode == 1) W = phase_word_exact(g, v);
    else if (mode == 4) {
        // the one-launch C tile's remainder words (W bits [1,64)) reassembled into W, bit 0 clear
        uint32_t s3lo, s4lo, s5, tlo;
        phase_columns_dp(g, pack_lo16(g.x2, g.x3), v, pack_lo16(v.x3, v.x2), s3lo, s4lo, s5, tlo);
        PhaseRem r = phase_rem_w_cols(s3lo, s4lo, s5);
        if (phase_carry_case_w(s3lo))
            r = phase_rem_w_of(phase_word_fixed(g, v, s3lo, s4lo, s5, tlo));
        W = ((uint64_t)(r.off >> 3) << 53) | ((uint64_t)(r.rhi & 0xFFFFFu) << 33) |
            ((uint64_t)r.rlo << 1);
    }
    else if (mode == 2) {
        uint32_t hi, lo;
        phase_window<K2_TAB32_BITS>(g, v, hi, lo);
        W = ((((uint64_t)hi << 32) | lo) - (1ull << (45 - K2_TAB32_BITS))) & ((1ull << 46) - 1);
    } else {
        uint32_t hi, lo;
        float t3x;
        phase_window_f<K2_TAB32_BITS>(g, pack_lo16(g.x2, g.x3), zgf[z0 + j], v,
                                       pack_lo16(v.x3, v.x2), window_vf(v), hi, lo, t3x);
        // column 3 back in units of 2^128: (t3x + 3/4 step) / (2 pi 2^-46)
        constexpr double unit = 6.283185307179586 / 70368744177664.0;   // 2 pi 2^-46 rad
        const double t3 = ((double)t3x + (double)WindowF<K2_TAB32_BITS>::tq) / unit;
        const uint64_t c3 = (uint64_t)llrint(t3);   // units of 2^128
        W = ((((uint64_t)hi << 32) | lo) - (3ull << (44 - K2_TAB32_BITS)) + c3) &
            ((1ull << 46) - 1);
    }
    out[idx] = W;
}

// ---- debug: the evaluators' sin/cos of phase words ---------------------------------------------
// mode 0: FP64 1,024-point table path (enclosure, batched, Z-tile and weight-vector kernels),
// 1: FP32 table path of the mixed kernel (the window of W: bits [18,64) plus half a table step),
// 2: FP64 2,048-point split table path of the one-launch C tile
__global__ __launch_bounds__(256) void k2_sincos_words_kernel(const uint64_t* __restrict__ W,
                                                              int n, int mode,
                                                              const Tab64* __restrict__ tab64,
                                                              const Tab32* __restrict__ tab32,
                                                              double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mode == 0) {
        double s, c;
        sincos_tab64(W[i], tab64, s, c);
        out[2 * i] = s; out[2 * i + 1] = c;
    } else if (mode == 2) {
        double s, c;
        sincos_tabw(W[i], reinterpret_cast<const double*>(tab64 + K2_TAB), s, c);
        out[2 * i] = s; out[2 * i + 1] = c;
    } else {
        float s, c;
        const uint64_t win = (W[i] >> 18) + (1ull << (45 - K2_TAB32_BITS));
        sincos_window<K2_TAB32_BITS>((uint32_t)(win >> 32), (uint32_t)win, tab32, s, c);
        out[2 * i] = s; out[2 * i + 1] = c;
    }
}

// ---- Newton update on 64.64 points (polish) ---------------------------------------------------
// Rule of search/path1_eval_fast.cpp:120-128: require |K''| > 1e-12, step = K'/K'' clamped to
// +-clamp, v -= step, stop once |step| < tol (the reference's 1e-13; the precision form stops at
// its own noise floor, see k2_polish_device).  A candidate whose flag is 0 has stopped.
__global__ __launch_bounds__(256) void k2_newton_kernel(
    uint4* __restrict__ cand, const double* __restrict__ dK, const double* __restrict__ d2K,
    int n, double clamp, double tol, int* __restrict__ active, int* __restrict__ steps) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || !active[c]) return;
    const double fpp = d2K[c];
    if (!(fabs(fpp) > 1e-12)) { active[c] = 0; return; }
    double step = dK[c] / fpp;
    step = fmin(fmax(step, -clamp), clamp);
    const long long d = llrint(ldexp(step, 64));          // |step| <= clamp < 2^-1
    uint4 v = cand[c];
    unsigned __int128 x = ((unsigned __int128)(((uint64_t)v.w << 32) | v.z) << 64) |
                          (((uint64_t)v.y << 32) | v.x);
    x -= (unsigned __int128)(__int128)d;
    const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    cand[c] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
    steps[c] += 1;
    if (fabs(step) < tol) active[c] = 0;
}

// ---- polish compaction: the candidates still moving, gathered and their sums scattered back -----
__global__ __launch_bounds__(256) void k2_gather_kernel(const uint4* __restrict__ V,
                                                        const int* __restrict__ idx, int n,
                                                        uint4* __restrict__ Vc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Vc[i] = V[idx[i]];
}

__global__ __launch_bounds__(256) void k2_scatter3_kernel(
    const int* __restrict__ idx, int n, const double* __restrict__ Kc,
    const double* __restrict__ dKc, const double* __restrict__ d2Kc, double* __restrict__ K,
    double* __restrict__ dK, double* __restrict__ d2K) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = idx[i];
    K[c] = Kc[i];
    dK[c] = dKc[i];
    d2K[c] = d2Kc[i];
}

// ---- reference-format compute_peak_AB (search/path1_eval_fast.cpp:15-53) ---------------------
// one lane per (peak, zero); ph = xh*ghi + xl*glo - psi mod 2^128; th = top64 * 2pi/2^64 + g dy0
__global__ __launch_bounds__(256) void k2_ref_peak_ab_kernel(
    int n_peaks, const uint64_t* __restrict__ x_lo, const uint64_t* __restrict__ x_hi,
    const double* __restrict__ dy0, int n_zeros, const double* __restrict__ alpha,
    const double* __restrict__ gamma, const uint64_t* __restrict__ ghi_lo,
    const uint64_t* __restrict__ ghi_hi, const uint64_t* __restrict__ glo_lo,
    const uint64_t* __restrict__ glo_hi, const uint64_t* __restrict__ psi_lo,
    const uint64_t* __restrict__ psi_hi, double* __restrict__ A, double* __restrict__ B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (i >= n_zeros || p >= n_peaks) return;
    typedef unsigned __int128 u128;
    const u128 x = ((u128)x_hi[p] << 64) | x_lo[p];
    const uint64_t xh = (uint64_t)(x >> 48), xl = (uint64_t)(x & ((1ull << 48) - 1));
    const u128 gh = ((u128)ghi_hi[i] << 64) | ghi_lo[i];
    const u128 gl = ((u128)glo_hi[i] << 64) | glo_lo[i];
    const u128 ps = ((u128)psi_hi[i] << 64) | psi_lo[i];
    const u128 ph = (u128)xh * gh + (u128)xl * gl - ps;
    const double th = (double)(uint64_t)(ph >> 64) * K2_TWO_PI_2M64 + gamma[i] * dy0[p];
    double s, c;
    sincos_general(th, s, c);
    const double ta = 2.0 * alpha[i];
    A[(size_t)p * n_zeros + i] = ta * c;
    B[(size_t)p * n_zeros + i] = ta * s;
}

// ---- launchers -------------------------------------------------------------------------------------
namespace k2impl {

template <int NC, int FL, bool F32>
static hipError_t launch_eval_t(const LaunchArgs& a, hipStream_t st) {
    const ZeroView zv{a.zg, a.za, a.zb, a.ze, a.zaf, a.zbf, a.zef, a.zgf};
    if constexpr (NC <= 2) {
        if (a.nt == 512) {
            hipLaunchKernelGGL((k2_eval_kernel<NC, FL, F32, 512>), dim3(a.n_blocks), dim3(512), 0,
                               st, zv, a.z_begin, a.z_end, a.zps, a.n_long, a.zps_long, a.cand,
                               a.n_cand, a.n_ctiles, a.partial, a.row_base, a.n_cand_pad,
                               a.tab64, a.tab32);
            return hipGetLastError();
        }
    }
    if (a.nt != 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k2_eval_kernel<NC, FL, F32, 256>), dim3(a.n_blocks), dim3(256), 0, st, zv,
                       a.z_begin, a.z_end, a.zps, a.n_long, a.zps_long, a.cand, a.n_cand,
                       a.n_ctiles, a.partial, a.row_base, a.n_cand_pad, a.tab64, a.tab32);
    return hipGetLastError();
}

template <int NC, bool F32>
static hipError_t dispatch_fl(const LaunchArgs& a, hipStream_t st) {
    switch (a.flags & (K2_D1 | K2_D2)) {
        case 0: return launch_eval_t<NC, 0, F32>(a, st);
        case K2_D1: return launch_eval_t<NC, K2_D1, F32>(a, st);
        case K2_D2: return launch_eval_t<NC, K2_D2, F32>(a, st);
        default: return launch_eval_t<NC, K2_D1 | K2_D2, F32>(a, st);
    }
}

template <int NC>
static hipError_t dispatch_prec(const LaunchArgs& a, bool mixed, hipStream_t st) {
    return mixed ? dispatch_fl<NC, true>(a, st) : dispatch_fl<NC, false>(a, st);
}

hipError_t launch_eval(const LaunchArgs& a, bool mixed, hipStream_t st) {
    switch (a.nc) {
        case 1: return dispatch_prec<1>(a, mixed, st);
        case 2: return dispatch_prec<2>(a, mixed, st);
        case 4: return dispatch_prec<4>(a, mixed, st);
        case 8: return dispatch_prec<8>(a, mixed, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_weights(const double* gamma, const double* w, int n, int n_pad, double k,
                          WeightArrays& out, hipStream_t st) {
    const int nb = (n_pad + 255) / 256;
    hipLaunchKernelGGL(k2_weights_kernel, dim3(nb), dim3(256), 0, st, gamma, (const double2*)w, n,
                       n_pad, k, out.a, out.b, out.e, out.af, out.bf, out.ef);
    return hipGetLastError();
}

template <int NC, int FL>
static hipError_t launch_mixed_t(const LaunchArgs& a, int head_end, int zps_head, int sp_head,
                                 int zps_tail, hipStream_t st) {
    const ZeroView zv{a.zg, a.za, a.zb, a.ze, a.zaf, a.zbf, a.zef, a.zgf};
    hipLaunchKernelGGL((k2_eval_mixed_kernel<NC, FL>), dim3(a.n_blocks), dim3(K2_ZT), 0, st, zv,
                       head_end, zps_head, sp_head, a.z_end, zps_tail, a.cand, a.n_cand,
                       a.n_ctiles, a.partial, a.n_cand_pad, a.tab64, a.tab32);
    return hipGetLastError();
}

template <int NC>
static hipError_t launch_mixed_fl(const LaunchArgs& a, int head_end, int zps_head, int sp_head,
                                  int zps_tail, hipStream_t st) {
    switch (a.flags & (K2_D1 | K2_D2)) {
        case 0: return launch_mixed_t<NC, 0>(a, head_end, zps_head, sp_head, zps_tail, st);
        case K2_D1: return launch_mixed_t<NC, K2_D1>(a, head_end, zps_head, sp_head, zps_tail, st);
        case K2_D2: return launch_mixed_t<NC, K2_D2>(a, head_end, zps_head, sp_head, zps_tail, st);
        default:
            return launch_mixed_t<NC, K2_D1 | K2_D2>(a, head_end, zps_head, sp_head, zps_tail, st);
    }
}

hipError_t launch_eval_mixed(const LaunchArgs& a, int head_end, int zps_head, int sp_head,
                             int zps_tail, hipStream_t st) {
    if (a.nt != 256) return hipErrorInvalidValue;
    switch (a.nc) {
        case 1: return launch_mixed_fl<1>(a, head_end, zps_head, sp_head, zps_tail, st);
        case 2: return launch_mixed_fl<2>(a, head_end, zps_head, sp_head, zps_tail, st);
        default: return hipErrorInvalidValue;
    }
}

template <int FL, bool F32>
static hipError_t launch_eval_z_t(const LaunchArgs& a, int n_groups, int n_chunks, int zpc,
                                  hipStream_t st) {
    const ZeroView zv{a.zg, a.za, a.zb, a.ze, a.zaf, a.zbf, a.zef, a.zgf};
    hipLaunchKernelGGL((k2_eval_z_kernel<FL, F32>), dim3(n_groups * n_chunks), dim3(K2_ZT), 0, st,
                       zv, a.z_begin, a.z_end, zpc, a.cand, a.n_cand, n_groups, a.partial,
                       a.row_base, a.tab64, a.tab32);
    return hipGetLastError();
}

template <bool F32>
static hipError_t launch_eval_z_f(const LaunchArgs& a, int n_groups, int n_chunks, int zpc,
                                  hipStream_t st) {
    switch (a.flags & (K2_D1 | K2_D2)) {
        case 0: return launch_eval_z_t<0, F32>(a, n_groups, n_chunks, zpc, st);
        case K2_D1: return launch_eval_z_t<K2_D1, F32>(a, n_groups, n_chunks, zpc, st);
        case K2_D2: return launch_eval_z_t<K2_D2, F32>(a, n_groups, n_chunks, zpc, st);
        default: return launch_eval_z_t<K2_D1 | K2_D2, F32>(a, n_groups, n_chunks, zpc, st);
    }
}

hipError_t launch_eval_z(const LaunchArgs& a, bool f32, int n_groups, int n_chunks, int zpc,
                         hipStream_t st) {
    return f32 ? launch_eval_z_f<true>(a, n_groups, n_chunks, zpc, st)
               : launch_eval_z_f<false>(a, n_groups, n_chunks, zpc, st);
}

hipError_t launch_reduce_z(const double* partial, int n_rows, int n_groups, int n_cand,
                           unsigned flags, double* K, double* dK, double* d2K, hipStream_t st) {
    hipLaunchKernelGGL(k2_reduce_z_kernel, dim3((n_cand + 3) / 4), dim3(256), 0, st, partial,
                       n_rows, n_groups, n_cand, flags, K, dK, d2K);
    return hipGetLastError();
}

hipError_t launch_gather(const uint4* V, const int* idx, int n, uint4* Vc, hipStream_t st) {
    hipLaunchKernelGGL(k2_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, st, V, idx, n, Vc);
    return hipGetLastError();
}

hipError_t launch_scatter3(const int* idx, int n, const double* Kc, const double* dKc,
                           const double* d2Kc, double* K, double* dK, double* d2K,
                           hipStream_t st) {
    hipLaunchKernelGGL(k2_scatter3_kernel, dim3((n + 255) / 256), dim3(256), 0, st, idx, n, Kc,
                       dKc, d2Kc, K, dK, d2K);
    return hipGetLastError();
}

hipError_t launch_abs_blocks(const double2* a, int n_pad, double* blk, hipStream_t st) {
    hipLaunchKernelGGL(k2_abs_blocks_kernel, dim3(n_pad / K2_ABS_BLOCK), dim3(256), 0, st, a, blk);
    return hipGetLastError();
}

template <int SG>
static hipError_t launch_reduce_t(const double* partial, int n_splits, int n_cand, int n_cand_pad,
                                  unsigned flags, double* K, double* dK, double* d2K,
                                  hipStream_t st) {
    constexpr int CB = 256 / SG;
    hipLaunchKernelGGL((k2_reduce_kernel<SG>), dim3((n_cand + CB - 1) / CB), dim3(256), 0, st,
                       partial, n_splits, n_cand, n_cand_pad, flags, K, dK, d2K);
    return hipGetLastError();
}

hipError_t launch_reduce(const double* partial, int n_splits, int n_cand, int n_cand_pad,
                         unsigned flags, double* K, double* dK, double* d2K, hipStream_t st) {
    if (n_splits >= 256)
        return launch_reduce_t<16>(partial, n_splits, n_cand, n_cand_pad, flags, K, dK, d2K, st);
    if (n_splits >= 32)
        return launch_reduce_t<4>(partial, n_splits, n_cand, n_cand_pad, flags, K, dK, d2K, st);
    return launch_reduce_t<1>(partial, n_splits, n_cand, n_cand_pad, flags, K, dK, d2K, st);
}

hipError_t launch_phase_words(const uint4* zg, const float4* zgf, int z0, int nz,
                              const uint4* cand, int n_cand, int mode, uint64_t* out,
                              hipStream_t st) {
    const long long n = (long long)n_cand * nz;
    const long long nb = (n + 255) / 256;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k2_phase_words_kernel, dim3((unsigned)nb), dim3(256), 0, st, zg, zgf, z0,
                       nz, cand, n_cand, mode, out);
    return hipGetLastError();
}

hipError_t launch_sincos_words(const uint64_t* W, int n, int mode, const Tab64* tab64,
                               const Tab32* tab32, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k2_sincos_words_kernel, dim3((n + 255) / 256), dim3(256), 0, st, W, n,
                       mode, tab64, tab32, out);
    return hipGetLastError();
}

hipError_t launch_newton(uint4* cand, const double* dK, const double* d2K, int n, double clamp,
                         double tol, int* active, int* steps, hipStream_t st) {
    hipLaunchKernelGGL(k2_newton_kernel, dim3((n + 255) / 256), dim3(256), 0, st, cand, dK, d2K, n,
                       clamp, tol, active, steps);
    return hipGetLastError();
}

hipError_t launch_ref_peak_ab(int n_peaks, const uint64_t* const* p, const double* dy0,
                              int n_zeros, const double* alpha, const double* gamma, double* A,
                              double* B, hipStream_t st) {
    // p: x_lo, x_hi, ghi_lo, ghi_hi, glo_lo, glo_hi, psi_lo, psi_hi
    dim3 grid((n_zeros + 255) / 256, n_peaks);
    hipLaunchKernelGGL(k2_ref_peak_ab_kernel, grid, dim3(256), 0, st, n_peaks, p[0], p[1], dy0,
                       n_zeros, alpha, gamma, p[2], p[3], p[4], p[5], p[6], p[7], A, B);
    return hipGetLastError();
}

}  // namespace k2impl
