// Enclosure mode: K2 (and K2') over the first N zeros with a rigorous error radius, in the style
// of the ball sums of certification/certify_K2.py:112-121 (term 2 e^{-k g^2} Re(w e^{i g v}),
// running sum S and magnitude sum absS = sum 2 e^{-k g^2} |w| bounding the zeros left out,
// :163-165,186-192).  The exact quantity is the sum over the table's definitions: g_j =
// 2 pi G_j / 2^110 exactly, w_j the stored complex128 values, k and v as given (v exact in
// 64.64).  The result is a midpoint K and radius R with |K - K2_exact| <= R.
//
// Error budget (u = 2^-53 the unit roundoff, every rounding at most u relative; x_j = k g_j^2;
// per zero, independent of the candidate; first order, higher orders in the final 1.01):
//   g_j     s = fl(G_hi) + G_lo 2^-64 (conversion u, sum u; the low word's conversion and the
//           power-of-two scaling add < 2^-63 u), times C = fl(2 pi / 2^46) (0.35 u: the rounding
//           of 2 pi), one product (u)                                         rel <= 3.5 u
//   x_j     (k g) g: two products and 2 x g's error                          rel <= 9 u
//           -> e^{-x} moves by e^{-x} x 9 u
//   e^{-x}  exp_neg_bounded                                                   rel <= 2.5 u
//   a_j     2 e^{-x} w (the factor 2 exact, one product per component)
//                                                            |a - A| <= |A| (3.5 + 9 x) u
//   b_j     g a (one product per component, g's 3.5 u)       |b - B| <= |B| (8 + 9 x) u
//   phase   W = floor(G V / 2^110) mod 2^64 exact; the remainder drops W bit 0 and the
//           fraction below W: < 2 pi 2^-63 rad = 6.8e-19 (0.006 u)
//   sin/cos sincos_tabw_at (2,048-point split table): table rounding 0.5 u, remainder
//           polynomial 1.15 u (cos's rounding 0.5 u, sin's dropped x^5/120 <= 0.64 u at the
//           interval edges), combine 0.55 u, the dropped bit 0.006 u -> |S - sin|, |C - cos|
//           <= 2.3 u (bounded here by 3 u; checked over 2^20 words and both edges of every
//           interval in tests/test_gpu_phase.py)
//   term    the exact products a.x C, a.y S (K2) and b.x S, b.y C (K2') at the computed S, C:
//           against a.x cos - a.y sin, sqrt(2) |a| 3 u = 4.25 |a| u
//   group   the products of K2_ENC_GROUP = 8 consecutive zeros enter a group sum by FMAs, 16
//           roundings: at most gamma_16 = 16 u / (1 - 16 u) times the sum of their magnitudes,
//           |a.x| |C| + |a.y| |S| <= sqrt(2) |a| (1 + 3 u), i.e. 22.7 |a| u per zero (the same
//           with b for K2')
//   =>      |term_j - T_j| <= |A_j| (30.5 + 9 x_j) u            (K2: 4.25 + 3.5 + 22.7 + 9 x)
//           |term'_j - T'_j| <= g_j |A_j| (35 + 9 x_j) u         (K2': 4.25 + 8 + 22.7 + 9 x)
//   sums    TwoSum (Knuth) of each group sum per lane and across splits: every addition's
//           rounding error is kept exactly in a second word; only the second word's own sum
//           rounds, which adds at most (m u)^2 sum|a_j| (m = zeros + 2 splits + 2 bounds the
//           additions), and the final hi + lo rounds by u |K|.
//   Zeros with x_j > 708 get weight 0 and contribute 2 |w_j| 1e-307 to the bound; 1e-300 covers
//   every absolute (subnormal) rounding.  The bound sums are formed in FP64 and the whole radius
//   is scaled by 1.01, which covers their own rounding (< 1e-9 relative) and |A| vs |a|.
// The radius is R = 1.01 (u B + (m u)^2 sum|a| + u |K|) + 1e-300 with B = sum_{j<N} |a_j|
// (30.5 + 9 x_j); the zeros of the table past N contribute at most tail = 1.01 sum_{j>=N} |a_j|.
// Summing each term by TwoSum (two FMAs and six additions per sum and zero) gave a radius a
// third of this size at 29% more time (221 against 171 ms on 65,536 x 2,325,572).
// tests/test_enclose_host.py replays g, x, a and b on the host and checks each line's bound
// against mpmath for x in [0, 708].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "k2_device.h"
#include "k2_internal.h"
#include "k2_tiles.h"

using namespace k2dev;

#define K2_ENC_Q 6   // bound sums: B, B', sum|a|, sum g|a| over j < N; sum|a|, sum g|a| over j >= N
#define K2_ENC_U 1.1102230246251565e-16   // 2^-53
#define K2_ENC_SAFE 1.01
#define K2_ENC_ABS 1e-300

// fixed-order sum over the 256 threads of a workgroup; result valid in thread 0
__device__ __forceinline__ double enc_block_sum(double x, double* red) {
    x = wave_sum(x);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) red[tid >> 6] = x;
    __syncthreads();
    const double s = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return s;
}

// a_j = 2 e^{-k g^2} w_j and b_j = g_j a_j from the exact words, and per-block bound sums
__global__ __launch_bounds__(256) void k2_enclose_weights_kernel(
    const uint4* __restrict__ zg, const double2* __restrict__ w, int n, int n_pad, int N, double k,
    double2* __restrict__ a, double2* __restrict__ b, double* __restrict__ part) {
    __shared__ double red[4];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double q[K2_ENC_Q] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double2 va = make_double2(0.0, 0.0), vb = va;
    if (j < n) {
        const uint4 gg = zg[j];
        const uint64_t ghi = ((uint64_t)gg.w << 32) | gg.z;
        const uint64_t glo = ((uint64_t)gg.y << 32) | gg.x;
        const double s = (double)ghi + (double)glo * 5.421010862427522e-20;   // G / 2^64
        const double g = s * 8.928943354902097e-14;                           // 2 pi / 2^46
        const double x = k * g * g;
        const double E = exp_neg_bounded(x);
        const double2 ww = w[j];
        double mag, magb, m, mb;
        if (E > 0.0) {
            const double f = 2.0 * E;
            va = make_double2(f * ww.x, f * ww.y);
            vb = make_double2(g * va.x, g * va.y);
            mag = sqrt(va.x * va.x + va.y * va.y);
            magb = g * mag;
            m = mag * (30.5 + 9.0 * x);
            mb = magb * (35.0 + 9.0 * x);
        } else {
            mag = 2.0 * sqrt(ww.x * ww.x + ww.y * ww.y) * 1e-307;
            magb = g * mag;
            m = mag;
            mb = magb;
        }
        if (j < N) {
            q[0] = m; q[1] = mb; q[2] = mag; q[3] = magb;
        } else {
            q[4] = mag; q[5] = magb;
        }
    }
    if (j < n_pad) {
        a[j] = va;
        b[j] = vb;
    }
#pragma unroll
    for (int i = 0; i < K2_ENC_Q; ++i) {
        const double s = enc_block_sum(q[i], red);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * K2_ENC_Q + i] = s;
    }
}

// one workgroup: bound sums over the block partials in a fixed order; tail outputs (optional)
__global__ __launch_bounds__(256) void k2_enclose_bounds_kernel(const double* __restrict__ part,
                                                                int n_blocks,
                                                                double* __restrict__ bounds,
                                                                double* __restrict__ tail) {
    __shared__ double red[4];
#pragma unroll
    for (int i = 0; i < K2_ENC_Q; ++i) {
        double s = 0.0;
        for (int blk = threadIdx.x; blk < n_blocks; blk += blockDim.x)
            s += part[(size_t)blk * K2_ENC_Q + i];
        s = enc_block_sum(s, red);
        if (threadIdx.x == 0) bounds[i] = s;
    }
    if (threadIdx.x == 0 && tail) {
        tail[0] = K2_ENC_SAFE * bounds[4] + K2_ENC_ABS;
        tail[1] = K2_ENC_SAFE * bounds[5] + K2_ENC_ABS;
    }
}

// C tiles over (split, candidate tile) workgroups (uniform or guided splits, block_split) with
// error-free accumulation; split s writes rows [4 s, 4 s + 4) of partial: hi, lo of K2, then
// hi, lo of K2'
// (the zero records come into scalar registers, ctile_f64s; LDS holds the split sin/cos table)
template <int NC, int FL>
__global__ __launch_bounds__(256, 4) void k2_enclose_kernel(
    ZeroView zv, int nz, int zps, int n_long, int zps_long, const uint4* __restrict__ cand,
    int n_cand, int n_ctiles, double* __restrict__ partial, int n_cand_pad,
    const Tab64* __restrict__ tab64) {
    __shared__ double s_tab[K2_TABW_DOUBLES];
    const BlockSplit bs = block_split(n_ctiles, 0, nz, zps, n_long, zps_long);
    if (bs.z0 >= nz) return;   // a padding split: no row
    const int split = bs.split, ct = bs.ct, z0 = bs.z0, z1 = bs.z1;
    const int tid = threadIdx.x;
    const double* tw = reinterpret_cast<const double*>(tab64 + K2_TAB);
    for (int i = tid; i < K2_TABW_DOUBLES; i += K2_WG) s_tab[i] = tw[i];
    Limbs4 v[NC];
    AccGD acc[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = ct * (K2_WG * NC) + i * K2_WG + tid;
        uint4 cv = make_uint4(0, 0, 0, 0);
        if (c < n_cand) cv = cand[c];
        v[i] = Limbs4{cv.x, cv.y, cv.z, cv.w};
        acc[i].zero();
    }
    __syncthreads();   // the table
    ctile_f64s<NC, FL>(zv.g, zv.a, zv.b, zv.e, z0, z1, v, acc, s_tab);
    double* prow = partial + (size_t)split * 4 * n_cand_pad;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = ct * (K2_WG * NC) + i * K2_WG + tid;
        prow[c] = acc[i].h0;
        prow[n_cand_pad + c] = acc[i].l0;
        if (FL & K2_D1) {
            prow[2 * n_cand_pad + c] = acc[i].h1;
            prow[3 * n_cand_pad + c] = acc[i].l1;
        }
    }
}

// midpoint and radius per candidate; splits combined by TwoSum in a fixed order
__global__ __launch_bounds__(256) void k2_enclose_reduce_kernel(
    const double* __restrict__ partial, int n_splits, int n_cand, int n_cand_pad, int n_q,
    const double* __restrict__ bounds, double m_terms, double* __restrict__ K,
    double* __restrict__ K_rad, double* __restrict__ dK, double* __restrict__ dK_rad) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cand) return;
    const double mu = m_terms * K2_ENC_U;
    for (int q = 0; q < n_q; ++q) {
        double h = 0.0, l = 0.0;
        for (int s = 0; s < n_splits; ++s) {
            const double* row = partial + (size_t)s * 4 * n_cand_pad + (size_t)(2 * q) * n_cand_pad;
            two_sum_acc(h, l, row[c]);
            l += row[n_cand_pad + c];
        }
        const double val = h + l;
        const double rad =
            K2_ENC_SAFE * (K2_ENC_U * bounds[q] + mu * mu * bounds[2 + q] + K2_ENC_U * fabs(val)) +
            K2_ENC_ABS;
        if (q == 0) { K[c] = val; K_rad[c] = rad; }
        else { dK[c] = val; dK_rad[c] = rad; }
    }
}

namespace k2impl {

void free_enclose(k2_table* t) {
    (void)hipFree(t->enc_a); (void)hipFree(t->enc_b); (void)hipFree(t->enc_part);
    (void)hipFree(t->enc_bounds); (void)hipFree(t->enc_partial);
    t->enc_a = t->enc_b = nullptr;
    t->enc_part = t->enc_bounds = t->enc_partial = nullptr;
    t->enc_partial_bytes = 0;
}

template <int NC, int FL>
static hipError_t launch_enclose_t(dim3 grid, const ZeroView& zv, int nz, const SplitPlan& sp,
                                   const uint4* cand, int n_cand, int n_ctiles, double* partial,
                                   int n_cand_pad, const Tab64* tab64, hipStream_t st) {
    hipLaunchKernelGGL((k2_enclose_kernel<NC, FL>), grid, dim3(K2_WG), 0, st, zv, nz, sp.zps,
                       sp.n_long, sp.zps_long, cand, n_cand, n_ctiles, partial, n_cand_pad, tab64);
    return hipGetLastError();
}

}  // namespace k2impl

// ---- C ABI -------------------------------------------------------------------------------------

using namespace k2impl;

extern "C" int k2_exp_neg_host(uint32_t n, const double* x, double* out) {
    if (n && (!x || !out)) return set_error(K2_E_ARG, "k2_exp_neg_host: null pointer");
    for (uint32_t i = 0; i < n; ++i) out[i] = exp_neg_bounded(x[i]);
    return K2_OK;
}

extern "C" int k2_enclose_device(k2_table* t, uint32_t n_cand, const uint64_t* V_dev, double k,
                                 uint32_t n_zeros, uint32_t flags, double* K_dev,
                                 double* K_rad_dev, double* dK_dev, double* dK_rad_dev,
                                 double* tail_dev, void* stream) {
    if (!t) return set_error(K2_E_ARG, "k2_enclose_device: null table");
    if (flags & ~(K2_VALUE | K2_D1))
        return set_error(K2_E_ARG, "k2_enclose_device: the enclosure mode computes K2 and K2' only");
    const bool d1 = (flags & K2_D1) != 0;
    if (n_cand && (!V_dev || !K_dev || !K_rad_dev || (d1 && (!dK_dev || !dK_rad_dev))))
        return set_error(K2_E_ARG, "k2_enclose_device: null pointer");
    if (n_cand > (1u << 30)) return set_error(K2_E_ARG, "k2_enclose_device: too many candidates");
    if (n_zeros == 0) n_zeros = t->n;
    if (n_zeros > t->n) return set_error(K2_E_ARG, "k2_enclose_device: n_zeros exceeds the table");
    if (!(k >= 0.0) || !(k < 1e300)) return set_error(K2_E_ARG, "k2_enclose_device: k must be >= 0");
    std::lock_guard<std::mutex> lk(t->mu);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipSetDevice(t->device);
    CallOrder order(t, st);   // the enclosure buffers are shared with earlier calls on the table
    if (e == hipSuccess) e = order.err();
    const int n_blk = (int)(t->n_pad / K2_WG);
    if (e == hipSuccess && !t->enc_a) e = hipMalloc((void**)&t->enc_a, (size_t)t->n_pad * sizeof(double2));
    if (e == hipSuccess && !t->enc_b) e = hipMalloc((void**)&t->enc_b, (size_t)t->n_pad * sizeof(double2));
    if (e == hipSuccess && !t->enc_part)
        e = hipMalloc((void**)&t->enc_part, (size_t)n_blk * K2_ENC_Q * sizeof(double));
    if (e == hipSuccess && !t->enc_bounds) e = hipMalloc((void**)&t->enc_bounds, 256);
    if (e != hipSuccess) {
        free_enclose(t);
        return set_error(K2_E_NOMEM, std::string("k2_enclose_device: ") + hipGetErrorString(e));
    }
    const int nc = t->tiling.nc <= 1 ? 1 : 2;
    const int n_ctiles = std::max(1, (int)((n_cand + 256u * nc - 1) / (256u * nc)));
    const int n_cand_pad = n_ctiles * 256 * nc;
    SplitPlan sp;
    if (n_cand) guided_splits((int)n_zeros, n_ctiles, t->tiling.target_blocks, K2_SPLIT_LONG, &sp);
    const int n_splits = sp.n;
    const size_t need = (size_t)std::max(1, n_splits) * 4 * n_cand_pad * sizeof(double);
    if (n_cand && need > t->enc_partial_bytes) {
        (void)hipFree(t->enc_partial);
        t->enc_partial = nullptr;
        t->enc_partial_bytes = 0;
        e = hipMalloc((void**)&t->enc_partial, need);
        if (e != hipSuccess) return set_error(K2_E_NOMEM, "k2_enclose_device: partial sums");
        t->enc_partial_bytes = need;
    }
    hipLaunchKernelGGL(k2_enclose_weights_kernel, dim3(n_blk), dim3(K2_WG), 0, st, t->g,
                       (const double2*)t->w, (int)t->n, (int)t->n_pad, (int)n_zeros, k, t->enc_a,
                       t->enc_b, t->enc_part);
    e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k2_enclose_bounds_kernel, dim3(1), dim3(K2_WG), 0, st, t->enc_part,
                           n_blk, t->enc_bounds, tail_dev);
        e = hipGetLastError();
    }
    if (e == hipSuccess && n_cand && n_splits > 0) {
        ZeroView zv{t->g, t->enc_a, t->enc_b, nullptr, nullptr, nullptr, nullptr, nullptr};
        const dim3 grid(ctile_blocks(n_ctiles, n_splits, sp.n_long));
        const uint4* cand = (const uint4*)V_dev;
        if (nc == 1)
            e = d1 ? launch_enclose_t<1, K2_D1>(grid, zv, n_zeros, sp, cand, n_cand, n_ctiles,
                                                 t->enc_partial, n_cand_pad, t->tab64, st)
                   : launch_enclose_t<1, 0>(grid, zv, n_zeros, sp, cand, n_cand, n_ctiles,
                                            t->enc_partial, n_cand_pad, t->tab64, st);
        else
            e = d1 ? launch_enclose_t<2, K2_D1>(grid, zv, n_zeros, sp, cand, n_cand, n_ctiles,
                                                 t->enc_partial, n_cand_pad, t->tab64, st)
                   : launch_enclose_t<2, 0>(grid, zv, n_zeros, sp, cand, n_cand, n_ctiles,
                                            t->enc_partial, n_cand_pad, t->tab64, st);
    }
    if (e == hipSuccess && n_cand) {
        const double m_terms = (double)n_zeros + 2.0 * n_splits + 2.0;
        hipLaunchKernelGGL(k2_enclose_reduce_kernel, dim3((n_cand + 255) / 256), dim3(256), 0, st,
                           t->enc_partial, n_splits, (int)n_cand, n_cand_pad, d1 ? 2 : 1,
                           t->enc_bounds, m_terms, K_dev, K_rad_dev, dK_dev, dK_rad_dev);
        e = hipGetLastError();
    }
    if (e != hipSuccess)
        return set_error(K2_E_HIP, std::string("k2_enclose_device: ") + hipGetErrorString(e));
    return K2_OK;
}
