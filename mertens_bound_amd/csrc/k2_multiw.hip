// Several real weight vectors in one pass over the zeros:
//
//   K_m(v) = sum_{j < n_zeros} wt_m[j] * 2 Re(w_j e^{i g_j v}),   m < n_w <= 8
//
// (and K_m' = -sum wt_m g 2 Im(...), K_m'' = -sum wt_m g^2 2 Re(...)), the quantity the
// reference forms as h = A @ wt per weight vector after compute_peak_AB (A_j = 2 Re(w_j e^{i g_j v}),
// B_j = 2 Im(...); search/eval_tech1_600k.py:243-267 with the Gaussian, JP-kernel, raw and
// truncated vectors).  The exact phase and the sin/cos are computed once per candidate and zero
// and shared by all vectors: each further vector costs one FMA per requested sum.
//
// C tiles of 256 lanes, NC candidates per lane (4 for up to 5 sums x vectors, 2 for up to 10);
// every lane handles the same zero at a time, so the zero's record (G, w, g, the M weights) is
// read with scalar loads into scalar registers (ld_const) and the VALU instructions take it as
// their scalar operand, as in the one-launch FP64 C tile (ctile_f64s); LDS holds the split sin/cos
// table.  The sums take wt_m as stored and are doubled when written (exact: a power of two).
// Split partial rows [split][q*M + m][cand].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "k2_device.h"
#include "k2_internal.h"
#include "k2_tiles.h"

using namespace k2dev;

#define K2_MW_MAX 8      // weight vectors per call

template <int M, int FL, int NC>
__global__ __launch_bounds__(256, 3) void k2_multiw_kernel(
    const uint4* __restrict__ zg, const double2* __restrict__ zw, const double* __restrict__ zgam,
    const double* __restrict__ wt, int ld, int nz, k2impl::SplitPlan sp,
    const uint4* __restrict__ cand, int n_cand, int n_ctiles, double* __restrict__ partial,
    int n_cand_pad, const Tab64* __restrict__ tab64) {
    constexpr bool D1 = (FL & K2_D1) != 0;
    constexpr bool D2 = (FL & K2_D2) != 0;
    constexpr bool DER = D1 || D2;
    constexpr int NQ = 1 + (D1 ? 1 : 0) + (D2 ? 1 : 0);
    __shared__ double s_tab[K2_TABW_DOUBLES];
    const BlockSplit bs = block_split(n_ctiles, 0, nz, sp.zps, sp.n_long, sp.zps_long);
    if (bs.z0 >= nz) return;   // a padding split: no row
    const int split = bs.split, ct = bs.ct, z0 = bs.z0, z1 = bs.z1;
    const int tid = threadIdx.x;
    const double* tw = reinterpret_cast<const double*>(tab64 + K2_TAB);
    for (int i = tid; i < K2_TABW_DOUBLES; i += 256) s_tab[i] = tw[i];
    Limbs4 v[NC];
    uint32_t vp[NC];
    double acc[NC][NQ][M];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = ct * 256 * NC + i * 256 + tid;
        uint4 cv = make_uint4(0, 0, 0, 0);
        if (c < n_cand) cv = cand[c];
        v[i] = Limbs4{cv.x, cv.y, cv.z, cv.w};
        vp[i] = pack_lo16(cv.w, cv.z);
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int m = 0; m < M; ++m) acc[i][q][m] = 0.0;
    }
    __syncthreads();   // the table
    const SinCosCoefW kc = sincos_coef_w();
    struct Rec { uint4 g; double2 w; double gam; double wt[M]; };
    auto load = [&](int j) -> Rec {
        Rec r;
        r.g = ld_const(zg, j);
        r.w = ld_const(zw, j);
        r.gam = DER ? ld_const(zgam, j) : 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) r.wt[m] = ld_const(wt + (size_t)m * ld, j);
        return r;
    };
    if (z1 > z0) {
        Rec cur = load(z0);
        for (int j = z0; j < z1; ++j) {
            const Rec nxt = load(min(j + 1, z1 - 1));   // in flight while zero j is summed
            double S[NC], C[NC];
            ctw_sincos<NC, true>(cur.g, pack_lo16_s(cur.g.z, cur.g.w), v, vp, s_tab, kc, S, C);
            const double2 ww = cur.w;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double A = fma(ww.x, C[i], -ww.y * S[i]);          // Re(w e^{i th})
#pragma unroll
                for (int m = 0; m < M; ++m) acc[i][0][m] = fma(cur.wt[m], A, acc[i][0][m]);
                if constexpr (D1) {
                    const double gB = cur.gam * fma(ww.x, S[i], ww.y * C[i]);   // g Im(w e^{i th})
#pragma unroll
                    for (int m = 0; m < M; ++m) acc[i][1][m] = fma(-cur.wt[m], gB, acc[i][1][m]);
                }
                if constexpr (D2) {
                    const double ggA = cur.gam * cur.gam * A;
#pragma unroll
                    for (int m = 0; m < M; ++m)
                        acc[i][NQ - 1][m] = fma(-cur.wt[m], ggA, acc[i][NQ - 1][m]);
                }
            }
            cur = nxt;
        }
    }
    double* prow = partial + (size_t)split * NQ * M * n_cand_pad;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = ct * 256 * NC + i * 256 + tid;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int m = 0; m < M; ++m)
                prow[(size_t)(q * M + m) * n_cand_pad + c] = 2.0 * acc[i][q][m];
    }
}

// row q*M + m of the split partials, summed over splits in order, into K (q = 0), then K' and
// K'' as requested; each output is [m][cand]
__global__ __launch_bounds__(256) void k2_multiw_reduce_kernel(
    const double* __restrict__ partial, int n_splits, int M, int nq, int has_d1, int n_cand,
    int n_cand_pad, double* __restrict__ K, double* __restrict__ dK, double* __restrict__ d2K) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int rows = nq * M;
    if (i >= (long long)rows * n_cand) return;
    const int r = (int)(i / n_cand), c = (int)(i - (long long)r * n_cand);
    double s = 0.0;
    for (int sp = 0; sp < n_splits; ++sp) s += partial[((size_t)sp * rows + r) * n_cand_pad + c];
    const int q = r / M, m = r - q * M;
    double* dst = q == 0 ? K : ((q == 1 && has_d1) ? dK : d2K);
    dst[(size_t)m * n_cand + c] = s;
}

namespace k2impl {

// candidates per lane for M vectors and these sums
static int mw_nc(int M, unsigned fl) {
    const int nq = 1 + ((fl & K2_D1) ? 1 : 0) + ((fl & K2_D2) ? 1 : 0);
    return nq * M <= 5 ? 4 : (nq * M <= 10 ? 2 : 1);
}

template <int M, int FL>
static hipError_t launch_mw_t(dim3 grid, const k2_table* t, const double* wt, int ld, int nz,
                              const SplitPlan& sp, const uint4* cand, int n_cand, int n_ctiles,
                              double* partial, int n_cand_pad, hipStream_t st) {
    constexpr int NQ = 1 + ((FL & K2_D1) ? 1 : 0) + ((FL & K2_D2) ? 1 : 0);
    constexpr int NC = NQ * M <= 5 ? 4 : (NQ * M <= 10 ? 2 : 1);
    hipLaunchKernelGGL((k2_multiw_kernel<M, FL, NC>), grid, dim3(256), 0, st, t->g,
                       (const double2*)t->w, t->gamma, wt, ld, nz, sp, cand, n_cand, n_ctiles,
                       partial, n_cand_pad, t->tab64);
    return hipGetLastError();
}

template <int M>
static hipError_t launch_mw_fl(unsigned fl, dim3 grid, const k2_table* t, const double* wt,
                               int ld, int nz, const SplitPlan& sp, const uint4* cand, int n_cand,
                               int n_ctiles, double* partial, int n_cand_pad, hipStream_t st) {
    switch (fl & (K2_D1 | K2_D2)) {
        case 0: return launch_mw_t<M, 0>(grid, t, wt, ld, nz, sp, cand, n_cand, n_ctiles, partial, n_cand_pad, st);
        case K2_D1: return launch_mw_t<M, K2_D1>(grid, t, wt, ld, nz, sp, cand, n_cand, n_ctiles, partial, n_cand_pad, st);
        case K2_D2: return launch_mw_t<M, K2_D2>(grid, t, wt, ld, nz, sp, cand, n_cand, n_ctiles, partial, n_cand_pad, st);
        default: return launch_mw_t<M, K2_D1 | K2_D2>(grid, t, wt, ld, nz, sp, cand, n_cand, n_ctiles, partial, n_cand_pad, st);
    }
}

static hipError_t launch_mw(int M, unsigned fl, dim3 grid, const k2_table* t, const double* wt,
                            int ld, int nz, const SplitPlan& sp, const uint4* cand, int n_cand,
                            int n_ctiles, double* partial, int n_cand_pad, hipStream_t st) {
    switch (M) {
#define K2_MW(X) case X: return launch_mw_fl<X>(fl, grid, t, wt, ld, nz, sp, cand, n_cand, n_ctiles, partial, n_cand_pad, st);
        K2_MW(1) K2_MW(2) K2_MW(3) K2_MW(4) K2_MW(5) K2_MW(6) K2_MW(7) K2_MW(8)
#undef K2_MW
        default: return hipErrorInvalidValue;
    }
}

}  // namespace k2impl

// ---- C ABI -------------------------------------------------------------------------------------

using namespace k2impl;

extern "C" int k2_eval_weights_device(k2_table* t, uint32_t n_cand, const uint64_t* V_dev,
                                      uint32_t n_w, const double* wt_dev, uint32_t ld,
                                      uint32_t n_zeros, uint32_t flags, double* K_dev,
                                      double* dK_dev, double* d2K_dev, void* stream) {
    if (!t) return set_error(K2_E_ARG, "k2_eval_weights_device: null table");
    if (n_w < 1 || n_w > K2_MW_MAX)
        return set_error(K2_E_ARG, "k2_eval_weights_device: 1..8 weight vectors");
    if (flags & ~(K2_VALUE | K2_D1 | K2_D2))
        return set_error(K2_E_ARG, "k2_eval_weights_device: flags are K2_VALUE, K2_D1, K2_D2");
    if (n_zeros == 0) n_zeros = t->n;
    if (n_zeros > t->n || ld < n_zeros)
        return set_error(K2_E_ARG, "k2_eval_weights_device: n_zeros exceeds the table or ld");
    if (n_cand && (!V_dev || !wt_dev || !K_dev || ((flags & K2_D1) && !dK_dev) ||
                   ((flags & K2_D2) && !d2K_dev)))
        return set_error(K2_E_ARG, "k2_eval_weights_device: null pointer");
    if (n_cand > (1u << 30)) return set_error(K2_E_ARG, "k2_eval_weights_device: too many candidates");
    if (n_cand == 0) return K2_OK;
    std::lock_guard<std::mutex> lk(t->mu);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipSetDevice(t->device);
    CallOrder order(t, st);   // the partial sums are shared with earlier calls on the table
    if (e == hipSuccess) e = order.err();
    const int cw = 256 * mw_nc((int)n_w, flags);   // candidates per C tile
    const int n_ctiles = (int)((n_cand + cw - 1) / cw);
    const int n_cand_pad = n_ctiles * cw;
    const int nq = 1 + ((flags & K2_D1) ? 1 : 0) + ((flags & K2_D2) ? 1 : 0);
    const int rows = nq * (int)n_w;
    // split rows of nq * n_w sums each: at most about 1 GiB of partial sums
    const double row_bytes = (double)rows * n_cand_pad * sizeof(double);
    const int target = (int)std::max(1.0, std::min((double)t->tiling.target_blocks,
                                                   (double)(1u << 30) / row_bytes * n_ctiles));
    SplitPlan sp;
    guided_splits((int)n_zeros, n_ctiles, target, K2_SPLIT_LONG_MW, &sp);
    const int n_splits = sp.n;
    const size_t need = (size_t)std::max(1, n_splits) * rows * n_cand_pad * sizeof(double);
    if (e == hipSuccess && need > t->partial_bytes) {
        (void)hipFree(t->partial);
        t->partial = nullptr;
        t->partial_bytes = 0;
        e = hipMalloc((void**)&t->partial, need);
        if (e == hipSuccess) t->partial_bytes = need;
    }
    if (e == hipSuccess && n_splits > 0)
        e = launch_mw((int)n_w, flags, dim3(ctile_blocks(n_ctiles, n_splits, sp.n_long)), t,
                      wt_dev, (int)ld, (int)n_zeros, sp, (const uint4*)V_dev, (int)n_cand, n_ctiles, t->partial,
                      n_cand_pad, st);
    if (e == hipSuccess) {
        const long long cnt = (long long)rows * n_cand;
        hipLaunchKernelGGL(k2_multiw_reduce_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256),
                           0, st, t->partial, n_splits, (int)n_w, nq, (flags & K2_D1) ? 1 : 0,
                           (int)n_cand, n_cand_pad, K_dev, dK_dev, d2K_dev);
        e = hipGetLastError();
    }
    if (e != hipSuccess)
        return set_error(K2_E_HIP, std::string("k2_eval_weights_device: ") + hipGetErrorString(e));
    return K2_OK;
}
