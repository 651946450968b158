// Staged pre-screen of candidates on the first zeros of the table, with shifted copies.
//
// Semantics of the reference's GPU leaf screen (search/gpu_mertens_stream.cu:186-250): for each
// candidate v, the sums h_m = sum_{j<n} 2 e^{-k g_j^2} Re(w_j e^{i g_j (v + m d)}) over the first
// n zeros at 2M+1 shifts m = -M..M (reference: M = 3, d = 0.0014), the stage value max_m |h_m|,
// and staged cutoffs: stage s continues only if the value after n_s zeros reaches thr_s
// (reference: 64/128/256 zeros, 0.78/0.855/0.892).  Here the phase is the precision form's window
// (columns 4..5 of G V exact in integers, column 3 in FP32: within 2^-34 turn, not a wrapped
// 64-bit approximation) and sin/cos come from the FP32 LDS table (not a 12-bit cos lookup); the
// shifted terms use the per-zero rotations e^{i m g_j d}, formed once per call in FP64:
// Re(q e^{+-i m g d}) is two FMAs each.
//
// One lane per candidate.  The zero records of all stages (<= K2_SCREEN_MAX: G and its limbs as
// scaled floats, the FP32 weight 2 e^{-k g^2} w and the rotations e^{i m g d}, m = 1..M) are formed once
// per call by k2_screen_records_kernel; every lane reads the same record at the same time, so the
// screen kernel takes them through the scalar cache, and LDS holds only the 16 KiB sin/cos table.
// A lane that fails a stage stops accumulating; its later stage values stay -1.  The stage zero
// counts and thresholds are kernel arguments: the call does not synchronise.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "k2_device.h"
#include "k2_internal.h"
#include "k2_tiles.h"

using namespace k2dev;

#define K2_SCREEN_MAX 1024   // zeros the screen can stage
#define K2_SCREEN_STAGES 4
#define K2_SCREEN_MAXM 8     // shifts -M..M, M <= 8

// stage zero counts and thresholds, passed by value (kernel arguments)
struct ScreenStages {
    int nz[K2_SCREEN_STAGES];
    float thr[K2_SCREEN_STAGES];
};

// one zero of the screen: g = (column 5's packed limbs pack_lo16(g3, g2), g1, g2, g3) and
// gvf = g0..g3 times 2^-34 as floats, the operands phase_window_f takes on its candidate side
// (the window's columns are symmetric in G and V, so the uniform G takes the scalar operand
// slots and the lane's V the other side); r[m-1] = e^{i m g d} (r padded to an odd count:
// 16-byte records)
template <int M>
struct ScreenRec {
    uint4 g;
    float4 gvf;
    float2 a;     // 2 e^{-k g^2} w (FP32)
    float2 r[M | 1];
};
#define K2_SCREEN_REC_MAX sizeof(ScreenRec<K2_SCREEN_MAXM>)

// the zero records of one call, one thread per zero
template <int M>
__global__ __launch_bounds__(256) void k2_screen_records_kernel(
    const uint4* __restrict__ zg, const double* __restrict__ gamma, const double2* __restrict__ w,
    double k, double d, int n_zeros, ScreenRec<M>* __restrict__ rec) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_zeros) return;
    const double g = gamma[j];
    const double f = 2.0 * exp(-k * g * g);
    const double2 ww = w[j];
    ScreenRec<M> z;
    const uint4 gg = zg[j];
    z.g = make_uint4(pack_lo16(gg.w, gg.z), gg.y, gg.z, gg.w);
    z.gvf = window_vf(Limbs4{gg.x, gg.y, gg.z, gg.w});
    z.a = make_float2((float)(f * ww.x), (float)(f * ww.y));
#pragma unroll
    for (int m = 0; m < (M | 1); ++m) {
        double sr = 0.0, cr = 1.0;
        if (m < M) sincos_general((double)(m + 1) * g * d, sr, cr);
        z.r[m] = make_float2((float)cr, (float)sr);
    }
    rec[j] = z;
}

template <int M>
__global__ __launch_bounds__(256) void k2_screen_kernel(
    const ScreenRec<M>* __restrict__ rec, const uint4* __restrict__ cand, int n_cand,
    int n_stages, ScreenStages sp,
    const Tab32* __restrict__ tab, float* __restrict__ out_val, int* __restrict__ out_shift,
    int* __restrict__ out_pass) {
    __shared__ Tab32 s_tab[K2_TAB32];
    const int tid = threadIdx.x;
    for (int i = tid; i < K2_TAB32; i += blockDim.x) s_tab[i] = tab[i];
    __syncthreads();
    const int c = blockIdx.x * blockDim.x + tid;
    if (c >= n_cand) return;
    const uint4 cv = cand[c];
    const Limbs4 v{cv.x, cv.y, cv.z, cv.w};
    // the lane's side of the window: V's limbs as floats (unscaled) and column 5's packed limbs
    const float4 vf = make_float4((float)v.x0, (float)v.x1, (float)v.x2, (float)v.x3);
    const uint32_t vp = pack_lo16(v.x2, v.x3);
    float acc[2 * M + 1];
#pragma unroll
    for (int m = 0; m < 2 * M + 1; ++m) acc[m] = 0.0f;
    int j = 0, best = M, passed = 1;
    for (int s = 0; s < n_stages; ++s) {
        const int nz = sp.nz[s];
        for (; j < nz; ++j) {
            const ScreenRec<M> z = rec[j];   // the same j in every lane
            uint32_t hi, lo;
            float t3;
            phase_window_f<K2_TAB32_BITS, true>(v, vp, vf, Limbs4{0u, z.g.y, z.g.z, z.g.w},
                                                z.g.x, z.gvf, hi, lo, t3);
            float S, C;
            sincos_window_f<K2_TAB32_BITS>(hi, lo, t3, s_tab, S, C);
            // q = a e^{i th}
            const float qr = fmaf(z.a.x, C, -z.a.y * S);
            const float qi = fmaf(z.a.x, S, z.a.y * C);
            acc[M] += qr;
#pragma unroll
            for (int m = 1; m <= M; ++m) {   // Re(q e^{+i m g d}), Re(q e^{-i m g d})
                const float2 r = z.r[m - 1];
                acc[M + m] = fmaf(-qi, r.y, fmaf(qr, r.x, acc[M + m]));
                acc[M - m] = fmaf(qi, r.y, fmaf(qr, r.x, acc[M - m]));
            }
        }
        float mv = -1.0f;
#pragma unroll
        for (int m = 0; m < 2 * M + 1; ++m) {
            const float a = fabsf(acc[m]);
            if (a > mv) { mv = a; best = m; }
        }
        out_val[(size_t)c * K2_SCREEN_STAGES + s] = mv;
        if (mv < sp.thr[s]) {
            passed = 0;
            for (int t = s + 1; t < K2_SCREEN_STAGES; ++t)
                out_val[(size_t)c * K2_SCREEN_STAGES + t] = -1.0f;
            break;
        }
    }
    for (int t = n_stages; t < K2_SCREEN_STAGES; ++t)
        out_val[(size_t)c * K2_SCREEN_STAGES + t] = -1.0f;
    out_shift[c] = best - M;
    out_pass[c] = passed;
}

namespace k2impl {

// records of n_zeros zeros into rec (room for K2_SCREEN_MAX records of ScreenRec<M>), then the
// screen
template <int M>
static hipError_t launch_screen_m(const uint4* zg, const double* gamma,
                                  const double2* w, double k, double d, int n_zeros, void* rec,
                                  const uint4* cand, int n_cand, int n_stages, ScreenStages sp,
                                  const Tab32* tab, float* val, int* shift, int* pass,
                                  hipStream_t st) {
    hipLaunchKernelGGL((k2_screen_records_kernel<M>), dim3((n_zeros + 255) / 256), dim3(256), 0,
                       st, zg, gamma, w, k, d, n_zeros, (ScreenRec<M>*)rec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k2_screen_kernel<M>), dim3((n_cand + 255) / 256), dim3(256), 0, st,
                       (const ScreenRec<M>*)rec, cand, n_cand, n_stages, sp, tab, val, shift,
                       pass);
    return hipGetLastError();
}

hipError_t launch_screen(int M, const uint4* zg, const double* gamma,
                         const double2* w, double k, double d, int n_zeros, void* rec,
                         const uint4* cand, int n_cand, int n_stages, ScreenStages sp,
                         const Tab32* tab, float* val, int* shift, int* pass, hipStream_t st) {
    switch (M) {
#define K2_M(X) case X: return launch_screen_m<X>(zg, gamma, w, k, d, n_zeros, rec, cand, \
                                                  n_cand, n_stages, sp, tab, val, shift, pass, \
                                                  st);
        K2_M(0) K2_M(1) K2_M(2) K2_M(3) K2_M(4) K2_M(5) K2_M(6) K2_M(7) K2_M(8)
#undef K2_M
        default: return hipErrorInvalidValue;
    }
}

}  // namespace k2impl

// ---- C ABI -------------------------------------------------------------------------------------

using namespace k2impl;

extern "C" int k2_screen_device(k2_table* t, uint32_t n_cand, const uint64_t* V_dev, double k,
                                double shift_step, int n_shift, int n_stages,
                                const int32_t* stage_zeros, const float* stage_thresholds,
                                float* val_dev, int32_t* best_shift_dev, int32_t* pass_dev,
                                void* stream) {
    if (!t || (n_cand && (!V_dev || !val_dev || !best_shift_dev || !pass_dev)) || !stage_zeros ||
        !stage_thresholds)
        return set_error(K2_E_ARG, "k2_screen_device: null pointer");
    if (n_stages < 1 || n_stages > K2_SCREEN_STAGES || n_shift < 0 || n_shift > K2_SCREEN_MAXM)
        return set_error(K2_E_ARG, "k2_screen_device: 1..4 stages and 0..8 shifts per side");
    int prev = 0;
    for (int s = 0; s < n_stages; ++s) {
        if (stage_zeros[s] <= prev || stage_zeros[s] > K2_SCREEN_MAX ||
            (uint32_t)stage_zeros[s] > t->n)
            return set_error(K2_E_ARG,
                             "k2_screen_device: stage zero counts must increase, <= 1024");
        prev = stage_zeros[s];
    }
    if (!(k >= 0.0) || !(shift_step == shift_step))
        return set_error(K2_E_ARG, "k2_screen_device: bad k or shift step");
    if (n_cand > (1u << 30)) return set_error(K2_E_ARG, "k2_screen_device: too many candidates");
    if (n_cand == 0) return K2_OK;
    std::lock_guard<std::mutex> lk(t->mu);
    hipError_t e = hipSetDevice(t->device);
    hipStream_t st = (hipStream_t)stream;
    CallOrder order(t, st);   // the zero records are shared with earlier calls on the table
    if (e == hipSuccess) e = order.err();
    // the zero records live in a device buffer owned by the table; the stage parameters travel
    // as kernel arguments, so the call stays asynchronous
    if (e == hipSuccess && t->screen_params == nullptr)
        e = hipMalloc((void**)&t->screen_params, K2_SCREEN_MAX * K2_SCREEN_REC_MAX);
    ScreenStages sp{};
    for (int s = 0; s < n_stages; ++s) {
        sp.nz[s] = stage_zeros[s];
        sp.thr[s] = stage_thresholds[s];
    }
    if (e == hipSuccess)
        e = launch_screen(n_shift, t->g, t->gamma, (const double2*)t->w, k, shift_step,
                          stage_zeros[n_stages - 1], t->screen_params, (const uint4*)V_dev,
                          (int)n_cand, n_stages, sp, t->tab32, val_dev, best_shift_dev, pass_dev,
                          st);
    if (e != hipSuccess)
        return set_error(K2_E_HIP, std::string("k2_screen_device: ") + hipGetErrorString(e));
    return K2_OK;
}
