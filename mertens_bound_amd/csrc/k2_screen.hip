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
// Each stage is its own launch over the candidates still in play: the first over all candidates
// (one lane each), the later ones over the survivors of the stage before, which each launch
// appends to lists in device memory (the lists' order does not matter, since every candidate
// writes only its own outputs).  A later stage starts its candidates from the first zero again
// (the earlier stages' sums are recomputed, bit for bit the same) and its waves are full however
// few candidates survive; the alternative, one launch in which a failed lane idles while its
// wave's survivors continue, leaves most lanes of the later stages masked off when survivors are
// rare.  The survivors go to K2_SCREEN_NSUB sub-lists, each with its own counter on its own
// cache line: in the first stage a wavefront appends to sub-list (wave index mod NSUB) with one
// atomic, and a later stage gives each sub-list K2_SCREEN_PER_SUB workgroups whose survivors stay
// in that sub-list (so a sub-list never outgrows its first-stage size).  One counter for all
// would take one atomic per wavefront at a single address (262,144 of them for 2^24
// candidates), which serialises.  A candidate's later stage values stay -1 once it fails.  The
// stage zero counts and thresholds are kernel arguments and the list lengths stay on the
// device: the call does not synchronise.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <vector>

#include "k2_device.h"
#include "k2_internal.h"
#include "k2_tiles.h"

using namespace k2dev;

#define K2_SCREEN_MAX 1024   // zeros the screen can stage
#define K2_SCREEN_STAGES 4
#define K2_SCREEN_MAXM 8     // shifts -M..M, M <= 8
#define K2_SCREEN_NSUB 256   // survivor sub-lists (and counters) per stage
#define K2_SCREEN_PER_SUB 8  // workgroups per sub-list in the later stages
#define K2_SCREEN_CSTRIDE 16 // ints between counters (one 64-byte line each)

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
    double k, double d, int n_zeros, ScreenRec<M>* __restrict__ rec, int* __restrict__ counts) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    // survivor counts of this call's stages (the launch has at least this many threads)
    if (j < K2_SCREEN_STAGES * K2_SCREEN_NSUB) counts[j * K2_SCREEN_CSTRIDE] = 0;
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

// Stage s of the screen.  s = 0: candidate i = lane i of the grid, for i < n_cand.  s > 0:
// workgroup b walks sub-list b mod NSUB of the stage before, entries (b / NSUB) 256 + lane in
// steps of PER_SUB 256.  Each candidate sums zeros [0, nz[s]) at all shifts, writes its stage-s
// value, and either stops (pass 0, later values -1), passes the last stage (pass 1) or joins its
// sub-list for stage s + 1 (sub-list q holds entries [q cap, q cap + counts[q])).  The best shift
// is that of its last stage.
template <int M>
__global__ __launch_bounds__(256) void k2_screen_kernel(
    const ScreenRec<M>* __restrict__ rec, const uint4* __restrict__ cand, int n_cand,
    int n_stages, ScreenStages sp, int s, const int* __restrict__ list_in,
    const int* __restrict__ count_in, int* __restrict__ list_out, int* __restrict__ count_out,
    int cap, const Tab32* __restrict__ tab, float* __restrict__ out_val,
    int* __restrict__ out_shift, int* __restrict__ out_pass) {
    __shared__ Tab32 s_tab[K2_TAB32];
    const int tid = threadIdx.x;
    for (int i = tid; i < K2_TAB32; i += blockDim.x) s_tab[i] = tab[i];
    __syncthreads();
    int q, i0, step, n;
    if (s == 0) {
        i0 = blockIdx.x * blockDim.x + tid;
        step = gridDim.x * blockDim.x;
        n = n_cand;
        q = 0;   // per wavefront below
    } else {
        q = blockIdx.x % K2_SCREEN_NSUB;
        i0 = (blockIdx.x / K2_SCREEN_NSUB) * blockDim.x + tid;
        step = K2_SCREEN_PER_SUB * blockDim.x;
        n = count_in[q * K2_SCREEN_CSTRIDE];
    }
    const int nz = sp.nz[s];
    for (int i = i0; i < n; i += step) {
        const int c = s == 0 ? i : list_in[(size_t)q * cap + i];
        const uint4 cv = cand[c];
        const Limbs4 v{cv.x, cv.y, cv.z, cv.w};
        // the lane's side of the window: V's limbs as floats (unscaled) and column 5's packed limbs
        const float4 vf = make_float4((float)v.x0, (float)v.x1, (float)v.x2, (float)v.x3);
        const uint32_t vp = pack_lo16(v.x2, v.x3);
        float acc[2 * M + 1];
#pragma unroll
        for (int m = 0; m < 2 * M + 1; ++m) acc[m] = 0.0f;
        for (int j = 0; j < nz; ++j) {
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
        int best = M;
#pragma unroll
        for (int m = 0; m < 2 * M + 1; ++m) {
            const float a = fabsf(acc[m]);
            if (a > mv) { mv = a; best = m; }
        }
        float* val = out_val + (size_t)c * K2_SCREEN_STAGES;
        val[s] = mv;
        const bool pass = mv >= sp.thr[s];
        const bool next = pass && s + 1 < n_stages;
        if (!next) {
            for (int t = s + 1; t < K2_SCREEN_STAGES; ++t) val[t] = -1.0f;
            out_shift[c] = best - M;
            out_pass[c] = pass ? 1 : 0;
        }
        // survivors join a sub-list of the next stage: one atomic per wavefront
        const uint64_t bal = __ballot(next);
        if (bal) {
            const int qo = s == 0 ? (i >> 6) % K2_SCREEN_NSUB : q;
            const int lane = __lane_id();
            const int leader = __ffsll((unsigned long long)bal) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(count_out + qo * K2_SCREEN_CSTRIDE, __popcll(bal));
            base = __shfl(base, leader, 64);
            if (next) list_out[(size_t)qo * cap + base + __popcll(bal & ((1ull << lane) - 1ull))] = c;
        }
    }
}

namespace k2impl {

// records of n_zeros zeros into rec (room for K2_SCREEN_MAX records of ScreenRec<M>), then the
// screen
// the screen's device buffer: the zero records, the stages' survivor counts, two survivor lists
// (stage s reads list (s - 1) & 1 and writes list s & 1)
struct ScreenBufs {
    void* rec;
    int* counts;    // [stage][sub-list], K2_SCREEN_CSTRIDE ints apart
    int* list[2];   // [sub-list][cap]
    int cap;        // entries per sub-list
};

// records of n_zeros zeros into rec (room for K2_SCREEN_MAX records of ScreenRec<M>), then one
// launch per stage
template <int M>
static hipError_t launch_screen_m(const uint4* zg, const double* gamma,
                                  const double2* w, double k, double d, int n_zeros,
                                  const ScreenBufs& B, const uint4* cand, int n_cand,
                                  int n_stages, ScreenStages sp, const Tab32* tab, float* val,
                                  int* shift, int* pass, hipStream_t st) {
    // at least one thread per survivor counter (they are reset here)
    const int rgrid = std::max((n_zeros + 255) / 256, K2_SCREEN_STAGES * K2_SCREEN_NSUB / 256);
    hipLaunchKernelGGL((k2_screen_records_kernel<M>), dim3(rgrid), dim3(256), 0, st, zg, gamma,
                       w, k, d, n_zeros, (ScreenRec<M>*)B.rec, B.counts);
    hipError_t e = hipGetLastError();
    const int cs = K2_SCREEN_NSUB * K2_SCREEN_CSTRIDE;   // ints of one stage's counters
    for (int s = 0; s < n_stages && e == hipSuccess; ++s) {
        const int grid = s == 0 ? (n_cand + 255) / 256 : K2_SCREEN_NSUB * K2_SCREEN_PER_SUB;
        hipLaunchKernelGGL((k2_screen_kernel<M>), dim3(grid), dim3(256), 0, st,
                           (const ScreenRec<M>*)B.rec, cand, n_cand, n_stages, sp, s,
                           s > 0 ? B.list[(s - 1) & 1] : nullptr,
                           s > 0 ? B.counts + (s - 1) * cs : nullptr, B.list[s & 1],
                           B.counts + s * cs, B.cap, tab, val, shift, pass);
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_screen(int M, const uint4* zg, const double* gamma,
                         const double2* w, double k, double d, int n_zeros, const ScreenBufs& B,
                         const uint4* cand, int n_cand, int n_stages, ScreenStages sp,
                         const Tab32* tab, float* val, int* shift, int* pass, hipStream_t st) {
    switch (M) {
#define K2_M(X) case X: return launch_screen_m<X>(zg, gamma, w, k, d, n_zeros, B, cand, \
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
    // the zero records, survivor counts and lists live in a device buffer owned by the table
    // (grown to the largest call); the stage parameters travel as kernel arguments, so the call
    // stays asynchronous
    // a first-stage sub-list takes the survivors of every NSUB-th wavefront: at most cap entries
    const int cap = (int)(((size_t)n_cand + 64 * K2_SCREEN_NSUB - 1) / (64 * K2_SCREEN_NSUB) * 64);
    const size_t off_cnt = K2_SCREEN_MAX * K2_SCREEN_REC_MAX;
    const size_t off_l0 = off_cnt + (size_t)K2_SCREEN_STAGES * K2_SCREEN_NSUB * K2_SCREEN_CSTRIDE * 4;
    const size_t list_bytes = (size_t)K2_SCREEN_NSUB * cap * 4;
    const size_t need = off_l0 + 2 * list_bytes;
    if (e == hipSuccess && need > t->screen_bytes) {
        (void)hipFree(t->screen_params);
        t->screen_params = nullptr;
        t->screen_bytes = 0;
        e = hipMalloc((void**)&t->screen_params, need);
        if (e == hipSuccess) t->screen_bytes = need;
    }
    ScreenStages sp{};
    for (int s = 0; s < n_stages; ++s) {
        sp.nz[s] = stage_zeros[s];
        sp.thr[s] = stage_thresholds[s];
    }
    char* base = (char*)t->screen_params;
    const ScreenBufs B{base, (int*)(base + off_cnt),
                       {(int*)(base + off_l0), (int*)(base + off_l0 + list_bytes)}, cap};
    if (e == hipSuccess)
        e = launch_screen(n_shift, t->g, t->gamma, (const double2*)t->w, k, shift_step,
                          stage_zeros[n_stages - 1], B, (const uint4*)V_dev, (int)n_cand,
                          n_stages, sp, t->tab32, val_dev, best_shift_dev, pass_dev, st);
    if (e != hipSuccess)
        return set_error(K2_E_HIP, std::string("k2_screen_device: ") + hipGetErrorString(e));
    return K2_OK;
}
