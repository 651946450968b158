// Staged pre-screen of candidates on the first zeros of the table, with shifted copies.
//
// Semantics of the reference's GPU leaf screen (search/gpu_mertens_stream.cu:186-250): for each
// candidate v, the sums h_m = sum_{j<n} 2 e^{-k g_j^2} Re(w_j e^{i g_j (v + m d)}) over the first
// n zeros at 2M+1 shifts m = -M..M (reference: M = 3, d = 0.0014), the stage value max_m |h_m|,
// and staged cutoffs: stage s continues only if the value after n_s zeros reaches thr_s
// (reference: 64/128/256 zeros, 0.78/0.855/0.892).  Here the phase is the precision form's window
// (columns 4..5 of G V exact in integers, column 3 in FP32: within 2^-34 turn, not a wrapped
// 64-bit approximation) and sin/cos come from the FP32 LDS table (not a 12-bit cos lookup); the
// shifted terms use the per-zero rotations e^{i m g_j d}, formed once per call in FP64.  With
// q_j = a_j e^{i g_j v}, Re(q e^{+-i m g d}) = Re q cos(m g d) -+ Im q sin(m g d), so a candidate
// keeps a0 = sum Re q, A_m = sum Re q cos(m g d) and B_m = sum Im q sin(m g d) (one FMA each per
// zero, 2M + 1 in all) and forms h_{+-m} = A_m -+ B_m once at the end, where summing each shifted
// term separately takes two FMAs per shift.
//
// The zero records of all stages (<= K2_SCREEN_MAX: G and its limbs as scaled floats, the FP32
// weight 2 e^{-k g^2} w and the rotations e^{i m g d}, m = 1..M) are formed once per call by
// k2_screen_records_kernel.  Every lane of a wave reads the same record at the same time, so the
// kernels take them through the scalar cache, and LDS holds the 16 KiB sin/cos table.  Each stage
// is its own launch over the candidates still in play:
// * the first over all candidates, one lane each;
// * the later ones over the survivors of the stage before, which each launch appends to lists in
//   device memory (the lists' order does not matter, since every candidate writes only its own
//   outputs).  A later stage starts its candidates from the first zero again (the earlier
//   stages' sums are recomputed, bit for bit the same in every run), and its waves are full
//   however few candidates survive: the four waves of a workgroup share 64 survivors, each wave
//   summing a quarter of the zeros, so that few survivors still make enough waves to fill the
//   GPU, and the workgroups take the survivors of all sub-lists as one list, so that survivors
//   that cluster in some sub-lists (those near one peak do) still spread over all of them.
// The survivors go to K2_SCREEN_NSUB sub-lists, each with its own counter on its own cache line:
// a wavefront of the first stage appends to sub-list (wave index mod NSUB), pass P of a later
// stage to sub-list P mod NSUB, with one atomic per wavefront.  One counter for all would take
// one atomic per wavefront at a single address (262,144 of them for 2^24 candidates), which
// serialises.  A candidate's later stage values stay -1 once it fails.  The stage zero counts and
// thresholds are kernel arguments and the list lengths stay on the device: the call does not
// synchronise.
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
#define K2_SCREEN_CSTRIDE 16 // ints between counters (one 64-byte line each)
#define K2_SCREEN_GRID 2048  // workgroups of a later stage (each loops over passes)

// stage zero counts and thresholds, passed by value (kernel arguments)
struct ScreenStages {
    int nz[K2_SCREEN_STAGES];
    float thr[K2_SCREEN_STAGES];
};

// one zero of the screen: g = (column 5's packed limbs pack_lo16(g3, g2), g1, g2, g3) and
// gvf = g0..g3 times 2 pi / 2^78 as floats, the operands phase_window_f takes on its
// candidate side (the window's columns are symmetric in G and V, so the uniform G takes the
// scalar operand slots and the lane's V the other side); r[m-1] = e^{i m g d} (r padded to an
// odd count: 16-byte records)
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
    z.gvf = window_gx(Limbs4{gg.x, gg.y, gg.z, gg.w});
    z.a = make_float2((float)(f * ww.x), (float)(f * ww.y));
#pragma unroll
    for (int m = 0; m < (M | 1); ++m) {
        double sr = 0.0, cr = 1.0;
        if (m < M) sincos_general((double)(m + 1) * g * d, sr, cr);
        z.r[m] = make_float2((float)cr, (float)sr);
    }
    rec[j] = z;
}

// a candidate's shifted sums: a0 = sum Re q_j, A[m-1] = sum Re q_j cos(m g_j d),
// B[m-1] = sum Im q_j sin(m g_j d)
template <int M>
struct ShiftSums {
    float a0;
    float A[M > 0 ? M : 1], B[M > 0 ? M : 1];
    __device__ __forceinline__ void zero() {
        a0 = 0.0f;
#pragma unroll
        for (int m = 0; m < M; ++m) A[m] = B[m] = 0.0f;
    }
};

// one zero's term for candidate v (v's window operands vp, vf); SB: the record is uniform over
// the wave and sits in scalar registers
template <int M, bool SB>
__device__ __forceinline__ void screen_term(const Limbs4& v, uint32_t vp, const float4& vf,
                                            const ScreenRec<M>& z, const Tab32* s_tab,
                                            ShiftSums<M>& acc) {
    uint32_t hi, lo;
    float t3;
    phase_window_f<K2_TAB32_BITS, SB>(v, vp, vf, Limbs4{0u, z.g.y, z.g.z, z.g.w}, z.g.x, z.gvf,
                                       hi, lo, t3);
    float S, C;
    sincos_window_f<K2_TAB32_BITS>(hi, lo, t3, s_tab, S, C);
    // q = a e^{i th}
    const float qr = fmaf(z.a.x, C, -z.a.y * S);
    const float qi = fmaf(z.a.x, S, z.a.y * C);
    acc.a0 += qr;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        acc.A[m] = fmaf(qr, z.r[m].x, acc.A[m]);
        acc.B[m] = fmaf(qi, z.r[m].y, acc.B[m]);
    }
}

// zeros [z0, z1) in order, two per step: the second zero's phase and sin/cos do not wait for the
// first one's sums, so the two chains interleave (the sums take the terms in the same order).
// Measured against one zero per step: 0.79 against 0.84 ms for the first stage over 2^24
// candidates; four per step took 0.83 ms there (and gained 10 us in the later stages).
template <int M>
__device__ __forceinline__ void screen_zeros(const Limbs4& v, uint32_t vp, const float4& vf,
                                             const ScreenRec<M>* __restrict__ rec, int z0, int z1,
                                             const Tab32* s_tab, ShiftSums<M>& acc) {
    int j = z0;
    for (; j + 1 < z1; j += 2) {
        const ScreenRec<M> r0 = rec[j], r1 = rec[j + 1];
        screen_term<M, true>(v, vp, vf, r0, s_tab, acc);
        screen_term<M, true>(v, vp, vf, r1, s_tab, acc);
    }
    if (j < z1) screen_term<M, true>(v, vp, vf, rec[j], s_tab, acc);
}

// the stage value max_m |h_m| and the first shift that reaches it (m = -M..M in order)
template <int M>
__device__ __forceinline__ float stage_value(const ShiftSums<M>& acc, int& best) {
    // h_{-m} = A + B, h_{+m} = A - B; shifts in the order -M..M with a strict comparison: the
    // first maximum wins
    float mv = -1.0f;
    best = 0;
#pragma unroll
    for (int i = 0; i < 2 * M + 1; ++i) {
        const int m = i - M;
        const float h = m == 0 ? acc.a0
                               : (m < 0 ? acc.A[-m - 1] + acc.B[-m - 1]
                                        : acc.A[m - 1] - acc.B[m - 1]);
        const float a = fabsf(h);
        if (a > mv) { mv = a; best = m; }
    }
    return mv;
}

// a candidate's stage-s outputs; returns whether it goes on to stage s + 1
__device__ __forceinline__ bool screen_outputs(int c, int s, int n_stages, const ScreenStages& sp,
                                               float mv, int best, float* __restrict__ out_val,
                                               int* __restrict__ out_shift,
                                               int* __restrict__ out_pass) {
    float* val = out_val + (size_t)c * K2_SCREEN_STAGES;
    val[s] = mv;
    const bool pass = mv >= sp.thr[s];
    const bool next = pass && s + 1 < n_stages;
    if (!next) {
        for (int t = s + 1; t < K2_SCREEN_STAGES; ++t) val[t] = -1.0f;
        out_shift[c] = best;
        out_pass[c] = pass ? 1 : 0;
    }
    return next;
}

// survivors join sub-list qo of the next stage: one atomic per wavefront (lanes with take set)
__device__ __forceinline__ void screen_append(bool take, int c, int qo, int* __restrict__ list_out,
                                              int* __restrict__ count_out, int cap) {
    const uint64_t bal = __ballot(take);
    if (!bal) return;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)bal) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count_out + qo * K2_SCREEN_CSTRIDE, __popcll(bal));
    base = __shfl(base, leader, 64);
    if (take) list_out[(size_t)qo * cap + base + __popcll(bal & ((1ull << lane) - 1ull))] = c;
}

// Stage 0: candidate i = lane i of the grid, for i < n_cand, summing zeros [0, nz[0]) at all
// shifts; it writes its stage value and either stops (pass 0, later values -1), passes the last
// stage (pass 1) or joins sub-list (wave index mod NSUB) of stage 1 (sub-list q holds entries
// [q cap, q cap + counts[q])).  The best shift is that of its last stage.
template <int M>
__global__ __launch_bounds__(256) void k2_screen_kernel(
    const ScreenRec<M>* __restrict__ rec, const uint4* __restrict__ cand, int n_cand,
    int n_stages, ScreenStages sp, int* __restrict__ list_out, int* __restrict__ count_out,
    int cap, const Tab32* __restrict__ tab, float* __restrict__ out_val,
    int* __restrict__ out_shift, int* __restrict__ out_pass) {
    __shared__ Tab32 s_tab[K2_TAB32];
    const int tid = threadIdx.x;
    for (int i = tid; i < K2_TAB32; i += blockDim.x) s_tab[i] = tab[i];
    __syncthreads();
    const int nz = sp.nz[0];
    for (int i = blockIdx.x * blockDim.x + tid; i < n_cand; i += gridDim.x * blockDim.x) {
        const uint4 cv = cand[i];
        const Limbs4 v{cv.x, cv.y, cv.z, cv.w};
        // the lane's side of the window: V's limbs as floats (unscaled) and column 5's packed limbs
        const float4 vf = make_float4((float)v.x0, (float)v.x1, (float)v.x2, (float)v.x3);
        const uint32_t vp = pack_lo16(v.x2, v.x3);
        ShiftSums<M> acc;
        acc.zero();
        screen_zeros<M>(v, vp, vf, rec, 0, nz, s_tab, acc);
        int best;
        const float mv = stage_value<M>(acc, best);
        const bool next = screen_outputs(i, 0, n_stages, sp, mv, best, out_val, out_shift,
                                         out_pass);
        screen_append(next, i, (i >> 6) % K2_SCREEN_NSUB, list_out, count_out, cap);
    }
}

// Stage s > 0, over the survivors of stage s - 1 in its NSUB sub-lists, taken as one list: entry
// e of the concatenation is entry e - pre[q] of sub-list q, pre the prefix sums of the sub-list
// counts (formed by every workgroup from the counters).  Workgroup pass P (P = b, b + grid, ...)
// takes entries [64 P, 64 P + 64), one per lane, and its four waves share them: wave w sums zeros
// [w Z, w Z + Z) of [0, nz[s]), Z = ceil(nz / 4), with the record uniform over the wave (scalar
// loads, as in stage 0), and wave 0 adds the other waves' sums (in wave order, through LDS) to
// its own.  Survivors of pass P join sub-list P mod NSUB of stage s + 1: a sub-list receives at
// most 64 entries from each of ceil(n / (64 NSUB)) passes, within its capacity.  Candidates that
// cluster in a few sub-lists of the stage before (survivors near one peak do) still spread over
// every workgroup.
template <int M>
__global__ __launch_bounds__(256) void k2_screen_later_kernel(
    const ScreenRec<M>* __restrict__ rec, const uint4* __restrict__ cand, int n_stages,
    ScreenStages sp, int s, const int* __restrict__ list_in, const int* __restrict__ count_in,
    int* __restrict__ list_out, int* __restrict__ count_out, int cap,
    const Tab32* __restrict__ tab, float* __restrict__ out_val, int* __restrict__ out_shift,
    int* __restrict__ out_pass) {
    constexpr int NW = 4;   // waves per workgroup
    constexpr int NS = 2 * M + 1;
    static_assert(K2_SCREEN_NSUB == NW * 64, "one sub-list counter per thread");
    __shared__ Tab32 s_tab[K2_TAB32];
    __shared__ int s_pre[K2_SCREEN_NSUB + 1];
    __shared__ int s_wsum[NW];
    __shared__ float s_part[NW - 1][NS][64];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int i = tid; i < K2_TAB32; i += blockDim.x) s_tab[i] = tab[i];
    // exclusive prefix sums of the sub-list counts: a scan within each wave, then the wave totals
    int x = count_in[tid * K2_SCREEN_CSTRIDE];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wsum[w] = x;
    __syncthreads();
    int off = 0;
    for (int u = 0; u < w; ++u) off += s_wsum[u];
    s_pre[tid + 1] = x + off;
    if (tid == 0) s_pre[0] = 0;
    const int n = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
    __syncthreads();
    const int nz = sp.nz[s];
    const int zq = (nz + NW - 1) / NW;
    const int z_lo = min(nz, w * zq), z_hi = min(nz, z_lo + zq);
    for (int P = blockIdx.x; P * 64 < n; P += gridDim.x) {   // uniform over the workgroup
        const int e = P * 64 + lane;
        const bool have = e < n;
        int c = 0;
        if (have) {
            int lo = 0, hi = K2_SCREEN_NSUB;   // the sub-list q with pre[q] <= e < pre[q + 1]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_pre[mid] <= e) lo = mid; else hi = mid;
            }
            c = list_in[(size_t)lo * cap + (e - s_pre[lo])];
        }
        const uint4 cv = cand[c];
        const Limbs4 v{cv.x, cv.y, cv.z, cv.w};
        const float4 vf = make_float4((float)v.x0, (float)v.x1, (float)v.x2, (float)v.x3);
        const uint32_t vp = pack_lo16(v.x2, v.x3);
        ShiftSums<M> acc;
        acc.zero();
        screen_zeros<M>(v, vp, vf, rec, z_lo, z_hi, s_tab, acc);
        if (w > 0) {
            s_part[w - 1][0][lane] = acc.a0;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                s_part[w - 1][1 + m][lane] = acc.A[m];
                s_part[w - 1][1 + M + m][lane] = acc.B[m];
            }
        }
        __syncthreads();
        if (w == 0) {
#pragma unroll
            for (int u = 0; u < NW - 1; ++u) {
                acc.a0 += s_part[u][0][lane];
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    acc.A[m] += s_part[u][1 + m][lane];
                    acc.B[m] += s_part[u][1 + M + m][lane];
                }
            }
            bool next = false;
            if (have) {
                int best;
                const float mv = stage_value<M>(acc, best);
                next = screen_outputs(c, s, n_stages, sp, mv, best, out_val, out_shift, out_pass);
            }
            screen_append(next, c, P % K2_SCREEN_NSUB, list_out, count_out, cap);
        }
        __syncthreads();   // wave 0 has read s_part before the next pass writes it
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
    if (e == hipSuccess) {
        hipLaunchKernelGGL((k2_screen_kernel<M>), dim3((n_cand + 255) / 256), dim3(256), 0, st,
                           (const ScreenRec<M>*)B.rec, cand, n_cand, n_stages, sp, B.list[0],
                           B.counts, B.cap, tab, val, shift, pass);
        e = hipGetLastError();
    }
    for (int s = 1; s < n_stages && e == hipSuccess; ++s) {
        hipLaunchKernelGGL((k2_screen_later_kernel<M>),
                           dim3(std::min(K2_SCREEN_GRID, (n_cand + 63) / 64)), dim3(256), 0, st,
                           (const ScreenRec<M>*)B.rec, cand, n_stages, sp, s,
                           B.list[(s - 1) & 1], B.counts + (s - 1) * cs, B.list[s & 1],
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
