// Internal declarations shared by the kernels (k2_kernels.hip) and the C ABI (k2_api.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/k2eval.h"
#include "k2_device.h"

namespace k2impl {

// per-k weighted zero arrays (device), length n_pad
struct WeightArrays {
    double k = 0.0;
    double2* a = nullptr;   // 2 e^{-k g^2} w
    double2* b = nullptr;   // g a
    double2* e = nullptr;   // g^2 a
    float2* af = nullptr;
    float2* bf = nullptr;
    float2* ef = nullptr;
    double* blk = nullptr;            // sum |a| per K2_ABS_BLOCK zeros (device)
    std::vector<double> abs_blocks;   // the same on the host
    uint64_t last_use = 0;
};

#define K2_ABS_BLOCK 4096   // zeros per weight-magnitude block (a multiple of the table padding)
#define K2_MIXED_TAIL 0.16  // default FP32-tail magnitude of the mixed form
#define K2_MIXED_STEP_TOL 1e-11  // Newton stop rule of the precision form's polish (its noise floor)
#define K2_Z_CZ 8           // candidates per Z-tile workgroup of the one-launch path
#define K2_Z_MAX_CAND 384   // default: candidate counts up to this use Z tiles (measured crossover)
#define K2_Z_CHUNK 8192     // fewest zeros per Z-tile workgroup on a full grid
#define K2_Z_MIN_CHUNK 2048 // fewest zeros per Z-tile workgroup for one or two candidate groups
#define K2_MIN_SPLIT 32     // fewest zeros per C-tile split (short splits for small problems)
#define K2_FILL_BLOCKS 2048 // a C-tile grid below this many workgroups gets shorter splits
#define K2_FUSED_MIXED_EVALS 1e9   // mixed-form problems up to this many evaluations: one launch
#define K2_SLOTS 1024       // C-tile workgroups resident at once on MI355X (256 CUs x 4)
// guided splits (guided_splits): long splits of this many short ones first, in the one-launch
// kernels (FP64, the mixed form's head and FP32 tail) and the enclosure; the multi-weight kernel
// (three workgroups per CU) measured 0.7% slower with them and keeps uniform splits
#ifndef K2_SPLIT_LONG
#define K2_SPLIT_LONG 8
#endif
#ifndef K2_SPLIT_LONG_MW
#define K2_SPLIT_LONG_MW 1
#endif

struct LaunchArgs {
    int nc = 4;
    int nt = 256;                          // threads per workgroup of the C-tile kernel
    unsigned flags = 0;
    int n_blocks = 0;
    const uint4* zg = nullptr;
    const double2* za = nullptr;
    const double2* zb = nullptr;
    const double2* ze = nullptr;
    const float2* zaf = nullptr;
    const float2* zbf = nullptr;
    const float2* zef = nullptr;
    const float4* zgf = nullptr;
    int z_begin = 0, z_end = 0, zps = 0;   // zero range and zeros per split of this launch
    int n_long = 0, zps_long = 0;          // the first n_long splits hold zps_long zeros each
    const uint4* cand = nullptr;
    int n_cand = 0, n_ctiles = 0;
    double* partial = nullptr;
    int row_base = 0, n_cand_pad = 0;      // first partial row written by this launch
    const k2dev::Tab64* tab64 = nullptr;   // sin/cos tables (k2_device.h)
    const k2dev::Tab32* tab32 = nullptr;
};

// host-side sin/cos tables, rounded from long double: K2_TAB FP64 entries at the interval
// midpoints 2 pi (k + 1/2) / K2_TAB (see sincos_tab64_at) followed by the split K2_TABW-point
// table (sincos_tabw_at) in t64, which holds K2_TAB64_ENTRIES, and K2_TAB32 FP32 entries of
// 2 pi k / K2_TAB32
// Tab64-sized entries of the FP64 table buffer: K2_TAB interleaved, then K2_TABW_DOUBLES
#define K2_TAB64_ENTRIES (K2_TAB + (K2_TABW_DOUBLES + 1) / 2)
void sincos_tables(k2dev::Tab64* t64, k2dev::Tab32* t32);
// device copies of those tables
int upload_tables(k2dev::Tab64** t64, k2dev::Tab32** t32, hipStream_t st);
// records the calling thread's last error (k2_last_error) and returns code
int set_error(int code, const std::string& msg);
int alloc_weights(WeightArrays& w, size_t n_pad);
void free_weights(WeightArrays& w);
// a, b, e (FP64 and FP32) for k, then the magnitude blocks, copied to w.abs_blocks; synchronises st
int compute_weights(const double* gamma, const double* wz, int n, int n_pad, double k,
                    WeightArrays& w, hipStream_t st);
// FP64 head of the mixed form over zeros [0, nz): fixed_head if > 0, else the shortest multiple of
// K2_ABS_BLOCK zeros after which the magnitude blocks up to nz (whole blocks) sum to <= tail
uint32_t mixed_head_zeros(const std::vector<double>& blocks, uint32_t nz, int fixed_head,
                          double tail);

}  // namespace k2impl

namespace k2impl {
struct ProfPair {
    hipEvent_t a, b;
};

// Orders the device work of successive calls on one handle (a table or a batch plan) whose
// scratch buffers they share: every call records `ev` on its stream when it has queued its
// work, and a call arriving on another stream first makes that stream wait for `ev`.  Calls on
// one stream are ordered by the stream itself.
struct StreamOrder {
    hipEvent_t ev = nullptr;
    hipStream_t last = nullptr;
    bool valid = false;
    hipError_t wait(hipStream_t st) {
        if (!valid || last == st) return hipSuccess;
        return hipStreamWaitEvent(st, ev, 0);
    }
    hipError_t mark(hipStream_t st) {
        hipError_t e = hipEventRecord(ev, st);
        if (e == hipSuccess) { last = st; valid = true; }
        return e;
    }
};

}  // namespace k2impl

struct k2_table {
    using WeightArrays = k2impl::WeightArrays;
    using ProfPair = k2impl::ProfPair;
    int device = 0;
    uint32_t n = 0, n_pad = 0;
    uint4* g = nullptr;
    float4* gf = nullptr;   // the limbs of g times 2 pi / 2^44 as floats (column 3 of the FP32 window)
    double* gamma = nullptr;
    double* w = nullptr;
    k2dev::Tab64* tab64 = nullptr;   // sin/cos tables (sincos_tables)
    k2dev::Tab32* tab32 = nullptr;
    std::vector<WeightArrays> wsets;
    uint64_t use_clock = 0;
    double* partial = nullptr;
    size_t partial_bytes = 0;
    void* host_io = nullptr;   // device staging for host-buffer entry points
    size_t host_io_bytes = 0;
    int* polish_active = nullptr;
    void* polish_buf = nullptr;          // compaction: indices, gathered points, their sums
    size_t polish_buf_bytes = 0;
    int64_t polish_evals = 0;            // candidate evaluations of the last polish
    void* screen_params = nullptr;       // k2_screen_device: zero records, survivor lists
    size_t screen_bytes = 0;
    size_t polish_cap = 0;
    // enclosure mode (k2_enclose_device), allocated on first use: weights with a stated error
    // bound, per-block bound sums, the final bound sums, split partials (hi, lo pairs)
    double2* enc_a = nullptr;
    double2* enc_b = nullptr;
    double* enc_part = nullptr;
    double* enc_bounds = nullptr;
    double* enc_partial = nullptr;
    size_t enc_partial_bytes = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;             // second stream for the FP64 head of the mixed form
    hipEvent_t fork = nullptr, join = nullptr;
    k2impl::StreamOrder order;             // orders calls on different streams (shared scratch)
    k2_tiling tiling{2, 3072, 0, K2_MIXED_TAIL, K2_Z_MAX_CAND, 256};
    bool profile = false;
    std::vector<ProfPair> prof_pending;
    std::vector<hipEvent_t> ev_pool;
    double prof_ms = 0.0;
    int64_t prof_launches = 0;
    std::mutex mu;
};


namespace k2impl {

// RAII form of StreamOrder for one entry point on table t and stream st: the constructor makes st
// wait for the previous call's work when that ran on another stream (check err()), the
// destructor records this call's end on st
struct CallOrder {
    k2_table* t;
    hipStream_t st;
    hipError_t e;
    CallOrder(k2_table* t_, hipStream_t st_) : t(t_), st(st_), e(t_->order.wait(st_)) {}
    ~CallOrder() { (void)t->order.mark(st); }
    hipError_t err() const { return e; }
};

hipError_t launch_eval(const LaunchArgs& a, bool mixed, hipStream_t st);
// mixed form in one launch: splits [0, sp_head) FP64 over [0, head_end) in zps_head steps, the
// rest FP32 over [head_end, a.z_end) in zps_tail steps (nc 1 or 2, 256 threads)
hipError_t launch_eval_mixed(const LaunchArgs& a, int head_end, int zps_head, int sp_head,
                             int zps_tail, hipStream_t st);
// zeros per split (a multiple of 256) and split count for nz zeros, n_ctiles candidate tiles
// and a target workgroup count
void split_range(int nz, int n_ctiles, int target, int* n_splits, int* zps);
// guided splits over nz zeros: n splits in all, the first n_long of zps_long zeros (a multiple
// of zps), the rest of zps zeros (the last one shorter), zps as split_range gives it
struct SplitPlan {
    int n = 0, zps = 0, n_long = 0, zps_long = 0;
};
void guided_splits(int nz, int n_ctiles, int target, int factor, SplitPlan* sp);
// workgroups of a C-tile grid over sp splits (guided splits pad the split count to a multiple
// of 8: the kernels deal splits to the XCDs in turn)
inline int ctile_blocks(int n_ctiles, int sp, int n_long) {
    return n_ctiles * (n_long > 0 ? (sp + 7) / 8 * 8 : sp);
}
// releases the enclosure-mode buffers of a table
void free_enclose(k2_table* t);
hipError_t launch_weights(const double* gamma, const double* w, int n, int n_pad, double k,
                          WeightArrays& out, hipStream_t st);
hipError_t launch_abs_blocks(const double2* a, int n_pad, double* blk, hipStream_t st);
hipError_t launch_gather(const uint4* V, const int* idx, int n, uint4* Vc, hipStream_t st);
hipError_t launch_scatter3(const int* idx, int n, const double* Kc, const double* dKc,
                           const double* d2Kc, double* K, double* dK, double* d2K,
                           hipStream_t st);
// Z-tile one-launch path: workgroup (chunk, group) sums zeros [z_begin + chunk*zpc, ...) for
// candidates [group*K2_Z_CZ, ...) into partial row (row_base + chunk) * n_groups + group
hipError_t launch_eval_z(const LaunchArgs& a, bool f32, int n_groups, int n_chunks, int zpc,
                         hipStream_t st);
hipError_t launch_reduce_z(const double* partial, int n_rows, int n_groups, int n_cand,
                           unsigned flags, double* K, double* dK, double* d2K, hipStream_t st);
hipError_t launch_reduce(const double* partial, int n_splits, int n_cand, int n_cand_pad,
                         unsigned flags, double* K, double* dK, double* d2K, hipStream_t st);
hipError_t launch_phase_words(const uint4* zg, const float4* zgf, int z0, int nz,
                              const uint4* cand, int n_cand, int mode, uint64_t* out,
                              hipStream_t st);
hipError_t launch_sincos_words(const uint64_t* W, int n, int mode, const k2dev::Tab64* tab64,
                               const k2dev::Tab32* tab32, double* out, hipStream_t st);
hipError_t launch_newton(uint4* cand, const double* dK, const double* d2K, int n, double clamp,
                         double tol, int* active, int* steps, hipStream_t st);
hipError_t launch_ref_peak_ab(int n_peaks, const uint64_t* const* p, const double* dy0,
                              int n_zeros, const double* alpha, const double* gamma, double* A,
                              double* B, hipStream_t st);
hipError_t launch_ref_polish(int n_peaks, int n_zeros, const double* A, const double* B,
                             const double* gamma, const double* w, double half, int steps,
                             double* h_unpol, double* h_pol, double* ddy, hipStream_t st);

}  // namespace k2impl
