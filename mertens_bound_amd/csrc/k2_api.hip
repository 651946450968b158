// C ABI of libk2eval.so (declared in include/k2eval.h): table management, launch planning
// (candidate tiles x zero splits), per-k weight caching, profiling events and the
// reference-signature entry points.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "k2_device.h"
#include "k2_internal.h"

using namespace k2impl;

#define K2_VERSION 1
#define K2_PAD 4096            // zero arrays are padded to a multiple of this
#define K2_MAX_WEIGHT_SETS 4   // cached per-k weight arrays per table

static int fail(int code, const std::string& msg);

void k2impl::sincos_tables(k2dev::Tab64* t64, k2dev::Tab32* t32) {
    const long double two_pi = 6.283185307179586476925286766559005768L;
    // interval midpoints, shifted by the rounding of the remainder's offset constant:
    // (2^52 + 2^51) K - M is exact in long double (both are doubles of about the same size)
    const long double shift = 6755399441055744.0L * (long double)K2_REM_K - (long double)K2_REM_M;
    for (int k = 0; k < K2_TAB; ++k) {
        const long double a = two_pi * ((long double)k + 0.5L) / (long double)K2_TAB - shift;
        t64[k] = k2dev::Tab64{(double)cosl(a), (double)sinl(a)};
    }
    // the split 2,048-point table after them: K2_TABW cosines, a zero pad, K2_TABW sines
    double* tw = reinterpret_cast<double*>(t64 + K2_TAB);
    tw[K2_TABW] = 0.0;
    const long double shiftw =
        6755399441055744.0L * (long double)K2_REMW_K - (long double)K2_REMW_M;
    for (int k = 0; k < K2_TABW; ++k) {
        const long double a = two_pi * ((long double)k + 0.5L) / (long double)K2_TABW - shiftw;
        tw[k] = (double)cosl(a);
        tw[K2_TABW_SIN + k] = (double)sinl(a);
    }
    for (int k = 0; k < K2_TAB32; ++k) {
        const long double a = two_pi * (long double)k / (long double)K2_TAB32;
        t32[k] = k2dev::Tab32{(float)cosl(a), (float)sinl(a)};
    }
}

int k2impl::upload_tables(k2dev::Tab64** t64, k2dev::Tab32** t32, hipStream_t st) {
    std::vector<k2dev::Tab64> h64(K2_TAB64_ENTRIES);
    std::vector<k2dev::Tab32> h32(K2_TAB32);
    sincos_tables(h64.data(), h32.data());
    const size_t b64 = K2_TAB64_ENTRIES * sizeof(k2dev::Tab64);
    const size_t b32 = K2_TAB32 * sizeof(k2dev::Tab32);
    hipError_t e = hipMalloc((void**)t64, b64);
    if (e == hipSuccess) e = hipMalloc((void**)t32, b32);
    if (e == hipSuccess) e = hipMemcpyAsync(*t64, h64.data(), b64, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(*t32, h32.data(), b32, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? K2_OK : fail(K2_E_HIP, std::string("sin/cos table: ") + hipGetErrorString(e));
}

static thread_local std::string g_err;

int k2impl::set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

static int fail(int code, const std::string& msg) { return set_error(code, msg); }

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(K2_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

extern "C" const char* k2_last_error(void) { return g_err.c_str(); }
extern "C" int k2_version(void) { return K2_VERSION; }

// ---- static scratch of the allocation-free entry point (k2_eval_device_noalloc) ---------------
// Each GPU the library runs on holds one copy of this array until the process ends.  In host code
// `scratch` names a host-side stand-in, not the device array: the device address comes from
// hipGetSymbolAddress, taken once per device (with that device current) when the first table
// is created on it, and only that address is passed to kernels.
#define K2_SCRATCH_BYTES ((size_t)64 << 20)
#define K2_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[K2_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[K2_MAX_DEVICES];

static int scratch_init(int dev) {
    if (dev < 0 || dev >= K2_MAX_DEVICES) return fail(K2_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return K2_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(K2_E_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
    return K2_OK;
}

static void* scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (dev >= 0 && dev < K2_MAX_DEVICES) ? g_scratch_ptr[dev] : nullptr;
}

static int ensure_bytes(void** p, size_t* cap, size_t need) {
    if (*cap >= need) return K2_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t want = std::max(need, (size_t)1 << 20);
    hipError_t e = hipMalloc(p, want);
    if (e != hipSuccess) return fail(K2_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    *cap = want;
    return K2_OK;
}

extern "C" int k2_table_create(k2_table** out, uint32_t n_zeros, const uint64_t* G,
                               const double* gamma, const double* w) {
    if (!out || !G || !gamma || !w || n_zeros == 0 || n_zeros > 0x7fff0000u)
        return fail(K2_E_ARG, "k2_table_create: null pointer or n_zeros out of range");
    k2_table* t = new k2_table();
    HIPCHK(hipGetDevice(&t->device));
    t->n = n_zeros;
    t->n_pad = (n_zeros + K2_PAD - 1) / K2_PAD * K2_PAD;
    const size_t np = t->n_pad;
    hipError_t e = hipMalloc((void**)&t->g, np * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc((void**)&t->gf, np * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc((void**)&t->gamma, np * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&t->w, np * 2 * sizeof(double));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemsetAsync(t->g, 0, np * sizeof(uint4), t->stream);
    if (e == hipSuccess) e = hipMemsetAsync(t->gamma, 0, np * sizeof(double), t->stream);
    if (e == hipSuccess) e = hipMemsetAsync(t->w, 0, np * 2 * sizeof(double), t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t->g, G, (size_t)n_zeros * 16, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t->gamma, gamma, (size_t)n_zeros * 8, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t->w, w, (size_t)n_zeros * 16, hipMemcpyHostToDevice, t->stream);
    // the limbs times 2 pi / 2^44 as floats (the FP32 window's column 3: with the candidate's
    // limbs times 2^-34 the products are in radians), each rounded once from double
    std::vector<float4> gf(np, make_float4(0.f, 0.f, 0.f, 0.f));
    const double gscale = 6.283185307179586 / 17592186044416.0;   // 2 pi / 2^44
    for (uint32_t j = 0; j < n_zeros; ++j) {
        const uint64_t lo = G[2 * j], hi = G[2 * j + 1];
        gf[j] = make_float4((float)((double)(uint32_t)lo * gscale),
                            (float)((double)(uint32_t)(lo >> 32) * gscale),
                            (float)((double)(uint32_t)hi * gscale),
                            (float)((double)(uint32_t)(hi >> 32) * gscale));
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(t->gf, gf.data(), np * sizeof(float4), hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (e != hipSuccess) {
        k2_table_destroy(t);
        return fail(K2_E_HIP, std::string("k2_table_create: ") + hipGetErrorString(e));
    }
    if (upload_tables(&t->tab64, &t->tab32, t->stream) != K2_OK) {
        std::string msg = g_err;
        k2_table_destroy(t);
        return fail(K2_E_HIP, msg);
    }
    // the mixed form's second stream and its fork/join events, and the call-ordering event,
    // exist from here on (no entry point creates them later)
    e = hipStreamCreateWithFlags(&t->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&t->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&t->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&t->order.ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        k2_table_destroy(t);
        return fail(K2_E_HIP, std::string("k2_table_create: ") + hipGetErrorString(e));
    }
    int rc = scratch_init(t->device);
    if (rc) {
        std::string msg = g_err;
        k2_table_destroy(t);
        return fail(rc, msg);
    }
    *out = t;
    return K2_OK;
}

extern "C" int k2_table_destroy(k2_table* t) {
    if (!t) return K2_OK;
    (void)hipSetDevice(t->device);
    if (t->order.valid) (void)hipEventSynchronize(t->order.ev);   // the last call's work
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    for (auto& ws : t->wsets) free_weights(ws);
    for (auto& pp : t->prof_pending) { (void)hipEventDestroy(pp.a); (void)hipEventDestroy(pp.b); }
    for (auto ev : t->ev_pool) (void)hipEventDestroy(ev);
    (void)hipFree(t->g); (void)hipFree(t->gf); (void)hipFree(t->gamma); (void)hipFree(t->w);
    (void)hipFree(t->tab64); (void)hipFree(t->tab32);
    (void)hipFree(t->partial); (void)hipFree(t->host_io); (void)hipFree(t->polish_active);
    (void)hipFree(t->polish_buf);
    (void)hipFree(t->screen_params);
    free_enclose(t);
    if (t->aux) (void)hipStreamSynchronize(t->aux);
    if (t->aux) (void)hipStreamDestroy(t->aux);
    if (t->fork) (void)hipEventDestroy(t->fork);
    if (t->join) (void)hipEventDestroy(t->join);
    if (t->order.ev) (void)hipEventDestroy(t->order.ev);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
    return K2_OK;
}

extern "C" uint32_t k2_table_size(const k2_table* t) { return t ? t->n : 0; }

extern "C" int k2_set_tiling(k2_table* t, const k2_tiling* ti) {
    if (!t || !ti) return fail(K2_E_ARG, "k2_set_tiling: null pointer");
    k2_tiling v = *ti;
    if (v.nc == 0) v.nc = 2;
    if (v.target_blocks == 0) v.target_blocks = 3072;
    if (v.mixed_tail == 0.0) v.mixed_tail = K2_MIXED_TAIL;
    if (v.z_max_cand == 0) v.z_max_cand = K2_Z_MAX_CAND;
    if (v.z_max_cand < -1) return fail(K2_E_ARG, "k2_set_tiling: z_max_cand must be >= -1");
    if (v.wg_threads == 0) v.wg_threads = 256;
    if (v.wg_threads != 256 && !(v.wg_threads == 512 && v.nc <= 2))
        return fail(K2_E_ARG, "k2_set_tiling: wg_threads must be 256, or 512 with nc 1 or 2");
    if (v.nc != 1 && v.nc != 2 && v.nc != 4 && v.nc != 8)
        return fail(K2_E_ARG, "k2_set_tiling: nc must be 1, 2, 4 or 8");
    if (v.target_blocks < 1 || v.mixed_head < 0 || !(v.mixed_tail > 0.0))
        return fail(K2_E_ARG,
                    "k2_set_tiling: target_blocks >= 1, mixed_head >= 0 and mixed_tail > 0 required");
    std::lock_guard<std::mutex> lk(t->mu);
    t->tiling = v;
    return K2_OK;
}

extern "C" int k2_get_tiling(const k2_table* t, k2_tiling* ti) {
    if (!t || !ti) return fail(K2_E_ARG, "k2_get_tiling: null pointer");
    *ti = t->tiling;
    return K2_OK;
}

int k2impl::alloc_weights(WeightArrays& w, size_t np) {
    hipError_t e = hipMalloc((void**)&w.a, np * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc((void**)&w.blk, (np / K2_ABS_BLOCK) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&w.b, np * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc((void**)&w.e, np * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc((void**)&w.af, np * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc((void**)&w.bf, np * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc((void**)&w.ef, np * sizeof(float2));
    if (e != hipSuccess) {
        free_weights(w);
        return fail(K2_E_NOMEM, "weight arrays: hipMalloc failed");
    }
    return K2_OK;
}

void k2impl::free_weights(WeightArrays& w) {
    (void)hipFree(w.a); (void)hipFree(w.b); (void)hipFree(w.e);
    (void)hipFree(w.af); (void)hipFree(w.bf); (void)hipFree(w.ef); (void)hipFree(w.blk);
    w.a = w.b = w.e = nullptr;
    w.af = w.bf = w.ef = nullptr;
    w.blk = nullptr;
    w.abs_blocks.clear();
}

int k2impl::compute_weights(const double* gamma, const double* wz, int n, int n_pad, double k,
                            WeightArrays& w, hipStream_t st) {
    hipError_t e = launch_weights(gamma, wz, n, n_pad, k, w, st);
    if (e == hipSuccess) e = launch_abs_blocks(w.a, n_pad, w.blk, st);
    w.abs_blocks.assign(n_pad / K2_ABS_BLOCK, 0.0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(w.abs_blocks.data(), w.blk, w.abs_blocks.size() * sizeof(double),
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(K2_E_HIP, std::string("weight arrays: ") + hipGetErrorString(e));
    w.k = k;
    return K2_OK;
}

uint32_t k2impl::mixed_head_zeros(const std::vector<double>& blocks, uint32_t nz, int fixed_head,
                                  double tail) {
    if (fixed_head > 0) return std::min<uint32_t>(nz, (uint32_t)fixed_head);
    const uint32_t nb = std::min<uint32_t>((uint32_t)blocks.size(),
                                           (nz + K2_ABS_BLOCK - 1) / K2_ABS_BLOCK);
    double sum = 0.0;
    uint32_t hb = nb;   // head = hb whole blocks
    while (hb > 0 && sum + blocks[hb - 1] <= tail) {
        sum += blocks[hb - 1];
        --hb;
    }
    return std::min<uint32_t>(nz, hb * K2_ABS_BLOCK);
}

// weight arrays for k, computed on stream st on first use and waited for, so that work on any
// stream may read them afterwards (LRU of K2_MAX_WEIGHT_SETS)
static int get_weights(k2_table* t, double k, hipStream_t st, WeightArrays** out) {
    t->use_clock++;
    for (auto& ws : t->wsets)
        if (ws.k == k) { ws.last_use = t->use_clock; *out = &ws; return K2_OK; }
    WeightArrays* slot = nullptr;
    if ((int)t->wsets.size() < K2_MAX_WEIGHT_SETS) {
        WeightArrays w;
        int rc = alloc_weights(w, t->n_pad);
        if (rc) return rc;
        t->wsets.push_back(w);
        slot = &t->wsets.back();
    } else {
        slot = &*std::min_element(t->wsets.begin(), t->wsets.end(),
                                  [](const WeightArrays& x, const WeightArrays& y) {
                                      return x.last_use < y.last_use;
                                  });
    }
    slot->k = k;
    slot->last_use = t->use_clock;
    int rc = compute_weights(t->gamma, t->w, (int)t->n, (int)t->n_pad, k, *slot, st);
    if (rc) {
        slot->k = -1.0;   // never matches a valid k: recomputed on next use
        return rc;
    }
    *out = slot;
    return K2_OK;
}

static hipEvent_t pool_event(k2_table* t) {
    if (!t->ev_pool.empty()) {
        hipEvent_t e = t->ev_pool.back();
        t->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct Plan {
    int n_ctiles = 0, n_cand_pad = 0;
    int sp_head = 0, zps_head = 0;   // FP64 part
    int sp_tail = 0, zps_tail = 0;   // FP32 part (mixed only)
    int nl_head = 0, zl_head = 0;    // long splits leading the FP64 part (guided_splits)
    int nl_tail = 0, zl_tail = 0;    // and the FP32 part
    int head_end = 0;
};

void k2impl::split_range(int nz, int n_ctiles, int target, int* n_splits, int* zps) {
    if (nz <= 0) { *n_splits = 0; *zps = 0; return; }
    // splits of whole 256-zero LDS tiles, about target workgroups in all; when that leaves the
    // grid under K2_FILL_BLOCKS workgroups (few zeros), shorter splits (down to K2_MIN_SPLIT
    // zeros) fill the GPU instead
    int sp = std::max(1, std::min((target + n_ctiles - 1) / n_ctiles, (nz + 255) / 256));
    if ((int64_t)sp * n_ctiles < K2_FILL_BLOCKS)
        sp = std::max(sp, std::min((K2_FILL_BLOCKS + n_ctiles - 1) / n_ctiles,
                                   (nz + K2_MIN_SPLIT - 1) / K2_MIN_SPLIT));
    int z = (nz + sp - 1) / sp;
    z = z >= 256 ? (z + 255) / 256 * 256 : (z + K2_MIN_SPLIT - 1) / K2_MIN_SPLIT * K2_MIN_SPLIT;
    *zps = z;
    *n_splits = (nz + z - 1) / z;
}

// Guided splits: with short splits only, the grid that ends the launch evenly has many of them,
// and each split writes a partial row per candidate (1,515 splits of 1,536 zeros, 1.59 GB of rows
// for 65,536 candidates over the full table).  Long splits (factor times the short ones) first and
// short ones last keep the even end with a fifth of the rows: the last factor rounds of the
// resident workgroups (K2_SLOTS) run short splits, enough to fill every slot while the last long
// split finishes.  The kernels deal the splits to the 8 XCDs in turn (block_split), each XCD
// walking its own in order, and the long splits come in multiples of 8, so every XCD gets the
// same share of long splits and at most one short split more than another.  The splits stay
// contiguous in zero order and the reduction adds them in that order: a plan's results are the
// same on every run (another plan's differ by rounding only).
void k2impl::guided_splits(int nz, int n_ctiles, int target, int factor, SplitPlan* s) {
    *s = SplitPlan{};
    split_range(nz, n_ctiles, target, &s->n, &s->zps);
    if (factor <= 1 || s->zps < 256 || s->n == 0) return;
    const int64_t n_short = ((int64_t)factor * K2_SLOTS + n_ctiles - 1) / n_ctiles;
    const int64_t zl = (int64_t)factor * s->zps;
    const int64_t long_zeros = (int64_t)nz - n_short * s->zps;
    if (long_zeros < 8 * zl) return;
    // a multiple of 8 long splits, so that every XCD gets the same number of them
    const int nl = (int)(long_zeros / zl) / 8 * 8;
    const int rest = nz - nl * (int)zl;
    s->n_long = nl;
    s->zps_long = (int)zl;
    s->n = nl + (rest + s->zps - 1) / s->zps;
}


static Plan make_plan(const k2_table* t, const WeightArrays* ws, int n_cand, int nz, bool mixed) {
    Plan p;
    const int cw = t->tiling.wg_threads * t->tiling.nc;   // candidates per C tile
    p.n_ctiles = (n_cand + cw - 1) / cw;
    p.n_cand_pad = p.n_ctiles * cw;
    p.head_end = mixed ? (int)mixed_head_zeros(ws->abs_blocks, (uint32_t)nz, t->tiling.mixed_head,
                                               t->tiling.mixed_tail)
                       : nz;
    const int tb = t->tiling.target_blocks;
    if (mixed && p.head_end < nz) {
        // the FP64 head runs beside the FP32 tail on a second stream; give it workgroups in
        // proportion to its work (an FP64 zero costs about 2.1 FP32 zeros; the time hardly
        // depends on this share, since the two launches share the CUs), at least one split
        const int nt = nz - p.head_end;
        const double wh = 2.1 * p.head_end, wt = (double)nt;
        const int tb_head = std::max(p.n_ctiles, (int)std::lround(tb * wh / (wh + wt)));
        SplitPlan h, tl;
        guided_splits(p.head_end, p.n_ctiles, tb_head, K2_SPLIT_LONG, &h);
        guided_splits(nt, p.n_ctiles, std::max(1, tb - tb_head), K2_SPLIT_LONG, &tl);
        p.sp_head = h.n; p.zps_head = h.zps; p.nl_head = h.n_long; p.zl_head = h.zps_long;
        p.sp_tail = tl.n; p.zps_tail = tl.zps; p.nl_tail = tl.n_long; p.zl_tail = tl.zps_long;
    } else {
        SplitPlan h;
        guided_splits(nz, p.n_ctiles, tb, K2_SPLIT_LONG, &h);
        p.sp_head = h.n; p.zps_head = h.zps; p.nl_head = h.n_long; p.zl_head = h.zps_long;
        p.head_end = nz;
    }
    return p;
}

// zero chunks of the Z-tile path: about target / n_groups chunks of at least K2_Z_CHUNK zeros;
// when that leaves the grid under K2_FILL_BLOCKS workgroups (one or two candidate groups),
// shorter chunks (down to K2_Z_MIN_CHUNK zeros) fill the GPU instead
static void z_chunks(int len, int n_groups, int target, int* n_ch, int* zpc) {
    if (len <= 0) { *n_ch = 0; *zpc = 0; return; }
    const int want = std::max(1, target / n_groups);
    int z = std::max(K2_Z_CHUNK, (len + want - 1) / want);
    if ((int64_t)((len + z - 1) / z) * n_groups < K2_FILL_BLOCKS) {
        const int fill = (K2_FILL_BLOCKS + n_groups - 1) / n_groups;
        z = std::max(K2_Z_MIN_CHUNK, std::min(z, (len + fill - 1) / fill));
    }
    z = (z + 255) / 256 * 256;
    *zpc = z;
    *n_ch = (len + z - 1) / z;
}

static void fill_args(k2_table* t, WeightArrays* ws, const uint4* cand, int n_cand,
                      unsigned flags, LaunchArgs& a) {
    a.nc = t->tiling.nc;
    a.flags = flags;
    a.zg = t->g;
    a.zgf = t->gf;
    a.za = ws->a; a.zb = ws->b; a.ze = ws->e;
    a.zaf = ws->af; a.zbf = ws->bf; a.zef = ws->ef;
    a.cand = cand;
    a.n_cand = n_cand;
    a.partial = t->partial;
    a.tab64 = t->tab64;
    a.tab32 = t->tab32;
}

// Where the split partial sums go: the table's buffer, grown on demand (grow), or the static
// scratch array of the allocation-free entry point, of fixed size cap (the plan is then cut to
// fit it, and calls it cannot fit are rejected).
struct PartialBuf {
    double* p = nullptr;
    size_t cap = 0;
    bool grow = true;
};

static int partial_for(k2_table* t, PartialBuf& pb, size_t need, const char* who) {
    if (pb.grow) {
        int rc = ensure_bytes((void**)&t->partial, &t->partial_bytes, need);
        pb.p = t->partial;
        return rc;
    }
    if (need > pb.cap)
        return fail(K2_E_ARG, std::string(who) + ": the partial sums need " + std::to_string(need) +
                                  " bytes, more than the " + std::to_string(pb.cap) +
                                  "-byte scratch");
    return K2_OK;
}

// splits of at most max_rows rows over nz zeros (whole 256-zero tiles where the splits are that
// long, else multiples of K2_MIN_SPLIT)
static void fit_splits(int nz, int max_rows, int* n_splits, int* zps) {
    if (nz <= 0) { *n_splits = 0; *zps = 0; return; }
    int z = (nz + max_rows - 1) / max_rows;
    z = z >= 256 ? (z + 255) / 256 * 256 : (z + K2_MIN_SPLIT - 1) / K2_MIN_SPLIT * K2_MIN_SPLIT;
    *zps = z;
    *n_splits = (nz + z - 1) / z;
}

// few candidates: Z tiles (lanes over zeros, K2_Z_CZ candidates per workgroup), FP64 head chunks
// then FP32 tail chunks in the mixed form, one wavefront per candidate in the reduction
static int eval_core_z(k2_table* t, int n_cand, const uint4* cand, WeightArrays* ws, int nz,
                       unsigned flags, double* K, double* dK, double* d2K, hipStream_t st,
                       PartialBuf& pb) {
    // one candidate group runs in FP64 throughout: there the FP32 tail's second launch costs
    // more than it saves (1 candidate x 2,325,572 zeros: 0.033 ms in FP64, 0.055 ms mixed)
    const bool mixed = (flags & K2_MIXED) != 0 && n_cand > K2_Z_CZ;
    const int head = mixed ? (int)mixed_head_zeros(ws->abs_blocks, (uint32_t)nz,
                                                   t->tiling.mixed_head, t->tiling.mixed_tail)
                           : nz;
    const int n_groups = (n_cand + K2_Z_CZ - 1) / K2_Z_CZ;
    int ch_h, zpc_h, ch_t, zpc_t;
    z_chunks(head, n_groups, t->tiling.target_blocks, &ch_h, &zpc_h);
    z_chunks(nz - head, n_groups, t->tiling.target_blocks, &ch_t, &zpc_t);
    const int n_rows = std::max(1, ch_h + ch_t);
    const size_t row = (size_t)n_groups * 3 * K2_Z_CZ;
    int rc = partial_for(t, pb, n_rows * row * sizeof(double), "Z-tile evaluation");
    if (rc) return rc;
    LaunchArgs a;
    fill_args(t, ws, cand, n_cand, flags, a);
    a.partial = pb.p;
    const bool prof = t->profile && pb.grow;
    ProfPair pp{nullptr, nullptr};
    if (prof) {
        pp.a = pool_event(t);
        pp.b = pool_event(t);
        HIPCHK(hipEventRecord(pp.a, st));
    }
    if (ch_h > 0) {
        a.z_begin = 0; a.z_end = head; a.row_base = 0;
        HIPCHK(launch_eval_z(a, false, n_groups, ch_h, zpc_h, st));
    }
    if (ch_t > 0) {
        a.z_begin = head; a.z_end = nz; a.row_base = ch_h;
        HIPCHK(launch_eval_z(a, true, n_groups, ch_t, zpc_t, st));
    }
    if (prof) {
        HIPCHK(hipEventRecord(pp.b, st));
        t->prof_pending.push_back(pp);
    }
    if (ch_h + ch_t == 0) return fail(K2_E_ARG, "Z-tile evaluation: empty zero range");
    HIPCHK(launch_reduce_z(pb.p, n_rows, n_groups, n_cand, flags, K, dK, d2K, st));
    return K2_OK;
}

// evaluation of one candidate list over zeros [0, nz) (nz >= 1) with weights ws; outputs on
// device; partial sums in pb
static int eval_core(k2_table* t, int n_cand, const uint4* cand, WeightArrays* ws, int nz,
                     unsigned flags, double* K, double* dK, double* d2K, hipStream_t st,
                     PartialBuf& pb) {
    const int zmax = t->tiling.z_max_cand == 0 ? K2_Z_MAX_CAND : t->tiling.z_max_cand;
    if (n_cand > 0 && n_cand <= zmax)
        return eval_core_z(t, n_cand, cand, ws, nz, flags, K, dK, d2K, st, pb);
    if (nz <= 0) return fail(K2_E_ARG, "evaluation: empty zero range");
    const bool mixed = (flags & K2_MIXED) != 0;
    Plan p = make_plan(t, ws, n_cand, nz, mixed);
    const size_t row_bytes = (size_t)3 * p.n_cand_pad * sizeof(double);
    if (!pb.grow) {
        // fixed scratch: fewer, longer splits where the plan's rows do not fit
        const int64_t max_rows = (int64_t)(pb.cap / row_bytes);
        const int need = (p.sp_head > 0) + (p.sp_tail > 0);
        if (max_rows < need)
            return fail(K2_E_ARG, "evaluation: " + std::to_string(n_cand) + " candidates need " +
                                      std::to_string(need * row_bytes) + " bytes of partial sums, "
                                      "more than the " + std::to_string(pb.cap) + "-byte scratch");
        if (p.sp_head + p.sp_tail > max_rows) {
            p.nl_head = p.zl_head = p.nl_tail = p.zl_tail = 0;
            if (p.sp_tail > 0 && p.sp_head > 0) {
                int rh = (int)std::max<int64_t>(1, max_rows * p.sp_head / (p.sp_head + p.sp_tail));
                rh = (int)std::min<int64_t>(rh, max_rows - 1);
                fit_splits(p.head_end, rh, &p.sp_head, &p.zps_head);
                fit_splits(nz - p.head_end, (int)(max_rows - p.sp_head), &p.sp_tail, &p.zps_tail);
            } else if (p.sp_tail > 0) {
                fit_splits(nz - p.head_end, (int)max_rows, &p.sp_tail, &p.zps_tail);
            } else {
                fit_splits(p.head_end, (int)max_rows, &p.sp_head, &p.zps_head);
            }
        }
    }
    const int n_rows = std::max(1, p.sp_head + p.sp_tail);
    int rc = partial_for(t, pb, (size_t)n_rows * row_bytes, "evaluation");
    if (rc) return rc;
    LaunchArgs a;
    a.nc = t->tiling.nc;
    a.nt = t->tiling.wg_threads;
    a.flags = flags;
    a.zg = t->g;
    a.zgf = t->gf;
    a.za = ws->a; a.zb = ws->b; a.ze = ws->e;
    a.zaf = ws->af; a.zbf = ws->bf; a.zef = ws->ef;
    a.cand = cand;
    a.n_cand = n_cand;
    a.n_ctiles = p.n_ctiles;
    a.partial = pb.p;
    a.n_cand_pad = p.n_cand_pad;
    a.tab64 = t->tab64;
    a.tab32 = t->tab32;
    const bool prof = t->profile && pb.grow;
    ProfPair pp{nullptr, nullptr};
    if (prof) {
        pp.a = pool_event(t);
        pp.b = pool_event(t);
        HIPCHK(hipEventRecord(pp.a, st));
    }
    const bool fused = p.sp_tail > 0 && p.sp_head > 0 && a.nt == 256 && a.nc <= 2 &&
                       (double)n_cand * nz <= K2_FUSED_MIXED_EVALS && p.nl_head == 0 &&
                       p.nl_tail == 0;
    if (fused) {
        // small mixed problem: head and tail splits in one launch on st
        a.z_begin = 0; a.z_end = nz; a.row_base = 0;
        a.n_blocks = p.n_ctiles * (p.sp_head + p.sp_tail);
        HIPCHK(launch_eval_mixed(a, p.head_end, p.zps_head, p.sp_head, p.zps_tail, st));
    } else if (p.sp_tail > 0 && p.sp_head > 0) {
        // mixed form: FP64 head on the auxiliary stream, FP32 tail on st, joined before reducing
        HIPCHK(hipEventRecord(t->fork, st));
        HIPCHK(hipStreamWaitEvent(t->aux, t->fork, 0));
        a.z_begin = 0; a.z_end = p.head_end; a.zps = p.zps_head; a.row_base = 0;
        a.n_long = p.nl_head; a.zps_long = p.zl_head;
        a.n_blocks = ctile_blocks(p.n_ctiles, p.sp_head, p.nl_head);
        HIPCHK(launch_eval(a, false, t->aux));
        a.z_begin = p.head_end; a.z_end = nz; a.zps = p.zps_tail; a.row_base = p.sp_head;
        a.n_long = p.nl_tail; a.zps_long = p.zl_tail;
        a.n_blocks = ctile_blocks(p.n_ctiles, p.sp_tail, p.nl_tail);
        HIPCHK(launch_eval(a, true, st));
        HIPCHK(hipEventRecord(t->join, t->aux));
        HIPCHK(hipStreamWaitEvent(st, t->join, 0));
    } else if (p.sp_tail > 0) {
        a.z_begin = p.head_end; a.z_end = nz; a.zps = p.zps_tail; a.row_base = 0;
        a.n_long = p.nl_tail; a.zps_long = p.zl_tail;
        a.n_blocks = ctile_blocks(p.n_ctiles, p.sp_tail, p.nl_tail);
        HIPCHK(launch_eval(a, true, st));
    } else if (p.sp_head > 0) {
        a.z_begin = 0; a.z_end = p.head_end; a.zps = p.zps_head; a.row_base = 0;
        a.n_long = p.nl_head; a.zps_long = p.zl_head;
        a.n_blocks = ctile_blocks(p.n_ctiles, p.sp_head, p.nl_head);
        HIPCHK(launch_eval(a, false, st));
    }
    if (prof) {
        HIPCHK(hipEventRecord(pp.b, st));
        t->prof_pending.push_back(pp);
    }
    HIPCHK(launch_reduce(pb.p, n_rows, n_cand, p.n_cand_pad, flags, K, dK, d2K, st));
    return K2_OK;
}

static int check_eval_args(k2_table* t, uint32_t n_cand, const void* V, uint32_t* n_zeros,
                           const double* K, const char* who) {
    if (!t) return fail(K2_E_ARG, std::string(who) + ": null table");
    if (n_cand > 0 && (!V || !K)) return fail(K2_E_ARG, std::string(who) + ": null buffer");
    if (n_cand > 0x40000000u) return fail(K2_E_ARG, std::string(who) + ": n_cand too large");
    if (*n_zeros == 0) *n_zeros = t->n;
    if (*n_zeros > t->n) return fail(K2_E_ARG, std::string(who) + ": n_zeros exceeds table size");
    return K2_OK;
}

extern "C" int k2_eval_device(k2_table* t, uint32_t n_cand, const uint64_t* V_dev, double k,
                              uint32_t n_zeros, uint32_t flags, double* K_dev, double* dK_dev,
                              double* d2K_dev, void* stream) {
    int rc = check_eval_args(t, n_cand, V_dev, &n_zeros, K_dev, "k2_eval_device");
    if (rc) return rc;
    if (((uintptr_t)V_dev & 15) != 0) return fail(K2_E_ARG, "k2_eval_device: V not 16-byte aligned");
    if (((flags & K2_D1) && !dK_dev) || ((flags & K2_D2) && !d2K_dev))
        return fail(K2_E_ARG, "k2_eval_device: derivative requested without an output buffer");
    if (!(k >= 0.0) || !isfinite(k)) return fail(K2_E_ARG, "k2_eval_device: k must be finite and >= 0");
    if (n_cand == 0) return K2_OK;
    std::lock_guard<std::mutex> lk(t->mu);
    HIPCHK(hipSetDevice(t->device));
    hipStream_t st = (hipStream_t)stream;   // NULL is the default stream
    CallOrder order(t, st);
    HIPCHK(order.err());
    WeightArrays* ws = nullptr;
    rc = get_weights(t, k, st, &ws);
    if (rc) return rc;
    PartialBuf pb;
    return eval_core(t, (int)n_cand, (const uint4*)V_dev, ws, (int)n_zeros, flags, K_dev, dK_dev,
                     d2K_dev, st, pb);
}

extern "C" int k2_eval_device_noalloc(k2_table* t, uint32_t n_cand, const uint64_t* V_dev,
                                      double k, uint32_t n_zeros, uint32_t flags, double* K_dev,
                                      double* dK_dev, double* d2K_dev, void* stream) {
    int rc = check_eval_args(t, n_cand, V_dev, &n_zeros, K_dev, "k2_eval_device_noalloc");
    if (rc) return rc;
    if (((uintptr_t)V_dev & 15) != 0)
        return fail(K2_E_ARG, "k2_eval_device_noalloc: V not 16-byte aligned");
    if (((flags & K2_D1) && !dK_dev) || ((flags & K2_D2) && !d2K_dev))
        return fail(K2_E_ARG, "k2_eval_device_noalloc: derivative requested without an output buffer");
    if (!(k >= 0.0) || !isfinite(k))
        return fail(K2_E_ARG, "k2_eval_device_noalloc: k must be finite and >= 0");
    if (n_cand == 0) return K2_OK;
    std::lock_guard<std::mutex> lk(t->mu);
    HIPCHK(hipSetDevice(t->device));
    hipStream_t st = (hipStream_t)stream;   // NULL is the default stream
    WeightArrays* ws = nullptr;
    for (auto& w : t->wsets)   // lookup only: computing a new k's weights would allocate
        if (w.k == k) { ws = &w; break; }
    if (!ws)
        return fail(K2_E_ARG, "k2_eval_device_noalloc: the weights for this k are not resident "
                              "(a k2_mixed_head or k2_eval_device call with this k computes them)");
    t->use_clock++;
    ws->last_use = t->use_clock;
    PartialBuf pb;
    pb.p = (double*)scratch_ptr(t->device);
    pb.cap = K2_SCRATCH_BYTES;
    pb.grow = false;
    if (!pb.p) return fail(K2_E_HIP, "k2_eval_device_noalloc: no scratch address for this device");
    CallOrder order(t, st);
    HIPCHK(order.err());
    return eval_core(t, (int)n_cand, (const uint4*)V_dev, ws, (int)n_zeros, flags, K_dev, dK_dev,
                     d2K_dev, st, pb);
}

extern "C" int k2_eval(k2_table* t, uint32_t n_cand, const uint64_t* V, double k, uint32_t n_zeros,
                       uint32_t flags, double* K, double* dK, double* d2K) {
    int rc = check_eval_args(t, n_cand, V, &n_zeros, K, "k2_eval");
    if (rc) return rc;
    if (((flags & K2_D1) && !dK) || ((flags & K2_D2) && !d2K))
        return fail(K2_E_ARG, "k2_eval: derivative requested without an output buffer");
    if (!(k >= 0.0) || !isfinite(k)) return fail(K2_E_ARG, "k2_eval: k must be finite and >= 0");
    if (n_cand == 0) return K2_OK;
    std::lock_guard<std::mutex> lk(t->mu);
    HIPCHK(hipSetDevice(t->device));
    const size_t nb_v = (size_t)n_cand * 16, nb_o = (size_t)n_cand * 8;
    rc = ensure_bytes(&t->host_io, &t->host_io_bytes, nb_v + 3 * nb_o + 64);
    if (rc) return rc;
    CallOrder order(t, t->stream);
    HIPCHK(order.err());
    char* base = (char*)t->host_io;
    uint4* dV = (uint4*)base;
    double* dK0 = (double*)(base + nb_v);
    double* dK1 = dK0 + n_cand;
    double* dK2 = dK1 + n_cand;
    HIPCHK(hipMemcpyAsync(dV, V, nb_v, hipMemcpyHostToDevice, t->stream));
    WeightArrays* ws = nullptr;
    rc = get_weights(t, k, t->stream, &ws);
    if (rc) return rc;
    PartialBuf pb;
    rc = eval_core(t, (int)n_cand, dV, ws, (int)n_zeros, flags, dK0, dK1, dK2, t->stream, pb);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(K, dK0, nb_o, hipMemcpyDeviceToHost, t->stream));
    if (flags & K2_D1) HIPCHK(hipMemcpyAsync(dK, dK1, nb_o, hipMemcpyDeviceToHost, t->stream));
    if (flags & K2_D2) HIPCHK(hipMemcpyAsync(d2K, dK2, nb_o, hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    return K2_OK;
}

extern "C" int k2_polish_device(k2_table* t, uint32_t n_cand, uint64_t* V_dev, double k,
                                uint32_t n_zeros, int n_steps, double clamp, uint32_t flags,
                                double* K_dev, double* dK_dev, double* d2K_dev, int32_t* steps_dev,
                                double* K0_dev, void* stream) {
    if (!t || (n_cand > 0 && (!V_dev || !K_dev || !dK_dev || !d2K_dev || !steps_dev)))
        return fail(K2_E_ARG, "k2_polish_device: null pointer");
    if (n_steps < 0 || !(clamp > 0.0) || clamp > 0.25)
        return fail(K2_E_ARG, "k2_polish_device: n_steps >= 0 and 0 < clamp <= 0.25 required");
    t->polish_evals = 0;
    if (n_cand == 0) return K2_OK;
    hipStream_t st = (hipStream_t)stream;   // NULL is the default stream
    const unsigned fl = (flags & K2_MIXED) | K2_VALUE | K2_D1 | K2_D2;
    // compaction buffers: idx (int), gathered points (uint4), their K, K', K'' (double)
    const size_t off_v = ((size_t)n_cand * 4 + 255) / 256 * 256;
    const size_t off_k = off_v + (size_t)n_cand * 16;
    int* idx_d = nullptr;
    uint4* Vc = nullptr;
    double* Kc = nullptr;
    {
        std::lock_guard<std::mutex> lk(t->mu);
        HIPCHK(hipSetDevice(t->device));
        // after the previous call's work on the table (the evaluations inside order themselves,
        // and the last of them records the end of this call)
        HIPCHK(t->order.wait(st));
        int rc = ensure_bytes((void**)&t->polish_active, &t->polish_cap, (size_t)n_cand * 4);
        if (!rc) rc = ensure_bytes(&t->polish_buf, &t->polish_buf_bytes,
                                   off_k + (size_t)n_cand * 3 * sizeof(double));
        if (rc) return rc;
        idx_d = (int*)t->polish_buf;
        Vc = (uint4*)((char*)t->polish_buf + off_v);
        Kc = (double*)((char*)t->polish_buf + off_k);
        std::vector<int> ones(n_cand, 1);
        HIPCHK(hipMemcpyAsync(t->polish_active, ones.data(), (size_t)n_cand * 4,
                              hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(steps_dev, 0, (size_t)n_cand * 4, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    std::vector<int> act(n_cand), idx;
    int64_t evals = 0;
    for (int it = 0; it < n_steps; ++it) {
        int n_act = (int)n_cand;
        if (it > 0) {
            // only candidates still moving are evaluated again (the reference stops each peak
            // once its step falls below 1e-13 or its K'' vanishes; the precision form stops at
            // 1e-11, below which its FP32 rounding in K2' moves the point by as much)
            HIPCHK(hipMemcpyAsync(act.data(), t->polish_active, (size_t)n_cand * 4,
                                  hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            idx.clear();
            for (uint32_t c = 0; c < n_cand; ++c)
                if (act[c]) idx.push_back((int)c);
            n_act = (int)idx.size();
            if (n_act == 0) break;
        }
        if (n_act == (int)n_cand) {
            int rc = k2_eval_device(t, n_cand, V_dev, k, n_zeros, fl, K_dev, dK_dev, d2K_dev, st);
            if (rc) return rc;
            // the first pass is at the starting points: their K, without a pass of its own
            if (it == 0 && K0_dev)
                HIPCHK(hipMemcpyAsync(K0_dev, K_dev, (size_t)n_cand * 8, hipMemcpyDeviceToDevice,
                                      st));
        } else {
            HIPCHK(hipMemcpyAsync(idx_d, idx.data(), (size_t)n_act * 4, hipMemcpyHostToDevice, st));
            HIPCHK(launch_gather((const uint4*)V_dev, idx_d, n_act, Vc, st));
            int rc = k2_eval_device(t, (uint32_t)n_act, (const uint64_t*)Vc, k, n_zeros, fl, Kc,
                                    Kc + n_cand, Kc + 2 * (size_t)n_cand, st);
            if (rc) return rc;
            HIPCHK(launch_scatter3(idx_d, n_act, Kc, Kc + n_cand, Kc + 2 * (size_t)n_cand, K_dev,
                                   dK_dev, d2K_dev, st));
            HIPCHK(hipStreamSynchronize(st));   // idx is rebuilt on the host next step
        }
        evals += n_act;
        std::lock_guard<std::mutex> lk(t->mu);
        HIPCHK(launch_newton((uint4*)V_dev, dK_dev, d2K_dev, (int)n_cand, clamp,
                             (flags & K2_MIXED) ? K2_MIXED_STEP_TOL : 1e-13, t->polish_active,
                             steps_dev, st));
    }
    evals += n_cand;
    t->polish_evals = evals;
    int rc = k2_eval_device(t, n_cand, V_dev, k, n_zeros, fl, K_dev, dK_dev, d2K_dev, st);
    if (!rc && n_steps == 0 && K0_dev)   // no Newton pass: the final pass is at the start
        HIPCHK(hipMemcpyAsync(K0_dev, K_dev, (size_t)n_cand * 8, hipMemcpyDeviceToDevice, st));
    return rc;
}

extern "C" int k2_polish_stats(k2_table* t, int64_t* cand_evals) {
    if (!t || !cand_evals) return fail(K2_E_ARG, "k2_polish_stats: null pointer");
    *cand_evals = t->polish_evals;
    return K2_OK;
}

extern "C" int k2_mixed_head(k2_table* t, double k, uint32_t n_zeros, uint32_t* head) {
    if (!t || !head) return fail(K2_E_ARG, "k2_mixed_head: null pointer");
    if (!(k >= 0.0) || !std::isfinite(k)) return fail(K2_E_ARG, "k2_mixed_head: k must be >= 0");
    if (n_zeros == 0) n_zeros = t->n;
    if (n_zeros > t->n) return fail(K2_E_ARG, "k2_mixed_head: n_zeros exceeds the table");
    std::lock_guard<std::mutex> lk(t->mu);
    HIPCHK(hipSetDevice(t->device));
    CallOrder order(t, t->stream);
    HIPCHK(order.err());
    WeightArrays* ws = nullptr;
    int rc = get_weights(t, k, t->stream, &ws);
    if (rc) return rc;
    *head = mixed_head_zeros(ws->abs_blocks, n_zeros, t->tiling.mixed_head, t->tiling.mixed_tail);
    return K2_OK;
}

extern "C" int k2_split_plan(int32_t nz, int32_t n_ctiles, int32_t target, int32_t factor,
                             int32_t* out) {
    if (!out || nz < 0 || n_ctiles < 1 || target < 1 || factor < 1)
        return fail(K2_E_ARG, "k2_split_plan: nz >= 0, n_ctiles, target and factor >= 1 required");
    SplitPlan sp;
    guided_splits(nz, n_ctiles, target, factor, &sp);
    out[0] = sp.n;
    out[1] = sp.zps;
    out[2] = sp.n_long;
    out[3] = sp.zps_long;
    out[4] = ctile_blocks(n_ctiles, sp.n, sp.n_long);
    return K2_OK;
}

extern "C" int k2_phase_words(k2_table* t, uint32_t n_cand, const uint64_t* V, uint32_t z0,
                              uint32_t nz, int mode, uint64_t* out) {
    if (!t || (n_cand && nz && (!V || !out))) return fail(K2_E_ARG, "k2_phase_words: null pointer");
    if ((uint64_t)z0 + nz > t->n) return fail(K2_E_ARG, "k2_phase_words: zero range out of table");
    if (mode < 0 || mode > 4) return fail(K2_E_ARG, "k2_phase_words: mode must be 0..4");
    if ((uint64_t)n_cand * nz > ((uint64_t)1 << 31))
        return fail(K2_E_ARG, "k2_phase_words: n_cand * nz too large");
    if (n_cand == 0 || nz == 0) return K2_OK;
    std::lock_guard<std::mutex> lk(t->mu);
    HIPCHK(hipSetDevice(t->device));
    const size_t nb_v = (size_t)n_cand * 16, nb_o = (size_t)n_cand * nz * 8;
    int rc = ensure_bytes(&t->host_io, &t->host_io_bytes, nb_v + nb_o + 64);
    if (rc) return rc;
    CallOrder order(t, t->stream);
    HIPCHK(order.err());
    uint4* dV = (uint4*)t->host_io;
    uint64_t* dO = (uint64_t*)((char*)t->host_io + nb_v);
    HIPCHK(hipMemcpyAsync(dV, V, nb_v, hipMemcpyHostToDevice, t->stream));
    HIPCHK(launch_phase_words(t->g, t->gf, (int)z0, (int)nz, dV, (int)n_cand, mode, dO,
                              t->stream));
    HIPCHK(hipMemcpyAsync(out, dO, nb_o, hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    return K2_OK;
}

extern "C" int k2_sincos_words(uint32_t n, const uint64_t* W, int mode, double* out) {
    if (n && (!W || !out)) return fail(K2_E_ARG, "k2_sincos_words: null pointer");
    if (n == 0) return K2_OK;
    if (n > (1u << 28)) return fail(K2_E_ARG, "k2_sincos_words: n too large");
    if (mode < 0 || mode > 2) return fail(K2_E_ARG, "k2_sincos_words: mode must be 0, 1 or 2");
    uint64_t* dW = nullptr;
    double* dO = nullptr;
    k2dev::Tab64* t64 = nullptr;
    k2dev::Tab32* t32 = nullptr;
    int rc = upload_tables(&t64, &t32, nullptr);
    if (rc) return rc;
    hipError_t e = hipMalloc((void**)&dW, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&dO, (size_t)n * 16);
    if (e == hipSuccess) e = hipMemcpy(dW, W, (size_t)n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_sincos_words(dW, (int)n, mode, t64, t32, dO, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dO, (size_t)n * 16, hipMemcpyDeviceToHost);
    (void)hipFree(dW);
    (void)hipFree(dO);
    (void)hipFree(t64);
    (void)hipFree(t32);
    if (e != hipSuccess) return fail(K2_E_HIP, std::string("k2_sincos_words: ") + hipGetErrorString(e));
    return K2_OK;
}

extern "C" int k2_profile(k2_table* t, int enable) {
    if (!t) return fail(K2_E_ARG, "k2_profile: null table");
    std::lock_guard<std::mutex> lk(t->mu);
    HIPCHK(hipSetDevice(t->device));
    for (auto& pp : t->prof_pending) {
        HIPCHK(hipEventSynchronize(pp.b));
        t->ev_pool.push_back(pp.a);
        t->ev_pool.push_back(pp.b);
    }
    t->prof_pending.clear();
    t->prof_ms = 0.0;
    t->prof_launches = 0;
    t->profile = enable != 0;
    return K2_OK;
}

extern "C" int k2_kernel_time(k2_table* t, double* ms, int64_t* launches) {
    if (!t || !ms || !launches) return fail(K2_E_ARG, "k2_kernel_time: null pointer");
    std::lock_guard<std::mutex> lk(t->mu);
    HIPCHK(hipSetDevice(t->device));
    for (auto& pp : t->prof_pending) {
        HIPCHK(hipEventSynchronize(pp.b));
        float e = 0.f;
        HIPCHK(hipEventElapsedTime(&e, pp.a, pp.b));
        t->prof_ms += e;
        t->prof_launches += 1;
        t->ev_pool.push_back(pp.a);
        t->ev_pool.push_back(pp.b);
    }
    t->prof_pending.clear();
    *ms = t->prof_ms;
    *launches = t->prof_launches;
    return K2_OK;
}

// ---- reference signatures ---------------------------------------------------------------------
// State of the reference-signature entry points per device: one stream, one device buffer and
// two pinned host staging buffers, kept across calls (the device buffer grows to the largest
// chunk seen), so that a driver's call loop allocates and creates nothing after its first call.
// The reference functions hold no state and are reentrant; here concurrent calls on one device
// take turns on the context's mutex.
//
// A and B live in the caller's pageable host memory (numpy arrays in the reference's drivers),
// and they are the bulk of these calls' time (16 B per evaluation against a few ns of kernel per
// thousand evaluations).  A copy between device memory and pageable memory runs at well under the
// link's rate, so A and B move through the pinned staging buffers in pieces: the GPU copies piece
// k into one buffer while host threads copy piece k - 1 out of (or into) the other.
#define K2_COMPAT_CHUNK ((size_t)256 << 20)   // bytes of A/B rows per chunk on the device
#define K2_COMPAT_STAGE ((size_t)64 << 20)    // bytes per pinned staging buffer (two)
#define K2_COMPAT_THREADS 8                   // host threads of a staging copy, at most

struct CompatCtx {
    hipStream_t st = nullptr;       // kernels and the copies back
    hipStream_t up = nullptr;       // uploads of A and B (polish_peaks_with_weights)
    char* buf = nullptr;
    size_t cap = 0;
    char* pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};        // the pinned buffers' last copies
    hipEvent_t loaded[2] = {nullptr, nullptr};    // a device half's upload done
    hipEvent_t used[2] = {nullptr, nullptr};      // a device half's kernel done
    std::mutex mu;
};

static CompatCtx* compat_ctx(hipError_t* err) {
    static std::mutex g_mu;
    static CompatCtx* g_ctx[K2_MAX_DEVICES] = {};
    int dev = 0;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= K2_MAX_DEVICES) { *err = hipErrorInvalidDevice; return nullptr; }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx[dev]) {
        CompatCtx* c = new CompatCtx();
        *err = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
        if (*err == hipSuccess) *err = hipStreamCreateWithFlags(&c->up, hipStreamNonBlocking);
        for (int i = 0; i < 2 && *err == hipSuccess; ++i) {
            *err = hipEventCreateWithFlags(&c->ev[i], hipEventDisableTiming);
            if (*err == hipSuccess) *err = hipEventCreateWithFlags(&c->loaded[i], hipEventDisableTiming);
            if (*err == hipSuccess) *err = hipEventCreateWithFlags(&c->used[i], hipEventDisableTiming);
        }
        for (int i = 0; i < 2 && *err == hipSuccess; ++i)
            *err = hipHostMalloc((void**)&c->pin[i], K2_COMPAT_STAGE, hipHostMallocDefault);
        if (*err != hipSuccess) {
            for (int i = 0; i < 2; ++i) {
                if (c->pin[i]) (void)hipHostFree(c->pin[i]);
                if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
                if (c->loaded[i]) (void)hipEventDestroy(c->loaded[i]);
                if (c->used[i]) (void)hipEventDestroy(c->used[i]);
            }
            if (c->up) (void)hipStreamDestroy(c->up);
            if (c->st) (void)hipStreamDestroy(c->st);
            delete c;
            return nullptr;
        }
        g_ctx[dev] = c;
    }
    return g_ctx[dev];
}

static hipError_t compat_reserve(CompatCtx* c, size_t need) {
    if (c->cap >= need) return hipSuccess;
    if (c->buf) (void)hipFree(c->buf);
    c->buf = nullptr;
    c->cap = 0;
    hipError_t e = hipMalloc((void**)&c->buf, need);
    if (e == hipSuccess) c->cap = need;
    return e;
}

// host memcpy of n bytes, split over up to K2_COMPAT_THREADS threads for large copies
static void host_copy(char* dst, const char* src, size_t n) {
    unsigned t = std::max(1u, std::min<unsigned>(K2_COMPAT_THREADS, std::thread::hardware_concurrency()));
    if (n < ((size_t)4 << 20)) t = 1;
    if (t == 1) { memcpy(dst, src, n); return; }
    const size_t per = (n / t + 4095) / 4096 * 4096;
    std::vector<std::thread> th;
    th.reserve(t);
    size_t done = 0;   // bytes handed to threads; the rest is copied here
    try {
        for (unsigned i = 0; i < t; ++i) {
            const size_t a = std::min(n, i * per), b = std::min(n, a + per);
            if (a < b) th.emplace_back([=] { memcpy(dst + a, src + a, b - a); });
            done = b;
        }
    } catch (...) {   // a thread could not be started
    }
    if (done < n) memcpy(dst + done, src + done, n - done);
    for (auto& x : th) x.join();
}

// n bytes from device memory to (pageable) host memory through the two pinned buffers: piece k
// of K2_COMPAT_STAGE bytes goes to buffer k mod 2 on the stream; the host copies piece k - 1 out of
// the other buffer meanwhile.  Returns when all n bytes are in dst.
static hipError_t d2h_staged(CompatCtx* c, char* dst, const char* src, size_t n, hipStream_t st) {
    const size_t S = K2_COMPAT_STAGE, np = (n + S - 1) / S;
    hipError_t e = hipSuccess;
    auto drain = [&](size_t k) {
        const int b = (int)(k & 1);
        hipError_t r = hipEventSynchronize(c->ev[b]);
        if (r == hipSuccess) host_copy(dst + k * S, c->pin[b], std::min(S, n - k * S));
        return r;
    };
    for (size_t k = 0; k < np && e == hipSuccess; ++k) {
        const int b = (int)(k & 1);
        if (k >= 2) e = drain(k - 2);   // buffer b held piece k - 2
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->pin[b], src + k * S, std::min(S, n - k * S), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(c->ev[b], st);
    }
    for (size_t k = np >= 2 ? np - 2 : 0; k < np && e == hipSuccess; ++k) e = drain(k);
    return e;
}

// n bytes from (pageable) host memory to device memory through the pinned buffers: the host
// fills buffer k mod 2 with piece k once the GPU has read piece k - 2 out of it, and the GPU
// copies piece k - 1 meanwhile.  The copies are queued on the stream; the last two may still be
// in flight at return (the buffers are not reused before their events).
static hipError_t h2d_staged(CompatCtx* c, char* dst, const char* src, size_t n, hipStream_t st) {
    const size_t S = K2_COMPAT_STAGE, np = (n + S - 1) / S;
    hipError_t e = hipSuccess;
    for (size_t k = 0; k < np && e == hipSuccess; ++k) {
        const int b = (int)(k & 1);
        const size_t len = std::min(S, n - k * S);
        e = hipEventSynchronize(c->ev[b]);   // the GPU is done with what the buffer held
        if (e != hipSuccess) break;
        host_copy(c->pin[b], src + k * S, len);
        e = hipMemcpyAsync(dst + k * S, c->pin[b], len, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(c->ev[b], st);
    }
    return e;
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

static void fill_nan(double* p, size_t n) {
    if (!p) return;
    for (size_t i = 0; i < n; ++i) p[i] = NAN;
}

extern "C" void compute_peak_AB(int n_peaks, const uint64_t* x_lo_arr, const uint64_t* x_hi_arr,
                                const double* dy0_arr, int n_zeros, const double* alpha,
                                const double* gamma, const uint64_t* ghi_lo, const uint64_t* ghi_hi,
                                const uint64_t* glo_lo, const uint64_t* glo_hi,
                                const uint64_t* psi_lo, const uint64_t* psi_hi, double* out_A,
                                double* out_B) {
    if (n_peaks <= 0 || n_zeros <= 0) return;
    const size_t np = (size_t)n_peaks, nz = (size_t)n_zeros, nab = np * nz;
    // peaks in chunks of at most K2_COMPAT_CHUNK bytes of A/B (at least one row) and 65,535 rows
    // (the grid's second dimension); each chunk's rows are copied back before the next is formed
    const size_t rows = std::max<size_t>(1, std::min<size_t>({np, 65535, K2_COMPAT_CHUNK / (16 * nz)}));
    // device layout: 6 zero words, alpha, gamma (nz each) | x_lo, x_hi, dy0 (rows each) | A B
    const size_t off_x = align256(nz * 8 * 8), off_ab = off_x + align256(rows * 8 * 3);
    hipError_t e = hipSuccess;
    CompatCtx* c = compat_ctx(&e);
    if (c) {
        std::lock_guard<std::mutex> lk(c->mu);
        hipStream_t st = c->st;
        e = compat_reserve(c, off_ab + rows * nz * 16);
        uint64_t* zw = (uint64_t*)c->buf;
        double* dal = (double*)(zw + 6 * nz);
        double* dga = dal + nz;
        uint64_t* xl = (uint64_t*)(c->buf + off_x);
        uint64_t* xh = xl + rows;
        double* ddy = (double*)(xh + rows);
        double* dA = (double*)(c->buf + off_ab);
        double* dB = dA + rows * nz;
        const uint64_t* hs[6] = {ghi_lo, ghi_hi, glo_lo, glo_hi, psi_lo, psi_hi};
        for (int q = 0; q < 6 && e == hipSuccess; ++q)
            e = hipMemcpyAsync(zw + q * nz, hs[q], nz * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dal, alpha, nz * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dga, gamma, nz * 8, hipMemcpyHostToDevice, st);
        const uint64_t* ptrs[8] = {xl, xh, zw, zw + nz, zw + 2 * nz, zw + 3 * nz, zw + 4 * nz,
                                   zw + 5 * nz};
        for (size_t p0 = 0; p0 < np && e == hipSuccess; p0 += rows) {
            const size_t r = std::min(rows, np - p0);
            e = hipMemcpyAsync(xl, x_lo_arr + p0, r * 8, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipMemcpyAsync(xh, x_hi_arr + p0, r * 8, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipMemcpyAsync(ddy, dy0_arr + p0, r * 8, hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = launch_ref_peak_ab((int)r, ptrs, ddy, n_zeros, dal, dga, dA, dB, st);
            if (e == hipSuccess)
                e = d2h_staged(c, (char*)(out_A + p0 * nz), (const char*)dA, r * nz * 8, st);
            if (e == hipSuccess)
                e = d2h_staged(c, (char*)(out_B + p0 * nz), (const char*)dB, r * nz * 8, st);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        // a failed call leaves no copy or kernel behind that could still use the buffers
        if (e != hipSuccess) (void)hipStreamSynchronize(st);
    }
    if (e != hipSuccess) {
        fail(K2_E_HIP, std::string("compute_peak_AB: ") + hipGetErrorString(e));
        fill_nan(out_A, nab); fill_nan(out_B, nab);
    }
}

extern "C" void polish_peaks_with_weights(int n_peaks, int n_zeros, const double* A,
                                          const double* B, const double* gamma, const double* w,
                                          double grid_half_width, int grid_steps,
                                          double* out_h_unpol, double* out_h_pol,
                                          double* out_ddy_pol) {
    if (n_peaks <= 0) return;
    const size_t np = (size_t)n_peaks, nz = (size_t)std::max(n_zeros, 0);
    // peaks in chunks of at most K2_COMPAT_CHUNK bytes of A/B rows (at least one row), in two
    // device halves: the upload stream fills half (i mod 2) with chunk i while the kernel stream
    // runs chunk i - 1 out of the other half
    const size_t rows = nz ? std::max<size_t>(1, std::min<size_t>(np, K2_COMPAT_CHUNK / (16 * nz)))
                           : np;
    // device layout: gamma, w (nz each) | h_unpol, h_pol, ddy (np each) | 2 x (A B) (rows x nz)
    const size_t off_o = align256(nz * 16), off_ab = off_o + align256(np * 24);
    const size_t half = align256(rows * nz * 16);
    hipError_t e = hipSuccess;
    CompatCtx* c = compat_ctx(&e);
    if (c) {
        std::lock_guard<std::mutex> lk(c->mu);
        hipStream_t st = c->st, up = c->up;
        e = compat_reserve(c, off_ab + 2 * half + 64);
        double* dg = (double*)c->buf;
        double* dw = dg + nz;
        double* oh = (double*)(c->buf + off_o);
        double* op = oh + np;
        double* od = op + np;
        if (e == hipSuccess && nz) e = hipMemcpyAsync(dg, gamma, nz * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && nz) e = hipMemcpyAsync(dw, w, nz * 8, hipMemcpyHostToDevice, st);
        size_t i = 0;
        for (size_t p0 = 0; p0 < np && e == hipSuccess; p0 += rows, ++i) {
            const size_t r = std::min(rows, np - p0);
            const int h = (int)(i & 1);
            double* dA = (double*)(c->buf + off_ab + h * half);
            double* dB = dA + r * nz;
            if (i >= 2) e = hipStreamWaitEvent(up, c->used[h], 0);   // chunk i - 2 is done with it
            if (e == hipSuccess && nz)
                e = h2d_staged(c, (char*)dA, (const char*)(A + p0 * nz), r * nz * 8, up);
            if (e == hipSuccess && nz)
                e = h2d_staged(c, (char*)dB, (const char*)(B + p0 * nz), r * nz * 8, up);
            if (e == hipSuccess) e = hipEventRecord(c->loaded[h], up);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, c->loaded[h], 0);
            if (e == hipSuccess)
                e = launch_ref_polish((int)r, (int)nz, dA, dB, dg, dw, grid_half_width, grid_steps,
                                      oh + p0, op + p0, od + p0, st);
            if (e == hipSuccess) e = hipEventRecord(c->used[h], st);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(out_h_unpol, oh, np * 8, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(out_h_pol, op, np * 8, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(out_ddy_pol, od, np * 8, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        // a failed call leaves no upload or kernel behind that could still read the buffers
        if (e != hipSuccess) (void)hipStreamSynchronize(st);
        (void)hipStreamSynchronize(up);
    }
    if (e != hipSuccess) {
        fail(K2_E_HIP, std::string("polish_peaks_with_weights: ") + hipGetErrorString(e));
        fill_nan(out_h_unpol, np); fill_nan(out_h_pol, np); fill_nan(out_ddy_pol, np);
    }
}
