// Batched evaluation of uneven requests in one persistent launch, captured in a HIP graph.
//
// A request r has its own zero cutoff (the first n_r zeros), Gaussian constant k_r and candidate
// slice.  The host cuts the requests into work items:
//   * C items: a tile of 256*NC candidates (NC per lane) x a chunk of zeros, the same body as the
//     one-launch kernel.  A request's full tiles are its own; the candidates left over from all
//     requests with the same k are packed together, in order of decreasing cutoff, into shared
//     tiles (candidates fetched through an index list).  A shared tile's zero range is cut at
//     every member's cutoff, and a candidate's sum takes only the items below its own cutoff (the
//     lanes of members whose cutoff has passed still compute, and their slots are never read);
//     ordered by cutoff, neighbours in a tile have similar cutoffs, so few lanes idle that way;
//   * Z items: a tile of CZ candidates x a longer chunk of zeros, one zero per lane at a time,
//     only for the last shared tile of a k when it holds fewer than c_min_fill candidates.  Each
//     Z item re-reads its zero chunk once per pass over its candidates, so packing the leftovers
//     into shared C tiles (each zero read once per 256*NC candidates) is what keeps the launch's
//     memory traffic near the algorithmic bytes;
//   * with K2_MIXED, the zeros below the FP64 head and the FP32 tail are separate items (a shared
//     tile takes the largest head among its members: more FP64, never less).
// Items are sorted longest first and pulled from an atomic counter by a grid of resident
// workgroups; each item writes its partial sums to its own slot, and a reduction kernel adds a
// candidate's slots in zero order, so results do not depend on the execution order.
// The graph holds [counter reset kernel, persistent kernel, reduction kernel]: no memset nodes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "k2_device.h"
#include "k2_internal.h"
#include "k2_tiles.h"

using namespace k2dev;

struct BatchItem {
    uint32_t flags;     // bit0: Z tile, bit1: FP32, bit2: candidates through the index list,
                        // bits 8-9: k slot
    uint32_t z0, z1;
    uint32_t c_first;   // global index of the tile's first candidate (or of its first index)
    uint32_t c_count;
    uint32_t pad;
    uint64_t poff;      // partial[poff + q * width + pos]
};

struct CandDesc {
    uint64_t base;      // first item slot of the candidate's tile
    uint32_t n_items;   // items of the tile (zero chunks)
    uint16_t width;     // candidates per slot row (256*NC or CZ)
    uint16_t pos;
};

struct BatchViews {
    ZeroView zv[4];
};

__global__ void k2_counter_reset_kernel(int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

// LDS of the persistent kernel: one 40 KiB region, laid out per kind of item --
//   FP64 C items: the split 2,048-point table (sincos_tabw_at) at [0, 32 KiB + 8), records in
//                 scalar registers (ctile_f64s);
//   FP64 Z items: the 1,024-point FP64 table at [0, 16 KiB), the Z-tile scratch above it;
//   FP32 items:   the FP32 table at [0, 16 KiB), the C-tile records or the Z-tile scratch above;
// and the next item's index in its last word (inside the FP32 records' span, past the split
// table's), written and read only between items: every item ends with a barrier.  The table is
// reloaded when the kind of table changes (a few times per workgroup: microseconds against
// milliseconds per item), which keeps four workgroups per CU (40 KiB with K2 and K2').
constexpr size_t K2_BATCH_TAB_BYTES = sizeof(Tab64) * K2_TAB;   // 16 KiB
template <int FL, int CZ>
constexpr size_t batch_lds_bytes() {
    constexpr size_t rec = sizeof(CTileSmem32<FL>) > sizeof(ZTileSmem<CZ>)
                               ? sizeof(CTileSmem32<FL>) : sizeof(ZTileSmem<CZ>);
    constexpr size_t low = K2_BATCH_TAB_BYTES + rec;
    constexpr size_t wide = K2_TABW_DOUBLES * sizeof(double) + sizeof(int);
    return ((low > wide ? low : wide) + 15) / 16 * 16;
}

template <int NC, int CZ, int FL, bool ALLOW32>
__global__ __launch_bounds__(256, 4) void k2_batch_kernel(
    BatchViews views, const BatchItem* __restrict__ items, int n_items, int* counter,
    const uint4* __restrict__ cand, const uint32_t* __restrict__ cidx,
    double* __restrict__ partial, const Tab64* __restrict__ tab64,
    const Tab32* __restrict__ tab32) {
    constexpr size_t LDS = batch_lds_bytes<FL, CZ>();
    static_assert(sizeof(Tab32) * K2_TAB32 <= K2_BATCH_TAB_BYTES, "FP32 table fits");
    __shared__ __align__(16) unsigned char lds[LDS];
    Tab64* s_tab64 = reinterpret_cast<Tab64*>(lds);
    Tab32* s_tab32 = reinterpret_cast<Tab32*>(lds);
    double* s_tabw = reinterpret_cast<double*>(lds);
    auto& sm32 = *reinterpret_cast<CTileSmem32<FL>*>(lds + K2_BATCH_TAB_BYTES);
    auto& zsm = *reinterpret_cast<ZTileSmem<CZ>*>(lds + K2_BATCH_TAB_BYTES);
    int& s_item = *reinterpret_cast<int*>(lds + LDS - sizeof(int));
    const int tid = threadIdx.x;
    int loaded = -1;   // table in LDS: 0 FP64 1,024-point, 1 FP32, 2 FP64 split (uniform)
    constexpr int CW = K2_WG * NC;
    for (;;) {
        __syncthreads();   // every wave is done with the previous item's LDS before s_item lands
        if (tid == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);   // uniform: scalar loads below
        __syncthreads();
        if (it >= n_items) break;
        const BatchItem I = items[it];
        const ZeroView zv = views.zv[(I.flags >> 8) & 3];
        const bool f32 = ALLOW32 && (I.flags & 2);
        const bool zt = (I.flags & 1) != 0;
        const int kind = f32 ? 1 : (zt ? 0 : 2);
        if (kind != loaded) {
            // every lookup of the previous item is done: tiles end with a barrier
            if (kind == 1) {
                for (int i = tid; i < K2_TAB32; i += K2_WG) s_tab32[i] = tab32[i];
            } else if (kind == 0) {
                for (int i = tid; i < K2_TAB; i += K2_WG) s_tab64[i] = tab64[i];
            } else {
                const double* tw = reinterpret_cast<const double*>(tab64 + K2_TAB);
                for (int i = tid; i < K2_TABW_DOUBLES; i += K2_WG) s_tabw[i] = tw[i];
            }
            __syncthreads();
            loaded = kind;
        }
        if (zt) {
            double* out = partial + I.poff;
            if (f32) {
                if constexpr (ALLOW32)
                    ztile<CZ, FL, true>(zv, I.z0, I.z1, cand, I.c_first, I.c_count, out, zsm,
                                        s_tab64, s_tab32);
            } else {
                ztile<CZ, FL, false>(zv, I.z0, I.z1, cand, I.c_first, I.c_count, out, zsm, s_tab64,
                                     s_tab32);
            }
            continue;
        }
        Limbs4 v[NC];
        const bool indirect = (I.flags & 4) != 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const uint32_t c = i * K2_WG + tid;
            uint4 cv = make_uint4(0, 0, 0, 0);
            if (c < I.c_count) cv = cand[indirect ? cidx[I.c_first + c] : I.c_first + c];
            v[i] = Limbs4{cv.x, cv.y, cv.z, cv.w};
        }
        double* out = partial + I.poff;
        if (f32) {
            if constexpr (ALLOW32) {
                Acc<true> acc[NC];
#pragma unroll
                for (int i = 0; i < NC; ++i) acc[i].zero();
                ctile_f32<NC, FL>(zv, I.z0, I.z1, v, acc, sm32, s_tab32);
#pragma unroll
                for (int i = 0; i < NC; ++i)
#pragma unroll
                    for (int q = 0; q < 3; ++q) out[q * CW + i * K2_WG + tid] = acc[i].get(q);
            }
        } else {
            Acc<false> acc[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) acc[i].zero();
            ctile_f64s<NC, FL>(zv.g, zv.a, zv.b, zv.e, I.z0, I.z1, v, acc, s_tabw);
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
                for (int q = 0; q < 3; ++q) out[q * CW + i * K2_WG + tid] = acc[i].get(q);
        }
    }
}

// K[c] = sum over the candidate's tile items of its slot, in zero order
__global__ __launch_bounds__(256) void k2_batch_reduce_kernel(
    const CandDesc* __restrict__ desc, uint64_t n, const double* __restrict__ partial,
    double* __restrict__ K, double* __restrict__ dK, double* __restrict__ d2K) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const CandDesc d = desc[c];
    if (d.n_items == 0) return;   // not part of any request
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    const uint64_t stride = 3ull * d.width;
    for (uint32_t h = 0; h < d.n_items; ++h) {
        const double* row = partial + d.base + h * stride + d.pos;
        s0 += row[0];
        s1 += row[d.width];
        s2 += row[2 * d.width];
    }
    K[c] = s0;
    if (dK) dK[c] = s1;
    if (d2K) d2K[c] = s2;
}

// ---- host side ------------------------------------------------------------------------------

namespace k2impl {

struct BatchPlan {
    std::vector<BatchItem> items;
    std::vector<CandDesc> desc;
    std::vector<uint32_t> cidx;   // candidate index lists of the shared tiles
    uint64_t partial_doubles = 0;
    double useful_evals = 0, padded_evals = 0;
};

// requests -> items and candidate descriptors (see the file comment)
std::string plan_batch(uint32_t n_req, const k2_request* req, uint64_t n_cand_total,
                       uint32_t table_n, const std::vector<double>& kslots,
                       const std::vector<WeightArrays>& kw, bool mixed, const k2_tiling& et,
                       const k2_batch_tiling& ti, BatchPlan& P) {
    const int NC = ti.nc, CZ = ti.cz;
    const uint32_t CW = 256u * NC;
    P.desc.assign(n_cand_total, CandDesc{0, 0, 0, 0});
    std::vector<uint8_t> seen(n_cand_total, 0);
    // items of one tile over the zero range [0, nz_max): cut at every boundary in `cuts` (the
    // members' cutoffs and the FP64 head), chunked; returns the tile's first partial slot
    auto emit_tile = [&](uint32_t c_first, uint32_t c_count, uint32_t width, bool z, bool indirect,
                         int ks, uint32_t head, std::vector<uint32_t> cuts) -> uint64_t {
        const uint64_t chunk = z ? (uint64_t)ti.chunk * CW / CZ : (uint64_t)ti.chunk;
        cuts.push_back(head);
        std::sort(cuts.begin(), cuts.end());
        cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
        const uint64_t base = P.partial_doubles;
        uint32_t h = 0, a = 0;
        for (uint32_t b : cuts) {
            for (uint64_t zz = a; zz < b; zz += chunk) {
                BatchItem it{};
                const bool f32 = mixed && zz >= head;
                it.flags = (z ? 1u : 0u) | (f32 ? 2u : 0u) | (indirect ? 4u : 0u) |
                           ((uint32_t)ks << 8);
                it.z0 = (uint32_t)zz;
                it.z1 = (uint32_t)std::min<uint64_t>(b, zz + chunk);
                it.c_first = c_first;
                it.c_count = c_count;
                it.poff = base + (uint64_t)h * 3 * width;
                P.items.push_back(it);
                P.padded_evals += (double)width * (it.z1 - it.z0);
                ++h;
            }
            a = std::max(a, b);
        }
        P.partial_doubles += (uint64_t)h * 3 * width;
        return base;
    };
    // items of the tile emitted last whose zeros end at or below nz (a member's own items)
    auto items_below = [&](uint64_t base, uint32_t nz) {
        uint32_t n = 0;
        for (size_t i = P.items.size(); i-- > 0 && P.items[i].poff >= base;)
            if (P.items[i].z1 <= nz) ++n;
        return n;
    };
    // leftover candidates per k slot: (cutoff, head, candidate index)
    struct Left { uint32_t nz, head, c; };
    std::vector<std::vector<Left>> left(kslots.size());
    for (uint32_t r = 0; r < n_req; ++r) {
        const k2_request& q = req[r];
        if (q.n_cand == 0) continue;
        const uint32_t nz = q.n_zeros ? q.n_zeros : table_n;
        if (nz > table_n) return "request " + std::to_string(r) + ": n_zeros exceeds the table";
        if (q.cand_offset + q.n_cand > n_cand_total)
            return "request " + std::to_string(r) + ": candidates out of range";
        for (uint64_t c = q.cand_offset; c < q.cand_offset + q.n_cand; ++c) {
            if (seen[c]) return "request " + std::to_string(r) + ": candidates overlap another request";
            seen[c] = 1;
        }
        const int ks = (int)(std::find(kslots.begin(), kslots.end(), q.k) - kslots.begin());
        const uint32_t head =
            mixed ? mixed_head_zeros(kw[ks].abs_blocks, nz, et.mixed_head, et.mixed_tail) : nz;
        P.useful_evals += (double)q.n_cand * nz;
        // the request's full tiles
        const uint32_t n_full = q.n_cand / CW, rem = q.n_cand % CW;
        for (uint32_t i = 0; i < n_full; ++i) {
            const uint32_t c0 = (uint32_t)q.cand_offset + i * CW;
            const uint64_t base = emit_tile(c0, CW, CW, false, false, ks, head, {nz});
            const uint32_t H = items_below(base, nz);
            for (uint32_t j = 0; j < CW; ++j) P.desc[c0 + j] = CandDesc{base, H, (uint16_t)CW, (uint16_t)j};
        }
        for (uint32_t j = 0; j < rem; ++j)
            left[ks].push_back({nz, head, (uint32_t)q.cand_offset + n_full * CW + j});
    }
    // the leftovers of each k, by decreasing cutoff, in shared tiles
    for (size_t ks = 0; ks < left.size(); ++ks) {
        std::vector<Left>& L = left[ks];
        std::stable_sort(L.begin(), L.end(), [](const Left& x, const Left& y) { return x.nz > y.nz; });
        for (size_t t0 = 0; t0 < L.size(); t0 += CW) {
            const uint32_t n = (uint32_t)std::min<size_t>(CW, L.size() - t0);
            if (n < (uint32_t)ti.c_min_fill) {
                // a thin last tile: Z tiles over runs of candidates with one cutoff and contiguous
                // indices
                size_t i = t0;
                while (i < t0 + n) {
                    size_t e = i + 1;
                    while (e < t0 + n && e - i < (size_t)CZ && L[e].nz == L[i].nz &&
                           L[e].c == L[e - 1].c + 1)
                        ++e;
                    const uint32_t cnt = (uint32_t)(e - i);
                    const uint64_t base = emit_tile(L[i].c, cnt, (uint32_t)CZ, true, false, (int)ks,
                                                    L[i].head, {L[i].nz});
                    const uint32_t H = items_below(base, L[i].nz);
                    for (uint32_t j = 0; j < cnt; ++j)
                        P.desc[L[i].c + j] = CandDesc{base, H, (uint16_t)CZ, (uint16_t)j};
                    i = e;
                }
                continue;
            }
            const uint32_t i0 = (uint32_t)P.cidx.size();
            std::vector<uint32_t> cuts;
            uint32_t head = 0;
            for (uint32_t j = 0; j < n; ++j) {
                P.cidx.push_back(L[t0 + j].c);
                cuts.push_back(L[t0 + j].nz);
                head = std::max(head, L[t0 + j].head);
            }
            const uint64_t base = emit_tile(i0, n, CW, false, true, (int)ks, head, cuts);
            for (uint32_t j = 0; j < n; ++j) {
                const Left& m = L[t0 + j];
                P.desc[m.c] = CandDesc{base, items_below(base, m.nz), (uint16_t)CW, (uint16_t)j};
            }
        }
    }
    // longest first (an FP32 item takes about half the time of an FP64 one of the same size)
    auto cost = [&](const BatchItem& it) {
        const double w = (it.flags & 1) ? (double)CZ : (double)CW;
        return w * (it.z1 - it.z0) * ((it.flags & 2) ? 0.5 : 1.0);
    };
    std::stable_sort(P.items.begin(), P.items.end(),
                     [&](const BatchItem& x, const BatchItem& y) { return cost(x) > cost(y); });
    if (P.items.size() > 0x7fffffff) return "too many work items";
    return "";
}

template <int NC, int CZ, int FL, bool A32>
static const void* batch_kernel_ptr() {
    return reinterpret_cast<const void*>(&k2_batch_kernel<NC, CZ, FL, A32>);
}

const void* batch_kernel(int nc, int cz, unsigned flags, bool allow32) {
    const bool d2 = (flags & K2_D2) != 0;
#define K2_SEL(NC_, CZ_)                                                                         \
    if (nc == NC_ && cz == CZ_) {                                                                \
        if (d2) return allow32 ? batch_kernel_ptr<NC_, CZ_, K2_D1 | K2_D2, true>()               \
                               : batch_kernel_ptr<NC_, CZ_, K2_D1 | K2_D2, false>();             \
        return allow32 ? batch_kernel_ptr<NC_, CZ_, K2_D1, true>()                               \
                       : batch_kernel_ptr<NC_, CZ_, K2_D1, false>();                             \
    }
    K2_SEL(2, 8) K2_SEL(2, 16) K2_SEL(4, 8) K2_SEL(4, 16)
#undef K2_SEL
    return nullptr;
}

hipError_t launch_batch(const void* kernel, int grid, const BatchViews* views,
                        const BatchItem* items, int n_items, int* counter, const uint4* cand,
                        const uint32_t* cidx, double* partial, const Tab64* tab64,
                        const Tab32* tab32,
                        const CandDesc* desc, uint64_t n_cand, double* K, double* dK, double* d2K,
                        hipStream_t st) {
    hipLaunchKernelGGL(k2_counter_reset_kernel, dim3(1), dim3(64), 0, st, counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (n_items > 0 && grid > 0) {
        BatchViews v = *views;
        void* args[] = {&v, &items, &n_items, &counter, &cand, &cidx, &partial, &tab64, &tab32};
        e = hipLaunchKernel(kernel, dim3(grid), dim3(256), args, 0, st);
        if (e != hipSuccess) return e;
    }
    const unsigned nb = (unsigned)((n_cand + 255) / 256);
    if (nb > 0) {
        hipLaunchKernelGGL(k2_batch_reduce_kernel, dim3(nb), dim3(256), 0, st, desc, n_cand,
                           (const double*)partial, K, dK, d2K);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace k2impl

// ---- C ABI: batched plans --------------------------------------------------------------------

using namespace k2impl;

struct k2_batch {
    k2_table* t = nullptr;
    std::vector<WeightArrays> w;   // one per distinct k, owned (the table's cache may evict)
    BatchItem* d_items = nullptr;
    CandDesc* d_desc = nullptr;
    uint32_t* d_cidx = nullptr;
    double* d_partial = nullptr;
    int* d_counter = nullptr;
    int n_items = 0;
    uint64_t n_cand = 0;
    int grid = 0;
    const void* kernel = nullptr;
    BatchViews views{};
    const uint4* cand = nullptr;
    double *K = nullptr, *dK = nullptr, *d2K = nullptr;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    StreamOrder order;   // runs on different streams share the plan's partial sums and counter
    k2_batch_info info{};
};

static int berr(int code, const std::string& m) { return set_error(code, m); }

#define BCHK(expr)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            k2_batch_destroy(b);                                                           \
            return berr(K2_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
        }                                                                                  \
    } while (0)

extern "C" int k2_batch_destroy(k2_batch* b) {
    if (!b) return K2_OK;
    if (b->t) (void)hipSetDevice(b->t->device);
    if (b->cap) (void)hipStreamSynchronize(b->cap);
    if (b->order.valid) (void)hipEventSynchronize(b->order.ev);
    if (b->order.ev) (void)hipEventDestroy(b->order.ev);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto& w : b->w) free_weights(w);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_desc);
    (void)hipFree(b->d_cidx);
    (void)hipFree(b->d_partial);
    (void)hipFree(b->d_counter);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    delete b;
    return K2_OK;
}

extern "C" int k2_batch_create(k2_table* t, uint32_t n_req, const k2_request* req,
                               const uint64_t* V_dev, uint64_t n_cand_total, uint32_t flags,
                               const k2_batch_tiling* tiling, double* K_dev, double* dK_dev,
                               double* d2K_dev, k2_batch** out) {
    if (!t || !out || (n_req > 0 && (!req || !V_dev || !K_dev)))
        return berr(K2_E_ARG, "k2_batch_create: null pointer");
    if (((flags & K2_D1) && !dK_dev) || ((flags & K2_D2) && !d2K_dev))
        return berr(K2_E_ARG, "k2_batch_create: derivative requested without an output buffer");
    if (n_cand_total >= (1ull << 32)) return berr(K2_E_ARG, "k2_batch_create: n_cand_total >= 2^32");
    if (((uintptr_t)V_dev & 15) != 0) return berr(K2_E_ARG, "k2_batch_create: V not 16-byte aligned");
    k2_batch_tiling ti{4, 8, 8192, 512, 0};
    if (tiling) {
        ti = *tiling;
        if (ti.nc == 0) ti.nc = 4;
        if (ti.cz == 0) ti.cz = 8;
        if (ti.chunk == 0) ti.chunk = 8192;
        if (ti.c_min_fill == 0) ti.c_min_fill = 256 * ti.nc / 2;
    }
    if ((ti.nc != 2 && ti.nc != 4) || (ti.cz != 8 && ti.cz != 16) || ti.chunk < 256 ||
        ti.c_min_fill < 1 || ti.wg_per_cu < 0)
        return berr(K2_E_ARG, "k2_batch_create: tiling out of range (nc 2|4, cz 8|16, chunk >= 256)");
    std::vector<double> ks;
    for (uint32_t r = 0; r < n_req; ++r) {
        if (!(req[r].k >= 0.0) || !std::isfinite(req[r].k))
            return berr(K2_E_ARG, "k2_batch_create: request k must be finite and >= 0");
        if (std::find(ks.begin(), ks.end(), req[r].k) == ks.end()) ks.push_back(req[r].k);
    }
    if (ks.size() > 4) return berr(K2_E_ARG, "k2_batch_create: more than 4 distinct k values");
    const bool mixed = (flags & K2_MIXED) != 0;
    // requests are checked before anything is allocated (weights are not needed for that)
    {
        std::vector<WeightArrays> none(ks.size());
        for (auto& w : none) w.abs_blocks.assign(t->n_pad / K2_ABS_BLOCK, 0.0);
        BatchPlan Q;
        std::string msg = plan_batch(n_req, req, n_cand_total, t->n, ks, none, false, t->tiling,
                                     ti, Q);
        if (!msg.empty()) return berr(K2_E_ARG, "k2_batch_create: " + msg);
    }

    k2_batch* b = new k2_batch();
    b->t = t;
    b->n_cand = n_cand_total;
    b->cand = (const uint4*)V_dev;
    b->K = K_dev;
    b->dK = dK_dev;
    b->d2K = d2K_dev;
    BCHK(hipSetDevice(t->device));
    BCHK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    BCHK(hipEventCreateWithFlags(&b->order.ev, hipEventDisableTiming));
    for (double k : ks) {
        WeightArrays w;
        if (alloc_weights(w, t->n_pad) != K2_OK) {
            k2_batch_destroy(b);
            return K2_E_NOMEM;
        }
        b->w.push_back(w);
        if (compute_weights(t->gamma, t->w, (int)t->n, (int)t->n_pad, k, b->w.back(), b->cap)) {
            k2_batch_destroy(b);
            return K2_E_HIP;
        }
    }
    BatchPlan P;
    std::string msg = plan_batch(n_req, req, n_cand_total, t->n, ks, b->w, mixed, t->tiling, ti, P);
    if (!msg.empty()) {
        k2_batch_destroy(b);
        return berr(K2_E_ARG, "k2_batch_create: " + msg);
    }
    b->n_items = (int)P.items.size();
    for (size_t i = 0; i < 4; ++i) {
        const WeightArrays& w = b->w.empty() ? WeightArrays{} : b->w[std::min(i, b->w.size() - 1)];
        b->views.zv[i] = ZeroView{t->g, w.a, w.b, w.e, w.af, w.bf, w.ef, t->gf};
    }
    BCHK(hipMalloc((void**)&b->d_counter, 256));
    BCHK(hipMalloc((void**)&b->d_items, std::max<size_t>(1, P.items.size()) * sizeof(BatchItem)));
    BCHK(hipMalloc((void**)&b->d_desc, std::max<size_t>(1, P.desc.size()) * sizeof(CandDesc)));
    BCHK(hipMalloc((void**)&b->d_partial, std::max<uint64_t>(1, P.partial_doubles) * sizeof(double)));
    BCHK(hipMalloc((void**)&b->d_cidx, std::max<size_t>(1, P.cidx.size()) * sizeof(uint32_t)));
    if (!P.cidx.empty())
        BCHK(hipMemcpyAsync(b->d_cidx, P.cidx.data(), P.cidx.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, b->cap));
    if (!P.items.empty())
        BCHK(hipMemcpyAsync(b->d_items, P.items.data(), P.items.size() * sizeof(BatchItem),
                            hipMemcpyHostToDevice, b->cap));
    if (!P.desc.empty())
        BCHK(hipMemcpyAsync(b->d_desc, P.desc.data(), P.desc.size() * sizeof(CandDesc),
                            hipMemcpyHostToDevice, b->cap));
    BCHK(hipStreamSynchronize(b->cap));

    b->kernel = batch_kernel(ti.nc, ti.cz, flags, mixed);
    if (!b->kernel) {
        k2_batch_destroy(b);
        return berr(K2_E_ARG, "k2_batch_create: no kernel for this tiling");
    }
    int per_cu = 0, n_cu = 0;
    BCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, b->kernel, 256, 0));
    BCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, t->device));
    if (ti.wg_per_cu > 0) per_cu = std::min(per_cu, ti.wg_per_cu);
    b->grid = std::max(1, std::min(b->n_items, std::max(1, per_cu) * n_cu));

    b->info.n_items = b->n_items;
    b->info.grid = b->grid;
    b->info.n_k = (int)ks.size();
    b->info.useful_evals = P.useful_evals;
    b->info.padded_evals = P.padded_evals;
    b->info.partial_bytes = (int64_t)(P.partial_doubles * sizeof(double));
    *out = b;
    return K2_OK;
}

static hipError_t batch_launches(k2_batch* b, hipStream_t st) {
    k2_table* t = b->t;
    return launch_batch(b->kernel, b->grid, &b->views, b->d_items, b->n_items, b->d_counter, b->cand,
                        b->d_cidx, b->d_partial, t->tab64, t->tab32, b->d_desc, b->n_cand, b->K,
                        b->dK, b->d2K, st);
}

extern "C" int k2_batch_run_direct(k2_batch* b, void* stream) {
    if (!b) return berr(K2_E_ARG, "k2_batch_run_direct: null batch");
    hipError_t e = hipSetDevice(b->t->device);
    if (e == hipSuccess) e = b->order.wait((hipStream_t)stream);
    if (e == hipSuccess) e = batch_launches(b, (hipStream_t)stream);
    if (e == hipSuccess) e = b->order.mark((hipStream_t)stream);
    if (e != hipSuccess) return berr(K2_E_HIP, std::string("k2_batch_run_direct: ") + hipGetErrorString(e));
    return K2_OK;
}

extern "C" int k2_batch_run(k2_batch* b, void* stream) {
    if (!b) return berr(K2_E_ARG, "k2_batch_run: null batch");
    hipError_t e = hipSetDevice(b->t->device);
    if (e == hipSuccess && !b->exec) {
        // capture once on the batch's own stream; replays may go to any stream
        e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t el = batch_launches(b, b->cap);
            e = hipStreamEndCapture(b->cap, &b->graph);
            if (el != hipSuccess) e = el;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    }
    if (e == hipSuccess) e = b->order.wait((hipStream_t)stream);
    if (e == hipSuccess) e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    if (e == hipSuccess) e = b->order.mark((hipStream_t)stream);
    if (e != hipSuccess) return berr(K2_E_HIP, std::string("k2_batch_run: ") + hipGetErrorString(e));
    return K2_OK;
}

extern "C" int k2_batch_get_info(const k2_batch* b, k2_batch_info* info) {
    if (!b || !info) return berr(K2_E_ARG, "k2_batch_get_info: null pointer");
    *info = b->info;
    return K2_OK;
}

extern "C" int k2_eval_batched_device(k2_table* t, uint32_t n_req, const k2_request* req,
                                      const uint64_t* V_dev, uint64_t n_cand_total, uint32_t flags,
                                      double* K_dev, double* dK_dev, double* d2K_dev,
                                      void* stream) {
    k2_batch* b = nullptr;
    int rc = k2_batch_create(t, n_req, req, V_dev, n_cand_total, flags, nullptr, K_dev, dK_dev,
                             d2K_dev, &b);
    if (rc) return rc;
    rc = k2_batch_run_direct(b, stream);
    if (rc == K2_OK) {
        hipError_t e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) rc = berr(K2_E_HIP, std::string("k2_eval_batched_device: ") + hipGetErrorString(e));
    }
    k2_batch_destroy(b);
    return rc;
}
