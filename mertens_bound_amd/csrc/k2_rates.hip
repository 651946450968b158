// Issue-rate microbenchmark of the VALU instruction classes the evaluation kernels use, on the
// GPU the library runs on (debug entry k2_valu_rates).  The roofline of a streamed trig sum is
// the peak rate of each instruction class it executes; the vendor publishes the FP64 and FP32
// FMA peaks, not the rates of the 32x32->64 integer multiply-add (v_mad_u64_u32, the exact
// phase) or of the hardware transcendentals (v_sin_f32), so these are measured.
//
// Each lane runs 8 independent dependency chains of one instruction (enough to cover the
// pipeline latency at 8 waves per SIMD), 64 instructions per chain per loop pass, over a grid of
// 16 workgroups of 256 lanes per CU.  The result of every chain feeds one store per lane, so no
// instruction is dead.  Rates are lane operations per second: wave instructions x 64 / time.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "k2_internal.h"

namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 64;

__global__ __launch_bounds__(256) void rate_mad_u64_kernel(int iters, uint32_t a,
                                                           uint64_t* __restrict__ out) {
    uint64_t x[kChains];
    uint32_t b = threadIdx.x | 1u;
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = (uint64_t)(threadIdx.x + c) << 7;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) {
                // x = a * b + x over 64 bits (the product's high half is not dead: x feeds x)
                asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0"
                             : "+v"(x[c]) : "v"(a), "v"(b) : "vcc");
            }
        }
    }
    uint64_t s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s ^= x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void rate_fma_f64_kernel(int iters, double a,
                                                           double* __restrict__ out) {
    double x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = 1.0 + 1e-3 * (threadIdx.x + c);
    const double b = 1e-9 * threadIdx.x;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) x[c] = __builtin_fma(x[c], a, b);
        }
    }
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void rate_fma_f32_kernel(int iters, float a,
                                                           double* __restrict__ out) {
    float x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = 1.0f + 1e-3f * (threadIdx.x + c);
    const float b = 1e-9f * threadIdx.x;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) x[c] = __builtin_fmaf(x[c], a, b);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void rate_sin_f32_kernel(int iters, double* __restrict__ out) {
    float x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = 1e-3f * (threadIdx.x + c);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) x[c] = __builtin_amdgcn_sinf(x[c]);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

// out[0..3]: lane operations per second of v_mad_u64_u32, v_fma_f64, v_fma_f32, v_sin_f32
extern "C" int k2_valu_rates(double* out) {
    using k2impl::set_error;
    if (!out) return set_error(K2_E_ARG, "k2_valu_rates: null pointer");
    int dev = 0, n_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = 16 * (n_cu > 0 ? n_cu : 256), nt = 256;
    void* buf = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipMalloc(&buf, (size_t)grid * nt * 8);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const int iters = 64;
    for (int kind = 0; kind < 4 && e == hipSuccess; ++kind) {
        double best = 0.0;
        for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {   // first pass warms up
            e = hipEventRecord(e0, st);
            if (e != hipSuccess) break;
            if (kind == 0)
                rate_mad_u64_kernel<<<grid, nt, 0, st>>>(iters, 0x9E3779B9u, (uint64_t*)buf);
            else if (kind == 1)
                rate_fma_f64_kernel<<<grid, nt, 0, st>>>(iters, 0.999999, (double*)buf);
            else if (kind == 2)
                rate_fma_f32_kernel<<<grid, nt, 0, st>>>(iters, 0.999999f, (double*)buf);
            else
                rate_sin_f32_kernel<<<grid, nt, 0, st>>>(iters, (double*)buf);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            const double ops = (double)grid * nt * iters * kUnroll * kChains;
            if (e == hipSuccess && rep > 0 && ms > 0.f) best = std::max(best, ops / (ms * 1e-3));
        }
        out[kind] = best;
    }
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return set_error(K2_E_HIP, std::string("k2_valu_rates: ") + hipGetErrorString(e));
    return K2_OK;
}
