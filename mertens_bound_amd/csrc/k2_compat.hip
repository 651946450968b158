// Reference-format polish on the GPU: polish_peaks_with_weights (search/path1_eval_fast.cpp:58-143).
//
// One workgroup per peak.  Each pass over the peak's A/B row (h0, 81-point grid on the first
// 4096 zeros, 6 Newton iterations, final value) is a strided lane loop followed by a wave-shuffle
// + LDS reduction; lane 0 takes the scalar decisions of the reference (grid argmax with strict
// '>' in m order, step clamp +-1e-4, |f''| > 1e-12 guard, |step| < 1e-13 stop, keep the larger
// of |h0| and |h_final|).  Sums are combined in a different order than the reference's
// sequential loops, so results agree to rounding, not bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "k2_device.h"
#include "k2_internal.h"

using namespace k2dev;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;
}

// sums three values over the 256-lane workgroup; every lane receives the totals
__device__ void block_sum3(double& a, double& b, double& c, double* red) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    a = wave_sum(a); b = wave_sum(b); c = wave_sum(c);
    __syncthreads();
    if (lane == 0) { red[wid * 3] = a; red[wid * 3 + 1] = b; red[wid * 3 + 2] = c; }
    __syncthreads();
    a = red[0] + red[3] + red[6] + red[9];
    b = red[1] + red[4] + red[7] + red[10];
    c = red[2] + red[5] + red[8] + red[11];
}

__global__ __launch_bounds__(256) void k2_ref_polish_kernel(
    int n_zeros, const double* __restrict__ A, const double* __restrict__ B,
    const double* __restrict__ gamma, const double* __restrict__ w, double half, int steps,
    double* __restrict__ h_unpol, double* __restrict__ h_pol, double* __restrict__ ddy_pol) {
    __shared__ double red[12];
    const int p = blockIdx.x;
    const double* Ap = A + (size_t)p * n_zeros;
    const double* Bp = B + (size_t)p * n_zeros;
    const int tid = threadIdx.x;

    double h0 = 0.0, u1 = 0.0, u2 = 0.0;
    for (int i = tid; i < n_zeros; i += 256) h0 = fma(w[i], Ap[i], h0);
    block_sum3(h0, u1, u2, red);

    double best = 0.0, best_abs = fabs(h0);
    const int ks = n_zeros < 4096 ? n_zeros : 4096;
    if (steps > 0) {
        for (int m = -steps; m <= steps; ++m) {
            const double dd = (half * m) / steps;
            double hs = 0.0, x1 = 0.0, x2 = 0.0;
            for (int i = tid; i < ks; i += 256) {
                double s, c;
                sincos_general(gamma[i] * dd, s, c);
                hs = fma(w[i], Ap[i] * c - Bp[i] * s, hs);
            }
            block_sum3(hs, x1, x2, red);
            if (fabs(hs) > best_abs) { best_abs = fabs(hs); best = dd; }
        }
    }
    double dd = best;
    for (int it = 0; it < 6; ++it) {
        double f = 0.0, fp = 0.0, fpp = 0.0;
        for (int i = tid; i < n_zeros; i += 256) {
            const double g = gamma[i];
            double s, c;
            sincos_general(g * dd, s, c);
            const double val = Ap[i] * c - Bp[i] * s;
            const double dval = g * (-Ap[i] * s - Bp[i] * c);
            f = fma(w[i], val, f);
            fp = fma(w[i], dval, fp);
            fpp = fma(w[i], -g * g * val, fpp);
        }
        block_sum3(f, fp, fpp, red);
        if (fabs(fpp) > 1e-12) {
            double step = fp / fpp;
            if (step > 1e-4) step = 1e-4;
            if (step < -1e-4) step = -1e-4;
            dd -= step;
            if (fabs(step) < 1e-13) break;
        } else {
            break;
        }
    }
    double hf = 0.0, y1 = 0.0, y2 = 0.0;
    for (int i = tid; i < n_zeros; i += 256) {
        double s, c;
        sincos_general(gamma[i] * dd, s, c);
        hf = fma(w[i], Ap[i] * c - Bp[i] * s, hf);
    }
    block_sum3(hf, y1, y2, red);
    if (tid == 0) {
        h_unpol[p] = h0;
        if (fabs(h0) > fabs(hf)) { h_pol[p] = h0; ddy_pol[p] = 0.0; }
        else { h_pol[p] = hf; ddy_pol[p] = dd; }
    }
}

namespace k2impl {
hipError_t launch_ref_polish(int n_peaks, int n_zeros, const double* A, const double* B,
                             const double* gamma, const double* w, double half, int steps,
                             double* h_unpol, double* h_pol, double* ddy, hipStream_t st) {
    hipLaunchKernelGGL(k2_ref_polish_kernel, dim3(n_peaks), dim3(256), 0, st, n_zeros, A, B, gamma,
                       w, half, steps, h_unpol, h_pol, ddy);
    return hipGetLastError();
}
}  // namespace k2impl
