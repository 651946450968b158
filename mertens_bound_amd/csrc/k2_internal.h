// Internal declarations shared by the kernels (k2_kernels.hip) and the C ABI (k2_api.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/k2eval.h"

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
    uint64_t last_use = 0;
};

struct LaunchArgs {
    int nc = 4;
    unsigned flags = 0;
    int n_blocks = 0;
    const uint4* zg = nullptr;
    const double2* za = nullptr;
    const double2* zb = nullptr;
    const double2* ze = nullptr;
    const float2* zaf = nullptr;
    const float2* zbf = nullptr;
    const float2* zef = nullptr;
    int z_begin = 0, z_end = 0, zeros_per_split = 0;
    const uint4* cand = nullptr;
    int n_cand = 0, n_ctiles = 0;
    double* partial = nullptr;
    int split_base = 0, n_cand_pad = 0;
    const double2* tab64 = nullptr;   // (cos, sin)(2 pi k / K2_TAB)
    const float2* tab32 = nullptr;
};

// host-side (cos, sin)(2 pi k / n) tables, rounded from long double
void sincos_tables(int n, double2* t64, float2* t32);

hipError_t launch_eval(const LaunchArgs& a, bool mixed, hipStream_t st);
hipError_t launch_weights(const double* gamma, const double* w, int n, int n_pad, double k,
                          WeightArrays& out, hipStream_t st);
hipError_t launch_reduce(const double* partial, int n_splits, int n_cand, int n_cand_pad,
                         unsigned flags, double* K, double* dK, double* d2K, hipStream_t st);
hipError_t launch_phase_words(const uint4* zg, int z0, int nz, const uint4* cand, int n_cand,
                              int mode, uint64_t* out, hipStream_t st);
hipError_t launch_sincos_words(const uint64_t* W, int n, int mode, const double2* tab64,
                               const float2* tab32, double* out, hipStream_t st);
hipError_t launch_newton(uint4* cand, const double* dK, const double* d2K, int n, double clamp,
                         int* active, int* steps, hipStream_t st);
hipError_t launch_ref_peak_ab(int n_peaks, const uint64_t* const* p, const double* dy0,
                              int n_zeros, const double* alpha, const double* gamma, double* A,
                              double* B, hipStream_t st);
hipError_t launch_ref_polish(int n_peaks, int n_zeros, const double* A, const double* B,
                             const double* gamma, const double* w, double half, int steps,
                             double* h_unpol, double* h_pol, double* ddy, hipStream_t st);

}  // namespace k2impl
