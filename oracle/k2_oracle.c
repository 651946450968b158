/*
 * CPU oracle for the K2 evaluator -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
 * and only as the checker or as the timed CPU baseline.  The product path (the HIP library)
 * never calls it.
 *
 * It restates, in plain C, the arithmetic of the reference CPU evaluator
 * search/path1_eval_fast.cpp (reference repo) and extends it to the canonical encoding:
 *
 *   K2(v)   = 2 sum_j e^{-k g_j^2} Re(w_j e^{i g_j v})          (certification/certify_K2.py:116-118)
 *   K2'(v)  = 2 sum_j e^{-k g_j^2} g_j Re(i w_j e^{i g_j v})
 *   K2''(v) = -2 sum_j e^{-k g_j^2} g_j^2 Re(w_j e^{i g_j v})
 *
 * with g_j = 2 pi G_j / 2^110 (G_j: Q18.110 word) and v = V / 2^64 (V: 64.64 word).  The phase
 * frac(g_j v / 2 pi) = frac(G_j V / 2^174) is formed exactly in integers; the 64-bit phase word
 * is W = floor(G V / 2^110) mod 2^64 (the analogue of the reference's `top64`,
 * search/path1_eval_fast.cpp:45-46).
 *
 * Three evaluation tiers:
 *   k2o_phase_words      exact integer phase words (bit-exact target of the GPU)
 *   k2o_eval_f64         FP64 + libm cos/sin, sequential sum per candidate: the restatement of
 *                        compute_peak_AB + the weighted sum h0 = sum w A (path1_eval_fast.cpp:35-50,
 *                        77-81) fused, used as the CPU baseline ("port")
 *   k2o_eval_quad        binary128 (__float128, 113-bit) evaluation from the exact 174-bit phase
 *                        fraction: the multiprecision accuracy oracle standing in for the Arb
 *                        sum of certification/certify_K2.py:112-121 (python-flint is absent)
 * plus reference-format restatements used to cross-check against the compiled reference:
 *   k2o_ref_compute_peak_AB           path1_eval_fast.cpp:15-53
 *   k2o_ref_polish_peaks_with_weights path1_eval_fast.cpp:58-143
 * and the canonical Newton polish k2o_polish_f64 (Newton rule of path1_eval_fast.cpp:105-129,
 * 8 steps, re-evaluated exactly at each new 64.64 point).
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <string.h>
#include <omp.h>

typedef unsigned __int128 u128;

#define K2O_VALUE 1u
#define K2O_D1 2u
#define K2O_D2 4u

/* bits [0,64), [64,128) and [128,192) of G*V for 128-bit G and V (bits >= 192 dropped) */
static inline void prod_low192(uint64_t glo, uint64_t ghi, uint64_t vlo, uint64_t vhi,
                               uint64_t *w0, uint64_t *w1, uint64_t *w2) {
    u128 p00 = (u128)glo * vlo;
    u128 p01 = (u128)glo * vhi;
    u128 p10 = (u128)ghi * vlo;
    u128 p11 = (u128)ghi * vhi;
    u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;           /* < 3 * 2^64 */
    u128 top = (mid >> 64) + (p01 >> 64) + (p10 >> 64) + (uint64_t)p11;
    *w0 = (uint64_t)p00;
    *w1 = (uint64_t)mid;
    *w2 = (uint64_t)top;
}

/* W = floor(G V / 2^110) mod 2^64 */
uint64_t k2o_phase_word(uint64_t glo, uint64_t ghi, uint64_t vlo, uint64_t vhi) {
    uint64_t w0, w1, w2;
    prod_low192(glo, ghi, vlo, vhi, &w0, &w1, &w2);
    (void)w0;
    return (w1 >> 46) | (w2 << 18);
}

/* out[c * nz + j] = W(G_{z0+j}, V_c) */
void k2o_phase_words(int nc, const uint64_t *V, int nz, const uint64_t *G, uint64_t *out) {
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < nc; ++c)
        for (int j = 0; j < nz; ++j)
            out[(size_t)c * nz + j] =
                k2o_phase_word(G[2 * j], G[2 * j + 1], V[2 * c], V[2 * c + 1]);
}

/* FP64 tier: theta = W * 2pi / 2^64 exactly as the reference forms `th`
 * (path1_eval_fast.cpp:32,47, with dy0 = 0 since V carries the sub-grid offset), libm cos/sin,
 * sequential accumulation over zeros in table order (path1_eval_fast.cpp:77-81). */
void k2o_eval_f64(int nz, const uint64_t *G, const double *gamma, const double *w,
                  int nc, const uint64_t *V, double k, unsigned flags,
                  double *K, double *dK, double *d2K, int nthreads) {
    const double scale = 2.0 * M_PI * ldexp(1.0, -64);
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int c = 0; c < nc; ++c) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int j = 0; j < nz; ++j) {
            uint64_t W = k2o_phase_word(G[2 * j], G[2 * j + 1], V[2 * c], V[2 * c + 1]);
            double th = (double)W * scale;
            double cs = cos(th), sn = sin(th);
            double g = gamma[j];
            double a = 2.0 * exp(-k * g * g);
            double re = w[2 * j], im = w[2 * j + 1];
            double val = a * (re * cs - im * sn);
            s0 += val;
            if (flags & K2O_D1) s1 += a * g * (-re * sn - im * cs);
            if (flags & K2O_D2) s2 += -g * g * val;
        }
        if (K) K[c] = s0;
        if (dK) dK[c] = s1;
        if (d2K) d2K[c] = s2;
    }
}

/* exact phase fraction frac(G V / 2^174) rounded to binary128 */
static inline __float128 phase_frac_q(uint64_t glo, uint64_t ghi, uint64_t vlo, uint64_t vhi) {
    uint64_t w0, w1, w2;
    prod_low192(glo, ghi, vlo, vhi, &w0, &w1, &w2);
    uint64_t W = (w1 >> 46) | (w2 << 18);
    uint64_t r46 = w1 & ((1ull << 46) - 1);
    __float128 f = (__float128)W;
    f += ldexpq((__float128)r46, -46);
    f += ldexpq((__float128)w0, -110);
    return ldexpq(f, -64);
}

/* binary128 tier; the Gaussian weight uses the exact gamma = 2 pi G / 2^110. */
void k2o_eval_quad(int nz, const uint64_t *G, const double *w, int nc, const uint64_t *V,
                   double k, unsigned flags, double *K, double *dK, double *d2K,
                   double *absS, int nthreads) {
    const __float128 twopi = 2 * M_PIq;
    const __float128 kq = (__float128)k;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int c = 0; c < nc; ++c) {
        __float128 s0 = 0, s1 = 0, s2 = 0, sa = 0;
        for (int j = 0; j < nz; ++j) {
            __float128 g = twopi * ldexpq((__float128)G[2 * j + 1], -46) +
                           twopi * ldexpq((__float128)G[2 * j], -110);
            __float128 a = 2 * expq(-kq * g * g);
            __float128 th = twopi * phase_frac_q(G[2 * j], G[2 * j + 1], V[2 * c], V[2 * c + 1]);
            __float128 cs, sn;
            sincosq(th, &sn, &cs);
            __float128 re = (__float128)w[2 * j], im = (__float128)w[2 * j + 1];
            __float128 val = a * (re * cs - im * sn);
            s0 += val;
            if (flags & K2O_D1) s1 += a * g * (-re * sn - im * cs);
            if (flags & K2O_D2) s2 -= g * g * val;
            sa += a * sqrtq(re * re + im * im);
        }
        if (K) K[c] = (double)s0;
        if (dK) dK[c] = (double)s1;
        if (d2K) d2K[c] = (double)s2;
        if (absS) absS[c] = (double)sa;
    }
}

/* binary128 K2 returned as a (hi, lo) double pair, for finite-difference checks of the
 * derivative formulas at spacings far below double resolution. */
void k2o_eval_quad_pair(int nz, const uint64_t *G, const double *w, int nc, const uint64_t *V,
                        double k, double *K_hi, double *K_lo) {
    const __float128 twopi = 2 * M_PIq;
    const __float128 kq = (__float128)k;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < nc; ++c) {
        __float128 s0 = 0;
        for (int j = 0; j < nz; ++j) {
            __float128 g = twopi * ldexpq((__float128)G[2 * j + 1], -46) +
                           twopi * ldexpq((__float128)G[2 * j], -110);
            __float128 a = 2 * expq(-kq * g * g);
            __float128 th = twopi * phase_frac_q(G[2 * j], G[2 * j + 1], V[2 * c], V[2 * c + 1]);
            __float128 cs, sn;
            sincosq(th, &sn, &cs);
            s0 += a * ((__float128)w[2 * j] * cs - (__float128)w[2 * j + 1] * sn);
        }
        double hi = (double)s0;
        K_hi[c] = hi;
        K_lo[c] = (double)(s0 - (__float128)hi);
    }
}

/* ---------------------------------------------------------------------------------------
 * Newton polish on the 64.64 point (canonical form).  Rule of path1_eval_fast.cpp:105-129:
 * step = K'/K'' clamped to +-clamp, v -= step, stop when |step| < 1e-13, require |K''| > 1e-12.
 * Each iteration re-evaluates K, K', K'' exactly at the new point.  The final value replaces the
 * start only when it is larger in magnitude (path1_eval_fast.cpp:135-141).
 * ------------------------------------------------------------------------------------- */
static void add_step(uint64_t *vlo, uint64_t *vhi, double step) {
    /* V -= rint(step * 2^64); |step| <= clamp << 2^-1, so the increment fits in int64 */
    __int128 d = (__int128)llrint(ldexp(step, 64));
    u128 x = ((u128)*vhi << 64) | *vlo;
    x -= (u128)d;
    *vlo = (uint64_t)x;
    *vhi = (uint64_t)(x >> 64);
}

void k2o_polish_f64(int nz, const uint64_t *G, const double *gamma, const double *w,
                    int nc, const uint64_t *V0, double k, int n_steps, double clamp,
                    uint64_t *V_out, double *K_start, double *K_final, double *dK_final,
                    int *steps_taken) {
    for (int c = 0; c < nc; ++c) {
        uint64_t vlo = V0[2 * c], vhi = V0[2 * c + 1];
        double f, fp, fpp;
        uint64_t vv[2] = {vlo, vhi};
        k2o_eval_f64(nz, G, gamma, w, 1, vv, k, 7u, &f, &fp, &fpp, 0);
        double f0 = f;
        int it = 0;
        for (; it < n_steps; ++it) {
            if (!(fabs(fpp) > 1e-12)) break;
            double step = fp / fpp;
            if (step > clamp) step = clamp;
            if (step < -clamp) step = -clamp;
            add_step(&vlo, &vhi, step);
            vv[0] = vlo; vv[1] = vhi;
            k2o_eval_f64(nz, G, gamma, w, 1, vv, k, 7u, &f, &fp, &fpp, 0);
            if (fabs(step) < 1e-13) { ++it; break; }
        }
        if (fabs(f0) > fabs(f)) {
            V_out[2 * c] = V0[2 * c]; V_out[2 * c + 1] = V0[2 * c + 1];
            K_final[c] = f0;
        } else {
            V_out[2 * c] = vlo; V_out[2 * c + 1] = vhi;
            K_final[c] = f;
        }
        K_start[c] = f0;
        dK_final[c] = fp;
        steps_taken[c] = it;
    }
}

/* ---------------------------------------------------------------------------------------
 * Reference-format restatements (inputs exactly as the reference's ctypes callers pass them).
 * ------------------------------------------------------------------------------------- */

/* path1_eval_fast.cpp:15-53: x = x_hi:x_lo in Q98.30, split at bit 48; phase word
 * ph = xh*ghi + xl*glo - psi mod 2^128; th = top64 * 2pi/2^64 + gamma * dy0. */
void k2o_ref_compute_peak_AB(int n_peaks, const uint64_t *x_lo, const uint64_t *x_hi,
                             const double *dy0, int n_zeros, const double *alpha,
                             const double *gamma, const uint64_t *ghi_lo, const uint64_t *ghi_hi,
                             const uint64_t *glo_lo, const uint64_t *glo_hi,
                             const uint64_t *psi_lo, const uint64_t *psi_hi,
                             double *out_A, double *out_B) {
    const double scale = 2.0 * M_PI * ldexp(1.0, -64);
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < n_peaks; ++p) {
        u128 x = ((u128)x_hi[p] << 64) | x_lo[p];
        uint64_t xh = (uint64_t)(x >> 48);
        uint64_t xl = (uint64_t)(x & ((1ull << 48) - 1));
        double d = dy0[p];
        for (int i = 0; i < n_zeros; ++i) {
            u128 gh = ((u128)ghi_hi[i] << 64) | ghi_lo[i];
            u128 gl = ((u128)glo_hi[i] << 64) | glo_lo[i];
            u128 ps = ((u128)psi_hi[i] << 64) | psi_lo[i];
            u128 ph = (u128)xh * gh + (u128)xl * gl - ps;
            double th = (double)(uint64_t)(ph >> 64) * scale + gamma[i] * d;
            double ta = 2.0 * alpha[i];
            out_A[(size_t)p * n_zeros + i] = ta * cos(th);
            out_B[(size_t)p * n_zeros + i] = ta * sin(th);
        }
    }
}

/* path1_eval_fast.cpp:58-143 */
void k2o_ref_polish_peaks_with_weights(int n_peaks, int n_zeros, const double *A, const double *B,
                                       const double *gamma, const double *w,
                                       double grid_half_width, int grid_steps,
                                       double *out_h_unpol, double *out_h_pol, double *out_ddy_pol) {
    #pragma omp parallel for schedule(dynamic, 4)
    for (int p = 0; p < n_peaks; ++p) {
        const double *Ap = A + (size_t)p * n_zeros;
        const double *Bp = B + (size_t)p * n_zeros;
        double h0 = 0.0;
        for (int i = 0; i < n_zeros; ++i) h0 += w[i] * Ap[i];
        out_h_unpol[p] = h0;
        double best = 0.0, best_abs = fabs(h0);
        int ks = n_zeros < 4096 ? n_zeros : 4096;
        if (grid_steps > 0) {
            for (int m = -grid_steps; m <= grid_steps; ++m) {
                double dd = (grid_half_width * m) / grid_steps;
                double hs = 0.0;
                for (int i = 0; i < ks; ++i) {
                    double th = gamma[i] * dd;
                    hs += w[i] * (Ap[i] * cos(th) - Bp[i] * sin(th));
                }
                if (fabs(hs) > best_abs) { best_abs = fabs(hs); best = dd; }
            }
        }
        double dd = best;
        for (int it = 0; it < 6; ++it) {
            double f = 0.0, fp = 0.0, fpp = 0.0;
            for (int i = 0; i < n_zeros; ++i) {
                double g = gamma[i], th = g * dd;
                double c = cos(th), s = sin(th);
                double val = Ap[i] * c - Bp[i] * s;
                double dval = g * (-Ap[i] * s - Bp[i] * c);
                double ddval = -g * g * val;
                f += w[i] * val;
                fp += w[i] * dval;
                fpp += w[i] * ddval;
            }
            (void)f;
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
        double hf = 0.0;
        for (int i = 0; i < n_zeros; ++i) {
            double th = gamma[i] * dd;
            hf += w[i] * (Ap[i] * cos(th) - Bp[i] * sin(th));
        }
        if (fabs(h0) > fabs(hf)) { out_h_pol[p] = h0; out_ddy_pol[p] = 0.0; }
        else { out_h_pol[p] = hf; out_ddy_pol[p] = dd; }
    }
}

int k2o_max_threads(void) { return omp_get_max_threads(); }
