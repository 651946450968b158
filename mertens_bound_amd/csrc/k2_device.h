// Device-side arithmetic shared by the K2 kernels (gfx950 / CDNA4).
//
// Phase: W = floor(G * V / 2^110) mod 2^64 for G (gamma/2pi, Q18.110) and V (v, 64.64), both
// 128-bit words split into 32-bit limbs g0..g3 / v0..v3.  The product is formed with 32x32->64
// multiply-adds on the VALU (v_mad_u64_u32).  Limb products in columns 0 and 1 (bit weights
// 2^0 and 2^32) sum to less than 2^98 and can only move W by a carry into bit 110, which needs
// bits [98,110) of the rest of the product to be all ones; in that case (probability 2^-12) the
// carry is formed exactly from the three products of columns 0-1, so W is bit-exact for every
// input.
//
// sincos: the phase is in turns, so the split into a table point and a small remainder comes
// straight from the integer word, with no floating-point range reduction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace k2dev {

struct Limbs4 { uint32_t x0, x1, x2, x3; };

__device__ __forceinline__ uint64_t mul32(uint32_t a, uint32_t b) { return (uint64_t)a * b; }

// The multiply-add helpers below take b from an SGPR when SB is set: the Z tiles' candidate limbs
// are the same in every lane and live in SGPRs there, which frees VGPRs for more waves (a VALU
// instruction reads at most one SGPR, and these read only b).

// (a mod 2^24)(b mod 2^24) + c mod 2^32 (v_mad_u32_u24, full rate).  Written as asm: the
// compiler otherwise drops the operand masks when only low result bits are used and falls back
// to the quarter-rate v_mul_lo_u32.
template <bool SB = false>
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    if constexpr (SB) asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    else asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// a*b + c over 64 bits with the carry out of bit 64 added to k (v_mad_u64_u32's carry-out,
// then v_addc_co_u32; a VALU carry-out feeds the next VALU's carry-in without wait states, as in
// the compiler's own 64-bit add sequences).
template <bool SB = false>
__device__ __forceinline__ uint64_t mad_carry(uint32_t a, uint32_t b, uint64_t c, uint32_t& k) {
    uint64_t r;
    if constexpr (SB)
        asm("v_mad_u64_u32 %0, vcc, %2, %3, %4\n\t"
            "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
            : "=&v"(r), "+v"(k) : "v"(a), "s"(b), "v"(c) : "vcc");
    else
        asm("v_mad_u64_u32 %0, vcc, %2, %3, %4\n\t"
            "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
            : "=&v"(r), "+v"(k) : "v"(a), "v"(b), "v"(c) : "vcc");
    return r;
}

// a*b + c over 64 bits, returning the carry out of bit 64 in k (first sum of a column)
template <bool SB = false>
__device__ __forceinline__ uint64_t mad_carry0(uint32_t a, uint32_t b, uint64_t c, uint32_t& k) {
    uint64_t r;
    if constexpr (SB)
        asm("v_mad_u64_u32 %0, vcc, %2, %3, %4\n\t"
            "v_cndmask_b32_e64 %1, 0, 1, vcc"
            : "=&v"(r), "=v"(k) : "v"(a), "s"(b), "v"(c) : "vcc");
    else
        asm("v_mad_u64_u32 %0, vcc, %2, %3, %4\n\t"
            "v_cndmask_b32_e64 %1, 0, 1, vcc"
            : "=&v"(r), "=v"(k) : "v"(a), "v"(b), "v"(c) : "vcc");
    return r;
}

// a*b + c mod 2^64 (v_mad_u64_u32; as asm so that a chain of them keeps its order: the compiler
// would sum the products apart and add the incoming word with one more 64-bit add)
template <bool SB = false>
__device__ __forceinline__ uint64_t mad64(uint32_t a, uint32_t b, uint64_t c) {
    uint64_t r;
    if constexpr (SB)
        asm("v_mad_u64_u32 %0, vcc, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c) : "vcc");
    else
        asm("v_mad_u64_u32 %0, vcc, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c) : "vcc");
    return r;
}

// as mad_carry0, with the carry count starting at B (an inline constant, 0..63): k = B + carry
template <int B, bool SB = false>
__device__ __forceinline__ uint64_t mad_carry0_from(uint32_t a, uint32_t b, uint64_t c,
                                                    uint32_t& k) {
    static_assert(B >= 0 && B < 64, "inline constant");
    uint64_t r;
    if constexpr (SB)
        asm("v_mad_u64_u32 %0, vcc, %2, %3, %4\n\t"
            "v_cndmask_b32_e64 %1, %5, %6, vcc"
            : "=&v"(r), "=v"(k) : "v"(a), "s"(b), "v"(c), "i"(B), "i"(B + 1) : "vcc");
    else
        asm("v_mad_u64_u32 %0, vcc, %2, %3, %4\n\t"
            "v_cndmask_b32_e64 %1, %5, %6, vcc"
            : "=&v"(r), "=v"(k) : "v"(a), "v"(b), "v"(c), "i"(B), "i"(B + 1) : "vcc");
    return r;
}

// exact W via a straightforward 256-bit schoolbook product (the debug entry's reference path)
__device__ __noinline__ uint64_t phase_word_exact(Limbs4 g, Limbs4 v) {
    const uint32_t G[4] = {g.x0, g.x1, g.x2, g.x3};
    const uint32_t V[4] = {v.x0, v.x1, v.x2, v.x3};
    uint32_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < 4; ++j) {
            uint64_t t = mul32(G[i], V[j]) + r[i + j] + carry;
            r[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        r[i + 4] = (uint32_t)carry;
    }
    // bits [110,174): r[3] >> 14 | r[4] << 18 | r[5] << 50
    uint64_t lo = ((uint64_t)r[4] << 18) | (r[3] >> 14);
    uint64_t hi = (uint64_t)r[5] << 50;
    return lo | hi;
}

// Remainder words of W for sincos_tab64_at: r = W bits [2,54) as rhi:rlo (rhi bits [0,20) are
// W bits [34,54); its higher bits are not used) and off = 16 (W >> 54), the byte offset of the
// 16-byte table entry of the interval holding the phase
struct PhaseRem {
    uint32_t rlo, rhi, off;
};

__device__ __forceinline__ PhaseRem phase_rem_of(uint64_t W) {
    return PhaseRem{(uint32_t)(W >> 2), (uint32_t)(W >> 34), (uint32_t)(W >> 50) & 0x3FF0u};
}

// Column 5 (weight 2^160) is needed mod 2^14 only: g2 v3 + g3 v2 mod 2^16 is the dot product of
// the limbs' low halves packed in pairs, one v_dot2_u32_u16 (each 16x16-bit product exact mod
// 2^16) in place of two 24-bit multiply-adds.  pack_lo16(x, y) = (x mod 2^16) | (y << 16); a zero
// packs (g2, g3), a candidate (v3, v2).  The FP32 window uses it (3% off the FP32 tail's time,
// the packed zero word staged in the tile), and so does the one-launch FP64 C tile
// (phase_columns_dp, the packed word staged with each zero); packing each zero inside an FP64
// loop instead measured no faster than the two multiply-adds.
__device__ __forceinline__ uint32_t pack_lo16(uint32_t x, uint32_t y) {
    return __builtin_amdgcn_perm(y, x, 0x05040100u);
}

__device__ __forceinline__ uint32_t dot2_lo16(uint32_t gp, uint32_t vp, uint32_t c) {
    typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
    return __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, gp), __builtin_bit_cast(u16x2, vp), c,
                                  false);
}

// The product columns 2..5 of G V (columns 0-1 skipped): s3lo, s4lo, s5 are bits [96,128),
// [128,160) and [160,174) (s5 bits above 14 are not formed exactly).  In the carry case
// (phase_carry_case: bits [98,110) all ones) W must come from phase_word_exact instead.
template <bool SB = false>
__device__ __forceinline__ void phase_columns(Limbs4 g, Limbs4 v, uint32_t& s3lo, uint32_t& s4lo,
                                              uint32_t& s5, uint32_t& tlo) {
    // column 2 (weight 2^64): g0 v2 + g1 v1 + g2 v0; a carry out of a 64-bit add shows as the
    // new sum being below the old one
    uint32_t k2;
    uint64_t t = mul32(g.x0, v.x2);
    t = mad_carry0<SB>(g.x1, v.x1, t, k2);
    t = mad_carry<SB>(g.x2, v.x0, t, k2);
    tlo = (uint32_t)t;   // product bits [64,96), for the carry fix-up
    // column 3 (2^96): carry-in hi32(t) + k2 2^32
    uint64_t s3 = (t >> 32) | ((uint64_t)k2 << 32);
    uint32_t k3;
    s3 = mad_carry0<SB>(g.x0, v.x3, s3, k3);
    s3 = mad_carry<SB>(g.x1, v.x2, s3, k3);
    s3 = mad_carry<SB>(g.x2, v.x1, s3, k3);
    s3 = mad_carry<SB>(g.x3, v.x0, s3, k3);
    // column 4 (2^128): only bits [0,46) are needed, so wrap mod 2^64; chained onto the
    // column-3 carry word
    uint64_t s4 = (s3 >> 32) | ((uint64_t)k3 << 32);
    s4 = mad64<SB>(g.x1, v.x3, s4);
    s4 = mad64<SB>(g.x2, v.x2, s4);
    s4 = mad64<SB>(g.x3, v.x1, s4);
    // column 5 (2^160): only bits [0,14) are needed; 24x24-bit products are exact mod 2^24
    s5 = (uint32_t)(s4 >> 32);
    s5 = mad24<SB>(g.x2, v.x3, s5);
    s5 = mad24<SB>(g.x3, v.x2, s5);
    s3lo = (uint32_t)s3;
    s4lo = (uint32_t)s4;
}

__device__ __forceinline__ bool phase_carry_case(uint32_t s3lo) {
    return ((s3lo >> 2) & 0xFFFu) == 0xFFFu;
}

// The carry that columns 0-1, T01 = g0 v0 + (g0 v1 + g1 v0) 2^32 < 2^98, add at bit 110 of the
// partial product P (columns 2..5, zero below bit 64): floor((P mod 2^110 + T01) / 2^110) with
// P mod 2^110 = A 2^64, A = P bits [64,110) (tlo and s3lo bits [0,14)).  Since A is an integer,
// that is (A + floor(T01 / 2^64)) >> 46, and floor(T01 / 2^64) = (hi32(g0 v0) + g0 v1 + g1 v0)
// >> 32 (the low word of g0 v0 cannot carry past a 2^32 boundary of that integer sum).
__device__ __forceinline__ uint32_t phase_low_carry(Limbs4 g, Limbs4 v, uint32_t s3lo,
                                                    uint32_t tlo) {
    const uint64_t c0 = mul32(g.x0, v.x0), c1a = mul32(g.x0, v.x1), c1b = mul32(g.x1, v.x0);
    const uint64_t s = c1a + c1b;
    uint32_t k = s < c1a;
    const uint64_t s2 = s + (c0 >> 32);
    k += s2 < s;
    const uint64_t bint = (s2 >> 32) | ((uint64_t)k << 32);   // floor(T01 / 2^64) < 2^35
    const uint64_t A = ((uint64_t)(s3lo & 0x3FFFu) << 32) | tlo;
    return (uint32_t)((A + bint) >> 46);
}

// W with columns 0-1 skipped and the carry case fixed up exactly
template <bool SB = false>
__device__ __forceinline__ uint64_t phase_word(Limbs4 g, Limbs4 v) {
    uint32_t s3lo, s4lo, s5, tlo;
    phase_columns<SB>(g, v, s3lo, s4lo, s5, tlo);
    const uint32_t wlo = __builtin_amdgcn_alignbit(s4lo, s3lo, 14);
    const uint32_t whi = __builtin_amdgcn_alignbit(s5, s4lo, 14);
    uint64_t W = ((uint64_t)whi << 32) | wlo;
    if (__builtin_expect(phase_carry_case(s3lo), 0)) W += phase_low_carry(g, v, s3lo, tlo);
    return W;
}

// The remainder words of W (phase_rem_of) straight from the columns: W bits [2,54) are product
// bits [112,164), and W bits [54,64) are bits [4,14) of the column-5 word, so off = s5 & 0x3FF0
template <bool SB = false>
__device__ __forceinline__ PhaseRem phase_rem(Limbs4 g, Limbs4 v) {
    uint32_t s3lo, s4lo, s5, tlo;
    phase_columns<SB>(g, v, s3lo, s4lo, s5, tlo);
    PhaseRem r{__builtin_amdgcn_alignbit(s4lo, s3lo, 16), __builtin_amdgcn_alignbit(s5, s4lo, 16),
               s5 & 0x3FF0u};
    if (__builtin_expect(phase_carry_case(s3lo), 0)) {
        const uint64_t W = ((uint64_t)__builtin_amdgcn_alignbit(s5, s4lo, 14) << 32 |
                            __builtin_amdgcn_alignbit(s4lo, s3lo, 14)) +
                           phase_low_carry(g, v, s3lo, tlo);
        r = phase_rem_of(W);
    }
    return r;
}

// The columns of phase_columns with column 5 as one packed dot product: gp packs the zero's
// (g2, g3) and vp the candidate's (v3, v2) low halves (pack_lo16), and dot2_lo16 gives
// g2 v3 + g3 v2 mod 2^16 added to the column-4 carry word (bits [0,14) are needed).  The C tile
// stages gp with each zero, so the pair costs one instruction per evaluation instead of two.
// SG: the zero's limbs (and gp) are uniform over the wave and sit in scalar registers; they
// take the multiply-adds' scalar operand slot.
template <bool SG = false>
__device__ __forceinline__ void phase_columns_dp(Limbs4 g, uint32_t gp, Limbs4 v, uint32_t vp,
                                                 uint32_t& s3lo, uint32_t& s4lo, uint32_t& s5,
                                                 uint32_t& tlo) {
    uint32_t k2;
    uint64_t t = mul32(g.x0, v.x2);
    t = mad_carry0<SG>(v.x1, g.x1, t, k2);
    t = mad_carry<SG>(v.x0, g.x2, t, k2);
    tlo = (uint32_t)t;
    uint64_t s3 = (t >> 32) | ((uint64_t)k2 << 32);
    uint32_t k3;
    s3 = mad_carry0<SG>(v.x3, g.x0, s3, k3);
    s3 = mad_carry<SG>(v.x2, g.x1, s3, k3);
    s3 = mad_carry<SG>(v.x1, g.x2, s3, k3);
    s3 = mad_carry<SG>(v.x0, g.x3, s3, k3);
    uint64_t s4 = (s3 >> 32) | ((uint64_t)k3 << 32);
    s4 = mad64<SG>(v.x3, g.x1, s4);
    s4 = mad64<SG>(v.x2, g.x2, s4);
    s4 = mad64<SG>(v.x1, g.x3, s4);
    s5 = dot2_lo16(gp, vp, (uint32_t)(s4 >> 32));
    s3lo = (uint32_t)s3;
    s4lo = (uint32_t)s4;
}

// pack_lo16(x, y) on the scalar unit, for values uniform over the wave
__device__ __forceinline__ uint32_t pack_lo16_s(uint32_t x, uint32_t y) {
    uint32_t r;
    asm("s_pack_ll_b32_b16 %0, %1, %2" : "=s"(r) : "s"(x), "s"(y));
    return r;
}

// the split table's remainder words straight from the columns: W bits [1,33) are product bits
// [111,143), W bits [33,53) bits [143,163), and 8 (W bits [53,64)) = s5 & 0x3FF8
__device__ __forceinline__ PhaseRem phase_rem_w_cols(uint32_t s3lo, uint32_t s4lo, uint32_t s5) {
    return PhaseRem{__builtin_amdgcn_alignbit(s4lo, s3lo, 15),
                    __builtin_amdgcn_alignbit(s5, s4lo, 15), s5 & 0x3FF8u};
}

// The carry case for the split table's words, which drop W bit 0: a carry from columns 0-1
// (at most one unit at bit 98) changes W bits [1,64) only if product bits [98,111) are all ones
// (probability 2^-13, half the rate of phase_carry_case).
__device__ __forceinline__ bool phase_carry_case_w(uint32_t s3lo) {
    return (s3lo & 0x7FFCu) == 0x7FFCu;
}

// the carry case's exact W from the columns (phase_carry_case holds)
__device__ __forceinline__ uint64_t phase_word_fixed(Limbs4 g, Limbs4 v, uint32_t s3lo,
                                                     uint32_t s4lo, uint32_t s5, uint32_t tlo) {
    return ((uint64_t)__builtin_amdgcn_alignbit(s5, s4lo, 14) << 32 |
            __builtin_amdgcn_alignbit(s4lo, s3lo, 14)) +
           phase_low_carry(g, v, s3lo, tlo);
}

// ---- FP32 form: the phase window ------------------------------------------------------------
// The FP32 kernels need the phase to about 2^-40 turn, so they form only the product columns
// 3..5 (bit weights 2^96..2^160).  Columns 0..2 sum to less than 2^130 and move the phase
// (bits [0,174) of G V, in units of 2^-174 turn) by less than 2^-44 turn; the integer turns are
// still removed exactly, because the window keeps bits [128,174) of the partial product, which is
// the phase mod 1.  hi:lo = bits [160,192) : [128,160) of the partial product plus 2^(173-TB),
// half a step of a 2^TB-point table, so that the top TB phase bits of the window round to the
// nearest table point.  Only hi bits [0,24) are formed (the column-5 terms are 24x24-bit
// products, exact mod 2^24).
template <int TB, bool SB = false>
__device__ __forceinline__ void phase_window(Limbs4 g, Limbs4 v, uint32_t& hi, uint32_t& lo) {
    static_assert(TB >= 9 && TB <= 11, "table of 512..2048 points");
    uint64_t s3 = mul32(g.x0, v.x3);
    uint32_t k3;   // column-3 carries, starting at half a table step: 2^(173-TB) = 2^(13-TB) 2^160
    s3 = mad_carry0_from<(1 << (13 - TB)), SB>(g.x1, v.x2, s3, k3);
    s3 = mad_carry<SB>(g.x2, v.x1, s3, k3);
    s3 = mad_carry<SB>(g.x3, v.x0, s3, k3);
    uint64_t s4 = (s3 >> 32) | ((uint64_t)k3 << 32);
    s4 = mad64<SB>(g.x1, v.x3, s4);
    s4 = mad64<SB>(g.x2, v.x2, s4);
    s4 = mad64<SB>(g.x3, v.x1, s4);
    uint32_t s5 = (uint32_t)(s4 >> 32);
    s5 = mad24<SB>(g.x2, v.x3, s5);
    s5 = mad24<SB>(g.x3, v.x2, s5);
    hi = s5;
    lo = (uint32_t)s4;
}

// The FP32 kernels' window: columns 4..5 exact, column 3 in FP32 beside them.  Column 3 adds
// S3 2^96 with S3 < 2^66, less than 2^-12 turn: it never wraps, so it needs no exact integer sum,
// only precision well below the FP32 kernels' 2^-35-turn remainder.  gf holds g0..g3 times
// 2 pi / 2^44 as floats (rounded once from double on the host), vf holds v0..v3 times 2^-34 as
// floats (window_vf), so that four FP32 FMAs give column 3 straight in the remainder's units,
// radians (S3 2^96 / 2^174 turn), and the sum starts from minus three quarters of a table step:
// t3 = 2 pi S3 / 2^78 - 3/4 step, and sincos_window_f forms the remainder x = R step + t3 in one
// FMA.  Error of t3: 2 2^-24 relative on S3 from the limbs' roundings (2.0e3 units of 2^-46
// turn; each limb's own rounding, so no common bias), and the four FMAs' roundings at
// |t| < 2^-8 rad, 2^-33 rad each (5.2e3 units): with columns 0-2, at most 7.3e3 units,
// 6.5e-10 rad, ninety times below FP32 rounding of sin/cos.  hi:lo = columns 4..5 plus three quarters of a table step,
// 3 2^(172-TB): the table index rounds there because column 3 adds up to half a step, so that
// the remainder with it stays within three quarters of a step either side.  Column 5 is one
// packed dot product (dot2_lo16 of gp, vp); g.x0 is not read (the FP32 C tile stages gp in its
// place).  Against the exact column 3 this takes four FP32 FMAs (half-cost instructions), where
// that takes four 64-bit multiply-adds and five carry and register moves.
template <int TB>
struct WindowF {
    static constexpr float step = 6.28318530717958647692f / (float)(1ull << (32 + TB));
    static constexpr float tq = step * 3221225472.0f;   // 3/4 of a table step
};

template <int TB, bool SB = false>
__device__ __forceinline__ void phase_window_f(Limbs4 g, uint32_t gp, float4 gf, Limbs4 v,
                                               uint32_t vp, float4 vf, uint32_t& hi,
                                               uint32_t& lo, float& t3) {
    static_assert(TB >= 9 && TB <= 11, "table of 512..2048 points");
    float t = fmaf(gf.w, vf.x, -WindowF<TB>::tq);   // g3 v0
    t = fmaf(gf.z, vf.y, t);                        // g2 v1
    t = fmaf(gf.y, vf.z, t);                        // g1 v2
    t = fmaf(gf.x, vf.w, t);                        // g0 v3
    t3 = t;
    uint64_t s4 = mad64<SB>(g.x1, v.x3, 3ull << (44 - TB));
    s4 = mad64<SB>(g.x2, v.x2, s4);
    s4 = mad64<SB>(g.x3, v.x1, s4);
    hi = dot2_lo16(gp, vp, (uint32_t)(s4 >> 32));   // column 5, bits [0,16)
    lo = (uint32_t)s4;
}

// per-candidate operand of phase_window_f: v0..v3 times 2^-34 as floats
__device__ __forceinline__ float4 window_vf(Limbs4 v) {
    constexpr float s = 5.8207660913467407e-11f;   // 2^-34
    return make_float4((float)v.x0 * s, (float)v.x1 * s, (float)v.x2 * s, (float)v.x3 * s);
}

// a zero's limbs times 2 pi / 2^78 as floats, each rounded once (the screen's records: there
// the zero takes the scaled side of phase_window_f and the candidate's limbs are plain floats)
__device__ __forceinline__ float4 window_gx(Limbs4 g) {
    constexpr double s = 6.283185307179586 / 302231454903657293676544.0;   // 2 pi / 2^78
    return make_float4((float)((double)g.x0 * s), (float)((double)g.x1 * s),
                       (float)((double)g.x2 * s), (float)((double)g.x3 * s));
}

// ---- FP64 sin/cos of 2 pi W / 2^64 ---------------------------------------------------------
#define K2_TWO_PI_2M64 3.4061215800865545e-19   // 2 pi / 2^64 (correctly rounded)

__device__ __forceinline__ void sincos_poly(double x, double& s, double& c) {
    const double z = x * x;
    // sin x = x + x z P(z), P = sum (-1)^n z^(n-1)/(2n+1)!, n = 1..8
    double ps = 2.8114572543455206e-15;              //  1/17!
    ps = fma(ps, z, -7.6471637318198164e-13);        // -1/15!
    ps = fma(ps, z, 1.6059043836821613e-10);         //  1/13!
    ps = fma(ps, z, -2.5052108385441720e-08);        // -1/11!
    ps = fma(ps, z, 2.7557319223985893e-06);         //  1/9!
    ps = fma(ps, z, -1.9841269841269841e-04);        // -1/7!
    ps = fma(ps, z, 8.3333333333333333e-03);         //  1/5!
    ps = fma(ps, z, -1.6666666666666666e-01);        // -1/3!
    s = fma(x * z, ps, x);
    // cos x = 1 + z Q(z), Q = sum (-1)^n z^(n-1)/(2n)!, n = 1..8
    double pc = 4.7794773323873853e-14;              //  1/16!
    pc = fma(pc, z, -1.1470745597729725e-11);        // -1/14!
    pc = fma(pc, z, 2.0876756987868099e-09);         //  1/12!
    pc = fma(pc, z, -2.7557319223985888e-07);        // -1/10!
    pc = fma(pc, z, 2.4801587301587302e-05);         //  1/8!
    pc = fma(pc, z, -1.3888888888888889e-03);        // -1/6!
    pc = fma(pc, z, 4.1666666666666664e-02);         //  1/4!
    pc = fma(pc, z, -0.5);
    c = fma(z, pc, 1.0);
}

// ---- table-driven sin/cos used by the evaluation kernels ------------------------------------
// The angle 2 pi W / 2^64 is split into the midpoint 2 pi (k + 1/2) / 1024 of its table interval
// (k = W bits [54,64)) and a remainder |x| <= pi/1024 = 3.1e-3 given by r = W bits [2,54):
// x = K u with u = r - 2^51 and K = 2 pi / 2^62 (W bits [0,2) are dropped: < 1.4e-18 rad).  The
// double with high word 0x43300000 | r bits [32,52) and low word r bits [0,32) is d = 2^52 + r
// exactly, so x = fma(d, K, -M) with M = (2^52 + 2^51) K: one v_and_or_b32 and one FMA from the
// integer word to radians, no conversions.  K and M are doubles (K2_REM_K, K2_REM_M); the table
// points are shifted by (2^52 + 2^51) K - M, the part of M lost to rounding (zero for these
// constants), so the shift is exact.  Table entries hold (cos, sin) of the midpoints (16 bytes,
// K2_TAB entries, staged in LDS: one ds_read_b128 per evaluation); sin x = x - x^3/6 + x^5/120
// (truncation < 5e-22) and cos x = 1 - x^2/2 + x^4/24 (< 1.2e-18), and one complex rotation
// combines the two (measured max error 1.5 u against 2^-53-relative rounding u = 1.1e-16, over
// 2e7 random words).  No quadrant selects, no range reduction: the split comes straight from the
// integer word.
#define K2_TAB 1024
#define K2_REM_K 1.3624486320346218e-18   // 2 pi / 2^62, rounded to double
#define K2_REM_M 0.009203884727313847     // (2^52 + 2^51) K2_REM_K, rounded to double

struct Tab64 {
    double c, s;      // cos, sin of 2 pi (k + 1/2) / K2_TAB (shifted as above)
};

// FP32 table: cos, sin of 2 pi k / K2_TAB32 (8 bytes, so that the byte offset of entry k is
// window bits shifted in place)
#define K2_TAB32_BITS 11
#define K2_TAB32 (1 << K2_TAB32_BITS)

struct Tab32 {
    float c, s;
};

// constants of sincos_tab64_at; a loop that calls it many times takes them from sincos_coef()
// once, with -M and -1/6 held in VGPRs (an FMA reads at most one scalar operand, and x and the
// sine polynomial each take two of these constants)
struct SinCosCoef {
    double k, nm, s1, s2, c2;   // K, -M, -1/6, 1/120, 1/24
};

#define K2_SINCOS_COEF {K2_REM_K, -K2_REM_M, -1.0 / 6.0, 1.0 / 120.0, 1.0 / 24.0}

__device__ __forceinline__ SinCosCoef sincos_coef() {
    SinCosCoef k K2_SINCOS_COEF;
    asm volatile("" : "+v"(k.nm), "+v"(k.s1));
    return k;
}

// sin/cos of the phase from its remainder words (phase_rem / phase_rem_of)
__device__ __forceinline__ void sincos_tab64_at(const PhaseRem& r, const Tab64* tab, double& S,
                                                double& C, const SinCosCoef& kc) {
    // (0x43300000 | (rhi & 0xFFFFF)) : rlo = 2^52 + r in one v_and_or_b32 (as asm: the compiler
    // splits it into an AND and an OR, since neither constant is an inline one)
    uint32_t hi;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(hi) : "v"(r.rhi), "s"(0xFFFFFu), "v"(0x43300000u));
    const double x = fma(__hiloint2double((int)hi, (int)r.rlo), kc.k, kc.nm);
    // the four products are FMAs with a zero addend (the same rounding): on gfx950 v_fma_f64
    // issues in about 4.4 cycles per wave against 5.0 for v_mul_f64 (0.5% off the plain kernel)
    const double z = fma(x, x, 0.0);
    const double sp = fma(fma(x, z, 0.0), fma(z, kc.s2, kc.s1), x);
    const double cp = fma(z, fma(z, kc.c2, -0.5), 1.0);
    const Tab64 t = *reinterpret_cast<const Tab64*>(reinterpret_cast<const char*>(tab) + r.off);
    C = fma(t.c, cp, fma(-t.s, sp, 0.0));
    S = fma(t.s, cp, fma(t.c, sp, 0.0));
}

__device__ __forceinline__ void sincos_tab64(uint64_t W, const Tab64* tab, double& S, double& C) {
    sincos_tab64_at(phase_rem_of(W), tab, S, C, SinCosCoef K2_SINCOS_COEF);
}

// ---- FP64 sin/cos with a 2,048-point split table (the one-launch C tile) ----------------------
// The same construction with twice the intervals: k = W bits [53,64), r = W bits [1,53) (W bit 0
// dropped: < 3.4e-19 rad), x = fma(2^52 + r, KW, -MW) with KW = 2 pi / 2^63, |x| <= pi/2048 =
// 1.53e-3.  That halves x and drops the x^5 term: sin x = x - x^3/6 (truncation x^5/120 <= 7.1e-17
// = 0.64 u, odd in x; x is centred in its interval, so it averages out over the zeros) and
// cos x = 1 - x^2/2 + x^4/24 (< 1.8e-20) -- five FMAs where the 1,024-point table takes six.
// The table holds the K2_TABW cosines, one unused double, and the K2_TABW sines of the
// midpoints (8-byte entries, so the byte offset 8k is W bits [53,64) masked in place; the two
// halves are read at that offset and 16 KiB + 8 above it, by two ds_read_b64).  Without the
// extra 8 bytes the compiler joins the two reads into one ds_read2st64_b64 whose two addresses
// per lane fall in the same LDS banks (twice the bank-conflict cycles, measured slower); 16-byte
// (cos, sin) entries read by one ds_read_b128 at the doubled offset measured 4% slower still.
// 32 KiB + 8 in LDS (K2_TABW_DOUBLES).  It sits in the FP64 table buffer after the K2_TAB
// interleaved entries (sincos_tables).
#define K2_TABW_BITS 11
#define K2_TABW (1 << K2_TABW_BITS)
#define K2_TABW_SIN (K2_TABW + 1)             // index of the first sine
#define K2_TABW_DOUBLES (2 * K2_TABW + 1)
#define K2_REMW_K 6.8122431601731090e-19   // 2 pi / 2^63, rounded to double
#define K2_REMW_M 0.0046019423636569235    // (2^52 + 2^51) K2_REMW_K, rounded to double

struct SinCosCoefW {
    double k, nm, s1, c2;   // KW, -MW, -1/6, 1/24
};

#define K2_SINCOSW_COEF {K2_REMW_K, -K2_REMW_M, -1.0 / 6.0, 1.0 / 24.0}

__device__ __forceinline__ SinCosCoefW sincos_coef_w() {
    SinCosCoefW k K2_SINCOSW_COEF;
    asm volatile("" : "+v"(k.nm), "+v"(k.s1));
    return k;
}

// the remainder words for the split table: rlo = W bits [1,33), rhi bits [0,20) = W bits
// [33,53) (higher bits unused), off = 8 * (W bits [53,64))
__device__ __forceinline__ PhaseRem phase_rem_w_of(uint64_t W) {
    return PhaseRem{(uint32_t)(W >> 1), (uint32_t)(W >> 33), (uint32_t)(W >> 50) & 0x3FF8u};
}

// tab: the split table (K2_TABW cosines, a pad, K2_TABW sines)
__device__ __forceinline__ void sincos_tabw_at(const PhaseRem& r, const double* tab, double& S,
                                               double& C, const SinCosCoefW& kc) {
    uint32_t hi;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(hi) : "v"(r.rhi), "s"(0xFFFFFu), "v"(0x43300000u));
    const double x = fma(__hiloint2double((int)hi, (int)r.rlo), kc.k, kc.nm);
    const double z = fma(x, x, 0.0);
    const double sp = fma(fma(x, z, 0.0), kc.s1, x);
    const double cp = fma(z, fma(z, kc.c2, -0.5), 1.0);
    const char* tb = reinterpret_cast<const char*>(tab) + r.off;
    const double tc = *reinterpret_cast<const double*>(tb);
    const double ts = *reinterpret_cast<const double*>(tb + K2_TABW_SIN * sizeof(double));
    C = fma(tc, cp, fma(-ts, sp, 0.0));
    S = fma(ts, cp, fma(tc, sp, 0.0));
}

__device__ __forceinline__ void sincos_tabw(uint64_t W, const double* tab, double& S, double& C) {
    sincos_tabw_at(phase_rem_w_of(W), tab, S, C, SinCosCoefW K2_SINCOSW_COEF);
}

// FP32 sin/cos of the window's phase with a 2^TB-point table (every 2^(11-TB)-th entry of the
// K2_TAB32-point table, as staged in LDS).  Window bits [46-TB,46) are the index of the nearest
// point (hi bits [14-TB,14)); the next 32 bits R give the remainder
// x = 2 pi (R - 2^31) 2^-(32+TB), |x| <= pi/2^TB, with R rounded to FP32 (error <= 2^-(24+TB)
// turn).  cos x = 1 - x^2/2 and sin x = x hold to x^4/24 and |x|^3/6 (<= 6.0e-10 at TB = 11,
// 4.8e-9 at TB = 10, against FP32 rounding of 6e-8); the x^3 error is odd in x, so it averages
// out over the zeros instead of biasing the sum.  cos(t + x) = c - x (s + c x/2) and
// sin(t + x) = s + x (c - s x/2) take four FMAs.
template <int TB>
__device__ __forceinline__ void sincos_window(uint32_t hi, uint32_t lo, const Tab32* tab,
                                              float& S, float& C) {
    constexpr uint32_t mask = ((1u << TB) - 1u) << 3;
    const uint32_t off = (TB == 11 ? hi : hi >> (11 - TB)) & mask;   // 8 * index
    const Tab32 t = *reinterpret_cast<const Tab32*>(reinterpret_cast<const char*>(tab) + off);
    const uint32_t R = __builtin_amdgcn_alignbit(hi, lo, 14 - TB);
    constexpr float step = 6.28318530717958647692f / (float)(1ull << (32 + TB));
    constexpr float half = step * 2147483648.0f;
    const float u = (float)R;
    const float x = fmaf(u, step, -half);
    const float h = fmaf(u, 0.5f * step, -0.5f * half);   // x / 2
    C = fmaf(-x, fmaf(h, t.c, t.s), t.c);
    S = fmaf(x, fmaf(-h, t.s, t.c), t.s);
}

// sin/cos of the FP32 kernels' window (phase_window_f): as sincos_window, with the index rounded
// at three quarters of a step and column 3 (t3, radians, with the offset) added to the
// remainder: x = R step + t3 = 2 pi ((R 2^-(32+TB) + S3 2^-78) - 3/4 2^-TB), one FMA,
// |x| <= 3/4 2 pi 2^-TB (2.3e-3 at TB = 11: |x|^3/6 <= 2.0e-9 and x^4/24 <= 1.2e-12, against
// FP32 rounding of 6e-8; the x^3 error is odd in x, and x is centred, so it averages out over the
// zeros).  x/2 is one multiply.
template <int TB>
__device__ __forceinline__ void sincos_window_f(uint32_t hi, uint32_t lo, float t3,
                                                const Tab32* tab, float& S, float& C) {
    constexpr uint32_t mask = ((1u << TB) - 1u) << 3;
    const uint32_t off = (TB == 11 ? hi : hi >> (11 - TB)) & mask;   // 8 * index
    const Tab32 t = *reinterpret_cast<const Tab32*>(reinterpret_cast<const char*>(tab) + off);
    const uint32_t R = __builtin_amdgcn_alignbit(hi, lo, 14 - TB);
    const float x = fmaf((float)R, WindowF<TB>::step, t3);
    const float h = 0.5f * x;
    C = fmaf(-x, fmaf(h, t.c, t.s), t.c);
    S = fmaf(x, fmaf(-h, t.s, t.c), t.s);
}

// ---- general FP64 sincos (|x| < 2^20 pi/2) for the reference-format entry points -----------
// Cody-Waite reduction by pi/2 with a three-part constant (W. J. Cody and W. Waite,
// "Software Manual for the Elementary Functions", 1980).  PIO2_1 and PIO2_2 carry 33 bits each,
// so n * PIO2_1 and n * PIO2_2 are exact for |n| < 2^20, i.e. |x| < 2^20 pi/2 ~ 1.65e6; larger
// arguments (none in the reference's use: |theta| <= 2 pi + gamma |dy0| there) go to the device
// math library's sincos, which reduces any double.
__device__ __forceinline__ void sincos_general(double x, double& s, double& c) {
    if (!(fabs(x) < 1.6e6)) {
        ::sincos(x, &s, &c);
        return;
    }
    const double PIO2_1 = 1.5707963267341256e+00;   // first 33 bits of pi/2
    const double PIO2_2 = 6.0771005063039660e-11;   // next 33 bits
    const double PIO2_3 = 2.0222662487959506e-21;   // remainder
    const double n = rint(x * 6.3661977236758138e-01);
    double r = fma(-n, PIO2_1, x);
    r = fma(-n, PIO2_2, r);
    r = fma(-n, PIO2_3, r);
    const int q = ((int)(long long)n) & 3;
    double sp, cp;
    sincos_poly(r, sp, cp);
    const double cs = (q & 1) ? -sp : cp;
    const double sn = (q & 1) ? cp : sp;
    const bool neg = (q & 2) != 0;
    c = neg ? -cs : cs;
    s = neg ? -sn : sn;
}

// ---- e^{-x} with a stated error bound (enclosure mode) ----------------------------------------
// For 0 <= x <= 708 the result is within 2.5 * 2^-53 relative of e^{-x}; larger x return 0
// (e^{-x} < 3.1e-308 there, which the caller bounds separately).  Reduction x = n ln2 + r with
// ln2 = L1 + L2, L1 carrying 32 significant bits, so n L1 is exact for n < 2^21 and x - n L1 is
// exact (Cody and Waite 1980); |r| <= ln2/2 + 2^-40.  e^{-r} = sum_k (-r)^k / k! to degree 13
// (truncation < 4.3e-18 relative), by Horner in FP64 (< 2 * 2^-53); r's rounding adds < 0.2 *
// 2^-53; scaling by 2^-n is exact above the subnormal range.  Host and device run the same
// operations (explicit fma), so host builds reproduce device results bit for bit.
__host__ __device__ inline double exp_neg_bounded(double x) {
    if (!(x <= 708.0)) return 0.0;
    const double L1 = 6.9314718036912382e-01;   // ln 2 to 32 significant bits
    const double L2 = 1.9082149292705877e-10;   // ln 2 - L1
    const double n = rint(x * 1.4426950408889634);
    const double r1 = fma(-n, L1, x);
    const double r = fma(-n, L2, r1);
    const double mr = -r;
    double p = 1.6059043836821613e-10;           // 1/13!
    p = fma(p, mr, 2.08767569878681e-09);         // 1/12!
    p = fma(p, mr, 2.505210838544172e-08);        // 1/11!
    p = fma(p, mr, 2.755731922398589e-07);        // 1/10!
    p = fma(p, mr, 2.7557319223985893e-06);       // 1/9!
    p = fma(p, mr, 2.48015873015873e-05);         // 1/8!
    p = fma(p, mr, 1.984126984126984e-04);        // 1/7!
    p = fma(p, mr, 1.388888888888889e-03);        // 1/6!
    p = fma(p, mr, 8.333333333333333e-03);        // 1/5!
    p = fma(p, mr, 4.1666666666666664e-02);       // 1/4!
    p = fma(p, mr, 1.6666666666666666e-01);       // 1/3!
    p = fma(p, mr, 0.5);
    p = fma(p, mr, 1.0);
    p = fma(p, mr, 1.0);
    return ldexp(p, -(int)n);
}

}  // namespace k2dev
