// Exact arithmetic mod q = 3329 on FP32 values, for the Kyber transforms.
//
// Every value is an integer held in a float.  A product of a value below 2^12
// and a twiddle below 2^11 fits in the 24-bit significand, and v_fma_f32 forms
// a * b + c from the exact product with a single rounding, so a product reduced
// by the right multiple of q is exact.  The multiple comes from a second fma:
// fma(x, z/q, M) with M = 1.5 * 2^23 rounds x * z / q to the nearest integer k
// (the float spacing is 1 between 2^23 and 2^24) and M is subtracted again.
// Any representative of the residue serves: the results leave this arithmetic
// through compress (below), whose value depends only on the residue.
//
// The FP32 forms issue at the fast VALU rate on gfx950 (fma / mul / add /
// sub), where the integer Montgomery and Barrett reductions take v_mul_lo_u32,
// v_mad_i32_i24 and v_bfe, which issue at about half that rate.
//
// Bounds (each is checked exhaustively on the CPU by tests/c/f32_arith_test.cpp):
//   mulc(x, z)      |x| <= kMulcMax, |z| <= q/2 centred        -> |result| <= 1665
//   mulc_wide(x, z) |x| <= kMulcWideMax                        -> |result| <= 1665
//   reduce_f(x)     |x| <= kReduceMax                          -> |result| <= 1665
//   compress_f<D>   |c| <= kCompressMax(D)                     -> the integer compress_d of (c mod q)
//   canon_f(x)      |x| <= kCanonMax                           -> x mod q in [0, q), as an integer
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PQC_F32_HD __host__ __device__
#else
#define PQC_F32_HD
#endif

namespace pqc {
namespace kyber {
namespace f32 {

constexpr int kQ = 3329;
constexpr float kQf = 3329.0f;
constexpr float kQinv = 1.0f / 3329.0f;
constexpr float kMagic = 12582912.0f;  // 1.5 * 2^23
constexpr int kMulcMax = 10000;
constexpr int kMulcWideMax = 20000;
constexpr int kReduceMax = 1 << 24;
constexpr int kOut = 1665;  // bound of every reduced result
constexpr int kCompressMax(int d) { return d >= 11 ? 4000 : 8000; }

PQC_F32_HD inline uint32_t bits_of(float x) { return __builtin_bit_cast(uint32_t, x); }

// round(x * zq): the integer nearest to x * zq, as a float.
PQC_F32_HD inline float round_mul(float x, float zq) { return __builtin_fmaf(x, zq, kMagic) - kMagic; }

// x * z mod q for a centred twiddle z (|z| <= 1664) and zq = fl(z / q).
// k = round(x z / q) is within 5039 of 0, so k * -q is exact, and the fma
// forms x z - k q exactly: |x z / q - k| <= 1/2 + |x| 2^-26.
PQC_F32_HD inline float mulc(float x, float z, float zq) {
  const float k = round_mul(x, zq);
  return __builtin_fmaf(x, z, k * -kQf);
}

// The same for larger |x|: k * 3328 = 13 k 2^8 stays exact for any k here, and
// x z - 3328 k = (x z - q k) + k is small enough to be exact.
PQC_F32_HD inline float mulc_wide(float x, float z, float zq) {
  const float k = round_mul(x, zq);
  return __builtin_fmaf(x, z, k * -3328.0f) - k;
}

// Centred representative of x mod q.
PQC_F32_HD inline float reduce_f(float x) {
  const float k = round_mul(x, kQinv);
  return __builtin_fmaf(k, -kQf, x);
}

// compress_d (round(x 2^D / q) mod 2^D, poly.cu:34 / polyvec.cu:33) from any
// representative c of x.  c 2^D / q + k 2^D for integer k rounds to the same
// integer mod 2^D, and no representative lies on a tie (q is odd), so the
// rounding of c * fl(2^D / q) by the magic constant gives the value once the
// error |c| 2^D / q 2^-24 stays below the 1/(2q) distance to the nearest tie.
// The rounded integer n sits in the low significand bits of n + M.
template <int D>
PQC_F32_HD inline uint32_t compress_f(float c) {
  constexpr float kd = (float)(1 << D) / kQf;
  return bits_of(__builtin_fmaf(c, kd, kMagic)) & ((1u << D) - 1);
}

// Canonical residue of x in [0, q), as an integer, for |x| <= kCanonMax (the
// values that are serialised: poly_tobytes packs csubq'd residues).  k =
// floor(x / q) is the rounding of x / q - 1/2 + 1/(2q): x / q is at least 1/q
// away from the next integer above unless it is one, so the shifted value never
// sits on a tie, and the float error (below 2^-20) stays far inside that 1/(2q)
// margin.  x + M - k q = M + r is exact and r sits in the low significand bits.
constexpr int kCanonMax = 1 << 16;
PQC_F32_HD inline uint32_t canon_f(float x) {
  constexpr float kBias = -0.5f + 0.5f / kQf;
  const float k = (__builtin_fmaf(x, kQinv, kBias) + kMagic) - kMagic;
  return bits_of(__builtin_fmaf(k, -kQf, x + kMagic)) & 0xFFFu;
}

// ---- schedules: which layers need the wide product or a reduction ------------
//
// Forward (Cooley-Tukey) layer li multiplies values bounded by b0 + li * 1665,
// where b0 bounds the input: each layer adds one reduced product.
constexpr unsigned fwd_wide_mask(int b0) {
  unsigned m = 0;
  for (int li = 0; li < 7; ++li)
    if (b0 + li * kOut > kMulcMax) m |= 1u << li;
  return m;
}
constexpr bool fwd_schedule_ok(int b0, unsigned wide) {
  for (int li = 0; li < 7; ++li) {
    const int b = b0 + li * kOut;
    if (b > (((wide >> li) & 1) ? kMulcWideMax : kMulcMax)) return false;
  }
  return true;
}
constexpr int fwd_out_bound(int b0) { return b0 + 7 * kOut; }

// Inverse (Gentleman-Sande) layer li: the difference x - y is multiplied and
// the sum x + y is kept, reduced on the layers in kInvReduce.  Starting from
// values bounded by 1665, sums reach 3330 on odd layers and 6660 before each
// reduction, and differences stay below kMulcMax.
constexpr unsigned kInvReduce = (1u << 1) | (1u << 3) | (1u << 5);
constexpr bool inv_schedule_ok(int b0, unsigned red) {
  int b = b0;
  for (int li = 0; li < 7; ++li) {
    if (2 * b > kMulcMax) return false;
    b = ((red >> li) & 1) ? kOut : (2 * b > kOut ? 2 * b : kOut);
  }
  return b <= kMulcMax;  // the final scaling multiplies values bounded by b
}
static_assert(inv_schedule_ok(kOut, kInvReduce), "inverse schedule");

// Base multiplication with accumulation: with every operand reduced (|a|, |b|,
// |b z| <= 1665) one pair product a0 b0 + a1 b1z is below 2 * 1665^2, and up to
// kBasemulTerms of them accumulate exactly before the sum must be reduced.
constexpr int kBasemulTerms = 3;
static_assert(kBasemulTerms * 2 * kOut * kOut < (1 << 24), "basemul accumulation exact");

}  // namespace f32
}  // namespace kyber
}  // namespace pqc
