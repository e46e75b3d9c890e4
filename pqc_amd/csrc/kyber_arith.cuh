// Kyber (round 2) scalar arithmetic and parameter sets for the HIP kernels.
//
// Integer semantics follow Kyber_GPU_Batched/reduce.cu:21-66 and
// ntt.cu:81-84.  Values are carried in 32-bit registers; every intermediate of
// the keypair/encaps/decaps data flow stays inside the int16 range the
// reference stores it in (forward NTT outputs < 8q, basemul accumulations
// < 8q, inverse-NTT sums < 2q), so the 16-bit truncations of the reference
// never change a value and the results are identical.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pqc {
namespace kyber {

constexpr int N = 256;
constexpr int Q = 3329;
constexpr int QINV = 62209;  // q^-1 mod 2^16 (reduce.h:11)
constexpr int MONT = 2285;   // 2^16 mod q (reduce.h:11)
constexpr int SYM = 32;

// Round-2 parameter sets (params.h:20-45 for the layout; the widths du/dv
// for K = 2, 3 follow the round-2 specification, the reference fixes K = 4).
template <int K>
struct Params {
  static constexpr int k = K;
  static constexpr int du = (K == 4) ? 11 : 10;
  static constexpr int dv = (K == 2) ? 3 : (K == 3) ? 4 : 5;
  static constexpr int polyvec_bytes = K * 384;
  static constexpr int polyvec_cbytes = K * 32 * du;
  static constexpr int poly_cbytes = 32 * dv;
  static constexpr int ct_bytes = polyvec_cbytes + poly_cbytes;
  static constexpr int pk_bytes = polyvec_bytes + SYM;
  static constexpr int sk_cpa_bytes = polyvec_bytes;
  static constexpr int sk_bytes = sk_cpa_bytes + pk_bytes + 2 * SYM;
  static constexpr int hpk_off = sk_bytes - 2 * SYM;  // H(pk) inside sk
  static constexpr int z_off = sk_bytes - SYM;        // implicit-rejection z
};

// montgomery_reduce (reduce.cu:21-31): a * 2^-16 mod q, |result| < q for
// |a| < q 2^15.  The low-16-bit product is taken signed as int16 u.
__device__ __forceinline__ int32_t mont_reduce(int32_t a) {
  const int32_t u = (int32_t)(int16_t)(a * QINV);
  return (a - u * Q) >> 16;
}

__device__ __forceinline__ int32_t fqmul(int32_t a, int32_t b) { return mont_reduce(a * b); }

// barrett_reduce (reduce.cu:43-51): representative in {0..q} for int16 input.
__device__ __forceinline__ int32_t barrett(int32_t a) {
  constexpr int32_t v = (1 << 26) / Q + 1;
  const int32_t t = (v * a) >> 26;
  return a - t * Q;
}

// csubq (reduce.cu:62-66) for inputs in [0, 2q).
__device__ __forceinline__ int32_t csubq(int32_t a) {
  a -= Q;
  return a + ((a >> 31) & Q);
}

// floor(n / q) for n <= kDivQMax as (n * m) >> 33, m = ceil(2^33 / q), from
// the 24-bit high product: two instructions where the compiler's division by
// a constant takes five.  Exact: with e = m q - 2^33, n m / 2^33 = n / q +
// n e / (q 2^33), and n e < 2^33 keeps the floor.  Written as the instruction
// itself because the compiler does not form v_mul_hi_u32_u24 from C.
constexpr uint32_t kDivQM = (uint32_t)(((1ull << 33) + Q - 1) / Q);
constexpr uint32_t kDivQMax = (uint32_t)((Q - 1) << 11) + Q / 2;  // compress_d's largest numerator (d <= 11)
static_assert(kDivQM < (1u << 24) && kDivQMax < (1u << 24), "24-bit operands");
static_assert((uint64_t)((uint64_t)kDivQM * Q - (1ull << 33)) * kDivQMax < (1ull << 33), "exact division");
static_assert(kDivQM == 0x275F6Fu, "literal below");

__device__ __forceinline__ uint32_t div_q(uint32_t n) {
  uint32_t hi;
  asm("v_mul_hi_u32_u24 %0, 0x275f6f, %1" : "=v"(hi) : "v"(n));
  return hi >> 1;
}

// round(x * 2^d / q) mod 2^d for x in [0, q) (poly.cu:34, polyvec.cu:33).
template <int D>
__device__ __forceinline__ uint32_t compress_d(int32_t x) {
  static_assert(D >= 1 && D <= 11, "numerator bound");
  return div_q(((uint32_t)x << D) + Q / 2) & ((1u << D) - 1);
}

// (t q + 2^(d-1)) >> d (poly.cu:59-66, polyvec.cu:71-78).
template <int D>
__device__ __forceinline__ int32_t decompress_d(uint32_t t) {
  return (int32_t)((t * (uint32_t)Q + (1u << (D - 1))) >> D);
}

// ---- twiddles: the generator of ntt.cu:11-47, evaluated at compile time ----
constexpr unsigned brv7(unsigned i) {
  unsigned r = 0;
  for (int b = 0; b < 7; b++) r |= ((i >> b) & 1u) << (6 - b);
  return r;
}

struct Zetas {
  int16_t fwd[128];
  int16_t inv[128];
};

constexpr Zetas make_zetas() {
  Zetas z{};
  int32_t tmp[128] = {};
  tmp[0] = MONT;
  for (int i = 1; i < 128; i++) tmp[i] = (int32_t)((int64_t)tmp[i - 1] * 17 % Q);
  for (int i = 0; i < 128; i++) z.fwd[i] = (int16_t)tmp[brv7(i)];
  int k = 0;
  for (int i = 64; i >= 1; i >>= 1)
    for (int j = i; j < 2 * i; j++) z.inv[k++] = (int16_t)((Q - tmp[128 - brv7(j)]) % Q);
  z.inv[127] = (int16_t)((int64_t)MONT * ((int64_t)MONT * (Q - 1) % Q * ((Q - 1) / 128) % Q) % Q);
  return z;
}

static constexpr Zetas kZetas = make_zetas();

// ---- wave64 polynomial layout --------------------------------------------
// One polynomial per wave, 4 coefficients per lane.  A coefficient index has
// bits b7..b0.  Register slot bit s0 always holds b0 (so basemul pairs are
// lane-local), slot bit s1 and the six lane bits hold the rest:
//   NAT (serialisation order): s1 = b1, lane bit i = b(i+2); lane l holds
//       coefficients 4l..4l+3.
//   NTT (after the forward transform): s1 = b1, lane bits 0..4 = b3..b7,
//       lane bit 5 = b2; lane l holds the 4-coefficient group
//       ((l & 31) << 1) | (l >> 5).
// A butterfly layer on bit b runs locally once b sits in slot s1; reaching it
// swaps s1 with the lane bit that holds b (two cross-lane moves per lane).
// Forward (CT, len 128 -> 2) swap masks: 32 16 8 4 2 1 32.
// Inverse (GS, len 2 -> 128) swap masks: - 32 1 2 4 8 16, then 32 back to NAT.
constexpr int kFwdSwapMask[7] = {32, 16, 8, 4, 2, 1, 32};
constexpr int kInvSwapMask[8] = {0, 32, 1, 2, 4, 8, 16, 32};

struct LaneTwiddles {
  int16_t fwd[7][64];  // zeta of forward layer b = 7 - i for each lane
  int16_t inv[7][64];  // zeta of inverse layer b = i + 1 for each lane
  int16_t bm[64];      // basemul zeta of the lane's group (sign per slot s1)
};

// Tracks which index bit each lane bit holds through the swap sequence and
// evaluates the reference's twiddle index for the lower butterfly input j:
// forward zetas[2^(7-b) + (j >> (b+1))], inverse zetas_inv[128 - 2^(8-b) + (j >> (b+1))].
constexpr LaneTwiddles make_lane_twiddles() {
  LaneTwiddles t{};
  for (int lane = 0; lane < 64; lane++) {
    int holds[6] = {2, 3, 4, 5, 6, 7};  // NAT
    int slot1 = 1;
    for (int li = 0; li < 7; li++) {
      const int b = 7 - li;
      int L = 0;
      for (int q = 0; q < 6; q++)
        if (holds[q] == b) L = q;
      holds[L] = slot1;
      slot1 = b;
      unsigned j = 0;
      for (int q = 0; q < 6; q++) j |= (unsigned)((lane >> q) & 1) << holds[q];
      t.fwd[li][lane] = kZetas.fwd[(1u << (7 - b)) + (j >> (b + 1))];
    }
    // now NTT layout: holds = {3,4,5,6,7,2}, slot1 = 1
    for (int li = 0; li < 7; li++) {
      const int b = li + 1;
      if (slot1 != b) {
        int L = 0;
        for (int q = 0; q < 6; q++)
          if (holds[q] == b) L = q;
        holds[L] = slot1;
        slot1 = b;
      }
      unsigned j = 0;
      for (int q = 0; q < 6; q++) j |= (unsigned)((lane >> q) & 1) << holds[q];
      t.inv[li][lane] = kZetas.inv[(128u - (1u << (8 - b))) + (j >> (b + 1))];
    }
    const int grp = ((lane & 31) << 1) | (lane >> 5);
    t.bm[lane] = kZetas.fwd[64 + grp];
  }
  return t;
}

static constexpr LaneTwiddles kLaneTw = make_lane_twiddles();

__device__ __forceinline__ int nat_group(int lane) { return lane; }
__device__ __forceinline__ int ntt_group(int lane) { return ((lane & 31) << 1) | (lane >> 5); }

}  // namespace kyber
}  // namespace pqc
