// Keccak-f[1600] and the sponge modes used by the lattice KEMs, one sponge
// state per lane.  The 25 64-bit lanes live in VGPRs (the compiler splits each
// into a 32-bit pair; 64-bit rotates become two v_alignbit_b32).
//
// Follows the permutation and padding of Kyber_GPU_Batched/fips202.cu:393-530
// (the table-driven rho/pi loop there walks the same constants).  Round
// constants, rho offsets and the pi lane map are derived at compile time from
// their FIPS 202 definitions (Algorithms 2, 3 and 5), not tabulated.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pqc {

// FIPS 202 Algorithm 5: rc(t) from the LFSR x^8 + x^6 + x^5 + x^4 + 1.
constexpr int keccak_rc_bit(unsigned t) {
  unsigned r = 1;
  t %= 255;
  for (unsigned i = 0; i < t; i++) {
    r <<= 1;
    if (r & 0x100) r ^= 0x171;
  }
  return (int)(r & 1);
}

constexpr uint64_t keccak_rc(unsigned ir) {
  uint64_t rc = 0;
  for (unsigned j = 0; j <= 6; j++)
    if (keccak_rc_bit(j + 7 * ir)) rc |= (uint64_t)1 << ((1u << j) - 1);
  return rc;
}

struct KeccakLaneMaps {
  int rho[25];     // rotation of lane x + 5y (Algorithm 2)
  int pi_dst[25];  // pi moves lane (x, y) to (y, 2x + 3y) (Algorithm 3)
};

constexpr KeccakLaneMaps keccak_lane_maps() {
  KeccakLaneMaps m{};
  for (int i = 0; i < 25; i++) m.rho[i] = 0;
  unsigned x = 1, y = 0;
  for (unsigned t = 0; t < 24; t++) {
    m.rho[x + 5 * y] = (int)(((t + 1) * (t + 2) / 2) % 64);
    unsigned nx = y, ny = (2 * x + 3 * y) % 5;
    x = nx;
    y = ny;
  }
  for (int xx = 0; xx < 5; xx++)
    for (int yy = 0; yy < 5; yy++) m.pi_dst[xx + 5 * yy] = yy + 5 * ((2 * xx + 3 * yy) % 5);
  return m;
}

static constexpr KeccakLaneMaps kKeccakMaps = keccak_lane_maps();

__constant__ const uint64_t c_keccak_rc[24] = {
    keccak_rc(0),  keccak_rc(1),  keccak_rc(2),  keccak_rc(3),  keccak_rc(4),  keccak_rc(5),
    keccak_rc(6),  keccak_rc(7),  keccak_rc(8),  keccak_rc(9),  keccak_rc(10), keccak_rc(11),
    keccak_rc(12), keccak_rc(13), keccak_rc(14), keccak_rc(15), keccak_rc(16), keccak_rc(17),
    keccak_rc(18), keccak_rc(19), keccak_rc(20), keccak_rc(21), keccak_rc(22), keccak_rc(23)};

// The permutation works on explicit 32-bit halves so that every step maps to
// one gfx950 instruction per half:
//   three-input XOR and chi's a ^ (~b & c) ... v_bitop3_b32 (truth tables 0x96
//     and 0xD2 over operand patterns S0 = 0xF0, S1 = 0xCC, S2 = 0xAA)
//   64-bit rotation .......................... two v_alignbit_b32
// giving 180 VALU instructions per round (theta 80, rho 48, chi 50, iota 2).
struct Lane2 {
  uint32_t lo, hi;
};

__device__ __forceinline__ uint32_t xor3_32(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ uint32_t chi_32(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0xD2);  // a ^ (~b & c)
}

template <int N>
__device__ __forceinline__ Lane2 rotl_2(Lane2 x) {
  if constexpr (N == 0) {
    return x;
  } else if constexpr (N < 32) {
    return Lane2{__builtin_amdgcn_alignbit(x.lo, x.hi, 32 - N), __builtin_amdgcn_alignbit(x.hi, x.lo, 32 - N)};
  } else if constexpr (N == 32) {
    return Lane2{x.hi, x.lo};
  } else {
    return Lane2{__builtin_amdgcn_alignbit(x.hi, x.lo, 64 - N), __builtin_amdgcn_alignbit(x.lo, x.hi, 64 - N)};
  }
}

template <int I>
__device__ __forceinline__ void keccak_theta_rho_pi(Lane2 (&B)[25], const Lane2 (&A)[25], const Lane2 (&Cm)[5],
                                                    const Lane2 (&Rp)[5]) {
  // theta folded into one XOR3: A ^ C[x-1] ^ rot(C[x+1], 1)
  constexpr int x = I % 5;
  const Lane2 t{xor3_32(A[I].lo, Cm[x].lo, Rp[x].lo), xor3_32(A[I].hi, Cm[x].hi, Rp[x].hi)};
  B[kKeccakMaps.pi_dst[I]] = rotl_2<kKeccakMaps.rho[I]>(t);
  if constexpr (I + 1 < 25) keccak_theta_rho_pi<I + 1>(B, A, Cm, Rp);
}

// One Keccak-f[1600] permutation of a register-resident state.  The round loop
// stays rolled (the round constant is a uniform scalar load) so that every
// call site costs one copy of the round body.
__device__ __forceinline__ void keccakf1600(uint64_t (&A64)[25]) {
  Lane2 A[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = Lane2{(uint32_t)A64[i], (uint32_t)(A64[i] >> 32)};
#pragma unroll 1
  for (int round = 0; round < 24; ++round) {
    Lane2 C[5], Cm[5], Rp[5], B[25];
#pragma unroll
    for (int x = 0; x < 5; ++x) {
      C[x].lo = xor3_32(xor3_32(A[x].lo, A[x + 5].lo, A[x + 10].lo), A[x + 15].lo, A[x + 20].lo);
      C[x].hi = xor3_32(xor3_32(A[x].hi, A[x + 5].hi, A[x + 10].hi), A[x + 15].hi, A[x + 20].hi);
    }
#pragma unroll
    for (int x = 0; x < 5; ++x) {
      Cm[x] = C[(x + 4) % 5];
      Rp[x] = rotl_2<1>(C[(x + 1) % 5]);
    }
    keccak_theta_rho_pi<0>(B, A, Cm, Rp);
#pragma unroll
    for (int y = 0; y < 5; ++y) {
#pragma unroll
      for (int x = 0; x < 5; ++x) {
        const Lane2 a = B[x + 5 * y], b = B[(x + 1) % 5 + 5 * y], c = B[(x + 2) % 5 + 5 * y];
        A[x + 5 * y] = Lane2{chi_32(a.lo, b.lo, c.lo), chi_32(a.hi, b.hi, c.hi)};
      }
    }
    const uint64_t rc = c_keccak_rc[round];
    A[0].lo ^= (uint32_t)rc;
    A[0].hi ^= (uint32_t)(rc >> 32);
  }
#pragma unroll
  for (int i = 0; i < 25; ++i) A64[i] = (uint64_t)A[i].lo | ((uint64_t)A[i].hi << 32);
}

// ---- split state: one sponge per pair of adjacent lanes ------------------------
// The even lane holds the low 32-bit halves of the 25 state words, the odd lane
// the high halves.  A 64-bit rotation needs the partner's half: one DPP
// quad_perm(1,0,3,2) move, then one v_alignbit_b32 whose operand order is the
// same in both lanes.  Theta's parity and chi are lane-local.  120 VALU
// instructions per round per lane (vs 180 for a whole state in one lane), so a
// long sequential sponge (Frodo's 231-block noise stream) finishes ~1.5x sooner.
__device__ __forceinline__ uint32_t pair_swap(uint32_t v) {
  // quad_perm(1,0,3,2); every source lane is valid, so no "old" value is needed
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
}

template <int N>
__device__ __forceinline__ uint32_t rotl_pair(uint32_t own) {
  if constexpr (N == 0) {
    return own;
  } else {
    const uint32_t partner = pair_swap(own);
    if constexpr (N < 32)
      return __builtin_amdgcn_alignbit(own, partner, 32 - N);
    else if constexpr (N == 32)
      return partner;
    else
      return __builtin_amdgcn_alignbit(partner, own, 64 - N);
  }
}

template <int I>
__device__ __forceinline__ void keccak_theta_rho_pi_pair(uint32_t (&B)[25], const uint32_t (&A)[25],
                                                         const uint32_t (&Cm)[5], const uint32_t (&Rp)[5]) {
  constexpr int x = I % 5;
  B[kKeccakMaps.pi_dst[I]] = rotl_pair<kKeccakMaps.rho[I]>(xor3_32(A[I], Cm[x], Rp[x]));
  if constexpr (I + 1 < 25) keccak_theta_rho_pi_pair<I + 1>(B, A, Cm, Rp);
}

// Both lanes of every pair must be active.  `odd` is lane & 1.  UNROLL rounds
// per loop trip: 1 keeps the code small; 24 drops the loop's round-constant load
// and branch (for a call site that carries a long serial chain).
template <int UNROLL = 1>
__device__ __forceinline__ void keccakf1600_pair(uint32_t (&A)[25], uint32_t odd) {
  static_assert(24 % UNROLL == 0, "rounds per trip");
#pragma unroll UNROLL
  for (int round = 0; round < 24; ++round) {
    uint32_t C[5], Cm[5], Rp[5], B[25];
#pragma unroll
    for (int x = 0; x < 5; ++x) C[x] = xor3_32(xor3_32(A[x], A[x + 5], A[x + 10]), A[x + 15], A[x + 20]);
#pragma unroll
    for (int x = 0; x < 5; ++x) {
      Cm[x] = C[(x + 4) % 5];
      Rp[x] = rotl_pair<1>(C[(x + 1) % 5]);
    }
    keccak_theta_rho_pi_pair<0>(B, A, Cm, Rp);
#pragma unroll
    for (int y = 0; y < 5; ++y)
#pragma unroll
      for (int x = 0; x < 5; ++x)
        A[x + 5 * y] = chi_32(B[x + 5 * y], B[(x + 1) % 5 + 5 * y], B[(x + 2) % 5 + 5 * y]);
    const uint64_t rc = c_keccak_rc[round];
    A[0] ^= odd ? (uint32_t)(rc >> 32) : (uint32_t)rc;
  }
}

// ---- bit-interleaved lane pairs ------------------------------------------------
// A 64-bit lane x as two 32-bit words, E (even lane of the pair) = bits 0, 2, ..,
// 62 of x and O (odd lane) = bits 1, 3, .., 63.  rot64(x, 2k) is then rot32 of
// each word by k (lane-local: no exchange), and rot64(x, 2k + 1) is
// E' = rot32(O, k + 1), O' = rot32(E, k): the partner's word rotated by
// k + 1 - half.  Of rho's 24 rotations 12 are even, so a round needs 12 fewer
// cross-lane moves than the lo/hi split (keccakf1600_pair).
constexpr uint32_t il_bits(uint64_t x) {  // bits 0, 2, .., 62 of x
  uint32_t r = 0;
  for (int i = 0; i < 32; ++i) r |= (uint32_t)((x >> (2 * i)) & 1u) << i;
  return r;
}

struct IlRc {
  uint32_t e[24], o[24];
};
constexpr IlRc il_round_constants() {
  IlRc t{};
  for (unsigned r = 0; r < 24; ++r) {
    t.e[r] = il_bits(keccak_rc(r));
    t.o[r] = il_bits(keccak_rc(r) >> 1);
  }
  return t;
}
static constexpr IlRc kIlRc = il_round_constants();

// this lane's interleaved word of the 64-bit value x (half = lane & 1)
__device__ __forceinline__ uint32_t il_word(uint64_t x, uint32_t half) {
  x = (x >> half) & 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  return (uint32_t)(x | (x >> 16));
}

template <int N>
__device__ __forceinline__ uint32_t rotl_il(uint32_t own, uint32_t half) {
  if constexpr (N == 0) {
    return own;
  } else if constexpr (N % 2 == 0) {
    return __builtin_amdgcn_alignbit(own, own, 32 - N / 2);
  } else {
    const uint32_t partner = pair_swap(own);
    // rotate left by N / 2 + 1 - half = right by 31 - N / 2 + half (mod 32)
    return __builtin_amdgcn_alignbit(partner, partner, 31u - (uint32_t)(N / 2) + half);
  }
}

template <int I>
__device__ __forceinline__ void keccak_theta_rho_pi_il(uint32_t (&B)[25], const uint32_t (&A)[25],
                                                       const uint32_t (&Cm)[5], const uint32_t (&Rp)[5],
                                                       uint32_t half) {
  constexpr int x = I % 5;
  B[kKeccakMaps.pi_dst[I]] = rotl_il<kKeccakMaps.rho[I]>(xor3_32(A[I], Cm[x], Rp[x]), half);
  if constexpr (I + 1 < 25) keccak_theta_rho_pi_il<I + 1>(B, A, Cm, Rp, half);
}

// Keccak-f[1600] on a bit-interleaved lane pair; half = lane & 1, both lanes active.
template <int UNROLL = 1>
__device__ __forceinline__ void keccakf1600_pair_il(uint32_t (&A)[25], uint32_t half) {
  static_assert(24 % UNROLL == 0, "rounds per trip");
#pragma unroll UNROLL
  for (int round = 0; round < 24; ++round) {
    uint32_t C[5], Cm[5], Rp[5], B[25];
#pragma unroll
    for (int x = 0; x < 5; ++x) C[x] = xor3_32(xor3_32(A[x], A[x + 5], A[x + 10]), A[x + 15], A[x + 20]);
#pragma unroll
    for (int x = 0; x < 5; ++x) {
      Cm[x] = C[(x + 4) % 5];
      Rp[x] = rotl_il<1>(C[(x + 1) % 5], half);
    }
    keccak_theta_rho_pi_il<0>(B, A, Cm, Rp, half);
#pragma unroll
    for (int y = 0; y < 5; ++y)
#pragma unroll
      for (int x = 0; x < 5; ++x)
        A[x + 5 * y] = chi_32(B[x + 5 * y], B[(x + 1) % 5 + 5 * y], B[(x + 2) % 5 + 5 * y]);
    A[0] ^= half ? kIlRc.o[round] : kIlRc.e[round];
  }
}

__device__ __forceinline__ void keccak_zero(uint64_t (&A)[25]) {
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = 0;
}

// Rates in 64-bit words and the domain bytes of the sponge modes
// (fips202.cu:543-689): SHAKE128 168 B / 0x1F, SHAKE256 136 B / 0x1F,
// SHA3-256 136 B / 0x06, SHA3-512 72 B / 0x06.
enum : int { kRateShake128 = 21, kRateShake256 = 17, kRateSha3_256 = 17, kRateSha3_512 = 9 };

// Hash a message of nwords 64-bit little-endian words read from global memory
// (8-byte aligned).  Every block, including the padded final one, goes through
// one keccakf1600 call site.  Returns the state after the final permutation.
template <int RATE>
__device__ __forceinline__ void sponge_absorb_words(uint64_t (&A)[25], const uint64_t* __restrict__ p,
                                                    int nwords, uint8_t domain) {
  keccak_zero(A);
  const int nblocks = nwords / RATE + 1;
  for (int b = 0; b < nblocks; ++b) {
    const int rem = nwords - b * RATE;  // words left, >= 0
#pragma unroll
    for (int w = 0; w < RATE; ++w) {
      if (w < rem) A[w] ^= p[w];
      if (w == rem) A[w] ^= (uint64_t)domain;
    }
    if (b == nblocks - 1) A[RATE - 1] ^= 0x8000000000000000ull;
    keccakf1600(A);
    p += RATE;
  }
}

// The same, with block b + 1's words loaded while block b is permuted: for a
// kernel that runs one wave per SIMD, where nothing else hides the load latency
// of a long absorb (2 RATE extra registers).
template <int RATE>
__device__ __forceinline__ void sponge_absorb_words_prefetch(uint64_t (&A)[25], const uint64_t* __restrict__ p,
                                                             int nwords, uint8_t domain) {
  keccak_zero(A);
  const int nblocks = nwords / RATE + 1;
  uint64_t nxt[RATE];
#pragma unroll
  for (int w = 0; w < RATE; ++w) nxt[w] = w < nwords ? p[w] : 0;
  for (int b = 0; b < nblocks; ++b) {
    const int rem = nwords - b * RATE;  // words left, >= 0
    uint64_t cur[RATE];
#pragma unroll
    for (int w = 0; w < RATE; ++w) cur[w] = nxt[w];
#pragma unroll
    for (int w = 0; w < RATE; ++w) nxt[w] = w < rem - RATE ? p[(b + 1) * RATE + w] : 0;
#pragma unroll
    for (int w = 0; w < RATE; ++w) {
      if (w < rem) A[w] ^= cur[w];
      if (w == rem) A[w] ^= (uint64_t)domain;
    }
    if (b == nblocks - 1) A[RATE - 1] ^= 0x8000000000000000ull;
    keccakf1600(A);
  }
}

// Blocks [b0, b1) of a rate-17-word absorb (SHA3-256 / SHAKE256) of nwords
// 64-bit words on a lane pair: the even lane holds the low 32-bit halves of
// the state words, the odd lane the high halves.  b0 = 0 starts from zero.
__device__ __forceinline__ void absorb_range_pair(uint32_t (&S)[25], const uint8_t* __restrict__ msg, int nwords,
                                                  int b0, int b1, uint32_t domain, uint32_t half) {
  constexpr int R = kRateSha3_256;
  if (b0 == 0) {
#pragma unroll
    for (int w = 0; w < 25; ++w) S[w] = 0;
  }
  const int nblocks = nwords / R + 1;
  const uint32_t* p = reinterpret_cast<const uint32_t*>(msg) + half;
  for (int b = b0; b < b1 && b < nblocks; ++b) {
    const int rem = nwords - b * R;  // words left, >= 0
#pragma unroll
    for (int w = 0; w < R; ++w) {
      if (w < rem) S[w] ^= p[2 * (b * R + w)];
      if (w == rem && !half) S[w] ^= domain;
    }
    if (b == nblocks - 1 && half) S[R - 1] ^= 0x80000000u;
    keccakf1600_pair(S, half);
  }
}

}  // namespace pqc
