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

template <int N>
__device__ __forceinline__ uint64_t rotl64(uint64_t x) {
  if constexpr (N == 0) {
    return x;
  } else {
    return (x << N) | (x >> (64 - N));
  }
}

template <int I>
__device__ __forceinline__ void keccak_rho_pi(uint64_t (&B)[25], const uint64_t (&A)[25],
                                              const uint64_t (&D)[5]) {
  B[kKeccakMaps.pi_dst[I]] = rotl64<kKeccakMaps.rho[I]>(A[I] ^ D[I % 5]);
  if constexpr (I + 1 < 25) keccak_rho_pi<I + 1>(B, A, D);
}

// One Keccak-f[1600] permutation of a register-resident state.  The round loop
// stays rolled (the round constant is a uniform scalar load) so that every
// call site costs one copy of the round body.
__device__ __forceinline__ void keccakf1600(uint64_t (&A)[25]) {
#pragma unroll 1
  for (int round = 0; round < 24; ++round) {
    uint64_t C[5], D[5], B[25];
#pragma unroll
    for (int x = 0; x < 5; ++x) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
#pragma unroll
    for (int x = 0; x < 5; ++x) D[x] = C[(x + 4) % 5] ^ rotl64<1>(C[(x + 1) % 5]);
    keccak_rho_pi<0>(B, A, D);
#pragma unroll
    for (int y = 0; y < 5; ++y) {
#pragma unroll
      for (int x = 0; x < 5; ++x) {
        const uint64_t a = B[x + 5 * y], b = B[(x + 1) % 5 + 5 * y], c = B[(x + 2) % 5 + 5 * y];
        // a ^ (~b & c) written as a bit-select: b ? a : a ^ c  (v_bfi_b32)
        A[x + 5 * y] = (b & a) | (~b & (a ^ c));
      }
    }
    A[0] ^= c_keccak_rc[round];
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

}  // namespace pqc
