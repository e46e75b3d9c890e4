// Short-message sponge helpers for one state per lane.
#pragma once
#include "keccak.cuh"

namespace pqc {

// Absorb W full little-endian words plus one extra word `tail` (the message
// tail bytes with the domain byte already placed), pad the last rate byte, and
// permute once.  Requires W + 1 <= RATE.
template <int RATE, int W>
__device__ __forceinline__ void absorb_fixed(uint64_t (&A)[25], const uint64_t (&m)[W], uint64_t tail) {
  static_assert(W + 1 <= RATE, "message must fit one block");
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = (i < W) ? m[i] : 0;
  A[W] ^= tail;
  A[RATE - 1] ^= 0x8000000000000000ull;
  keccakf1600(A);
}

__device__ __forceinline__ uint64_t load_u64(const uint8_t* p) {
  return *reinterpret_cast<const uint64_t*>(p);
}

// The same load from global memory, as a global (not flat) load: it counts in
// vmcnt only, so waits for LDS results do not also wait for it.
__device__ __forceinline__ uint64_t load_u64_global(const uint8_t* p) {
  return *(const __attribute__((address_space(1))) uint64_t*)p;
}

__device__ __forceinline__ void store_u64(uint8_t* p, uint64_t v) {
  *reinterpret_cast<uint64_t*>(p) = v;
}

}  // namespace pqc
