// Wave64 polynomial primitives for Kyber: one polynomial per wave, four int16
// coefficients per lane (kept in 32-bit registers).  Layouts NAT and NTT and
// the swap schedule are described in kyber_arith.cuh.
//
// Butterflies follow ntt_p / invntt_p (Kyber_GPU_Batched/ntt.cu:152-224);
// basemul follows basemul2 / poly_basemul (ntt.cu:249-259, poly.cu:204-225).
#pragma once
#include "kyber_arith.cuh"

namespace pqc {
namespace kyber {

struct P4 {
  int32_t v[4];  // v[(s1 << 1) | s0]
};

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Swap register slot s1 with lane bit `MASK`.  After the swap, for lanes whose
// bit is 0: v[s0] keeps its value and v[2|s0] receives the partner's v[s0];
// for lanes whose bit is 1: v[2|s0] keeps its value and v[s0] receives the
// partner's v[2|s0].
template <int MASK>
__device__ __forceinline__ void swap_s1(P4& p) {
  if constexpr (MASK == 32) {
#pragma unroll
    for (int s0 = 0; s0 < 2; ++s0) {
      auto r = __builtin_amdgcn_permlane32_swap(p.v[s0], p.v[2 | s0], false, false);
      p.v[s0] = r[0];
      p.v[2 | s0] = r[1];
    }
  } else if constexpr (MASK == 16) {
#pragma unroll
    for (int s0 = 0; s0 < 2; ++s0) {
      auto r = __builtin_amdgcn_permlane16_swap(p.v[s0], p.v[2 | s0], false, false);
      p.v[s0] = r[0];
      p.v[2 | s0] = r[1];
    }
  } else {
    // DPP row shifts inside a 16-lane row: row_shl:n reads lane i+n, row_shr:n reads lane i-n.
    // A lane whose bit is 0 reads i+n (same row), one whose bit is 1 reads i-n;
    // the lanes whose source falls outside the row discard the value, so
    // bound_ctrl (read 0 there) spares initialising the destination.
    constexpr int SHL = 0x100 | MASK, SHR = 0x110 | MASK;
    const bool hi = (lane_id() & MASK) != 0;
#pragma unroll
    for (int s0 = 0; s0 < 2; ++s0) {
      const int32_t from_up = __builtin_amdgcn_mov_dpp(p.v[s0], SHL, 0xF, 0xF, true);
      const int32_t from_dn = __builtin_amdgcn_mov_dpp(p.v[2 | s0], SHR, 0xF, 0xF, true);
      const int32_t nlo = hi ? from_dn : p.v[s0];
      const int32_t nhi = hi ? p.v[2 | s0] : from_up;
      p.v[s0] = nlo;
      p.v[2 | s0] = nhi;
    }
  }
}

// Cooley-Tukey butterfly on the slot-s1 pairs (ntt.cu:180-182).
__device__ __forceinline__ void ct_layer(P4& p, int32_t zeta) {
#pragma unroll
  for (int s0 = 0; s0 < 2; ++s0) {
    const int32_t t = fqmul(zeta, p.v[2 | s0]);
    const int32_t x = p.v[s0];
    p.v[2 | s0] = x - t;
    p.v[s0] = x + t;
  }
}

// Gentleman-Sande butterfly (ntt.cu:212-215).
__device__ __forceinline__ void gs_layer(P4& p, int32_t zeta) {
#pragma unroll
  for (int s0 = 0; s0 < 2; ++s0) {
    const int32_t x = p.v[s0], y = p.v[2 | s0];
    p.v[s0] = barrett(x + y);
    p.v[2 | s0] = fqmul(zeta, x - y);
  }
}

// Per-lane twiddles, loaded once per kernel.
struct LaneZ {
  int32_t f[7], i[7], bm;
};

__device__ __forceinline__ void load_lane_z(LaneZ& z, const LaneTwiddles* tw) {
  const int l = lane_id();
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    z.f[k] = tw->fwd[k][l];
    z.i[k] = tw->inv[k][l];
  }
  z.bm = tw->bm[l];
}

// poly_ntt (poly.cu:153-157): NAT in, NTT layout out, Barrett-reduced.
__device__ __forceinline__ void ntt(P4& p, const LaneZ& z) {
  swap_s1<32>(p); ct_layer(p, z.f[0]);
  swap_s1<16>(p); ct_layer(p, z.f[1]);
  swap_s1<8>(p);  ct_layer(p, z.f[2]);
  swap_s1<4>(p);  ct_layer(p, z.f[3]);
  swap_s1<2>(p);  ct_layer(p, z.f[4]);
  swap_s1<1>(p);  ct_layer(p, z.f[5]);
  swap_s1<32>(p); ct_layer(p, z.f[6]);
#pragma unroll
  for (int r = 0; r < 4; ++r) p.v[r] = barrett(p.v[r]);
}

// poly_invntt (poly.cu:178-181): NTT layout in, NAT out (times 2^16 mod q).
__device__ __forceinline__ void invntt(P4& p, const LaneZ& z) {
  gs_layer(p, z.i[0]);
  swap_s1<32>(p); gs_layer(p, z.i[1]);
  swap_s1<1>(p);  gs_layer(p, z.i[2]);
  swap_s1<2>(p);  gs_layer(p, z.i[3]);
  swap_s1<4>(p);  gs_layer(p, z.i[4]);
  swap_s1<8>(p);  gs_layer(p, z.i[5]);
  swap_s1<16>(p); gs_layer(p, z.i[6]);
  swap_s1<32>(p);
  constexpr int32_t f = kZetas.inv[127];
#pragma unroll
  for (int r = 0; r < 4; ++r) p.v[r] = fqmul(p.v[r], f);
}

// r = a o b in the NTT domain (basemul2 over the lane's two pairs).
__device__ __forceinline__ void basemul(P4& r, const P4& a, const P4& b, int32_t zeta) {
#pragma unroll
  for (int s1 = 0; s1 < 2; ++s1) {
    const int32_t zz = s1 ? -zeta : zeta;
    const int32_t a0 = a.v[2 * s1], a1 = a.v[2 * s1 + 1];
    const int32_t b0 = b.v[2 * s1], b1 = b.v[2 * s1 + 1];
    r.v[2 * s1] = fqmul(fqmul(a1, b1), zz) + fqmul(a0, b0);
    r.v[2 * s1 + 1] = fqmul(a0, b1) + fqmul(a1, b0);
  }
}

__device__ __forceinline__ void basemul_acc(P4& r, const P4& a, const P4& b, int32_t zeta) {
  P4 t;
  basemul(t, a, b, zeta);
#pragma unroll
  for (int i = 0; i < 4; ++i) r.v[i] += t.v[i];
}

__device__ __forceinline__ void reduce(P4& p) {
#pragma unroll
  for (int r = 0; r < 4; ++r) p.v[r] = barrett(p.v[r]);
}

// ---- serialisation of one lane's 4-coefficient group ----------------------

// Read `nbits` (<= 48) bits at bit offset `bit` from a 4-byte-aligned buffer
// whose length is a multiple of 4 bytes (every read dword holds a wanted byte).
__device__ __forceinline__ uint64_t read_bits(const uint8_t* base, uint32_t bit, int nbits) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(base) + (bit >> 5);
  const int s = (int)(bit & 31);
  const uint64_t lo = (uint64_t)w[0] | ((uint64_t)w[1] << 32);  // needs s + nbits > 32: always, nbits >= 12 ... see callers
  uint64_t v = lo >> s;
  if (s + nbits > 64) v |= (uint64_t)w[2] << (64 - s);
  return v & ((1ull << nbits) - 1);
}

// Unpack 4 coefficients of D bits for group g.
template <int D>
__device__ __forceinline__ void unpack_group(uint32_t (&t)[4], const uint8_t* base, int g) {
  const uint32_t bit = (uint32_t)(4 * D) * (uint32_t)g;
  uint64_t v;
  if constexpr (4 * D <= 32) {
    // fits in two dwords at most; avoid reading a second dword that may lie
    // beyond the buffer end
    const uint32_t* w = reinterpret_cast<const uint32_t*>(base) + (bit >> 5);
    const int s = (int)(bit & 31);
    v = (uint64_t)w[0] >> s;
    if (s + 4 * D > 32) v |= (uint64_t)w[1] << (32 - s);
  } else {
    v = read_bits(base, bit, 4 * D);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) t[i] = (uint32_t)(v >> (D * i)) & ((1u << D) - 1);
}

// 12-bit unpack of a group (poly_frombytes, poly.cu:106-116): 6 bytes at 6g.
__device__ __forceinline__ void load_poly12(P4& p, const uint8_t* base, int g) {
  const uint16_t* h = reinterpret_cast<const uint16_t*>(base + 6 * g);
  const uint32_t w0 = h[0], w1 = h[1], w2 = h[2];
  p.v[0] = (int32_t)(w0 & 0xFFF);
  p.v[1] = (int32_t)((w0 >> 12) | ((w1 & 0xFF) << 4));
  p.v[2] = (int32_t)((w1 >> 8) | ((w2 & 0xF) << 8));
  p.v[3] = (int32_t)(w2 >> 4);
}

// The same unpackings split into a load (issued early, e.g. for the next
// request while the current one computes) and the arithmetic on the loaded words.
//
// The loads are written so that nothing but the load itself happens where they
// are issued: the words are kept as loaded and taken apart only where they are
// used.  (A load whose value is masked, selected or split at the load site makes
// the wave wait for it right there, which turns a prefetch into a stall.)  They
// go through global-address-space pointers: global loads count in vmcnt only,
// while flat loads also count in lgkmcnt, so every wait for an LDS result would
// also wait for the prefetches in flight.
template <class T>
using global_t = __attribute__((address_space(1))) T;

template <class T>
__device__ __forceinline__ const global_t<T>* gptr(const T* p) {
  return (const global_t<T>*)p;
}

// Words 0..2 of group g's bits.  Word 1 (2) is only needed when the group's bits
// reach it; otherwise the load reads word 0 again, so no dword outside the ones
// holding the group's bytes is ever read, and no data-dependent select is needed
// (unpack_loaded never takes a bit from a word the group does not reach).
template <int D>
__device__ __forceinline__ void load_group(uint32_t (&w)[3], const uint8_t* base, int g) {
  const uint32_t bit = (uint32_t)(4 * D) * (uint32_t)g;
  const global_t<uint32_t>* p = gptr(reinterpret_cast<const uint32_t*>(base)) + (bit >> 5);
  const int s = (int)(bit & 31);
  w[0] = p[0];
  w[1] = p[(s + 4 * D > 32) ? 1 : 0];
  w[2] = p[(s + 4 * D > 64) ? 2 : 0];
}

template <int D>
__device__ __forceinline__ void unpack_loaded(uint32_t (&t)[4], const uint32_t (&w)[3], int g) {
  const int s = (int)(((uint32_t)(4 * D) * (uint32_t)g) & 31);
  uint64_t v = ((uint64_t)w[0] | ((uint64_t)w[1] << 32)) >> s;
  if (s + 4 * D > 64) v |= (uint64_t)w[2] << (64 - s);
#pragma unroll
  for (int i = 0; i < 4; ++i) t[i] = (uint32_t)(v >> (D * i)) & ((1u << D) - 1);
}

// 12-bit group g (6 bytes at 6g, poly_frombytes) as the two aligned dwords that
// hold it: bytes 6g .. 6g + 5 start at byte 0 or 2 of the first (base is 4-byte
// aligned, and 6g + 5 < 384 keeps both dwords inside the polynomial).
__device__ __forceinline__ void load_h12(uint32_t (&h)[2], const uint8_t* base, int g) {
  const global_t<uint32_t>* p = gptr(reinterpret_cast<const uint32_t*>(base + ((6 * g) & ~3)));
  h[0] = p[0];
  h[1] = p[1];
}

__device__ __forceinline__ void poly12_loaded(P4& p, const uint32_t (&h)[2], int g) {
  const uint64_t v = (((uint64_t)h[1] << 32) | h[0]) >> (8 * ((6 * g) & 3));
#pragma unroll
  for (int r = 0; r < 4; ++r) p.v[r] = (int32_t)((uint32_t)(v >> (12 * r)) & 0xFFF);
}

// CBD (eta = 2) of the group's 16 noise bits (cbd.cu:42-63): coefficient c uses
// bits 4c..4c+3 of the PRF output; a = b0 + b1, b = b2 + b3.
__device__ __forceinline__ void cbd_group(P4& p, uint32_t bits16) {
  const uint32_t d = (bits16 & 0x5555u) + ((bits16 >> 1) & 0x5555u);
#pragma unroll
  for (int j = 0; j < 4; ++j)
    p.v[j] = (int32_t)((d >> (4 * j)) & 3) - (int32_t)((d >> (4 * j + 2)) & 3);
}

}  // namespace kyber
}  // namespace pqc
