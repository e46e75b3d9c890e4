// Wave64 polynomial arithmetic on exact FP32 values (kyber_f32.h): the same
// layouts, swap schedule and butterfly network as kyber_wave.cuh, with the
// Montgomery and Barrett reductions replaced by FP32 products reduced by the
// rounded quotient.  Results equal the integer forms' mod q; they leave through
// compress_f, whose value depends only on the residue, so the bytes produced are
// the same.
//
// Twiddles: the integer forms multiply by Montgomery-form zetas (fqmul(zeta R, x)
// = zeta x), so the FP32 forms take the plain zetas, centred.  The integer
// base multiplication carries a factor R^-1 (two fqmuls per product) that the
// inverse transform's final fqmul by f = kZetas.inv[127] (another R^-1) cancels
// together with R^2 inside f; here the product carries no factor, so the final
// scaling is f R^-2 mod q.
#pragma once
#include "kyber_f32.h"
#include "kyber_wave.cuh"

namespace pqc {
namespace kyber {

constexpr int kRinv = 169;  // 2^-16 mod q
static_assert((MONT * kRinv) % Q == 1, "R^-1");

constexpr int mod_q(long long a) { return (int)(((a % Q) + Q) % Q); }
constexpr int centred(long long a) {
  const int r = mod_q(a);
  return r > Q / 2 ? r - Q : r;
}
// plain centred value of a Montgomery-form twiddle m (m R^-1 mod q)
constexpr int plain_tw(int m) { return centred((long long)m * kRinv); }

struct LaneTwiddlesF {
  float fwd[7][64], fwdq[7][64];
  float inv[7][64], invq[7][64];
  float bm[64], bmq[64];
};

constexpr LaneTwiddlesF make_lane_twiddles_f() {
  LaneTwiddlesF t{};
  for (int li = 0; li < 7; ++li)
    for (int l = 0; l < 64; ++l) {
      const int zf = plain_tw(kLaneTw.fwd[li][l]), zi = plain_tw(kLaneTw.inv[li][l]);
      t.fwd[li][l] = (float)zf;
      t.fwdq[li][l] = (float)zf / f32::kQf;
      t.inv[li][l] = (float)zi;
      t.invq[li][l] = (float)zi / f32::kQf;
    }
  for (int l = 0; l < 64; ++l) {
    const int zb = plain_tw(kLaneTw.bm[l]);
    t.bm[l] = (float)zb;
    t.bmq[l] = (float)zb / f32::kQf;
  }
  return t;
}

static constexpr LaneTwiddlesF kLaneTwF = make_lane_twiddles_f();

// final scaling of the inverse transform: f R^-2 (f = kZetas.inv[127])
constexpr int kInvScale = centred((long long)kZetas.inv[127] * kRinv % Q * kRinv);
constexpr float kInvScaleF = (float)kInvScale;
constexpr float kInvScaleQ = (float)kInvScale / f32::kQf;

struct F4 {
  float v[4];  // v[(s1 << 1) | s0], as P4
};

struct LaneZF {
  float f[7], fq[7], i[7], iq[7], bm, bmq;
};

__device__ __forceinline__ void load_lane_zf(LaneZF& z, const LaneTwiddlesF* tw) {
  const int l = lane_id();
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    z.f[k] = tw->fwd[k][l];
    z.fq[k] = tw->fwdq[k][l];
    z.i[k] = tw->inv[k][l];
    z.iq[k] = tw->invq[k][l];
  }
  z.bm = tw->bm[l];
  z.bmq = tw->bmq[l];
}

template <int MASK>
__device__ __forceinline__ void swap_f(F4& p) {
  P4 q;
#pragma unroll
  for (int r = 0; r < 4; ++r) q.v[r] = __float_as_int(p.v[r]);
  swap_s1<MASK>(q);
#pragma unroll
  for (int r = 0; r < 4; ++r) p.v[r] = __int_as_float(q.v[r]);
}

template <bool WIDE>
__device__ __forceinline__ void ct_layer_f(F4& p, float z, float zq) {
#pragma unroll
  for (int s0 = 0; s0 < 2; ++s0) {
    const float t = WIDE ? f32::mulc_wide(p.v[2 | s0], z, zq) : f32::mulc(p.v[2 | s0], z, zq);
    const float x = p.v[s0];
    p.v[2 | s0] = x - t;
    p.v[s0] = x + t;
  }
}

template <bool RED>
__device__ __forceinline__ void gs_layer_f(F4& p, float z, float zq) {
#pragma unroll
  for (int s0 = 0; s0 < 2; ++s0) {
    const float x = p.v[s0], y = p.v[2 | s0];
    const float s = x + y;
    p.v[s0] = RED ? f32::reduce_f(s) : s;
    p.v[2 | s0] = f32::mulc(x - y, z, zq);
  }
}

// Forward transform (poly_ntt without its final Barrett reduction) of values
// bounded by B0: NAT in, NTT layout out, results bounded by fwd_out_bound(B0).
template <int B0>
__device__ __forceinline__ void ntt_f(F4& p, const LaneZF& z) {
  constexpr unsigned W = f32::fwd_wide_mask(B0);
  static_assert(f32::fwd_schedule_ok(B0, W), "forward schedule");
  swap_f<32>(p); ct_layer_f<((W >> 0) & 1) != 0>(p, z.f[0], z.fq[0]);
  swap_f<16>(p); ct_layer_f<((W >> 1) & 1) != 0>(p, z.f[1], z.fq[1]);
  swap_f<8>(p);  ct_layer_f<((W >> 2) & 1) != 0>(p, z.f[2], z.fq[2]);
  swap_f<4>(p);  ct_layer_f<((W >> 3) & 1) != 0>(p, z.f[3], z.fq[3]);
  swap_f<2>(p);  ct_layer_f<((W >> 4) & 1) != 0>(p, z.f[4], z.fq[4]);
  swap_f<1>(p);  ct_layer_f<((W >> 5) & 1) != 0>(p, z.f[5], z.fq[5]);
  swap_f<32>(p); ct_layer_f<((W >> 6) & 1) != 0>(p, z.f[6], z.fq[6]);
}

// Inverse transform (poly_invntt) of values bounded by 1665: NTT layout in,
// NAT out, results bounded by 1665.
__device__ __forceinline__ void invntt_f(F4& p, const LaneZF& z) {
  constexpr unsigned R = f32::kInvReduce;
  gs_layer_f<((R >> 0) & 1) != 0>(p, z.i[0], z.iq[0]);
  swap_f<32>(p); gs_layer_f<((R >> 1) & 1) != 0>(p, z.i[1], z.iq[1]);
  swap_f<1>(p);  gs_layer_f<((R >> 2) & 1) != 0>(p, z.i[2], z.iq[2]);
  swap_f<2>(p);  gs_layer_f<((R >> 3) & 1) != 0>(p, z.i[3], z.iq[3]);
  swap_f<4>(p);  gs_layer_f<((R >> 4) & 1) != 0>(p, z.i[4], z.iq[4]);
  swap_f<8>(p);  gs_layer_f<((R >> 5) & 1) != 0>(p, z.i[5], z.iq[5]);
  swap_f<16>(p); gs_layer_f<((R >> 6) & 1) != 0>(p, z.i[6], z.iq[6]);
  swap_f<32>(p);
#pragma unroll
  for (int r = 0; r < 4; ++r) p.v[r] = f32::mulc(p.v[r], kInvScaleF, kInvScaleQ);
}

__device__ __forceinline__ void reduce_f4(F4& p) {
#pragma unroll
  for (int r = 0; r < 4; ++r) p.v[r] = f32::reduce_f(p.v[r]);
}

// b1 zeta for the lane's two pairs (the basemul zeta's sign follows slot s1),
// from a reduced b.
struct Bz {
  float v[2];
};
__device__ __forceinline__ Bz basemul_twist(const F4& b, const LaneZF& z) {
  Bz r;
  r.v[0] = f32::mulc(b.v[1], z.bm, z.bmq);
  r.v[1] = f32::mulc(b.v[3], -z.bm, -z.bmq);
  return r;
}

// acc += a o b (basemul2 over the lane's two pairs, poly.cu:204-225), without
// reduction: a, b and bz reduced (|.| <= 1665); at most kBasemulTerms calls
// accumulate into one acc that starts at 0 before it is reduced.
__device__ __forceinline__ void basemul_acc_f(F4& acc, const F4& a, const F4& b, const Bz& bz) {
#pragma unroll
  for (int s1 = 0; s1 < 2; ++s1) {
    const float a0 = a.v[2 * s1], a1 = a.v[2 * s1 + 1];
    const float b0 = b.v[2 * s1], b1 = b.v[2 * s1 + 1];
    acc.v[2 * s1] = __builtin_fmaf(a1, bz.v[s1], __builtin_fmaf(a0, b0, acc.v[2 * s1]));
    acc.v[2 * s1 + 1] = __builtin_fmaf(a0, b1, __builtin_fmaf(a1, b0, acc.v[2 * s1 + 1]));
  }
}

// sum_j a_j o b_j for j < K, reduced; a_j supplied reduced by get_a(j).
template <int K, class GetA>
__device__ __forceinline__ F4 basemul_sum_f(GetA&& get_a, const F4 (&b)[K], const Bz (&bz)[K]) {
  F4 acc = {{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j > 0 && j % f32::kBasemulTerms == 0) reduce_f4(acc);
    basemul_acc_f(acc, get_a(j), b[j], bz[j]);
  }
  reduce_f4(acc);
  return acc;
}

__device__ __forceinline__ F4 to_f4(const P4& p) {
  F4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.v[i] = (float)p.v[i];
  return r;
}

}  // namespace kyber
}  // namespace pqc
