// Fused persistent Kyber encapsulation / decapsulation.
//
// One workgroup = one wave64; a persistent grid walks tiles of up to 64
// requests.  Inside a tile the work is split two ways:
//   * per-request sponge chains (H(ct), G, PRF, KDF, and for encaps H(seed),
//     H(pk), H(ct)) run one request per lane, the Keccak state in VGPRs;
//   * the matrix A^T is generated in rounds of up to 64 SHAKE128 entries, one
//     entry per lane, rejection-sampled straight from the state registers into
//     LDS, and consumed by the wave-cooperative polynomial arithmetic
//     (one request at a time, 4 coefficients per lane, NTT butterflies through
//     register/lane-bit swaps; kyber_wave.cuh).
// Keccak lanes stay >= 98% busy: every request contributes K*K entries per
// round and one lane per sequential chain.
//
// Data flow per request follows the round-2 FO decapsulation assembled from
// the reference's pieces: indcpa_dec (Kyber_GPU_Batched/indcpa.cu:691-733),
// G = sha3_512 (fips202.cu:672-689), indcpa_enc (indcpa.cu:578-679),
// verify + cmov (verify.cu:15-50, here a wave-wide OR of coefficient
// mismatches, no branch on it), H = sha3_256 (fips202.cu:615-632) and
// KDF = SHAKE256 (symmetric.h:33).  Comparing compressed coefficients is
// comparing ciphertext bytes: every width packs 256 coefficients into whole
// bytes with no unused bits.
#include <algorithm>
#include <mutex>

#include "internal.h"
#include "req_check.h"
#include "kyber_wave.cuh"
#include "kyber_wavef.cuh"
#include "sponge.cuh"

namespace pqc {
namespace kyber {
namespace fused {

__constant__ LaneTwiddles c_lane_tw = make_lane_twiddles();
__constant__ LaneTwiddlesF c_lane_twf = kLaneTwF;

constexpr int kAStride = 262;   // int16 slots per A entry: 256 + guard, odd dword count
static_assert(kAStride >= N + 4, "guard slots for a word's four candidates past the 256th accept");
constexpr int kMaxBlocks = 32;  // SHAKE128 blocks per entry before giving up (never reached)

// A tile is 64 requests, one per lane in the per-request phases.  Its 64*K*K
// matrix entries are generated in K*K rounds of 64 (one per lane).  Entry
// E = t*K*K + i*K + j lives in ring slot E mod RING; a round's 64 entries plus
// the at most K*K-1 earlier entries of a request that straddles two rounds are
// exactly RING consecutive values of E, so no live slot is ever overwritten.
template <int K>
struct Tile {
  static constexpr int T = 64;                                  // requests per tile
  static constexpr int KK = K * K;
  static constexpr int NR = KK;                                 // rounds: 64*K*K / 64
  static constexpr int RING = (64 % KK == 0) ? 64 : 64 + KK - 1;  // 64, 72, 64 slots
  static constexpr int PRFB = (2 * K + 1) * 128;                // PRF bytes per request
};

template <int RING>
struct Lds {
  int16_t a[RING * kAStride];
  uint8_t msg[64][32];
  uint32_t fail[64];
};

// ---- wave-cooperative pieces -------------------------------------------------

// indcpa_dec inputs of one request, as this lane's words: the ciphertext's u and v
// groups and the secret key's 12-bit groups.
template <int K>
struct DecIn {
  uint32_t u[K][3], s[K][2], v[3];
};

template <int K>
__device__ __forceinline__ void dec_load(DecIn<K>& in, const uint8_t* ct, const uint8_t* sk) {
  using P = Params<K>;
  const int l = lane_id();
#pragma unroll
  for (int i = 0; i < K; ++i) {
    load_group<P::du>(in.u[i], ct + 32 * P::du * i, nat_group(l));
    load_h12(in.s[i], sk + 384 * i, ntt_group(l));
  }
  load_group<P::dv>(in.v, ct + P::polyvec_cbytes, nat_group(l));
}

// Pack the lane's 4 D-bit values with its odd neighbour's: even lanes write D bytes.
template <int D>
__device__ __forceinline__ void pack_pair(uint8_t* dst, const uint32_t (&t)[4]) {
  const int l = lane_id();
  uint64_t mine = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) mine |= (uint64_t)t[r] << (D * r);
  const uint64_t nb = (uint64_t)__shfl_xor((long long)mine, 1);
  if ((l & 1) == 0) {
    // 8 values, 8*D bits: low 4*D bits mine, next 4*D bits the neighbour's
    const uint64_t lo = mine | (nb << (4 * D));
    const uint64_t hi = (4 * D + 4 * D > 64) ? (nb >> (64 - 4 * D)) : 0;
    uint8_t* o = dst + D * (l >> 1);
#pragma unroll
    for (int b = 0; b < D; ++b) o[b] = (uint8_t)(b < 8 ? (lo >> (8 * b)) : (hi >> (8 * (b - 8))));
  }
}

// Re-encryption inputs of one request, as this lane's words: its 16-bit PRF
// groups (2K+1 noise polynomials), the public key's 12-bit groups and, for
// decapsulation, the ciphertext's u and v groups to compare with.
template <int K, bool DEC>
struct EncIn {
  uint32_t prf[2 * K + 1];
  uint32_t t[K][2];
  uint32_t cu[DEC ? K : 1][3], cv[3];
};

// PRF_LDS: the PRF rows are in LDS (the per-request kernels); otherwise in the
// global scratch, read through a global-address-space pointer.
template <int K, bool DEC, bool PRF_LDS = false>
__device__ __forceinline__ void enc_load(EncIn<K, DEC>& in, const uint8_t* prf, const uint8_t* pk, const uint8_t* ct) {
  using P = Params<K>;
  const int l = lane_id();
  if constexpr (PRF_LDS) {
    const uint16_t* prf16 = reinterpret_cast<const uint16_t*>(prf);
#pragma unroll
    for (int j = 0; j < 2 * K + 1; ++j) in.prf[j] = prf16[64 * j + l];
  } else {
    const global_t<uint16_t>* prf16 = gptr(reinterpret_cast<const uint16_t*>(prf));
#pragma unroll
    for (int j = 0; j < 2 * K + 1; ++j) in.prf[j] = prf16[64 * j + l];
  }
#pragma unroll
  for (int j = 0; j < K; ++j) load_h12(in.t[j], pk + 384 * j, ntt_group(l));
  if constexpr (DEC) {
#pragma unroll
    for (int i = 0; i < K; ++i) load_group<P::du>(in.cu[i], ct + 32 * P::du * i, nat_group(l));
    load_group<P::dv>(in.cv, ct + P::polyvec_cbytes, nat_group(l));
  }
}

// Sources of the A^T entries for the re-encryption: at.get(i, j, w0, w1) gives
// this lane's group of entry (i, j) as two words of two coefficients.
template <int K>
struct AtRing {  // the LDS ring of the 64-request tiles: entry e0 + i K + j in slot mod RING
  const int16_t* ring;
  int e0;
  __device__ __forceinline__ void get(int i, int j, uint32_t& w0, uint32_t& w1) const {
    const int slot = (e0 + i * K + j) % Tile<K>::RING;
    const uint32_t* e32 = reinterpret_cast<const uint32_t*>(ring + slot * kAStride + 4 * ntt_group(lane_id()));
    w0 = e32[0];
    w1 = e32[1];
  }
};

// ---- indcpa_dec and the re-encryption on exact FP32 values (kyber_wavef.cuh) -----
// Every polynomial value equals the reference's int16 value mod q (the
// reference's arithmetic never leaves the int16 range, so its values are exact
// residues too), and the outputs leave through compress_f, which reads only the
// residue: the message bits, the ciphertext and the mismatch bits are the
// reference's.

// indcpa_dec (indcpa.cu:691-733, poly_tomsg) for one request; the message is written to
// msg (LDS, 32 bytes).
template <int K>
__device__ __forceinline__ void coop_indcpa_dec_f(const DecIn<K>& in, const LaneZF& z, uint8_t* msg) {
  using P = Params<K>;
  const int l = lane_id();
  F4 u[K];
  Bz uz[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    uint32_t t[4];
    unpack_loaded<P::du>(t, in.u[i], nat_group(l));
#pragma unroll
    for (int r = 0; r < 4; ++r) u[i].v[r] = (float)decompress_d<P::du>(t[r]);  // [0, q)
    ntt_f<Q>(u[i], z);
    reduce_f4(u[i]);
    uz[i] = basemul_twist(u[i], z);
  }
  F4 mp = basemul_sum_f<K>(
      [&](int i) {
        P4 s;
        poly12_loaded(s, in.s[i], ntt_group(l));  // 12-bit, unreduced
        F4 sf = to_f4(s);
        reduce_f4(sf);
        return sf;
      },
      u, uz);
  invntt_f(mp, z);
  uint32_t tv[4];
  unpack_loaded<P::dv>(tv, in.v, nat_group(l));
  uint32_t nib = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) nib |= f32::compress_f<1>((float)decompress_d<P::dv>(tv[r]) - mp.v[r]) << r;
  const uint32_t other = (uint32_t)__shfl_xor((int)nib, 1);
  if ((l & 1) == 0) msg[l >> 1] = (uint8_t)(nib | (other << 4));
}

// Re-encryption of one request (indcpa_enc from sp-sampling on).  DEC: compare
// with ct and return the lane's mismatch bits.  ENC: write ct.
template <int K, bool DEC, class AT>
__device__ __forceinline__ uint32_t coop_reencrypt_f(const AT& at, const EncIn<K, DEC>& in, const uint8_t* msg,
                                                     uint8_t* ct, const LaneZF& z) {
  using P = Params<K>;
  const int l = lane_id();
  const int gn = ntt_group(l);
  F4 sp[K];
  Bz spz[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    P4 c;
    cbd_group(c, in.prf[j]);  // [-2, 2]
    sp[j] = to_f4(c);
    ntt_f<2>(sp[j], z);
    reduce_f4(sp[j]);
    spz[j] = basemul_twist(sp[j], z);
  }
  uint32_t mismatch = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    F4 acc = basemul_sum_f<K>(
        [&](int j) {
          uint32_t w0, w1;
          at.get(i, j, w0, w1);  // accepted candidates, [0, 19q)
          F4 a = {{(float)(w0 & 0xFFFF), (float)(w0 >> 16), (float)(w1 & 0xFFFF), (float)(w1 >> 16)}};
          reduce_f4(a);
          return a;
        },
        sp, spz);
    invntt_f(acc, z);
    P4 e;
    cbd_group(e, in.prf[K + i]);
    uint32_t c[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) c[r] = f32::compress_f<P::du>(acc.v[r] + (float)e.v[r]);
    if constexpr (DEC) {
      uint32_t t[4];
      unpack_loaded<P::du>(t, in.cu[i], nat_group(l));
#pragma unroll
      for (int r = 0; r < 4; ++r) mismatch |= c[r] ^ t[r];
    } else {
      pack_pair<P::du>(ct + 32 * P::du * i, c);
    }
  }
  {
    F4 acc = basemul_sum_f<K>(
        [&](int j) {
          P4 t;
          poly12_loaded(t, in.t[j], gn);
          F4 tf = to_f4(t);
          reduce_f4(tf);
          return tf;
        },
        sp, spz);
    invntt_f(acc, z);
    P4 e;
    cbd_group(e, in.prf[2 * K]);
    const uint32_t nib = (msg[l >> 1] >> (4 * (l & 1))) & 0xF;
    uint32_t c[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float kq = ((nib >> r) & 1) ? (float)((Q + 1) / 2) : 0.f;
      c[r] = f32::compress_f<P::dv>(acc.v[r] + (float)e.v[r] + kq);
    }
    if constexpr (DEC) {
      uint32_t t[4];
      unpack_loaded<P::dv>(t, in.cv, nat_group(l));
#pragma unroll
      for (int r = 0; r < 4; ++r) mismatch |= c[r] ^ t[r];
    } else {
      pack_pair<P::dv>(ct + P::polyvec_cbytes, c);
    }
  }
  return mismatch;
}

// ---- per-lane sponge pieces ---------------------------------------------------

// SHAKE128(rho || i || j) rejection-sampled into an LDS entry (gen_matrix_n,
// indcpa.cu:339-403, rej_uniform indcpa.cu:290-322).  The entry keeps the
// accepted 16-bit candidates themselves (< 19q): the reference stores
// v - (v >> 12) q, the same residue, and every consumer reduces the entry mod q
// when it converts it to FP32.  Every candidate is written without a branch:
// at the running position p, which advances past accepted ones only: a
// rejected candidate is overwritten by the next one, and whatever lies at or
// past the final count is never read.  p is clamped to slot 256 after each
// word's four candidates, so writes past the 256th land in slots 256..259
// (kAStride leaves 262).  Returns the accept count, clamped (256 unless the
// block cap was hit).
__device__ __forceinline__ int sample_entry(int16_t* dst, const uint64_t (&rho)[4], int x, int y, bool valid) {
  uint64_t A[25];
#pragma unroll
  for (int t = 0; t < 25; ++t) A[t] = (t < 4) ? rho[t] : 0;
  A[4] = (uint64_t)x | ((uint64_t)y << 8) | (0x1Full << 16);
  A[kRateShake128 - 1] ^= 0x8000000000000000ull;
  int16_t* p = dst;
  int16_t* const end = dst + N;
  for (int blk = 0; blk < kMaxBlocks; ++blk) {
    if (blk >= 4 && !__any(valid && p < end)) break;
    keccakf1600(A);
#pragma unroll
    for (int w = 0; w < kRateShake128; ++w) {
      const uint32_t half[2] = {(uint32_t)A[w], (uint32_t)(A[w] >> 32)};
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const uint32_t val = (h & 1) ? (half[h >> 1] >> 16) : (half[h >> 1] & 0xFFFFu);
        *p = (int16_t)val;
        p += val < 19u * Q ? 1 : 0;
      }
      p = p < end ? p : end;
    }
  }
  return (int)(p - dst);
}

// Where the requests of a tile live.  DenseSrc: element-contiguous arrays of
// one parameter set (element r at r * size).  MixedSrc: requests of mixed
// parameter sets named by descriptors (pqc_kem_req: alg + byte offsets into
// four arenas), bucketed by K into `order` by k_mixed_classify; index r of the
// K bucket is request order[(K - 2) * cap + r].
struct ReqP {
  uint8_t* ss;
  uint8_t* ct;
  const uint8_t* key;
  const uint8_t* coins;
  uint8_t* prf;
};

struct DenseSrc {
  uint8_t* ss;
  uint8_t* ct;
  const uint8_t* key;
  const uint8_t* coins;
  uint8_t* prf;
  template <int K, bool DEC>
  __device__ __forceinline__ ReqP req(int r) const {
    using P = Params<K>;
    return ReqP{ss + (size_t)r * 32, ct + (size_t)r * P::ct_bytes,
                key + (size_t)r * (DEC ? P::sk_bytes : P::pk_bytes), coins ? coins + (size_t)r * 32 : nullptr,
                prf + (size_t)r * Tile<K>::PRFB};
  }
};

struct MixedSrc {
  const pqc_kem_req* reqs;
  const uint32_t* order;
  int cap;
  uint8_t* ss;
  uint8_t* ct;
  const uint8_t* key;
  const uint8_t* coins;
  uint8_t* prf;       // prf_stride bytes per request index
  size_t prf_stride;  // >= Tile<K>::PRFB of every K the context holds
  template <int K, bool DEC>
  __device__ __forceinline__ ReqP req(int r) const {
    const uint32_t idx = order[(size_t)(K - 2) * cap + r];
    const pqc_kem_req d = reqs[idx];
    return ReqP{ss + d.ss_off, ct + d.ct_off, key + d.key_off, coins ? coins + d.coins_off : nullptr,
                prf + (size_t)idx * prf_stride};
  }
};

// Request t of the tile, from the lane that holds it (t wave-uniform): no
// descriptor reload for the cooperative phases.
__device__ __forceinline__ ReqP lane_req(const ReqP& me, int t) {
  auto rd = [t](const void* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, t), hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), t);
    return reinterpret_cast<uint8_t*>(((uint64_t)hi << 32) | lo);
  };
  return ReqP{rd(me.ss), rd(me.ct), rd(me.key), rd(me.coins), rd(me.prf)};
}

// One tile of up to 64 requests (indices base .. base + 63 of the source, n of
// them valid) through the whole FO encapsulation (DEC false) or decapsulation.
// The cooperative phases load request t + 1's inputs while request t computes
// (one wave per SIMD: nothing else hides the loads' latency).
template <int K, bool DEC, class Src>
__device__ __forceinline__ void fused_tile(const Src& src, int base, int n, Lds<Tile<K>::RING>& sh, const LaneZF& z,
                                           uint32_t* __restrict__ status) {
  using P = Params<K>;
  using TL = Tile<K>;
  constexpr int PK_OFF = DEC ? P::sk_cpa_bytes : 0;  // public key inside the key buffer
  const int lane = threadIdx.x;
  const int r = base + lane;
  const bool active = lane < TL::T && r < n;
  const int rc = active ? r : min(base, n - 1);  // clamped request for address math
  const ReqP me = src.template req<K, DEC>(rc);
  uint64_t kbar[4], hct[4];

  if constexpr (DEC) {
    // ---- H(ct), one request per lane
    {
      uint64_t A[25];
      sponge_absorb_words_prefetch<kRateSha3_256>(A, reinterpret_cast<const uint64_t*>(me.ct), P::ct_bytes / 8, 0x06);
#pragma unroll
      for (int w = 0; w < 4; ++w) hct[w] = A[w];
    }
    // ---- indcpa_dec, wave-cooperative, one request at a time
    const int nt = min(TL::T, n - base);
    // request t's words were loaded during request t - 1: the copy at the top of
    // the iteration is the only wait for them, and the prefetch issued after it
    // is not waited for until the next iteration
    DecIn<K> cur, nxt;
    {
      const ReqP q = lane_req(me, 0);
      dec_load<K>(nxt, q.ct, q.key);
    }
    for (int t = 0; t < nt; ++t) {
      cur = nxt;
      if (t + 1 < nt) {
        const ReqP q = lane_req(me, t + 1);
        dec_load<K>(nxt, q.ct, q.key);
      }
      coop_indcpa_dec_f<K>(cur, z, sh.msg[t]);
    }
    __syncthreads();
  } else {
    // ---- m = H(seed) ("don't release system RNG output"), H(pk)
    uint64_t A[25];
    uint64_t c4[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) c4[w] = load_u64(me.coins + 8 * w);
    absorb_fixed<kRateSha3_256, 4>(A, c4, 0x06);
    if (active) {
#pragma unroll
      for (int w = 0; w < 4; ++w) store_u64(sh.msg[lane] + 8 * w, A[w]);
    }
    sponge_absorb_words_prefetch<kRateSha3_256>(A, reinterpret_cast<const uint64_t*>(me.key), P::pk_bytes / 8, 0x06);
#pragma unroll
    for (int w = 0; w < 4; ++w) hct[w] = A[w];  // holds H(pk) until G
    __syncthreads();
  }

  // ---- (Kbar, coins) = G(m || H(pk)); PRF(coins, 0..2K) -> scratch
  {
    uint64_t m[8], A[25];
    const uint8_t* msg = sh.msg[lane < TL::T ? lane : 0];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      m[w] = load_u64(msg + 8 * w);
      m[4 + w] = DEC ? load_u64_global(me.key + P::hpk_off + 8 * w) : hct[w];
    }
    absorb_fixed<kRateSha3_512, 8>(A, m, 0x06);
    uint64_t rc4[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      kbar[w] = A[w];
      rc4[w] = A[4 + w];
    }
    for (int nonce = 0; nonce < 2 * K + 1; ++nonce) {
      absorb_fixed<kRateShake256, 4>(A, rc4, (uint64_t)nonce | (0x1Full << 8));
      if (active) {
#pragma unroll
        for (int w = 0; w < 16; ++w) store_u64(me.prf + 128 * nonce + 8 * w, A[w]);
      }
    }
  }
  __syncthreads();

  // ---- A^T rounds + re-encryption
  const int nt = min(TL::T, n - base);
  int t_done = 0;  // requests whose re-encryption is finished
  EncIn<K, DEC> cur, nxt;
  // rho of this lane's entry in round g: the public seed of request E / KK, its
  // key pointer taken from the lane that holds it (invalid entries use lane 0's)
  auto load_rho = [&](uint64_t (&rho)[4], int g) {
    const int t = (64 * g + lane) / TL::KK;
    const uint8_t* key = reinterpret_cast<const uint8_t*>(__shfl((long long)me.key, t < nt ? t : 0));
#pragma unroll
    for (int w = 0; w < 4; ++w) rho[w] = load_u64_global(key + PK_OFF + P::polyvec_bytes + 8 * w);
  };
  uint64_t rho_next[4];
  load_rho(rho_next, 0);
  // a partial tile needs only the rounds that hold its nt * K^2 entries (a lone
  // request: one round instead of K^2, each at least 4 Keccak-f)
  const int ng = min(TL::NR, (nt * TL::KK + 63) / 64);
  for (int g = 0; g < ng; ++g) {
    // requests all of whose entries are sampled by the end of this round
    const int t_end = min(nt, (64 * (g + 1)) / TL::KK);
    if (t_done < t_end) {  // the round's first request loads during the sampling
      const ReqP q = lane_req(me, t_done);
      enc_load<K, DEC>(cur, q.prf, q.key + PK_OFF, q.ct);
    }
    {
      const int E = 64 * g + lane;
      const int t = E / TL::KK;
      const int ij = E % TL::KK;
      const bool valid = t < nt;
      uint64_t rho[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) rho[w] = rho_next[w];
      const int got = sample_entry(sh.a + (E % TL::RING) * kAStride, rho, ij / K, ij % K, valid);
      if (valid && got < N) atomicOr(status, 1u);
    }
    __syncthreads();
    // the next round's rho loads during this round's re-encryptions
    if (g + 1 < ng) load_rho(rho_next, g + 1);
    for (int t = t_done; t < t_end; ++t) {
      if (t + 1 < t_end) {
        const ReqP q = lane_req(me, t + 1);
        enc_load<K, DEC>(nxt, q.prf, q.key + PK_OFF, q.ct);
      }
      const ReqP q = lane_req(me, t);
      const uint32_t mm = coop_reencrypt_f<K, DEC>(AtRing<K>{sh.a, t * TL::KK}, cur, sh.msg[t], q.ct, z);
      cur = nxt;
      if constexpr (DEC) {
        const uint32_t f = __any(mm != 0) ? 1u : 0u;
        if (lane == 0) sh.fail[t] = f;
      }
    }
    t_done = max(t_done, t_end);
    __syncthreads();
  }

  // ---- KDF
  {
    uint64_t m[8], A[25];
    if constexpr (DEC) {
      const uint64_t mask = 0 - (uint64_t)(sh.fail[lane < TL::T ? lane : 0] & 1u);
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const uint64_t zz = load_u64_global(me.key + P::z_off + 8 * w);
        m[w] = kbar[w] ^ (mask & (zz ^ kbar[w]));
        m[4 + w] = hct[w];
      }
    } else {
      sponge_absorb_words_prefetch<kRateSha3_256>(A, reinterpret_cast<const uint64_t*>(me.ct), P::ct_bytes / 8, 0x06);
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        m[w] = kbar[w];
        m[4 + w] = A[w];
      }
    }
    absorb_fixed<kRateShake256, 8>(A, m, 0x1F);
    if (active) {
#pragma unroll
      for (int w = 0; w < 4; ++w) store_u64(me.ss + 8 * w, A[w]);
    }
  }
  __syncthreads();
}

// ---- decapsulation in 32-request tiles, two waves per SIMD --------------------------
// fused_tile runs one wave per SIMD: a batch of 65536 is 1024 tiles, one per SIMD,
// and its A^T ring takes 40 KB of LDS, so four waves fill a CU's 160 KB.  A lone
// wave issues a VALU instruction every ~4.8 cycles; two waves per SIMD issue the
// same Keccak mix at ~3.9 cycles per instruction per SIMD (measured on MI355X).
// This tile serves 32 requests, so 65536 requests are 2048 waves, two per SIMD,
// and its A^T entries go to a per-wave global scratch (written and read back by
// the same wave while the lines are in L2) instead of LDS: ~1.2 KB of LDS per
// wave.  So that the per-request sponge chains still keep all 64 lanes busy with
// 32 requests, lanes 0..31 absorb H(ct) of request `lane` while lanes 32..63
// compute G and the 2K + 1 PRF outputs of request `lane - 32` in the same
// permutations (H(ct) is the longer chain for every K).  Same bytes as fused_tile.
template <int K>
struct Tile32 {
  static constexpr int T = 32;
  static constexpr int KK = K * K;
  static constexpr int NR = (T * KK + 63) / 64;                              // sampling rounds: 2, 5, 8
  static constexpr int HB = Params<K>::ct_bytes / 8 / kRateSha3_256 + 1;     // H(ct) blocks: 6, 9, 12
  static constexpr int ENTRIES = 64 * NR;                                    // scratch entries per wave
  static_assert(HB >= 2 * K + 2, "the H(ct) lanes run the longest chain of the sponge phase");
};
constexpr int kAtStride = 264;  // int16 per A^T entry in the scratch: 256 + guard slots, 16-byte multiple
static_assert(kAtStride >= N + 4 && (kAtStride * 2) % 16 == 0, "entry stride");

// Sampling stage: one row per lane.  A block's candidates go to the row at their
// entry position minus base8 (the entry's last 8-aligned position before the
// block), so that whole 16-byte chunks can be copied to the global entry; the
// row holds the 0..7 earlier coefficients of that chunk and the block's <= 84
// candidates written at the running position (the last write at most at index
// 7 + 83), and the chunk copies read up to the next multiple of 8.  A 208-byte
// row (13 x 16 B) spreads the lanes' rows over 16 bank offsets.
constexpr int kStageRow = 104;  // int16 per lane row
static_assert(7 + 84 + 7 <= kStageRow && kStageRow % 8 == 0 && (kStageRow / 8) % 2 == 1, "stage row");

struct Lds32 {
  alignas(16) int16_t stage[64][kStageRow];
  uint8_t msg[32][32];
  uint32_t fail[32];
};

// SHAKE128(rho || x || y) rejection-sampled into the global entry `dst` (as
// sample_entry: the accepted candidates themselves), through this lane's LDS
// stage row: after each block the chunks
// [base8, round_up_8(ctr)) are copied to dst with 16-byte stores, and the last,
// partial chunk moves to the front of the row for the next block.  So every
// global store is a whole 16 bytes of one entry (a 2-byte store per candidate
// would be one L2 request each).  Returns the accept count, clamped at 256.
__device__ __forceinline__ int sample_entry_staged(global_t<int16_t>* dst, int16_t* row, const uint64_t (&rho)[4],
                                                   int x, int y, bool valid) {
  uint64_t A[25];
#pragma unroll
  for (int t = 0; t < 25; ++t) A[t] = (t < 4) ? rho[t] : 0;
  A[4] = (uint64_t)x | ((uint64_t)y << 8) | (0x1Full << 16);
  A[kRateShake128 - 1] ^= 0x8000000000000000ull;
  int ctr = 0, base8 = 0;
  for (int blk = 0; blk < kMaxBlocks; ++blk) {
    if (blk >= 4 && !__any(valid && ctr < N)) break;
    keccakf1600(A);
    int16_t* p = row + (ctr - base8);
    int16_t* const pend = row + (N - base8);  // position N: the clamp
#pragma unroll
    for (int w = 0; w < kRateShake128; ++w) {
      const uint32_t half[2] = {(uint32_t)A[w], (uint32_t)(A[w] >> 32)};
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const uint32_t val = (h & 1) ? (half[h >> 1] >> 16) : (half[h >> 1] & 0xFFFFu);
        *p = (int16_t)val;  // overwritten by the next candidate unless accepted
        p += val < 19u * Q ? 1 : 0;
      }
      p = p < pend ? p : pend;
    }
    ctr = base8 + (int)(p - row);
    // whole 16-byte chunks of [base8, ctr): at most (7 + 84 + 7) / 8 = 12
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const int nchunks = (ctr - base8 + 7) >> 3;
#pragma unroll
    for (int c = 0; c < (7 + 84 + 7) / 8; ++c) {
      if (c < nchunks) {
        const u32x4 v = *reinterpret_cast<const u32x4*>(row + 8 * c);
        *(global_t<u32x4>*)(dst + base8 + 8 * c) = v;
      }
    }
    // the partial chunk that holds position ctr starts the row for the next block
    const int keep = (ctr & ~7) - base8;
    const u32x4 t = *reinterpret_cast<const u32x4*>(row + keep);
    *reinterpret_cast<u32x4*>(row) = t;
    base8 = ctr & ~7;
  }
  return ctr;
}

// This lane's group of the A^T entries of one request, loaded from the scratch
// ahead of their use (the entries of request t start at entry t K^2).
template <int K>
struct AtRegs {
  uint32_t w[K * K][2];
  __device__ __forceinline__ void load(const global_t<int16_t>* at, int t) {
    const global_t<uint32_t>* p =
        (const global_t<uint32_t>*)(at + (size_t)t * K * K * kAtStride + 4 * ntt_group(lane_id()));
#pragma unroll
    for (int e = 0; e < K * K; ++e) {
      w[e][0] = p[e * (kAtStride / 2)];
      w[e][1] = p[e * (kAtStride / 2) + 1];
    }
  }
  __device__ __forceinline__ void get(int i, int j, uint32_t& w0, uint32_t& w1) const {
    w0 = w[i * K + j][0];
    w1 = w[i * K + j][1];
  }
};

// The A^T rounds of a 32-request tile: round g samples entries 64 g .. 64 g + 63
// (entry E = request E / K^2, position E % K^2) into the scratch, then re-encrypts
// every request whose entries are all sampled (request t's words load during
// request t - 1).  Decapsulation compares with the ciphertext and sets sh.fail[t];
// encapsulation writes the ciphertext.
template <int K, bool DEC>
__device__ __forceinline__ void at_reencrypt32(const ReqP& me, int nt, Lds32& sh, global_t<int16_t>* at,
                                               const LaneZF& z, uint32_t* __restrict__ status) {
  using P = Params<K>;
  using TL = Tile32<K>;
  constexpr int PK_OFF = DEC ? P::sk_cpa_bytes : 0;  // public key inside the key buffer
  const int lane = threadIdx.x;
  const int ng = (nt * TL::KK + 63) / 64;
  int t_done = 0;
  auto load_rho = [&](uint64_t (&rho)[4], int g) {
    const int t = (64 * g + lane) / TL::KK;
    const uint8_t* key = reinterpret_cast<const uint8_t*>(__shfl((long long)me.key, t < nt ? t : 0));
#pragma unroll
    for (int w = 0; w < 4; ++w) rho[w] = load_u64_global(key + PK_OFF + P::polyvec_bytes + 8 * w);
  };
  uint64_t rho_next[4];
  load_rho(rho_next, 0);
  for (int g = 0; g < ng; ++g) {
    {
      const int E = 64 * g + lane;
      const int t = E / TL::KK;
      const int ij = E % TL::KK;
      const bool valid = t < nt;
      uint64_t rho[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) rho[w] = rho_next[w];
      const int got = sample_entry_staged(at + (size_t)E * kAtStride, sh.stage[lane], rho, ij / K, ij % K, valid);
      if (valid && got < N) atomicOr(status, 1u);
    }
    __syncthreads();  // the round's entries are in the scratch for every lane
    if (g + 1 < ng) load_rho(rho_next, g + 1);
    const int t_end = min(nt, (64 * (g + 1)) / TL::KK);
    if (t_done < t_end) {
      EncIn<K, DEC> cur, nxt;
      {
        const ReqP q = lane_req(me, t_done);
        enc_load<K, DEC>(nxt, q.prf, q.key + PK_OFF, q.ct);
      }
      for (int t = t_done; t < t_end; ++t) {
        cur = nxt;
        if (t + 1 < t_end) {
          const ReqP q = lane_req(me, t + 1);
          enc_load<K, DEC>(nxt, q.prf, q.key + PK_OFF, q.ct);
        }
        AtRegs<K> a;
        a.load(at, t);
        if constexpr (DEC) {
          const uint32_t mm = coop_reencrypt_f<K, true>(a, cur, sh.msg[t], nullptr, z);
          const uint32_t f = __any(mm != 0) ? 1u : 0u;
          if (lane == 0) sh.fail[t] = f;
        } else {
          coop_reencrypt_f<K, false>(a, cur, sh.msg[t], lane_req(me, t).ct, z);
        }
      }
      t_done = t_end;
    }
  }
}

template <int K, class Src>
__device__ __forceinline__ void dec_tile32(const Src& src, int base, int n, Lds32& sh, global_t<int16_t>* at,
                                           const LaneZF& z, uint32_t* __restrict__ status) {
  using P = Params<K>;
  using TL = Tile32<K>;
  const int lane = threadIdx.x;
  const int rq = lane & 31;  // the request this lane serves in the sponge phases (lanes l and l + 32 alike)
  const int nt = min(TL::T, n - base);
  const bool act = rq < nt;
  const ReqP me = src.template req<K, true>(act ? base + rq : base);
  const bool hl = lane < 32;  // H(ct) lane; the others compute G and the PRF outputs

  // ---- indcpa_dec, wave-cooperative, one request at a time; request t's words are
  // loaded during request t - 1
  {
    DecIn<K> cur, nxt;
    {
      const ReqP q = lane_req(me, 0);
      dec_load<K>(nxt, q.ct, q.key);
    }
    for (int t = 0; t < nt; ++t) {
      cur = nxt;
      if (t + 1 < nt) {
        const ReqP q = lane_req(me, t + 1);
        dec_load<K>(nxt, q.ct, q.key);
      }
      coop_indcpa_dec_f<K>(cur, z, sh.msg[t]);
    }
  }
  __syncthreads();

  // ---- one sponge phase: lanes 0..31 H(ct) (SHA3-256, HB blocks); lanes 32..63
  // step 0: (Kbar, coins) = G(m || H(pk)), steps 1 .. 2K+1: PRF(coins, step - 1)
  uint64_t hct[4], kbar[4] = {0, 0, 0, 0};
  {
    constexpr int W = P::ct_bytes / 8;
    constexpr int R = kRateSha3_256;
    static_assert(W >= R, "the first H(ct) block is whole");
    const global_t<uint64_t>* cw = gptr(reinterpret_cast<const uint64_t*>(me.ct));
    uint64_t A[25], cur[R], coins[4];
#pragma unroll
    for (int i = 0; i < 25; ++i) A[i] = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) coins[w] = 0;
    if (hl) {
#pragma unroll
      for (int w = 0; w < R; ++w) cur[w] = cw[w];
    } else {
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        A[w] = load_u64(sh.msg[rq] + 8 * w);
        A[4 + w] = load_u64_global(me.key + P::hpk_off + 8 * w);
      }
      A[8] ^= 0x06;
      A[kRateSha3_512 - 1] ^= 0x8000000000000000ull;
    }
    for (int st = 0; st < TL::HB; ++st) {
      if (hl) {
        const int rem = W - st * R;  // message words left (uniform), >= 0
#pragma unroll
        for (int w = 0; w < R; ++w) {
          if (w < rem) A[w] ^= cur[w];
          if (w == rem) A[w] ^= 0x06;
        }
        if (st == TL::HB - 1) A[R - 1] ^= 0x8000000000000000ull;
        // the next block's words load during this permutation
#pragma unroll
        for (int w = 0; w < R; ++w)
          if (w < rem - R) cur[w] = cw[(st + 1) * R + w];
      } else if (st >= 1 && st <= 2 * K + 1) {
#pragma unroll
        for (int i = 0; i < 25; ++i) A[i] = i < 4 ? coins[i] : 0;
        A[4] = (uint64_t)(st - 1) | (0x1Full << 8);
        A[kRateShake256 - 1] ^= 0x8000000000000000ull;
      }
      keccakf1600(A);
      if (!hl) {
        if (st == 0) {
#pragma unroll
          for (int w = 0; w < 4; ++w) {
            kbar[w] = A[w];
            coins[w] = A[4 + w];
          }
        } else if (st <= 2 * K + 1 && act) {
#pragma unroll
          for (int w = 0; w < 16; ++w) store_u64(me.prf + 128 * (st - 1) + 8 * w, A[w]);
        }
      }
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) hct[w] = A[w];  // lanes 0..31
  }
  __syncthreads();

  // ---- A^T rounds of 64 entries into the scratch, then the re-encryption of every
  // request whose entries are complete
  at_reencrypt32<K, true>(me, nt, sh, at, z, status);
  __syncthreads();

  // ---- KDF(Kbar or z, H(ct)) on lanes 0..31; Kbar comes from lane + 32
  {
    uint64_t m[8], A[25];
    const uint64_t mask = 0 - (uint64_t)(sh.fail[rq] & 1u);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const uint64_t kb = (uint64_t)__shfl_xor((long long)kbar[w], 32);
      const uint64_t zz = load_u64_global(me.key + P::z_off + 8 * w);
      m[w] = kb ^ (mask & (zz ^ kb));
      m[4 + w] = hct[w];
    }
    absorb_fixed<kRateShake256, 8>(A, m, 0x1F);
    if (hl && act) {
#pragma unroll
      for (int w = 0; w < 4; ++w) store_u64(me.ss + 8 * w, A[w]);
    }
  }
  __syncthreads();
}

// Encapsulation in the same 32-request tiles.  The first sponge phase runs
// H(pk) (the longest chain) on lanes 0..31 while lanes 32..63 compute
// m = H(coins) ("don't release system RNG output") in its first permutation;
// then both lanes of a pair compute G(m || H(pk)) and split the 2K + 1 PRF
// outputs (lane l the even nonces, lane l + 32 the odd ones); the A^T rounds
// and re-encryptions write the ciphertexts; H(ct) and the KDF end it.  Same
// bytes as fused_tile<K, false>.
template <int K>
struct EncTile32 {
  static constexpr int HB = Params<K>::pk_bytes / 8 / kRateSha3_256 + 1;  // H(pk) blocks: 6, 9, 12
  static_assert(Params<K>::pk_bytes / 8 >= kRateSha3_256, "the first H(pk) block is whole");
};

template <int K, class Src>
__device__ __forceinline__ void enc_tile32(const Src& src, int base, int n, Lds32& sh, global_t<int16_t>* at,
                                           const LaneZF& z, uint32_t* __restrict__ status) {
  using P = Params<K>;
  using TL = Tile32<K>;
  const int lane = threadIdx.x;
  const int rq = lane & 31;  // the request this lane serves in the sponge phases (lanes l and l + 32 alike)
  const int nt = min(TL::T, n - base);
  const bool act = rq < nt;
  const ReqP me = src.template req<K, false>(act ? base + rq : base);
  const bool hl = lane < 32;  // H(pk) lane; the others compute m = H(coins)

  // ---- lanes 0..31: H(pk) (SHA3-256, HB blocks); lanes 32..63: m = H(coins)
  uint64_t hpk[4];
  {
    constexpr int W = P::pk_bytes / 8;
    constexpr int R = kRateSha3_256;
    const global_t<uint64_t>* pw = gptr(reinterpret_cast<const uint64_t*>(me.key));
    uint64_t A[25], cur[R];
#pragma unroll
    for (int i = 0; i < 25; ++i) A[i] = 0;
    if (hl) {
#pragma unroll
      for (int w = 0; w < R; ++w) cur[w] = pw[w];
    } else {
#pragma unroll
      for (int w = 0; w < 4; ++w) A[w] = load_u64_global(me.coins + 8 * w);
      A[4] ^= 0x06;
      A[R - 1] ^= 0x8000000000000000ull;
    }
    for (int st = 0; st < EncTile32<K>::HB; ++st) {
      if (hl) {
        const int rem = W - st * R;  // message words left (uniform), >= 0
#pragma unroll
        for (int w = 0; w < R; ++w) {
          if (w < rem) A[w] ^= cur[w];
          if (w == rem) A[w] ^= 0x06;
        }
        if (st == EncTile32<K>::HB - 1) A[R - 1] ^= 0x8000000000000000ull;
        // the next block's words load during this permutation
#pragma unroll
        for (int w = 0; w < R; ++w)
          if (w < rem - R) cur[w] = pw[(st + 1) * R + w];
      }
      keccakf1600(A);
      if (!hl && st == 0 && act) {
#pragma unroll
        for (int w = 0; w < 4; ++w) store_u64(sh.msg[rq] + 8 * w, A[w]);
      }
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) hpk[w] = A[w];  // lanes 0..31
  }
  __syncthreads();

  // ---- (Kbar, coins) = G(m || H(pk)) on both lanes of a pair, then PRF(coins,
  // 2 i + (lane >= 32)) for i = 0 .. K (lane l + 32's last one, nonce 2K + 1, unused)
  uint64_t kbar[4];
  {
    uint64_t m[8], A[25], coins[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      m[w] = load_u64(sh.msg[rq] + 8 * w);
      m[4 + w] = (uint64_t)__shfl((long long)hpk[w], rq);
    }
    absorb_fixed<kRateSha3_512, 8>(A, m, 0x06);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      kbar[w] = A[w];
      coins[w] = A[4 + w];
    }
    for (int i = 0; i <= K; ++i) {
      const int nonce = 2 * i + (hl ? 0 : 1);
      absorb_fixed<kRateShake256, 4>(A, coins, (uint64_t)nonce | (0x1Full << 8));
      if (act && nonce <= 2 * K) {
#pragma unroll
        for (int w = 0; w < 16; ++w) store_u64(me.prf + 128 * nonce + 8 * w, A[w]);
      }
    }
  }
  __syncthreads();

  // ---- A^T rounds and the re-encryptions, which write the ciphertexts
  at_reencrypt32<K, false>(me, nt, sh, at, z, status);
  __syncthreads();

  // ---- KDF(Kbar, H(ct)); both lanes of a pair compute it, lanes 0..31 store
  {
    uint64_t m[8], A[25];
    sponge_absorb_words_prefetch<kRateSha3_256>(A, reinterpret_cast<const uint64_t*>(me.ct), P::ct_bytes / 8, 0x06);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      m[w] = kbar[w];
      m[4 + w] = A[w];
    }
    absorb_fixed<kRateShake256, 8>(A, m, 0x1F);
    if (hl && act) {
#pragma unroll
      for (int w = 0; w < 4; ++w) store_u64(me.ss + 8 * w, A[w]);
    }
  }
  __syncthreads();
}

template <int K>
__global__ __launch_bounds__(64, 2) void k_fused_dec32(int n, int ntiles, DenseSrc src, int16_t* at_scratch,
                                                       uint32_t* __restrict__ status) {
  __shared__ Lds32 sh;
  LaneZF z;
  load_lane_zf(z, &c_lane_twf);
  global_t<int16_t>* at = (global_t<int16_t>*)(at_scratch + (size_t)blockIdx.x * Tile32<K>::ENTRIES * kAtStride);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) dec_tile32<K>(src, tile * 32, n, sh, at, z, status);
}

template <int K>
__global__ __launch_bounds__(64, 2) void k_fused_enc32(int n, int ntiles, DenseSrc src, int16_t* at_scratch,
                                                       uint32_t* __restrict__ status) {
  __shared__ Lds32 sh;
  LaneZF z;
  load_lane_zf(z, &c_lane_twf);
  global_t<int16_t>* at = (global_t<int16_t>*)(at_scratch + (size_t)blockIdx.x * Tile32<K>::ENTRIES * kAtStride);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) enc_tile32<K>(src, tile * 32, n, sh, at, z, status);
}

// Mixed batches: the bucketed requests (k_mixed_classify) in 32-request tiles at
// two waves per SIMD, K = 4 tiles first (the longest), then K = 3, then K = 2;
// each wave's A^T scratch region holds `region` int16 (sized for the largest K
// the context holds).
template <bool DEC>
__device__ __forceinline__ void mixed32_body(const MixedSrc& src, const uint32_t* __restrict__ counts,
                                             int16_t* at_scratch, size_t region, uint32_t* __restrict__ status) {
  __shared__ Lds32 sh;
  LaneZF z;
  load_lane_zf(z, &c_lane_twf);
  global_t<int16_t>* at = (global_t<int16_t>*)(at_scratch + (size_t)blockIdx.x * region);
  const int n2 = (int)counts[0], n3 = (int)counts[1], n4 = (int)counts[2];
  const int t4 = (n4 + 31) / 32, t3 = (n3 + 31) / 32, t2 = (n2 + 31) / 32;
  for (int tile = blockIdx.x; tile < t4 + t3 + t2; tile += gridDim.x) {
    if (tile < t4) {
      if constexpr (DEC) dec_tile32<4>(src, tile * 32, n4, sh, at, z, status);
      else enc_tile32<4>(src, tile * 32, n4, sh, at, z, status);
    } else if (tile < t4 + t3) {
      if constexpr (DEC) dec_tile32<3>(src, (tile - t4) * 32, n3, sh, at, z, status);
      else enc_tile32<3>(src, (tile - t4) * 32, n3, sh, at, z, status);
    } else {
      if constexpr (DEC) dec_tile32<2>(src, (tile - t4 - t3) * 32, n2, sh, at, z, status);
      else enc_tile32<2>(src, (tile - t4 - t3) * 32, n2, sh, at, z, status);
    }
  }
}

__global__ __launch_bounds__(64, 2) void k_fused_mixed_dec32(MixedSrc src, const uint32_t* __restrict__ counts,
                                                            int16_t* at_scratch, size_t region,
                                                            uint32_t* __restrict__ status) {
  mixed32_body<true>(src, counts, at_scratch, region, status);
}

__global__ __launch_bounds__(64, 2) void k_fused_mixed_enc32(MixedSrc src, const uint32_t* __restrict__ counts,
                                                            int16_t* at_scratch, size_t region,
                                                            uint32_t* __restrict__ status) {
  mixed32_body<false>(src, counts, at_scratch, region, status);
}

// ---- key generation -----------------------------------------------------------------
// indcpa_keypair (Kyber_GPU_Batched/indcpa.cu:474-564) + the KEM tail: (rho, sigma)
// = G(d); s_j = CBD(PRF(sigma, j)), e_i = CBD(PRF(sigma, K + i)), both through
// poly_ntt (ntt + Barrett); t_i = frommont(sum_j A[i][j] o s_j) + e_i, Barrett;
// pk = t || rho, sk = s || pk || H(pk) || z, polynomials packed 12 bits per
// coefficient after csubq (polyvec_tobytes).  A is not transposed:
// A[i][j] = SHAKE128(rho || j || i) (gen_matrix_n with transposed = 0).

// 12-bit pack of a group in its NTT-order slot (the inverse of load_poly12)
__device__ __forceinline__ void store_poly12(uint8_t* base, int g, const P4& p) {
  uint16_t* h = reinterpret_cast<uint16_t*>(base + 6 * g);
  const uint32_t c0 = (uint32_t)p.v[0], c1 = (uint32_t)p.v[1], c2 = (uint32_t)p.v[2], c3 = (uint32_t)p.v[3];
  h[0] = (uint16_t)(c0 | (c1 << 12));
  h[1] = (uint16_t)((c1 >> 4) | (c2 << 8));
  h[2] = (uint16_t)((c2 >> 8) | (c3 << 4));
}

template <int K>
struct KeyIn {
  uint32_t prf[2 * K];
};

template <int K>
__device__ __forceinline__ void key_load(KeyIn<K>& in, const uint8_t* prf) {
  const uint16_t* prf16 = reinterpret_cast<const uint16_t*>(prf);
#pragma unroll
  for (int j = 0; j < 2 * K; ++j) in.prf[j] = prf16[64 * j + lane_id()];
}

// one request's s and t, written into sk_cpa, pk and sk's copy of pk, on the
// exact FP32 arithmetic (kyber_wavef.cuh).  The reference computes
// t = frommont(A o s) + e (basemul carries R^-1, frommont multiplies by R^2 R^-1)
// and packs csubq'd residues; here A o s carries no factor, and canon_f gives
// the same residues in [0, q).
template <int K, class AT>
__device__ __forceinline__ void coop_keypair(const AT& at, const KeyIn<K>& in, uint8_t* pk, uint8_t* sk,
                                             const LaneZF& z) {
  using P = Params<K>;
  const int gn = ntt_group(lane_id());
  F4 sv[K];
  Bz svz[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    P4 c;
    cbd_group(c, in.prf[j]);  // [-2, 2]
    sv[j] = to_f4(c);
    ntt_f<2>(sv[j], z);
    reduce_f4(sv[j]);
    svz[j] = basemul_twist(sv[j], z);
#pragma unroll
    for (int r = 0; r < 4; ++r) c.v[r] = (int32_t)f32::canon_f(sv[j].v[r]);
    store_poly12(sk + 384 * j, gn, c);
  }
#pragma unroll
  for (int i = 0; i < K; ++i) {
    F4 acc = basemul_sum_f<K>(
        [&](int j) {
          uint32_t w0, w1;
          at.get(i, j, w0, w1);  // A[i][j]: accepted candidates, [0, 19q)
          F4 a = {{(float)(w0 & 0xFFFF), (float)(w0 >> 16), (float)(w1 & 0xFFFF), (float)(w1 >> 16)}};
          reduce_f4(a);
          return a;
        },
        sv, svz);
    P4 e;
    cbd_group(e, in.prf[K + i]);
    F4 ef = to_f4(e);
    ntt_f<2>(ef, z);
    static_assert(f32::fwd_out_bound(2) + f32::kOut <= f32::kCanonMax, "t = A o s + e within canon_f's range");
    P4 c;
#pragma unroll
    for (int r = 0; r < 4; ++r) c.v[r] = (int32_t)f32::canon_f(acc.v[r] + ef.v[r]);
    store_poly12(pk + 384 * i, gn, c);
    store_poly12(sk + P::sk_cpa_bytes + 384 * i, gn, c);
  }
}

struct KeySrc {
  uint8_t* pk;
  uint8_t* sk;
  const uint8_t* coins;  // 64 bytes per request: d || z
  uint8_t* prf;
};

template <int K>
__device__ __forceinline__ void fused_keypair_tile(const KeySrc& src, int base, int n, Lds<Tile<K>::RING>& sh,
                                                   const LaneZF& z, uint32_t* __restrict__ status) {
  using P = Params<K>;
  using TL = Tile<K>;
  const int lane = threadIdx.x;
  const int r = base + lane;
  const bool active = r < n;
  const int rc = active ? r : base;  // clamped request for address math
  uint8_t* pk = src.pk + (size_t)rc * P::pk_bytes;
  uint8_t* sk = src.sk + (size_t)rc * P::sk_bytes;
  uint8_t* prf = src.prf + (size_t)rc * TL::PRFB;
  const uint8_t* coins = src.coins + (size_t)rc * 64;
  // ---- (rho, sigma) = G(d); PRF(sigma, 0 .. 2K-1) -> scratch; rho -> LDS and both pk copies
  {
    uint64_t d[4], A[25], sg[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) d[w] = load_u64(coins + 8 * w);
    absorb_fixed<kRateSha3_512, 4>(A, d, 0x06);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      sg[w] = A[4 + w];
      store_u64(sh.msg[lane] + 8 * w, A[w]);
      if (active) {
        store_u64(pk + P::polyvec_bytes + 8 * w, A[w]);
        store_u64(sk + P::sk_cpa_bytes + P::polyvec_bytes + 8 * w, A[w]);
      }
    }
    for (int nonce = 0; nonce < 2 * K; ++nonce) {
      absorb_fixed<kRateShake256, 4>(A, sg, (uint64_t)nonce | (0x1Full << 8));
      if (active) {
#pragma unroll
        for (int w = 0; w < 16; ++w) store_u64(prf + 128 * nonce + 8 * w, A[w]);
      }
    }
  }
  __syncthreads();
  // ---- A rounds + t, s
  const int nt = min(TL::T, n - base);
  int t_done = 0;
  KeyIn<K> cur, nxt;
  const int ng = min(TL::NR, (nt * TL::KK + 63) / 64);  // rounds that hold the tile's entries
  for (int g = 0; g < ng; ++g) {
    const int t_end = min(nt, (64 * (g + 1)) / TL::KK);
    if (t_done < t_end) key_load<K>(cur, src.prf + (size_t)(base + t_done) * TL::PRFB);
    {
      const int E = 64 * g + lane;
      const int t = E / TL::KK;
      const int ij = E % TL::KK;
      const bool valid = t < nt;
      uint64_t rho[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) rho[w] = load_u64(sh.msg[valid ? t : 0] + 8 * w);
      const int got = sample_entry(sh.a + (E % TL::RING) * kAStride, rho, ij % K, ij / K, valid);
      if (valid && got < N) atomicOr(status, 1u);
    }
    __syncthreads();
    for (int t = t_done; t < t_end; ++t) {
      if (t + 1 < t_end) key_load<K>(nxt, src.prf + (size_t)(base + t + 1) * TL::PRFB);
      coop_keypair<K>(AtRing<K>{sh.a, t * TL::KK}, cur, src.pk + (size_t)(base + t) * P::pk_bytes,
                      src.sk + (size_t)(base + t) * P::sk_bytes, z);
      cur = nxt;
    }
    t_done = max(t_done, t_end);
    __syncthreads();
  }
  // ---- H(pk) over the packed key (written by the whole wave), z
  {
    uint64_t A[25];
    sponge_absorb_words_prefetch<kRateSha3_256>(A, reinterpret_cast<const uint64_t*>(pk), P::pk_bytes / 8, 0x06);
    if (active) {
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        store_u64(sk + P::hpk_off + 8 * w, A[w]);
        store_u64(sk + P::z_off + 8 * w, load_u64(coins + 32 + 8 * w));
      }
    }
  }
  __syncthreads();
}

// Key generation in the 32-request tiles of decapsulation and encapsulation
// (two waves per SIMD, A through the per-wave global scratch): both lanes of a
// pair compute G(d) and split the 2K PRF outputs (lane l the even nonces, lane
// l + 32 the odd ones); the rounds of A entries, each followed by the key
// computation of the requests it completes; H(pk) on lanes 0..31.  Same bytes
// as fused_keypair_tile.
template <int K>
__device__ __forceinline__ void keypair_tile32(const KeySrc& src, int base, int n, Lds32& sh, global_t<int16_t>* at,
                                               const LaneZF& z, uint32_t* __restrict__ status) {
  using P = Params<K>;
  using TL = Tile32<K>;
  constexpr size_t PRFB = Tile<K>::PRFB;
  const int lane = threadIdx.x;
  const int rq = lane & 31;  // the request this lane serves in the sponge phases (lanes l and l + 32 alike)
  const int nt = min(TL::T, n - base);
  const bool act = rq < nt;
  const bool hl = lane < 32;
  const int rc = act ? base + rq : base;  // clamped request for address math
  uint8_t* pk = src.pk + (size_t)rc * P::pk_bytes;
  uint8_t* sk = src.sk + (size_t)rc * P::sk_bytes;
  uint8_t* prf = src.prf + (size_t)rc * PRFB;
  const uint8_t* coins = src.coins + (size_t)rc * 64;
  // ---- (rho, sigma) = G(d) on both lanes of a pair; PRF(sigma, 2 i + (lane >= 32)),
  // i = 0 .. K - 1 -> scratch; rho -> LDS and both pk copies
  {
    uint64_t d[4], A[25], sg[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) d[w] = load_u64_global(coins + 8 * w);
    absorb_fixed<kRateSha3_512, 4>(A, d, 0x06);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      sg[w] = A[4 + w];
      if (hl) store_u64(sh.msg[rq] + 8 * w, A[w]);
      if (hl && act) {
        store_u64(pk + P::polyvec_bytes + 8 * w, A[w]);
        store_u64(sk + P::sk_cpa_bytes + P::polyvec_bytes + 8 * w, A[w]);
      }
    }
    for (int i = 0; i < K; ++i) {
      const int nonce = 2 * i + (hl ? 0 : 1);
      absorb_fixed<kRateShake256, 4>(A, sg, (uint64_t)nonce | (0x1Full << 8));
      if (act) {
#pragma unroll
        for (int w = 0; w < 16; ++w) store_u64(prf + 128 * nonce + 8 * w, A[w]);
      }
    }
  }
  __syncthreads();
  // ---- rounds of 64 A entries into the scratch (entry E: request E / K^2, A[i][j]
  // at position i K + j), then t and s of every request the round completes
  const int ng = (nt * TL::KK + 63) / 64;
  int t_done = 0;
  for (int g = 0; g < ng; ++g) {
    {
      const int E = 64 * g + lane;
      const int t = E / TL::KK;
      const int ij = E % TL::KK;
      const bool valid = t < nt;
      uint64_t rho[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) rho[w] = load_u64(sh.msg[valid ? t : 0] + 8 * w);
      const int got = sample_entry_staged(at + (size_t)E * kAtStride, sh.stage[lane], rho, ij % K, ij / K, valid);
      if (valid && got < N) atomicOr(status, 1u);
    }
    __syncthreads();  // the round's entries are in the scratch for every lane
    const int t_end = min(nt, (64 * (g + 1)) / TL::KK);
    if (t_done < t_end) {
      KeyIn<K> cur, nxt;
      key_load<K>(nxt, src.prf + (size_t)(base + t_done) * PRFB);
      for (int t = t_done; t < t_end; ++t) {
        cur = nxt;
        if (t + 1 < t_end) key_load<K>(nxt, src.prf + (size_t)(base + t + 1) * PRFB);
        AtRegs<K> a;
        a.load(at, t);
        coop_keypair<K>(a, cur, src.pk + (size_t)(base + t) * P::pk_bytes, src.sk + (size_t)(base + t) * P::sk_bytes,
                        z);
      }
      t_done = t_end;
    }
  }
  __syncthreads();
  // ---- H(pk) over the packed key (written by the whole wave), z; lanes 0..31 store
  {
    uint64_t A[25];
    sponge_absorb_words_prefetch<kRateSha3_256>(A, reinterpret_cast<const uint64_t*>(pk), P::pk_bytes / 8, 0x06);
    if (hl && act) {
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        store_u64(sk + P::hpk_off + 8 * w, A[w]);
        store_u64(sk + P::z_off + 8 * w, load_u64_global(coins + 32 + 8 * w));
      }
    }
  }
  __syncthreads();
}

template <int K>
__global__ __launch_bounds__(64, 2) void k_fused_keypair32(int n, int ntiles, KeySrc src, int16_t* at_scratch,
                                                           uint32_t* __restrict__ status) {
  __shared__ Lds32 sh;
  LaneZF z;
  load_lane_zf(z, &c_lane_twf);
  global_t<int16_t>* at = (global_t<int16_t>*)(at_scratch + (size_t)blockIdx.x * Tile32<K>::ENTRIES * kAtStride);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) keypair_tile32<K>(src, tile * 32, n, sh, at, z, status);
}

template <int K>
__global__ __launch_bounds__(64) void k_fused_keypair(int n, int ntiles, KeySrc src, uint32_t* __restrict__ status) {
  __shared__ Lds<Tile<K>::RING> sh;
  LaneZF z;
  load_lane_zf(z, &c_lane_twf);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) fused_keypair_tile<K>(src, tile * 64, n, sh, z, status);
}

template <int K, bool DEC>
__global__ __launch_bounds__(64) void k_fused(int n, int ntiles, DenseSrc src, uint32_t* __restrict__ status) {
  __shared__ Lds<Tile<K>::RING> sh;
  LaneZF z;
  load_lane_zf(z, &c_lane_twf);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) fused_tile<K, DEC>(src, tile * 64, n, sh, z, status);
}

// ---- low latency: one workgroup of four waves per request -------------------------
// A tile's latency is one lane's serial sponge chain (H(ct), G, the 2K + 1 PRF
// outputs, the A^T entries, KDF) plus the wave's cooperative work for all its
// requests in turn, about 1 ms for a full tile whatever the batch size.  For
// small batches (the single-instance drop-ins, a serving loop's batches) these
// kernels spread one request over four waves and many lanes.  Decapsulation:
//   phase 1  wave 0: indcpa_dec (wave-cooperative), then G and the 2K + 1 PRF
//                    outputs on 2K + 1 lanes at once;
//            wave 1: the K^2 A^T entries on K^2 lanes at once;
//            wave 2: the first blocks of H(ct) on a lane pair (keccakf1600_pair,
//                    about 2/3 of the single-lane latency per block);
//   phase 2  wave 0: re-encryption and compare; wave 2: the rest of H(ct), its
//                    sponge state kept in registers across the barrier;
//   phase 3  one lane: KDF.
// Encapsulation: see lowlat_enc_body.  Results are byte-identical to the tile
// kernel's.

template <int K>
struct alignas(16) LowLds {
  uint8_t prf[(2 * K + 1) * 128];                 // PRF(coins, nonce) at 128 nonce
  alignas(16) uint8_t msg[32];
  uint32_t kbar[8], hct[8], hpk[8];               // 32-byte values as words
  uint32_t fail;
  alignas(16) uint8_t ct[Params<K>::ct_bytes];    // encapsulation: the ciphertext before it is stored
  alignas(16) int16_t a[K * K * kAStride];        // A^T entry i K + j in slot i K + j
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// K^2 A^T entries of one request on lanes 0 .. K^2 - 1 of the calling wave
template <int K>
__device__ __forceinline__ void lowlat_sample(LowLds<K>& sh, const uint8_t* pk, int l, uint32_t* status) {
  using P = Params<K>;
  if (l < K * K) {
    // entry ij = i K + j on lane ij: A^T[i][j] = SHAKE128(rho || i || j)
    uint64_t rho[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) rho[w] = load_u64(pk + P::polyvec_bytes + 8 * w);
    const int got = sample_entry(sh.a + l * kAStride, rho, l / K, l % K, true);
    if (got < N) atomicOr(status, 1u);
  }
}

// (Kbar, coins) = G(msg || h) on lanes 0 .. 2K; lane j then writes PRF(coins, j)
// to sh.prf and lane 0 Kbar to sh.kbar.  Every lane recomputes G (same inputs)
// instead of waiting for a broadcast.
template <int K>
__device__ __forceinline__ void lowlat_g_prf(LowLds<K>& sh, const uint64_t (&h)[4], int l) {
  if (l < 2 * K + 1) {
    uint64_t m[8], A[25], rc4[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      m[w] = load_u64(sh.msg + 8 * w);
      m[4 + w] = h[w];
    }
    absorb_fixed<kRateSha3_512, 8>(A, m, 0x06);
#pragma unroll
    for (int w = 0; w < 4; ++w) rc4[w] = A[4 + w];
    if (l == 0) {
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        sh.kbar[2 * w] = (uint32_t)A[w];
        sh.kbar[2 * w + 1] = (uint32_t)(A[w] >> 32);
      }
    }
    absorb_fixed<kRateShake256, 4>(A, rc4, (uint64_t)l | (0x1Full << 8));
#pragma unroll
    for (int w = 0; w < 16; ++w) store_u64(sh.prf + 128 * l + 8 * w, A[w]);
  }
}

// KDF(key || H(c)) -> ss on the calling lane
__device__ __forceinline__ void lowlat_kdf(uint8_t* ss, const uint32_t (&kb)[8], const uint32_t (&hc)[8]) {
  uint64_t m[8], A[25];
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    m[w] = (uint64_t)kb[2 * w] | ((uint64_t)kb[2 * w + 1] << 32);
    m[4 + w] = (uint64_t)hc[2 * w] | ((uint64_t)hc[2 * w + 1] << 32);
  }
  absorb_fixed<kRateShake256, 8>(A, m, 0x1F);
#pragma unroll
  for (int w = 0; w < 4; ++w) store_u64(ss + 8 * w, A[w]);
}

// One decapsulation by the four waves of the workgroup (wave = threadIdx.x / 64)
template <int K>
__device__ __forceinline__ void lowlat_dec_body(uint8_t* ss, const uint8_t* c, const uint8_t* key, LowLds<K>& sh,
                                                const LaneZF& z, int wave, int l, uint32_t* status) {
  using P = Params<K>;
  constexpr int HW = P::ct_bytes / 8;         // H(ct) message words
  constexpr int HB = HW / kRateSha3_256 + 1;  // H(ct) blocks, padding included
  constexpr int HB1 = (2 * HB + 2) / 3;       // of them absorbed during phase 1
  const uint8_t* pk = key + P::sk_cpa_bytes;
  uint32_t S[25];  // wave 2's H(ct) state, held across the first barrier
  // ---- phase 1
  if (wave == 0) {
    {
      DecIn<K> in;
      dec_load<K>(in, c, key);
      coop_indcpa_dec_f<K>(in, z, sh.msg);
    }
    wave_sync();
    uint64_t h[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) h[w] = load_u64(key + P::hpk_off + 8 * w);
    lowlat_g_prf<K>(sh, h, l);
  } else if (wave == 1) {
    lowlat_sample<K>(sh, pk, l, status);
  } else if (wave == 2 && l < 2) {
    absorb_range_pair(S, c, HW, 0, HB1, 0x06, (uint32_t)l);
  }
  __syncthreads();
  // ---- phase 2
  if (wave == 0) {
    EncIn<K, true> in;
    enc_load<K, true, true>(in, sh.prf, pk, c);
    const uint32_t mm = coop_reencrypt_f<K, true>(AtRing<K>{sh.a, 0}, in, sh.msg, nullptr, z);
    const uint32_t f = __any(mm != 0) ? 1u : 0u;
    if (l == 0) sh.fail = f;
  } else if (wave == 2 && l < 2) {
    absorb_range_pair(S, c, HW, HB1, HB, 0x06, (uint32_t)l);
#pragma unroll
    for (int w = 0; w < 4; ++w) sh.hct[2 * w + l] = S[w];
  }
  __syncthreads();
  // ---- phase 3: KDF(Kbar or z, H(ct)), z chosen by mask
  if (wave == 0 && l == 0) {
    const uint32_t mask = 0u - (sh.fail & 1u);
    uint32_t kb[8], hc[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      const uint32_t zz = reinterpret_cast<const uint32_t*>(key + P::z_off)[w];
      kb[w] = sh.kbar[w] ^ (mask & (zz ^ sh.kbar[w]));
      hc[w] = sh.hct[w];
    }
    lowlat_kdf(ss, kb, hc);
  }
}

// One encapsulation by the four waves of the workgroup
//   phase 1  wave 0: m = H(coins); wave 1: A^T; wave 2: H(pk) on a lane pair
//   phase 2  wave 0: G and the PRF outputs, then the encryption into LDS
//   phase 3  wave 0: ct to memory; wave 2: H(ct) on a lane pair
//   phase 4  one lane: KDF(Kbar || H(ct))
template <int K>
__device__ __forceinline__ void lowlat_enc_body(uint8_t* ss, uint8_t* c, const uint8_t* pk, const uint8_t* coins,
                                                LowLds<K>& sh, const LaneZF& z, int wave, int l, uint32_t* status) {
  using P = Params<K>;
  // ---- phase 1
  if (wave == 0 && l == 0) {
    uint64_t A[25], c4[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) c4[w] = load_u64(coins + 8 * w);
    absorb_fixed<kRateSha3_256, 4>(A, c4, 0x06);  // "don't release system RNG output"
#pragma unroll
    for (int w = 0; w < 4; ++w) store_u64(sh.msg + 8 * w, A[w]);
  } else if (wave == 1) {
    lowlat_sample<K>(sh, pk, l, status);
  } else if (wave == 2 && l < 2) {
    uint32_t S[25];
    constexpr int W = P::pk_bytes / 8;
    absorb_range_pair(S, pk, W, 0, W / kRateSha3_256 + 1, 0x06, (uint32_t)l);
#pragma unroll
    for (int w = 0; w < 4; ++w) sh.hpk[2 * w + l] = S[w];
  }
  __syncthreads();
  // ---- phase 2
  if (wave == 0) {
    uint64_t h[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) h[w] = (uint64_t)sh.hpk[2 * w] | ((uint64_t)sh.hpk[2 * w + 1] << 32);
    lowlat_g_prf<K>(sh, h, l);
    wave_sync();
    EncIn<K, false> in;
    enc_load<K, false, true>(in, sh.prf, pk, nullptr);
    (void)coop_reencrypt_f<K, false>(AtRing<K>{sh.a, 0}, in, sh.msg, sh.ct, z);
  }
  __syncthreads();
  // ---- phase 3
  if (wave == 0) {
    for (int w = l; w < P::ct_bytes / 8; w += 64) store_u64(c + 8 * w, load_u64(sh.ct + 8 * w));
  } else if (wave == 2 && l < 2) {
    uint32_t S[25];
    constexpr int W = P::ct_bytes / 8;
    absorb_range_pair(S, sh.ct, W, 0, W / kRateSha3_256 + 1, 0x06, (uint32_t)l);
#pragma unroll
    for (int w = 0; w < 4; ++w) sh.hct[2 * w + l] = S[w];
  }
  __syncthreads();
  // ---- phase 4
  if (wave == 0 && l == 0) {
    uint32_t kb[8], hc[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      kb[w] = sh.kbar[w];
      hc[w] = sh.hct[w];
    }
    lowlat_kdf(ss, kb, hc);
  }
}

// Dense batches: request x at x * size in each buffer
template <int K, bool DEC>
__global__ __launch_bounds__(256) void k_lowlat(int n, uint8_t* __restrict__ ss, uint8_t* __restrict__ ct,
                                                const uint8_t* __restrict__ key, const uint8_t* __restrict__ coins,
                                                uint32_t* __restrict__ status) {
  using P = Params<K>;
  __shared__ LowLds<K> sh;
  const int x = blockIdx.x;
  if (x >= n) return;  // whole workgroup
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int l = threadIdx.x & 63;
  LaneZF z;
  load_lane_zf(z, &c_lane_twf);
  if constexpr (DEC)
    lowlat_dec_body<K>(ss + (size_t)x * 32, ct + (size_t)x * P::ct_bytes, key + (size_t)x * P::sk_bytes, sh, z, wave,
                       l, status);
  else
    lowlat_enc_body<K>(ss + (size_t)x * 32, ct + (size_t)x * P::ct_bytes, key + (size_t)x * P::pk_bytes,
                       coins + (size_t)x * 32, sh, z, wave, l, status);
}

// Mixed batches: one descriptor per workgroup, checked as k_mixed_classify
// checks it (a rejected one sets status bit 1 and is skipped)
union LowLdsMixed {
  LowLds<2> k2;
  LowLds<3> k3;
  LowLds<4> k4;
};

template <bool DEC>
__global__ __launch_bounds__(256) void k_lowlat_mixed(int n, unsigned alg_mask, const pqc_kem_req* __restrict__ reqs,
                                                      uint8_t* ct, uint64_t ct_bytes, const uint8_t* key,
                                                      uint64_t key_bytes, uint8_t* ss, uint64_t ss_bytes,
                                                      const uint8_t* coins, uint64_t coins_bytes,
                                                      uint32_t* __restrict__ status) {
  __shared__ LowLdsMixed sh;
  const int x = blockIdx.x;
  if (x >= n) return;
  const pqc_kem_req d = reqs[x];
  const int k = pqc_req_k(d, DEC, alg_mask, ct_bytes, key_bytes, ss_bytes, coins_bytes);
  if (!k) {  // uniform: every lane read the same descriptor
    if (threadIdx.x == 0) atomicOr(status, 2u);
    return;
  }
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int l = threadIdx.x & 63;
  LaneZF z;
  load_lane_zf(z, &c_lane_twf);
  uint8_t* s_ = ss + d.ss_off;
  uint8_t* c_ = ct + d.ct_off;
  const uint8_t* k_ = key + d.key_off;
  const uint8_t* r_ = DEC ? nullptr : coins + d.coins_off;
  switch (k) {
    case 2:
      if constexpr (DEC) lowlat_dec_body<2>(s_, c_, k_, sh.k2, z, wave, l, status);
      else lowlat_enc_body<2>(s_, c_, k_, r_, sh.k2, z, wave, l, status);
      break;
    case 3:
      if constexpr (DEC) lowlat_dec_body<3>(s_, c_, k_, sh.k3, z, wave, l, status);
      else lowlat_enc_body<3>(s_, c_, k_, r_, sh.k3, z, wave, l, status);
      break;
    default:
      if constexpr (DEC) lowlat_dec_body<4>(s_, c_, k_, sh.k4, z, wave, l, status);
      else lowlat_enc_body<4>(s_, c_, k_, r_, sh.k4, z, wave, l, status);
      break;
  }
}

// ---- mixed parameter sets ------------------------------------------------------
// Bucket n descriptors by K (Kyber-512/768/1024 -> K 2/3/4) into order[(K-2) * cap
// + slot], counts[K-2] = bucket sizes.  One wave-aggregated global atomic per
// bucket per wave; the order inside a bucket is whatever the atomics give (the
// results do not depend on it).  Descriptors with another algorithm id, an
// unaligned offset, a buffer outside its arena or a parameter set the context
// was not created for are skipped and set status bit 1.
__global__ void k_mixed_zero(uint32_t* counts) {
  if (threadIdx.x < 3) counts[threadIdx.x] = 0;
}

// One thread per descriptor, 1024 per workgroup.  A request's place inside its
// bucket does not matter, so each workgroup reserves its places with one atomic
// per K (the same three counters taken once per wave made the launch ~26 us at
// 65536 descriptors), and each wave's share follows from the waves before it.
constexpr int kClassifyBlock = 1024;
__global__ __launch_bounds__(kClassifyBlock) void k_mixed_classify(
    int n, int cap, int dec, unsigned alg_mask, const pqc_kem_req* __restrict__ reqs, uint64_t ct_bytes,
    uint64_t key_bytes, uint64_t ss_bytes, uint64_t coins_bytes, uint32_t* __restrict__ order,
    uint32_t* __restrict__ counts, uint32_t* __restrict__ status) {
  constexpr int kWaves = kClassifyBlock / 64;
  __shared__ uint32_t wcount[3][kWaves], wbase[3][kWaves];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int k = 0;
  if (i < n) {
    const pqc_kem_req d = reqs[i];
    k = pqc_req_k(d, dec != 0, alg_mask, ct_bytes, key_bytes, ss_bytes, coins_bytes);
    if (!k) atomicOr(status, 2u);
  }
  uint64_t mine = 0;  // the ballot of this lane's K
#pragma unroll
  for (int kk = 2; kk <= 4; ++kk) {
    const uint64_t mask = __ballot(k == kk);
    if (k == kk) mine = mask;
    if (lane == 0) wcount[kk - 2][wave] = (uint32_t)__popcll(mask);
  }
  __syncthreads();
  if (threadIdx.x < 3) {  // thread j: bucket K = j + 2
    const int j = threadIdx.x;
    uint32_t run = 0;
    for (int w = 0; w < kWaves; ++w) {
      wbase[j][w] = run;
      run += wcount[j][w];
    }
    const uint32_t g = run ? atomicAdd(&counts[j], run) : 0u;
    for (int w = 0; w < kWaves; ++w) wbase[j][w] += g;
  }
  __syncthreads();
  if (k) {
    const uint32_t below = (uint32_t)__popcll(mine & ((1ull << lane) - 1));
    order[(size_t)(k - 2) * cap + wbase[k - 2][wave] + below] = (uint32_t)i;
  }
}

// Tile i of a mixed batch -> (K, tile index inside the K bucket).  order 0:
// K = 4, 3, 2 (longest tiles first keeps the tail short); 1: K = 2, 3, 4;
// 2: round-robin over the buckets that still have tiles, in the order 4, 3, 2.
struct TilePick {
  int k, idx;
};

__device__ __forceinline__ TilePick pick_tile(int i, int order, int t4, int t3, int t2) {
  if (order == 0) {
    if (i < t4) return {4, i};
    if (i < t4 + t3) return {3, i - t4};
    return {2, i - t4 - t3};
  }
  if (order == 1) {
    if (i < t2) return {2, i};
    if (i < t2 + t3) return {3, i - t2};
    return {4, i - t2 - t3};
  }
  // round-robin: phases in which 3, 2, then 1 buckets remain
  int ks[3] = {4, 3, 2}, ts[3] = {t4, t3, t2};
  int done = 0;  // tiles per bucket consumed by earlier phases
  int live = (t4 > 0) + (t3 > 0) + (t2 > 0);
  while (live > 0) {
    int m = 1 << 30;  // rounds in this phase: the smallest remaining bucket
    for (int b = 0; b < 3; ++b)
      if (ts[b] > done && ts[b] - done < m) m = ts[b] - done;
    if (m == (1 << 30)) break;
    if (i < live * m) {
      const int round = i / live, pos = i % live;
      int seen = 0;
      for (int b = 0; b < 3; ++b)
        if (ts[b] > done) {
          if (seen == pos) return {ks[b], done + round};
          ++seen;
        }
    }
    i -= live * m;
    done += m;
    live = 0;
    for (int b = 0; b < 3; ++b) live += ts[b] > done ? 1 : 0;
  }
  return {0, 0};  // not reached for i below t4 + t3 + t2
}

// Persistent walk over the tiles of the three buckets.  LDS is sized for the
// largest ring (K = 3).
union LdsMixed {
  Lds<Tile<2>::RING> k2;
  Lds<Tile<3>::RING> k3;
  Lds<Tile<4>::RING> k4;
};

// Requests per tile of each bucket.  Order 3 (split) halves the K = 4 and K = 2
// tiles and walks K 4, 3, 2: with one tile per wave the launch lasts as long as
// its heaviest tile, and a K = 4 half tile plus a K = 2 half tile on one wave is
// about one K = 3 tile (the per-request sponges of a half tile run on half the
// lanes, its Aᵀ rounds and cooperative work halve).  The autotuner measures it
// against the other orders; a tile of fewer requests is a partial tile.
__host__ __device__ __forceinline__ int mixed_tile_req(int order, int k) { return (order == 3 && k != 3) ? 32 : 64; }

template <bool DEC>
__global__ __launch_bounds__(64) void k_fused_mixed(MixedSrc src, const uint32_t* __restrict__ counts, int order,
                                                    uint32_t* __restrict__ status) {
  __shared__ LdsMixed sh;
  LaneZF z;
  load_lane_zf(z, &c_lane_twf);
  const int n2 = (int)counts[0], n3 = (int)counts[1], n4 = (int)counts[2];
  const int s4 = mixed_tile_req(order, 4), s3 = mixed_tile_req(order, 3), s2 = mixed_tile_req(order, 2);
  const int t4 = (n4 + s4 - 1) / s4, t3 = (n3 + s3 - 1) / s3, t2 = (n2 + s2 - 1) / s2;
  const int walk = order == 3 ? 0 : order;
  for (int tile = blockIdx.x; tile < t4 + t3 + t2; tile += gridDim.x) {
    const TilePick p = pick_tile(tile, walk, t4, t3, t2);
    if (p.k == 4)
      fused_tile<4, DEC>(src, p.idx * s4, min(n4, (p.idx + 1) * s4), sh.k4, z, status);
    else if (p.k == 3)
      fused_tile<3, DEC>(src, p.idx * s3, min(n3, (p.idx + 1) * s3), sh.k3, z, status);
    else if (p.k == 2)
      fused_tile<2, DEC>(src, p.idx * s2, min(n2, (p.idx + 1) * s2), sh.k2, z, status);
  }
}

// ---- component kernels -------------------------------------------------------

// poly_ntt / poly_invntt on polynomials in coefficient order, one per wave.
__global__ __launch_bounds__(64) void k_poly_transform(int op, int n, int16_t* polys) {
  const int l = threadIdx.x;
  LaneZ z;
  load_lane_z(z, &c_lane_tw);
  for (int p = blockIdx.x; p < n; p += gridDim.x) {
    int16_t* poly = polys + (size_t)p * N;
    const int gin = op == 0 ? nat_group(l) : ntt_group(l);
    const int gout = op == 0 ? ntt_group(l) : nat_group(l);
    P4 x;
#pragma unroll
    for (int r = 0; r < 4; ++r) x.v[r] = poly[4 * gin + r];
    if (op == 0)
      ntt(x, z);
    else
      invntt(x, z);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) poly[4 * gout + r] = (int16_t)x.v[r];
    __syncthreads();
  }
}

// n fixed-length messages, one per lane: SHA3-256 / SHA3-512 / SHAKE128 / SHAKE256.
template <int RATE>
__device__ void keccak_one(uint8_t* out, size_t outlen, const uint8_t* in, int words, uint8_t domain) {
  uint64_t A[25];
  sponge_absorb_words<RATE>(A, reinterpret_cast<const uint64_t*>(in), words, domain);
  size_t done = 0;
  while (true) {
#pragma unroll
    for (int w = 0; w < RATE; ++w) {
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const size_t o = done + 8 * w + b;
        if (o < outlen) out[o] = (uint8_t)(A[w] >> (8 * b));
      }
    }
    done += 8 * RATE;
    if (done >= outlen) break;
    keccakf1600(A);
  }
}

__global__ void k_keccak_batch(int op, int n, uint8_t* out, size_t outlen, const uint8_t* in, size_t inlen) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint8_t* m = in + (size_t)x * inlen;
  uint8_t* o = out + (size_t)x * outlen;
  const int words = (int)(inlen / 8);
  switch (op) {
    case 0: keccak_one<kRateSha3_256>(o, outlen, m, words, 0x06); break;
    case 1: keccak_one<kRateSha3_512>(o, outlen, m, words, 0x06); break;
    case 2: keccak_one<kRateShake128>(o, outlen, m, words, 0x1F); break;
    default: keccak_one<kRateShake256>(o, outlen, m, words, 0x1F); break;
  }
}

template <int K, bool DEC>
static int launch_fused(pqc_ctx* c, size_t n, uint8_t* ss, uint8_t* ct, const uint8_t* key, const uint8_t* coins,
                        hipStream_t s) {
  if (n == 0) return PQC_OK;
  const int ntiles = (int)((n + Tile<K>::T - 1) / Tile<K>::T);
  int grid = c->fused_grid_override > 0 ? c->fused_grid_override : c->cu_count * c->tune.fused_waves_per_cu;
  if (grid > ntiles) grid = ntiles;
  if (grid < 1) grid = 1;
  const DenseSrc src{ss, ct, key, coins, c->kprf};
  hipLaunchKernelGGL((k_fused<K, DEC>), dim3(grid), dim3(64), 0, s, (int)n, ntiles, src, c->status);
  return hip_err(hipGetLastError());
}

// Kyber batches (dense or mixed) of at most this many requests take the
// one-workgroup-per-request kernels.  Measured on MI355X (Kyber-768): a tile
// launch takes ~1.0 ms for any batch up to 65536; the per-request kernels run
// ~10.6 M decaps/s and ~5.4 M encaps/s with ~0.12 / 0.2 ms latency, so they win
// below ~10 K decapsulations and ~5 K encapsulations.  tuning.lowlat_max
// overrides both (-1 = never).
constexpr int kLowLatDefaultDec = 8192, kLowLatDefaultEnc = 4096;

inline bool use_lowlat(const pqc_ctx* c, size_t n, bool dec) {
  const int lim = c->tune.lowlat_max != 0 ? c->tune.lowlat_max : dec ? kLowLatDefaultDec : kLowLatDefaultEnc;
  return n > 0 && (long long)n <= (long long)lim;
}

template <int K, bool DEC>
static int launch_lowlat(pqc_ctx* c, size_t n, uint8_t* ss, uint8_t* ct, const uint8_t* key, const uint8_t* coins,
                         hipStream_t s) {
  hipLaunchKernelGGL((k_lowlat<K, DEC>), dim3((unsigned)n), dim3(256), 0, s, (int)n, ss, ct, key, coins, c->status);
  return hip_err(hipGetLastError());
}

template <int K>
static int launch_dec32(pqc_ctx* c, size_t n, uint8_t* ss, uint8_t* ct, const uint8_t* sk, hipStream_t s) {
  const int ntiles = (int)((n + Tile32<K>::T - 1) / Tile32<K>::T);
  // two waves per SIMD at the default tuning (fused_waves_per_cu 4 -> 8 waves per CU)
  int grid = c->fused_grid_override > 0 ? c->fused_grid_override : c->cu_count * 2 * c->tune.fused_waves_per_cu;
  grid = std::min(grid, ntiles);
  grid = std::min<long long>(grid, (long long)c->kat_waves);  // one scratch region per wave
  if (grid < 1) return PQC_ERR_INVALID_ARG;
  const DenseSrc src{ss, ct, sk, nullptr, c->kprf};
  hipLaunchKernelGGL((k_fused_dec32<K>), dim3(grid), dim3(64), 0, s, (int)n, ntiles, src, c->kat, c->status);
  return hip_err(hipGetLastError());
}

template <int K>
static int launch_enc32(pqc_ctx* c, size_t n, uint8_t* ss, uint8_t* ct, const uint8_t* pk, const uint8_t* coins,
                        hipStream_t s) {
  const int ntiles = (int)((n + Tile32<K>::T - 1) / Tile32<K>::T);
  int grid = c->fused_grid_override > 0 ? c->fused_grid_override : c->cu_count * 2 * c->tune.fused_waves_per_cu;
  grid = std::min(grid, ntiles);
  grid = std::min<long long>(grid, (long long)c->kat_waves);  // one scratch region per wave
  if (grid < 1) return PQC_ERR_INVALID_ARG;
  const DenseSrc src{ss, ct, pk, coins, c->kprf};
  hipLaunchKernelGGL((k_fused_enc32<K>), dim3(grid), dim3(64), 0, s, (int)n, ntiles, src, c->kat, c->status);
  return hip_err(hipGetLastError());
}

template <int K>
static int launch_fused_keypair(pqc_ctx* c, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* coins, hipStream_t s) {
  if (n == 0) return PQC_OK;
  if (c->tune.fused_key_tile != 64 && c->kat) {
    // 32-request tiles, two waves per SIMD, one scratch region per wave
    const int ntiles = (int)((n + Tile32<K>::T - 1) / Tile32<K>::T);
    int grid = c->fused_grid_override > 0 ? c->fused_grid_override : c->cu_count * 2 * c->tune.fused_waves_per_cu;
    grid = std::min(grid, ntiles);
    grid = std::min<long long>(grid, (long long)c->kat_waves);
    if (grid < 1) return PQC_ERR_INVALID_ARG;
    const KeySrc src{pk, sk, coins, c->kprf};
    hipLaunchKernelGGL((k_fused_keypair32<K>), dim3(grid), dim3(64), 0, s, (int)n, ntiles, src, c->kat, c->status);
    return hip_err(hipGetLastError());
  }
  const int ntiles = (int)((n + Tile<K>::T - 1) / Tile<K>::T);
  int grid = c->fused_grid_override > 0 ? c->fused_grid_override : c->cu_count * c->tune.fused_waves_per_cu;
  if (grid > ntiles) grid = ntiles;
  if (grid < 1) grid = 1;
  const KeySrc src{pk, sk, coins, c->kprf};
  hipLaunchKernelGGL((k_fused_keypair<K>), dim3(grid), dim3(64), 0, s, (int)n, ntiles, src, c->status);
  return hip_err(hipGetLastError());
}

}  // namespace fused
}  // namespace kyber

int kyber_fused_keypair(pqc_ctx* c, int k, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* coins, hipStream_t s) {
  switch (k) {
    case 2: return kyber::fused::launch_fused_keypair<2>(c, n, pk, sk, coins, s);
    case 3: return kyber::fused::launch_fused_keypair<3>(c, n, pk, sk, coins, s);
    case 4: return kyber::fused::launch_fused_keypair<4>(c, n, pk, sk, coins, s);
    default: return PQC_ERR_UNSUPPORTED;
  }
}

size_t tile32_scratch_bytes(int k) {
  using namespace kyber::fused;
  const int entries = k == 2 ? Tile32<2>::ENTRIES : k == 3 ? Tile32<3>::ENTRIES : Tile32<4>::ENTRIES;
  return (size_t)entries * kAtStride * sizeof(int16_t);
}

int kyber_fused_dec(pqc_ctx* c, int k, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk, hipStream_t s) {
  uint8_t* ctm = const_cast<uint8_t*>(ct);  // read-only in the decaps instantiation
  if (n > 0 && !kyber::fused::use_lowlat(c, n, true) && c->tune.fused_dec_tile != 64 && c->kat) {
    switch (k) {
      case 2: return kyber::fused::launch_dec32<2>(c, n, ss, ctm, sk, s);
      case 3: return kyber::fused::launch_dec32<3>(c, n, ss, ctm, sk, s);
      case 4: return kyber::fused::launch_dec32<4>(c, n, ss, ctm, sk, s);
      default: return PQC_ERR_UNSUPPORTED;
    }
  }
  if (kyber::fused::use_lowlat(c, n, true)) {
    switch (k) {
      case 2: return kyber::fused::launch_lowlat<2, true>(c, n, ss, ctm, sk, nullptr, s);
      case 3: return kyber::fused::launch_lowlat<3, true>(c, n, ss, ctm, sk, nullptr, s);
      case 4: return kyber::fused::launch_lowlat<4, true>(c, n, ss, ctm, sk, nullptr, s);
      default: return PQC_ERR_UNSUPPORTED;
    }
  }
  switch (k) {
    case 2: return kyber::fused::launch_fused<2, true>(c, n, ss, ctm, sk, nullptr, s);
    case 3: return kyber::fused::launch_fused<3, true>(c, n, ss, ctm, sk, nullptr, s);
    case 4: return kyber::fused::launch_fused<4, true>(c, n, ss, ctm, sk, nullptr, s);
    default: return PQC_ERR_UNSUPPORTED;
  }
}

int kyber_fused_enc(pqc_ctx* c, int k, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk, const uint8_t* coins,
                    hipStream_t s) {
  if (n > 0 && !kyber::fused::use_lowlat(c, n, false) && c->tune.fused_enc_tile != 64 && c->kat) {
    switch (k) {
      case 2: return kyber::fused::launch_enc32<2>(c, n, ss, ct, pk, coins, s);
      case 3: return kyber::fused::launch_enc32<3>(c, n, ss, ct, pk, coins, s);
      case 4: return kyber::fused::launch_enc32<4>(c, n, ss, ct, pk, coins, s);
      default: return PQC_ERR_UNSUPPORTED;
    }
  }
  if (kyber::fused::use_lowlat(c, n, false)) {
    switch (k) {
      case 2: return kyber::fused::launch_lowlat<2, false>(c, n, ss, ct, pk, coins, s);
      case 3: return kyber::fused::launch_lowlat<3, false>(c, n, ss, ct, pk, coins, s);
      case 4: return kyber::fused::launch_lowlat<4, false>(c, n, ss, ct, pk, coins, s);
      default: return PQC_ERR_UNSUPPORTED;
    }
  }
  switch (k) {
    case 2: return kyber::fused::launch_fused<2, false>(c, n, ss, ct, pk, coins, s);
    case 3: return kyber::fused::launch_fused<3, false>(c, n, ss, ct, pk, coins, s);
    case 4: return kyber::fused::launch_fused<4, false>(c, n, ss, ct, pk, coins, s);
    default: return PQC_ERR_UNSUPPORTED;
  }
}

// Mixed Kyber-512/768/1024 batch: classify into buckets, then one persistent
// launch.  Every launch has a fixed shape (the bucket sizes are read on the
// device), so the sequence can be captured into a graph and replayed.
int kyber_fused_mixed(pqc_ctx* c, int dec, size_t n, const pqc_kem_req* reqs, uint8_t* ct, size_t ct_bytes,
                      const uint8_t* key, size_t key_bytes, uint8_t* ss, size_t ss_bytes, const uint8_t* coins,
                      size_t coins_bytes, hipStream_t s) {
  using namespace kyber::fused;
  if (n == 0) return PQC_OK;
  if (!c->kmix_order || !c->kprf) return PQC_ERR_UNSUPPORTED;
  if (use_lowlat(c, n, dec != 0)) {  // one workgroup per descriptor, no bucketing
    if (dec)
      hipLaunchKernelGGL(k_lowlat_mixed<true>, dim3((unsigned)n), dim3(256), 0, s, (int)n, c->alg_mask, reqs, ct,
                         (uint64_t)ct_bytes, key, (uint64_t)key_bytes, ss, (uint64_t)ss_bytes, coins,
                         (uint64_t)coins_bytes, c->status);
    else
      hipLaunchKernelGGL(k_lowlat_mixed<false>, dim3((unsigned)n), dim3(256), 0, s, (int)n, c->alg_mask, reqs, ct,
                         (uint64_t)ct_bytes, key, (uint64_t)key_bytes, ss, (uint64_t)ss_bytes, coins,
                         (uint64_t)coins_bytes, c->status);
    return hip_err(hipGetLastError());
  }
  const int cap = (int)c->max_batch;
  hipLaunchKernelGGL(k_mixed_zero, dim3(1), dim3(64), 0, s, c->kmix_counts);
  hipLaunchKernelGGL(k_mixed_classify, dim3((unsigned)((n + kClassifyBlock - 1) / kClassifyBlock)),
                     dim3(kClassifyBlock), 0, s, (int)n, cap, dec,
                     c->alg_mask, reqs,
                     (uint64_t)ct_bytes, (uint64_t)key_bytes, (uint64_t)ss_bytes, (uint64_t)coins_bytes,
                     c->kmix_order, c->kmix_counts, c->status);
  const MixedSrc src{reqs, c->kmix_order, cap, ss, ct, key, coins, c->kprf, c->kprf_per_elem};
  const int max_tiles = (int)((n + 31) / 32) + 2;  // half tiles at most; buckets round up separately
  if (c->kat && (dec ? c->tune.fused_dec_tile : c->tune.fused_enc_tile) != 64) {
    // 32-request tiles, two waves per SIMD, one A^T scratch region per wave
    int grid = c->fused_grid_override > 0 ? c->fused_grid_override : c->cu_count * 2 * c->tune.fused_waves_per_cu;
    grid = std::min(grid, max_tiles);
    grid = std::min<long long>(grid, (long long)c->kat_waves);
    if (grid < 1) return PQC_ERR_INVALID_ARG;
    if (dec)
      hipLaunchKernelGGL(k_fused_mixed_dec32, dim3(grid), dim3(64), 0, s, src, (const uint32_t*)c->kmix_counts,
                         c->kat, c->kat_region, c->status);
    else
      hipLaunchKernelGGL(k_fused_mixed_enc32, dim3(grid), dim3(64), 0, s, src, (const uint32_t*)c->kmix_counts,
                         c->kat, c->kat_region, c->status);
    return hip_err(hipGetLastError());
  }
  int grid = c->fused_grid_override > 0 ? c->fused_grid_override : c->cu_count * c->tune.fused_waves_per_cu;
  if (grid > max_tiles) grid = max_tiles;
  if (grid < 1) grid = 1;
  if (dec)
    hipLaunchKernelGGL(k_fused_mixed<true>, dim3(grid), dim3(64), 0, s, src, (const uint32_t*)c->kmix_counts,
                       c->tune.mixed_order, c->status);
  else
    hipLaunchKernelGGL(k_fused_mixed<false>, dim3(grid), dim3(64), 0, s, src, (const uint32_t*)c->kmix_counts,
                       c->tune.mixed_order, c->status);
  return hip_err(hipGetLastError());
}

// ---- allocation-free decapsulation ------------------------------------------------
// pqc_kyber_dec_batch_noalloc runs the same kernels as crypto_kem_dec_batch on
// the default tuning, with the tile kernel's PRF scratch and the status word in
// `scratch`, a file-scope device array (one copy per GPU, held until the process
// ends) instead of a context's allocation.  The host never names `scratch`
// directly (in host code it is only a stand-in symbol): the device address is
// looked up once per device with hipGetSymbolAddress and that address is what
// the kernels receive.  Nothing in the scratch must start at zero: every PRF
// row is written by the same tile before it is read, and the status word only
// collects bits for callers that pass no status word of their own.
constexpr size_t kNoAllocMax = PQC_NOALLOC_MAX_BATCH;
constexpr size_t kNoAllocPrf = kNoAllocMax * kyber::fused::Tile<4>::PRFB;  // K = 4 rows are the largest
constexpr size_t kNoAllocBytes = 256 + kNoAllocPrf;                        // status word, then the PRF rows
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
struct NoAllocDev {
  uint8_t* ptr = nullptr;  // device address of `scratch` on this device
  int cus = 0;
  bool ready = false;
};
std::mutex g_noalloc_mu;
NoAllocDev g_noalloc[kMaxDevices];

// Set-up on the first call on a device, with that device current.
int noalloc_device(NoAllocDev* out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return PQC_ERR_NO_DEVICE;
  if (dev < 0 || dev >= kMaxDevices) return PQC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  NoAllocDev& d = g_noalloc[dev];
  if (!d.ready) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PQC_ERR_DEVICE;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      return PQC_ERR_DEVICE;
    d.ptr = static_cast<uint8_t*>(p);
    d.cus = cus;
    d.ready = true;
  }
  *out = d;
  return PQC_OK;
}

template <int K>
int noalloc_dec(const NoAllocDev& d, size_t n, uint8_t* ss, uint8_t* ct, const uint8_t* sk, uint32_t* status,
                hipStream_t s) {
  using namespace kyber::fused;
  if (n <= (size_t)kLowLatDefaultDec) {
    hipLaunchKernelGGL((k_lowlat<K, true>), dim3((unsigned)n), dim3(256), 0, s, (int)n, ss, ct, sk,
                       (const uint8_t*)nullptr, status);
  } else {
    const int ntiles = (int)((n + Tile<K>::T - 1) / Tile<K>::T);
    const int grid = std::max(1, std::min(ntiles, d.cus * 4));  // the default fused_waves_per_cu
    const DenseSrc src{ss, ct, sk, nullptr, d.ptr + 256};
    hipLaunchKernelGGL((k_fused<K, true>), dim3(grid), dim3(64), 0, s, (int)n, ntiles, src, status);
  }
  return hip_err(hipGetLastError());
}
}  // namespace

int kyber_poly_transform(int op, size_t n, int16_t* polys, hipStream_t s) {
  if (op != 0 && op != 1) return PQC_ERR_INVALID_ARG;
  if (n == 0) return PQC_OK;
  const int grid = (int)(n < 4096 ? n : 4096);
  hipLaunchKernelGGL(kyber::fused::k_poly_transform, dim3(grid), dim3(64), 0, s, op, (int)n, polys);
  return hip_err(hipGetLastError());
}

int keccak_batch(int op, size_t n, uint8_t* out, size_t outlen, const uint8_t* in, size_t inlen, hipStream_t s) {
  if (op < 0 || op > 3 || (inlen % 8) != 0) return PQC_ERR_INVALID_ARG;
  if ((op == 0 && outlen != 32) || (op == 1 && outlen != 64) || outlen == 0) return PQC_ERR_INVALID_ARG;
  if (n == 0) return PQC_OK;
  const int grid = (int)((n + 255) / 256);
  hipLaunchKernelGGL(kyber::fused::k_keccak_batch, dim3(grid), dim3(256), 0, s, op, (int)n, out, outlen, in, inlen);
  return hip_err(hipGetLastError());
}

}  // namespace pqc

extern "C" int pqc_kyber_dec_batch_noalloc(int alg, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk,
                                           uint32_t* status, void* stream) {
  if (alg != PQC_ALG_KYBER512 && alg != PQC_ALG_KYBER768 && alg != PQC_ALG_KYBER1024) return PQC_ERR_UNSUPPORTED;
  if (n > pqc::kNoAllocMax) return PQC_ERR_BATCH_TOO_LARGE;
  if (n == 0) return PQC_OK;
  if (!ss || !ct || !sk || (((uintptr_t)ss | (uintptr_t)ct | (uintptr_t)sk) & 7) || ((uintptr_t)status & 3))
    return PQC_ERR_INVALID_ARG;
  pqc::NoAllocDev d;
  const int rc = pqc::noalloc_device(&d);
  if (rc) return rc;
  uint32_t* st = status ? status : reinterpret_cast<uint32_t*>(d.ptr);
  uint8_t* ctm = const_cast<uint8_t*>(ct);  // read-only in the decapsulation kernels
  hipStream_t s = (hipStream_t)stream;
  switch (alg) {
    case PQC_ALG_KYBER512: return pqc::noalloc_dec<2>(d, n, ss, ctm, sk, st, s);
    case PQC_ALG_KYBER768: return pqc::noalloc_dec<3>(d, n, ss, ctm, sk, st, s);
    default: return pqc::noalloc_dec<4>(d, n, ss, ctm, sk, st, s);
  }
}
