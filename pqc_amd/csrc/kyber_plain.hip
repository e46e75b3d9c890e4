// Plain per-op batched Kyber kernels: the reference's execution structure
// (Kyber_GPU_Batched/indcpa.cu:474-733) on MI355X.  One kernel per reference
// operation, one thread per instance, every intermediate polynomial in a
// coefficient-major SoA array in HBM (poly.h:12-22: coefficient c of element x
// at c * ld + x, coalesced across threads).  The CCA glue the reference lacks
// (kem.cu, Q4) is added as further per-op kernels: G, H, verify, cmov, KDF.
//
// This path is the comparison baseline for the fused persistent kernels
// (kyber_fused.hip); both produce identical bytes.
#include "internal.h"
#include "kyber_arith.cuh"
#include "sponge.cuh"

namespace pqc {
namespace kyber {
namespace plain {


__constant__ int16_t c_zetas[128];
__constant__ int16_t c_zetas_inv[128];

struct Pv {
  int16_t* p;
  int ld;
  __device__ __forceinline__ int16_t& at(int c, int x) const { return p[(size_t)c * ld + x]; }
};

// d-bit field idx of a little-endian bit stream; reads only bytes it needs.
template <int D>
__device__ __forceinline__ uint32_t get_field(const uint8_t* base, int idx) {
  const uint32_t bit = (uint32_t)idx * D;
  const uint8_t* b = base + (bit >> 3);
  const int sh = (int)(bit & 7);
  uint32_t v = b[0];
  if (sh + D > 8) v |= (uint32_t)b[1] << 8;
  if (sh + D > 16) v |= (uint32_t)b[2] << 16;
  return (v >> sh) & ((1u << D) - 1);
}

// Pack 256 D-bit fields produced by `f(c)` LSB-first into out.
template <int D, typename F>
__device__ __forceinline__ void put_fields(uint8_t* out, F f) {
  uint64_t acc = 0;
  int nb = 0, o = 0;
  for (int c = 0; c < N; ++c) {
    acc |= (uint64_t)f(c) << nb;
    nb += D;
    while (nb >= 8) {
      out[o++] = (uint8_t)acc;
      acc >>= 8;
      nb -= 8;
    }
  }
}

// ---- symmetric kernels ---------------------------------------------------

// H: SHA3-256 of a message of `words` 64-bit words at in + x * stride.
__global__ void k_hash_h(int n, uint8_t* out, size_t out_stride, const uint8_t* in, size_t stride,
                         int words) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t A[25];
  sponge_absorb_words<kRateSha3_256>(A, reinterpret_cast<const uint64_t*>(in + (size_t)x * stride),
                                     words, 0x06);
  uint8_t* o = out + (size_t)x * out_stride;
#pragma unroll
  for (int w = 0; w < 4; ++w) store_u64(o + 8 * w, A[w]);
}

// G: SHA3-512 of (a || b), 32 + 32 bytes (fips202.cu:672-689).  With b == null,
// SHA3-512 of a alone (keypair: G(d)).
__global__ void k_hash_g(int n, uint8_t* out, const uint8_t* a, size_t a_stride, const uint8_t* b,
                         size_t b_stride) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t A[25];
  const uint8_t* pa = a + (size_t)x * a_stride;
  if (b) {
    const uint8_t* pb = b + (size_t)x * b_stride;
    uint64_t m[8];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      m[w] = load_u64(pa + 8 * w);
      m[4 + w] = load_u64(pb + 8 * w);
    }
    absorb_fixed<kRateSha3_512, 8>(A, m, 0x06);
  } else {
    uint64_t m[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) m[w] = load_u64(pa + 8 * w);
    absorb_fixed<kRateSha3_512, 4>(A, m, 0x06);
  }
  uint8_t* o = out + (size_t)x * 64;
#pragma unroll
  for (int w = 0; w < 8; ++w) store_u64(o + 8 * w, A[w]);
}

// KDF: SHAKE256(kr[0:32] || h, 32) (symmetric.h:33).
__global__ void k_kdf(int n, uint8_t* ss, const uint8_t* kr, const uint8_t* h) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t m[8], A[25];
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    m[w] = load_u64(kr + (size_t)x * 64 + 8 * w);
    m[4 + w] = load_u64(h + (size_t)x * 32 + 8 * w);
  }
  absorb_fixed<kRateShake256, 8>(A, m, 0x1F);
#pragma unroll
  for (int w = 0; w < 4; ++w) store_u64(ss + (size_t)x * 32 + 8 * w, A[w]);
}

// verify (verify.cu:15-28, restated without the early exit): fail = ct != ct'.
__global__ void k_verify(int n, uint32_t* fail, const uint8_t* ct, const uint8_t* ct2, int words) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint64_t* a = reinterpret_cast<const uint64_t*>(ct + (size_t)x * words * 8);
  const uint64_t* b = reinterpret_cast<const uint64_t*>(ct2 + (size_t)x * words * 8);
  uint64_t r = 0;
  for (int w = 0; w < words; ++w) r |= a[w] ^ b[w];
  fail[x] = (uint32_t)((0 - r) >> 63) | (uint32_t)(r >> 63);
}

// cmov (verify.cu:43-50): kr[0:32] = fail ? z : kr[0:32].
__global__ void k_cmov(int n, uint8_t* kr, const uint8_t* sk, size_t sk_stride, int z_off,
                       const uint32_t* fail) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint64_t mask = 0 - (uint64_t)(fail[x] & 1);
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    uint64_t r = load_u64(kr + (size_t)x * 64 + 8 * w);
    const uint64_t z = load_u64(sk + (size_t)x * sk_stride + z_off + 8 * w);
    r ^= mask & (z ^ r);
    store_u64(kr + (size_t)x * 64 + 8 * w, r);
  }
}

// ---- sampling kernels ------------------------------------------------------

// poly_getnoise (poly.cu:129-142): CBD(SHAKE256(key || nonce, 128)).
__global__ void k_getnoise(int n, Pv r, const uint8_t* key, size_t key_stride, uint8_t nonce) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t m[4], A[25];
#pragma unroll
  for (int w = 0; w < 4; ++w) m[w] = load_u64(key + (size_t)x * key_stride + 8 * w);
  absorb_fixed<kRateShake256, 4>(A, m, (uint64_t)nonce | (0x1Full << 8));
#pragma unroll
  for (int w = 0; w < 16; ++w) {
    const uint64_t t64 = A[w];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t t = (uint32_t)(t64 >> (32 * h));
      const uint32_t d = (t & 0x55555555u) + ((t >> 1) & 0x55555555u);
#pragma unroll
      for (int j = 0; j < 8; ++j)
        r.at(16 * w + 8 * h + j, x) = (int16_t)((int)((d >> (4 * j)) & 3) - (int)((d >> (4 * j + 2)) & 3));
    }
  }
}

// gen_matrix_n (indcpa.cu:339-403) with rej_uniform (indcpa.cu:290-322): one
// thread per instance generates all K*K entries; the squeezed block is sampled
// straight out of the state registers.
template <int K>
__global__ void k_gen_matrix(int n, Pv a0, size_t poly_stride, const uint8_t* seed, size_t seed_stride,
                             int transposed) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t rho[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) rho[w] = load_u64(seed + (size_t)x * seed_stride + 8 * w);
  for (int i = 0; i < K; ++i) {
    for (int j = 0; j < K; ++j) {
      const uint64_t xi = transposed ? i : j, yi = transposed ? j : i;
      uint64_t A[25];
#pragma unroll
      for (int t = 0; t < 25; ++t) A[t] = (t < 4) ? rho[t] : 0;
      A[4] = xi | (yi << 8) | (0x1Full << 16);
      A[kRateShake128 - 1] ^= 0x8000000000000000ull;
      Pv a{a0.p + (size_t)(i * K + j) * poly_stride, a0.ld};
      int ctr = 0;
      // at most 32 blocks: 2688 candidates for 256 accepts at a 96.5% accept
      // rate is never short; the cap only guarantees termination
      for (int blk = 0; blk < 32 && ctr < N; ++blk) {
        keccakf1600(A);
#pragma unroll
        for (int w = 0; w < kRateShake128; ++w) {
#pragma unroll
          for (int h = 0; h < 4; ++h) {
            const uint32_t val = (uint32_t)(A[w] >> (16 * h)) & 0xFFFFu;
            if (val < 19u * Q && ctr < N) {
              a.at(ctr, x) = (int16_t)(val - (val >> 12) * Q);
              ++ctr;
            }
          }
        }
      }
    }
  }
}

// ---- polynomial kernels ------------------------------------------------------

__device__ __forceinline__ int16_t fq(int16_t a, int16_t b) { return (int16_t)fqmul(a, b); }

// polyvec_ntt_n (polyvec.cu:124-133): ntt_p + poly_reduce on `cnt` polys.
__global__ void k_ntt(int n, Pv r0, size_t poly_stride, int cnt) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int p = 0; p < cnt; ++p) {
    Pv r{r0.p + (size_t)p * poly_stride, r0.ld};
    int k = 1;
    for (int len = 128; len >= 2; len >>= 1) {
      for (int start = 0; start < N; start += 2 * len) {
        const int16_t zeta = c_zetas[k++];
        for (int j = start; j < start + len; ++j) {
          const int16_t t = fq(zeta, r.at(j + len, x));
          const int16_t a = r.at(j, x);
          r.at(j + len, x) = (int16_t)(a - t);
          r.at(j, x) = (int16_t)(a + t);
        }
      }
    }
    for (int c = 0; c < N; ++c) r.at(c, x) = (int16_t)barrett(r.at(c, x));
  }
}

// polyvec_invntt_n / poly_invntt_n (invntt_p, ntt.cu:197-224).
__global__ void k_invntt(int n, Pv r0, size_t poly_stride, int cnt) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int p = 0; p < cnt; ++p) {
    Pv r{r0.p + (size_t)p * poly_stride, r0.ld};
    int k = 0;
    for (int len = 2; len <= 128; len <<= 1) {
      for (int start = 0; start < N; start += 2 * len) {
        const int16_t zeta = c_zetas_inv[k++];
        for (int j = start; j < start + len; ++j) {
          const int16_t t = r.at(j, x);
          const int16_t u = r.at(j + len, x);
          r.at(j, x) = (int16_t)barrett((int16_t)(t + u));
          r.at(j + len, x) = fq(zeta, (int16_t)(t - u));
        }
      }
    }
    const int16_t f = c_zetas_inv[127];
    for (int c = 0; c < N; ++c) r.at(c, x) = fq(r.at(c, x), f);
  }
}

// polyvec_pointwise_acc_n (polyvec.cu:161-177): r = reduce(sum_i a_i o b_i).
template <int K>
__global__ void k_pointwise_acc(int n, Pv r, Pv a0, Pv b0, size_t poly_stride) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int g = 0; g < N / 4; ++g) {
    const int16_t zeta = c_zetas[64 + g];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int16_t z = s ? (int16_t)-zeta : zeta;
      const int c = 4 * g + 2 * s;
      int16_t r0 = 0, r1 = 0;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        Pv a{a0.p + (size_t)i * poly_stride, a0.ld}, b{b0.p + (size_t)i * poly_stride, b0.ld};
        const int16_t a_0 = a.at(c, x), a_1 = a.at(c + 1, x), b_0 = b.at(c, x), b_1 = b.at(c + 1, x);
        const int16_t t0 = (int16_t)(fq(fq(a_1, b_1), z) + fq(a_0, b_0));
        const int16_t t1 = (int16_t)(fq(a_0, b_1) + fq(a_1, b_0));
        r0 = (int16_t)(r0 + t0);
        r1 = (int16_t)(r1 + t1);
      }
      r.at(c, x) = (int16_t)barrett(r0);
      r.at(c + 1, x) = (int16_t)barrett(r1);
    }
  }
}

// poly_frommont_n (poly.cu:235-247): multiply by 2^32 mod q, Montgomery-reduce.
__global__ void k_frommont(int n, Pv r) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int16_t f = (int16_t)((1ull << 32) % Q);
  for (int c = 0; c < N; ++c) r.at(c, x) = (int16_t)mont_reduce((int32_t)r.at(c, x) * f);
}

// poly_add_n / poly_sub_n / polyvec_add_n (poly.cu:316-347, polyvec.cu:233-241)
__global__ void k_add(int n, Pv r0, Pv a0, Pv b0, size_t poly_stride, int cnt, int subtract) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int p = 0; p < cnt; ++p) {
    Pv r{r0.p + (size_t)p * poly_stride, r0.ld}, a{a0.p + (size_t)p * poly_stride, a0.ld},
        b{b0.p + (size_t)p * poly_stride, b0.ld};
    for (int c = 0; c < N; ++c)
      r.at(c, x) = subtract ? (int16_t)(a.at(c, x) - b.at(c, x)) : (int16_t)(a.at(c, x) + b.at(c, x));
  }
}

// poly_reduce_n / polyvec_reduce_n (poly.cu:257-267, polyvec.cu:197-205)
__global__ void k_reduce(int n, Pv r0, size_t poly_stride, int cnt) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int p = 0; p < cnt; ++p) {
    Pv r{r0.p + (size_t)p * poly_stride, r0.ld};
    for (int c = 0; c < N; ++c) r.at(c, x) = (int16_t)barrett(r.at(c, x));
  }
}

// ---- (de)serialisation kernels -------------------------------------------------

// unpack_ciphertext_n (indcpa.cu:265-275) with round-2 widths.
template <int K>
__global__ void k_unpack_ct(int n, Pv bp0, size_t poly_stride, Pv v, const uint8_t* ct) {
  using P = Params<K>;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint8_t* c0 = ct + (size_t)x * P::ct_bytes;
  for (int i = 0; i < K; ++i) {
    Pv b{bp0.p + (size_t)i * poly_stride, bp0.ld};
    for (int c = 0; c < N; ++c)
      b.at(c, x) = (int16_t)decompress_d<P::du>(get_field<P::du>(c0 + 32 * P::du * i, c));
  }
  for (int c = 0; c < N; ++c)
    v.at(c, x) = (int16_t)decompress_d<P::dv>(get_field<P::dv>(c0 + P::polyvec_cbytes, c));
}

// unpack_sk_n / unpack_pk_n (indcpa.cu:178-230): 12-bit polyvec, optional seed.
template <int K>
__global__ void k_unpack_polyvec(int n, Pv r0, size_t poly_stride, const uint8_t* src, size_t stride,
                                 uint8_t* seed_out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint8_t* s = src + (size_t)x * stride;
  for (int i = 0; i < K; ++i) {
    Pv r{r0.p + (size_t)i * poly_stride, r0.ld};
    for (int c = 0; c < N; ++c) r.at(c, x) = (int16_t)get_field<12>(s + 384 * i, c);
  }
  if (seed_out) {
#pragma unroll
    for (int w = 0; w < 4; ++w) store_u64(seed_out + (size_t)x * 32 + 8 * w, load_u64(s + 384 * K + 8 * w));
  }
}

// pack_sk_n / pack_pk_n (indcpa.cu:153-210): csubq then 12-bit polyvec, optional seed.
template <int K>
__global__ void k_pack_polyvec(int n, uint8_t* dst, size_t stride, Pv r0, size_t poly_stride,
                               const uint8_t* seed, size_t seed_stride) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint8_t* d = dst + (size_t)x * stride;
  for (int i = 0; i < K; ++i) {
    Pv r{r0.p + (size_t)i * poly_stride, r0.ld};
    put_fields<12>(d + 384 * i, [&](int c) { return (uint32_t)csubq(r.at(c, x)); });
  }
  if (seed) {
#pragma unroll
    for (int w = 0; w < 4; ++w)
      store_u64(d + 384 * K + 8 * w, load_u64(seed + (size_t)x * seed_stride + 8 * w));
  }
}

// pack_ciphertext_n (indcpa.cu:243-253): csubq, compress du/dv, pack.
template <int K>
__global__ void k_pack_ct(int n, uint8_t* ct, Pv bp0, size_t poly_stride, Pv v) {
  using P = Params<K>;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint8_t* c0 = ct + (size_t)x * P::ct_bytes;
  for (int i = 0; i < K; ++i) {
    Pv b{bp0.p + (size_t)i * poly_stride, bp0.ld};
    put_fields<P::du>(c0 + 32 * P::du * i, [&](int c) { return compress_d<P::du>(csubq(b.at(c, x))); });
  }
  put_fields<P::dv>(c0 + P::polyvec_cbytes, [&](int c) { return compress_d<P::dv>(csubq(v.at(c, x))); });
}

// poly_frommsg_n (poly.cu:367-386)
__global__ void k_frommsg(int n, Pv r, const uint8_t* msg) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int i = 0; i < 32; ++i) {
    const uint32_t b = msg[(size_t)x * 32 + i];
    for (int j = 0; j < 8; ++j) r.at(8 * i + j, x) = (int16_t)((0u - ((b >> j) & 1)) & ((Q + 1) / 2));
  }
}

// poly_tomsg_n (poly.cu:396-417)
__global__ void k_tomsg(int n, uint8_t* msg, Pv a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int i = 0; i < 32; ++i) {
    uint32_t byte = 0;
    for (int j = 0; j < 8; ++j) byte |= compress_d<1>(csubq(a.at(8 * i + j, x))) << j;
    msg[(size_t)x * 32 + i] = (uint8_t)byte;
  }
}

// KEM keypair tail: sk = sk_cpa || pk || H(pk) || z.
template <int K>
__global__ void k_sk_tail(int n, uint8_t* sk, const uint8_t* pk, const uint8_t* coins) {
  using P = Params<K>;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint8_t* s = sk + (size_t)x * P::sk_bytes;
  const uint8_t* p = pk + (size_t)x * P::pk_bytes;
  for (int w = 0; w < P::pk_bytes / 8; ++w) store_u64(s + P::sk_cpa_bytes + 8 * w, load_u64(p + 8 * w));
#pragma unroll
  for (int w = 0; w < 4; ++w) store_u64(s + P::z_off + 8 * w, load_u64(coins + (size_t)x * 64 + 32 + 8 * w));
}

// ---- host orchestration ------------------------------------------------------

struct Scratch {
  pqc_ctx* c;
  size_t ps;  // elements per poly = 256 * ld
  int block;  // threads per workgroup (tuning)
  Pv poly(int i) const { return Pv{c->kpolys + (size_t)i * ps, (int)c->ld}; }
  // field region `off` holds that field for every element; ld is a multiple of
  // 64, so every region start stays 8-byte aligned
  uint8_t* bytes(size_t off) const { return c->kbytes + off * c->ld; }
};

// byte-buffer offsets (per element)
constexpr size_t kMsg = 0, kKr = 32, kHc = 96, kSeed = 128, kBuf64 = 160, kFail = 224, kCt2 = 228;

static bool g_tables_ready[64] = {false};

static int ensure_tables(int dev) {
  if (dev < 0 || dev >= 64) return PQC_ERR_DEVICE;
  if (g_tables_ready[dev]) return PQC_OK;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_zetas), kZetas.fwd, sizeof(kZetas.fwd)) != hipSuccess)
    return PQC_ERR_DEVICE;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_zetas_inv), kZetas.inv, sizeof(kZetas.inv)) != hipSuccess)
    return PQC_ERR_DEVICE;
  g_tables_ready[dev] = true;
  return PQC_OK;
}

#define LAUNCH(kern, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(S.block), 0, s, __VA_ARGS__)

// indcpa_enc (indcpa.cu:578-679) for n elements: m (32 B/elem at msg),
// pk (stride pk_stride), coins (32 B at coins + x*coins_stride) -> ct.
template <int K>
static void indcpa_enc(const Scratch& S, int n, uint8_t* ct, const uint8_t* msg, const uint8_t* pk,
                       size_t pk_stride, const uint8_t* coins, size_t coins_stride, hipStream_t s) {
  const int grid = (n + S.block - 1) / S.block;
  const size_t ps = S.ps;
  // poly slots: at[0..K*K), sp, ep, epp, pkpv, kpoly, bp, v
  const int AT = 0, SP = K * K, EP = SP + K, EPP = EP + K, PKPV = EPP + 1, KP = PKPV + K, BP = KP + 1,
            V = BP + K;
  LAUNCH(k_unpack_polyvec<K>, n, S.poly(PKPV), ps, pk, pk_stride, S.bytes(kSeed));
  LAUNCH(k_frommsg, n, S.poly(KP), msg);
  LAUNCH(k_gen_matrix<K>, n, S.poly(AT), ps, (const uint8_t*)S.bytes(kSeed), (size_t)32, 1);
  uint8_t nonce = 0;
  for (int i = 0; i < K; ++i) LAUNCH(k_getnoise, n, S.poly(SP + i), coins, coins_stride, nonce++);
  for (int i = 0; i < K; ++i) LAUNCH(k_getnoise, n, S.poly(EP + i), coins, coins_stride, nonce++);
  LAUNCH(k_getnoise, n, S.poly(EPP), coins, coins_stride, nonce++);
  LAUNCH(k_ntt, n, S.poly(SP), ps, K);
  for (int i = 0; i < K; ++i) LAUNCH(k_pointwise_acc<K>, n, S.poly(BP + i), S.poly(AT + K * i), S.poly(SP), ps);
  LAUNCH(k_pointwise_acc<K>, n, S.poly(V), S.poly(PKPV), S.poly(SP), ps);
  LAUNCH(k_invntt, n, S.poly(BP), ps, K);
  LAUNCH(k_invntt, n, S.poly(V), ps, 1);
  LAUNCH(k_add, n, S.poly(BP), S.poly(BP), S.poly(EP), ps, K, 0);
  LAUNCH(k_add, n, S.poly(V), S.poly(V), S.poly(EPP), ps, 1, 0);
  LAUNCH(k_add, n, S.poly(V), S.poly(V), S.poly(KP), ps, 1, 0);
  LAUNCH(k_reduce, n, S.poly(BP), ps, K);
  LAUNCH(k_reduce, n, S.poly(V), ps, 1);
  LAUNCH(k_pack_ct<K>, n, ct, S.poly(BP), ps, S.poly(V));
}

template <int K>
static int plain_dec(pqc_ctx* c, size_t n_, uint8_t* ss, const uint8_t* ct, const uint8_t* sk, hipStream_t s) {
  using P = Params<K>;
  const int n = (int)n_;
  Scratch S{c, (size_t)N * c->ld, c->tune.plain_block};
  const int grid = (n + S.block - 1) / S.block;
  const size_t ps = S.ps;
  // decryption slots live after the encryption slots
  const int base = K * K + 5 * K + 4;
  const int BPD = base, VD = BPD + K, SKPV = VD + 1, MP = SKPV + K;
  LAUNCH(k_unpack_ct<K>, n, S.poly(BPD), ps, S.poly(VD), ct);
  LAUNCH(k_unpack_polyvec<K>, n, S.poly(SKPV), ps, sk, (size_t)P::sk_bytes, (uint8_t*)nullptr);
  LAUNCH(k_ntt, n, S.poly(BPD), ps, K);
  LAUNCH(k_pointwise_acc<K>, n, S.poly(MP), S.poly(SKPV), S.poly(BPD), ps);
  LAUNCH(k_invntt, n, S.poly(MP), ps, 1);
  LAUNCH(k_add, n, S.poly(MP), S.poly(VD), S.poly(MP), ps, 1, 1);
  LAUNCH(k_reduce, n, S.poly(MP), ps, 1);
  LAUNCH(k_tomsg, n, S.bytes(kMsg), S.poly(MP));
  // (Kbar, r) = G(m' || H(pk))
  LAUNCH(k_hash_g, n, S.bytes(kKr), (const uint8_t*)S.bytes(kMsg), (size_t)32, sk + P::hpk_off, (size_t)P::sk_bytes);
  uint8_t* ct2 = S.bytes(kCt2);
  indcpa_enc<K>(S, n, ct2, S.bytes(kMsg), sk + P::sk_cpa_bytes, P::sk_bytes, S.bytes(kKr) + 32, 64, s);
  LAUNCH(k_hash_h, n, S.bytes(kHc), (size_t)32, ct, (size_t)P::ct_bytes, P::ct_bytes / 8);
  LAUNCH(k_verify, n, reinterpret_cast<uint32_t*>(S.bytes(kFail)), ct, (const uint8_t*)ct2, P::ct_bytes / 8);
  LAUNCH(k_cmov, n, S.bytes(kKr), sk, (size_t)P::sk_bytes, P::z_off, reinterpret_cast<const uint32_t*>(S.bytes(kFail)));
  LAUNCH(k_kdf, n, ss, (const uint8_t*)S.bytes(kKr), (const uint8_t*)S.bytes(kHc));
  return hip_err(hipGetLastError());
}

template <int K>
static int plain_enc(pqc_ctx* c, size_t n_, uint8_t* ct, uint8_t* ss, const uint8_t* pk, const uint8_t* coins,
                     hipStream_t s) {
  using P = Params<K>;
  const int n = (int)n_;
  Scratch S{c, (size_t)N * c->ld, c->tune.plain_block};
  const int grid = (n + S.block - 1) / S.block;
  // buf64 = m || H(pk); m = H(coins) ("don't release system RNG output")
  LAUNCH(k_hash_h, n, S.bytes(kBuf64), (size_t)64, coins, (size_t)32, 4);
  LAUNCH(k_hash_h, n, S.bytes(kBuf64) + 32, (size_t)64, pk, (size_t)P::pk_bytes, P::pk_bytes / 8);
  LAUNCH(k_hash_g, n, S.bytes(kKr), (const uint8_t*)S.bytes(kBuf64), (size_t)64,
         (const uint8_t*)S.bytes(kBuf64) + 32, (size_t)64);
  // indcpa_enc reads m at stride 32: copy m out of buf64
  if (hipMemcpy2DAsync(S.bytes(kMsg), 32, S.bytes(kBuf64), 64, 32, n, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return PQC_ERR_DEVICE;
  indcpa_enc<K>(S, n, ct, S.bytes(kMsg), pk, P::pk_bytes, S.bytes(kKr) + 32, 64, s);
  LAUNCH(k_hash_h, n, S.bytes(kHc), (size_t)32, (const uint8_t*)ct, (size_t)P::ct_bytes, P::ct_bytes / 8);
  LAUNCH(k_kdf, n, ss, (const uint8_t*)S.bytes(kKr), (const uint8_t*)S.bytes(kHc));
  return hip_err(hipGetLastError());
}

template <int K>
static int plain_keypair(pqc_ctx* c, size_t n_, uint8_t* pk, uint8_t* sk, const uint8_t* coins, hipStream_t s) {
  using P = Params<K>;
  const int n = (int)n_;
  Scratch S{c, (size_t)N * c->ld, c->tune.plain_block};
  const int grid = (n + S.block - 1) / S.block;
  const size_t ps = S.ps;
  const int A = 0, SKPV = K * K, E = SKPV + K, PKPV = E + K;
  // (publicseed, noiseseed) = G(d)
  LAUNCH(k_hash_g, n, S.bytes(kBuf64), coins, (size_t)64, (const uint8_t*)nullptr, (size_t)0);
  LAUNCH(k_gen_matrix<K>, n, S.poly(A), ps, (const uint8_t*)S.bytes(kBuf64), (size_t)64, 0);
  uint8_t nonce = 0;
  for (int i = 0; i < K; ++i) LAUNCH(k_getnoise, n, S.poly(SKPV + i), (const uint8_t*)S.bytes(kBuf64) + 32, (size_t)64, nonce++);
  for (int i = 0; i < K; ++i) LAUNCH(k_getnoise, n, S.poly(E + i), (const uint8_t*)S.bytes(kBuf64) + 32, (size_t)64, nonce++);
  LAUNCH(k_ntt, n, S.poly(SKPV), ps, K);
  LAUNCH(k_ntt, n, S.poly(E), ps, K);
  for (int i = 0; i < K; ++i) {
    LAUNCH(k_pointwise_acc<K>, n, S.poly(PKPV + i), S.poly(A + K * i), S.poly(SKPV), ps);
    LAUNCH(k_frommont, n, S.poly(PKPV + i));
  }
  LAUNCH(k_add, n, S.poly(PKPV), S.poly(PKPV), S.poly(E), ps, K, 0);
  LAUNCH(k_reduce, n, S.poly(PKPV), ps, K);
  LAUNCH(k_pack_polyvec<K>, n, sk, (size_t)P::sk_bytes, S.poly(SKPV), ps, (const uint8_t*)nullptr, (size_t)0);
  LAUNCH(k_pack_polyvec<K>, n, pk, (size_t)P::pk_bytes, S.poly(PKPV), ps, (const uint8_t*)S.bytes(kBuf64), (size_t)64);
  LAUNCH(k_sk_tail<K>, n, sk, (const uint8_t*)pk, coins);
  LAUNCH(k_hash_h, n, sk + P::hpk_off, (size_t)P::sk_bytes, (const uint8_t*)pk, (size_t)P::pk_bytes, P::pk_bytes / 8);
  return hip_err(hipGetLastError());
}

}  // namespace plain
}  // namespace kyber

size_t kyber_plain_polys(int kmax) {
  // encryption slots K*K + 5K + 4 (at, sp, ep, epp, pkpv, kpoly, bp, v) plus
  // decryption slots 3K + 2 (bp, v, skpv, mp)
  return (size_t)kmax * kmax + 8 * kmax + 6;
}

size_t kyber_plain_bytes_per_elem(int kmax) {
  (void)kmax;
  return kyber::plain::kCt2 + 1568 + 28;
}

#define DISPATCH_K(k, fn, ...)                          \
  switch (k) {                                          \
    case 2: return kyber::plain::fn<2>(__VA_ARGS__);    \
    case 3: return kyber::plain::fn<3>(__VA_ARGS__);    \
    case 4: return kyber::plain::fn<4>(__VA_ARGS__);    \
    default: return PQC_ERR_UNSUPPORTED;                \
  }

int kyber_plain_keypair(pqc_ctx* c, int k, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* coins,
                        hipStream_t s) {
  int rc = kyber::plain::ensure_tables(c->device);
  if (rc) return rc;
  DISPATCH_K(k, plain_keypair, c, n, pk, sk, coins, s);
}

int kyber_plain_enc(pqc_ctx* c, int k, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk,
                    const uint8_t* coins, hipStream_t s) {
  int rc = kyber::plain::ensure_tables(c->device);
  if (rc) return rc;
  DISPATCH_K(k, plain_enc, c, n, ct, ss, pk, coins, s);
}

int kyber_plain_dec(pqc_ctx* c, int k, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk,
                    hipStream_t s) {
  int rc = kyber::plain::ensure_tables(c->device);
  if (rc) return rc;
  DISPATCH_K(k, plain_dec, c, n, ss, ct, sk, s);
}

}  // namespace pqc
