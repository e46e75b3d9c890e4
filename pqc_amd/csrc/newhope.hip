// NewHope-1024 CPA-PKE as plain batched kernels (NewHope_NIST_GPU/cpapke.cu:
// 265-476): one kernel per reference operation, one thread per instance,
// polynomials in coefficient-major SoA arrays (coefficient c of element x at
// c * ld + x) so that every load and store is coalesced across the wave.
//
// The 16-bit storage and 32-bit intermediate arithmetic of the reference are
// kept exactly (lazy even NTT levels stored as uint16, Barrett on a uint16
// argument, Montgomery with R = 2^18 in uint32), so results are identical for
// every input, including out-of-range ones.
#include "internal.h"
#include "sponge.cuh"

namespace pqc {
namespace newhope {

constexpr int N = 1024;
constexpr uint32_t Q = 12289;
constexpr int SYM = 32;
constexpr int POLYBYTES = 14 * N / 8;  // 1792
constexpr int CBYTES = 3 * N / 8;      // 384
constexpr int PK = POLYBYTES + SYM;    // 1824
constexpr int SK = POLYBYTES;          // 1792
constexpr int CT = POLYBYTES + CBYTES; // 2176

// ---- tables: the Pari/GP definitions of precomp.cu:5-72 (n = 1024, g = 7,
// mont = 2^18), evaluated at compile time -------------------------------------
constexpr uint32_t powmod(uint32_t b, uint32_t e) {
  uint64_t r = 1, x = b % Q;
  while (e) {
    if (e & 1) r = r * x % Q;
    x = x * x % Q;
    e >>= 1;
  }
  return (uint32_t)r;
}

constexpr uint32_t brv10(uint32_t i) {
  uint32_t r = 0;
  for (int b = 0; b < 10; b++) r |= ((i >> b) & 1u) << (9 - b);
  return r;
}

struct Tables {
  uint16_t omegas_inv[N / 2];
  uint16_t gammas[N];
  uint16_t gammas_inv[N];
  uint16_t brv[N];
};

constexpr Tables make_tables() {
  Tables t{};
  const uint32_t g = 7, mont = (1u << 18) % Q;
  const uint32_t g2inv = powmod(g * g % Q, Q - 2), ginv = powmod(g, Q - 2), ninv = powmod(N, Q - 2);
  for (uint32_t k = 0; k < N / 2; k++) t.omegas_inv[k] = (uint16_t)((uint64_t)powmod(g2inv, brv10(2 * k)) * mont % Q);
  for (uint32_t k = 0; k < N; k++) {
    t.gammas[k] = (uint16_t)((uint64_t)powmod(g, brv10(k)) * mont % Q);
    t.gammas_inv[k] = (uint16_t)((uint64_t)powmod(ginv, k) * ninv % Q * mont % Q);
    t.brv[k] = (uint16_t)brv10(k);
  }
  return t;
}

__constant__ Tables c_tab = make_tables();

// reduce.cu:22-31
__device__ __forceinline__ uint32_t mont(uint32_t a) {
  uint32_t u = (a * 12287u) & ((1u << 18) - 1);
  u *= Q;
  a = a + u;
  return (a >> 18) & 0xFFFFu;
}

// reduce.cu:33-47 on a uint16 argument
__device__ __forceinline__ uint32_t barrett(uint32_t a16) {
  const uint32_t a = a16 & 0xFFFFu;
  const uint32_t u = ((a * 5) >> 16) * Q;
  return (a - u) & 0xFFFFu;
}

struct Pv {
  uint16_t* p;
  int ld;
  __device__ __forceinline__ uint16_t& at(int c, int x) const { return p[(size_t)c * ld + x]; }
};

// ---- sponge kernels --------------------------------------------------------

// keypair seed expansion: z = SHAKE256(seed, 64) (cpapke.cu:283-298)
__global__ void k_seed_expand(int n, uint8_t* z, const uint8_t* seed) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t m[4], A[25];
#pragma unroll
  for (int w = 0; w < 4; ++w) m[w] = load_u64(seed + (size_t)x * 32 + 8 * w);
  absorb_fixed<kRateShake256, 4>(A, m, 0x1F);
#pragma unroll
  for (int w = 0; w < 8; ++w) store_u64(z + (size_t)x * 64 + 8 * w, A[w]);
}

// gen_a / poly_uniform (cpapke.cu:236-247, poly.cu:332-366)
__global__ void k_uniform(int n, Pv a, const uint8_t* seed, size_t stride) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t s[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) s[w] = load_u64(seed + (size_t)x * stride + 8 * w);
  for (int i = 0; i < N / 64; ++i) {
    uint64_t A[25];
#pragma unroll
    for (int t = 0; t < 25; ++t) A[t] = t < 4 ? s[t] : 0;
    A[4] = (uint64_t)i | (0x1Full << 8);
    A[kRateShake128 - 1] ^= 0x8000000000000000ull;
    int ctr = 0;
    for (int blk = 0; blk < 16 && ctr < 64; ++blk) {
      keccakf1600(A);
#pragma unroll
      for (int w = 0; w < kRateShake128; ++w) {
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          const uint32_t val = (uint32_t)(A[w] >> (16 * h)) & 0xFFFFu;
          if (val < 5 * Q && ctr < 64) {
            a.at(i * 64 + ctr, x) = (uint16_t)val;
            ++ctr;
          }
        }
      }
    }
  }
}

// poly_sample (poly.cu:396-444): hw(a) + q - hw(b) over SHAKE256(seed || nonce || i, 128)
__global__ void k_sample(int n, Pv r, const uint8_t* seed, size_t stride, int nonce) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t s[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) s[w] = load_u64(seed + (size_t)x * stride + 8 * w);
  for (int i = 0; i < N / 64; ++i) {
    uint64_t A[25];
    absorb_fixed<kRateShake256, 4>(A, s, (uint64_t)nonce | ((uint64_t)i << 8) | (0x1Full << 16));
#pragma unroll
    for (int w = 0; w < 16; ++w) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t ab = (uint32_t)(A[w] >> (16 * k)) & 0xFFFFu;
        r.at(64 * i + 4 * w + k, x) = (uint16_t)(__popc(ab & 0xFF) + Q - __popc(ab >> 8));
      }
    }
  }
}

// ---- NTT kernels ----------------------------------------------------------------

// ntt (ntt.cu:195-234) on one SoA polynomial
__device__ __forceinline__ void ntt_core(const Pv& a, int x, const uint16_t* omega) {
  for (int i = 0; i < 10; i += 2) {
    int distance = 1 << i;
    for (int start = 0; start < distance; start++) {
      int jt = 0;
      for (int j = start; j < N - 1; j += 2 * distance) {
        const uint32_t W = omega[jt++];
        const uint32_t temp = a.at(j, x), hi = a.at(j + distance, x);
        a.at(j, x) = (uint16_t)(temp + hi);
        a.at(j + distance, x) = (uint16_t)mont(W * (temp + 3 * Q - hi));
      }
    }
    distance <<= 1;
    for (int start = 0; start < distance; start++) {
      int jt = 0;
      for (int j = start; j < N - 1; j += 2 * distance) {
        const uint32_t W = omega[jt++];
        const uint32_t temp = a.at(j, x), hi = a.at(j + distance, x);
        a.at(j, x) = (uint16_t)barrett(temp + hi);
        a.at(j + distance, x) = (uint16_t)mont(W * (temp + 3 * Q - hi));
      }
    }
  }
}

// poly_ntt (poly.cu:519-527)
__global__ void k_ntt(int n, Pv r) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int c = 0; c < N; ++c) r.at(c, x) = (uint16_t)mont((uint32_t)r.at(c, x) * c_tab.gammas[c]);
  ntt_core(r, x, c_tab.gammas);
}

// poly_invntt (poly.cu:538-547)
__global__ void k_invntt(int n, Pv r) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int i = 0; i < N; ++i) {
    const int j = c_tab.brv[i];
    if (i < j) {
      const uint16_t t = r.at(i, x);
      r.at(i, x) = r.at(j, x);
      r.at(j, x) = t;
    }
  }
  ntt_core(r, x, c_tab.omegas_inv);
  for (int c = 0; c < N; ++c) r.at(c, x) = (uint16_t)mont((uint32_t)r.at(c, x) * c_tab.gammas_inv[c]);
}

// poly_mul_pointwise (poly.cu:455-468)
__global__ void k_pointwise(int n, Pv r, Pv a, Pv b) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int c = 0; c < N; ++c) {
    const uint32_t t = mont(3186u * b.at(c, x));
    r.at(c, x) = (uint16_t)mont((uint32_t)a.at(c, x) * t);
  }
}

// poly_add / poly_sub (poly.cu:479-508)
__global__ void k_addsub(int n, Pv r, Pv a, Pv b, int subtract) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int c = 0; c < N; ++c) {
    const uint32_t av = a.at(c, x), bv = b.at(c, x);
    r.at(c, x) = (uint16_t)((subtract ? av + 3 * Q - bv : av + bv) % Q);
  }
}

// ---- (de)serialisation --------------------------------------------------------

// coeff_freeze (poly.cu:21-33)
__device__ __forceinline__ uint32_t freeze(uint32_t x) {
  const uint32_t r = x % Q;
  const uint32_t m = (r - Q) & 0xFFFFu;
  const uint32_t c = (uint32_t)((int32_t)(int16_t)m >> 15) & 0xFFFFu;
  return (m ^ ((r ^ m) & c)) & 0xFFFFu;
}

// flipabs (poly.cu:44-52)
__device__ __forceinline__ uint32_t flipabs(uint32_t x) {
  int32_t r = (int16_t)freeze(x);
  r = (int16_t)(r - (int32_t)(Q / 2));
  const int32_t m = (int16_t)(r >> 15);
  return (uint32_t)(uint16_t)((r + m) ^ m);
}

// poly_tobytes (poly.cu:105-127) into dst, optional seed after the polynomial
__global__ void k_tobytes(int n, uint8_t* dst, size_t stride, Pv p, const uint8_t* seed, size_t seed_stride) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint8_t* r = dst + (size_t)x * stride;
  for (int i = 0; i < N / 4; i++) {
    const uint32_t t0 = freeze(p.at(4 * i, x)), t1 = freeze(p.at(4 * i + 1, x)), t2 = freeze(p.at(4 * i + 2, x)),
                   t3 = freeze(p.at(4 * i + 3, x));
    r[7 * i + 0] = (uint8_t)(t0 & 0xff);
    r[7 * i + 1] = (uint8_t)((t0 >> 8) | (t1 << 6));
    r[7 * i + 2] = (uint8_t)(t1 >> 2);
    r[7 * i + 3] = (uint8_t)((t1 >> 10) | (t2 << 4));
    r[7 * i + 4] = (uint8_t)(t2 >> 4);
    r[7 * i + 5] = (uint8_t)((t2 >> 12) | (t3 << 2));
    r[7 * i + 6] = (uint8_t)(t3 >> 6);
  }
  if (seed) {
#pragma unroll
    for (int w = 0; w < 4; ++w) store_u64(r + POLYBYTES + 8 * w, load_u64(seed + (size_t)x * seed_stride + 8 * w));
  }
}

// poly_frombytes (poly.cu:62-74), optional seed copy-out (decode_pk)
__global__ void k_frombytes(int n, Pv r, const uint8_t* src, size_t stride, uint8_t* seed_out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint8_t* a = src + (size_t)x * stride;
  for (int i = 0; i < N / 4; i++) {
    r.at(4 * i + 0, x) = (uint16_t)(a[7 * i + 0] | (((uint32_t)a[7 * i + 1] & 0x3f) << 8));
    r.at(4 * i + 1, x) = (uint16_t)((a[7 * i + 1] >> 6) | ((uint32_t)a[7 * i + 2] << 2) | (((uint32_t)a[7 * i + 3] & 0x0f) << 10));
    r.at(4 * i + 2, x) = (uint16_t)((a[7 * i + 3] >> 4) | ((uint32_t)a[7 * i + 4] << 4) | (((uint32_t)a[7 * i + 5] & 0x03) << 12));
    r.at(4 * i + 3, x) = (uint16_t)((a[7 * i + 5] >> 2) | ((uint32_t)a[7 * i + 6] << 6));
  }
  if (seed_out) {
#pragma unroll
    for (int w = 0; w < 4; ++w) store_u64(seed_out + (size_t)x * 32 + 8 * w, load_u64(a + POLYBYTES + 8 * w));
  }
}

// poly_compress (poly.cu:165-186) into ct + POLYBYTES
__global__ void k_compress(int n, uint8_t* ct, Pv p) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint8_t* r = ct + (size_t)x * CT + POLYBYTES;
  for (int i = 0, k = 0; i < N; i += 8, k += 3) {
    uint32_t t[8];
#pragma unroll
    for (int j = 0; j < 8; j++) t[j] = (((freeze(p.at(i + j, x)) << 3) + Q / 2) / Q) & 7;
    r[k] = (uint8_t)(t[0] | (t[1] << 3) | (t[2] << 6));
    r[k + 1] = (uint8_t)((t[2] >> 2) | (t[3] << 1) | (t[4] << 4) | (t[5] << 7));
    r[k + 2] = (uint8_t)((t[5] >> 1) | (t[6] << 2) | (t[7] << 5));
  }
}

// poly_decompress (poly.cu:197-216) from ct + POLYBYTES
__global__ void k_decompress(int n, Pv r, const uint8_t* ct) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint8_t* a = ct + (size_t)x * CT + POLYBYTES;
  for (int i = 0; i < N; i += 8, a += 3) {
    uint32_t t[8];
    t[0] = a[0] & 7;
    t[1] = (a[0] >> 3) & 7;
    t[2] = (a[0] >> 6) | ((a[1] << 2) & 4);
    t[3] = (a[1] >> 1) & 7;
    t[4] = (a[1] >> 4) & 7;
    t[5] = (a[1] >> 7) | ((a[2] << 1) & 6);
    t[6] = (a[2] >> 2) & 7;
    t[7] = (a[2] >> 5);
#pragma unroll
    for (int j = 0; j < 8; j++) r.at(i + j, x) = (uint16_t)((t[j] * Q + 4) >> 3);
  }
}

// poly_frommsg_n (poly.cu:226-251)
__global__ void k_frommsg(int n, Pv r, const uint8_t* msg) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int i = 0; i < 32; i++) {
    const uint32_t b = msg[(size_t)x * 32 + i];
    for (int j = 0; j < 8; j++) {
      const uint16_t v = (uint16_t)((0u - ((b >> j) & 1)) & (Q / 2));
#pragma unroll
      for (int q = 0; q < 4; ++q) r.at(8 * i + j + 256 * q, x) = v;
    }
  }
}

// poly_tomsg_n (poly.cu:261-291)
__global__ void k_tomsg(int n, uint8_t* msg, Pv p) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int i = 0; i < 32; i++) {
    uint32_t byte = 0;
    for (int j = 0; j < 8; j++) {
      const int c = 8 * i + j;
      uint32_t t = flipabs(p.at(c, x)) + flipabs(p.at(c + 256, x)) + flipabs(p.at(c + 512, x)) +
                   flipabs(p.at(c + 768, x));
      t = (t - Q) & 0xFFFFu;
      byte |= (t >> 15) << j;
    }
    msg[(size_t)x * 32 + i] = (uint8_t)byte;
  }
}

// ---- orchestration (cpapke.cu:265-476) -------------------------------------------

struct Scratch {
  pqc_ctx* c;
  Pv poly(int i) const { return Pv{c->nhpolys + (size_t)i * N * c->ld, (int)c->ld}; }
};

#define LAUNCH(kern, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, __VA_ARGS__)

}  // namespace newhope

int newhope_keypair(pqc_ctx* c, size_t n_, uint8_t* pk, uint8_t* sk, const uint8_t* seed, hipStream_t s) {
  using namespace newhope;
  const int block = c->tune.newhope_block;
  const int n = (int)n_, grid = (n + block - 1) / block;
  Scratch S{c};
  uint8_t* z = c->nhbytes;  // 64 bytes per element: publicseed || noiseseed
  const int AHAT = 0, EHAT = 1, AS = 2, BHAT = 3, SHAT = 4;
  LAUNCH(k_seed_expand, n, z, seed);
  LAUNCH(k_uniform, n, S.poly(AHAT), (const uint8_t*)z, (size_t)64);
  LAUNCH(k_sample, n, S.poly(SHAT), (const uint8_t*)z + 32, (size_t)64, 0);
  LAUNCH(k_ntt, n, S.poly(SHAT));
  LAUNCH(k_sample, n, S.poly(EHAT), (const uint8_t*)z + 32, (size_t)64, 1);
  LAUNCH(k_ntt, n, S.poly(EHAT));
  LAUNCH(k_pointwise, n, S.poly(AS), S.poly(SHAT), S.poly(AHAT));
  LAUNCH(k_addsub, n, S.poly(BHAT), S.poly(EHAT), S.poly(AS), 0);
  LAUNCH(k_tobytes, n, sk, (size_t)SK, S.poly(SHAT), (const uint8_t*)nullptr, (size_t)0);
  LAUNCH(k_tobytes, n, pk, (size_t)PK, S.poly(BHAT), (const uint8_t*)z, (size_t)64);
  return hip_err(hipGetLastError());
}

int newhope_enc(pqc_ctx* c, size_t n_, uint8_t* ct, const uint8_t* m, const uint8_t* pk, const uint8_t* coin,
                hipStream_t s) {
  using namespace newhope;
  const int block = c->tune.newhope_block;
  const int n = (int)n_, grid = (n + block - 1) / block;
  Scratch S{c};
  uint8_t* seed = c->nhbytes;  // publicseed copy, 32 bytes per element (stride 64)
  const int SP = 0, EP = 1, VP = 2, AHAT = 3, BHAT = 4, EPP = 5, UHAT = 6, V = 7;
  LAUNCH(k_frommsg, n, S.poly(V), m);
  LAUNCH(k_frombytes, n, S.poly(BHAT), pk, (size_t)PK, seed);
  LAUNCH(k_uniform, n, S.poly(AHAT), (const uint8_t*)seed, (size_t)32);
  LAUNCH(k_sample, n, S.poly(SP), coin, (size_t)32, 0);
  LAUNCH(k_sample, n, S.poly(EP), coin, (size_t)32, 1);
  LAUNCH(k_sample, n, S.poly(EPP), coin, (size_t)32, 2);
  LAUNCH(k_ntt, n, S.poly(SP));
  LAUNCH(k_ntt, n, S.poly(EP));
  LAUNCH(k_pointwise, n, S.poly(UHAT), S.poly(AHAT), S.poly(SP));
  LAUNCH(k_addsub, n, S.poly(UHAT), S.poly(UHAT), S.poly(EP), 0);
  LAUNCH(k_pointwise, n, S.poly(VP), S.poly(BHAT), S.poly(SP));
  LAUNCH(k_invntt, n, S.poly(VP));
  LAUNCH(k_addsub, n, S.poly(VP), S.poly(VP), S.poly(EPP), 0);
  LAUNCH(k_addsub, n, S.poly(VP), S.poly(VP), S.poly(V), 0);
  LAUNCH(k_tobytes, n, ct, (size_t)CT, S.poly(UHAT), (const uint8_t*)nullptr, (size_t)0);
  LAUNCH(k_compress, n, ct, S.poly(VP));
  return hip_err(hipGetLastError());
}

int newhope_dec(pqc_ctx* c, size_t n_, uint8_t* m, const uint8_t* ct, const uint8_t* sk, hipStream_t s) {
  using namespace newhope;
  const int block = c->tune.newhope_block;
  const int n = (int)n_, grid = (n + block - 1) / block;
  Scratch S{c};
  const int VP = 0, UHAT = 1, TMP = 2, SHAT = 3;
  LAUNCH(k_frombytes, n, S.poly(SHAT), sk, (size_t)SK, (uint8_t*)nullptr);
  LAUNCH(k_frombytes, n, S.poly(UHAT), ct, (size_t)CT, (uint8_t*)nullptr);
  LAUNCH(k_decompress, n, S.poly(VP), ct);
  LAUNCH(k_pointwise, n, S.poly(TMP), S.poly(SHAT), S.poly(UHAT));
  LAUNCH(k_invntt, n, S.poly(TMP));
  LAUNCH(k_addsub, n, S.poly(TMP), S.poly(TMP), S.poly(VP), 1);
  LAUNCH(k_tomsg, n, m, S.poly(TMP));
  return hip_err(hipGetLastError());
}

}  // namespace pqc
