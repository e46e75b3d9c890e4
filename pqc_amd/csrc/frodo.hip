// FrodoKEM-976-SHAKE on MI355X: batched encapsulation / decapsulation under
// one device-resident key, with the shared-A products on int8 MFMA.
//
// Flow follows FrodoKEM_GPU/kem.c:64-214 (the reference's CPU KEM), the
// arithmetic follows the reference's matrix kernels
// (frodo_macrify_reference.cu:164-361), packing util.cu:18-113, sampling
// noise.c:10-29 with CDF_TABLE frodo976.cu:21.
//
// All requests share A and B, so S'.A (B'' = S'A + E') and S'.B
// (V = S'B + E'') of every request are one GEMM
//     [B'' | V] (8n x 984) = S' (8n x 976) . [A | B] (976 x 984) + [E' | E''] mod 2^16.
// It runs on v_mfma_i32_16x16x64_i8: [A | B] = 256*hi + lo with lo, hi signed
// int8 digits (exact mod 2^16), |S'| <= 10 fits int8, and every int32
// accumulation stays below 976 * 10 * 128 < 2^31.  K is padded 976 -> 1024
// and N 984 -> 1024 with zeros.
#include "internal.h"
#include "sponge.cuh"

namespace pqc {
namespace frodo {

constexpr int N = 976, NB = 8, SS = 24, SEEDA = 16, MU = 24;
constexpr int PK = 15632, CT = 15744, SK = 31296;
constexpr int C1 = 2 * N * NB;        // 15616: B' bytes in ct, B bytes in pk
constexpr int SK_S = SS + PK;         // S^T (LE int16) inside sk
constexpr int SK_PKH = SS + PK + C1;  // H(pk) inside sk
constexpr int KSTEPS = 16;            // K = 1024 / 64
constexpr int NT = 64;                // 16-column tiles of [A | B] (1024 columns)
constexpr int OUTC = 1024;            // row stride of the GEMM output (uint16)
constexpr int NSAMP = (2 * N + NB) * NB;  // 15680 samples in the S'/E'/E'' stream

typedef int v4i __attribute__((ext_vector_type(4)));

__constant__ uint16_t c_cdf[11] = {5638, 15915, 23689, 28571, 31116, 32217, 32613, 32731, 32760, 32766, 32767};

// frodo_sample_n (noise.c:10-29) on one 16-bit value
__device__ __forceinline__ uint32_t sample(uint32_t r) {
  const uint32_t prnd = (r & 0xFFFFu) >> 1, sign = r & 1;
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < 10; ++j) s += ((c_cdf[j] - prnd) >> 15) & 1;
  return (((0u - sign) ^ s) + sign) & 0xFFFFu;
}

__device__ __forceinline__ uint32_t bswap16(uint32_t v) { return ((v & 0xFF) << 8) | ((v >> 8) & 0xFF); }

// ---- key expansion -----------------------------------------------------------

// generateMatA2 (frodo_macrify_reference.cu:164-188): row i =
// SHAKE128(le16(i) || seedA, 1952 bytes) read as little-endian uint16.
__global__ void k_gen_a(uint16_t* A, const uint8_t* seedA) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint64_t a0 = load_u64(seedA), a1 = load_u64(seedA + 8);
  uint64_t S[25];
  keccak_zero(S);
  S[0] = (uint64_t)i | (a0 << 16);
  S[1] = (a0 >> 48) | (a1 << 16);
  S[2] = (a1 >> 48) | (0x1Full << 16);
  S[kRateShake128 - 1] ^= 0x8000000000000000ull;
  uint64_t* row = reinterpret_cast<uint64_t*>(A + (size_t)i * N);
  for (int blk = 0; blk < 12; ++blk) {
    keccakf1600(S);
#pragma unroll
    for (int w = 0; w < kRateShake128; ++w) {
      const int idx = blk * kRateShake128 + w;  // 64-bit word = 4 columns
      if (idx < N / 4) row[idx] = S[w];
    }
  }
}

// Build the MFMA B-operand fragments of [A | B] as signed int8 digits.
// Fragment (d, nt, ks), lane l holds 16 bytes: digit d of [A|B][k][n] for
// n = 16 nt + (l & 15), k = 64 ks + 16 (l >> 4) + b, b = 0..15.
__global__ void k_prep_frags(int8_t* frag, const uint16_t* A, const uint8_t* pk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * NT * KSTEPS * 64) return;
  const int lane = t & 63, ks = (t >> 6) & (KSTEPS - 1), nt = (t >> 10) & (NT - 1), d = t >> 16;
  const int n = 16 * nt + (lane & 15);
  uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
  for (int b = 0; b < 16; ++b) {
    const int k = 64 * ks + 16 * (lane >> 4) + b;
    uint32_t v = 0;
    if (k < N) {
      if (n < N)
        v = A[(size_t)k * N + n];
      else if (n < N + NB)
        v = bswap16(*reinterpret_cast<const uint16_t*>(pk + SEEDA + 2 * (NB * k + (n - N))));
    }
    const int lo = (int)(int8_t)(v & 0xFF);
    const int hi = (int)(int8_t)((((int)v - lo) >> 8) & 0xFF);
    const uint32_t byte = (uint32_t)(d ? hi : lo) & 0xFF;
    w[b >> 2] |= byte << (8 * (b & 3));
  }
  uint4* o = reinterpret_cast<uint4*>(frag) + t;
  *o = make_uint4(w[0], w[1], w[2], w[3]);
}

// ---- decapsulation: W = B'S, mu' = decode(C - W) -------------------------------
// One wave per request; lane (i, j) computes W[i][j] (frodo_mul_bs_kernel), the
// 8 lanes of row i assemble 24 bits of mu' (frodo_key_decode_kernel).
__global__ __launch_bounds__(64) void k_w_decode(int n, const uint8_t* ct, const uint8_t* sk, uint8_t* mu) {
  const int r = blockIdx.x;
  if (r >= n) return;
  const int l = threadIdx.x, i = l >> 3, j = l & 7;
  const uint8_t* bp = ct + (size_t)r * CT + 2 * N * i;
  const uint8_t* st = sk + SK_S + 2 * N * j;
  uint32_t acc = 0;
  for (int k = 0; k < N; k += 4) {
    const uint64_t bw = load_u64(bp + 2 * k), sw = load_u64(st + 2 * k);
#pragma unroll
    for (int q = 0; q < 4; ++q) acc += bswap16((uint32_t)(bw >> (16 * q))) * ((uint32_t)(sw >> (16 * q)) & 0xFFFF);
  }
  const uint32_t c = bswap16(*reinterpret_cast<const uint16_t*>(ct + (size_t)r * CT + C1 + 2 * (8 * i + j)));
  const uint32_t m = (c - acc) & 0xFFFF;
  uint32_t word = ((m + (1u << 12)) >> 13 & 7) << (3 * j);
  word |= (uint32_t)__shfl_xor((int)word, 1);
  word |= (uint32_t)__shfl_xor((int)word, 2);
  word |= (uint32_t)__shfl_xor((int)word, 4);
  if (j == 0) {
    uint8_t* o = mu + (size_t)r * 32 + 3 * i;
    o[0] = (uint8_t)word;
    o[1] = (uint8_t)(word >> 8);
    o[2] = (uint8_t)(word >> 16);
  }
}

// ---- G2 + the S'/E'/E'' SHAKE256 stream (one lane per request) ---------------------
// Lanes [0, n): (seedSE || k) = SHAKE256(pkh || mu, 48); stream
// SHAKE256(0x96 || seedSE) sampled into S' (int8, MFMA A-operand fragments),
// E' ([8n][976] uint16) and E'' ([n][64] uint16).
// Decapsulation adds lanes [npad, npad + n): the first 115 SHAKE256 blocks of
// F(ct || key), which cover only ciphertext bytes, saved as a sponge state.
__global__ void k_shake(int n, int npad, int with_f, const uint8_t* pkh, const uint8_t* mu, int8_t* sfrag,
                        uint16_t* e1, uint16_t* e2, uint8_t* kout, const uint8_t* ct, uint64_t* fstate) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < npad) {
    const int r = t;
    if (r >= n) return;
    uint64_t S[25], g[6];
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      g[w] = load_u64(pkh + 8 * w);
      g[3 + w] = load_u64(mu + (size_t)r * 32 + 8 * w);
    }
    absorb_fixed<kRateShake256, 6>(S, g, 0x1F);
    const uint64_t s0 = S[0], s1 = S[1], s2 = S[2];
#pragma unroll
    for (int w = 0; w < 3; ++w) store_u64(kout + (size_t)r * 32 + 8 * w, S[3 + w]);
    keccak_zero(S);
    S[0] = 0x96 | (s0 << 8);
    S[1] = (s0 >> 56) | (s1 << 8);
    S[2] = (s1 >> 56) | (s2 << 8);
    S[3] = (s2 >> 56) | (0x1Full << 8);
    S[kRateShake256 - 1] ^= 0x8000000000000000ull;
    for (int blk = 0; blk * 68 < NSAMP; ++blk) {
      keccakf1600(S);
#pragma unroll
      for (int w = 0; w < kRateShake256; ++w) {
        const int g0 = blk * 68 + 4 * w;  // first sample of this group of 4
        if (g0 >= NSAMP) break;
        const uint64_t raw = S[w];
        uint32_t s4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) s4[q] = sample((uint32_t)(raw >> (16 * q)));
        if (g0 < N * NB) {
          const int k = g0 / N, j = g0 % N, m = 8 * r + k;
          const int lane = (m & 15) + 16 * ((j >> 4) & 3);
          const size_t off = ((((size_t)(m >> 4) * KSTEPS + (j >> 6)) * 64 + lane) * 16) + (j & 15);
          const uint32_t packed = (s4[0] & 0xFF) | ((s4[1] & 0xFF) << 8) | ((s4[2] & 0xFF) << 16) | ((s4[3] & 0xFF) << 24);
          *reinterpret_cast<uint32_t*>(sfrag + off) = packed;
        } else {
          const uint64_t v = (uint64_t)s4[0] | ((uint64_t)s4[1] << 16) | ((uint64_t)s4[2] << 32) | ((uint64_t)s4[3] << 48);
          if (g0 < 2 * N * NB) {
            const int g1 = g0 - N * NB, k = g1 / N, j = g1 % N;
            *reinterpret_cast<uint64_t*>(e1 + ((size_t)(8 * r + k) * N + j)) = v;
          } else {
            *reinterpret_cast<uint64_t*>(e2 + ((size_t)r * 64 + (g0 - 2 * N * NB))) = v;
          }
        }
      }
    }
  } else if (with_f) {
    const int r = t - npad;
    if (r >= n) return;
    uint64_t S[25];
    keccak_zero(S);
    const uint64_t* p = reinterpret_cast<const uint64_t*>(ct + (size_t)r * CT);
    for (int blk = 0; blk < (CT + SS) / 136; ++blk) {  // 115 full blocks of ciphertext bytes
#pragma unroll
      for (int w = 0; w < kRateShake256; ++w) S[w] ^= p[w];
      keccakf1600(S);
      p += kRateShake256;
    }
#pragma unroll
    for (int w = 0; w < 25; ++w) fstate[(size_t)r * 25 + w] = S[w];
  }
}

// ---- the shared-A GEMM on int8 MFMA ----------------------------------------------
// Workgroup = 4 waves: 256 rows (16 m-tiles, 64 per wave) x 64 columns (4 n-tiles),
// both digits.  Per 64-deep k-step the 8 B-operand fragments go through LDS
// (double-buffered, each wave loads 2), the A-operand fragments straight to
// registers.  Workgroups are ordered so that the 16 column blocks of one row
// block run on the same XCD and share its S' rows in that XCD's L2.
constexpr int GEMM_NB = NT / 4;  // 16 column blocks

__global__ __launch_bounds__(256) void k_gemm(int mtiles, int mblocks, int nrows, const v4i* __restrict__ sfrag,
                                              const v4i* __restrict__ bfrag, const uint16_t* __restrict__ e1,
                                              const uint16_t* __restrict__ e2, uint16_t* __restrict__ out) {
  __shared__ v4i lds[2][8][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // XCD-aware order: dispatch slot s on XCD x -> row block 8 (s / 16) + x, column block s % 16
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int mb = 8 * (slot / GEMM_NB) + xcd, nb = slot % GEMM_NB;
  if (mb >= mblocks) return;
  const int mt0 = mb * 16 + wave * 4;
  v4i acc[4][4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int d = 0; d < 2; ++d) acc[a][b][d] = v4i{0, 0, 0, 0};

  // this wave's two B fragments: f = 2 wave + q -> (ni = f >> 1, d = f & 1)
  auto bidx = [&](int f, int ks) {
    const int ni = f >> 1, d = f & 1;
    return (((size_t)(d * NT + nb * 4 + ni)) * KSTEPS + ks) * 64 + lane;
  };
  auto aidx = [&](int mi, int ks) { return (((size_t)(mt0 + mi)) * KSTEPS + ks) * 64 + lane; };
  const v4i zero = {0, 0, 0, 0};

  v4i bn[2], ac[4], an[4];
#pragma unroll
  for (int q = 0; q < 2; ++q) bn[q] = bfrag[bidx(2 * wave + q, 0)];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi) ac[mi] = (mt0 + mi < mtiles) ? sfrag[aidx(mi, 0)] : zero;
#pragma unroll
  for (int q = 0; q < 2; ++q) lds[0][2 * wave + q][lane] = bn[q];
  __syncthreads();

  for (int ks = 0; ks < KSTEPS; ++ks) {
    const int cur = ks & 1;
    if (ks + 1 < KSTEPS) {
#pragma unroll
      for (int q = 0; q < 2; ++q) bn[q] = bfrag[bidx(2 * wave + q, ks + 1)];
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) an[mi] = (mt0 + mi < mtiles) ? sfrag[aidx(mi, ks + 1)] : zero;
    }
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const v4i b = lds[cur][2 * ni + d][lane];
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) acc[mi][ni][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ac[mi], b, acc[mi][ni][d], 0, 0, 0);
      }
    }
    if (ks + 1 < KSTEPS) {
#pragma unroll
      for (int q = 0; q < 2; ++q) lds[cur ^ 1][2 * wave + q][lane] = bn[q];
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) ac[mi] = an[mi];
    }
    __syncthreads();
  }

  // epilogue: out = lo + 256 hi + E (mod 2^16); lane holds rows 4 (l >> 4) + r, column l & 15
#pragma unroll
  for (int mi = 0; mi < 4; ++mi) {
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const int col = 16 * (nb * 4 + ni) + (lane & 15);
      if (col >= N + NB) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * (mt0 + mi) + 4 * (lane >> 4) + r;
        if (row >= nrows) continue;
        uint32_t v = (uint32_t)acc[mi][ni][0][r] + ((uint32_t)acc[mi][ni][1][r] << 8);
        v += (col < N) ? e1[(size_t)row * N + col] : e2[(size_t)row * NB + (col - N)];
        out[(size_t)row * OUTC + col] = (uint16_t)v;
      }
    }
  }
}

// ---- decapsulation: compare, then F's last block --------------------------------

// key_encode of mu (frodo_key_encode_kernel) for coefficient (k, i)
__device__ __forceinline__ uint32_t encode_mu(const uint8_t* mu, int k, int i) {
  const uint32_t t = mu[3 * k] | ((uint32_t)mu[3 * k + 1] << 8) | ((uint32_t)mu[3 * k + 2] << 16);
  return ((t >> (3 * i)) & 7) << 13;
}

// One wave per request: fail = (B'' != B') or (V + encode(mu') != C), kem.c:191-201
// without the branch.
__global__ __launch_bounds__(64) void k_compare(int n, const uint8_t* ct, const uint16_t* out, const uint8_t* mu,
                                                uint32_t* fail) {
  const int r = blockIdx.x;
  if (r >= n) return;
  const int l = threadIdx.x;
  const uint8_t* c = ct + (size_t)r * CT;
  uint32_t mism = 0;
  for (int idx = l; idx < N * NB; idx += 64) {
    const int k = idx / N, j = idx % N;
    const uint32_t bpp = out[(size_t)(8 * r + k) * OUTC + j];
    mism |= bpp ^ bswap16(*reinterpret_cast<const uint16_t*>(c + 2 * idx));
  }
  {
    const int k = l >> 3, i = l & 7;
    const uint32_t v = out[(size_t)(8 * r + k) * OUTC + N + i];
    const uint32_t cc = (v + encode_mu(mu + (size_t)r * 32, k, i)) & 0xFFFF;
    mism |= cc ^ bswap16(*reinterpret_cast<const uint16_t*>(c + C1 + 2 * (8 * k + i)));
  }
  const uint32_t f = __any(mism != 0) ? 1u : 0u;
  if (l == 0) fail[r] = f;
}

// One lane per request: absorb ct[15640:15744] || key (k' or s by mask) into
// the saved sponge state, pad, permute: ss = F(ct || key) (kem.c:195-202).
__global__ void k_f_final(int n, uint8_t* ss, const uint8_t* ct, const uint64_t* fstate, const uint32_t* fail,
                          const uint8_t* kp, const uint8_t* s) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint64_t S[25];
#pragma unroll
  for (int w = 0; w < 25; ++w) S[w] = fstate[(size_t)r * 25 + w];
  const uint64_t* p = reinterpret_cast<const uint64_t*>(ct + (size_t)r * CT + 115 * 136);
#pragma unroll
  for (int w = 0; w < 13; ++w) S[w] ^= p[w];
  const uint64_t mask = 0 - (uint64_t)(fail[r] & 1);
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    const uint64_t a = load_u64(kp + (size_t)r * 32 + 8 * w), b = load_u64(s + 8 * w);
    S[13 + w] ^= (a & ~mask) | (b & mask);
  }
  S[16] ^= 0x1Full | 0x8000000000000000ull;
  keccakf1600(S);
  for (int w = 0; w < 3; ++w) store_u64(ss + (size_t)r * 24 + 8 * w, S[w]);
}

// ---- encapsulation: pack ct, then F(ct || k) --------------------------------------
__global__ __launch_bounds__(64) void k_enc_pack(int n, uint8_t* ct, const uint16_t* out, const uint8_t* mu) {
  const int r = blockIdx.x;
  if (r >= n) return;
  const int l = threadIdx.x;
  uint8_t* c = ct + (size_t)r * CT;
  for (int idx = l; idx < N * NB; idx += 64) {
    const int k = idx / N, j = idx % N;
    *reinterpret_cast<uint16_t*>(c + 2 * idx) = (uint16_t)bswap16(out[(size_t)(8 * r + k) * OUTC + j]);
  }
  const int k = l >> 3, i = l & 7;
  const uint32_t v = (out[(size_t)(8 * r + k) * OUTC + N + i] + encode_mu(mu + (size_t)r * MU, k, i)) & 0xFFFF;
  *reinterpret_cast<uint16_t*>(c + C1 + 2 * (8 * k + i)) = (uint16_t)bswap16(v);
}

__global__ void k_f_full(int n, uint8_t* ss, const uint8_t* ct, const uint8_t* kp) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint64_t S[25];
  keccak_zero(S);
  const uint64_t* p = reinterpret_cast<const uint64_t*>(ct + (size_t)r * CT);
  for (int blk = 0; blk < 115; ++blk) {
#pragma unroll
    for (int w = 0; w < kRateShake256; ++w) S[w] ^= p[w];
    keccakf1600(S);
    p += kRateShake256;
  }
#pragma unroll
  for (int w = 0; w < 13; ++w) S[w] ^= p[w];
#pragma unroll
  for (int w = 0; w < 3; ++w) S[13 + w] ^= load_u64(kp + (size_t)r * 32 + 8 * w);
  S[16] ^= 0x1Full | 0x8000000000000000ull;
  keccakf1600(S);
  for (int w = 0; w < 3; ++w) store_u64(ss + (size_t)r * 24 + 8 * w, S[w]);
}

// copy n mu values (24 bytes each) to 32-byte strided scratch
__global__ void k_stride_mu(int n, uint8_t* dst, const uint8_t* src) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * MU) return;
  dst[(size_t)(t / MU) * 32 + t % MU] = src[t];
}

// pkh = SHAKE256(pk, 24) for a loaded public key (lane 0)
__global__ void k_pkh(uint8_t* pkh, const uint8_t* pk) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t S[25];
  sponge_absorb_words<kRateShake256>(S, reinterpret_cast<const uint64_t*>(pk), PK / 8, 0x1F);
  for (int w = 0; w < 3; ++w) store_u64(pkh + 8 * w, S[w]);
}

// ---- key generation (kem.c:13-61), one key at a time ----------------------------------
// k_kp_stream: lane per key: seedA = SHAKE256(z, 16), S||E from SHAKE256(0x5F || seedSE).
__global__ void k_kp_stream(int n, const uint8_t* rnd, uint8_t* seedA, uint16_t* se) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint8_t* rr = rnd + (size_t)r * 64;
  uint64_t S[25], z[2] = {load_u64(rr + 48), load_u64(rr + 56)};
  absorb_fixed<kRateShake256, 2>(S, z, 0x1F);
  store_u64(seedA + (size_t)r * 16, S[0]);
  store_u64(seedA + (size_t)r * 16 + 8, S[1]);
  const uint64_t s0 = load_u64(rr + 24), s1 = load_u64(rr + 32), s2 = load_u64(rr + 40);
  keccak_zero(S);
  S[0] = 0x5F | (s0 << 8);
  S[1] = (s0 >> 56) | (s1 << 8);
  S[2] = (s1 >> 56) | (s2 << 8);
  S[3] = (s2 >> 56) | (0x1Full << 8);
  S[kRateShake256 - 1] ^= 0x8000000000000000ull;
  uint16_t* o = se + (size_t)r * 2 * N * NB;
  for (int blk = 0; blk * 68 < 2 * N * NB; ++blk) {
    keccakf1600(S);
#pragma unroll
    for (int w = 0; w < kRateShake256; ++w) {
      const int g0 = blk * 68 + 4 * w;
      if (g0 >= 2 * N * NB) break;
      const uint64_t raw = S[w];
      uint64_t v = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) v |= (uint64_t)sample((uint32_t)(raw >> (16 * q))) << (16 * q);
      *reinterpret_cast<uint64_t*>(o + g0) = v;
    }
  }
}

// B = A S + E (frodo_mul_add_as_plus_e) for key r, written packed into pk and the sk tail
__global__ void k_kp_finish(int r, uint8_t* pk, uint8_t* sk, const uint16_t* A, const uint16_t* se,
                            const uint8_t* seedA, const uint8_t* rnd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // (i, k) output of B
  const uint16_t* St = se + (size_t)r * 2 * N * NB;
  const uint16_t* E = St + N * NB;
  uint8_t* p = pk + (size_t)r * PK;
  uint8_t* s = sk + (size_t)r * SK;
  if (t < N * NB) {
    const int i = t / NB, k = t % NB;
    uint32_t acc = E[t];
    for (int j = 0; j < N; ++j) acc += (uint32_t)A[(size_t)i * N + j] * St[k * N + j];
    const uint16_t be = (uint16_t)bswap16(acc & 0xFFFF);
    *reinterpret_cast<uint16_t*>(p + SEEDA + 2 * t) = be;
    *reinterpret_cast<uint16_t*>(s + SS + SEEDA + 2 * t) = be;
    // S^T into the secret key as little-endian int16
    *reinterpret_cast<uint16_t*>(s + SK_S + 2 * t) = St[t];
  }
  if (t < SEEDA) {
    p[t] = seedA[(size_t)r * 16 + t];
    s[SS + t] = seedA[(size_t)r * 16 + t];
  }
  if (t < SS) s[t] = rnd[(size_t)r * 64 + t];
}

// sk tail: H(pk) for key r
__global__ void k_kp_pkh(int r, uint8_t* sk, const uint8_t* pk) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t S[25];
  sponge_absorb_words<kRateShake256>(S, reinterpret_cast<const uint64_t*>(pk + (size_t)r * PK), PK / 8, 0x1F);
  uint8_t* s = sk + (size_t)r * SK + SK_PKH;
  for (int w = 0; w < 3; ++w) store_u64(s + 8 * w, S[w]);
}

}  // namespace frodo

// ---- host orchestration -------------------------------------------------------------

using namespace frodo;

size_t frodo_bytes_per_elem() {
  // E' 15616, E'' 128, GEMM output 16384, mu 32, k 32, F state 200, fail 4
  // (key generation reuses E'..output for its 31232-byte S||E samples and mu
  // for seedA); the S' fragments live in their own 8 KB-per-element buffer
  return 15616 + 128 + 16384 + 32 + 32 + 200 + 4;
}

static int frodo_need(pqc_ctx* c) { return (c->fr_frag && c->fr_sfrag) ? PQC_OK : PQC_ERR_UNSUPPORTED; }

// Load a key: sk (is_sk = 1) or pk.  Expands A into the MFMA digit fragments.
int frodo_load_key(pqc_ctx* c, const uint8_t* key, int is_sk, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  if (is_sk) {
    if (hipMemcpyAsync(c->fr_sk, key, SK, hipMemcpyDeviceToDevice, s) != hipSuccess) return PQC_ERR_DEVICE;
    if (hipMemcpyAsync(c->fr_pk, key + SS, PK, hipMemcpyDeviceToDevice, s) != hipSuccess) return PQC_ERR_DEVICE;
    if (hipMemcpyAsync(c->fr_pkh, key + SK_PKH, SS, hipMemcpyDeviceToDevice, s) != hipSuccess) return PQC_ERR_DEVICE;
  } else {
    if (hipMemcpyAsync(c->fr_pk, key, PK, hipMemcpyDeviceToDevice, s) != hipSuccess) return PQC_ERR_DEVICE;
    hipLaunchKernelGGL(k_pkh, dim3(1), dim3(64), 0, s, c->fr_pkh, (const uint8_t*)c->fr_pk);
  }
  c->fr_has_sk = is_sk;
  hipLaunchKernelGGL(k_gen_a, dim3((N + 63) / 64), dim3(64), 0, s, c->fr_a16, (const uint8_t*)c->fr_pk);
  hipLaunchKernelGGL(k_prep_frags, dim3(2 * NT * KSTEPS * 64 / 256), dim3(256), 0, s, c->fr_frag,
                     (const uint16_t*)c->fr_a16, (const uint8_t*)c->fr_pk);
  c->fr_loaded = 1;
  return hip_err(hipGetLastError());
}

struct FrScratch {
  int8_t* sfrag;
  uint16_t *e1, *e2, *out;
  uint8_t *mu, *kp;
  uint64_t* fstate;
  uint32_t* fail;
};

static FrScratch fr_scratch(pqc_ctx* c) {
  const size_t ld = c->ld;
  uint8_t* p = c->fr_elem;
  FrScratch f;
  f.e1 = (uint16_t*)p;
  p += (size_t)8 * N * 2 * ld;
  f.e2 = (uint16_t*)p;
  p += (size_t)128 * ld;
  f.out = (uint16_t*)p;
  p += (size_t)8 * OUTC * 2 * ld;
  f.mu = p;
  p += (size_t)32 * ld;
  f.kp = p;
  p += (size_t)32 * ld;
  f.fstate = (uint64_t*)p;
  p += (size_t)200 * ld;
  f.fail = (uint32_t*)p;
  f.sfrag = c->fr_sfrag;
  return f;
}

static int launch_gemm(pqc_ctx* c, size_t n, const FrScratch& f, hipStream_t s) {
  const int mtiles = (int)((8 * n + 15) / 16);
  const int mblocks = (mtiles + 15) / 16;
  const int mblocks8 = (mblocks + 7) / 8 * 8;
  hipLaunchKernelGGL(k_gemm, dim3(mblocks8 * GEMM_NB), dim3(256), 0, s, mtiles, mblocks, (int)(8 * n),
                     reinterpret_cast<const v4i*>(f.sfrag), reinterpret_cast<const v4i*>(c->fr_frag),
                     (const uint16_t*)f.e1, (const uint16_t*)f.e2, f.out);
  return hip_err(hipGetLastError());
}

int frodo_dec(pqc_ctx* c, size_t n, uint8_t* ss, const uint8_t* ct, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  if (!c->fr_loaded || !c->fr_has_sk) return PQC_ERR_INVALID_ARG;
  if (n == 0) return PQC_OK;
  FrScratch f = fr_scratch(c);
  const int ni = (int)n, npad = (ni + 63) / 64 * 64;
  hipLaunchKernelGGL(k_w_decode, dim3(ni), dim3(64), 0, s, ni, ct, (const uint8_t*)c->fr_sk, f.mu);
  hipLaunchKernelGGL(k_shake, dim3((npad + ni + 63) / 64), dim3(64), 0, s, ni, npad, 1, (const uint8_t*)c->fr_pkh,
                     (const uint8_t*)f.mu, f.sfrag, f.e1, f.e2, f.kp, ct, f.fstate);
  rc = launch_gemm(c, n, f, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_compare, dim3(ni), dim3(64), 0, s, ni, ct, (const uint16_t*)f.out, (const uint8_t*)f.mu, f.fail);
  hipLaunchKernelGGL(k_f_final, dim3((ni + 63) / 64), dim3(64), 0, s, ni, ss, ct, (const uint64_t*)f.fstate,
                     (const uint32_t*)f.fail, (const uint8_t*)f.kp, (const uint8_t*)c->fr_sk);
  return hip_err(hipGetLastError());
}

int frodo_enc(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* mu, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  if (!c->fr_loaded) return PQC_ERR_INVALID_ARG;
  if (n == 0) return PQC_OK;
  FrScratch f = fr_scratch(c);
  const int ni = (int)n, npad = (ni + 63) / 64 * 64;
  hipLaunchKernelGGL(k_stride_mu, dim3((ni * MU + 255) / 256), dim3(256), 0, s, ni, f.mu, mu);
  hipLaunchKernelGGL(k_shake, dim3(npad / 64), dim3(64), 0, s, ni, npad, 0, (const uint8_t*)c->fr_pkh,
                     (const uint8_t*)f.mu, f.sfrag, f.e1, f.e2, f.kp, (const uint8_t*)nullptr, (uint64_t*)nullptr);
  rc = launch_gemm(c, n, f, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_enc_pack, dim3(ni), dim3(64), 0, s, ni, ct, (const uint16_t*)f.out, mu);
  hipLaunchKernelGGL(k_f_full, dim3((ni + 63) / 64), dim3(64), 0, s, ni, ss, (const uint8_t*)ct, (const uint8_t*)f.kp);
  return hip_err(hipGetLastError());
}

int frodo_keypair(pqc_ctx* c, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* rnd, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  if (n == 0) return PQC_OK;
  // per-key S||E samples and seeds go through the E' region of the scratch
  FrScratch f = fr_scratch(c);
  uint16_t* se = f.e1;
  uint8_t* seedA = f.mu;
  const int ni = (int)n;
  hipLaunchKernelGGL(k_kp_stream, dim3((ni + 63) / 64), dim3(64), 0, s, ni, rnd, seedA, se);
  for (int r = 0; r < ni; ++r) {
    hipLaunchKernelGGL(k_gen_a, dim3((N + 63) / 64), dim3(64), 0, s, c->fr_a16, (const uint8_t*)seedA + (size_t)r * 16);
    hipLaunchKernelGGL(k_kp_finish, dim3((N * NB + 255) / 256), dim3(256), 0, s, r, pk, sk, (const uint16_t*)c->fr_a16,
                       (const uint16_t*)se, (const uint8_t*)seedA, rnd);
    hipLaunchKernelGGL(k_kp_pkh, dim3(1), dim3(64), 0, s, r, sk, (const uint8_t*)pk);
  }
  c->fr_loaded = 0;  // the A buffer now holds the last generated key's A
  return hip_err(hipGetLastError());
}

int frodo_products(pqc_ctx* c, size_t n, uint16_t* bpp, uint16_t* v, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  FrScratch f = fr_scratch(c);
  if (bpp && hipMemcpy2DAsync(bpp, N * 2, f.out, OUTC * 2, N * 2, 8 * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return PQC_ERR_DEVICE;
  if (v && hipMemcpy2DAsync(v, NB * 2, f.out + N, OUTC * 2, NB * 2, 8 * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return PQC_ERR_DEVICE;
  return PQC_OK;
}

}  // namespace pqc
