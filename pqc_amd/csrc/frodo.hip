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
#include <algorithm>

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

// frodo_sample_n on the two 16-bit halves of raw at once, in packed 16-bit
// arithmetic (v_pk_sub_u16 / v_pk_lshrrev_b16 / v_pk_add_u16), the same
// (CDF[j] - prnd) >> 15 count and sign fold as the scalar form above.
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
constexpr uint16_t kCdf[10] = {5638, 15915, 23689, 28571, 31116, 32217, 32613, 32731, 32760, 32766};

__device__ __forceinline__ uint32_t sample2(uint32_t raw) {
  const u16x2 r = __builtin_bit_cast(u16x2, raw);
  const u16x2 one = {1, 1}, fifteen = {15, 15}, zero = {0, 0};
  const u16x2 sign = r & one;
  // hide prnd's range (< 2^15) from the compiler, which would otherwise turn the
  // sign-bit count into a compare + select per half
  uint32_t pr = __builtin_bit_cast(uint32_t, (u16x2)(r >> one));
  asm volatile("" : "+v"(pr));
  const u16x2 prnd = __builtin_bit_cast(u16x2, pr);
  u16x2 s = zero;
#pragma unroll
  for (int j = 0; j < 10; ++j) {
    const u16x2 c = {kCdf[j], kCdf[j]};
    s += (u16x2)(c - prnd) >> fifteen;
  }
  return __builtin_bit_cast(uint32_t, ((zero - sign) ^ s) + sign);
}

// the low bytes of two 16-bit samples, i.e. two int8 values in [-10, 10]
__device__ __forceinline__ uint32_t bytes2(uint32_t s2) { return __builtin_amdgcn_perm(0u, s2, 0x0C0C0200u); }

// Where each request's key lives: one key for the whole batch (strides 0), one
// per request (kid null, strides = element sizes), or a key store (kid[r] =
// the key index of request r).
struct FrKeys {
  const uint8_t* sk;  // secret keys (s || pk || S^T || pkh), decapsulation
  size_t sk_stride;
  const uint8_t* pkh;  // H(pk)
  size_t pkh_stride;
  const uint8_t* pk;  // public keys (seedA || B)
  size_t pk_stride;
  const uint16_t* A;  // expanded A, 976 x 976 row-major per key (many-key products)
  size_t a_stride;    // uint16 elements
  const uint32_t* kid;
  __device__ __forceinline__ size_t idx(int r) const { return kid ? (size_t)kid[r] : (size_t)r; }
  __device__ __forceinline__ const uint8_t* sk_of(int r) const { return sk + idx(r) * sk_stride; }
  __device__ __forceinline__ const uint8_t* pkh_of(int r) const { return pkh + idx(r) * pkh_stride; }
  __device__ __forceinline__ const uint8_t* pk_of(int r) const { return pk + idx(r) * pk_stride; }
  __device__ __forceinline__ const uint16_t* a_of(int r) const { return A + idx(r) * a_stride; }
};

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

// The same for n matrices (request r's seedA at seeds + r * seed_stride), one
// thread per (matrix, row); matrix r at A + r * 976 * 976.
__global__ void k_gen_a_many(int n, uint16_t* A, const uint8_t* seeds, size_t seed_stride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * N) return;
  const int r = t / N, i = t % N;
  const uint8_t* seedA = seeds + (size_t)r * seed_stride;
  uint64_t a0 = 0, a1 = 0;
  for (int b = 0; b < 8; ++b) {  // byte loads: seeds sit at any offset inside a key
    a0 |= (uint64_t)seedA[b] << (8 * b);
    a1 |= (uint64_t)seedA[8 + b] << (8 * b);
  }
  uint64_t S[25];
  keccak_zero(S);
  S[0] = (uint64_t)i | (a0 << 16);
  S[1] = (a0 >> 48) | (a1 << 16);
  S[2] = (a1 >> 48) | (0x1Full << 16);
  S[kRateShake128 - 1] ^= 0x8000000000000000ull;
  uint64_t* row = reinterpret_cast<uint64_t*>(A + ((size_t)r * N + i) * N);
  for (int blk = 0; blk < 12; ++blk) {
    keccakf1600(S);
#pragma unroll
    for (int w = 0; w < kRateShake128; ++w) {
      const int idx = blk * kRateShake128 + w;
      if (idx < N / 4) row[idx] = S[w];
    }
  }
}

// k of byte b of lane l's 16-byte fragment in k-step ks, or -1 for a byte no MFMA
// reads.  Steps 0..14 are 64 deep (v_mfma_i32_16x16x64_i8: lane quarter l >> 4
// holds k 16 q .. 16 q + 15); the last step holds the 16-column tail k = 960 .. 975
// (976 = 15 x 64 + 16) for v_mfma_i32_16x16x32_i8 (quarter q holds k 960 + 8 q ..
// +7 in bytes 0..7; quarters 2 and 3 are k >= 976, zero).
__device__ __forceinline__ int frag_k(int ks, int lane, int b) {
  if (ks < KSTEPS - 1) return 64 * ks + 16 * (lane >> 4) + b;
  return b < 8 ? 64 * ks + 8 * (lane >> 4) + b : -1;
}

// Build the MFMA B-operand fragments of [A | B] as signed int8 digits.
// Fragment (d, nt, ks), lane l holds 16 bytes: digit d of [A|B][k][n] for
// n = 16 nt + (l & 15), k = frag_k(ks, l, b), b = 0..15.
__global__ void k_prep_frags(int8_t* frag, const uint16_t* A, const uint8_t* pk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * NT * KSTEPS * 64) return;
  const int lane = t & 63, ks = (t >> 6) & (KSTEPS - 1), nt = (t >> 10) & (NT - 1), d = t >> 16;
  const int n = 16 * nt + (lane & 15);
  uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
  for (int b = 0; b < 16; ++b) {
    const int k = frag_k(ks, lane, b);
    uint32_t v = 0;
    if (k >= 0 && k < N) {
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
// frodo_mul_bs_kernel + frodo_key_decode_kernel, one wave per request.  The
// request's S^T (8 rows of 976 int16) is staged in LDS as k-pairs: entry kp holds
// (S[2kp][j], S[2kp + 1][j]) for j = 0..7, padded to kWdStride words so that the
// 8 entries a wave reads at once fall in different banks.  Lane (i, c) takes the
// 16-byte chunks c, c + 8, c + 16, ... of B' row i (the 8 lanes of a row read
// 128 contiguous bytes per load) and keeps W[i][0..7] as packed 16-bit pairs
// (even and odd k): only W mod 2^16 is used, so the 16-bit wrap is exact.  The 8
// lanes of a row then sum their partial rows, and lane (i, j) decodes W[i][j].
// A workgroup takes requests r, r + G, ...; under one key for all requests the
// S^T staging happens once per workgroup, else once per request.
constexpr int kWdChunks = N / 8;  // 16-byte chunks (8 coefficients, 4 k-pairs) per row
constexpr int kWdStride = 10;     // words per k-pair entry: 8 + 2 pad
static_assert(N % 8 == 0, "whole chunks");

__device__ __forceinline__ void w_stage(uint32_t* spair, const uint8_t* sk, int l) {
  // S^T row j, chunk q: k-pairs 4q .. 4q + 3 (S^T sits 8-byte aligned in sk)
  const uint8_t* st = sk + SK_S;
  for (int t = l; t < NB * kWdChunks; t += 64) {
    const int j = t / kWdChunks, q = t - j * kWdChunks;
    const uint8_t* src = st + 2 * N * j + 16 * q;
    const uint2 lo = *reinterpret_cast<const uint2*>(src), hi = *reinterpret_cast<const uint2*>(src + 8);
    uint32_t* e = spair + 4 * q * kWdStride + j;
    e[0] = lo.x;
    e[kWdStride] = lo.y;
    e[2 * kWdStride] = hi.x;
    e[3 * kWdStride] = hi.y;
  }
}

__device__ __forceinline__ void w_decode_one(const uint32_t* spair, int r, const uint8_t* ct, uint8_t* mu, int l) {
  const int i = l >> 3, c = l & 7;
  typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
  u16x2 acc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) acc[j] = u16x2{0, 0};
  const uint8_t* bp = ct + (size_t)r * CT + 2 * N * i;
  for (int q = c; q < kWdChunks; q += 8) {
    const uint4 b = *reinterpret_cast<const uint4*>(bp + 16 * q);
    const uint32_t bw[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      // two big-endian B' coefficients -> two native uint16
      const u16x2 b2 = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(0u, bw[p], 0x02030001u));
      const uint32_t* e = spair + (4 * q + p) * kWdStride;
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[j] += b2 * __builtin_bit_cast(u16x2, e[j]);
    }
  }
  uint32_t w[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    w[j] = (uint32_t)acc[j].x + (uint32_t)acc[j].y;
#pragma unroll
    for (int sft = 1; sft < 8; sft <<= 1) w[j] += (uint32_t)__shfl_xor((int)w[j], sft);
  }
  // lane (i, j): W[i][j]
  uint32_t wij = w[0];
#pragma unroll
  for (int j = 1; j < NB; ++j) wij = c == j ? w[j] : wij;
  const int j = c;
  const uint32_t cv = bswap16(*reinterpret_cast<const uint16_t*>(ct + (size_t)r * CT + C1 + 2 * (8 * i + j)));
  const uint32_t m = (cv - wij) & 0xFFFF;
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

__global__ __launch_bounds__(64) void k_w_decode(int n, const uint8_t* ct, FrKeys keys, uint8_t* mu) {
  __shared__ uint32_t spair[(N / 2) * kWdStride];
  const int l = threadIdx.x;
  const bool one_key = keys.sk_stride == 0 && keys.kid == nullptr;  // uniform
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    if (!one_key || r == (int)blockIdx.x) {
      __syncthreads();  // the previous request's LDS reads are done
      w_stage(spair, keys.sk_of(r), l);
      __syncthreads();
    }
    w_decode_one(spair, r, ct, mu, l);
  }
}

static inline int w_decode_grid(const pqc_ctx* c, int n) { return std::max(1, std::min(n, c->cu_count * 8)); }

// ---- G2 + the S'/E'/E'' SHAKE256 stream (one lane pair per request) ----------------
// Pairs [0, npad): (seedSE || k) = SHAKE256(pkh || mu, 48), then the stream
// SHAKE256(0x96 || seedSE, 31360 B) sampled into S' (int8 MFMA operand
// fragments) and E' | E'' (int8, in the GEMM output order: see eoff) (kem.c:87-98,153-166).
// Each lane of a pair samples the two 16-bit values of its half of every word.
// Decapsulation adds pairs [npad, npad + n): the first 115 SHAKE256 blocks of
// F(ct || key), which hold only ciphertext bytes, saved as a sponge state.
__device__ __forceinline__ void absorb_ct_pair(uint32_t (&S)[25], const uint8_t* c, int half) {
#pragma unroll
  for (int w = 0; w < 25; ++w) S[w] = 0;
  const uint32_t* p = reinterpret_cast<const uint32_t*>(c) + half;
  for (int blk = 0; blk < (CT + SS) / 136; ++blk) {  // 115 full blocks
#pragma unroll
    for (int w = 0; w < kRateShake256; ++w) S[w] ^= p[2 * w];
    keccakf1600_pair(S, (uint32_t)half);
    p += 2 * kRateShake256;
  }
}

// Byte offset of [E' | E''] entry (m, j) (row m of the 8n stacked rows, column j
// < 984, E'' at j = 976 + i) in the layout the GEMM epilogue reads: per 16-row
// m-tile and 64-column block, 64 lanes x 16 bytes; lane (m & 15) + 16 ((j >> 2) & 3)
// holds columns 16 ni + 4 (lane >> 4) + r at byte 4 ni + r -- the four n-tiles of
// the block, four consecutive columns each, as the swapped-operand MFMA output has them.
__device__ __forceinline__ size_t eoff(int m, int j) {
  const int lane = (m & 15) + 16 * ((j >> 2) & 3);
  return ((((size_t)(m >> 4) * 16 + (j >> 6)) * 64 + lane) * 4 + ((j >> 4) & 3)) * 4 + (j & 3);
}

// The stream is written raw (RAWB bytes per request: each 8-byte SHAKE256
// output word as its even-bit then its odd-bit 32-bit word) and sampled by
// k_sample: the serial sponge chain of a lane pair is then one permutation plus
// 17 dword stores per block, with no sampling or index arithmetic on it.
constexpr int RAWB = 31424;  // 231 blocks x 136 bytes, rounded up to 16
static_assert(231 * 136 <= RAWB && 231 * 68 >= NSAMP && RAWB % 16 == 0, "raw stream size");

__global__ __launch_bounds__(256) void k_shake(int n, int npad, int with_f, FrKeys keys, const uint8_t* mu,
                                              int mu_stride, uint8_t* raw, uint8_t* kout, const uint8_t* ct,
                                              uint32_t* fstate) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int pr = t >> 1, half = t & 1;
  uint32_t S[25];
  if (pr < npad) {
    const int r = pr;
    if (r >= n) return;
#pragma unroll
    for (int w = 0; w < 25; ++w) S[w] = 0;
    const uint32_t* ph = reinterpret_cast<const uint32_t*>(keys.pkh_of(r));
    const uint32_t* pm = reinterpret_cast<const uint32_t*>(mu + (size_t)r * mu_stride);
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      S[w] = ph[2 * w + half];
      S[3 + w] = pm[2 * w + half];
    }
    S[6] ^= half ? 0u : 0x1Fu;
    S[kRateShake256 - 1] ^= half ? 0x80000000u : 0u;
    keccakf1600_pair(S, (uint32_t)half);
    uint32_t* pk = reinterpret_cast<uint32_t*>(kout + (size_t)r * 32);
#pragma unroll
    for (int w = 0; w < 3; ++w) pk[2 * w + half] = S[3 + w];
    // seedSE as whole 64-bit words in both lanes, shifted behind the 0x96 byte
    uint64_t sd[3];
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      const uint32_t other = pair_swap(S[w]);
      sd[w] = half ? ((uint64_t)other | ((uint64_t)S[w] << 32)) : ((uint64_t)S[w] | ((uint64_t)other << 32));
    }
    const uint64_t init[4] = {0x96 | (sd[0] << 8), (sd[0] >> 56) | (sd[1] << 8), (sd[1] >> 56) | (sd[2] << 8),
                              (sd[2] >> 56) | (0x1Full << 8)};
    // the stream's state bit-interleaved (keccakf1600_pair_il): this lane holds
    // the even (half 0) or odd (half 1) bits of every 64-bit lane
#pragma unroll
    for (int w = 0; w < 25; ++w) S[w] = w < 4 ? il_word(init[w < 4 ? w : 0], (uint32_t)half) : 0u;
    S[kRateShake256 - 1] ^= half ? 0x80000000u : 0u;  // pad bit 63 = odd bit 31
    // lane `half` of stream word w of block b (bytes 136 b + 8 w .. + 7) goes to
    // 136 b + 8 w + 4 half: the even-bit word, then the odd-bit word
    uint32_t* out = reinterpret_cast<uint32_t*>(raw + (size_t)r * RAWB) + half;
    for (int blk = 0; blk < 231; ++blk) {
      keccakf1600_pair_il<24>(S, (uint32_t)half);
#pragma unroll
      for (int w = 0; w < kRateShake256; ++w) out[2 * w] = S[w];
      out += 2 * kRateShake256;
    }
  } else if (with_f) {
    const int r = pr - npad;
    if (r >= n) return;
    absorb_ct_pair(S, ct + (size_t)r * CT, half);
#pragma unroll
    for (int w = 0; w < 25; ++w) fstate[(size_t)r * 50 + 2 * w + half] = S[w];
  }
}

// The same G2, stream and F prefix with one lane per request (the throughput form):
// a lane pair finishes a request's chain sooner, but costs 2 x 107 against 180 VALU
// instructions per round, so calls that fill every SIMD several waves deep run
// faster on single lanes.  The stream is written as plain little-endian words.
// Lanes [0, nlane) are requests; [nlane, nlane + n) the F prefixes (nlane is a
// multiple of 64, so no wave mixes the two).
__global__ __launch_bounds__(256) void k_shake_lane(int n, int nlane, int with_f, FrKeys keys, const uint8_t* mu,
                                                   int mu_stride, uint8_t* raw, uint8_t* kout, const uint8_t* ct,
                                                   uint64_t* fstate) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t S[25];
#pragma unroll
  for (int w = 0; w < 25; ++w) S[w] = 0;
  if (t < nlane) {
    const int r = t;
    if (r >= n) return;
    const uint8_t* ph = keys.pkh_of(r);
    const uint8_t* pm = mu + (size_t)r * mu_stride;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      S[w] = load_u64(ph + 8 * w);
      S[3 + w] = load_u64(pm + 8 * w);
    }
    S[6] ^= 0x1Full;
    S[kRateShake256 - 1] ^= 0x8000000000000000ull;
    keccakf1600(S);
#pragma unroll
    for (int w = 0; w < 3; ++w) store_u64(kout + (size_t)r * 32 + 8 * w, S[3 + w]);
    // 0x96 || seedSE, padded
    const uint64_t sd0 = S[0], sd1 = S[1], sd2 = S[2];
#pragma unroll
    for (int w = 0; w < 25; ++w) S[w] = 0;
    S[0] = 0x96 | (sd0 << 8);
    S[1] = (sd0 >> 56) | (sd1 << 8);
    S[2] = (sd1 >> 56) | (sd2 << 8);
    S[3] = (sd2 >> 56) | (0x1Full << 8);
    S[kRateShake256 - 1] ^= 0x8000000000000000ull;
    uint64_t* out = reinterpret_cast<uint64_t*>(raw + (size_t)r * RAWB);
    for (int blk = 0; blk < 231; ++blk) {
      keccakf1600(S);
#pragma unroll
      for (int w = 0; w < kRateShake256; ++w) out[w] = S[w];
      out += kRateShake256;
    }
  } else if (with_f) {
    const int r = t - nlane;
    if (r >= n) return;
    const uint64_t* p = reinterpret_cast<const uint64_t*>(ct + (size_t)r * CT);
    for (int blk = 0; blk < (CT + SS) / 136; ++blk) {  // 115 full blocks
#pragma unroll
      for (int w = 0; w < kRateShake256; ++w) S[w] ^= p[w];
      keccakf1600(S);
      p += kRateShake256;
    }
#pragma unroll
    for (int w = 0; w < 25; ++w) fstate[(size_t)r * 25 + w] = S[w];
  }
}

// CDF sampling of the raw stream (frodo_sample_n, frodo976.cu:693-715; kem.c:94-98):
// one thread per 8 consecutive samples (16 stream bytes).  Sample g of request r
// is S' row g / 976, column g % 976 for g < 7808, then E' likewise, then E''
// (8 x 8).  976 is a multiple of 8, so the 8 samples of a thread share a row.
// ROWS = false: S' into the MFMA fragments, E' | E'' at eoff (shared-A GEMM).
// ROWS = true: S' and E' as plain int8 rows [8n][976] in sfrag / efrag, E'' after
// E' as [n][64] (the many-key products).
constexpr int kSampleThreads = NSAMP / 8;  // per request

// bits 0..15 of the 16-bit half `h` of v spread to the even bit positions
__device__ __forceinline__ uint32_t il_spread(uint32_t v, int h) {
  v = __builtin_amdgcn_perm(0u, v, h ? 0x0C030C02u : 0x0C010C00u);  // byte b of the half -> byte 2 b
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  return (v | (v << 1)) & 0x55555555u;
}

// 32-bit half h of the 64-bit word whose even bits are e and odd bits are o
__device__ __forceinline__ uint32_t il_join(uint32_t e, uint32_t o, int h) {
  return il_spread(e, h) | (il_spread(o, h) << 1);
}
static_assert(NSAMP % 8 == 0 && (N * NB) % 8 == 0, "sample groups");

// IL: the stream as k_shake writes it (each word as its even-bit then odd-bit
// halves); otherwise plain little-endian words (k_shake_lane).
template <bool ROWS, bool IL>
__global__ __launch_bounds__(256) void k_sample(int n, const uint8_t* raw, int8_t* sfrag, int8_t* efrag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * kSampleThreads) return;
  const int r = t / kSampleThreads, g = 8 * (t - r * kSampleThreads);
  // two stream words
  const uint4 v = *reinterpret_cast<const uint4*>(raw + (size_t)r * RAWB + 2 * g);
  uint32_t w0 = v.x, w1 = v.y, w2 = v.z, w3 = v.w;
  if constexpr (IL) {
    w0 = il_join(v.x, v.y, 0);
    w1 = il_join(v.x, v.y, 1);
    w2 = il_join(v.z, v.w, 0);
    w3 = il_join(v.z, v.w, 1);
  }
  const uint32_t lo = bytes2(sample2(w0)) | (bytes2(sample2(w1)) << 16);
  const uint32_t hi = bytes2(sample2(w2)) | (bytes2(sample2(w3)) << 16);
  if constexpr (ROWS) {
    int8_t* dst = g < N * NB       ? sfrag + (size_t)r * N * NB + g
                  : g < 2 * N * NB ? efrag + (size_t)r * N * NB + (g - N * NB)
                                   : efrag + (size_t)n * N * NB + (size_t)r * 64 + (g - 2 * N * NB);
    *reinterpret_cast<uint2*>(dst) = make_uint2(lo, hi);
  } else if (g < N * NB) {
    // column j of S' row m: k-step j >> 6, placed as frag_k lays the step out
    const int k = g / N, j = g % N, m = 8 * r + k;
    const int tail = j >= 64 * (KSTEPS - 1);
    const int lane = (m & 15) + 16 * (tail ? ((j >> 3) & 1) : ((j >> 4) & 3));
    const size_t off = ((((size_t)(m >> 4) * KSTEPS + (j >> 6)) * 64 + lane) * 16) + (tail ? 0 : (j & 15));
    *reinterpret_cast<uint2*>(sfrag + off) = make_uint2(lo, hi);
  } else {
    int m, j;  // the group's first entry; entries j .. j+3 and j+4 .. j+7 each sit in one lane's word
    if (g < 2 * N * NB) {
      const int g1 = g - N * NB;
      m = 8 * r + g1 / N;
      j = g1 % N;
    } else {
      const int g2 = g - 2 * N * NB;
      m = 8 * r + (g2 >> 3);
      j = N + (g2 & 7);
    }
    *reinterpret_cast<uint32_t*>(efrag + eoff(m, j)) = lo;
    *reinterpret_cast<uint32_t*>(efrag + eoff(m, j + 4)) = hi;
  }
}

// key_encode of mu (frodo_key_encode_kernel) for coefficient (k, i)
__device__ __forceinline__ uint32_t encode_mu(const uint8_t* mu, int k, int i) {
  const uint32_t t = mu[3 * k] | ((uint32_t)mu[3 * k + 1] << 8) | ((uint32_t)mu[3 * k + 2] << 16);
  return ((t >> (3 * i)) & 7) << 13;
}

// ---- the shared-A GEMM on int8 MFMA ----------------------------------------------
// Workgroup = 4 waves, tile = 256 rows (16 m-tiles, 4 per wave) x 64 columns
// (4 n-tiles) x 2 digits.  Per 64-deep k-step, the 8 [A | B] fragments, which
// all four waves read, go global -> LDS by global_load_lds_dwordx4 (2 per wave)
// into a 3-slot ring; a wave's own 4 S' fragments, which no other wave reads, go
// straight into a 3-slot register ring by plain 16-byte loads.  Both are issued
// two steps ahead of the MFMAs.  (S' through LDS cost a DMA write and an LDS
// read of 16 KB per workgroup and step for data no other wave uses: 3-5% of the
// launch.)  Every LDS slot is its own __shared__ object so the compiler's LDS-DMA
// tracking never makes an LDS read wait for an unrelated slot.  Per step: this
// wave's loads of the step done (explicit vmcnt), LDS reads of the previous step
// returned (lgkmcnt 0), one s_barrier.
//
// The MFMA takes [A | B] as its first operand, so the output is the transposed
// tile: lane l holds row 16 mt + (l & 15), columns 16 nt + 4 (l >> 4) + r,
// i.e. four consecutive uint16 of a ciphertext row -- the epilogue reads and
// writes 8-byte words.  [E' | E''] (16 bytes per m-tile) is the lo digit's
// starting value, loaded first.  Decapsulation's ciphertext words are loaded
// during steps 12..15, one m-tile per step, when the S' register ring drains
// (earlier, the 32 VGPRs they hold spill); issued at once, every workgroup's
// burst of HBM misses queued ahead of its next loads and stalled the step after.
//
// Epilogue per mode:
//   products: [B'' | V] (uint16) to the output buffer;
//   encaps:   B'' big-endian into ct, C = V + encode(mu) likewise
//             (frodo_pack, util.cu:18-64);
//   decaps:   compare with ct's B' and C (kem.c:191-201) and store one flag
//             byte per (request, column block) -- no atomics, no branch on data.
//
// Workgroups are ordered so that the 16 column blocks of one row block run on
// one XCD and share its S' rows in that XCD's L2.  S' and E' are allocated in
// whole 256-row blocks, so no load needs a bounds check; rows past 8n are never
// stored.  N-tiles 62 and 63 (columns 992..1023) hold only padding: skipped.
constexpr int GEMM_NB = NT / 4;  // 16 column blocks
constexpr int GEMM_NTILES = (N + NB + 15) / 16;  // 62 n-tiles carry data
enum { kProducts = 0, kEncaps = 1, kDecaps = 2 };

__shared__ v4i g_sb0[8][64], g_sb1[8][64], g_sb2[8][64];  // [A | B] fragments of one k-step

template <int ST>
__device__ __forceinline__ v4i (*ring_b())[64] {
  if constexpr (ST == 0) return g_sb0;
  else if constexpr (ST == 1) return g_sb1;
  else return g_sb2;
}

struct GemmArgs {
  const v4i* sfrag;
  const v4i* bfrag;
  int mt0, nb, wave, lane;
};

// this wave's S' fragments (4 m-tiles) of three k-steps
struct SRing {
  v4i a[3][4];
};

template <int KS>
__device__ __forceinline__ void gemm_issue(const GemmArgs& g, SRing& sr) {
#if defined(__HIP_DEVICE_COMPILE__)  // the LDS-DMA builtin exists only in the device pass
  constexpr int st = KS % 3;
  // a wave-uniform fragment address plus the lane's 32-bit offset (the SGPR-base form)
  const uint32_t lo = (uint32_t)g.lane * 16u;
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
    sr.a[st][mi] = *reinterpret_cast<const v4i*>(
        reinterpret_cast<const uint8_t*>(g.sfrag + ((size_t)(g.mt0 + mi) * KSTEPS + KS) * 64) + lo);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int f = 2 * g.wave + q, ni = f >> 1, d = f & 1;
    __builtin_amdgcn_global_load_lds(
        reinterpret_cast<const uint8_t*>(g.bfrag + ((size_t)(d * NT + g.nb * 4 + ni) * KSTEPS + KS) * 64) + lo,
        &ring_b<st>()[f][0], 16, 0, 0);
  }
#endif
}

// s_waitcnt immediates: vmcnt(v), expcnt(7), lgkmcnt(0)
template <int V>
__device__ __forceinline__ void wait_vm_lgkm0() {
  static_assert(V < 64, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((V & 15) | ((V >> 4) << 14) | 0x0070);
}

constexpr int LOADS_PER_STEP = 6;  // loads issued per wave per k-step: 4 S' (registers) + 2 [A | B] (LDS DMA)

template <int MODE>
struct EpiRegs {
  uint2 c[MODE == kDecaps ? 4 : 1][4];     // ciphertext words (decapsulation)
  static constexpr int kLoads = MODE == kDecaps ? 16 : 0;
};

template <int MODE>
__device__ __forceinline__ uint8_t* ct_word(uint8_t* ct, int row, int n0) {
  const int req = row >> 3, k = row & 7;
  uint8_t* c = ct + (size_t)req * CT;
  if (n0 < N) return c + 2 * (k * N + n0);
  if (n0 < N + NB) return c + C1 + 2 * (8 * k + (n0 - N));
  return c;  // padding columns: a valid address, never compared or stored
}

// the ciphertext words of m-tile MI (decapsulation): EpiRegs::kLoads / 4 loads
template <int MODE, int MI>
__device__ __forceinline__ void epi_issue(EpiRegs<MODE>& ep, const GemmArgs& g, uint8_t* ct, int nrows) {
  if constexpr (MODE == kDecaps) {
    const int row = min(16 * (g.mt0 + MI) + (g.lane & 15), nrows - 1);
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const int n0 = 16 * (4 * g.nb + ni) + 4 * (g.lane >> 4);
      ep.c[MI][ni] = *reinterpret_cast<const uint2*>(ct_word<MODE>(ct, row, n0));
    }
  }
}

// the steps that issue the ciphertext words (m-tile ks - kEpiStep0 at step ks)
constexpr int kEpiStep0 = 12;
static_assert(kEpiStep0 + 4 <= KSTEPS, "every m-tile's words are loaded inside the k loop");
template <int MODE>
constexpr int epi_loads_at(int ks) {
  return ks >= kEpiStep0 && ks < kEpiStep0 + 4 ? EpiRegs<MODE>::kLoads / 4 : 0;
}

template <int MODE, int KS>
__device__ __forceinline__ void gemm_step(v4i (&acc)[4][4][2], const GemmArgs& g, EpiRegs<MODE>& ep, uint8_t* ct,
                                          int nrows, SRing& sr) {
  // this wave's loads of step KS done; what may stay in flight is what was issued
  // after them, at steps KS - 2 and KS - 1: the loads of step KS + 1 and the
  // ciphertext words of those steps.  (Within a step the S' loads and the DMAs
  // may issue in either order; the s_barrier asm below keeps each step's loads
  // in its own step.)
  constexpr int allowed =
      (KS + 1 < KSTEPS ? LOADS_PER_STEP : 0) + epi_loads_at<MODE>(KS - 2) + epi_loads_at<MODE>(KS - 1);
  wait_vm_lgkm0<allowed>();
  asm volatile("s_barrier" ::: "memory");  // a workgroup barrier no memory access moves across
  if constexpr (KS + 2 < KSTEPS) gemm_issue<KS + 2>(g, sr);
  if constexpr (KS >= kEpiStep0 && KS < kEpiStep0 + 4) {
    // the vmcnt values above count these loads after this step's DMAs: keep the
    // scheduler from moving them across
    __builtin_amdgcn_sched_barrier(0);
    epi_issue<MODE, KS - kEpiStep0>(ep, g, ct, nrows);
    __builtin_amdgcn_sched_barrier(0);
  }
  constexpr int st = KS % 3;
  if constexpr (KS + 1 < KSTEPS) {
    v4i a[4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) a[mi] = sr.a[st][mi];
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      if (4 * g.nb + ni >= GEMM_NTILES) continue;  // uniform: padding n-tiles
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const v4i b = ring_b<st>()[2 * ni + d][g.lane];
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
          acc[mi][ni][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(b, a[mi], acc[mi][ni][d], 0, 0, 0);
      }
    }
  } else {
    // the 16-deep k tail (frag_k): 32-deep MFMAs on the first 8 bytes of each lane
    long a[4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) a[mi] = (long)(uint32_t)sr.a[st][mi][0] | ((long)sr.a[st][mi][1] << 32);
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      if (4 * g.nb + ni >= GEMM_NTILES) continue;
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const long b = *reinterpret_cast<const long*>(&ring_b<st>()[2 * ni + d][g.lane]);
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
          acc[mi][ni][d] = __builtin_amdgcn_mfma_i32_16x16x32_i8(b, a[mi], acc[mi][ni][d], 0, 0, 0);
      }
    }
  }
  if constexpr (KS + 1 < KSTEPS) gemm_step<MODE, KS + 1>(acc, g, ep, ct, nrows, sr);
}

__device__ __forceinline__ uint32_t bswap16x2(uint32_t v) { return __builtin_amdgcn_perm(0u, v, 0x02030001u); }
__device__ __forceinline__ uint32_t sx8(uint32_t w, int i) { return (uint32_t)(((int32_t)(w << (24 - 8 * i))) >> 24); }
__device__ __forceinline__ uint32_t word_of(const uint4& v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }

// one flag byte per (request, 64-column block): lanes with (l & 15) < 8 hold
// request 2 mt, the others 2 mt + 1
__device__ __forceinline__ void gemm_flags(uint32_t mism, const GemmArgs& g, int mi, int n, uint8_t* flags) {
  const uint64_t bal = __ballot(mism != 0);
  const int req0 = 2 * (g.mt0 + mi);
  if (g.lane == 0) {
    if (req0 < n) flags[(size_t)req0 * GEMM_NB + g.nb] = (uint8_t)((bal & 0x00FF00FF00FF00FFull) != 0);
    if (req0 + 1 < n) flags[(size_t)(req0 + 1) * GEMM_NB + g.nb] = (uint8_t)((bal & 0xFF00FF00FF00FF00ull) != 0);
  }
}

template <int MODE>
__global__ __launch_bounds__(256, 2) void k_gemm(int mblocks, int n, const v4i* __restrict__ sfrag,
                                                 const v4i* __restrict__ bfrag, const int8_t* __restrict__ efrag,
                                                 uint16_t* __restrict__ out, uint8_t* __restrict__ ct,
                                                 const uint8_t* __restrict__ mu, int mu_stride,
                                                 uint8_t* __restrict__ flags) {
  // wave-uniform in an SGPR, so every index derived from it stays scalar
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // XCD-aware order: dispatch slot s on XCD x -> row block 8 (s / 16) + x, column block s % 16
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int mb = 8 * (slot / GEMM_NB) + xcd, nb = slot % GEMM_NB;
  if (mb >= mblocks) return;
  const GemmArgs g{sfrag, bfrag, mb * 16 + wave * 4, nb, wave, lane};
  const int nrows = 8 * n;
  // [E' | E''] of this wave's tile (16 bytes per m-tile, in the output's lane
  // order) is the lo digit's starting value: the epilogue then only combines digits
  uint4 e[4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
    e[mi] = *reinterpret_cast<const uint4*>(efrag + (((size_t)(g.mt0 + mi) * 16 + nb) * 64 + lane) * 16);
  SRing sr;
  gemm_issue<0>(g, sr);
  // the explicit vmcnt values count the loads in step order: step 1's plain loads
  // must not move above step 0's LDS DMAs
  asm volatile("" ::: "memory");
  gemm_issue<1>(g, sr);
  v4i acc[4][4][2];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        acc[mi][ni][0][r] = (int)sx8(word_of(e[mi], ni), r);
        acc[mi][ni][1][r] = 0;
      }
  EpiRegs<MODE> ep;
  gemm_step<MODE, 0>(acc, g, ep, ct, nrows, sr);

  // epilogue: value = acc_lo + 256 acc_hi (mod 2^16) (acc_lo started at E); lane
  // holds row 16 mt + (l & 15), columns n0 .. n0 + 3 with n0 = 16 nt + 4 (l >> 4)
  if (16 * (g.mt0 + 4) <= nrows && nb < GEMM_NB - 1) {
    // every row is a request's and every column a B'' column (uniform)
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      const int row = 16 * (g.mt0 + mi) + (lane & 15);
      uint32_t mism = 0;
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) {
        const int n0 = 16 * (4 * nb + ni) + 4 * (lane >> 4);
        uint32_t v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = (uint32_t)acc[mi][ni][0][r] + ((uint32_t)acc[mi][ni][1][r] << 8);
        if constexpr (MODE == kProducts) {
          const uint32_t p0 = __builtin_amdgcn_perm(v[1], v[0], 0x05040100u), p1 = __builtin_amdgcn_perm(v[3], v[2], 0x05040100u);
          *reinterpret_cast<uint2*>(out + (size_t)row * OUTC + n0) = make_uint2(p0, p1);
        } else {
          // two uint16 big-endian per word, as the ciphertext holds them
          const uint32_t p0 = __builtin_amdgcn_perm(v[1], v[0], 0x04050001u), p1 = __builtin_amdgcn_perm(v[3], v[2], 0x04050001u);
          if constexpr (MODE == kEncaps)
            *reinterpret_cast<uint2*>(ct_word<MODE>(ct, row, n0)) = make_uint2(p0, p1);
          else
            mism |= (p0 ^ ep.c[mi][ni].x) | (p1 ^ ep.c[mi][ni].y);
        }
      }
      if constexpr (MODE == kDecaps) gemm_flags(mism, g, mi, n, flags);
    }
    return;
  }
#pragma unroll
  for (int mi = 0; mi < 4; ++mi) {
    const int row = 16 * (g.mt0 + mi) + (lane & 15), req = row >> 3, k = row & 7;
    uint32_t mism = 0;
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const int n0 = 16 * (4 * nb + ni) + 4 * (lane >> 4);
      if (row >= nrows || n0 >= N + NB) continue;
      uint32_t v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        v[r] = (uint32_t)acc[mi][ni][0][r] + ((uint32_t)acc[mi][ni][1][r] << 8);
        if constexpr (MODE != kProducts)
          if (n0 >= N) v[r] += encode_mu(mu + (size_t)req * mu_stride, k, n0 - N + r);
      }
      const uint32_t p0 = (v[0] & 0xFFFF) | (v[1] << 16), p1 = (v[2] & 0xFFFF) | (v[3] << 16);
      if constexpr (MODE == kProducts) {
        *reinterpret_cast<uint2*>(out + (size_t)row * OUTC + n0) = make_uint2(p0, p1);
      } else if constexpr (MODE == kEncaps) {
        *reinterpret_cast<uint2*>(ct_word<MODE>(ct, row, n0)) = make_uint2(bswap16x2(p0), bswap16x2(p1));
      } else {
        mism |= (p0 ^ bswap16x2(ep.c[mi][ni].x)) | (p1 ^ bswap16x2(ep.c[mi][ni].y));
      }
    }
    if constexpr (MODE == kDecaps) gemm_flags(mism, g, mi, n, flags);
  }
}

// ---- decapsulation: F's last block ----------------------------------------------
// One lane per request: absorb ct[15640:15744] || key (k' or s, by the OR of the
// request's 16 compare flags, as a mask) into the saved sponge state, pad,
// permute: ss = F(ct || key) (kem.c:195-202).
__global__ void k_f_final(int n, uint8_t* ss, const uint8_t* ct, const uint64_t* fstate, const uint8_t* flags,
                          const uint8_t* kp, FrKeys keys) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint64_t S[25];
#pragma unroll
  for (int w = 0; w < 25; ++w) S[w] = fstate[(size_t)r * 25 + w];
  const uint64_t* p = reinterpret_cast<const uint64_t*>(ct + (size_t)r * CT + 115 * 136);
#pragma unroll
  for (int w = 0; w < 13; ++w) S[w] ^= p[w];
  const uint4 f = *reinterpret_cast<const uint4*>(flags + (size_t)r * GEMM_NB);
  const uint64_t mask = 0 - (uint64_t)((f.x | f.y | f.z | f.w) != 0);
  const uint8_t* s = keys.sk_of(r);
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    const uint64_t a = load_u64(kp + (size_t)r * 32 + 8 * w), b = load_u64(s + 8 * w);
    S[13 + w] ^= (a & ~mask) | (b & mask);
  }
  S[16] ^= 0x1Full | 0x8000000000000000ull;
  keccakf1600(S);
  for (int w = 0; w < 3; ++w) store_u64(ss + (size_t)r * 24 + 8 * w, S[w]);
}

// ---- encapsulation: ss = F(ct || k), one lane pair per request ---------------------
__global__ __launch_bounds__(64) void k_f_full(int n, uint8_t* ss, const uint8_t* ct, const uint8_t* kp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = t >> 1, half = t & 1;
  if (r >= n) return;
  uint32_t S[25];
  absorb_ct_pair(S, ct + (size_t)r * CT, half);
  const uint32_t* p = reinterpret_cast<const uint32_t*>(ct + (size_t)r * CT + 115 * 136) + half;
#pragma unroll
  for (int w = 0; w < 13; ++w) S[w] ^= p[2 * w];
  const uint32_t* k = reinterpret_cast<const uint32_t*>(kp + (size_t)r * 32) + half;
#pragma unroll
  for (int w = 0; w < 3; ++w) S[13 + w] ^= k[2 * w];
  S[16] ^= half ? 0x80000000u : 0x1Fu;
  keccakf1600_pair(S, (uint32_t)half);
  uint32_t* o = reinterpret_cast<uint32_t*>(ss + (size_t)r * 24) + half;
#pragma unroll
  for (int w = 0; w < 3; ++w) o[2 * w] = S[w];
}

// ---- many keys: products on packed 16-bit VALU ---------------------------------------
// [B'' | V] = S'.[A | B] + [E' | E''] (mod 2^16) for one request per workgroup
// pair: workgroup (r, y) computes columns 512 y + 2 t, +1 for thread t with
// v_pk_mad_u16 (two 16-bit multiply-adds mod 2^16 per lane), reading row k of
// the request's own A (key cache or per-request expansion) and S'[.][k] as a
// broadcast from LDS.  Then, as k_gemm's epilogue, compare with ct (decaps,
// one flag byte per workgroup among the request's 16) or pack into ct (encaps).
// A request's key has its own A, so there is no shared operand for MFMA here.
typedef uint16_t u16x2v __attribute__((ext_vector_type(2)));

template <int MODE>
__global__ __launch_bounds__(256) void k_mk_product(int n, const int8_t* __restrict__ srow,
                                                    const int8_t* __restrict__ erow, FrKeys keys,
                                                    uint8_t* __restrict__ ct, const uint8_t* __restrict__ mu,
                                                    int mu_stride, uint8_t* __restrict__ flags) {
  __shared__ uint4 st[N];  // S'[0..7][k] as 8 uint16 per k
  __shared__ uint32_t wmis[4];
  const int r = blockIdx.x, y = blockIdx.y, tid = threadIdx.x;
  if (r >= n) return;
  for (int k = tid; k < N; k += blockDim.x) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t lo = (uint32_t)(int32_t)srow[((size_t)r * NB + 2 * q) * N + k] & 0xFFFF;
      const uint32_t hi = (uint32_t)(int32_t)srow[((size_t)r * NB + 2 * q + 1) * N + k] & 0xFFFF;
      w[q] = lo | (hi << 16);
    }
    st[k] = make_uint4(w[0], w[1], w[2], w[3]);
  }
  __syncthreads();
  const int c0 = 512 * y + 2 * tid;  // this thread's columns c0, c0 + 1
  const uint16_t* A = keys.a_of(r);
  const uint8_t* pk = keys.pk_of(r);
  u16x2v acc[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) acc[i] = u16x2v{0, 0};
  if (c0 < N + NB) {
#pragma unroll 4
    for (int k = 0; k < N; ++k) {
      uint32_t a2;
      if (c0 < N)
        a2 = *reinterpret_cast<const uint32_t*>(A + (size_t)k * N + c0);
      else  // B columns from the public key, big-endian uint16
        a2 = __builtin_amdgcn_perm(0u, *reinterpret_cast<const uint32_t*>(pk + SEEDA + 2 * (NB * k + (c0 - N))),
                                   0x02030001u);
      const u16x2v av = __builtin_bit_cast(u16x2v, a2);
      const uint4 sv = st[k];
      const uint32_t sw[4] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint16_t s0 = (uint16_t)sw[q], s1 = (uint16_t)(sw[q] >> 16);
        acc[2 * q] += av * u16x2v{s0, s0};
        acc[2 * q + 1] += av * u16x2v{s1, s1};
      }
    }
  }
  // epilogue: + E' | E'', then compare / pack
  uint32_t mism = 0;
  if (c0 < N + NB) {
    const int8_t* e1 = erow + (size_t)r * N * NB;
    const int8_t* e2 = erow + (size_t)n * N * NB + (size_t)r * 64;
    uint8_t* c = ct + (size_t)r * CT;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      uint32_t v0 = acc[i][0], v1 = acc[i][1];
      if (c0 < N) {
        v0 += (uint32_t)(int32_t)e1[i * N + c0];
        v1 += (uint32_t)(int32_t)e1[i * N + c0 + 1];
      } else {
        const int j = c0 - N;
        v0 += (uint32_t)(int32_t)e2[8 * i + j] + encode_mu(mu + (size_t)r * mu_stride, i, j);
        v1 += (uint32_t)(int32_t)e2[8 * i + j + 1] + encode_mu(mu + (size_t)r * mu_stride, i, j + 1);
      }
      const uint32_t p = (v0 & 0xFFFF) | (v1 << 16);
      uint32_t* cw = reinterpret_cast<uint32_t*>(c0 < N ? c + 2 * (i * N + c0) : c + C1 + 2 * (8 * i + (c0 - N)));
      if constexpr (MODE == kEncaps) *cw = bswap16x2(p);
      else mism |= p ^ bswap16x2(*cw);
    }
  }
  if constexpr (MODE == kDecaps) {
    const uint64_t bal = __ballot(mism != 0);
    if ((tid & 63) == 0) wmis[tid >> 6] = bal != 0;
    __syncthreads();
    if (tid < 8) {  // bytes 8 y .. 8 y + 7 of the request's 16 flag bytes
      const uint8_t f = (uint8_t)(tid == 0 ? (wmis[0] | wmis[1] | wmis[2] | wmis[3]) : 0);
      flags[(size_t)r * GEMM_NB + 8 * y + tid] = f;
    }
  }
}

// pkh = SHAKE256(pk, 24) for n public keys, one lane pair per key; key r's
// hash at pkh + r * out_stride
__global__ __launch_bounds__(64) void k_pkh_many(int n, uint8_t* pkh, size_t out_stride, const uint8_t* pk,
                                                 size_t pk_stride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = t >> 1, half = t & 1;
  if (r >= n) return;
  uint32_t S[25];
#pragma unroll
  for (int w = 0; w < 25; ++w) S[w] = 0;
  const uint32_t* p = reinterpret_cast<const uint32_t*>(pk + (size_t)r * pk_stride) + half;
  constexpr int W = PK / 8;  // 1954 words: 114 full blocks + 16 words
  for (int blk = 0; blk < W / kRateShake256; ++blk) {
#pragma unroll
    for (int w = 0; w < kRateShake256; ++w) S[w] ^= p[2 * w];
    keccakf1600_pair(S, (uint32_t)half);
    p += 2 * kRateShake256;
  }
#pragma unroll
  for (int w = 0; w < W % kRateShake256; ++w) S[w] ^= p[2 * w];
  S[W % kRateShake256] ^= half ? 0u : 0x1Fu;
  S[kRateShake256 - 1] ^= half ? 0x80000000u : 0u;
  keccakf1600_pair(S, (uint32_t)half);
  uint32_t* o = reinterpret_cast<uint32_t*>(pkh + (size_t)r * out_stride) + half;
#pragma unroll
  for (int w = 0; w < 3; ++w) o[2 * w] = S[w];
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


}  // namespace frodo

// ---- host orchestration -------------------------------------------------------------

using namespace frodo;

size_t frodo_bytes_per_elem() {
  // 31232: key generation's S || E samples (uint16), or [E' | E''] int8 fragments
  // 8192 + GEMM output uint16 16384; then mu 32, k 32, F state 200, compare flags 16.
  // The S' fragments live in their own 8 KB-per-element buffer.
  // Then the raw S'/E'/E'' SHAKE256 stream, RAWB bytes.
  return 31232 + 32 + 32 + 200 + 16 + RAWB;
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
  int8_t *sfrag, *efrag;
  uint16_t *out, *se;
  uint8_t *mu, *kp, *flags, *raw;
  uint32_t* fstate;
};

static FrScratch fr_scratch(pqc_ctx* c) {
  const size_t ld = c->ld;  // a multiple of 64, so every region below is 16-byte aligned
  uint8_t* p = c->fr_elem;
  FrScratch f;
  f.se = (uint16_t*)p;
  f.efrag = (int8_t*)p;                              // 8 rows x 1024 bytes per element
  f.out = (uint16_t*)(f.efrag + (size_t)8192 * ld);  // 8 rows x 1024 uint16 per element
  p += (size_t)31232 * ld;
  f.mu = p;
  p += (size_t)32 * ld;
  f.kp = p;
  p += (size_t)32 * ld;
  f.fstate = (uint32_t*)p;
  p += (size_t)200 * ld;
  f.flags = p;
  p += (size_t)16 * ld;
  f.raw = p;
  f.sfrag = c->fr_sfrag;
  return f;
}

// The scratch of elements [e0, e0 + m) as a scratch of its own: every region is
// element-linear (S' and E' fragments: 8192 bytes per element, whole 16-row
// m-tiles for even e0), so a chunk of a batch runs its kernels on these pointers
// with element indices relative to e0.
static FrScratch fr_scratch_at(pqc_ctx* c, size_t e0) {
  FrScratch f = fr_scratch(c);
  f.se += e0 * (31232 / 2);
  f.efrag += e0 * 8192;
  f.out += e0 * 8 * OUTC;
  f.mu += e0 * 32;
  f.kp += e0 * 32;
  f.fstate += e0 * 50;
  f.flags += e0 * GEMM_NB;
  f.raw += e0 * RAWB;
  f.sfrag += e0 * 8192;
  return f;
}

template <int MODE>
static int launch_gemm(pqc_ctx* c, size_t n, const FrScratch& f, uint8_t* ct, const uint8_t* mu, int mu_stride,
                       hipStream_t s) {
  const int mtiles = (int)((8 * n + 15) / 16);
  const int mblocks = (mtiles + 15) / 16;  // whole 256-row blocks fit the S' buffer (ld is a multiple of 64)
  const int mblocks8 = (mblocks + 7) / 8 * 8;
  hipLaunchKernelGGL(k_gemm<MODE>, dim3(mblocks8 * GEMM_NB), dim3(256), 0, s, mblocks, (int)n,
                     reinterpret_cast<const v4i*>(f.sfrag), reinterpret_cast<const v4i*>(c->fr_frag),
                     (const int8_t*)f.efrag, f.out, ct, mu, mu_stride, f.flags);
  return hip_err(hipGetLastError());
}

static inline int pairs_pad(int n) { return (n + 31) / 32 * 32; }

constexpr int kFrChunk = 128;  // per-element-key requests per A expansion (A buffer: 128 x 1.9 MB)

// the loaded key as a key source for every request (strides 0)
static FrKeys loaded_key(pqc_ctx* c) {
  return FrKeys{c->fr_sk, 0, c->fr_pkh, 0, c->fr_pk, 0, c->fr_a16, 0, nullptr};
}

// G2 and the S'/E'/E'' stream of n requests (lane pairs; with_f adds the F prefix
// of each ciphertext), then its sampling into the ROWS or fragment layouts.
template <bool ROWS>
static void launch_stream(pqc_ctx* c, const FrScratch& f, int ni, int with_f, const FrKeys& keys, const uint8_t* mu,
                          int mu_stride, const uint8_t* ct, hipStream_t s) {
  const int sb = c->tune.frodo_stream_block, nt = ni * kSampleThreads;
  // shared-key decapsulation: its F-prefix lanes share the SIMDs with the stream
  // lanes.  Encapsulation on single lanes runs one wave per SIMD (4.4 against 3.4 ms
  // on pairs at 32768, no faster at 65536); the keyed paths run at most a server batch
  if (!ROWS && with_f && frodo_lanes(c->tune, (uint64_t)ni)) {
    const int nlane = (ni + 63) / 64 * 64, lanes = nlane + (with_f ? ni : 0);
    hipLaunchKernelGGL(k_shake_lane, dim3((lanes + sb - 1) / sb), dim3(sb), 0, s, ni, nlane, with_f, keys, mu,
                       mu_stride, f.raw, f.kp, ct, with_f ? reinterpret_cast<uint64_t*>(f.fstate) : nullptr);
    hipLaunchKernelGGL((k_sample<false, false>), dim3((nt + 255) / 256), dim3(256), 0, s, ni, (const uint8_t*)f.raw,
                       f.sfrag, f.efrag);
    return;
  }
  const int npad = pairs_pad(ni);
  const int pairs = npad + (with_f ? ni : 0);
  hipLaunchKernelGGL(k_shake, dim3((2 * pairs + sb - 1) / sb), dim3(sb), 0, s, ni, npad, with_f, keys, mu, mu_stride,
                     f.raw, f.kp, ct, with_f ? f.fstate : (uint32_t*)nullptr);
  hipLaunchKernelGGL((k_sample<ROWS, true>), dim3((nt + 255) / 256), dim3(256), 0, s, ni, (const uint8_t*)f.raw,
                     f.sfrag, f.efrag);
}

// ---- batches under the loaded key, optionally cut into chunks over streams ------
// One request's S'/E'/E'' stream is a serial chain of 231 permutations on a lane
// pair (plus a pair for the F prefix), so a batch of 4096 keeps only a quarter of
// the SIMDs busy while k_shake runs; 16384 requests put one wave on every SIMD.
// Batches above tune.frodo_chunk requests are cut into chunks of that many, and
// chunk i runs its whole chain (w-decode, stream, sampling, shared-A GEMM, F) on
// stream i mod 4: the caller's stream or one of the context's three side
// streams, forked from and joined back to the caller's stream by events.  Every
// chunk has scratch of its own (fr_scratch_at), so the bytes are those of the
// unsplit batch.  Measured on MI355X, four concurrent 4096-request chains slow
// each other's k_shake to 1.6-3.2 ms (1.1 ms alone), so one chain over the whole
// batch is faster than chunks of 4096; the default therefore splits only above
// 65536 requests, and the knob stays for the autotuner.
constexpr size_t kFrPipeChunkDefault = 65536;
static_assert(kFrPipeChunkDefault % 32 == 0, "chunks start on whole 256-row GEMM blocks");

template <class F>
static int frodo_pipelined(pqc_ctx* c, size_t n, hipStream_t s, F chunk) {
  const size_t kFrPipeChunk = c->tune.frodo_chunk > 0 ? (size_t)c->tune.frodo_chunk : kFrPipeChunkDefault;
  if (n <= kFrPipeChunk || !c->fr_side[0]) return chunk((size_t)0, n, s);
  const size_t nchunks = (n + kFrPipeChunk - 1) / kFrPipeChunk;
  const int nside = (int)std::min<size_t>(kFrSideStreams, nchunks - 1);
  if (hipEventRecord(c->fr_fork, s) != hipSuccess) return PQC_ERR_DEVICE;
  for (int i = 0; i < nside; ++i)
    if (hipStreamWaitEvent(c->fr_side[i], c->fr_fork, 0) != hipSuccess) return PQC_ERR_DEVICE;
  int rc = PQC_OK;
  for (size_t k = 0; k < nchunks && !rc; ++k) {
    const int lane = (int)(k % (size_t)(nside + 1));
    hipStream_t t = lane == 0 ? s : c->fr_side[lane - 1];
    const size_t e0 = k * kFrPipeChunk;
    rc = chunk(e0, std::min(kFrPipeChunk, n - e0), t);
  }
  // join every side stream back, also after an error, so none is left forked
  for (int i = 0; i < nside; ++i)
    if (hipEventRecord(c->fr_join[i], c->fr_side[i]) != hipSuccess ||
        hipStreamWaitEvent(s, c->fr_join[i], 0) != hipSuccess)
      rc = rc ? rc : PQC_ERR_DEVICE;
  return rc;
}

int frodo_dec(pqc_ctx* c, size_t n, uint8_t* ss, const uint8_t* ct, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  if (!c->fr_loaded || !c->fr_has_sk) return PQC_ERR_INVALID_ARG;
  if (n == 0) return PQC_OK;
  const FrKeys keys = loaded_key(c);
  return frodo_pipelined(c, n, s, [&](size_t e0, size_t m, hipStream_t t) {
    const FrScratch f = fr_scratch_at(c, e0);
    const int ni = (int)m;
    const uint8_t* ctc = ct + e0 * CT;
    uint8_t* ssc = ss + e0 * SS;
    hipLaunchKernelGGL(k_w_decode, dim3(w_decode_grid(c, ni)), dim3(64), 0, t, ni, ctc, keys, f.mu);
    launch_stream<false>(c, f, ni, 1, keys, f.mu, 32, ctc, t);
    int r = launch_gemm<kDecaps>(c, m, f, const_cast<uint8_t*>(ctc), f.mu, 32, t);
    if (r) return r;
    hipLaunchKernelGGL(k_f_final, dim3((ni + 63) / 64), dim3(64), 0, t, ni, ssc, ctc, (const uint64_t*)f.fstate,
                       (const uint8_t*)f.flags, (const uint8_t*)f.kp, keys);
    return hip_err(hipGetLastError());
  });
}

// ---- many keys ------------------------------------------------------------------------
// Decapsulation / encapsulation of n requests whose keys (and expanded A) come
// from `keys`; the products run on k_mk_product.
int frodo_dec_keyed(pqc_ctx* c, size_t n, uint8_t* ss, const uint8_t* ct, const FrKeys& keys, hipStream_t s) {
  if (!c->fr_elem || !c->fr_sfrag) return PQC_ERR_UNSUPPORTED;
  if (n == 0) return PQC_OK;
  FrScratch f = fr_scratch(c);
  const int ni = (int)n;
  hipLaunchKernelGGL(k_w_decode, dim3(w_decode_grid(c, ni)), dim3(64), 0, s, ni, ct, keys, f.mu);
  launch_stream<true>(c, f, ni, 1, keys, f.mu, 32, ct, s);
  hipLaunchKernelGGL(k_mk_product<kDecaps>, dim3(ni, 2), dim3(256), 0, s, ni, (const int8_t*)f.sfrag,
                     (const int8_t*)f.efrag, keys, const_cast<uint8_t*>(ct), (const uint8_t*)f.mu, 32, f.flags);
  hipLaunchKernelGGL(k_f_final, dim3((ni + 63) / 64), dim3(64), 0, s, ni, ss, ct, (const uint64_t*)f.fstate,
                     (const uint8_t*)f.flags, (const uint8_t*)f.kp, keys);
  return hip_err(hipGetLastError());
}

int frodo_enc_keyed(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* mu, const FrKeys& keys,
                    hipStream_t s) {
  if (!c->fr_elem || !c->fr_sfrag) return PQC_ERR_UNSUPPORTED;
  if (n == 0) return PQC_OK;
  FrScratch f = fr_scratch(c);
  const int ni = (int)n;
  launch_stream<true>(c, f, ni, 0, keys, mu, MU, nullptr, s);
  hipLaunchKernelGGL(k_mk_product<kEncaps>, dim3(ni, 2), dim3(256), 0, s, ni, (const int8_t*)f.sfrag,
                     (const int8_t*)f.efrag, keys, ct, mu, MU, (uint8_t*)nullptr);
  hipLaunchKernelGGL(k_f_full, dim3((2 * ni + 63) / 64), dim3(64), 0, s, ni, ss, (const uint8_t*)ct,
                     (const uint8_t*)f.kp);
  return hip_err(hipGetLastError());
}

// Per-element keys (crypto_kem_enc/dec_batch): A is expanded per request into
// the context's A buffer, kFrChunk requests at a time.
int frodo_dec_batch(pqc_ctx* c, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk, hipStream_t s) {
  if (!c->fr_abuf) return PQC_ERR_UNSUPPORTED;
  for (size_t off = 0; off < n; off += kFrChunk) {
    const int m = (int)(n - off < (size_t)kFrChunk ? n - off : (size_t)kFrChunk);
    const uint8_t* skc = sk + off * SK;
    hipLaunchKernelGGL(k_gen_a_many, dim3((m * N + 255) / 256), dim3(256), 0, s, m, c->fr_abuf, skc + SS,
                       (size_t)SK);
    const FrKeys keys{skc, SK, skc + SK_PKH, SK, skc + SS, SK, c->fr_abuf, (size_t)N * N, nullptr};
    int rc = frodo_dec_keyed(c, m, ss + off * SS, ct + off * CT, keys, s);
    if (rc) return rc;
  }
  return PQC_OK;
}

int frodo_enc_batch(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk, const uint8_t* mu,
                    hipStream_t s) {
  return frodo_enc_batch_strided(c, n, ct, ss, pk, PK, mu, s);
}

// public key of element r at pk + r * pk_stride (8-byte aligned)
int frodo_enc_batch_strided(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk, size_t pk_stride,
                            const uint8_t* mu, hipStream_t s) {
  if (!c->fr_abuf) return PQC_ERR_UNSUPPORTED;
  uint8_t* pkh = (uint8_t*)fr_scratch(c).out;  // 24 bytes per request, the products' output is unused here
  for (size_t off = 0; off < n; off += kFrChunk) {
    const int m = (int)(n - off < (size_t)kFrChunk ? n - off : (size_t)kFrChunk);
    const uint8_t* pkc = pk + off * pk_stride;
    hipLaunchKernelGGL(k_gen_a_many, dim3((m * N + 255) / 256), dim3(256), 0, s, m, c->fr_abuf, pkc, pk_stride);
    hipLaunchKernelGGL(k_pkh_many, dim3((2 * m + 63) / 64), dim3(64), 0, s, m, pkh, (size_t)24, pkc, pk_stride);
    const FrKeys keys{nullptr, 0, pkh, 24, pkc, pk_stride, c->fr_abuf, (size_t)N * N, nullptr};
    int rc = frodo_enc_keyed(c, m, ct + off * CT, ss + off * SS, mu + off * MU, keys, s);
    if (rc) return rc;
  }
  return PQC_OK;
}

int frodo_enc(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* mu, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  if (!c->fr_loaded) return PQC_ERR_INVALID_ARG;
  if (n == 0) return PQC_OK;
  const FrKeys keys = loaded_key(c);
  return frodo_pipelined(c, n, s, [&](size_t e0, size_t m, hipStream_t t) {
    const FrScratch f = fr_scratch_at(c, e0);
    const int ni = (int)m;
    uint8_t* ctc = ct + e0 * CT;
    const uint8_t* muc = mu + e0 * MU;
    launch_stream<false>(c, f, ni, 0, keys, muc, MU, nullptr, t);
    int r = launch_gemm<kEncaps>(c, m, f, ctc, muc, MU, t);
    if (r) return r;
    hipLaunchKernelGGL(k_f_full, dim3((2 * ni + 63) / 64), dim3(64), 0, t, ni, ss + e0 * SS, (const uint8_t*)ctc,
                       (const uint8_t*)f.kp);
    return hip_err(hipGetLastError());
  });
}

int frodo_keypair(pqc_ctx* c, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* rnd, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  if (n == 0) return PQC_OK;
  FrScratch f = fr_scratch(c);
  uint8_t* seedA = f.mu;
  const int ni = (int)n;
  hipLaunchKernelGGL(k_kp_stream, dim3((ni + 63) / 64), dim3(64), 0, s, ni, rnd, seedA, f.se);
  for (int r = 0; r < ni; ++r) {
    hipLaunchKernelGGL(k_gen_a, dim3((N + 63) / 64), dim3(64), 0, s, c->fr_a16, (const uint8_t*)seedA + (size_t)r * 16);
    hipLaunchKernelGGL(k_kp_finish, dim3((N * NB + 255) / 256), dim3(256), 0, s, r, pk, sk, (const uint16_t*)c->fr_a16,
                       (const uint16_t*)f.se, (const uint8_t*)seedA, rnd);
  }
  // sk tail: H(pk) of every key in one launch
  hipLaunchKernelGGL(k_pkh_many, dim3((2 * ni + 63) / 64), dim3(64), 0, s, ni, sk + SK_PKH, (size_t)SK,
                     (const uint8_t*)pk, (size_t)PK);
  c->fr_loaded = 0;  // the A buffer now holds the last generated key's A
  return hip_err(hipGetLastError());
}

// Key store: n_keys secret keys (pk and H(pk) inside) and their expanded A.
int frodo_expand_a(size_t n_keys, uint16_t* a_out, const uint8_t* sk_store, hipStream_t s) {
  if (n_keys == 0) return PQC_OK;
  const int m = (int)n_keys;
  hipLaunchKernelGGL(k_gen_a_many, dim3((unsigned)(((size_t)m * N + 255) / 256)), dim3(256), 0, s, m, a_out,
                     sk_store + SS, (size_t)SK);
  return hip_err(hipGetLastError());
}

int frodo_dec_store(pqc_ctx* c, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk_store,
                    const uint16_t* a_store, const uint32_t* kid, hipStream_t s) {
  const FrKeys keys{sk_store, SK, sk_store + SK_PKH, SK, sk_store + SS, SK, a_store, (size_t)N * N, kid};
  return frodo_dec_keyed(c, n, ss, ct, keys, s);
}

int frodo_enc_store(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* mu, const uint8_t* sk_store,
                    const uint16_t* a_store, const uint32_t* kid, hipStream_t s) {
  const FrKeys keys{sk_store, SK, sk_store + SK_PKH, SK, sk_store + SS, SK, a_store, (size_t)N * N, kid};
  return frodo_enc_keyed(c, n, ct, ss, mu, keys, s);
}

int frodo_gemm(pqc_ctx* c, size_t n, const uint8_t* ct, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  if (!c->fr_loaded) return PQC_ERR_INVALID_ARG;
  FrScratch f = fr_scratch(c);
  if (ct) return launch_gemm<kDecaps>(c, n, f, const_cast<uint8_t*>(ct), f.mu, 32, s);
  return launch_gemm<kProducts>(c, n, f, nullptr, nullptr, 0, s);
}

int frodo_products(pqc_ctx* c, size_t n, uint16_t* bpp, uint16_t* v, hipStream_t s) {
  int rc = frodo_need(c);
  if (rc) return rc;
  if (!c->fr_loaded) return PQC_ERR_INVALID_ARG;
  FrScratch f = fr_scratch(c);
  rc = launch_gemm<kProducts>(c, n, f, nullptr, nullptr, 0, s);
  if (rc) return rc;
  if (bpp && hipMemcpy2DAsync(bpp, N * 2, f.out, OUTC * 2, N * 2, 8 * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return PQC_ERR_DEVICE;
  if (v && hipMemcpy2DAsync(v, NB * 2, f.out + N, OUTC * 2, NB * 2, 8 * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return PQC_ERR_DEVICE;
  return PQC_OK;
}

}  // namespace pqc
