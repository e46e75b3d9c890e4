// NewHope CPA-PKE (N = 1024 and N = 512) as plain batched kernels
// (NewHope_NIST_GPU/cpapke.cu:265-476): one kernel per reference operation,
// and the NewHope CCA-KEM over them.  Polynomials are element-major (element
// x's N uint16 coefficients contiguous at x * N), and each kernel spreads an
// element over as many lanes as its data flow allows:
//   - NTT / inverse NTT: one wave per polynomial, N / 64 coefficients per lane
//     in VGPRs, the butterfly layers in register-resident groups (N = 1024:
//     index bits 0-3, 4-7, 8-9; N = 512: bits 0-2, 3-5, 6-8) with the
//     polynomial re-laid out through LDS between groups;
//   - poly_uniform / poly_sample: one lane per 64-coefficient SHAKE stream
//     (N / 64 per polynomial);
//   - pointwise, add/sub, (de)serialisation, (de)compression, message coding:
//     flat passes, one lane per 4 or 8 coefficients, 8- or 16-byte accesses.
//
// The 16-bit storage and 32-bit intermediate arithmetic of the reference are
// kept exactly (lazy even NTT levels stored as uint16, Barrett on a uint16
// argument at N = 1024 and % q at N = 512 on the odd levels, Montgomery with
// R = 2^18 in uint32), so results are identical for every input, including
// out-of-range ones.  N = 512 follows the reference's NEWHOPE_N == 512 branches
// (precomp.cu:75-190 tables, ntt.cu:130-180, poly.cu:244,279), which its
// batched GPU path never compiled.
#include "internal.h"
#include "sponge.cuh"

namespace pqc {
namespace newhope {

constexpr uint32_t Q = 12289;
constexpr int SYM = 32;

// sizes of NewHope-N (params.h:26-39)
template <int N>
struct Par {
  static constexpr int POLYBYTES = 14 * N / 8;          // 1792 / 896
  static constexpr int CBYTES = 3 * N / 8;              // 384 / 192
  static constexpr int PK = POLYBYTES + SYM;            // 1824 / 928
  static constexpr int SK = POLYBYTES;                  // 1792 / 896
  static constexpr int CT = POLYBYTES + CBYTES;         // 2176 / 1088
  static constexpr int CCA_SK = SK + PK + 2 * SYM;      // 3680 / 1888: sk_cpa || pk || H(pk) || s
  static constexpr int CCA_CT = CT + SYM;               // 2208 / 1120: ct_cpa || d
  static constexpr int LOG = N == 1024 ? 10 : 9;
  static constexpr uint32_t G = N == 1024 ? 7 : 10968;  // precomp.cu: g = Mod(7 | 10968, q)
  static constexpr int STREAMS = N / 64;                // SHAKE streams per polynomial
  static constexpr int COPIES = N / 256;                // coefficients per message bit
};

// ---- tables: the Pari/GP definitions of precomp.cu:5-72 (mont = 2^18),
// evaluated at compile time ---------------------------------------------------
constexpr uint32_t powmod(uint32_t b, uint32_t e) {
  uint64_t r = 1, x = b % Q;
  while (e) {
    if (e & 1) r = r * x % Q;
    x = x * x % Q;
    e >>= 1;
  }
  return (uint32_t)r;
}

constexpr uint32_t bitrev(uint32_t i, int bits) {
  uint32_t r = 0;
  for (int b = 0; b < bits; b++) r |= ((i >> b) & 1u) << (bits - 1 - b);
  return r;
}

template <int N>
struct alignas(16) Tables {  // 16-byte aligned rows: read as uint4 by the NTT kernels
  uint16_t omegas_inv[N / 2];
  uint16_t gammas[N];
  uint16_t gammas_inv[N];
  uint16_t brv[N];
};

template <int N>
constexpr Tables<N> make_tables() {
  Tables<N> t{};
  constexpr int B = Par<N>::LOG;
  const uint32_t g = Par<N>::G, mont = (1u << 18) % Q;
  const uint32_t g2inv = powmod(g * g % Q, Q - 2), ginv = powmod(g, Q - 2), ninv = powmod(N, Q - 2);
  for (uint32_t k = 0; k < N / 2; k++) t.omegas_inv[k] = (uint16_t)((uint64_t)powmod(g2inv, bitrev(2 * k, B)) * mont % Q);
  for (uint32_t k = 0; k < N; k++) {
    t.gammas[k] = (uint16_t)((uint64_t)powmod(g, bitrev(k, B)) * mont % Q);
    t.gammas_inv[k] = (uint16_t)((uint64_t)powmod(ginv, k) * ninv % Q * mont % Q);
    t.brv[k] = (uint16_t)bitrev(k, B);
  }
  return t;
}

__constant__ Tables<1024> c_tab1024 = make_tables<1024>();
__constant__ Tables<512> c_tab512 = make_tables<512>();

template <int N>
__device__ __forceinline__ const Tables<N>& tab();
template <>
__device__ __forceinline__ const Tables<1024>& tab<1024>() { return c_tab1024; }
template <>
__device__ __forceinline__ const Tables<512>& tab<512>() { return c_tab512; }

// floor(n / q) for n < 2^21: the high word of the 24-bit product n m with
// m = ceil(2^32 / q), one v_mul_hi_u32_u24 where the compiler's division by a
// constant takes five instructions.  Exact: n m / 2^32 = n / q + n e / (q 2^32)
// with e = m q - 2^32, and n e < 2^32.  Every operand here is below 2^17 (sums of
// two uint16 values, or of a uint16 value and 3q).  Written as the instruction
// itself because the compiler does not form it from C.
constexpr uint32_t kDivQM = (uint32_t)(((1ull << 32) + Q - 1) / Q);
static_assert(kDivQM < (1u << 24), "24-bit operand");
static_assert((((uint64_t)kDivQM * Q - (1ull << 32)) << 21) < (1ull << 32), "exact below 2^21");
static_assert(kDivQM == 0x55539u, "literal below");

__device__ __forceinline__ uint32_t div_q(uint32_t n) {
  uint32_t r;
  asm("v_mul_hi_u32_u24 %0, 0x55539, %1" : "=v"(r) : "v"(n));
  return r;
}

__device__ __forceinline__ uint32_t mod_q(uint32_t n) { return n - div_q(n) * Q; }

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

template <int N>
struct Pv {
  uint16_t* p;  // element x's coefficients at p + N x
  __device__ __forceinline__ uint16_t* el(int x) const { return p + (size_t)x * N; }
};

// ---- sponge kernels --------------------------------------------------------

// keypair seed expansion: z = SHAKE256(seed, 64) (cpapke.cu:283-298)
__global__ void k_seed_expand(int n, uint8_t* z, const uint8_t* seed, size_t seed_stride) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t m[4], A[25];
#pragma unroll
  for (int w = 0; w < 4; ++w) m[w] = load_u64(seed + (size_t)x * seed_stride + 8 * w);
  absorb_fixed<kRateShake256, 4>(A, m, 0x1F);
#pragma unroll
  for (int w = 0; w < 8; ++w) store_u64(z + (size_t)x * 64 + 8 * w, A[w]);
}

// gen_a / poly_uniform (cpapke.cu:236-247, poly.cu:332-366): lane (x, i) runs
// SHAKE128(seed || i) and keeps 16-bit candidates < 5q for coefficients 64 i ..
// 64 i + 63.  Accepted values land at data-dependent positions, so each lane
// collects its 64 in its own LDS row (66 uint16: 33 words, rows on distinct
// banks) and then writes them out as eight 16-byte stores.  Fixed 128-lane
// workgroups (16.5 KB of LDS).
constexpr int kUniformBlock = 128;
constexpr int kUniformRow = 66;

template <int N>
__global__ __launch_bounds__(kUniformBlock) void k_uniform(int n, Pv<N> a, const uint8_t* seed, size_t stride) {
  constexpr int S = Par<N>::STREAMS;
  __shared__ uint16_t stage[kUniformBlock * kUniformRow];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = t / S, i = t % S;
  if (x >= n) return;  // no barrier below: every lane reads back only its own row
  uint16_t* row = stage + threadIdx.x * kUniformRow;
  uint64_t A[25];
#pragma unroll
  for (int w = 0; w < 25; ++w) A[w] = w < 4 ? load_u64(seed + (size_t)x * stride + 8 * w) : 0;
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
        if (val < 5 * Q && ctr < 64) row[ctr++] = (uint16_t)val;
      }
    }
  }
  const uint32_t* rw = reinterpret_cast<const uint32_t*>(row);
  uint4* dst = reinterpret_cast<uint4*>(a.el(x) + 64 * i);
#pragma unroll
  for (int q = 0; q < 8; ++q) dst[q] = make_uint4(rw[4 * q], rw[4 * q + 1], rw[4 * q + 2], rw[4 * q + 3]);
}

// poly_sample (poly.cu:396-444): lane (x, i) = coefficients 64 i .. 64 i + 63,
// hw(a) + q - hw(b) over SHAKE256(seed || nonce || i, 128)
template <int N>
__global__ void k_sample(int n, Pv<N> r, const uint8_t* seed, size_t stride, int nonce) {
  constexpr int S = Par<N>::STREAMS;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = t / S, i = t % S;
  if (x >= n) return;
  uint64_t s[4], A[25];
#pragma unroll
  for (int w = 0; w < 4; ++w) s[w] = load_u64(seed + (size_t)x * stride + 8 * w);
  absorb_fixed<kRateShake256, 4>(A, s, (uint64_t)nonce | ((uint64_t)i << 8) | (0x1Full << 16));
  uint4* dst = reinterpret_cast<uint4*>(r.el(x) + 64 * i);
#pragma unroll
  for (int w = 0; w < 16; w += 2) {
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t word = A[w + (k >> 1)];
      const uint32_t ab0 = (uint32_t)(word >> (32 * (k & 1))) & 0xFFFFu;
      const uint32_t ab1 = (uint32_t)(word >> (32 * (k & 1) + 16)) & 0xFFFFu;
      const uint32_t c0 = (__popc(ab0 & 0xFF) + Q - __popc(ab0 >> 8)) & 0xFFFFu;
      const uint32_t c1 = (__popc(ab1 & 0xFF) + Q - __popc(ab1 >> 8)) & 0xFFFFu;
      o[k] = c0 | (c1 << 16);
    }
    dst[w >> 1] = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// ---- NTT kernels: one wave per polynomial ------------------------------------------

// One butterfly of ntt on 16-bit values: even layers keep the sum unreduced
// (stored as uint16), odd layers reduce it (MODE 1: Barrett, ntt.cu:195-234 at
// N = 1024; MODE 2: % q, ntt.cu:130-180 at N = 512); the difference is
// multiplied by W in Montgomery form.
template <int MODE>
__device__ __forceinline__ void bfly(uint32_t& a, uint32_t& b, uint32_t W) {
  const uint32_t t = a, h = b;
  a = MODE == 0 ? ((t + h) & 0xFFFFu) : MODE == 1 ? barrett(t + h) : mod_q(t + h);
  b = mont(W * (t + 3 * Q - h));
}

// Layer L (distance 2^L) on register pairs (r, r + 2^RB), RB the register bit
// that holds index bit L; the twiddle of index idx is omega[idx >> (L + 1)].
template <int R, int L, int RB, int ODD, typename IDX>
__device__ __forceinline__ void ntt_layer(uint32_t (&v)[R], const uint16_t* omega, IDX idx) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (r & (1 << RB)) continue;
    const uint32_t W = omega[idx(r) >> (L + 1)];
    bfly<(L & 1) ? ODD : 0>(v[r], v[r + (1 << RB)], W);
  }
}

// ntt_core at N = 1024 on a polynomial that sits in LDS (sm) in ntt input
// order; returns the output in v with v[r] = coefficient 4 l + (r & 3) + 256 (r >> 2).
__device__ __forceinline__ void ntt_wave(uint32_t (&v)[16], uint16_t* sm, int l, const uint16_t* omega) {
  // layers 0-3: v[r] = index 16 l + r
  {
    const uint4 p0 = *reinterpret_cast<const uint4*>(sm + 16 * l), p1 = *reinterpret_cast<const uint4*>(sm + 16 * l + 8);
    const uint32_t w[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      v[2 * k] = w[k] & 0xFFFFu;
      v[2 * k + 1] = w[k] >> 16;
    }
  }
  auto ia = [l](int r) { return 16 * l + r; };
  ntt_layer<16, 0, 0, 1>(v, omega, ia);
  ntt_layer<16, 1, 1, 1>(v, omega, ia);
  ntt_layer<16, 2, 2, 1>(v, omega, ia);
  ntt_layer<16, 3, 3, 1>(v, omega, ia);
  {
    uint32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = v[2 * k] | (v[2 * k + 1] << 16);
    __syncthreads();
    *reinterpret_cast<uint4*>(sm + 16 * l) = make_uint4(w[0], w[1], w[2], w[3]);
    *reinterpret_cast<uint4*>(sm + 16 * l + 8) = make_uint4(w[4], w[5], w[6], w[7]);
    __syncthreads();
  }
  // layers 4-7: v[r] = index (l & 15) + 16 r + 256 (l >> 4)
  auto ib = [l](int r) { return (l & 15) + 16 * r + 256 * (l >> 4); };
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = sm[ib(r)];
  ntt_layer<16, 4, 0, 1>(v, omega, ib);
  ntt_layer<16, 5, 1, 1>(v, omega, ib);
  ntt_layer<16, 6, 2, 1>(v, omega, ib);
  ntt_layer<16, 7, 3, 1>(v, omega, ib);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) sm[ib(r)] = (uint16_t)v[r];
  __syncthreads();
  // layers 8-9: v[r] = index 4 l + (r & 3) + 256 (r >> 2)
  auto ic = [l](int r) { return 4 * l + (r & 3) + 256 * (r >> 2); };
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint2 p = *reinterpret_cast<const uint2*>(sm + 4 * l + 256 * q);
    v[4 * q] = p.x & 0xFFFFu;
    v[4 * q + 1] = p.x >> 16;
    v[4 * q + 2] = p.y & 0xFFFFu;
    v[4 * q + 3] = p.y >> 16;
  }
  ntt_layer<16, 8, 2, 1>(v, omega, ic);
  ntt_layer<16, 9, 3, 1>(v, omega, ic);
}

// ntt_core at N = 512 (nine layers): returns v[r] = coefficient l + 64 r.
__device__ __forceinline__ void ntt_wave(uint32_t (&v)[8], uint16_t* sm, int l, const uint16_t* omega) {
  // layers 0-2: v[r] = index 8 l + r
  {
    const uint4 p = *reinterpret_cast<const uint4*>(sm + 8 * l);
    const uint32_t w[4] = {p.x, p.y, p.z, p.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[2 * k] = w[k] & 0xFFFFu;
      v[2 * k + 1] = w[k] >> 16;
    }
  }
  auto ia = [l](int r) { return 8 * l + r; };
  ntt_layer<8, 0, 0, 2>(v, omega, ia);
  ntt_layer<8, 1, 1, 2>(v, omega, ia);
  ntt_layer<8, 2, 2, 2>(v, omega, ia);
  {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = v[2 * k] | (v[2 * k + 1] << 16);
    __syncthreads();
    *reinterpret_cast<uint4*>(sm + 8 * l) = make_uint4(w[0], w[1], w[2], w[3]);
    __syncthreads();
  }
  // layers 3-5: v[r] = index (l & 7) + 8 r + 64 (l >> 3)
  auto ib = [l](int r) { return (l & 7) + 8 * r + 64 * (l >> 3); };
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = sm[ib(r)];
  ntt_layer<8, 3, 0, 2>(v, omega, ib);
  ntt_layer<8, 4, 1, 2>(v, omega, ib);
  ntt_layer<8, 5, 2, 2>(v, omega, ib);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) sm[ib(r)] = (uint16_t)v[r];
  __syncthreads();
  // layers 6-8: v[r] = index l + 64 r
  auto ic = [l](int r) { return l + 64 * r; };
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = sm[ic(r)];
  ntt_layer<8, 6, 0, 2>(v, omega, ic);
  ntt_layer<8, 7, 1, 2>(v, omega, ic);
  ntt_layer<8, 8, 2, 2>(v, omega, ic);
}

// output position of register r of lane l after ntt_wave
template <int N>
__device__ __forceinline__ int out_index(int l, int r) {
  return N == 1024 ? 4 * l + (r & 3) + 256 * (r >> 2) : l + 64 * r;
}

template <int N>
__device__ __forceinline__ void store_out(const uint32_t (&v)[N / 64], uint16_t* el, int l) {
  if constexpr (N == 1024) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      *reinterpret_cast<uint2*>(el + 4 * l + 256 * q) =
          make_uint2(v[4 * q] | (v[4 * q + 1] << 16), v[4 * q + 2] | (v[4 * q + 3] << 16));
  } else {
#pragma unroll
    for (int r = 0; r < 8; ++r) el[l + 64 * r] = (uint16_t)v[r];
  }
}

// poly_ntt (poly.cu:519-527): mul_coefficients by gammas, then ntt with the same
// table.  Workgroup = one wave = one polynomial; lane l starts with
// coefficients (N / 64) l .. (N / 64) l + N / 64 - 1.
template <int N>
__global__ __launch_bounds__(64) void k_ntt(int n, Pv<N> r) {
  constexpr int C = N / 64;
  __shared__ __align__(16) uint16_t sm[N];
  const int x = blockIdx.x, l = threadIdx.x;
  if (x >= n) return;
  uint16_t* el = r.el(x);
  const uint4* src = reinterpret_cast<const uint4*>(el + C * l);
  const uint4* gam = reinterpret_cast<const uint4*>(tab<N>().gammas + C * l);
#pragma unroll
  for (int h = 0; h < C / 8; ++h) {
    const uint4 a = src[h], g = gam[h];
    const uint32_t aw[4] = {a.x, a.y, a.z, a.w}, gw[4] = {g.x, g.y, g.z, g.w};
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      o[k] = mont((aw[k] & 0xFFFFu) * (gw[k] & 0xFFFFu)) | (mont((aw[k] >> 16) * (gw[k] >> 16)) << 16);
    *reinterpret_cast<uint4*>(sm + C * l + 8 * h) = make_uint4(o[0], o[1], o[2], o[3]);
  }
  __syncthreads();
  uint32_t v[C];
  ntt_wave(v, sm, l, tab<N>().gammas);
  store_out<N>(v, el, l);
}

// poly_invntt (poly.cu:538-547): bitrev_vector, ntt with omegas_inv, then
// mul_coefficients by gammas_inv
template <int N>
__global__ __launch_bounds__(64) void k_invntt(int n, Pv<N> r) {
  constexpr int C = N / 64;
  __shared__ __align__(16) uint16_t sm[N];
  const int x = blockIdx.x, l = threadIdx.x;
  if (x >= n) return;
  uint16_t* el = r.el(x);
  const uint4* src = reinterpret_cast<const uint4*>(el + C * l);
  uint16_t raw[C];
#pragma unroll
  for (int h = 0; h < C / 8; ++h) {
    const uint4 a = src[h];
    const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      raw[8 * h + 2 * k] = (uint16_t)w[k];
      raw[8 * h + 2 * k + 1] = (uint16_t)(w[k] >> 16);
    }
  }
  // the coefficient at ntt input position i is the stored one at brv(i): write
  // stored coefficient c to position brv(c)
#pragma unroll
  for (int k = 0; k < C; ++k) sm[tab<N>().brv[C * l + k]] = raw[k];
  __syncthreads();
  uint32_t v[C];
  ntt_wave(v, sm, l, tab<N>().omegas_inv);
#pragma unroll
  for (int k = 0; k < C; ++k) v[k] = mont(v[k] * tab<N>().gammas_inv[out_index<N>(l, k)]);
  store_out<N>(v, el, l);
}

// ---- flat passes -------------------------------------------------------------------

// poly_mul_pointwise (poly.cu:455-468), 8 coefficients per lane
template <int N>
__global__ void k_pointwise(int n, Pv<N> r, Pv<N> a, Pv<N> b) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * (N / 8)) return;
  const uint4 av = reinterpret_cast<const uint4*>(a.p)[t], bv = reinterpret_cast<const uint4*>(b.p)[t];
  const uint32_t aw[4] = {av.x, av.y, av.z, av.w}, bw[4] = {bv.x, bv.y, bv.z, bv.w};
  uint32_t o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t t0 = mont(3186u * (bw[k] & 0xFFFFu)), t1 = mont(3186u * (bw[k] >> 16));
    o[k] = mont((aw[k] & 0xFFFFu) * t0) | (mont((aw[k] >> 16) * t1) << 16);
  }
  reinterpret_cast<uint4*>(r.p)[t] = make_uint4(o[0], o[1], o[2], o[3]);
}

// poly_add / poly_sub (poly.cu:479-508), 8 coefficients per lane
template <int N>
__global__ void k_addsub(int n, Pv<N> r, Pv<N> a, Pv<N> b, int subtract) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * (N / 8)) return;
  const uint4 av = reinterpret_cast<const uint4*>(a.p)[t], bv = reinterpret_cast<const uint4*>(b.p)[t];
  const uint32_t aw[4] = {av.x, av.y, av.z, av.w}, bw[4] = {bv.x, bv.y, bv.z, bv.w};
  uint32_t o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint32_t h[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const uint32_t x = (aw[k] >> (16 * e)) & 0xFFFFu, y = (bw[k] >> (16 * e)) & 0xFFFFu;
      h[e] = mod_q(subtract ? x + 3 * Q - y : x + y);
    }
    o[k] = h[0] | (h[1] << 16);
  }
  reinterpret_cast<uint4*>(r.p)[t] = make_uint4(o[0], o[1], o[2], o[3]);
}

// ---- (de)serialisation --------------------------------------------------------

// coeff_freeze (poly.cu:21-33)
__device__ __forceinline__ uint32_t freeze(uint32_t x) {
  const uint32_t r = mod_q(x & 0xFFFFu);
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

// poly_tobytes (poly.cu:105-127): lane (x, i) packs coefficients 4 i .. 4 i + 3
// into bytes 7 i .. 7 i + 6; lanes i < 4 also copy seed word i after the
// polynomial (encode_pk)
template <int N>
__global__ void k_tobytes(int n, uint8_t* dst, size_t stride, Pv<N> p, const uint8_t* seed, size_t seed_stride) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * (N / 4)) return;
  const int x = (int)(t / (N / 4)), i = (int)(t % (N / 4));
  const uint2 c = reinterpret_cast<const uint2*>(p.el(x))[i];
  const uint32_t t0 = freeze(c.x & 0xFFFFu), t1 = freeze(c.x >> 16), t2 = freeze(c.y & 0xFFFFu), t3 = freeze(c.y >> 16);
  uint8_t* r = dst + (size_t)x * stride + 7 * i;
  r[0] = (uint8_t)(t0 & 0xff);
  r[1] = (uint8_t)((t0 >> 8) | (t1 << 6));
  r[2] = (uint8_t)(t1 >> 2);
  r[3] = (uint8_t)((t1 >> 10) | (t2 << 4));
  r[4] = (uint8_t)(t2 >> 4);
  r[5] = (uint8_t)((t2 >> 12) | (t3 << 2));
  r[6] = (uint8_t)(t3 >> 6);
  if (seed && i < 4)
    store_u64(dst + (size_t)x * stride + Par<N>::POLYBYTES + 8 * i, load_u64(seed + (size_t)x * seed_stride + 8 * i));
}

// poly_frombytes (poly.cu:62-74): lane (x, i) unpacks bytes 7 i .. 7 i + 6;
// lanes i < 4 copy seed word i out (decode_pk)
template <int N>
__global__ void k_frombytes(int n, Pv<N> r, const uint8_t* src, size_t stride, uint8_t* seed_out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * (N / 4)) return;
  const int x = (int)(t / (N / 4)), i = (int)(t % (N / 4));
  const uint8_t* a = src + (size_t)x * stride + 7 * i;
  const uint32_t c0 = a[0] | (((uint32_t)a[1] & 0x3f) << 8);
  const uint32_t c1 = (a[1] >> 6) | ((uint32_t)a[2] << 2) | (((uint32_t)a[3] & 0x0f) << 10);
  const uint32_t c2 = (a[3] >> 4) | ((uint32_t)a[4] << 4) | (((uint32_t)a[5] & 0x03) << 12);
  const uint32_t c3 = (a[5] >> 2) | ((uint32_t)a[6] << 6);
  reinterpret_cast<uint2*>(r.el(x))[i] = make_uint2((c0 & 0xFFFFu) | (c1 << 16), (c2 & 0xFFFFu) | ((c3 & 0xFFFFu) << 16));
  if (seed_out && i < 4)
    store_u64(seed_out + (size_t)x * 32 + 8 * i, load_u64(src + (size_t)x * stride + Par<N>::POLYBYTES + 8 * i));
}

// poly_compress (poly.cu:165-186): lane (x, g) packs coefficients 8 g .. 8 g + 7
// into ct + POLYBYTES + 3 g
template <int N>
__global__ void k_compress(int n, uint8_t* ct, size_t ct_stride, Pv<N> p) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * (N / 8)) return;
  const int x = (int)(t / (N / 8)), g = (int)(t % (N / 8));
  const uint4 c = reinterpret_cast<const uint4*>(p.el(x))[g];
  const uint32_t cw[4] = {c.x, c.y, c.z, c.w};
  uint32_t v[8];
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = div_q((freeze((cw[j >> 1] >> (16 * (j & 1))) & 0xFFFFu) << 3) + Q / 2) & 7;
  uint8_t* r = ct + (size_t)x * ct_stride + Par<N>::POLYBYTES + 3 * g;
  r[0] = (uint8_t)(v[0] | (v[1] << 3) | (v[2] << 6));
  r[1] = (uint8_t)((v[2] >> 2) | (v[3] << 1) | (v[4] << 4) | (v[5] << 7));
  r[2] = (uint8_t)((v[5] >> 1) | (v[6] << 2) | (v[7] << 5));
}

// poly_decompress (poly.cu:197-216): lane (x, g) = coefficients 8 g .. 8 g + 7
template <int N>
__global__ void k_decompress(int n, Pv<N> r, const uint8_t* ct, size_t ct_stride) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * (N / 8)) return;
  const int x = (int)(t / (N / 8)), g = (int)(t % (N / 8));
  const uint8_t* a = ct + (size_t)x * ct_stride + Par<N>::POLYBYTES + 3 * g;
  const uint32_t a0 = a[0], a1 = a[1], a2 = a[2];
  uint32_t v[8];
  v[0] = a0 & 7;
  v[1] = (a0 >> 3) & 7;
  v[2] = (a0 >> 6) | ((a1 << 2) & 4);
  v[3] = (a1 >> 1) & 7;
  v[4] = (a1 >> 4) & 7;
  v[5] = (a1 >> 7) | ((a2 << 1) & 6);
  v[6] = (a2 >> 2) & 7;
  v[7] = (a2 >> 5);
  uint32_t o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    o[k] = (((v[2 * k] * Q + 4) >> 3) & 0xFFFFu) | ((((v[2 * k + 1] * Q + 4) >> 3) & 0xFFFFu) << 16);
  reinterpret_cast<uint4*>(r.el(x))[g] = make_uint4(o[0], o[1], o[2], o[3]);
}

// poly_frommsg_n (poly.cu:226-251): lane (x, i) spreads message byte i over
// coefficients 8 i + j + 256 q, q < N / 256
template <int N>
__global__ void k_frommsg(int n, Pv<N> r, const uint8_t* msg, size_t msg_stride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 32) return;
  const int x = t >> 5, i = t & 31;
  const uint32_t b = msg[(size_t)x * msg_stride + i];
  uint32_t o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t v0 = (0u - ((b >> (2 * k)) & 1)) & (Q / 2), v1 = (0u - ((b >> (2 * k + 1)) & 1)) & (Q / 2);
    o[k] = v0 | (v1 << 16);
  }
#pragma unroll
  for (int q = 0; q < Par<N>::COPIES; ++q)
    reinterpret_cast<uint4*>(r.el(x) + 8 * i + 256 * q)[0] = make_uint4(o[0], o[1], o[2], o[3]);
}

// poly_tomsg_n (poly.cu:261-291): lane (x, i) decodes message byte i from the
// sum of flipabs over the N / 256 copies, compared with q (N = 1024) or q / 2
template <int N>
__global__ void k_tomsg(int n, uint8_t* msg, size_t msg_stride, Pv<N> p) {
  constexpr int CP = Par<N>::COPIES;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 32) return;
  const int x = t >> 5, i = t & 31;
  uint32_t c[CP][8];
#pragma unroll
  for (int q = 0; q < CP; ++q) {
    const uint4 w = reinterpret_cast<const uint4*>(p.el(x) + 8 * i + 256 * q)[0];
    const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) c[q][j] = (ww[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
  }
  uint32_t byte = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    uint32_t v = 0;
#pragma unroll
    for (int q = 0; q < CP; ++q) v += flipabs(c[q][j]);
    v = (v - (N == 1024 ? Q : Q / 2)) & 0xFFFFu;
    byte |= (v >> 15) << j;
  }
  msg[(size_t)x * msg_stride + i] = (uint8_t)byte;
}

// ---- fused per-element kernels: one wave per element ------------------------------
// Keypair, encryption and decryption each as one kernel.  The element's SHAKE
// streams run on the wave's lanes (one stream per lane), its polynomials stay in
// LDS or in VGPRs in ntt_wave's output layout, and HBM sees only the element's
// inputs and outputs.  Every value is produced by the same operations, in the
// same order and with the same 16-bit storage points as the per-operation
// kernels above, so both paths give identical outputs.

// gen_a rows in LDS: 64 coefficients of one stream, padded to 66 so that the
// lanes' scattered accept stores fall on different banks
constexpr int kArow = 66;

__device__ __forceinline__ int arow_index(int c) { return (c >> 6) * kArow + (c & 63); }

template <int N>
__device__ __forceinline__ int brv_index(int c) {
  return (int)(__builtin_bitreverse32((uint32_t)c) >> (32 - Par<N>::LOG));
}

// One SHAKE stream on this lane: gen_a stream i (SHAKE128(seed || i), 16-bit
// candidates < 5q kept unreduced, poly.cu:332-366) when `uniform`, else
// poly_sample stream i (SHAKE256(seed || nonce || i, 128 bytes), hw(a) + q -
// hw(b), poly.cu:396-444).  Writes the stream's 64 coefficients to row[0..63].
// Lanes that take no part pass active = false; every lane of the wave reaches
// the block loop's vote.
__device__ __forceinline__ void lane_stream(uint16_t* row, const uint64_t (&seed)[4], bool active, bool uniform,
                                            int nonce, int i) {
  uint64_t A[25];
#pragma unroll
  for (int w = 0; w < 25; ++w) A[w] = w < 4 ? seed[w] : 0;
  A[4] = uniform ? ((uint64_t)i | (0x1Full << 8)) : ((uint64_t)nonce | ((uint64_t)i << 8) | (0x1Full << 16));
  A[kRateShake128 - 1] ^= uniform ? 0x8000000000000000ull : 0;
  A[kRateShake256 - 1] ^= uniform ? 0 : 0x8000000000000000ull;
  keccakf1600(A);
  if (active && !uniform) {
#pragma unroll
    for (int w = 0; w < 16; w += 2) {
      uint32_t o[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint64_t word = A[w + (k >> 1)];
        const uint32_t ab0 = (uint32_t)(word >> (32 * (k & 1))) & 0xFFFFu;
        const uint32_t ab1 = (uint32_t)(word >> (32 * (k & 1) + 16)) & 0xFFFFu;
        const uint32_t c0 = (__popc(ab0 & 0xFF) + Q - __popc(ab0 >> 8)) & 0xFFFFu;
        const uint32_t c1 = (__popc(ab1 & 0xFF) + Q - __popc(ab1 >> 8)) & 0xFFFFu;
        o[k] = c0 | (c1 << 16);
      }
      reinterpret_cast<uint4*>(row)[w >> 1] = make_uint4(o[0], o[1], o[2], o[3]);
    }
  }
  int ctr = (active && uniform) ? 0 : 64;
  for (int blk = 0; blk < 16; ++blk) {
    if (blk > 0) {
      if (!__any(ctr < 64)) break;
      keccakf1600(A);
    }
    if (ctr < 64) {
#pragma unroll
      for (int w = 0; w < kRateShake128; ++w) {
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          const uint32_t val = (uint32_t)(A[w] >> (16 * h)) & 0xFFFFu;
          if (val < 5 * Q && ctr < 64) row[ctr++] = (uint16_t)val;
        }
      }
    }
  }
}

// poly_ntt of the natural-order polynomial in sm (clobbered): v[r] = output
// coefficient out_index(l, r), as k_ntt
template <int N>
__device__ __forceinline__ void ntt_lds(uint32_t (&v)[N / 64], uint16_t* sm, int l) {
  constexpr int C = N / 64;
  uint4* p = reinterpret_cast<uint4*>(sm + C * l);
  const uint4* gam = reinterpret_cast<const uint4*>(tab<N>().gammas + C * l);
#pragma unroll
  for (int h = 0; h < C / 8; ++h) {
    const uint4 a = p[h], g = gam[h];
    const uint32_t aw[4] = {a.x, a.y, a.z, a.w}, gw[4] = {g.x, g.y, g.z, g.w};
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      o[k] = mont((aw[k] & 0xFFFFu) * (gw[k] & 0xFFFFu)) | (mont((aw[k] >> 16) * (gw[k] >> 16)) << 16);
    p[h] = make_uint4(o[0], o[1], o[2], o[3]);
  }
  __syncthreads();
  ntt_wave(v, sm, l, tab<N>().gammas);
}

// poly_invntt of the polynomial held as v[r] = coefficient out_index(l, r), in
// place, through sm (clobbered), as k_invntt
template <int N>
__device__ __forceinline__ void invntt_regs(uint32_t (&v)[N / 64], uint16_t* sm, int l) {
  __syncthreads();
#pragma unroll
  for (int r = 0; r < N / 64; ++r) sm[brv_index<N>(out_index<N>(l, r))] = (uint16_t)v[r];
  __syncthreads();
  ntt_wave(v, sm, l, tab<N>().omegas_inv);
#pragma unroll
  for (int k = 0; k < N / 64; ++k) v[k] = mont(v[k] * tab<N>().gammas_inv[out_index<N>(l, k)]);
}

// natural-order LDS polynomial <-> registers in ntt_wave's output layout
template <int N>
__device__ __forceinline__ void put_nat(uint16_t* sm, const uint32_t (&v)[N / 64], int l) {
#pragma unroll
  for (int r = 0; r < N / 64; ++r) sm[out_index<N>(l, r)] = (uint16_t)v[r];
}

// poly_frombytes of the 14-bit packing at src (8-byte aligned) into sm in
// natural order: group g (coefficients 4 g .. 4 g + 3) is bits 0..55 of the 7
// bytes at 7 g, read as one or two aligned words (never a word past the
// polynomial: only groups that straddle a word boundary read the next one)
template <int N>
__device__ __forceinline__ void unpack14(uint16_t* sm, const uint8_t* src, int l) {
  const uint64_t* s64 = reinterpret_cast<const uint64_t*>(src);
#pragma unroll
  for (int g = l; g < N / 4; g += 64) {
    const uint32_t bo = 7u * (uint32_t)g, sh = 8u * (bo & 7u);
    const uint64_t lo = s64[bo >> 3];
    const uint64_t hi = (bo & 7u) > 1u ? s64[(bo >> 3) + 1] : 0;
    const uint64_t w = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
    const uint32_t c0 = (uint32_t)w & 0x3FFFu, c1 = (uint32_t)(w >> 14) & 0x3FFFu;
    const uint32_t c2 = (uint32_t)(w >> 28) & 0x3FFFu, c3 = (uint32_t)(w >> 42) & 0x3FFFu;
    reinterpret_cast<uint2*>(sm)[g] = make_uint2(c0 | (c1 << 16), c2 | (c3 << 16));
  }
}

// poly_tobytes of the natural-order polynomial in sm into LDS bytes bb
template <int N>
__device__ __forceinline__ void pack14(uint8_t* bb, const uint16_t* sm, int l) {
#pragma unroll
  for (int g = l; g < N / 4; g += 64) {
    const uint2 c = reinterpret_cast<const uint2*>(sm)[g];
    const uint32_t t0 = freeze(c.x & 0xFFFFu), t1 = freeze(c.x >> 16), t2 = freeze(c.y & 0xFFFFu),
                   t3 = freeze(c.y >> 16);
    uint8_t* r = bb + 7 * g;
    r[0] = (uint8_t)(t0 & 0xff);
    r[1] = (uint8_t)((t0 >> 8) | (t1 << 6));
    r[2] = (uint8_t)(t1 >> 2);
    r[3] = (uint8_t)((t1 >> 10) | (t2 << 4));
    r[4] = (uint8_t)(t2 >> 4);
    r[5] = (uint8_t)((t2 >> 12) | (t3 << 2));
    r[6] = (uint8_t)(t3 >> 6);
  }
}

// poly_compress of the natural-order polynomial in sm into LDS bytes bb
template <int N>
__device__ __forceinline__ void compress3(uint8_t* bb, const uint16_t* sm, int l) {
#pragma unroll
  for (int g = l; g < N / 8; g += 64) {
    const uint4 c = reinterpret_cast<const uint4*>(sm)[g];
    const uint32_t cw[4] = {c.x, c.y, c.z, c.w};
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = div_q((freeze((cw[j >> 1] >> (16 * (j & 1))) & 0xFFFFu) << 3) + Q / 2) & 7;
    uint8_t* r = bb + 3 * g;
    r[0] = (uint8_t)(v[0] | (v[1] << 3) | (v[2] << 6));
    r[1] = (uint8_t)((v[2] >> 2) | (v[3] << 1) | (v[4] << 4) | (v[5] << 7));
    r[2] = (uint8_t)((v[5] >> 1) | (v[6] << 2) | (v[7] << 5));
  }
}

// poly_decompress of the 3-bit packing at src into sm in natural order
template <int N>
__device__ __forceinline__ void decompress3(uint16_t* sm, const uint8_t* src, int l) {
#pragma unroll
  for (int g = l; g < N / 8; g += 64) {
    const uint8_t* a = src + 3 * g;
    const uint32_t a0 = a[0], a1 = a[1], a2 = a[2];
    uint32_t v[8];
    v[0] = a0 & 7;
    v[1] = (a0 >> 3) & 7;
    v[2] = (a0 >> 6) | ((a1 << 2) & 4);
    v[3] = (a1 >> 1) & 7;
    v[4] = (a1 >> 4) & 7;
    v[5] = (a1 >> 7) | ((a2 << 1) & 6);
    v[6] = (a2 >> 2) & 7;
    v[7] = (a2 >> 5);
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      o[k] = (((v[2 * k] * Q + 4) >> 3) & 0xFFFFu) | ((((v[2 * k + 1] * Q + 4) >> 3) & 0xFFFFu) << 16);
    reinterpret_cast<uint4*>(sm)[g] = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// `bytes` (a multiple of 8) from LDS to global memory (8-byte aligned), one
// word per lane per step
__device__ __forceinline__ void copy_out(uint8_t* dst, const uint8_t* bb, int bytes, int l) {
  for (int w = l; w < bytes / 8; w += 64) store_u64(dst + 8 * w, reinterpret_cast<const uint64_t*>(bb)[w]);
}

// cpapke_keypair (cpapke.cu:265-338) for element blockIdx.x, from z =
// publicseed || noiseseed = SHAKE256(seed, 64) at z + 64 x (k_seed_expand: the
// one serial permutation stays on one lane per element).  Lanes 0..S-1 run the
// gen_a streams, S..2S-1 the streams of s (nonce 0), 2S..3S-1 those of e
// (nonce 1); S = N / 64.
template <int N>
__global__ __launch_bounds__(64) void k_keypair_fused(int n, uint8_t* pk, uint8_t* sk, size_t sk_stride,
                                                      const uint8_t* z) {
  using P = Par<N>;
  constexpr int C = N / 64, S = P::STREAMS;
  __shared__ __align__(16) uint16_t sa[S * kArow];
  __shared__ __align__(16) uint16_t ss[N];
  __shared__ __align__(16) uint16_t se[N];
  const int x = blockIdx.x, l = threadIdx.x;
  if (x >= n) return;
  const uint8_t* zx = z + (size_t)x * 64;
  {
    const int kind = l / S, i = l % S;
    const bool pub = kind == 0;
    uint64_t sd[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) sd[w] = load_u64(zx + (pub ? 0 : 32) + 8 * w);
    uint16_t* row = pub ? sa + i * kArow : (kind == 1 ? ss : se) + 64 * i;
    lane_stream(row, sd, kind < 3, pub, kind - 1, i);
  }
  __syncthreads();
  uint32_t vs[C], vb[C];
  ntt_lds<N>(vs, ss, l);
  ntt_lds<N>(vb, se, l);
  // b = e + s o a (k_pointwise(AS, SHAT, AHAT), k_addsub(BHAT, EHAT, AS))
#pragma unroll
  for (int r = 0; r < C; ++r) {
    const uint32_t ahat = sa[arow_index(out_index<N>(l, r))];
    vb[r] = mod_q(vb[r] + mont(vs[r] * mont(3186u * ahat)));
  }
  __syncthreads();
  put_nat<N>(ss, vs, l);
  put_nat<N>(se, vb, l);
  __syncthreads();
  pack14<N>(reinterpret_cast<uint8_t*>(sa), ss, l);
  __syncthreads();
  copy_out(sk + (size_t)x * sk_stride, reinterpret_cast<const uint8_t*>(sa), P::POLYBYTES, l);
  __syncthreads();
  pack14<N>(reinterpret_cast<uint8_t*>(sa), se, l);
  __syncthreads();
  uint8_t* pkx = pk + (size_t)x * P::PK;
  copy_out(pkx, reinterpret_cast<const uint8_t*>(sa), P::POLYBYTES, l);
  if (l < 4) store_u64(pkx + P::POLYBYTES + 8 * l, load_u64(zx + 8 * l));
}

// cpapke_enc (cpapke.cu:353-429) for element blockIdx.x.  Lanes 0..S-1 run the
// gen_a streams of the public seed, S..4S-1 the streams of s', e', e'' (coin,
// nonces 0, 1, 2).
template <int N>
__global__ __launch_bounds__(64) void k_enc_fused(int n, uint8_t* ct, size_t ct_stride, const uint8_t* m,
                                                  size_t m_stride, const uint8_t* pk, size_t pk_stride,
                                                  const uint8_t* coin, size_t coin_stride) {
  using P = Par<N>;
  constexpr int C = N / 64, S = P::STREAMS;
  __shared__ __align__(16) uint16_t sa[S * kArow];
  __shared__ __align__(16) uint16_t s1[N];
  __shared__ __align__(16) uint16_t s2[N];
  __shared__ __align__(16) uint16_t s3[N];
  const int x = blockIdx.x, l = threadIdx.x;
  if (x >= n) return;
  const uint8_t* pkx = pk + (size_t)x * pk_stride;
  {
    const int kind = l / S, i = l % S;
    const bool pub = kind == 0;
    const uint8_t* src = pub ? pkx + P::POLYBYTES : coin + (size_t)x * coin_stride;
    uint64_t sd[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) sd[w] = load_u64(src + 8 * w);
    uint16_t* row = pub ? sa + i * kArow : (kind == 1 ? s1 : kind == 2 ? s2 : s3) + 64 * i;
    lane_stream(row, sd, kind < 4, pub, kind - 1, i);
  }
  __syncthreads();
  uint32_t vsp[C], vv[C];
  ntt_lds<N>(vsp, s1, l);
  {
    // u = a o s' + e' (k_pointwise(UHAT, AHAT, SP), k_addsub(UHAT, UHAT, EP)),
    // parked in natural order over the gen_a rows once every lane has read them
    uint32_t vu[C];
    ntt_lds<N>(vu, s2, l);
#pragma unroll
    for (int r = 0; r < C; ++r) {
      const uint32_t ahat = sa[arow_index(out_index<N>(l, r))];
      vu[r] = mod_q(mont(ahat * mont(3186u * vsp[r])) + vu[r]);
    }
    __syncthreads();
    put_nat<N>(sa, vu, l);
  }
  unpack14<N>(s1, pkx, l);  // b (decode_pk)
  __syncthreads();
  // v' = invntt(b o s') (k_pointwise(VP, BHAT, SP), k_invntt)
#pragma unroll
  for (int r = 0; r < C; ++r) vv[r] = mont((uint32_t)s1[out_index<N>(l, r)] * mont(3186u * vsp[r]));
  invntt_regs<N>(vv, s2, l);
  // v = v' + e'' + frommsg(m) (k_addsub twice, k_frommsg)
  uint64_t mw[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) mw[w] = load_u64(m + (size_t)x * m_stride + 8 * w);
#pragma unroll
  for (int r = 0; r < C; ++r) {
    const int c = out_index<N>(l, r), b = c & 255;
    const uint64_t word = (b >> 6) == 0 ? mw[0] : (b >> 6) == 1 ? mw[1] : (b >> 6) == 2 ? mw[2] : mw[3];
    const uint32_t bit = (uint32_t)(word >> (b & 63)) & 1u;
    const uint32_t t = mod_q(vv[r] + s3[c]);
    vv[r] = mod_q(t + ((0u - bit) & (Q / 2)));
  }
  __syncthreads();
  put_nat<N>(s3, vv, l);
  __syncthreads();
  pack14<N>(reinterpret_cast<uint8_t*>(s2), sa, l);
  compress3<N>(reinterpret_cast<uint8_t*>(s1), s3, l);
  __syncthreads();
  uint8_t* ctx_ = ct + (size_t)x * ct_stride;
  copy_out(ctx_, reinterpret_cast<const uint8_t*>(s2), P::POLYBYTES, l);
  copy_out(ctx_ + P::POLYBYTES, reinterpret_cast<const uint8_t*>(s1), P::CBYTES, l);
}

// cpapke_dec (cpapke.cu:443-476) for element blockIdx.x
template <int N>
__global__ __launch_bounds__(64) void k_dec_fused(int n, uint8_t* m, size_t m_stride, const uint8_t* ct,
                                                  size_t ct_stride, const uint8_t* sk, size_t sk_stride) {
  using P = Par<N>;
  constexpr int C = N / 64;
  __shared__ __align__(16) uint16_t s1[N];
  __shared__ __align__(16) uint16_t s2[N];
  __shared__ __align__(16) uint16_t s3[N];
  const int x = blockIdx.x, l = threadIdx.x;
  if (x >= n) return;
  const uint8_t* ctx_ = ct + (size_t)x * ct_stride;
  unpack14<N>(s1, sk + (size_t)x * sk_stride, l);
  unpack14<N>(s2, ctx_, l);
  decompress3<N>(s3, ctx_ + P::POLYBYTES, l);
  __syncthreads();
  // tmp = invntt(s o u) - v (k_pointwise(TMP, SHAT, UHAT), k_invntt, k_addsub(TMP, TMP, VP, 1))
  uint32_t vt[C];
#pragma unroll
  for (int r = 0; r < C; ++r) {
    const int c = out_index<N>(l, r);
    vt[r] = mont((uint32_t)s1[c] * mont(3186u * (uint32_t)s2[c]));
  }
  invntt_regs<N>(vt, s1, l);
#pragma unroll
  for (int r = 0; r < C; ++r) vt[r] = mod_q(vt[r] + 3 * Q - (uint32_t)s3[out_index<N>(l, r)]);
  __syncthreads();
  put_nat<N>(s2, vt, l);
  __syncthreads();
  // poly_tomsg: byte i from the N / 256 copies of coefficients 8 i + j
  if (l < 32) {
    uint32_t byte = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      uint32_t v = 0;
#pragma unroll
      for (int q = 0; q < Par<N>::COPIES; ++q) v += flipabs(s2[8 * l + j + 256 * q]);
      v = (v - (N == 1024 ? Q : Q / 2)) & 0xFFFFu;
      byte |= (v >> 15) << j;
    }
    m[(size_t)x * m_stride + l] = (uint8_t)byte;
  }
}

// ---- CCA-KEM glue: one lane per element for the sponges, flat copies -------------

// out[x] = SHAKE256(in[x], 8 words bytes), first 8 OW bytes
template <int OW>
__global__ void k_shake256_words(int n, uint8_t* out, size_t out_stride, const uint8_t* in, size_t in_stride,
                                 int words) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t A[25];
  sponge_absorb_words<kRateShake256>(A, reinterpret_cast<const uint64_t*>(in + (size_t)x * in_stride), words, 0x1F);
#pragma unroll
  for (int w = 0; w < OW; ++w) store_u64(out + (size_t)x * out_stride + 8 * w, A[w]);
}

// The same on a lane pair per element (lane 2x + h holds half h of every state
// word, keccakf1600_pair): the latency form for the long serial absorbs of a
// public key or a ciphertext (14 to 17 Keccak-f), with twice the lanes at work
template <int OW>
__global__ void k_shake256_pair(int n, uint8_t* out, size_t out_stride, const uint8_t* in, size_t in_stride,
                                int words) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = t >> 1;
  const uint32_t half = (uint32_t)(t & 1);
  if (x >= n) return;  // both lanes of a pair leave together (2n lanes, even)
  uint32_t S[25];
  absorb_range_pair(S, in + (size_t)x * in_stride, words, 0, words / kRateShake256 + 1, 0x1F, half);
  uint32_t* o = reinterpret_cast<uint32_t*>(out + (size_t)x * out_stride);
#pragma unroll
  for (int w = 0; w < OW; ++w) o[2 * w + half] = S[w];
}

// out[x] = SHAKE256(a[x] (32 bytes) || b[x] (32 bytes)), first 8 OW bytes
template <int OW>
__global__ void k_shake256_2x32(int n, uint8_t* out, size_t out_stride, const uint8_t* a, size_t a_stride,
                                const uint8_t* b, size_t b_stride) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t m[8], A[25];
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    m[w] = load_u64(a + (size_t)x * a_stride + 8 * w);
    m[4 + w] = load_u64(b + (size_t)x * b_stride + 8 * w);
  }
  absorb_fixed<kRateShake256, 8>(A, m, 0x1F);
#pragma unroll
  for (int w = 0; w < OW; ++w) store_u64(out + (size_t)x * out_stride + 8 * w, A[w]);
}

// dst[x][0 .. 8 words) = src[x][0 .. 8 words), one lane per word
__global__ void k_copy_words(int n, uint8_t* dst, size_t dst_stride, const uint8_t* src, size_t src_stride,
                             int words) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * words) return;
  const size_t x = t / words, w = t % words;
  store_u64(dst + x * dst_stride + 8 * w, load_u64(src + x * src_stride + 8 * w));
}

// fail[x] = 1 if a[x] and b[x] differ anywhere in `words` words, else 0: one wave
// per element, no early exit (verify.cu:15-28 without its data-dependent branch)
__global__ __launch_bounds__(64) void k_verify_words(int n, uint32_t* fail, size_t fail_stride, const uint8_t* a,
                                                    size_t a_stride, const uint8_t* b, size_t b_stride, int words) {
  const int x = blockIdx.x;
  if (x >= n) return;
  const uint64_t* pa = reinterpret_cast<const uint64_t*>(a + (size_t)x * a_stride);
  const uint64_t* pb = reinterpret_cast<const uint64_t*>(b + (size_t)x * b_stride);
  uint64_t d = 0;
  for (int w = threadIdx.x; w < words; w += 64) d |= pa[w] ^ pb[w];
  const bool any = __any(d != 0);
  if (threadIdx.x == 0) *reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(fail) + (size_t)x * fail_stride) = any;
}

// k[x] = fail ? z[x] : k[x], 32 bytes, by mask (verify.cu:40-50 cmov)
__global__ void k_cmov32(int n, uint8_t* k, size_t k_stride, const uint8_t* z, size_t z_stride, const uint32_t* fail,
                         size_t fail_stride) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * 4) return;
  const size_t x = t >> 2, w = t & 3;
  const uint64_t f = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const uint8_t*>(fail) + x * fail_stride);
  const uint64_t mask = 0ull - (f & 1);
  uint8_t* pk = k + x * k_stride + 8 * w;
  store_u64(pk, (load_u64(pk) & ~mask) | (load_u64(z + x * z_stride + 8 * w) & mask));
}

// ---- orchestration (cpapke.cu:265-476) -------------------------------------------

template <int N>
struct Scratch {
  pqc_ctx* c;
  Pv<N> poly(int i) const { return Pv<N>{c->nhpolys + (size_t)i * N * c->ld}; }
};

// grid for `threads` lanes at the tuned block size
struct Launch {
  int block;
  dim3 grid(size_t threads) const { return dim3((unsigned)((threads + block - 1) / block)); }
};
#define LAUNCH(kern, threads, ...) hipLaunchKernelGGL(kern, L.grid(threads), dim3(L.block), 0, s, __VA_ARGS__)
#define LAUNCH_NTT(kern, ...) hipLaunchKernelGGL(kern, dim3(n), dim3(64), 0, s, __VA_ARGS__)

// cpapke_keypair (cpapke.cu:265-338) with pk at stride PK and sk at sk_stride
// (the CCA-KEM's sk holds sk_cpa first), seeds at seed_stride
template <int N>
int keypair_impl(pqc_ctx* c, int n, uint8_t* pk, uint8_t* sk, size_t sk_stride, const uint8_t* seed,
                 size_t seed_stride, bool plain, hipStream_t s) {
  const Launch L{c->tune.newhope_block};
  if (!plain) {
    const size_t n1 = n;
    LAUNCH(k_seed_expand, n1, n, c->nhbytes, seed, seed_stride);
    hipLaunchKernelGGL(k_keypair_fused<N>, dim3((unsigned)n), dim3(64), 0, s, n, pk, sk, sk_stride,
                       (const uint8_t*)c->nhbytes);
    return hip_err(hipGetLastError());
  }
  const size_t n1 = n, nS = (size_t)Par<N>::STREAMS * n, n8 = (size_t)n * (N / 8), n4 = (size_t)n * (N / 4);
  Scratch<N> S{c};
  uint8_t* z = c->nhbytes;  // 64 bytes per element: publicseed || noiseseed
  const int AHAT = 0, EHAT = 1, AS = 2, BHAT = 3, SHAT = 4;
  LAUNCH(k_seed_expand, n1, n, z, seed, seed_stride);
  hipLaunchKernelGGL(k_uniform<N>, dim3((unsigned)((nS + kUniformBlock - 1) / kUniformBlock)), dim3(kUniformBlock), 0, s,
                     n, S.poly(AHAT), (const uint8_t*)z, (size_t)64);
  LAUNCH(k_sample<N>, nS, n, S.poly(SHAT), (const uint8_t*)z + 32, (size_t)64, 0);
  LAUNCH_NTT(k_ntt<N>, n, S.poly(SHAT));
  LAUNCH(k_sample<N>, nS, n, S.poly(EHAT), (const uint8_t*)z + 32, (size_t)64, 1);
  LAUNCH_NTT(k_ntt<N>, n, S.poly(EHAT));
  LAUNCH(k_pointwise<N>, n8, n, S.poly(AS), S.poly(SHAT), S.poly(AHAT));
  LAUNCH(k_addsub<N>, n8, n, S.poly(BHAT), S.poly(EHAT), S.poly(AS), 0);
  LAUNCH(k_tobytes<N>, n4, n, sk, sk_stride, S.poly(SHAT), (const uint8_t*)nullptr, (size_t)0);
  LAUNCH(k_tobytes<N>, n4, n, pk, (size_t)Par<N>::PK, S.poly(BHAT), (const uint8_t*)z, (size_t)64);
  return hip_err(hipGetLastError());
}

// cpapke_enc (cpapke.cu:353-429), every buffer at its own stride
template <int N>
int enc_impl(pqc_ctx* c, int n, uint8_t* ct, size_t ct_stride, const uint8_t* m, size_t m_stride, const uint8_t* pk,
             size_t pk_stride, const uint8_t* coin, size_t coin_stride, bool plain, hipStream_t s) {
  if (!plain) {
    hipLaunchKernelGGL(k_enc_fused<N>, dim3((unsigned)n), dim3(64), 0, s, n, ct, ct_stride, m, m_stride, pk, pk_stride,
                       coin, coin_stride);
    return hip_err(hipGetLastError());
  }
  const Launch L{c->tune.newhope_block};
  const size_t nS = (size_t)Par<N>::STREAMS * n, n32 = 32 * (size_t)n, n8 = (size_t)n * (N / 8), n4 = (size_t)n * (N / 4);
  Scratch<N> S{c};
  uint8_t* seed = c->nhbytes;  // publicseed copy, 32 bytes per element
  const int SP = 0, EP = 1, VP = 2, AHAT = 3, BHAT = 4, EPP = 5, UHAT = 6, V = 7;
  LAUNCH(k_frommsg<N>, n32, n, S.poly(V), m, m_stride);
  LAUNCH(k_frombytes<N>, n4, n, S.poly(BHAT), pk, pk_stride, seed);
  hipLaunchKernelGGL(k_uniform<N>, dim3((unsigned)((nS + kUniformBlock - 1) / kUniformBlock)), dim3(kUniformBlock), 0, s,
                     n, S.poly(AHAT), (const uint8_t*)seed, (size_t)32);
  LAUNCH(k_sample<N>, nS, n, S.poly(SP), coin, coin_stride, 0);
  LAUNCH(k_sample<N>, nS, n, S.poly(EP), coin, coin_stride, 1);
  LAUNCH(k_sample<N>, nS, n, S.poly(EPP), coin, coin_stride, 2);
  LAUNCH_NTT(k_ntt<N>, n, S.poly(SP));
  LAUNCH_NTT(k_ntt<N>, n, S.poly(EP));
  LAUNCH(k_pointwise<N>, n8, n, S.poly(UHAT), S.poly(AHAT), S.poly(SP));
  LAUNCH(k_addsub<N>, n8, n, S.poly(UHAT), S.poly(UHAT), S.poly(EP), 0);
  LAUNCH(k_pointwise<N>, n8, n, S.poly(VP), S.poly(BHAT), S.poly(SP));
  LAUNCH_NTT(k_invntt<N>, n, S.poly(VP));
  LAUNCH(k_addsub<N>, n8, n, S.poly(VP), S.poly(VP), S.poly(EPP), 0);
  LAUNCH(k_addsub<N>, n8, n, S.poly(VP), S.poly(VP), S.poly(V), 0);
  LAUNCH(k_tobytes<N>, n4, n, ct, ct_stride, S.poly(UHAT), (const uint8_t*)nullptr, (size_t)0);
  LAUNCH(k_compress<N>, n8, n, ct, ct_stride, S.poly(VP));
  return hip_err(hipGetLastError());
}

// cpapke_dec (cpapke.cu:443-476), every buffer at its own stride
template <int N>
int dec_impl(pqc_ctx* c, int n, uint8_t* m, size_t m_stride, const uint8_t* ct, size_t ct_stride, const uint8_t* sk,
             size_t sk_stride, bool plain, hipStream_t s) {
  if (!plain) {
    hipLaunchKernelGGL(k_dec_fused<N>, dim3((unsigned)n), dim3(64), 0, s, n, m, m_stride, ct, ct_stride, sk, sk_stride);
    return hip_err(hipGetLastError());
  }
  const Launch L{c->tune.newhope_block};
  const size_t n32 = 32 * (size_t)n, n8 = (size_t)n * (N / 8), n4 = (size_t)n * (N / 4);
  Scratch<N> S{c};
  const int VP = 0, UHAT = 1, TMP = 2, SHAT = 3;
  LAUNCH(k_frombytes<N>, n4, n, S.poly(SHAT), sk, sk_stride, (uint8_t*)nullptr);
  LAUNCH(k_frombytes<N>, n4, n, S.poly(UHAT), ct, ct_stride, (uint8_t*)nullptr);
  LAUNCH(k_decompress<N>, n8, n, S.poly(VP), ct, ct_stride);
  LAUNCH(k_pointwise<N>, n8, n, S.poly(TMP), S.poly(SHAT), S.poly(UHAT));
  LAUNCH_NTT(k_invntt<N>, n, S.poly(TMP));
  LAUNCH(k_addsub<N>, n8, n, S.poly(TMP), S.poly(TMP), S.poly(VP), 1);
  LAUNCH(k_tomsg<N>, n32, n, m, m_stride, S.poly(TMP));
  return hip_err(hipGetLastError());
}

// Per-element CCA scratch row (one size for both N): mu or m' | H(pk) |
// K || coin || d | H(c) | fail | re-encrypted ct
constexpr int KM_MU = 0, KM_HPK = 32, KM_KCD = 64, KM_HC = 160, KM_FAIL = 192, KM_CT2 = 224;
static_assert(KM_CT2 + Par<1024>::CCA_CT == kNewhopeKemBytes, "scratch row size");

// NewHope CCA-KEM, the NewHope specification's transform over the CPA-PKE
// (SHAKE256 for every hash).  Keypair coins: CPA seed || s, 64 bytes.
template <int N>
int cca_keypair(pqc_ctx* c, size_t n_, uint8_t* pk, uint8_t* sk, const uint8_t* coins, bool plain, hipStream_t s) {
  using P = Par<N>;
  const Launch L{c->tune.newhope_block};
  const int n = (int)n_;
  int rc = keypair_impl<N>(c, n, pk, sk, P::CCA_SK, coins, 64, plain, s);
  if (rc) return rc;
  LAUNCH(k_copy_words, n_ * (P::PK / 8), n, sk + P::SK, (size_t)P::CCA_SK, (const uint8_t*)pk, (size_t)P::PK, P::PK / 8);
  LAUNCH(k_shake256_pair<4>, 2 * n_, n, sk + P::SK + P::PK, (size_t)P::CCA_SK, (const uint8_t*)pk, (size_t)P::PK,
         P::PK / 8);
  LAUNCH(k_copy_words, n_ * 4, n, sk + P::CCA_SK - SYM, (size_t)P::CCA_SK, coins + SYM, (size_t)64, 4);
  return hip_err(hipGetLastError());
}

template <int N>
int cca_enc(pqc_ctx* c, size_t n_, uint8_t* ct, uint8_t* ss, const uint8_t* pk, const uint8_t* coins, bool plain,
            hipStream_t s) {
  using P = Par<N>;
  const Launch L{c->tune.newhope_block};
  const int n = (int)n_;
  uint8_t* km = c->nhkem;
  const size_t kb = kNewhopeKemBytes;
  LAUNCH(k_shake256_words<4>, n_, n, km + KM_MU, kb, coins, (size_t)32, 4);              // mu = H(coins)
  LAUNCH(k_shake256_pair<4>, 2 * n_, n, km + KM_HPK, kb, pk, (size_t)P::PK, P::PK / 8);  // H(pk)
  LAUNCH(k_shake256_2x32<12>, n_, n, km + KM_KCD, kb, (const uint8_t*)km + KM_MU, kb, (const uint8_t*)km + KM_HPK, kb);
  int rc = enc_impl<N>(c, n, ct, P::CCA_CT, km + KM_MU, kb, pk, P::PK, km + KM_KCD + SYM, kb, plain, s);
  if (rc) return rc;
  LAUNCH(k_copy_words, n_ * 4, n, ct + P::CT, (size_t)P::CCA_CT, (const uint8_t*)km + KM_KCD + 2 * SYM, kb, 4);  // d
  LAUNCH(k_shake256_pair<4>, 2 * n_, n, km + KM_HC, kb, (const uint8_t*)ct, (size_t)P::CCA_CT, P::CCA_CT / 8);  // H(c)
  LAUNCH(k_shake256_2x32<4>, n_, n, ss, (size_t)32, (const uint8_t*)km + KM_KCD, kb, (const uint8_t*)km + KM_HC, kb);
  return hip_err(hipGetLastError());
}

template <int N>
int cca_dec(pqc_ctx* c, size_t n_, uint8_t* ss, const uint8_t* ct, const uint8_t* sk, bool plain, hipStream_t s) {
  using P = Par<N>;
  const Launch L{c->tune.newhope_block};
  const int n = (int)n_;
  uint8_t* km = c->nhkem;
  const size_t kb = kNewhopeKemBytes;
  // H(c) depends only on the input ciphertext: it runs on the context's side
  // stream (two lanes per element, half a wave per SIMD at 16384 elements)
  // beside the decryption and re-encryption, and joins before the key selection
  if (hipEventRecord(c->nh_fork, s) != hipSuccess || hipStreamWaitEvent(c->nh_side, c->nh_fork, 0) != hipSuccess)
    return PQC_ERR_DEVICE;
  hipLaunchKernelGGL(k_shake256_pair<4>, L.grid(2 * n_), dim3(L.block), 0, c->nh_side, n, km + KM_HC, kb, ct,
                     (size_t)P::CCA_CT, P::CCA_CT / 8);  // H(c)
  if (hipGetLastError() != hipSuccess || hipEventRecord(c->nh_join, c->nh_side) != hipSuccess) return PQC_ERR_DEVICE;
  int rc = dec_impl<N>(c, n, km + KM_MU, kb, ct, P::CCA_CT, sk, P::CCA_SK, plain, s);  // m'
  if (rc) {
    (void)hipStreamWaitEvent(s, c->nh_join, 0);
    return rc;
  }
  LAUNCH(k_copy_words, n_ * 4, n, km + KM_HPK, kb, sk + P::SK + P::PK, (size_t)P::CCA_SK, 4);
  LAUNCH(k_shake256_2x32<12>, n_, n, km + KM_KCD, kb, (const uint8_t*)km + KM_MU, kb, (const uint8_t*)km + KM_HPK, kb);
  rc = enc_impl<N>(c, n, km + KM_CT2, kb, km + KM_MU, kb, sk + P::SK, P::CCA_SK, km + KM_KCD + SYM, kb, plain, s);
  if (rc) {
    (void)hipStreamWaitEvent(s, c->nh_join, 0);
    return rc;
  }
  LAUNCH(k_copy_words, n_ * 4, n, km + KM_CT2 + P::CT, kb, (const uint8_t*)km + KM_KCD + 2 * SYM, kb, 4);  // d'
  hipLaunchKernelGGL(k_verify_words, dim3((unsigned)n), dim3(64), 0, s, n, reinterpret_cast<uint32_t*>(km + KM_FAIL),
                     kb, ct, (size_t)P::CCA_CT, (const uint8_t*)km + KM_CT2, kb, P::CCA_CT / 8);
  if (hipStreamWaitEvent(s, c->nh_join, 0) != hipSuccess) return PQC_ERR_DEVICE;
  LAUNCH(k_cmov32, n_ * 4, n, km + KM_KCD, kb, sk + P::CCA_SK - SYM, (size_t)P::CCA_SK,
         reinterpret_cast<const uint32_t*>(km + KM_FAIL), kb);
  LAUNCH(k_shake256_2x32<4>, n_, n, ss, (size_t)32, (const uint8_t*)km + KM_KCD, kb, (const uint8_t*)km + KM_HC, kb);
  return hip_err(hipGetLastError());
}

}  // namespace newhope

#define NH_DISPATCH(N, call) ((N) == 1024 ? newhope::call<1024> : newhope::call<512>)

int newhope_keypair(pqc_ctx* c, int N, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* seed, bool plain,
                    hipStream_t s) {
  if (N != 512 && N != 1024) return PQC_ERR_INVALID_ARG;
  const size_t sks = N == 1024 ? newhope::Par<1024>::SK : newhope::Par<512>::SK;
  return NH_DISPATCH(N, keypair_impl)(c, (int)n, pk, sk, sks, seed, 32, plain, s);
}

int newhope_enc(pqc_ctx* c, int N, size_t n, uint8_t* ct, const uint8_t* m, const uint8_t* pk, const uint8_t* coin,
                bool plain, hipStream_t s) {
  if (N != 512 && N != 1024) return PQC_ERR_INVALID_ARG;
  const size_t cts = N == 1024 ? newhope::Par<1024>::CT : newhope::Par<512>::CT;
  const size_t pks = N == 1024 ? newhope::Par<1024>::PK : newhope::Par<512>::PK;
  return NH_DISPATCH(N, enc_impl)(c, (int)n, ct, cts, m, 32, pk, pks, coin, 32, plain, s);
}

int newhope_dec(pqc_ctx* c, int N, size_t n, uint8_t* m, const uint8_t* ct, const uint8_t* sk, bool plain,
                hipStream_t s) {
  if (N != 512 && N != 1024) return PQC_ERR_INVALID_ARG;
  const size_t cts = N == 1024 ? newhope::Par<1024>::CT : newhope::Par<512>::CT;
  const size_t sks = N == 1024 ? newhope::Par<1024>::SK : newhope::Par<512>::SK;
  return NH_DISPATCH(N, dec_impl)(c, (int)n, m, 32, ct, cts, sk, sks, plain, s);
}

int newhope_cca_keypair(pqc_ctx* c, int N, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* coins, bool plain,
                        hipStream_t s) {
  if (N != 512 && N != 1024) return PQC_ERR_INVALID_ARG;
  return NH_DISPATCH(N, cca_keypair)(c, n, pk, sk, coins, plain, s);
}

int newhope_cca_enc(pqc_ctx* c, int N, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk, const uint8_t* coins,
                    bool plain, hipStream_t s) {
  if (N != 512 && N != 1024) return PQC_ERR_INVALID_ARG;
  return NH_DISPATCH(N, cca_enc)(c, n, ct, ss, pk, coins, plain, s);
}

int newhope_cca_dec(pqc_ctx* c, int N, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk, bool plain,
                    hipStream_t s) {
  if (N != 512 && N != 1024) return PQC_ERR_INVALID_ARG;
  return NH_DISPATCH(N, cca_dec)(c, n, ss, ct, sk, plain, s);
}

int newhope_poly_sample(size_t n, uint16_t* polys, const uint8_t* seeds, size_t seed_stride, int nonce,
                        hipStream_t s) {
  using namespace newhope;
  if (nonce < 0 || nonce > 255 || seed_stride < 32 || seed_stride % 8) return PQC_ERR_INVALID_ARG;
  if (n == 0) return PQC_OK;
  if (n > (size_t)INT32_MAX / 16) return PQC_ERR_BATCH_TOO_LARGE;
  const size_t threads = 16 * n;
  hipLaunchKernelGGL(k_sample<1024>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, (int)n,
                     Pv<1024>{polys}, seeds, seed_stride, nonce);
  return hip_err(hipGetLastError());
}

int newhope_poly_transform(int op, size_t n, uint16_t* polys, hipStream_t s) {
  using namespace newhope;
  if (op != 0 && op != 1) return PQC_ERR_INVALID_ARG;
  if (n == 0) return PQC_OK;
  if (n > (size_t)INT32_MAX) return PQC_ERR_BATCH_TOO_LARGE;
  if (op == 0)
    hipLaunchKernelGGL(k_ntt<1024>, dim3((unsigned)n), dim3(64), 0, s, (int)n, Pv<1024>{polys});
  else
    hipLaunchKernelGGL(k_invntt<1024>, dim3((unsigned)n), dim3(64), 0, s, (int)n, Pv<1024>{polys});
  return hip_err(hipGetLastError());
}

}  // namespace pqc
