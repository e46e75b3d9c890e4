// C ABI runtime: contexts, dispatch to the plain / fused kernels, host-pointer
// staging, the NIST single-instance drop-ins, randomness and device plumbing.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <sys/random.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "internal.h"

namespace {

struct AlgInfo {
  int kind;  // 1 = Kyber (k = K), 2 = NewHope CPA, 3 = Frodo, 4 = NewHope CCA (k = N)
  int k;
  size_t pk, sk, ct, ss;
};

bool alg_info(int alg, AlgInfo* a) {
  switch (alg) {
    case PQC_ALG_KYBER512: *a = {1, 2, 800, 1632, 736, 32}; return true;
    case PQC_ALG_KYBER768: *a = {1, 3, 1184, 2400, 1088, 32}; return true;
    case PQC_ALG_KYBER1024: *a = {1, 4, 1568, 3168, 1568, 32}; return true;
    case PQC_ALG_NEWHOPE1024_CPA: *a = {2, 0, 1824, 1792, 2176, 32}; return true;
    case PQC_ALG_FRODO976: *a = {3, 0, 15632, 31296, 15744, 24}; return true;
    case PQC_ALG_NEWHOPE1024: *a = {4, 1024, 1824, 3680, 2208, 32}; return true;
    case PQC_ALG_NEWHOPE512: *a = {4, 512, 928, 1888, 1120, 32}; return true;
    case PQC_ALG_NEWHOPE512_CPA: *a = {2, 512, 928, 896, 1088, 32}; return true;
    default: return false;
  }
}

constexpr unsigned kNewhopeCca = PQC_ALG_MASK(PQC_ALG_NEWHOPE1024) | PQC_ALG_MASK(PQC_ALG_NEWHOPE512);
constexpr unsigned kNewhopeAny =
    kNewhopeCca | PQC_ALG_MASK(PQC_ALG_NEWHOPE1024_CPA) | PQC_ALG_MASK(PQC_ALG_NEWHOPE512_CPA);

int check_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PQC_ERR_NO_DEVICE;
  return PQC_OK;
}

int ctx_alloc(pqc_ctx* c, void** p, size_t bytes) {
  if (bytes == 0) {
    *p = nullptr;
    return PQC_OK;
  }
  if (c->hbm_budget && c->allocated + bytes > c->hbm_budget) return PQC_ERR_NO_MEMORY;
  if (hipMalloc(p, bytes) != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return PQC_ERR_NO_MEMORY;
  }
  c->allocated += bytes;
  return PQC_OK;
}

bool aligned8(const void* p) { return ((uintptr_t)p & 7u) == 0; }

// ---- randomness ----------------------------------------------------------------
std::mutex g_rng_mu;
bool g_rng_det = false;
uint64_t g_rng_seed = 0, g_rng_index = 0;

// The deterministic DRBG is SHAKE256("pqc-mi355x" || le64 seed || le64 stream
// || le64 index); it is computed on the host by a small portable Keccak.
uint64_t rotl(uint64_t x, int n) { return n ? (x << n) | (x >> (64 - n)) : x; }

void host_keccakf(uint64_t a[25]) {
  static uint64_t rc[24];
  static int rho[25];
  static bool init = false;
  if (!init) {
    for (int ir = 0; ir < 24; ir++) {
      uint64_t v = 0;
      for (int j = 0; j <= 6; j++) {
        unsigned t = (unsigned)(j + 7 * ir) % 255, r = 1;
        for (unsigned i = 0; i < t; i++) {
          r <<= 1;
          if (r & 0x100) r ^= 0x171;
        }
        if (r & 1) v |= 1ull << ((1 << j) - 1);
      }
      rc[ir] = v;
    }
    rho[0] = 0;
    unsigned x = 1, y = 0;
    for (unsigned t = 0; t < 24; t++) {
      rho[x + 5 * y] = (int)(((t + 1) * (t + 2) / 2) % 64);
      unsigned nx = y, ny = (2 * x + 3 * y) % 5;
      x = nx;
      y = ny;
    }
    init = true;
  }
  uint64_t b[25], c[5], d[5];
  for (int ir = 0; ir < 24; ir++) {
    for (int i = 0; i < 5; i++) c[i] = a[i] ^ a[i + 5] ^ a[i + 10] ^ a[i + 15] ^ a[i + 20];
    for (int i = 0; i < 5; i++) d[i] = c[(i + 4) % 5] ^ rotl(c[(i + 1) % 5], 1);
    for (int i = 0; i < 25; i++) a[i] ^= d[i % 5];
    for (int xx = 0; xx < 5; xx++)
      for (int yy = 0; yy < 5; yy++) b[yy + 5 * ((2 * xx + 3 * yy) % 5)] = rotl(a[xx + 5 * yy], rho[xx + 5 * yy]);
    for (int yy = 0; yy < 5; yy++)
      for (int xx = 0; xx < 5; xx++)
        a[xx + 5 * yy] = b[xx + 5 * yy] ^ (~b[(xx + 1) % 5 + 5 * yy] & b[(xx + 2) % 5 + 5 * yy]);
    a[0] ^= rc[ir];
  }
}

void drbg(uint8_t* out, size_t len, uint64_t seed, uint64_t stream, uint64_t index) {
  uint8_t in[136] = {0};
  memcpy(in, "pqc-mi355x", 10);
  for (int i = 0; i < 8; i++) {
    in[10 + i] = (uint8_t)(seed >> (8 * i));
    in[18 + i] = (uint8_t)(stream >> (8 * i));
    in[26 + i] = (uint8_t)(index >> (8 * i));
  }
  in[34] = 0x1F;
  in[135] |= 0x80;
  uint64_t a[25] = {0};
  for (int w = 0; w < 17; w++)
    for (int b = 0; b < 8; b++) a[w] |= (uint64_t)in[8 * w + b] << (8 * b);
  while (len) {
    host_keccakf(a);
    for (int w = 0; w < 17 && len; w++)
      for (int b = 0; b < 8 && len; b++, len--) *out++ = (uint8_t)(a[w] >> (8 * b));
  }
}

// ---- library-owned contexts for the NIST drop-ins ---------------------------------
// One pool per device.  A drop-in call checks a context out for its whole
// duration and returns it afterwards, so two host threads never share a
// context's staging buffer, pinned double buffers or status word.  The pool
// grows to the largest number of concurrent callers and is kept until exit.
std::mutex g_default_mu;
std::vector<pqc_ctx*> g_default_free[64];

}  // namespace

void pqc::host_drbg(uint8_t* out, size_t len, uint64_t seed, uint64_t stream, uint64_t index) {
  drbg(out, len, seed, stream, index);
}

namespace {

pqc_ctx* default_ctx_acquire(int* dev_out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  *dev_out = dev;
  {
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default_free[dev].empty()) {
      pqc_ctx* c = g_default_free[dev].back();
      g_default_free[dev].pop_back();
      return c;
    }
  }
  pqc_ctx* c = nullptr;
  unsigned mask = PQC_ALG_MASK(PQC_ALG_KYBER512) | PQC_ALG_MASK(PQC_ALG_KYBER768) | PQC_ALG_MASK(PQC_ALG_KYBER1024) |
                  PQC_ALG_MASK(PQC_ALG_FRODO976) | PQC_ALG_MASK(PQC_ALG_NEWHOPE1024) |
                  PQC_ALG_MASK(PQC_ALG_NEWHOPE512);
  if (pqc_ctx_create(&c, mask, 64, 0) != PQC_OK) return nullptr;
  return c;
}

void default_ctx_release(int dev, pqc_ctx* c) {
  std::lock_guard<std::mutex> lk(g_default_mu);
  g_default_free[dev].push_back(c);
}

int stage_ensure(pqc_ctx* c, size_t bytes) {
  if (c->stage_bytes >= bytes) return PQC_OK;
  if (c->stage) {
    (void)hipFree(c->stage);
    c->allocated -= c->stage_bytes;
    c->stage = nullptr;
    c->stage_bytes = 0;
  }
  void* p = nullptr;
  int rc = ctx_alloc(c, &p, bytes);
  if (rc) return rc;
  c->stage = (uint8_t*)p;
  c->stage_bytes = bytes;
  return PQC_OK;
}

enum Op { OP_KEYPAIR, OP_ENC, OP_DEC };

// One buffer argument of a batched call: `per` bytes per element.
struct Buf {
  const void* p;
  size_t per;
  bool out;
};

// Validates pointers, stages host buffers through the context when
// PQC_F_HOST_PTRS is set, runs `body` on device pointers, copies outputs back
// and synchronises.  Device-pointer calls stay asynchronous on the stream.
template <size_t NB, typename F>
int staged(pqc_ctx* c, size_t n, unsigned flags, hipStream_t s, const Buf (&bufs)[NB], F body) {
  uint8_t* dev[NB];
  for (size_t i = 0; i < NB; ++i) {
    if (!bufs[i].p) return PQC_ERR_INVALID_ARG;
    dev[i] = (uint8_t*)bufs[i].p;
  }
  const bool host = (flags & PQC_F_HOST_PTRS) != 0;
  // Host-pointer calls report only the status bits they raise themselves:
  // bits left by earlier device-pointer calls stay for pqc_ctx_take_status.
  uint32_t st0 = 0;
  if (host) {
    if (hipMemcpyAsync(&st0, c->status, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return PQC_ERR_DEVICE;
    size_t total = 0;
    for (size_t i = 0; i < NB; ++i) total += ((n * bufs[i].per + 255) & ~(size_t)255);
    int rc = stage_ensure(c, total);
    if (rc) return rc;
    uint8_t* p = c->stage;
    for (size_t i = 0; i < NB; ++i) {
      dev[i] = p;
      p += (n * bufs[i].per + 255) & ~(size_t)255;
      if (!bufs[i].out && (rc = pqc::host_to_device(c, dev[i], bufs[i].p, n * bufs[i].per, s))) return rc;
    }
  } else {
    for (size_t i = 0; i < NB; ++i)
      if (!aligned8(bufs[i].p)) return PQC_ERR_INVALID_ARG;
  }
  int rc = body(dev);
  if (rc) return rc;
  if (host) {
    for (size_t i = 0; i < NB; ++i)
      if (bufs[i].out && (rc = pqc::device_to_host(c, (void*)bufs[i].p, dev[i], n * bufs[i].per, s))) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) return PQC_ERR_DEVICE;
    uint32_t st = 0;
    if (hipMemcpy(&st, c->status, 4, hipMemcpyDeviceToHost) != hipSuccess) return PQC_ERR_DEVICE;
    if (st & ~st0) return PQC_ERR_DEVICE;
  }
  return PQC_OK;
}

int check_call(pqc_ctx* c, int alg, size_t n) {
  if (!c) return PQC_ERR_INVALID_ARG;
  if (!(c->alg_mask & PQC_ALG_MASK(alg))) return PQC_ERR_UNSUPPORTED;
  if (n > c->max_batch) return PQC_ERR_BATCH_TOO_LARGE;
  return PQC_OK;
}

// Kyber batched entry: o0/o1 outputs, i0/i1 inputs (see the switch for sizes).
int kem_call(pqc_ctx* c, int alg, Op op, size_t n, uint8_t* o0, uint8_t* o1, const uint8_t* i0,
             const uint8_t* i1, unsigned flags, void* stream) {
  AlgInfo a;
  if (!alg_info(alg, &a) || (a.kind != 1 && a.kind != 3 && a.kind != 4)) return PQC_ERR_UNSUPPORTED;
  int rc = check_call(c, alg, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const bool plain = (flags & PQC_F_PLAIN) != 0;
  if (a.kind == 4) {  // NewHope CCA-KEM, N = a.k
    switch (op) {
      case OP_KEYPAIR: {
        const Buf b[3] = {{o0, a.pk, true}, {o1, a.sk, true}, {i0, 64, false}};
        return staged(c, n, flags, s, b,
                      [&](uint8_t** d) { return pqc::newhope_cca_keypair(c, a.k, n, d[0], d[1], d[2], plain, s); });
      }
      case OP_ENC: {
        const Buf b[4] = {{o0, a.ct, true}, {o1, a.ss, true}, {i0, a.pk, false}, {i1, 32, false}};
        return staged(c, n, flags, s, b,
                      [&](uint8_t** d) { return pqc::newhope_cca_enc(c, a.k, n, d[0], d[1], d[2], d[3], plain, s); });
      }
      default: {
        const Buf b[3] = {{o0, a.ss, true}, {i0, a.ct, false}, {i1, a.sk, false}};
        return staged(c, n, flags, s, b,
                      [&](uint8_t** d) { return pqc::newhope_cca_dec(c, a.k, n, d[0], d[1], d[2], plain, s); });
      }
    }
  }
  if (a.kind == 3) {  // FrodoKEM-976 with one key per element
    switch (op) {
      case OP_KEYPAIR: {
        const Buf b[3] = {{o0, a.pk, true}, {o1, a.sk, true}, {i0, 64, false}};
        return staged(c, n, flags, s, b, [&](uint8_t** d) { return pqc::frodo_keypair(c, n, d[0], d[1], d[2], s); });
      }
      case OP_ENC: {
        const Buf b[4] = {{o0, a.ct, true}, {o1, a.ss, true}, {i0, a.pk, false}, {i1, 24, false}};
        return staged(c, n, flags, s, b,
                      [&](uint8_t** d) { return pqc::frodo_enc_batch(c, n, d[0], d[1], d[2], d[3], s); });
      }
      default: {
        const Buf b[3] = {{o0, a.ss, true}, {i0, a.ct, false}, {i1, a.sk, false}};
        return staged(c, n, flags, s, b, [&](uint8_t** d) { return pqc::frodo_dec_batch(c, n, d[0], d[1], d[2], s); });
      }
    }
  }
  const int k = a.k;
  switch (op) {
    case OP_KEYPAIR: {
      const Buf b[3] = {{o0, a.pk, true}, {o1, a.sk, true}, {i0, 64, false}};
      return staged(c, n, flags, s, b, [&](uint8_t** d) {
        return plain ? pqc::kyber_plain_keypair(c, k, n, d[0], d[1], d[2], s)
                     : pqc::kyber_fused_keypair(c, k, n, d[0], d[1], d[2], s);
      });
    }
    case OP_ENC: {
      const Buf b[4] = {{o0, a.ct, true}, {o1, a.ss, true}, {i0, a.pk, false}, {i1, 32, false}};
      return staged(c, n, flags, s, b, [&](uint8_t** d) {
        return plain ? pqc::kyber_plain_enc(c, k, n, d[0], d[1], d[2], d[3], s)
                     : pqc::kyber_fused_enc(c, k, n, d[0], d[1], d[2], d[3], s);
      });
    }
    default: {
      const Buf b[3] = {{o0, a.ss, true}, {i0, a.ct, false}, {i1, a.sk, false}};
      return staged(c, n, flags, s, b, [&](uint8_t** d) {
        return plain ? pqc::kyber_plain_dec(c, k, n, d[0], d[1], d[2], s)
                     : pqc::kyber_fused_dec(c, k, n, d[0], d[1], d[2], s);
      });
    }
  }
}

int single_on(pqc_ctx* c, int alg, Op op, uint8_t* o0, uint8_t* o1, const uint8_t* i0, const uint8_t* i1) {
  // Coins are drawn call by call as the NIST submissions do, so that the
  // known-answer-test DRBG reproduces their outputs: Kyber and NewHope keypair
  // CPA seed then z / s (cpapke/indcpa keypair, then crypto_kem_keypair's own
  // draw), encaps 32 bytes; Frodo
  // keypair s || seedSE || z in one draw, encaps mu (FrodoKEM_GPU/kem.c:33,90).
  uint8_t coins[64];
  const bool frodo = alg == PQC_ALG_FRODO976;
  if (op == OP_KEYPAIR) {
    if (frodo ? pqc_randombytes(coins, 64) : (pqc_randombytes(coins, 32) || pqc_randombytes(coins + 32, 32)))
      return PQC_ERR_DEVICE;
    i0 = coins;
    i1 = coins;
  } else if (op == OP_ENC) {
    if (pqc_randombytes(coins, frodo ? 24 : 32)) return PQC_ERR_DEVICE;
    i1 = coins;
  }
  return kem_call(c, alg, op, 1, o0, o1, i0, i1, PQC_F_HOST_PTRS, nullptr);
}

int single(int alg, Op op, uint8_t* o0, uint8_t* o1, const uint8_t* i0, const uint8_t* i1) {
  int dev = 0;
  pqc_ctx* c = default_ctx_acquire(&dev);
  if (!c) return PQC_ERR_NO_DEVICE;
  const int rc = single_on(c, alg, op, o0, o1, i0, i1);
  default_ctx_release(dev, c);
  return rc;
}

}  // namespace

extern "C" {

int pqc_ctx_create(pqc_ctx** out, unsigned alg_mask, size_t max_batch, size_t hbm_budget) {
  if (!out || max_batch == 0 || max_batch > (1u << 26)) return PQC_ERR_INVALID_ARG;
  *out = nullptr;
  int rc = check_device();
  if (rc) return rc;
  const unsigned known = PQC_ALG_MASK(PQC_ALG_KYBER512) | PQC_ALG_MASK(PQC_ALG_KYBER768) |
                         PQC_ALG_MASK(PQC_ALG_KYBER1024) | PQC_ALG_MASK(PQC_ALG_NEWHOPE1024_CPA) |
                         PQC_ALG_MASK(PQC_ALG_FRODO976) | PQC_ALG_MASK(PQC_ALG_NEWHOPE1024) |
                         PQC_ALG_MASK(PQC_ALG_NEWHOPE512) | PQC_ALG_MASK(PQC_ALG_NEWHOPE512_CPA);
  if (alg_mask == 0 || (alg_mask & ~known)) return PQC_ERR_UNSUPPORTED;
  pqc_ctx* c = new pqc_ctx();
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return PQC_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
    delete c;
    return PQC_ERR_DEVICE;
  }
  c->cu_count = prop.multiProcessorCount;
  c->alg_mask = alg_mask;
  c->max_batch = max_batch;
  c->hbm_budget = hbm_budget;
  c->ld = (max_batch + 63) & ~(size_t)63;
  int kmax = 0;
  for (int alg = PQC_ALG_KYBER512; alg <= PQC_ALG_KYBER1024; ++alg)
    if (alg_mask & PQC_ALG_MASK(alg)) kmax = alg + 1;
  void* p = nullptr;
  if (kmax) {
    c->kpolys_count = (int)pqc::kyber_plain_polys(kmax);
    rc = ctx_alloc(c, &p, (size_t)c->kpolys_count * 256 * c->ld * sizeof(int16_t));
    if (rc) goto fail;
    c->kpolys = (int16_t*)p;
    c->kbytes_per_elem = pqc::kyber_plain_bytes_per_elem(kmax);
    rc = ctx_alloc(c, &p, c->kbytes_per_elem * c->ld);
    if (rc) goto fail;
    c->kbytes = (uint8_t*)p;
    c->kprf_per_elem = (size_t)(2 * kmax + 1) * 128;
    rc = ctx_alloc(c, &p, c->kprf_per_elem * c->ld);
    if (rc) goto fail;
    c->kprf = (uint8_t*)p;
    // A^T scratch of the 32-request decapsulation tiles: one region per wave of the
    // largest grid (two waves per SIMD) the batch size can use
    c->kat_waves = std::min<size_t>((size_t)c->cu_count * 8, (max_batch + 31) / 32);
    c->kat_region = pqc::tile32_scratch_bytes(kmax) / sizeof(int16_t);
    rc = ctx_alloc(c, &p, pqc::tile32_scratch_bytes(kmax) * c->kat_waves);
    if (rc) goto fail;
    c->kat = (int16_t*)p;
    rc = ctx_alloc(c, &p, (size_t)3 * c->ld * sizeof(uint32_t) + 256);
    if (rc) goto fail;
    c->kmix_order = (uint32_t*)p;
    c->kmix_counts = (uint32_t*)((uint8_t*)p + (size_t)3 * c->ld * sizeof(uint32_t));
  }
  if (alg_mask & kNewhopeAny) {
    rc = ctx_alloc(c, &p, (size_t)8 * 1024 * c->ld * sizeof(uint16_t));
    if (rc) goto fail;
    c->nhpolys = (uint16_t*)p;
    rc = ctx_alloc(c, &p, (size_t)64 * c->ld);
    if (rc) goto fail;
    c->nhbytes = (uint8_t*)p;
  }
  if (alg_mask & kNewhopeCca) {
    rc = ctx_alloc(c, &p, (size_t)pqc::kNewhopeKemBytes * c->ld);
    if (rc) goto fail;
    c->nhkem = (uint8_t*)p;
    if (hipStreamCreateWithFlags(&c->nh_side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->nh_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->nh_join, hipEventDisableTiming) != hipSuccess) {
      rc = PQC_ERR_DEVICE;
      goto fail;
    }
  }
  if (alg_mask & PQC_ALG_MASK(PQC_ALG_FRODO976)) {
    rc = ctx_alloc(c, &p, (size_t)976 * 976 * 2);
    if (rc) goto fail;
    c->fr_a16 = (uint16_t*)p;
    rc = ctx_alloc(c, &p, (size_t)2 * 64 * 16 * 64 * 16);
    if (rc) goto fail;
    c->fr_frag = (int8_t*)p;
    rc = ctx_alloc(c, &p, (size_t)8192 * c->ld);
    if (rc) goto fail;
    c->fr_sfrag = (int8_t*)p;
    if (hipMemset(c->fr_sfrag, 0, (size_t)8192 * c->ld) != hipSuccess) {
      rc = PQC_ERR_DEVICE;
      goto fail;
    }
    rc = ctx_alloc(c, &p, 31296 + 15632 + 64);
    if (rc) goto fail;
    c->fr_sk = (uint8_t*)p;
    c->fr_pk = c->fr_sk + 31296;
    c->fr_pkh = c->fr_pk + 15632;
    rc = ctx_alloc(c, &p, pqc::frodo_bytes_per_elem() * c->ld);
    if (rc) goto fail;
    c->fr_elem = (uint8_t*)p;
    rc = ctx_alloc(c, &p, (size_t)std::min<size_t>(c->ld, 128) * 976 * 976 * 2);
    if (rc) goto fail;
    c->fr_abuf = (uint16_t*)p;
    bool ok = hipEventCreateWithFlags(&c->fr_fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < kFrSideStreams; ++i)
      ok = hipStreamCreateWithFlags(&c->fr_side[i], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&c->fr_join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      rc = PQC_ERR_DEVICE;
      goto fail;
    }
  }
  rc = ctx_alloc(c, &p, 256);
  if (rc) goto fail;
  c->status = (uint32_t*)p;
  if (hipMemset(c->status, 0, 256) != hipSuccess) {
    rc = PQC_ERR_DEVICE;
    goto fail;
  }
  *out = c;
  return PQC_OK;
fail:
  pqc_ctx_destroy(c);
  return rc;
}

void pqc_ctx_destroy(pqc_ctx* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  pqc::pqc_tuner_free(c);
  if (c->kpolys) (void)hipFree(c->kpolys);
  if (c->kbytes) (void)hipFree(c->kbytes);
  if (c->kprf) (void)hipFree(c->kprf);
  if (c->kat) (void)hipFree(c->kat);
  if (c->kmix_order) (void)hipFree(c->kmix_order);
  if (c->fr_a16) (void)hipFree(c->fr_a16);
  if (c->fr_abuf) (void)hipFree(c->fr_abuf);
  if (c->fr_frag) (void)hipFree(c->fr_frag);
  if (c->fr_sfrag) (void)hipFree(c->fr_sfrag);
  if (c->fr_sk) (void)hipFree(c->fr_sk);
  if (c->fr_elem) (void)hipFree(c->fr_elem);
  if (c->nhpolys) (void)hipFree(c->nhpolys);
  if (c->nhbytes) (void)hipFree(c->nhbytes);
  if (c->nhkem) (void)hipFree(c->nhkem);
  if (c->nh_side) (void)hipStreamSynchronize(c->nh_side);
  if (c->nh_fork) (void)hipEventDestroy(c->nh_fork);
  if (c->nh_join) (void)hipEventDestroy(c->nh_join);
  if (c->nh_side) (void)hipStreamDestroy(c->nh_side);
  for (int i = 0; i < kFrSideStreams; ++i) {
    if (c->fr_side[i]) (void)hipStreamDestroy(c->fr_side[i]);
    if (c->fr_join[i]) (void)hipEventDestroy(c->fr_join[i]);
  }
  if (c->fr_fork) (void)hipEventDestroy(c->fr_fork);
  if (c->stage) (void)hipFree(c->stage);
  pqc::hostio_free(c);
  if (c->status) (void)hipFree(c->status);
  delete c;
}

size_t pqc_ctx_device_bytes(const pqc_ctx* c) { return c ? c->allocated : 0; }

int pqc_ctx_set_tuning(pqc_ctx* c, int waves_per_cu, int grid) {
  if (!c || waves_per_cu < 0 || waves_per_cu > 32 || grid < 0) return PQC_ERR_INVALID_ARG;
  if (waves_per_cu) c->tune.fused_waves_per_cu = waves_per_cu;
  c->fused_grid_override = grid;
  return PQC_OK;
}

int pqc_ctx_get_tuning(const pqc_ctx* c, pqc_tuning* t) {
  if (!c || !t) return PQC_ERR_INVALID_ARG;
  *t = c->tune;
  return PQC_OK;
}

int pqc_ctx_put_tuning(pqc_ctx* c, const pqc_tuning* t) {
  if (!c || !t) return PQC_ERR_INVALID_ARG;
  auto block_ok = [](int b) { return b >= 64 && b <= 1024 && b % 64 == 0; };
  if (!block_ok(t->plain_block) || !block_ok(t->newhope_block) || t->fused_waves_per_cu < 1 ||
      t->fused_waves_per_cu > 4 || t->mixed_order < 0 || t->mixed_order > 3 ||
      (t->frodo_stream_block != 64 && t->frodo_stream_block != 128 && t->frodo_stream_block != 256) ||
      t->lowlat_max < -1 || t->lowlat_max > 65536 || t->frodo_chunk < 0 || t->frodo_chunk > 65536 ||
      t->frodo_chunk % 32 != 0 || (t->fused_dec_tile != 0 && t->fused_dec_tile != 32 && t->fused_dec_tile != 64) ||
      (t->fused_enc_tile != 0 && t->fused_enc_tile != 32 && t->fused_enc_tile != 64) ||
      (t->fused_key_tile != 0 && t->fused_key_tile != 32 && t->fused_key_tile != 64) || t->frodo_lane_min < -1)
    return PQC_ERR_INVALID_ARG;
  c->tune = *t;
  return PQC_OK;
}

int pqc_alg_sizes(int alg, size_t* pk, size_t* sk, size_t* ct, size_t* ss) {
  AlgInfo a;
  if (!alg_info(alg, &a)) return PQC_ERR_UNSUPPORTED;
  if (pk) *pk = a.pk;
  if (sk) *sk = a.sk;
  if (ct) *ct = a.ct;
  if (ss) *ss = a.ss;
  return PQC_OK;
}

int crypto_kem_keypair_derand_batch(pqc_ctx* c, int alg, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* coins,
                                    unsigned flags, void* stream) {
  return kem_call(c, alg, OP_KEYPAIR, n, pk, sk, coins, coins, flags, stream);
}

int crypto_kem_enc_derand_batch(pqc_ctx* c, int alg, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk,
                                const uint8_t* coins, unsigned flags, void* stream) {
  return kem_call(c, alg, OP_ENC, n, ct, ss, pk, coins, flags, stream);
}

int crypto_kem_dec_batch(pqc_ctx* c, int alg, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk,
                         unsigned flags, void* stream) {
  return kem_call(c, alg, OP_DEC, n, ss, nullptr, ct, sk, flags, stream);
}

int crypto_kem_keypair_batch(pqc_ctx* c, int alg, size_t n, uint8_t* pk, uint8_t* sk, unsigned flags, void* stream) {
  if (!c || n > c->max_batch) return c ? PQC_ERR_BATCH_TOO_LARGE : PQC_ERR_INVALID_ARG;
  uint8_t* coins = new (std::nothrow) uint8_t[n * 64 + 1];
  if (!coins) return PQC_ERR_NO_MEMORY;
  int rc = pqc_randombytes(coins, n * 64);
  if (!rc) {
    if (flags & PQC_F_HOST_PTRS) {
      rc = kem_call(c, alg, OP_KEYPAIR, n, pk, sk, coins, coins, flags, stream);
    } else {
      void* d = nullptr;
      rc = hipMalloc(&d, n * 64) == hipSuccess ? PQC_OK : PQC_ERR_NO_MEMORY;
      if (!rc && hipMemcpy(d, coins, n * 64, hipMemcpyHostToDevice) != hipSuccess) rc = PQC_ERR_DEVICE;
      if (!rc) rc = kem_call(c, alg, OP_KEYPAIR, n, pk, sk, (uint8_t*)d, (uint8_t*)d, flags, stream);
      if (!rc && hipStreamSynchronize((hipStream_t)stream) != hipSuccess) rc = PQC_ERR_DEVICE;
      if (d) (void)hipFree(d);
    }
  }
  delete[] coins;
  return rc;
}

int crypto_kem_enc_batch(pqc_ctx* c, int alg, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk, unsigned flags,
                         void* stream) {
  if (!c || n > c->max_batch) return c ? PQC_ERR_BATCH_TOO_LARGE : PQC_ERR_INVALID_ARG;
  uint8_t* coins = new (std::nothrow) uint8_t[n * 32 + 1];
  if (!coins) return PQC_ERR_NO_MEMORY;
  int rc = pqc_randombytes(coins, n * 32);
  if (!rc) {
    if (flags & PQC_F_HOST_PTRS) {
      rc = kem_call(c, alg, OP_ENC, n, ct, ss, pk, coins, flags, stream);
    } else {
      void* d = nullptr;
      rc = hipMalloc(&d, n * 32) == hipSuccess ? PQC_OK : PQC_ERR_NO_MEMORY;
      if (!rc && hipMemcpy(d, coins, n * 32, hipMemcpyHostToDevice) != hipSuccess) rc = PQC_ERR_DEVICE;
      if (!rc) rc = kem_call(c, alg, OP_ENC, n, ct, ss, pk, (uint8_t*)d, flags, stream);
      if (!rc && hipStreamSynchronize((hipStream_t)stream) != hipSuccess) rc = PQC_ERR_DEVICE;
      if (d) (void)hipFree(d);
    }
  }
  delete[] coins;
  return rc;
}

int pqc_newhope1024_cpa_keypair_batch(pqc_ctx* c, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* seeds,
                                      unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_NEWHOPE1024_CPA, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[3] = {{pk, 1824, true}, {sk, 1792, true}, {seeds, 32, false}};
  const bool plain = (flags & PQC_F_PLAIN) != 0;
  return staged(c, n, flags, s, b,
                [&](uint8_t** d) { return pqc::newhope_keypair(c, 1024, n, d[0], d[1], d[2], plain, s); });
}

int pqc_newhope1024_cpa_enc_batch(pqc_ctx* c, size_t n, uint8_t* ct, const uint8_t* msgs, const uint8_t* pk,
                                  const uint8_t* coins, unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_NEWHOPE1024_CPA, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[4] = {{ct, 2176, true}, {msgs, 32, false}, {pk, 1824, false}, {coins, 32, false}};
  const bool plain = (flags & PQC_F_PLAIN) != 0;
  return staged(c, n, flags, s, b,
                [&](uint8_t** d) { return pqc::newhope_enc(c, 1024, n, d[0], d[1], d[2], d[3], plain, s); });
}

int pqc_newhope1024_cpa_dec_batch(pqc_ctx* c, size_t n, uint8_t* msgs, const uint8_t* ct, const uint8_t* sk,
                                  unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_NEWHOPE1024_CPA, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[3] = {{msgs, 32, true}, {ct, 2176, false}, {sk, 1792, false}};
  const bool plain = (flags & PQC_F_PLAIN) != 0;
  return staged(c, n, flags, s, b,
                [&](uint8_t** d) { return pqc::newhope_dec(c, 1024, n, d[0], d[1], d[2], plain, s); });
}

int pqc_newhope512_cpa_keypair_batch(pqc_ctx* c, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* seeds,
                                     unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_NEWHOPE512_CPA, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[3] = {{pk, 928, true}, {sk, 896, true}, {seeds, 32, false}};
  const bool plain = (flags & PQC_F_PLAIN) != 0;
  return staged(c, n, flags, s, b,
                [&](uint8_t** d) { return pqc::newhope_keypair(c, 512, n, d[0], d[1], d[2], plain, s); });
}

int pqc_newhope512_cpa_enc_batch(pqc_ctx* c, size_t n, uint8_t* ct, const uint8_t* msgs, const uint8_t* pk,
                                 const uint8_t* coins, unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_NEWHOPE512_CPA, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[4] = {{ct, 1088, true}, {msgs, 32, false}, {pk, 928, false}, {coins, 32, false}};
  const bool plain = (flags & PQC_F_PLAIN) != 0;
  return staged(c, n, flags, s, b,
                [&](uint8_t** d) { return pqc::newhope_enc(c, 512, n, d[0], d[1], d[2], d[3], plain, s); });
}

int pqc_newhope512_cpa_dec_batch(pqc_ctx* c, size_t n, uint8_t* msgs, const uint8_t* ct, const uint8_t* sk,
                                 unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_NEWHOPE512_CPA, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[3] = {{msgs, 32, true}, {ct, 1088, false}, {sk, 896, false}};
  const bool plain = (flags & PQC_F_PLAIN) != 0;
  return staged(c, n, flags, s, b,
                [&](uint8_t** d) { return pqc::newhope_dec(c, 512, n, d[0], d[1], d[2], plain, s); });
}

int pqc_frodo976_keypair_batch(pqc_ctx* c, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* randomness,
                               unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_FRODO976, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[3] = {{pk, 15632, true}, {sk, 31296, true}, {randomness, 64, false}};
  return staged(c, n, flags, s, b, [&](uint8_t** d) { return pqc::frodo_keypair(c, n, d[0], d[1], d[2], s); });
}

int pqc_frodo976_load_key(pqc_ctx* c, const uint8_t* key, int is_sk, unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_FRODO976, 1);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[1] = {{key, (size_t)(is_sk ? 31296 : 15632), false}};
  return staged(c, 1, flags, s, b, [&](uint8_t** d) { return pqc::frodo_load_key(c, d[0], is_sk, s); });
}

int pqc_frodo976_enc_batch(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* mu, unsigned flags,
                           void* stream) {
  int rc = check_call(c, PQC_ALG_FRODO976, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[3] = {{ct, 15744, true}, {ss, 24, true}, {mu, 24, false}};
  return staged(c, n, flags, s, b, [&](uint8_t** d) { return pqc::frodo_enc(c, n, d[0], d[1], d[2], s); });
}

int pqc_frodo976_dec_batch(pqc_ctx* c, size_t n, uint8_t* ss, const uint8_t* ct, unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_FRODO976, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[2] = {{ss, 24, true}, {ct, 15744, false}};
  return staged(c, n, flags, s, b, [&](uint8_t** d) { return pqc::frodo_dec(c, n, d[0], d[1], s); });
}

int pqc_frodo976_products(pqc_ctx* c, size_t n, uint16_t* bpp, uint16_t* v, unsigned flags, void* stream) {
  int rc = check_call(c, PQC_ALG_FRODO976, n);
  if (rc || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Buf b[2] = {{bpp, 976 * 8 * 2, true}, {v, 64 * 2, true}};
  return staged(c, n, flags, s, b, [&](uint8_t** d) {
    return pqc::frodo_products(c, n, (uint16_t*)d[0], (uint16_t*)d[1], s);
  });
}

int pqc_frodo976_gemm(pqc_ctx* c, size_t n, const uint8_t* ct, void* stream) {
  int rc = check_call(c, PQC_ALG_FRODO976, n);
  if (rc || n == 0) return rc;
  if (ct && !aligned8(ct)) return PQC_ERR_INVALID_ARG;
  return pqc::frodo_gemm(c, n, ct, (hipStream_t)stream);
}

int pqc_kem_enc_mixed(pqc_ctx* c, size_t n, const pqc_kem_req* reqs, uint8_t* ct, size_t ct_bytes, const uint8_t* pk,
                      size_t pk_bytes, uint8_t* ss, size_t ss_bytes, const uint8_t* coins, size_t coins_bytes,
                      void* stream) {
  if (!c) return PQC_ERR_INVALID_ARG;
  if (!c->kmix_order) return PQC_ERR_UNSUPPORTED;
  if (n > c->max_batch) return PQC_ERR_BATCH_TOO_LARGE;
  if (n == 0) return PQC_OK;
  if (!reqs || !ct || !pk || !ss || !coins || !aligned8(reqs) || !aligned8(ct) || !aligned8(pk) || !aligned8(ss) ||
      !aligned8(coins))
    return PQC_ERR_INVALID_ARG;
  return pqc::kyber_fused_mixed(c, 0, n, reqs, ct, ct_bytes, pk, pk_bytes, ss, ss_bytes, coins, coins_bytes,
                                (hipStream_t)stream);
}

int pqc_kem_dec_mixed(pqc_ctx* c, size_t n, const pqc_kem_req* reqs, const uint8_t* ct, size_t ct_bytes,
                      const uint8_t* sk, size_t sk_bytes, uint8_t* ss, size_t ss_bytes, void* stream) {
  if (!c) return PQC_ERR_INVALID_ARG;
  if (!c->kmix_order) return PQC_ERR_UNSUPPORTED;
  if (n > c->max_batch) return PQC_ERR_BATCH_TOO_LARGE;
  if (n == 0) return PQC_OK;
  if (!reqs || !ct || !sk || !ss || !aligned8(reqs) || !aligned8(ct) || !aligned8(sk) || !aligned8(ss))
    return PQC_ERR_INVALID_ARG;
  return pqc::kyber_fused_mixed(c, 1, n, reqs, const_cast<uint8_t*>(ct), ct_bytes, sk, sk_bytes, ss, ss_bytes,
                                nullptr, 0, (hipStream_t)stream);
}

int pqc_ctx_take_status(pqc_ctx* c, void* stream, uint32_t* status) {
  if (!c || !status) return PQC_ERR_INVALID_ARG;
  hipStream_t s = (hipStream_t)stream;
  uint32_t zero = 0;
  if (hipStreamSynchronize(s) != hipSuccess) return PQC_ERR_DEVICE;
  if (hipMemcpy(status, c->status, 4, hipMemcpyDeviceToHost) != hipSuccess) return PQC_ERR_DEVICE;
  if (hipMemcpy(c->status, &zero, 4, hipMemcpyHostToDevice) != hipSuccess) return PQC_ERR_DEVICE;
  return PQC_OK;
}

struct pqc_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int pqc_graph_capture_begin(void* stream) {
  if (!stream) return PQC_ERR_INVALID_ARG;  // the legacy default stream cannot be captured
  return hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal) == hipSuccess ? PQC_OK
                                                                                                     : PQC_ERR_DEVICE;
}

int pqc_graph_capture_end(void* stream, pqc_graph** out) {
  if (!stream || !out) return PQC_ERR_INVALID_ARG;
  *out = nullptr;
  auto* g = new (std::nothrow) pqc_graph;
  if (!g) return PQC_ERR_NO_MEMORY;
  if (hipStreamEndCapture((hipStream_t)stream, &g->graph) != hipSuccess || !g->graph) {
    delete g;
    return PQC_ERR_DEVICE;
  }
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    delete g;
    return PQC_ERR_DEVICE;
  }
  *out = g;
  return PQC_OK;
}

int pqc_graph_launch(pqc_graph* g, void* stream) {
  if (!g || !g->exec) return PQC_ERR_INVALID_ARG;
  return hipGraphLaunch(g->exec, (hipStream_t)stream) == hipSuccess ? PQC_OK : PQC_ERR_DEVICE;
}

void pqc_graph_destroy(pqc_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

int crypto_kem_keypair(unsigned char* pk, unsigned char* sk) { return single(PQC_ALG_KYBER768, OP_KEYPAIR, pk, sk, nullptr, nullptr); }
int crypto_kem_enc(unsigned char* ct, unsigned char* ss, const unsigned char* pk) { return single(PQC_ALG_KYBER768, OP_ENC, ct, ss, pk, nullptr); }
int crypto_kem_dec(unsigned char* ss, const unsigned char* ct, const unsigned char* sk) { return single(PQC_ALG_KYBER768, OP_DEC, ss, nullptr, ct, sk); }

#define PQC_NIST_SET(name, alg)                                                                        \
  int pqc_##name##_crypto_kem_keypair(unsigned char* pk, unsigned char* sk) {                          \
    return single(alg, OP_KEYPAIR, pk, sk, nullptr, nullptr);                                          \
  }                                                                                                    \
  int pqc_##name##_crypto_kem_enc(unsigned char* ct, unsigned char* ss, const unsigned char* pk) {     \
    return single(alg, OP_ENC, ct, ss, pk, nullptr);                                                   \
  }                                                                                                    \
  int pqc_##name##_crypto_kem_dec(unsigned char* ss, const unsigned char* ct, const unsigned char* sk) { \
    return single(alg, OP_DEC, ss, nullptr, ct, sk);                                                   \
  }
PQC_NIST_SET(kyber512, PQC_ALG_KYBER512)
PQC_NIST_SET(kyber768, PQC_ALG_KYBER768)
PQC_NIST_SET(kyber1024, PQC_ALG_KYBER1024)
PQC_NIST_SET(frodo976, PQC_ALG_FRODO976)
PQC_NIST_SET(newhope1024, PQC_ALG_NEWHOPE1024)
PQC_NIST_SET(newhope512, PQC_ALG_NEWHOPE512)

void pqc_randombytes_seed(uint64_t seed) {
  pqc::nist_drbg_disable();
  std::lock_guard<std::mutex> lk(g_rng_mu);
  g_rng_det = true;
  g_rng_seed = seed;
  g_rng_index = 0;
}

void pqc_randombytes_system(void) {
  pqc::nist_drbg_disable();
  std::lock_guard<std::mutex> lk(g_rng_mu);
  g_rng_det = false;
}

int pqc_randombytes(uint8_t* out, size_t n) {
  if (!out && n) return PQC_ERR_INVALID_ARG;
  if (pqc::nist_drbg_active()) {
    pqc::nist_drbg_generate(out, n);
    return PQC_OK;
  }
  std::lock_guard<std::mutex> lk(g_rng_mu);
  if (g_rng_det) {
    drbg(out, n, g_rng_seed, 0xFFFFFFFFull, g_rng_index++);
    return PQC_OK;
  }
  size_t got = 0;
  while (got < n) {
    ssize_t r = getrandom(out + got, n - got, 0);
    if (r <= 0) return PQC_ERR_DEVICE;
    got += (size_t)r;
  }
  return PQC_OK;
}

int pqc_device_count(int* count) {
  if (!count) return PQC_ERR_INVALID_ARG;
  *count = 0;
  if (hipGetDeviceCount(count) != hipSuccess) {
    *count = 0;
    return PQC_ERR_NO_DEVICE;
  }
  return PQC_OK;
}

int pqc_set_device(int device) { return pqc::hip_err(hipSetDevice(device)); }

int pqc_device_info(char* name, size_t name_len, int* cu_count, size_t* hbm_bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return PQC_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return PQC_ERR_DEVICE;
  if (name && name_len) snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
  return PQC_OK;
}

int pqc_device_malloc(void** ptr, size_t bytes) {
  if (!ptr) return PQC_ERR_INVALID_ARG;
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
    *ptr = nullptr;
    return PQC_ERR_NO_MEMORY;
  }
  return PQC_OK;
}

int pqc_device_free(void* ptr) { return pqc::hip_err(hipFree(ptr)); }

int pqc_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
  if (stream) return pqc::hip_err(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return pqc::hip_err(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

int pqc_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
  if (stream) return pqc::hip_err(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return pqc::hip_err(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

int pqc_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
  return pqc::hip_err(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
}

int pqc_stream_create(void** stream) {
  if (!stream) return PQC_ERR_INVALID_ARG;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PQC_ERR_DEVICE;
  *stream = s;
  return PQC_OK;
}

int pqc_stream_destroy(void* stream) { return pqc::hip_err(hipStreamDestroy((hipStream_t)stream)); }
int pqc_stream_synchronize(void* stream) { return pqc::hip_err(hipStreamSynchronize((hipStream_t)stream)); }

int pqc_event_create(void** ev) {
  if (!ev) return PQC_ERR_INVALID_ARG;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return PQC_ERR_DEVICE;
  *ev = e;
  return PQC_OK;
}

int pqc_event_destroy(void* ev) { return pqc::hip_err(hipEventDestroy((hipEvent_t)ev)); }
int pqc_event_record(void* ev, void* stream) { return pqc::hip_err(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream)); }
int pqc_event_synchronize(void* ev) { return pqc::hip_err(hipEventSynchronize((hipEvent_t)ev)); }
int pqc_event_elapsed_ms(float* ms, void* a, void* b) {
  if (!ms) return PQC_ERR_INVALID_ARG;
  return pqc::hip_err(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
}

int pqc_keccak_batch(int op, size_t n, uint8_t* out, size_t outlen, const uint8_t* in, size_t inlen, void* stream) {
  if (!out || !in) return PQC_ERR_INVALID_ARG;
  return pqc::keccak_batch(op, n, out, outlen, in, inlen, (hipStream_t)stream);
}

int pqc_kyber_poly_transform(int op, size_t n, int16_t* polys, void* stream) {
  if (!polys) return PQC_ERR_INVALID_ARG;
  return pqc::kyber_poly_transform(op, n, polys, (hipStream_t)stream);
}

int pqc_newhope1024_poly_transform(int op, size_t n, uint16_t* polys, void* stream) {
  if (!polys) return PQC_ERR_INVALID_ARG;
  return pqc::newhope_poly_transform(op, n, polys, (hipStream_t)stream);
}

int pqc_newhope1024_poly_sample(size_t n, uint16_t* polys, const uint8_t* seeds, size_t seed_stride, int nonce,
                                void* stream) {
  if (!polys || !seeds) return PQC_ERR_INVALID_ARG;
  return pqc::newhope_poly_sample(n, polys, seeds, seed_stride, nonce, (hipStream_t)stream);
}

const char* pqc_version(void) { return "pqc-amd 0.1 (gfx950)"; }

const char* pqc_strerror(int code) {
  switch (code) {
    case PQC_OK: return "ok";
    case PQC_ERR_INVALID_ARG: return "invalid argument";
    case PQC_ERR_UNSUPPORTED: return "algorithm not supported by this context";
    case PQC_ERR_BATCH_TOO_LARGE: return "batch larger than the context's max_batch";
    case PQC_ERR_NO_MEMORY: return "device memory budget exceeded";
    case PQC_ERR_DEVICE: return "device error";
    case PQC_ERR_NO_DEVICE: return "no device";
    default: return "unknown error";
  }
}

}  // extern "C"
