// Host-side internals shared by the runtime and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/pqc_api.h"

constexpr int kFrSideStreams = 3;

// FrodoKEM-976 shared-key calls of at least this many requests run the noise
// stream with one Keccak state per lane (tuning.frodo_lane_min = 0)
constexpr int kFrodoLaneMinDefault = 49152;
inline bool frodo_lanes(const pqc_tuning& t, uint64_t n) {
  const int m = t.frodo_lane_min == 0 ? kFrodoLaneMinDefault : t.frodo_lane_min;
  return m > 0 && n >= (uint64_t)m;
}

struct pqc_ctx {
  int device = 0;
  unsigned alg_mask = 0;
  size_t max_batch = 0;
  size_t hbm_budget = 0;
  size_t ld = 0;  // SoA leading dimension of the plain-path polynomials

  // plain Kyber path: SoA int16 polynomials and per-element byte buffers
  int16_t* kpolys = nullptr;
  int kpolys_count = 0;
  uint8_t* kbytes = nullptr;
  size_t kbytes_per_elem = 0;

  // fused Kyber path: PRF outputs, (2K+1) x 128 bytes per request (largest K held)
  uint8_t* kprf = nullptr;
  size_t kprf_per_elem = 0;
  // mixed Kyber batches: request indices bucketed by K (3 x max_batch) and the
  // three bucket sizes
  uint32_t* kmix_order = nullptr;
  uint32_t* kmix_counts = nullptr;

  // NewHope-1024 CPA plain path: 8 SoA uint16 polynomials of 1024 coefficients
  // and 64 bytes per element (seed expansion)
  uint16_t* nhpolys = nullptr;
  uint8_t* nhbytes = nullptr;
  // NewHope-1024 CCA-KEM: per-element row of hashes, flags and the re-encryption;
  // a side stream (default priority) and two events fork the ciphertext hash of
  // decapsulation off the caller's stream and join it back
  uint8_t* nhkem = nullptr;
  hipStream_t nh_side = nullptr;
  hipEvent_t nh_fork = nullptr, nh_join = nullptr;

  // FrodoKEM-976: device-resident key, expanded A (uint16 and MFMA digit
  // fragments), per-element scratch (S' fragments, E', E'', GEMM output, mu, k,
  // F sponge states, fail flags)
  uint16_t* fr_a16 = nullptr;
  uint16_t* fr_abuf = nullptr;  // A of up to 128 per-element keys (crypto_kem_*_batch)
  int8_t* fr_frag = nullptr;
  int8_t* fr_sfrag = nullptr;
  uint8_t* fr_sk = nullptr;
  uint8_t* fr_pk = nullptr;
  uint8_t* fr_pkh = nullptr;
  uint8_t* fr_elem = nullptr;
  int fr_loaded = 0, fr_has_sk = 0;
  // side streams (default priority) and events that pipeline batches under the
  // loaded key in chunks (frodo.hip)
  hipStream_t fr_side[3] = {nullptr, nullptr, nullptr};
  hipEvent_t fr_fork = nullptr, fr_join[3] = {nullptr, nullptr, nullptr};

  // staging for host-pointer calls: device buffers, and (hostio.cpp) the
  // pinned double buffers with their copy threads
  uint8_t* stage = nullptr;
  size_t stage_bytes = 0;
  void* hostio = nullptr;

  uint32_t* status = nullptr;  // device word: bit 0 = a sampler hit its block cap

  size_t allocated = 0;  // bytes of device memory owned by the context
  int cu_count = 0;
  int fused_grid_override = 0;
  pqc_tuning tune = {256, 128, 4, 0, 64, 0, 0, 0, 0, 0, 0};
  // fused decapsulation in 32-request tiles: per-wave A^T scratch for kat_waves waves
  int16_t* kat = nullptr;
  size_t kat_waves = 0;
  size_t kat_region = 0;  // int16 per wave region (sized for the largest K held)
  void* tuner = nullptr;  // autotuner samples and the learned cost model (tuner.cpp)
};

namespace pqc {

inline int hip_err(hipError_t e) { return e == hipSuccess ? PQC_OK : PQC_ERR_DEVICE; }

// host-buffer transfers (hostio.cpp): pageable memory is staged through pinned
// double buffers in chunks, pqc_host_alloc memory is copied directly; both
// are ordered on stream s, and the caller synchronises before reading dst
int host_to_device(pqc_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s);
int device_to_host(pqc_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s);
bool host_is_pinned(const void* p, size_t bytes);
void hostio_free(pqc_ctx* c);

// plain per-op Kyber path (kyber_plain.hip); all pointers are device pointers
int kyber_plain_keypair(pqc_ctx* c, int k, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* coins,
                        hipStream_t s);
int kyber_plain_enc(pqc_ctx* c, int k, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk,
                    const uint8_t* coins, hipStream_t s);
int kyber_plain_dec(pqc_ctx* c, int k, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk,
                    hipStream_t s);

// fused persistent Kyber path (kyber_fused.hip)
int kyber_fused_keypair(pqc_ctx* c, int k, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* coins, hipStream_t s);
int kyber_fused_enc(pqc_ctx* c, int k, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk,
                    const uint8_t* coins, hipStream_t s);
int kyber_fused_dec(pqc_ctx* c, int k, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk,
                    hipStream_t s);

// mixed Kyber-512/768/1024 batch through the fused kernels (kyber_fused.hip)
int kyber_fused_mixed(pqc_ctx* c, int dec, size_t n, const pqc_kem_req* reqs, uint8_t* ct, size_t ct_bytes,
                      const uint8_t* key, size_t key_bytes, uint8_t* ss, size_t ss_bytes, const uint8_t* coins,
                      size_t coins_bytes, hipStream_t s);

// NewHope-N CPA (newhope.hip), N = 512 or 1024: fused per-element kernels, or
// the per-operation kernels when plain
int newhope_keypair(pqc_ctx* c, int N, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* seed, bool plain,
                    hipStream_t s);
int newhope_enc(pqc_ctx* c, int N, size_t n, uint8_t* ct, const uint8_t* m, const uint8_t* pk, const uint8_t* coin,
                bool plain, hipStream_t s);
int newhope_dec(pqc_ctx* c, int N, size_t n, uint8_t* m, const uint8_t* ct, const uint8_t* sk, bool plain,
                hipStream_t s);
// NewHope-N CCA-KEM over the CPA kernels (newhope.hip); keypair coins 64 bytes;
// one scratch row of kNewhopeKemBytes per element serves both N
constexpr int kNewhopeKemBytes = 2432;
int newhope_cca_keypair(pqc_ctx* c, int N, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* coins, bool plain,
                        hipStream_t s);
int newhope_cca_enc(pqc_ctx* c, int N, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk, const uint8_t* coins,
                    bool plain, hipStream_t s);
int newhope_cca_dec(pqc_ctx* c, int N, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk, bool plain,
                    hipStream_t s);
int newhope_poly_transform(int op, size_t n, uint16_t* polys, hipStream_t s);
int newhope_poly_sample(size_t n, uint16_t* polys, const uint8_t* seeds, size_t seed_stride, int nonce,
                        hipStream_t s);

// FrodoKEM-976 (frodo.hip)
size_t frodo_bytes_per_elem();
int frodo_load_key(pqc_ctx* c, const uint8_t* key, int is_sk, hipStream_t s);
int frodo_keypair(pqc_ctx* c, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* rnd, hipStream_t s);
int frodo_enc(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* mu, hipStream_t s);
int frodo_dec(pqc_ctx* c, size_t n, uint8_t* ss, const uint8_t* ct, hipStream_t s);
int frodo_products(pqc_ctx* c, size_t n, uint16_t* bpp, uint16_t* v, hipStream_t s);
int frodo_gemm(pqc_ctx* c, size_t n, const uint8_t* ct, hipStream_t s);
int frodo_dec_batch(pqc_ctx* c, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk, hipStream_t s);
// key store: secret keys (pk and H(pk) inside) + expanded A; kid = key index per request
int frodo_expand_a(size_t n_keys, uint16_t* a_out, const uint8_t* sk_store, hipStream_t s);
int frodo_dec_store(pqc_ctx* c, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk_store,
                    const uint16_t* a_store, const uint32_t* kid, hipStream_t s);
int frodo_enc_store(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* mu, const uint8_t* sk_store,
                    const uint16_t* a_store, const uint32_t* kid, hipStream_t s);
// dst[r] = src[idx[r]] for n records of `bytes` (multiple of 8) bytes (gather.hip)
int zero_words(uint32_t* p, int n, hipStream_t s);
size_t tile32_scratch_bytes(int k);  // per wave, A^T scratch of the 32-request decapsulation tiles
int gather_records(size_t n, uint8_t* dst, size_t bytes, const uint8_t* src, size_t src_stride, const uint32_t* idx,
                   hipStream_t s);
// the library's deterministic generator: SHAKE256("pqc-mi355x" || seed || stream || index)
void host_drbg(uint8_t* out, size_t len, uint64_t seed, uint64_t stream, uint64_t index);

// NIST PQC known-answer-test DRBG (nist_drbg.cpp): active after randombytes_init
bool nist_drbg_active();
void nist_drbg_disable();
void nist_drbg_generate(uint8_t* x, size_t len);
int frodo_enc_batch(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk, const uint8_t* mu,
                    hipStream_t s);
int frodo_enc_batch_strided(pqc_ctx* c, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk, size_t pk_stride,
                            const uint8_t* mu, hipStream_t s);

// component kernels exposed for tests (kyber_fused.hip)
int kyber_poly_transform(int op, size_t n, int16_t* polys, hipStream_t s);
int keccak_batch(int op, size_t n, uint8_t* out, size_t outlen, const uint8_t* in, size_t inlen,
                 hipStream_t s);

void pqc_tuner_free(pqc_ctx* c);

size_t kyber_plain_polys(int kmax);
size_t kyber_plain_bytes_per_elem(int kmax);

}  // namespace pqc
