/*
 * CPU oracle for the batched lattice-KEM engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libpqc_amd.so) never links, loads or calls it.
 *
 * It is a plain, portable, per-element C restatement of the reference's
 * arithmetic (nainag/PQC): no SIMD intrinsics, no GPU.  Each function cites
 * the reference file:line it follows.  Reference quirks Q1..Q17 (races,
 * stride mismatches, non-standard SHAKE in the Frodo GPU sampler, secret-
 * dependent branches) are deliberately NOT reproduced.
 */
#ifndef PQC_ORACLE_H
#define PQC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FIPS 202 (Kyber_GPU_Batched/fips202.cu) ---------------------------- */
void oracle_keccakf1600(uint64_t st[25]);
void oracle_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void oracle_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void oracle_sha3_256(uint8_t out[32], const uint8_t *in, size_t inlen);
void oracle_sha3_512(uint8_t out[64], const uint8_t *in, size_t inlen);

/* ---- seeded deterministic DRBG ---------------------------------------------
 * out = SHAKE256("pqc-mi355x" || le64(seed) || le64(stream) || le64(index)) */
void oracle_drbg(uint8_t *out, size_t len, uint64_t seed, uint64_t stream, uint64_t index);

/* ---- Kyber round 2, K in {2,3,4} ------------------------------------------ */
typedef struct {
    int k, du, dv;
    size_t polyvec_bytes, polyvec_cbytes, poly_cbytes;
    size_t indcpa_pk, indcpa_sk, ct, pk, sk;
} oracle_kyber_params;

int oracle_kyber_params_get(int k, oracle_kyber_params *p);

/* internal pieces, exported for the pinning tests */
int16_t oracle_kyber_montgomery_reduce(int32_t a);
int16_t oracle_kyber_barrett_reduce(int16_t a);
int16_t oracle_kyber_csubq(int16_t a);
void oracle_kyber_zetas(int16_t zetas[128], int16_t zetas_inv[128]);
void oracle_kyber_ntt(int16_t r[256]);      /* ntt_p + poly_reduce (poly_ntt) */
void oracle_kyber_ntt_raw(int16_t r[256]);  /* ntt_p only */
void oracle_kyber_invntt(int16_t r[256]);
void oracle_kyber_basemul(int16_t r[256], const int16_t a[256], const int16_t b[256]);
void oracle_kyber_cbd(int16_t r[256], const uint8_t buf[128]);
unsigned oracle_kyber_rej_uniform(int16_t *r, unsigned len, const uint8_t *buf, unsigned buflen);
void oracle_kyber_gen_matrix(int k, int16_t *a /* k*k*256 */, const uint8_t seed[32], int transposed);
void oracle_kyber_poly_compress(uint8_t *r, const int16_t a[256], int d);
void oracle_kyber_poly_decompress(int16_t r[256], const uint8_t *a, int d);

int oracle_kyber_indcpa_keypair(int k, uint8_t *pk, uint8_t *sk, const uint8_t seed[32]);
int oracle_kyber_indcpa_enc(int k, uint8_t *ct, const uint8_t m[32], const uint8_t *pk,
                            const uint8_t coins[32]);
int oracle_kyber_indcpa_dec(int k, uint8_t m[32], const uint8_t *ct, const uint8_t *sk);

/* CCA KEM, explicit randomness (coins64 = d || z for keypair, coins32 for enc) */
int oracle_kyber_keypair_derand(int k, uint8_t *pk, uint8_t *sk, const uint8_t coins64[64]);
int oracle_kyber_enc_derand(int k, uint8_t *ct, uint8_t ss[32], const uint8_t *pk,
                            const uint8_t coins32[32]);
int oracle_kyber_dec(int k, uint8_t ss[32], const uint8_t *ct, const uint8_t *sk);

/* threaded batch drivers (the CPU baseline): element-contiguous buffers */
int oracle_kyber_dec_batch(int k, size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk,
                           int nthreads);
int oracle_kyber_enc_batch(int k, size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk,
                           const uint8_t *coins, int nthreads);
int oracle_kyber_keypair_batch(int k, size_t n, uint8_t *pk, uint8_t *sk, const uint8_t *coins,
                               int nthreads);

/* ---- NewHope-1024 CPA-PKE (NewHope_NIST_GPU) --------------------------------
 * pk 1824, sk 1792, ct 2176, message 32 bytes. */
void oracle_newhope_tables(uint16_t omegas_inv[512], uint16_t gammas[1024], uint16_t gammas_inv[1024]);
uint16_t oracle_newhope_montgomery_reduce(uint32_t a);
uint16_t oracle_newhope_barrett_reduce(uint16_t a);
void oracle_newhope_ntt(uint16_t *r);     /* poly_ntt: mul gammas, ntt */
void oracle_newhope_invntt(uint16_t *r);  /* poly_invntt */
void oracle_newhope_poly_uniform(uint16_t *a, const uint8_t seed[32]);
void oracle_newhope_poly_sample(uint16_t *r, const uint8_t seed[32], uint8_t nonce);
int oracle_newhope_cpa_keypair(uint8_t *pk, uint8_t *sk, const uint8_t seed[32]);
int oracle_newhope_cpa_enc(uint8_t *c, const uint8_t m[32], const uint8_t *pk, const uint8_t coin[32]);
int oracle_newhope_cpa_dec(uint8_t m[32], const uint8_t *c, const uint8_t *sk);
int oracle_newhope_cca_keypair(uint8_t *pk, uint8_t *sk, const uint8_t coins[64]);
int oracle_newhope_cca_enc(uint8_t *ct, uint8_t ss[32], const uint8_t *pk, const uint8_t coins[32]);
int oracle_newhope_cca_dec(uint8_t ss[32], const uint8_t *ct, const uint8_t *sk);
/* NewHope-512: the same functions at N = 512 (newhope512.c) */
void oracle_newhope512_tables(uint16_t omegas_inv[256], uint16_t gammas[512], uint16_t gammas_inv[512]);
uint16_t oracle_newhope512_montgomery_reduce(uint32_t a);
uint16_t oracle_newhope512_barrett_reduce(uint16_t a);
void oracle_newhope512_ntt(uint16_t *r);
void oracle_newhope512_invntt(uint16_t *r);
void oracle_newhope512_poly_uniform(uint16_t *a, const uint8_t seed[32]);
void oracle_newhope512_poly_sample(uint16_t *r, const uint8_t seed[32], uint8_t nonce);
int oracle_newhope512_cpa_keypair(uint8_t *pk, uint8_t *sk, const uint8_t seed[32]);
int oracle_newhope512_cpa_enc(uint8_t *c, const uint8_t m[32], const uint8_t *pk, const uint8_t coin[32]);
int oracle_newhope512_cpa_dec(uint8_t m[32], const uint8_t *c, const uint8_t *sk);
int oracle_newhope512_cca_keypair(uint8_t *pk, uint8_t *sk, const uint8_t coins[64]);
int oracle_newhope512_cca_enc(uint8_t *ct, uint8_t ss[32], const uint8_t *pk, const uint8_t coins[32]);
int oracle_newhope512_cca_dec(uint8_t ss[32], const uint8_t *ct, const uint8_t *sk);
int oracle_newhope512_triple_batch(size_t n, uint8_t *pk, uint8_t *sk, uint8_t *ct, uint8_t *m2,
                                   const uint8_t *seed, const uint8_t *msg, const uint8_t *coin, int nthreads);
int oracle_newhope_triple_batch(size_t n, uint8_t *pk, uint8_t *sk, uint8_t *ct, uint8_t *m2,
                                const uint8_t *seed, const uint8_t *msg, const uint8_t *coin, int nthreads);

/* ---- FrodoKEM-976-SHAKE (FrodoKEM_GPU) --------------------------------------
 * pk 15632, sk 31296, ct 15744, ss 24; n = 976, nbar = 8, q = 2^16. */
void oracle_frodo_sample_n(uint16_t *s, size_t n);
void oracle_frodo_pack(uint8_t *out, size_t outlen, const uint16_t *in, size_t inlen, unsigned lsb);
void oracle_frodo_unpack(uint16_t *out, size_t outlen, const uint8_t *in, size_t inlen, unsigned lsb);
void oracle_frodo_gen_a(uint16_t *A, const uint8_t seedA[16]);
void oracle_frodo_mul_add_sa_plus_e(uint16_t *out, const uint16_t *A, const uint16_t *S, const uint16_t *E);
int oracle_frodo_keypair_derand(uint8_t *pk, uint8_t *sk, const uint8_t randomness[64]);
int oracle_frodo_enc_derand(uint8_t *ct, uint8_t ss[24], const uint8_t *pk, const uint8_t mu[24]);
int oracle_frodo_enc_derand_a(uint8_t *ct, uint8_t ss[24], const uint8_t *pk, const uint8_t mu[24],
                              const uint16_t *A);
int oracle_frodo_dec(uint8_t ss[24], const uint8_t *ct, const uint8_t *sk);
int oracle_frodo_dec_a(uint8_t ss[24], const uint8_t *ct, const uint8_t *sk, const uint16_t *A);
int oracle_frodo_enc_batch(size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu, int nthreads);
int oracle_frodo_dec_batch(size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
