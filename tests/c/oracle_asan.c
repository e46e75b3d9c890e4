/*
 * Drives every scheme of the CPU oracle under AddressSanitizer and
 * UndefinedBehaviorSanitizer: Kyber K = 2, 3, 4, NewHope-1024/512 (CPA and
 * CCA) and FrodoKEM-976, single calls and the threaded batch drivers (the CPU
 * baseline's code), with honest and tampered ciphertexts.  Any out-of-bounds
 * access, use after free, signed overflow, misaligned access or shift out of
 * range aborts the run; exit status 0 means every round trip held.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int fails = 0;
#define CHECK(c)                                                \
  do {                                                          \
    if (!(c)) {                                                 \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      fails++;                                                  \
    }                                                           \
  } while (0)

static void fill(uint8_t *p, size_t n, uint64_t seed, uint64_t stream) { oracle_drbg(p, n, seed, stream, 0); }

static void kyber(int k) {
  oracle_kyber_params p;
  CHECK(oracle_kyber_params_get(k, &p) == 0);
  enum { N = 6 };
  uint8_t *pk = malloc(N * p.pk), *sk = malloc(N * p.sk), *ct = malloc(N * p.ct);
  uint8_t coins[N * 64], ecoins[N * 32], ss[N * 32], ss2[N * 32];
  fill(coins, sizeof coins, 1, k);
  fill(ecoins, sizeof ecoins, 2, k);
  CHECK(oracle_kyber_keypair_batch(k, N, pk, sk, coins, 3) == 0);
  CHECK(oracle_kyber_enc_batch(k, N, ct, ss, pk, ecoins, 3) == 0);
  CHECK(oracle_kyber_dec_batch(k, N, ss2, ct, sk, 3) == 0);
  CHECK(memcmp(ss, ss2, sizeof ss) == 0);
  /* single calls agree with the batch drivers; a tampered ct is rejected */
  uint8_t ss1[32];
  CHECK(oracle_kyber_dec(k, ss1, ct, sk) == 0 && memcmp(ss1, ss, 32) == 0);
  ct[p.ct - 1] ^= 1;
  CHECK(oracle_kyber_dec(k, ss1, ct, sk) == 0 && memcmp(ss1, ss, 32) != 0);
  int16_t a[256], b[256], r[256];
  for (int i = 0; i < 256; i++) a[i] = (int16_t)((i * 97) % 3329), b[i] = (int16_t)((i * 31 + 5) % 3329);
  oracle_kyber_ntt(a);
  oracle_kyber_ntt(b);
  oracle_kyber_basemul(r, a, b);
  oracle_kyber_invntt(r);
  uint8_t buf[512];
  for (int d = 1; d <= 11; d++) {
    int16_t back[256];
    for (int i = 0; i < 256; i++) a[i] = (int16_t)((i * 13) % 3329);
    oracle_kyber_poly_compress(buf, a, d);
    oracle_kyber_poly_decompress(back, buf, d);
  }
  free(pk), free(sk), free(ct);
}

static void newhope(int big) {
  const size_t pkb = big ? 1824 : 928, skb = big ? 3680 : 1888, ctb = big ? 2208 : 1120;
  uint8_t *pk = malloc(pkb), *sk = malloc(skb), *ct = malloc(ctb);
  uint8_t coins[64], ec[32], ss[32], ss2[32];
  fill(coins, 64, 3, big);
  fill(ec, 32, 4, big);
  if (big) {
    CHECK(oracle_newhope_cca_keypair(pk, sk, coins) == 0);
    CHECK(oracle_newhope_cca_enc(ct, ss, pk, ec) == 0);
    CHECK(oracle_newhope_cca_dec(ss2, ct, sk) == 0);
  } else {
    CHECK(oracle_newhope512_cca_keypair(pk, sk, coins) == 0);
    CHECK(oracle_newhope512_cca_enc(ct, ss, pk, ec) == 0);
    CHECK(oracle_newhope512_cca_dec(ss2, ct, sk) == 0);
  }
  CHECK(memcmp(ss, ss2, 32) == 0);
  ct[0] ^= 0x80;
  if (big) CHECK(oracle_newhope_cca_dec(ss2, ct, sk) == 0);
  else CHECK(oracle_newhope512_cca_dec(ss2, ct, sk) == 0);
  CHECK(memcmp(ss, ss2, 32) != 0);
  /* the CPA triple driver on threads */
  enum { N = 4 };
  const size_t cpk = big ? 1824 : 928, csk = big ? 1792 : 896, cct = big ? 2176 : 1088;
  uint8_t *tpk = malloc(N * cpk), *tsk = malloc(N * csk), *tct = malloc(N * cct);
  uint8_t seeds[N * 32], msg[N * 32], coin[N * 32], m2[N * 32];
  fill(seeds, sizeof seeds, 5, big);
  fill(msg, sizeof msg, 6, big);
  fill(coin, sizeof coin, 7, big);
  if (big) CHECK(oracle_newhope_triple_batch(N, tpk, tsk, tct, m2, seeds, msg, coin, 2) == 0);
  else CHECK(oracle_newhope512_triple_batch(N, tpk, tsk, tct, m2, seeds, msg, coin, 2) == 0);
  CHECK(memcmp(msg, m2, sizeof msg) == 0);
  free(pk), free(sk), free(ct), free(tpk), free(tsk), free(tct);
}

static void frodo(void) {
  enum { N = 2 };
  uint8_t *pk = malloc(15632), *sk = malloc(31296), *ct = malloc(N * 15744);
  uint8_t *sks = malloc(N * 31296);
  uint8_t coins[64], mu[N * 24], ss[N * 24], ss2[N * 24];
  fill(coins, 64, 8, 0);
  fill(mu, sizeof mu, 9, 0);
  CHECK(oracle_frodo_keypair_derand(pk, sk, coins) == 0);
  uint8_t *pks = malloc(N * 15632);
  for (int i = 0; i < N; i++) memcpy(pks + i * 15632, pk, 15632), memcpy(sks + i * 31296, sk, 31296);
  CHECK(oracle_frodo_enc_batch(N, ct, ss, pks, mu, 2) == 0);
  CHECK(oracle_frodo_dec_batch(N, ss2, ct, sks, 2) == 0);
  CHECK(memcmp(ss, ss2, sizeof ss) == 0);
  ct[100] ^= 1;
  CHECK(oracle_frodo_dec(ss2, ct, sk) == 0 && memcmp(ss, ss2, 24) != 0);
  uint16_t s[64];
  for (int i = 0; i < 64; i++) s[i] = (uint16_t)(i * 1021);
  oracle_frodo_sample_n(s, 64);
  free(pk), free(sk), free(ct), free(sks), free(pks);
}

int main(void) {
  uint8_t h[64];
  oracle_sha3_512(h, (const uint8_t *)"abc", 3);
  oracle_shake128(h, 64, (const uint8_t *)"", 0);
  for (int k = 2; k <= 4; k++) kyber(k);
  newhope(1);
  newhope(0);
  frodo();
  if (fails) {
    fprintf(stderr, "oracle_asan: %d checks failed\n", fails);
    return 1;
  }
  printf("oracle_asan: all checks passed under ASan+UBSan\n");
  return 0;
}
