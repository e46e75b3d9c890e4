/*
 * Drives the library's host code through every path that runs without a GPU,
 * for the sanitizer build (make -C tests/c asan-run): argument checks, the
 * size table, the NIST AES-256 CTR_DRBG (randombytes_init / randombytes) and
 * the SHAKE256 DRBG from several threads at once, the drop-ins' context pool
 * from several threads at once, context and server creation, and error strings.
 * Without a device every GPU entry point must return PQC_ERR_NO_DEVICE (or an
 * argument error) and never compute on the CPU.  Run only where no GPU is
 * visible; exit status 0 when every check holds.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pqc_api.h"

static int fails = 0;
#define CHECK(c)                                                            \
  do {                                                                      \
    if (!(c)) {                                                             \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      fails++;                                                              \
    }                                                                       \
  } while (0)

/* First 48-byte seed of NIST's PQC KAT files: DRBG seeded with bytes 0..47. */
static const char *NIST_COUNT0 =
    "061550234D158C5EC95595FE04EF7A25767F2E24CC2BC479D09D86DC9ABCFDE7056A8C266F9EF97ED08541DBD2E1FFA1";

static void *dropin_thread(void *arg) {
  int alg = (int)(size_t)arg;
  static unsigned char pk[15632], sk[31296];
  unsigned char ct[16], ss[32];
  int i, rc;
  for (i = 0; i < 20; i++) {
    switch (alg % 4) {
      case 0: rc = pqc_kyber512_crypto_kem_dec(ss, ct, sk); break;
      case 1: rc = pqc_frodo976_crypto_kem_keypair(pk, sk); break;
      case 2: rc = pqc_newhope1024_crypto_kem_enc(ct, ss, pk); break;
      default: rc = crypto_kem_dec(ss, ct, sk); break;
    }
    if (rc != PQC_ERR_NO_DEVICE) return (void *)1;
  }
  return NULL;
}

static void *drbg_thread(void *arg) {
  unsigned char buf[97];
  int i;
  (void)arg;
  for (i = 0; i < 200; i++)
    if (pqc_randombytes(buf, 1 + (size_t)(i % 97)) != PQC_OK) return (void *)1;
  return NULL;
}

int main(void) {
  int n = -1, i;
  size_t pk, sk, ct, ss;
  char hex[97];
  unsigned char e[48], seed[48];
  pthread_t th[8];

  CHECK(pqc_device_count(&n) == PQC_OK || n <= 0);
  if (n > 0) {
    fprintf(stderr, "a GPU is visible: this driver covers the no-device paths only\n");
    return 0;
  }

  /* sizes */
  CHECK(pqc_alg_sizes(PQC_ALG_KYBER768, &pk, &sk, &ct, &ss) == PQC_OK && pk == 1184 && sk == 2400 && ct == 1088 &&
        ss == 32);
  CHECK(pqc_alg_sizes(PQC_ALG_FRODO976, &pk, &sk, &ct, &ss) == PQC_OK && pk == 15632 && ss == 24);
  CHECK(pqc_alg_sizes(99, &pk, &sk, &ct, &ss) != PQC_OK);
  CHECK(pqc_alg_sizes(PQC_ALG_KYBER512, NULL, NULL, NULL, NULL) != PQC_OK || 1);

  /* NIST DRBG: the published count-0 seed */
  for (i = 0; i < 48; i++) e[i] = (unsigned char)i;
  randombytes_init(e, NULL, 256);
  CHECK(randombytes(seed, 48) == 0);
  for (i = 0; i < 48; i++) sprintf(hex + 2 * i, "%02X", seed[i]);
  CHECK(strcmp(hex, NIST_COUNT0) == 0);
  /* concurrent draws (the DRBG is one global state behind a lock) */
  for (i = 0; i < 4; i++) pthread_create(&th[i], NULL, drbg_thread, NULL);
  for (i = 0; i < 4; i++) {
    void *r;
    pthread_join(th[i], &r);
    CHECK(r == NULL);
  }
  pqc_randombytes_seed(7);
  for (i = 0; i < 4; i++) pthread_create(&th[i], NULL, drbg_thread, NULL);
  for (i = 0; i < 4; i++) {
    void *r;
    pthread_join(th[i], &r);
    CHECK(r == NULL);
  }
  pqc_randombytes_system();
  CHECK(pqc_randombytes(NULL, 4) == PQC_ERR_INVALID_ARG);
  CHECK(pqc_randombytes(NULL, 0) == PQC_OK);

  /* contexts and calls without a device */
  {
    pqc_ctx *c = (pqc_ctx *)0x1;
    CHECK(pqc_ctx_create(&c, PQC_ALG_MASK(PQC_ALG_KYBER768), 16, 0) == PQC_ERR_NO_DEVICE);
    CHECK(c == NULL);
    CHECK(pqc_ctx_create(NULL, PQC_ALG_MASK(PQC_ALG_KYBER768), 16, 0) == PQC_ERR_INVALID_ARG);
    CHECK(pqc_ctx_create(&c, PQC_ALG_MASK(PQC_ALG_KYBER768), 0, 0) == PQC_ERR_INVALID_ARG);
    pqc_ctx_destroy(NULL);
    CHECK(pqc_ctx_device_bytes(NULL) == 0);
    CHECK(crypto_kem_dec_batch(NULL, PQC_ALG_KYBER768, 1, NULL, NULL, NULL, 0, NULL) != PQC_OK);
    CHECK(pqc_ctx_take_status(NULL, NULL, NULL) != PQC_OK);
  }
  {
    pqc_tuning t;
    CHECK(pqc_ctx_get_tuning(NULL, &t) != PQC_OK);
    CHECK(pqc_ctx_put_tuning(NULL, &t) != PQC_OK);
  }
  /* the drop-ins' context pool from several threads at once */
  for (i = 0; i < 8; i++) pthread_create(&th[i], NULL, dropin_thread, (void *)(size_t)i);
  for (i = 0; i < 8; i++) {
    void *r;
    pthread_join(th[i], &r);
    CHECK(r == NULL);
  }
  {
    void *p = NULL;
    CHECK(pqc_host_alloc(64, &p) != PQC_OK || pqc_host_free(p) == PQC_OK);
    CHECK(pqc_host_free((void *)0x1234) == PQC_ERR_INVALID_ARG);
  }
  for (i = -7; i <= 0; i++) CHECK(pqc_strerror(i) != NULL && strlen(pqc_strerror(i)) > 0);
  CHECK(pqc_version() != NULL);

  if (fails) {
    fprintf(stderr, "abi_nodevice: %d checks failed\n", fails);
    return 1;
  }
  printf("abi_nodevice: all checks passed under ASan+UBSan\n");
  return 0;
}
