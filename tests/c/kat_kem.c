/*
 * A NIST-harness-shaped KEM caller, in C, built against include/pqc_api.h with
 * nothing but the names a caller of the reference's api.h uses:
 * CRYPTO_SECRETKEYBYTES, CRYPTO_PUBLICKEYBYTES, CRYPTO_CIPHERTEXTBYTES,
 * CRYPTO_BYTES, CRYPTO_ALGNAME, crypto_kem_keypair / enc / dec and the
 * harness's randombytes_init / randombytes
 * (Kyber_GPU_Batched/api.h:10-22, FrodoKEM_GPU/api.h:11-24; the call sequence
 * of Kyber_GPU_Batched/main.cu:23-64 and FrodoKEM_GPU/test_KEM976.cu:33-74,
 * with the known-answer draw order of NIST's PQCgenKAT_kem.c).
 *
 * The parameter set is chosen at compile time with -DPQC_NIST_ALG=<id>, the
 * only line that differs from a build against the reference's api.h.
 *
 * Usage: kat_kem <count>  - writes a .rsp file to stdout: the DRBG is seeded
 * with bytes 0..47, `count` 48-byte seeds are drawn, and for each seed the DRBG
 * is re-seeded and one keypair / encapsulation / decapsulation runs.  Exit
 * status 0 when every decapsulated secret equals the encapsulated one.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pqc_api.h"

#define MAX_COUNT 100

static void put_hex(const char *name, const unsigned char *p, size_t n) {
  size_t i;
  printf("%s = ", name);
  for (i = 0; i < n; i++) printf("%02X", p[i]);
  printf("\n");
}

int main(int argc, char **argv) {
  static unsigned char seeds[MAX_COUNT][48];
  unsigned char entropy[48];
  unsigned char pk[CRYPTO_PUBLICKEYBYTES];
  unsigned char sk[CRYPTO_SECRETKEYBYTES];
  unsigned char ct[CRYPTO_CIPHERTEXTBYTES];
  unsigned char ss[CRYPTO_BYTES], ss1[CRYPTO_BYTES];
  int count = argc > 1 ? atoi(argv[1]) : 1;
  int i, rc, bad = 0;

  if (count < 1 || count > MAX_COUNT) {
    fprintf(stderr, "count must be 1..%d\n", MAX_COUNT);
    return 2;
  }
  for (i = 0; i < 48; i++) entropy[i] = (unsigned char)i;
  randombytes_init(entropy, NULL, 256);
  for (i = 0; i < count; i++) randombytes(seeds[i], 48);

  printf("# %s\n\n", CRYPTO_ALGNAME);
  for (i = 0; i < count; i++) {
    randombytes_init(seeds[i], NULL, 256);
    if ((rc = crypto_kem_keypair(pk, sk)) != 0) {
      fprintf(stderr, "crypto_kem_keypair returned <%d>\n", rc);
      return 1;
    }
    if ((rc = crypto_kem_enc(ct, ss, pk)) != 0) {
      fprintf(stderr, "crypto_kem_enc returned <%d>\n", rc);
      return 1;
    }
    if ((rc = crypto_kem_dec(ss1, ct, sk)) != 0) {
      fprintf(stderr, "crypto_kem_dec returned <%d>\n", rc);
      return 1;
    }
    if (memcmp(ss, ss1, CRYPTO_BYTES) != 0) {
      fprintf(stderr, "crypto_kem_dec returned bad 'ss' value at count %d\n", i);
      bad = 1;
    }
    printf("count = %d\n", i);
    put_hex("seed", seeds[i], 48);
    put_hex("pk", pk, CRYPTO_PUBLICKEYBYTES);
    put_hex("sk", sk, CRYPTO_SECRETKEYBYTES);
    put_hex("ct", ct, CRYPTO_CIPHERTEXTBYTES);
    put_hex("ss", ss, CRYPTO_BYTES);
    printf("\n");
  }
  return bad;
}
