// Host-side check of the mixed-batch descriptor validation
// (pqc_amd/csrc/req_check.h), which the bucketing and low-latency kernels run
// on the device: offsets near 2^64 must be rejected, not wrap into the arena.
#include <stdio.h>

#include "../../pqc_amd/csrc/req_check.h"

static int fails = 0;
#define CHECK(c)                                                            \
  do {                                                                      \
    if (!(c)) {                                                             \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      fails++;                                                              \
    }                                                                       \
  } while (0)

int main() {
  const uint64_t M = ~0ull;
  // spans
  CHECK(pqc_span_ok(0, 32, 32));
  CHECK(!pqc_span_ok(8, 32, 32));
  CHECK(pqc_span_ok(32, 0, 32));
  CHECK(!pqc_span_ok(40, 0, 32));
  CHECK(!pqc_span_ok(4, 8, 64));                 // unaligned
  CHECK(!pqc_span_ok(M - 1023, 32, 1 << 20));    // 2^64 - 1024: wraps with off + need
  CHECK(!pqc_span_ok(M - 7, 1088, 1 << 20));     // 2^64 - 8
  CHECK(!pqc_span_ok(M - 7, 8, M));              // ends exactly at 2^64: outside an arena of 2^64 - 1 bytes
  CHECK(pqc_span_ok(M - 7 - 8, 8, M));
  // descriptors
  pqc_kem_req d = {PQC_ALG_KYBER768, 0, 0, 0, 0, 0};
  const unsigned all = PQC_ALG_MASK(PQC_ALG_KYBER512) | PQC_ALG_MASK(PQC_ALG_KYBER768) | PQC_ALG_MASK(PQC_ALG_KYBER1024);
  CHECK(pqc_req_k(d, true, all, 1088, 2400, 32, 0) == 3);
  CHECK(pqc_req_k(d, true, all, 1087, 2400, 32, 0) == 0);
  CHECK(pqc_req_k(d, false, all, 1088, 1184, 32, 32) == 3);
  CHECK(pqc_req_k(d, false, all, 1088, 1184, 32, 31) == 0);  // encapsulation needs coins
  CHECK(pqc_req_k(d, true, PQC_ALG_MASK(PQC_ALG_KYBER512), 1088, 2400, 32, 0) == 0);
  d.ss_off = M - 1023;
  CHECK(pqc_req_k(d, true, all, 1088, 2400, 1 << 20, 0) == 0);
  d.ss_off = 0;
  d.key_off = M - 7;
  CHECK(pqc_req_k(d, true, all, 1088, 1 << 20, 32, 0) == 0);
  d.key_off = 0;
  d.ct_off = M - 1023;
  CHECK(pqc_req_k(d, false, all, 1 << 20, 1184, 32, 32) == 0);
  d.ct_off = 0;
  d.coins_off = M - 1023;
  CHECK(pqc_req_k(d, false, all, 1088, 1184, 32, 1 << 20) == 0);
  CHECK(pqc_req_k(d, true, all, 1088, 2400, 32, 0) == 3);  // decapsulation ignores coins
  d.coins_off = 0;
  d.alg = PQC_ALG_FRODO976;
  CHECK(pqc_req_k(d, true, ~0u, 1 << 20, 1 << 20, 32, 0) == 0);
  d.alg = PQC_ALG_KYBER1024;
  CHECK(pqc_req_k(d, true, all, 1568, 3168, 32, 0) == 4);
  d.alg = PQC_ALG_KYBER512;
  CHECK(pqc_req_k(d, false, all, 736, 800, 32, 32) == 2);
  if (fails) {
    fprintf(stderr, "req_check_test: %d checks failed\n", fails);
    return 1;
  }
  printf("req_check_test: all checks passed\n");
  return 0;
}
