// Validation of one mixed-batch request descriptor (pqc_kem_req), shared by the
// bucketing kernel, the per-request low-latency kernel and host-side tests.
//
// An offset is accepted when it is 8-byte aligned and the whole buffer
// [off, off + need) lies inside its arena of `bytes` bytes.  The test is
// written as `off <= bytes && need <= bytes - off`, which cannot wrap for any
// 64-bit offset (the form `off + need <= bytes` wraps for offsets near 2^64 and
// would accept a buffer that starts below the arena).
#pragma once
#include <stdint.h>

#include "../../include/pqc_api.h"

#if defined(__HIPCC__)
#define PQC_HD __host__ __device__
#else
#define PQC_HD
#endif

PQC_HD inline bool pqc_span_ok(uint64_t off, uint64_t need, uint64_t bytes) {
  return (off & 7) == 0 && off <= bytes && need <= bytes - off;
}

// Returns K (2, 3, 4) for a valid descriptor, 0 for one to skip.  `dec` selects
// the key size (sk for decapsulation, pk for encapsulation); encapsulation also
// needs 32 coin bytes.  Sizes are the round-2 Kyber ones (ct 736 / 1088 / 1568,
// sk 1632 / 2400 / 3168, pk 800 / 1184 / 1568).
PQC_HD inline int pqc_req_k(const pqc_kem_req& d, bool dec, unsigned alg_mask, uint64_t ct_bytes,
                            uint64_t key_bytes, uint64_t ss_bytes, uint64_t coins_bytes) {
  const bool known = (d.alg == PQC_ALG_KYBER512 || d.alg == PQC_ALG_KYBER768 || d.alg == PQC_ALG_KYBER1024) &&
                     (alg_mask & PQC_ALG_MASK(d.alg));
  if (!known) return 0;
  const int k = d.alg + 1;
  const uint64_t ctb = k == 2 ? 736 : k == 3 ? 1088 : 1568;
  const uint64_t keyb = dec ? (k == 2 ? 1632 : k == 3 ? 2400 : 3168) : (k == 2 ? 800 : k == 3 ? 1184 : 1568);
  const bool ok = pqc_span_ok(d.ct_off, ctb, ct_bytes) && pqc_span_ok(d.key_off, keyb, key_bytes) &&
                  pqc_span_ok(d.ss_off, 32, ss_bytes) && (dec || pqc_span_ok(d.coins_off, 32, coins_bytes));
  return ok ? k : 0;
}
