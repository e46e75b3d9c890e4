// CPU check of the exact FP32 arithmetic mod q (pqc_amd/csrc/kyber_f32.h).
//
// The primitives use only IEEE-754 binary32 fma, multiply and add/subtract in
// round-to-nearest-even, which the host computes exactly as v_fma_f32 /
// v_mul_f32 / v_add_f32 do on the GPU, so checking them here checks the
// device arithmetic bit for bit.  Exhaustive where the domain is small:
//   - mulc, for every centred twiddle and every |x| <= kMulcMax: the result is
//     x z mod q and within 1665 of 0;
//   - mulc_wide, the same for every |x| <= kMulcWideMax;
//   - reduce_f, for every |x| <= 2^24;
//   - canon_f, for every |x| <= kCanonMax: the residue in [0, q);
//   - compress_f<D>, D = 1, 3, 4, 5, 10, 11, for every representative within
//     kCompressMax(D): equal to the integer compress of the residue, written
//     as in the reference (((x << D) + q / 2) / q) & (2^D - 1));
// and by simulation for the transform schedules: a scalar forward transform
// over the same seven layers (Cooley-Tukey with the wide product where
// fwd_wide_mask says) and a scalar inverse transform (Gentleman-Sande with
// kInvReduce), on extreme and random inputs, against the same transforms in
// exact 64-bit modular arithmetic, with every intermediate checked against
// the bound the schedule assumes.
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cmath>
#include <random>

#include "../../pqc_amd/csrc/kyber_f32.h"

using namespace pqc::kyber::f32;

static int fails = 0;
#define CHECK(c, ...)                                    \
  do {                                                   \
    if (!(c)) {                                          \
      if (fails < 20) {                                  \
        std::fprintf(stderr, "%s:%d: ", __FILE__, __LINE__); \
        std::fprintf(stderr, __VA_ARGS__);               \
        std::fprintf(stderr, "\n");                      \
      }                                                  \
      fails++;                                           \
    }                                                    \
  } while (0)

static int64_t modq(int64_t a) { return ((a % kQ) + kQ) % kQ; }
static bool is_int(float f) { return std::nearbyint(f) == f; }

static int64_t pw(int64_t b, int e) {
  int64_t r = 1;
  while (e--) r = r * b % kQ;
  return r;
}
static unsigned brv7(unsigned i) {
  unsigned r = 0;
  for (int b = 0; b < 7; b++) r |= ((i >> b) & 1u) << (6 - b);
  return r;
}
static int centred(int64_t a) {
  int r = (int)modq(a);
  return r > kQ / 2 ? r - kQ : r;
}

static void check_mulc() {
  // every centred twiddle value, and both signs
  for (int z = -kQ / 2; z <= kQ / 2; ++z) {
    const float zf = (float)z, zq = (float)z / kQf;
    for (int x = -kMulcWideMax; x <= kMulcWideMax; ++x) {
      const bool narrow = x >= -kMulcMax && x <= kMulcMax;
      if (narrow) {
        const float r = mulc((float)x, zf, zq);
        CHECK(is_int(r) && std::fabs(r) <= kOut && modq((int64_t)r - (int64_t)x * z) == 0, "mulc(%d, %d) = %f", x,
              z, r);
      }
      const float w = mulc_wide((float)x, zf, zq);
      CHECK(is_int(w) && std::fabs(w) <= kOut && modq((int64_t)w - (int64_t)x * z) == 0, "mulc_wide(%d, %d) = %f",
            x, z, w);
    }
  }
}

static void check_reduce() {
  for (int x = -kReduceMax; x <= kReduceMax; ++x) {
    const float r = reduce_f((float)x);
    if (!(is_int(r) && std::fabs(r) <= kOut && modq((int64_t)r - x) == 0)) {
      CHECK(false, "reduce_f(%d) = %f", x, r);
    }
  }
}

static void check_canon() {
  for (int x = -kCanonMax; x <= kCanonMax; ++x) {
    const uint32_t got = canon_f((float)x);
    CHECK(got == (uint32_t)modq(x), "canon_f(%d) = %u", x, got);
  }
}

template <int D>
static void check_compress() {
  const int m = kCompressMax(D);
  for (int c = -m; c <= m; ++c) {
    const uint32_t x = (uint32_t)modq(c);
    const uint32_t want = (((x << D) + kQ / 2) / kQ) & ((1u << D) - 1);
    const uint32_t got = compress_f<D>((float)c);
    CHECK(got == want, "compress_f<%d>(%d) = %u, want %u", D, c, got, want);
  }
}

// ---- transform schedules ----------------------------------------------------

static int zeta_fwd(int k) { return centred(pw(17, (int)brv7(k))); }           // plain forward zetas
static int zeta_inv(int k) { return centred(-pw(17, 128 - (int)brv7(k))); }    // plain inverse zetas (order of use)

// Forward: len 128 -> 2, as poly_ntt's loop; values bounded by b0 on input.
static void fwd_sim(const int (&in)[256], int b0) {
  const unsigned wide = fwd_wide_mask(b0);
  float f[256];
  int64_t e[256];
  for (int i = 0; i < 256; ++i) {
    f[i] = (float)in[i];
    e[i] = in[i];
  }
  int k = 1, li = 0;
  for (int len = 128; len >= 2; len >>= 1, ++li) {
    for (int start = 0; start < 256; start += 2 * len) {
      const int z = zeta_fwd(k++);
      const float zf = (float)z, zq = (float)z / kQf;
      for (int j = start; j < start + len; ++j) {
        const float lim = ((wide >> li) & 1) ? (float)kMulcWideMax : (float)kMulcMax;
        CHECK(std::fabs(f[j + len]) <= lim, "forward layer %d input %f over %f", li, f[j + len], lim);
        const float t = ((wide >> li) & 1) ? mulc_wide(f[j + len], zf, zq) : mulc(f[j + len], zf, zq);
        const int64_t te = modq(e[j + len] * z);
        f[j + len] = f[j] - t;
        f[j] = f[j] + t;
        e[j + len] = modq(e[j] - te);
        e[j] = modq(e[j] + te);
      }
    }
  }
  for (int i = 0; i < 256; ++i) {
    CHECK(is_int(f[i]) && std::fabs(f[i]) <= fwd_out_bound(b0), "forward output %f", f[i]);
    CHECK(modq((int64_t)f[i]) == e[i], "forward output %d: %f vs %lld", i, f[i], (long long)e[i]);
  }
}

// Inverse: len 2 -> 128 as poly_invntt's loop, then the final scaling.
static void inv_sim(const int (&in)[256]) {
  float f[256];
  int64_t e[256];
  for (int i = 0; i < 256; ++i) {
    f[i] = (float)in[i];
    e[i] = in[i];
  }
  int k = 0, li = 0;
  for (int len = 2; len <= 128; len <<= 1, ++li) {
    const bool red = (kInvReduce >> li) & 1;
    for (int start = 0; start < 256; start += 2 * len) {
      const int z = zeta_inv(k++);
      const float zf = (float)z, zq = (float)z / kQf;
      for (int j = start; j < start + len; ++j) {
        const float x = f[j], y = f[j + len];
        CHECK(std::fabs(x - y) <= kMulcMax, "inverse layer %d difference %f", li, x - y);
        const float s = x + y;
        f[j] = red ? reduce_f(s) : s;
        f[j + len] = mulc(x - y, zf, zq);
        const int64_t ex = e[j], ey = e[j + len];
        e[j] = modq(ex + ey);
        e[j + len] = modq((ex - ey) * z);
      }
    }
  }
  const int sc = centred(pw(128, kQ - 2));  // 1/128 mod q
  for (int i = 0; i < 256; ++i) {
    CHECK(std::fabs(f[i]) <= kMulcMax, "inverse scaling input %f", f[i]);
    const float r = mulc(f[i], (float)sc, (float)sc / kQf);
    CHECK(is_int(r) && std::fabs(r) <= kOut && modq((int64_t)r) == modq(e[i] * sc), "inverse output %d", i);
  }
}

static void check_schedules() {
  std::mt19937 rng(7);
  const int b0s[3] = {2, kQ, kOut};  // noise, decompressed [0, q], reduced
  for (int b0 : b0s) {
    int in[256];
    for (int v : {b0, -b0}) {  // all-extreme inputs
      for (int i = 0; i < 256; ++i) in[i] = (b0 == kQ && v < 0) ? 0 : v;
      fwd_sim(in, b0);
    }
    for (int rep = 0; rep < 2000; ++rep) {
      for (int i = 0; i < 256; ++i) {
        const int lo = b0 == kQ ? 0 : -b0;
        in[i] = lo + (int)(rng() % (uint32_t)(b0 - lo + 1));
        if (rep % 4 == 1) in[i] = (rng() & 1) ? b0 : lo;  // extremes mixed
      }
      fwd_sim(in, b0);
    }
  }
  int in[256];
  for (int v : {kOut, -kOut}) {
    for (int i = 0; i < 256; ++i) in[i] = v;
    inv_sim(in);
    for (int i = 0; i < 256; ++i) in[i] = (i & 1) ? v : -v;
    inv_sim(in);
  }
  for (int rep = 0; rep < 2000; ++rep) {
    for (int i = 0; i < 256; ++i) {
      in[i] = -kOut + (int)(rng() % (2 * kOut + 1));
      if (rep % 4 == 1) in[i] = (rng() & 1) ? kOut : -kOut;
    }
    inv_sim(in);
  }
}

int main() {
  check_mulc();
  check_reduce();
  check_canon();
  check_compress<1>();
  check_compress<3>();
  check_compress<4>();
  check_compress<5>();
  check_compress<10>();
  check_compress<11>();
  check_schedules();
  if (fails) {
    std::fprintf(stderr, "%d checks failed\n", fails);
    return 1;
  }
  std::printf("f32 arithmetic mod q: all checks hold\n");
  return 0;
}
