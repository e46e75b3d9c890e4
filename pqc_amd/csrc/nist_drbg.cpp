// The NIST PQC known-answer-test randomness source: CTR_DRBG over AES-256 with
// no derivation function (NIST SP 800-90A section 10.2.1), keyed by a 48-byte
// seed, exported under the names of the NIST harness's rng.h
// (randombytes_init / randombytes) so that PQCgenKAT_kem.c links against this
// library unchanged.  After randombytes_init the library's non-derand entry
// points (crypto_kem_keypair, crypto_kem_enc, the *_batch variants without
// coins) draw their coins from it, call by call, as the NIST harness does.
//
// AES-256 follows FIPS 197: the S-box is computed from its definition
// (multiplicative inverse in GF(2^8) mod x^8+x^4+x^3+x+1, then the affine map
// with constant 0x63), 14 rounds, 60-word key schedule.
#include <mutex>
#include <string.h>

#include "internal.h"

namespace pqc {
namespace {

struct Aes256 {
  uint8_t sbox[256];
  Aes256() {
    // log/antilog tables over the generator 3
    uint8_t exp[256], log[256];
    uint8_t x = 1;
    for (int i = 0; i < 255; ++i) {
      exp[i] = x;
      log[x] = (uint8_t)i;
      x = (uint8_t)(x ^ (x << 1) ^ ((x & 0x80) ? 0x1B : 0));  // x * 3
    }
    for (int v = 0; v < 256; ++v) {
      const uint8_t inv = v ? exp[(255 - log[v]) % 255] : 0;
      uint8_t s = inv;
      for (int r = 1; r <= 4; ++r) s ^= (uint8_t)((inv << r) | (inv >> (8 - r)));
      sbox[v] = (uint8_t)(s ^ 0x63);
    }
  }
  static uint8_t xt(uint8_t a) { return (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1B : 0)); }

  // round keys as 15 x 16 bytes in state order (byte r of column c at 4c + r)
  void expand(const uint8_t key[32], uint8_t rk[240]) const {
    memcpy(rk, key, 32);
    uint8_t rcon = 1;
    for (int i = 8; i < 60; ++i) {
      uint8_t t[4];
      memcpy(t, rk + 4 * (i - 1), 4);
      if (i % 8 == 0) {
        const uint8_t t0 = t[0];
        t[0] = (uint8_t)(sbox[t[1]] ^ rcon);
        t[1] = sbox[t[2]];
        t[2] = sbox[t[3]];
        t[3] = sbox[t0];
        rcon = xt(rcon);
      } else if (i % 8 == 4) {
        for (int j = 0; j < 4; ++j) t[j] = sbox[t[j]];
      }
      for (int j = 0; j < 4; ++j) rk[4 * i + j] = (uint8_t)(rk[4 * (i - 8) + j] ^ t[j]);
    }
  }

  void encrypt(const uint8_t rk[240], const uint8_t in[16], uint8_t out[16]) const {
    uint8_t s[16];
    for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(in[i] ^ rk[i]);
    for (int round = 1; round <= 14; ++round) {
      uint8_t t[16];
      // SubBytes + ShiftRows: row r of column c comes from column c + r
      for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) t[4 * c + r] = sbox[s[4 * ((c + r) & 3) + r]];
      if (round < 14) {
        for (int c = 0; c < 4; ++c) {
          const uint8_t a0 = t[4 * c], a1 = t[4 * c + 1], a2 = t[4 * c + 2], a3 = t[4 * c + 3];
          const uint8_t all = (uint8_t)(a0 ^ a1 ^ a2 ^ a3);
          // b_r = a_r ^ all ^ 2(a_r ^ a_{r+1}), the MixColumns matrix [2 3 1 1] rotated
          s[4 * c] = (uint8_t)(a0 ^ all ^ xt((uint8_t)(a0 ^ a1)));
          s[4 * c + 1] = (uint8_t)(a1 ^ all ^ xt((uint8_t)(a1 ^ a2)));
          s[4 * c + 2] = (uint8_t)(a2 ^ all ^ xt((uint8_t)(a2 ^ a3)));
          s[4 * c + 3] = (uint8_t)(a3 ^ all ^ xt((uint8_t)(a3 ^ a0)));
        }
      } else {
        memcpy(s, t, 16);
      }
      for (int i = 0; i < 16; ++i) s[i] ^= rk[16 * round + i];
    }
    memcpy(out, s, 16);
  }
};

const Aes256& aes() {
  static const Aes256 a;
  return a;
}

struct CtrDrbg {
  uint8_t key[32];
  uint8_t v[16];
  uint64_t reseed_counter;
};

void incr_v(uint8_t v[16]) {
  for (int j = 15; j >= 0; --j)
    if (++v[j]) break;
}

// CTR_DRBG_Update (SP 800-90A 10.2.1.2) with seedlen = 48 bytes.
void drbg_update(CtrDrbg& d, const uint8_t* provided) {
  uint8_t rk[240], tmp[48];
  aes().expand(d.key, rk);
  for (int i = 0; i < 3; ++i) {
    incr_v(d.v);
    aes().encrypt(rk, d.v, tmp + 16 * i);
  }
  if (provided)
    for (int i = 0; i < 48; ++i) tmp[i] ^= provided[i];
  memcpy(d.key, tmp, 32);
  memcpy(d.v, tmp + 32, 16);
}

std::mutex g_mu;
CtrDrbg g_drbg;
bool g_on = false;

}  // namespace

bool nist_drbg_active() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_on;
}

void nist_drbg_disable() {
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = false;
}

// CTR_DRBG_Generate (10.2.1.5.1) without additional input: one update per call.
void nist_drbg_generate(uint8_t* x, size_t len) {
  std::lock_guard<std::mutex> lk(g_mu);
  uint8_t rk[240], block[16];
  aes().expand(g_drbg.key, rk);
  while (len) {
    incr_v(g_drbg.v);
    aes().encrypt(rk, g_drbg.v, block);
    const size_t take = len < 16 ? len : 16;
    memcpy(x, block, take);
    x += take;
    len -= take;
  }
  drbg_update(g_drbg, nullptr);
  ++g_drbg.reseed_counter;
}

}  // namespace pqc

extern "C" {

void randombytes_init(unsigned char* entropy_input, unsigned char* personalization_string, int security_strength) {
  (void)security_strength;
  uint8_t seed[48];
  memcpy(seed, entropy_input, 48);
  if (personalization_string)
    for (int i = 0; i < 48; ++i) seed[i] ^= personalization_string[i];
  std::lock_guard<std::mutex> lk(pqc::g_mu);
  memset(pqc::g_drbg.key, 0, 32);
  memset(pqc::g_drbg.v, 0, 16);
  pqc::drbg_update(pqc::g_drbg, seed);
  pqc::g_drbg.reseed_counter = 1;
  pqc::g_on = true;
}

int randombytes(unsigned char* x, unsigned long long xlen) {
  if (!x && xlen) return PQC_ERR_INVALID_ARG;
  if (!pqc::nist_drbg_active()) return pqc_randombytes(x, (size_t)xlen);
  pqc::nist_drbg_generate(x, (size_t)xlen);
  return 0;
}

}  // extern "C"
