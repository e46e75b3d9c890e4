/*
 * FIPS 202 restatement for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Follows the sponge used by the reference:
 *   Keccak-f[1600] ........ Kyber_GPU_Batched/fips202.cu:393-447
 *   keccak_absorb ......... Kyber_GPU_Batched/fips202.cu:462-498 (zero state,
 *                           full-rate blocks, pad byte p, final 0x80)
 *   keccak_squeezeblocks .. Kyber_GPU_Batched/fips202.cu:513-530
 *   shake128/shake256 ..... Kyber_GPU_Batched/fips202.cu:543-604 (p = 0x1F)
 *   sha3_256/sha3_512 ..... Kyber_GPU_Batched/fips202.cu:615-689 (p = 0x06)
 * The round constants and rho offsets are derived here from their FIPS 202
 * definitions (the rc(t) LFSR of Algorithm 5 and the (t+1)(t+2)/2 offsets of
 * Algorithm 2) instead of being tabulated.
 */
#include <string.h>
#include "oracle.h"

static uint64_t g_rc[24];
static unsigned g_rho[25];   /* rotation offset of lane x + 5y */
static int g_init = 0;

/* FIPS 202 Algorithm 5: rc(t) from the LFSR x^8 + x^6 + x^5 + x^4 + 1 */
static int rc_bit(unsigned t) {
    unsigned r = 1;
    t %= 255;
    for (unsigned i = 0; i < t; i++) {
        r <<= 1;
        if (r & 0x100) r ^= 0x171;
    }
    return r & 1;
}

static void init_tables(void) {
    if (g_init) return;
    for (unsigned ir = 0; ir < 24; ir++) {
        uint64_t rc = 0;
        for (unsigned j = 0; j <= 6; j++)
            if (rc_bit(j + 7 * ir)) rc |= (uint64_t)1 << ((1u << j) - 1);
        g_rc[ir] = rc;
    }
    /* FIPS 202 Algorithm 2: (x,y) = (1,0); offset (t+1)(t+2)/2; (x,y) <- (y, 2x+3y) */
    memset(g_rho, 0, sizeof g_rho);
    unsigned x = 1, y = 0;
    for (unsigned t = 0; t < 24; t++) {
        g_rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        unsigned nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx; y = ny;
    }
    g_init = 1;
}

static uint64_t rotl64(uint64_t v, unsigned n) {
    n &= 63;
    return n ? (v << n) | (v >> (64 - n)) : v;
}

/* Keccak-f[1600] by the step mappings theta, rho, pi, chi, iota
 * (Kyber_GPU_Batched/fips202.cu:393-447). */
void oracle_keccakf1600(uint64_t a[25]) {
    init_tables();
    uint64_t b[25], c[5], d[5];
    for (int ir = 0; ir < 24; ir++) {
        for (int x = 0; x < 5; x++)
            c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; x++)
            d[x] = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) a[i] ^= d[i % 5];
        /* rho + pi: B[y, 2x+3y] = rot(A[x,y], r[x,y]) */
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(a[x + 5 * y], g_rho[x + 5 * y]);
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++)
                a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        a[0] ^= g_rc[ir];
    }
}

static uint64_t load64le(const uint8_t *x) {
    uint64_t r = 0;
    for (int i = 0; i < 8; i++) r |= (uint64_t)x[i] << (8 * i);
    return r;
}

static void store64le(uint8_t *x, uint64_t v) {
    for (int i = 0; i < 8; i++) x[i] = (uint8_t)(v >> (8 * i));
}

/* keccak_absorb (fips202.cu:462-498): zero state, absorb full blocks, then
 * the tail with domain byte p and the final 0x80 in the last rate byte. */
static void absorb(uint64_t s[25], unsigned rate, const uint8_t *m, size_t mlen, uint8_t p) {
    uint8_t t[200];
    memset(s, 0, 25 * sizeof(uint64_t));
    while (mlen >= rate) {
        for (unsigned i = 0; i < rate / 8; i++) s[i] ^= load64le(m + 8 * i);
        oracle_keccakf1600(s);
        m += rate;
        mlen -= rate;
    }
    memset(t, 0, rate);
    memcpy(t, m, mlen);
    t[mlen] = p;
    t[rate - 1] |= 0x80;
    for (unsigned i = 0; i < rate / 8; i++) s[i] ^= load64le(t + 8 * i);
}

/* keccak_squeezeblocks (fips202.cu:513-530) followed by a partial block, as
 * in shake256 (fips202.cu:577-604). */
static void squeeze(uint8_t *out, size_t outlen, uint64_t s[25], unsigned rate) {
    uint8_t t[200];
    while (outlen > 0) {
        oracle_keccakf1600(s);
        for (unsigned i = 0; i < rate / 8; i++) store64le(t + 8 * i, s[i]);
        size_t n = outlen < rate ? outlen : rate;
        memcpy(out, t, n);
        out += n;
        outlen -= n;
    }
}

void oracle_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
    uint64_t s[25];
    absorb(s, 168, in, inlen, 0x1F);
    squeeze(out, outlen, s, 168);
}

void oracle_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
    uint64_t s[25];
    absorb(s, 136, in, inlen, 0x1F);
    squeeze(out, outlen, s, 136);
}

void oracle_sha3_256(uint8_t out[32], const uint8_t *in, size_t inlen) {
    uint64_t s[25];
    absorb(s, 136, in, inlen, 0x06);
    squeeze(out, 32, s, 136);
}

void oracle_sha3_512(uint8_t out[64], const uint8_t *in, size_t inlen) {
    uint64_t s[25];
    absorb(s, 72, in, inlen, 0x06);
    squeeze(out, 64, s, 72);
}

void oracle_drbg(uint8_t *out, size_t len, uint64_t seed, uint64_t stream, uint64_t index) {
    uint8_t in[10 + 24];
    memcpy(in, "pqc-mi355x", 10);
    store64le(in + 10, seed);
    store64le(in + 18, stream);
    store64le(in + 26, index);
    oracle_shake256(out, len, in, sizeof in);
}
