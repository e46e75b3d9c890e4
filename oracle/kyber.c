/*
 * Kyber (round 2) restatement for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Per-element C following the reference's batched CUDA sources, with the
 * batch dimension removed and K made a run-time parameter:
 *   params ............ Kyber_GPU_Batched/params.h:20-45
 *   reduce ............ Kyber_GPU_Batched/reduce.cu:21-66
 *   ntt/invntt/basemul  Kyber_GPU_Batched/ntt.cu:152-259 (tables: generator
 *                       comment ntt.cu:11-47)
 *   cbd ............... Kyber_GPU_Batched/cbd.cu:42-63
 *   poly/polyvec ...... Kyber_GPU_Batched/poly.cu, polyvec.cu
 *   indcpa ............ Kyber_GPU_Batched/indcpa.cu:153-403 (pack, rej_uniform,
 *                       gen_matrix), 474-733 (keypair/enc/dec)
 *   symmetric ......... Kyber_GPU_Batched/symmetric-fips202.cu:20-79,
 *                       symmetric.h:26-33 (H=SHA3-256, G=SHA3-512,
 *                       PRF=SHAKE256(key||nonce), KDF=SHAKE256(.,32))
 *   verify/cmov ....... Kyber_GPU_Batched/verify.cu:15-50 (compare restated
 *                       branch-free; same result)
 * The reference hard-wires Kyber-1024 compression (Q2).  The round-2 widths
 * (du,dv) = (10,3), (10,4), (11,5) for K = 2,3,4 and the CCA KEM glue
 * (kem.cu is missing, Q4) follow the round-2 Kyber specification.  All the
 * packings are least-significant-bit-first, which is what the reference's
 * hand-unrolled 5- and 11-bit packers (poly.cu:22-69, polyvec.cu:18-82) do.
 */
#include <string.h>
#include <stdlib.h>
#include <pthread.h>
#include "oracle.h"

#define N 256
#define Q 3329
#define QINV 62209 /* q^-1 mod 2^16, reduce.h:11-12 */
#define MONT 2285  /* 2^16 mod q */
#define SYM 32

int oracle_kyber_params_get(int k, oracle_kyber_params *p) {
    if (k < 2 || k > 4) return -1;
    p->k = k;
    p->du = (k == 4) ? 11 : 10;
    p->dv = (k == 2) ? 3 : (k == 3) ? 4 : 5;
    p->polyvec_bytes = (size_t)k * 384;
    p->polyvec_cbytes = (size_t)k * 32 * p->du;
    p->poly_cbytes = (size_t)32 * p->dv;
    p->indcpa_pk = p->polyvec_bytes + SYM;
    p->indcpa_sk = p->polyvec_bytes;
    p->ct = p->polyvec_cbytes + p->poly_cbytes;
    p->pk = p->indcpa_pk;
    p->sk = p->indcpa_sk + p->indcpa_pk + 2 * SYM;
    return 0;
}

/* ---- reduce.cu ------------------------------------------------------------ */
int16_t oracle_kyber_montgomery_reduce(int32_t a) { /* reduce.cu:21-31 */
    int16_t u = (int16_t)(a * QINV);
    int32_t t = (int32_t)u * Q;
    t = a - t;
    t >>= 16;
    return (int16_t)t;
}

int16_t oracle_kyber_barrett_reduce(int16_t a) { /* reduce.cu:43-51 */
    const int32_t v = (1U << 26) / Q + 1;
    int32_t t = v * a;
    t >>= 26;
    t *= Q;
    return (int16_t)(a - t);
}

int16_t oracle_kyber_csubq(int16_t a) { /* reduce.cu:62-66 */
    a = (int16_t)(a - Q);
    a = (int16_t)(a + ((a >> 15) & Q));
    return a;
}

static int16_t fqmul(int16_t a, int16_t b) { /* ntt.cu:81-84 */
    return oracle_kyber_montgomery_reduce((int32_t)a * b);
}

/* ---- twiddles: the generator of ntt.cu:11-47 -------------------------------
 * tmp[i] = MONT * 17^i (Montgomery form), zetas[i] = tmp[brv7(i)],
 * zetas_inv[k] = -tmp[128 - brv7(j)] over the inverse-layer order, and
 * zetas_inv[127] = MONT^2 (q-1) ((q-1)/128) mod q.  All kept in [0, q). */
static int16_t g_zetas[128], g_zetas_inv[128];
static int g_zinit = 0;

static unsigned brv7(unsigned i) {
    unsigned r = 0;
    for (int b = 0; b < 7; b++) r |= ((i >> b) & 1u) << (6 - b);
    return r;
}

static void init_zetas(void) {
    if (g_zinit) return;
    int32_t tmp[128];
    tmp[0] = MONT;
    for (int i = 1; i < 128; i++) tmp[i] = (int32_t)((int64_t)tmp[i - 1] * 17 % Q);
    for (int i = 0; i < 128; i++) g_zetas[i] = (int16_t)tmp[brv7(i)];
    int k = 0;
    for (int i = 64; i >= 1; i >>= 1)
        for (int j = i; j < 2 * i; j++) g_zetas_inv[k++] = (int16_t)((Q - tmp[128 - brv7(j)]) % Q);
    g_zetas_inv[127] = (int16_t)((int64_t)MONT * ((int64_t)MONT * (Q - 1) % Q * ((Q - 1) / 128) % Q) % Q);
    g_zinit = 1;
}

void oracle_kyber_zetas(int16_t zetas[128], int16_t zetas_inv[128]) {
    init_zetas();
    memcpy(zetas, g_zetas, sizeof g_zetas);
    memcpy(zetas_inv, g_zetas_inv, sizeof g_zetas_inv);
}

/* ---- ntt.cu:152-224, poly.cu:153-190 -------------------------------------- */
void oracle_kyber_ntt_raw(int16_t r[256]) {
    init_zetas();
    unsigned k = 1;
    for (unsigned len = 128; len >= 2; len >>= 1) {
        for (unsigned start = 0; start < 256; start += 2 * len) {
            int16_t zeta = g_zetas[k++];
            for (unsigned j = start; j < start + len; j++) {
                int16_t t = fqmul(zeta, r[j + len]);
                r[j + len] = (int16_t)(r[j] - t);
                r[j] = (int16_t)(r[j] + t);
            }
        }
    }
}

static void poly_reduce(int16_t r[256]) { /* poly.cu:269-277 */
    for (int i = 0; i < N; i++) r[i] = oracle_kyber_barrett_reduce(r[i]);
}

static void poly_csubq(int16_t r[256]) { /* poly.cu:287-295 */
    for (int i = 0; i < N; i++) r[i] = oracle_kyber_csubq(r[i]);
}

void oracle_kyber_ntt(int16_t r[256]) { /* poly_ntt, poly.cu:153-157 */
    oracle_kyber_ntt_raw(r);
    poly_reduce(r);
}

void oracle_kyber_invntt(int16_t r[256]) { /* invntt_p, ntt.cu:197-224 */
    init_zetas();
    unsigned k = 0;
    for (unsigned len = 2; len <= 128; len <<= 1) {
        for (unsigned start = 0; start < 256; start += 2 * len) {
            int16_t zeta = g_zetas_inv[k++];
            for (unsigned j = start; j < start + len; j++) {
                int16_t t = r[j];
                r[j] = oracle_kyber_barrett_reduce((int16_t)(t + r[j + len]));
                r[j + len] = (int16_t)(t - r[j + len]);
                r[j + len] = fqmul(zeta, r[j + len]);
            }
        }
    }
    for (int j = 0; j < 256; j++) r[j] = fqmul(r[j], g_zetas_inv[127]);
}

/* basemul2 (ntt.cu:249-259) over the whole polynomial (poly_basemul,
 * poly.cu:204-225): pairs 4i use +zetas[64+i], pairs 4i+2 use -zetas[64+i]. */
static void basemul_pair(int16_t r[2], const int16_t a[2], const int16_t b[2], int16_t zeta) {
    r[0] = fqmul(a[1], b[1]);
    r[0] = fqmul(r[0], zeta);
    r[0] = (int16_t)(r[0] + fqmul(a[0], b[0]));
    r[1] = fqmul(a[0], b[1]);
    r[1] = (int16_t)(r[1] + fqmul(a[1], b[0]));
}

void oracle_kyber_basemul(int16_t r[256], const int16_t a[256], const int16_t b[256]) {
    init_zetas();
    for (int i = 0; i < N / 4; i++) {
        basemul_pair(r + 4 * i, a + 4 * i, b + 4 * i, g_zetas[64 + i]);
        basemul_pair(r + 4 * i + 2, a + 4 * i + 2, b + 4 * i + 2, (int16_t)-g_zetas[64 + i]);
    }
}

static void poly_add(int16_t r[256], const int16_t a[256], const int16_t b[256]) {
    for (int i = 0; i < N; i++) r[i] = (int16_t)(a[i] + b[i]);
}

static void poly_sub(int16_t r[256], const int16_t a[256], const int16_t b[256]) {
    for (int i = 0; i < N; i++) r[i] = (int16_t)(a[i] - b[i]);
}

/* polyvec_pointwise_acc (polyvec.cu:161-177) */
static void pointwise_acc(int k, int16_t r[256], const int16_t *a, const int16_t *b) {
    int16_t t[256];
    oracle_kyber_basemul(r, a, b);
    for (int i = 1; i < k; i++) {
        oracle_kyber_basemul(t, a + 256 * i, b + 256 * i);
        poly_add(r, r, t);
    }
    poly_reduce(r);
}

static void poly_frommont(int16_t r[256]) { /* poly.cu:235-247 */
    const int16_t f = (int16_t)((1ULL << 32) % Q);
    for (int i = 0; i < N; i++) r[i] = oracle_kyber_montgomery_reduce((int32_t)r[i] * f);
}

/* ---- cbd.cu:42-63 ----------------------------------------------------------- */
void oracle_kyber_cbd(int16_t r[256], const uint8_t buf[128]) {
    for (int i = 0; i < N / 8; i++) {
        uint32_t t = (uint32_t)buf[4 * i] | ((uint32_t)buf[4 * i + 1] << 8) |
                     ((uint32_t)buf[4 * i + 2] << 16) | ((uint32_t)buf[4 * i + 3] << 24);
        uint32_t d = t & 0x55555555u;
        d += (t >> 1) & 0x55555555u;
        for (int j = 0; j < 8; j++) {
            int16_t a = (int16_t)((d >> (4 * j)) & 3);
            int16_t b = (int16_t)((d >> (4 * j + 2)) & 3);
            r[8 * i + j] = (int16_t)(a - b);
        }
    }
}

/* poly_getnoise (poly.cu:129-142) with shake256_prf (symmetric-fips202.cu:67-79) */
static void poly_getnoise(int16_t r[256], const uint8_t seed[32], uint8_t nonce) {
    uint8_t ext[SYM + 1], buf[128];
    memcpy(ext, seed, SYM);
    ext[SYM] = nonce;
    oracle_shake256(buf, sizeof buf, ext, sizeof ext);
    oracle_kyber_cbd(r, buf);
}

/* ---- bit packing (poly.cu:22-116, polyvec.cu:18-115) ------------------------- */
static void pack_bits(uint8_t *r, const uint16_t *v, int n, int d) {
    size_t nbytes = ((size_t)n * d + 7) / 8;
    memset(r, 0, nbytes);
    size_t bit = 0;
    for (int i = 0; i < n; i++)
        for (int b = 0; b < d; b++, bit++)
            if ((v[i] >> b) & 1) r[bit >> 3] |= (uint8_t)(1u << (bit & 7));
}

static void unpack_bits(uint16_t *v, const uint8_t *a, int n, int d) {
    size_t bit = 0;
    for (int i = 0; i < n; i++) {
        uint16_t x = 0;
        for (int b = 0; b < d; b++, bit++) x |= (uint16_t)(((a[bit >> 3] >> (bit & 7)) & 1) << b);
        v[i] = x;
    }
}

/* poly_compress / polyvec_compress: csubq then round(x * 2^d / q) mod 2^d */
void oracle_kyber_poly_compress(uint8_t *r, const int16_t a_in[256], int d) {
    int16_t a[256];
    uint16_t t[256];
    memcpy(a, a_in, sizeof a);
    poly_csubq(a);
    for (int i = 0; i < N; i++)
        t[i] = (uint16_t)(((((uint32_t)a[i] << d) + Q / 2) / Q) & ((1u << d) - 1));
    pack_bits(r, t, N, d);
}

void oracle_kyber_poly_decompress(int16_t r[256], const uint8_t *a, int d) {
    uint16_t t[256];
    unpack_bits(t, a, N, d);
    for (int i = 0; i < N; i++) r[i] = (int16_t)(((uint32_t)t[i] * Q + (1u << (d - 1))) >> d);
}

static void poly_tobytes(uint8_t r[384], const int16_t a_in[256]) { /* poly.cu:79-95 */
    int16_t a[256];
    uint16_t t[256];
    memcpy(a, a_in, sizeof a);
    poly_csubq(a);
    for (int i = 0; i < N; i++) t[i] = (uint16_t)a[i];
    pack_bits(r, t, N, 12);
}

static void poly_frombytes(int16_t r[256], const uint8_t a[384]) { /* poly.cu:106-116 */
    uint16_t t[256];
    unpack_bits(t, a, N, 12);
    for (int i = 0; i < N; i++) r[i] = (int16_t)t[i];
}

static void poly_frommsg(int16_t r[256], const uint8_t msg[32]) { /* poly.cu:367-386 */
    for (int i = 0; i < SYM; i++)
        for (int j = 0; j < 8; j++) {
            uint16_t mask = (uint16_t)(-(int)((msg[i] >> j) & 1));
            r[8 * i + j] = (int16_t)(mask & ((Q + 1) / 2));
        }
}

static void poly_tomsg(uint8_t msg[32], const int16_t a_in[256]) { /* poly.cu:396-417 */
    int16_t a[256];
    memcpy(a, a_in, sizeof a);
    poly_csubq(a);
    for (int i = 0; i < SYM; i++) {
        msg[i] = 0;
        for (int j = 0; j < 8; j++) {
            uint16_t t = (uint16_t)(((((uint16_t)a[8 * i + j]) << 1) + Q / 2) / Q) & 1;
            msg[i] |= (uint8_t)(t << j);
        }
    }
}

/* ---- rej_uniform / gen_matrix (indcpa.cu:290-403) --------------------------- */
unsigned oracle_kyber_rej_uniform(int16_t *r, unsigned len, const uint8_t *buf, unsigned buflen) {
    unsigned ctr = 0, pos = 0;
    while (ctr < len && pos + 2 <= buflen) {
        uint16_t val = (uint16_t)(buf[pos] | ((uint16_t)buf[pos + 1] << 8));
        pos += 2;
        if (val < 19 * Q) {
            val = (uint16_t)(val - (val >> 12) * Q);
            r[ctr++] = (int16_t)val;
        }
    }
    return ctr;
}

void oracle_kyber_gen_matrix(int k, int16_t *a, const uint8_t seed[32], int transposed) {
    /* 4 blocks first ((530 + 168) / 168), then one block at a time.  The
     * fallback writes continue at r + ctr (the reference's pointer offset by
     * whole polynomials, Q3, is not reproduced). */
    const unsigned maxnblocks = (530 + 168) / 168;
    uint8_t buf[168 * 4];
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) {
            uint64_t s[25];
            uint8_t ext[SYM + 2];
            memcpy(ext, seed, SYM);
            ext[SYM] = (uint8_t)(transposed ? i : j);
            ext[SYM + 1] = (uint8_t)(transposed ? j : i);
            /* absorb then squeeze incrementally: the sponge state is kept */
            memset(s, 0, sizeof s);
            {
                uint8_t t[168];
                memset(t, 0, sizeof t);
                memcpy(t, ext, sizeof ext);
                t[sizeof ext] = 0x1F;
                t[167] |= 0x80;
                for (int w = 0; w < 21; w++) {
                    uint64_t v = 0;
                    for (int b = 0; b < 8; b++) v |= (uint64_t)t[8 * w + b] << (8 * b);
                    s[w] ^= v;
                }
            }
            int16_t *poly = a + (size_t)(i * k + j) * N;
            for (unsigned blk = 0; blk < maxnblocks; blk++) {
                oracle_keccakf1600(s);
                for (int w = 0; w < 21; w++)
                    for (int b = 0; b < 8; b++) buf[168 * blk + 8 * w + b] = (uint8_t)(s[w] >> (8 * b));
            }
            unsigned ctr = oracle_kyber_rej_uniform(poly, N, buf, maxnblocks * 168);
            while (ctr < N) {
                oracle_keccakf1600(s);
                for (int w = 0; w < 21; w++)
                    for (int b = 0; b < 8; b++) buf[8 * w + b] = (uint8_t)(s[w] >> (8 * b));
                ctr += oracle_kyber_rej_uniform(poly + ctr, N - ctr, buf, 168);
            }
        }
}

/* ---- indcpa (indcpa.cu:474-733) ----------------------------------------------- */
int oracle_kyber_indcpa_keypair(int k, uint8_t *pk, uint8_t *sk, const uint8_t seed[32]) {
    oracle_kyber_params p;
    if (oracle_kyber_params_get(k, &p)) return -1;
    uint8_t buf[64];
    int16_t *a = (int16_t *)malloc(sizeof(int16_t) * k * k * N);
    int16_t e[4 * 256], pkpv[4 * 256], skpv[4 * 256];
    if (!a) return -2;
    oracle_sha3_512(buf, seed, SYM); /* G(d) -> publicseed || noiseseed */
    const uint8_t *publicseed = buf, *noiseseed = buf + SYM;
    oracle_kyber_gen_matrix(k, a, publicseed, 0);
    uint8_t nonce = 0;
    for (int i = 0; i < k; i++) poly_getnoise(skpv + N * i, noiseseed, nonce++);
    for (int i = 0; i < k; i++) poly_getnoise(e + N * i, noiseseed, nonce++);
    for (int i = 0; i < k; i++) oracle_kyber_ntt(skpv + N * i);
    for (int i = 0; i < k; i++) oracle_kyber_ntt(e + N * i);
    for (int i = 0; i < k; i++) {
        pointwise_acc(k, pkpv + N * i, a + (size_t)N * k * i, skpv);
        poly_frommont(pkpv + N * i);
    }
    for (int i = 0; i < k; i++) poly_add(pkpv + N * i, pkpv + N * i, e + N * i);
    for (int i = 0; i < k; i++) poly_reduce(pkpv + N * i);
    for (int i = 0; i < k; i++) poly_tobytes(sk + 384 * i, skpv + N * i);
    for (int i = 0; i < k; i++) poly_tobytes(pk + 384 * i, pkpv + N * i);
    memcpy(pk + p.polyvec_bytes, publicseed, SYM);
    free(a);
    return 0;
}

int oracle_kyber_indcpa_enc(int k, uint8_t *c, const uint8_t m[32], const uint8_t *pk,
                            const uint8_t coins[32]) {
    oracle_kyber_params p;
    if (oracle_kyber_params_get(k, &p)) return -1;
    int16_t *at = (int16_t *)malloc(sizeof(int16_t) * k * k * N);
    int16_t sp[4 * 256], pkpv[4 * 256], ep[4 * 256], bp[4 * 256], v[256], kpoly[256], epp[256];
    uint8_t seed[SYM];
    if (!at) return -2;
    for (int i = 0; i < k; i++) poly_frombytes(pkpv + N * i, pk + 384 * i);
    memcpy(seed, pk + p.polyvec_bytes, SYM);
    poly_frommsg(kpoly, m);
    oracle_kyber_gen_matrix(k, at, seed, 1);
    uint8_t nonce = 0;
    for (int i = 0; i < k; i++) poly_getnoise(sp + N * i, coins, nonce++);
    for (int i = 0; i < k; i++) poly_getnoise(ep + N * i, coins, nonce++);
    poly_getnoise(epp, coins, nonce++);
    for (int i = 0; i < k; i++) oracle_kyber_ntt(sp + N * i);
    for (int i = 0; i < k; i++) pointwise_acc(k, bp + N * i, at + (size_t)N * k * i, sp);
    pointwise_acc(k, v, pkpv, sp);
    for (int i = 0; i < k; i++) oracle_kyber_invntt(bp + N * i);
    oracle_kyber_invntt(v);
    for (int i = 0; i < k; i++) poly_add(bp + N * i, bp + N * i, ep + N * i);
    poly_add(v, v, epp);
    poly_add(v, v, kpoly);
    for (int i = 0; i < k; i++) poly_reduce(bp + N * i);
    poly_reduce(v);
    for (int i = 0; i < k; i++) oracle_kyber_poly_compress(c + (size_t)32 * p.du * i, bp + N * i, p.du);
    oracle_kyber_poly_compress(c + p.polyvec_cbytes, v, p.dv);
    free(at);
    return 0;
}

int oracle_kyber_indcpa_dec(int k, uint8_t m[32], const uint8_t *c, const uint8_t *sk) {
    oracle_kyber_params p;
    if (oracle_kyber_params_get(k, &p)) return -1;
    int16_t bp[4 * 256], skpv[4 * 256], v[256], mp[256];
    for (int i = 0; i < k; i++) oracle_kyber_poly_decompress(bp + N * i, c + (size_t)32 * p.du * i, p.du);
    oracle_kyber_poly_decompress(v, c + p.polyvec_cbytes, p.dv);
    for (int i = 0; i < k; i++) poly_frombytes(skpv + N * i, sk + 384 * i);
    for (int i = 0; i < k; i++) oracle_kyber_ntt(bp + N * i);
    pointwise_acc(k, mp, skpv, bp);
    oracle_kyber_invntt(mp);
    poly_sub(mp, v, mp);
    poly_reduce(mp);
    poly_tomsg(m, mp);
    return 0;
}

/* ---- CCA KEM (round-2 Kyber FO transform; kem.cu is absent, Q4) ------------ */
static int verify_ct(const uint8_t *a, const uint8_t *b, size_t len) {
    /* same result as verify.cu:15-28, without the early exit */
    uint64_t r = 0;
    for (size_t i = 0; i < len; i++) r |= (uint64_t)(a[i] ^ b[i]);
    return (int)((-r) >> 63);
}

static void cmov(uint8_t *r, const uint8_t *x, size_t len, uint8_t b) { /* verify.cu:43-50 */
    b = (uint8_t)(-b);
    for (size_t i = 0; i < len; i++) r[i] ^= (uint8_t)(b & (x[i] ^ r[i]));
}

int oracle_kyber_keypair_derand(int k, uint8_t *pk, uint8_t *sk, const uint8_t coins64[64]) {
    oracle_kyber_params p;
    if (oracle_kyber_params_get(k, &p)) return -1;
    int rc = oracle_kyber_indcpa_keypair(k, pk, sk, coins64);
    if (rc) return rc;
    memcpy(sk + p.indcpa_sk, pk, p.pk);
    oracle_sha3_256(sk + p.sk - 2 * SYM, pk, p.pk);
    memcpy(sk + p.sk - SYM, coins64 + SYM, SYM);
    return 0;
}

int oracle_kyber_enc_derand(int k, uint8_t *ct, uint8_t ss[32], const uint8_t *pk,
                            const uint8_t coins32[32]) {
    oracle_kyber_params p;
    if (oracle_kyber_params_get(k, &p)) return -1;
    uint8_t buf[64], kr[64];
    oracle_sha3_256(buf, coins32, SYM);      /* m = H(random) */
    oracle_sha3_256(buf + SYM, pk, p.pk);    /* H(pk) */
    oracle_sha3_512(kr, buf, 2 * SYM);       /* (Kbar, r) = G(m || H(pk)) */
    int rc = oracle_kyber_indcpa_enc(k, ct, buf, pk, kr + SYM);
    if (rc) return rc;
    oracle_sha3_256(kr + SYM, ct, p.ct);     /* H(c) */
    oracle_shake256(ss, 32, kr, 2 * SYM);    /* KDF(Kbar || H(c)) */
    return 0;
}

int oracle_kyber_dec(int k, uint8_t ss[32], const uint8_t *ct, const uint8_t *sk) {
    oracle_kyber_params p;
    if (oracle_kyber_params_get(k, &p)) return -1;
    uint8_t buf[64], kr[64], cmp[1568];
    const uint8_t *pk = sk + p.indcpa_sk;
    int rc = oracle_kyber_indcpa_dec(k, buf, ct, sk);
    if (rc) return rc;
    memcpy(buf + SYM, sk + p.sk - 2 * SYM, SYM);
    oracle_sha3_512(kr, buf, 2 * SYM);
    rc = oracle_kyber_indcpa_enc(k, cmp, buf, pk, kr + SYM);
    if (rc) return rc;
    int fail = verify_ct(ct, cmp, p.ct);
    oracle_sha3_256(kr + SYM, ct, p.ct);
    cmov(kr, sk + p.sk - SYM, SYM, (uint8_t)fail);
    oracle_shake256(ss, 32, kr, 2 * SYM);
    return 0;
}

/* ---- threaded batch drivers: the CPU baseline ------------------------------ */
typedef struct {
    int op, k;
    size_t lo, hi;
    uint8_t *o0, *o1;
    const uint8_t *i0, *i1;
    int rc;
} job_t;

static void *run_job(void *arg) {
    job_t *j = (job_t *)arg;
    oracle_kyber_params p;
    oracle_kyber_params_get(j->k, &p);
    for (size_t e = j->lo; e < j->hi && !j->rc; e++) {
        if (j->op == 0)
            j->rc = oracle_kyber_keypair_derand(j->k, j->o0 + e * p.pk, j->o1 + e * p.sk, j->i0 + e * 64);
        else if (j->op == 1)
            j->rc = oracle_kyber_enc_derand(j->k, j->o0 + e * p.ct, j->o1 + e * 32, j->i0 + e * p.pk,
                                            j->i1 + e * 32);
        else
            j->rc = oracle_kyber_dec(j->k, j->o0 + e * 32, j->i0 + e * p.ct, j->i1 + e * p.sk);
    }
    return NULL;
}

static int run_batch(int op, int k, size_t n, uint8_t *o0, uint8_t *o1, const uint8_t *i0,
                     const uint8_t *i1, int nthreads) {
    oracle_kyber_params p;
    if (oracle_kyber_params_get(k, &p)) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    job_t jobs[256];
    uint8_t warm[32] = {0};
    init_zetas();
    oracle_sha3_256(warm, warm, 0); /* builds the Keccak tables before the threads start */
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (job_t){op, k, n * t / nthreads, n * (t + 1) / nthreads, o0, o1, i0, i1, 0};
        if (pthread_create(&th[t], NULL, run_job, &jobs[t])) { nthreads = t; break; }
    }
    int rc = 0;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].rc) rc = jobs[t].rc;
    }
    return rc;
}

int oracle_kyber_keypair_batch(int k, size_t n, uint8_t *pk, uint8_t *sk, const uint8_t *coins,
                               int nthreads) {
    return run_batch(0, k, n, pk, sk, coins, NULL, nthreads);
}

int oracle_kyber_enc_batch(int k, size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk,
                           const uint8_t *coins, int nthreads) {
    return run_batch(1, k, n, ct, ss, pk, coins, nthreads);
}

int oracle_kyber_dec_batch(int k, size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk,
                           int nthreads) {
    return run_batch(2, k, n, ss, NULL, ct, sk, nthreads);
}
