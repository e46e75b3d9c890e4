/*
 * NewHope-1024 (and, through newhope512.c, NewHope-512) CPA-PKE and CCA-KEM
 * restatement for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Per-element C following the batched CUDA sources of NewHope_NIST_GPU:
 *   params ........... params.h:9-39 (n = 1024, q = 12289, k = 8)
 *   reduce ........... reduce.cu:22-47 (Montgomery R = 2^18, qinv 12287;
 *                      Barrett a - ((5a) >> 16) q)
 *   ntt .............. ntt.cu:92-128 (bitrev_vector, mul_coefficients),
 *                      ntt.cu:195-234 (lazy even levels, Barrett odd levels)
 *   tables ........... precomp.cu:5-72 (Pari/GP formulas for
 *                      omegas_inv_bitrev_montgomery, gammas_bitrev_montgomery,
 *                      gammas_inv_montgomery), regenerated here
 *   poly ............. poly.cu:21-547 (freeze, flipabs, 14-bit packing,
 *                      3-bit compression, frommsg/tomsg, uniform, sample,
 *                      pointwise, add, sub, ntt, invntt)
 *   cpapke ........... cpapke.cu:140-476 (encode/decode, gen_a, keypair,
 *                      enc, dec)
 * uint16_t/uint32_t arithmetic is kept exactly as the reference stores it.
 * The keypair seed expansion is SHAKE256(seed, 64) -> publicseed ||
 * noiseseed (cpapke.cu:283-298) without the reference's stride mismatch (Q6).
 * The CCA-KEM at the end is the NewHope specification's transform over this
 * CPA-PKE (the reference has its sizes, params.h:37-39, but no code).
 */
#include <string.h>
#include <stdlib.h>
#include <pthread.h>
#include "oracle.h"

/* N = 1024 by default; newhope512.c compiles this file again with N = 512
 * (precomp.cu:75-190 tables, g = 10968; ntt.cu:130-180 NTT with 9 levels and
 * % q on the odd ones; poly.cu:244,279 message coding over 2 coefficients).
 * NH_FN names the exported functions of each build. */
#ifndef NH_N
#define NH_N 1024
#endif
#if NH_N == 1024
#define NH_FN(x) oracle_newhope_##x
#define NH_G 7
#define NH_LOG 10
#elif NH_N == 512
#define NH_FN(x) oracle_newhope512_##x
#define NH_G 10968
#define NH_LOG 9
#else
#error "NH_N must be 512 or 1024"
#endif
#define NH_Q 12289
#define NH_SYM 32
#define NH_POLYBYTES ((14 * NH_N) / 8)
#define NH_CBYTES ((3 * NH_N) / 8)

static uint16_t g_omegas_inv[NH_N / 2], g_gammas[NH_N], g_gammas_inv[NH_N], g_brv[NH_N];
static int g_nh_init = 0;

static uint32_t powmod(uint32_t b, uint32_t e) {
    uint64_t r = 1, x = b % NH_Q;
    while (e) {
        if (e & 1) r = r * x % NH_Q;
        x = x * x % NH_Q;
        e >>= 1;
    }
    return (uint32_t)r;
}

static uint32_t invmod(uint32_t a) { return powmod(a, NH_Q - 2); }

static void nh_init(void) {
    if (g_nh_init) return;
    const uint32_t g = NH_G, mont = (1u << 18) % NH_Q;
    for (unsigned i = 0; i < NH_N; i++) {
        unsigned r = 0;
        for (int b = 0; b < NH_LOG; b++) r |= ((i >> b) & 1u) << (NH_LOG - 1 - b);
        g_brv[i] = (uint16_t)r;
    }
    const uint32_t g2inv = invmod(g * g % NH_Q), ginv = invmod(g), ninv = invmod(NH_N);
    for (unsigned k = 0; k < NH_N / 2; k++)
        g_omegas_inv[k] = (uint16_t)((uint64_t)powmod(g2inv, g_brv[2 * k]) * mont % NH_Q);
    for (unsigned k = 0; k < NH_N; k++) {
        g_gammas[k] = (uint16_t)((uint64_t)powmod(g, g_brv[k]) * mont % NH_Q);
        g_gammas_inv[k] = (uint16_t)((uint64_t)powmod(ginv, k) * ninv % NH_Q * mont % NH_Q);
    }
    g_nh_init = 1;
}

void NH_FN(tables)(uint16_t omegas_inv[NH_N / 2], uint16_t gammas[NH_N], uint16_t gammas_inv[NH_N]) {
    nh_init();
    memcpy(omegas_inv, g_omegas_inv, sizeof g_omegas_inv);
    memcpy(gammas, g_gammas, sizeof g_gammas);
    memcpy(gammas_inv, g_gammas_inv, sizeof g_gammas_inv);
}

/* reduce.cu:22-31 */
uint16_t NH_FN(montgomery_reduce)(uint32_t a) {
    uint32_t u = a * 12287u;
    u &= (1u << 18) - 1;
    u *= NH_Q;
    a = a + u;
    return (uint16_t)(a >> 18);
}

/* reduce.cu:33-47 */
uint16_t NH_FN(barrett_reduce)(uint16_t a) {
    uint32_t u = ((uint32_t)a * 5) >> 16;
    u *= NH_Q;
    a = (uint16_t)(a - u);
    return a;
}

#define mont NH_FN(montgomery_reduce)
#define barr NH_FN(barrett_reduce)

/* ntt.cu:92-109 */
static void bitrev_vector(uint16_t *p) {
    for (unsigned i = 0; i < NH_N; i++) {
        unsigned r = g_brv[i];
        if (i < r) {
            uint16_t t = p[i];
            p[i] = p[r];
            p[r] = t;
        }
    }
}

/* ntt.cu:120-128 */
static void mul_coefficients(uint16_t *p, const uint16_t *f) {
    for (unsigned i = 0; i < NH_N; i++) p[i] = mont((uint32_t)p[i] * f[i]);
}

/* ntt.cu:195-234 (N = 1024): levels in pairs, even lazy, odd Barrett;
 * ntt.cu:130-180 (N = 512): 9 levels, even lazy, odd reduced with % q */
static void ntt_core(uint16_t *a, const uint16_t *omega) {
    for (int i = 0; i < NH_LOG; i += 2) {
        int distance = 1 << i;
        for (int start = 0; start < distance; start++) {
            int jt = 0;
            for (int j = start; j < NH_N - 1; j += 2 * distance) {
                uint16_t W = omega[jt++];
                uint16_t temp = a[j];
                a[j] = (uint16_t)(temp + a[j + distance]);
                a[j + distance] = mont((uint32_t)W * ((uint32_t)temp + 3 * NH_Q - a[j + distance]));
            }
        }
        if (i + 1 >= NH_LOG) break;
        distance <<= 1;
        for (int start = 0; start < distance; start++) {
            int jt = 0;
            for (int j = start; j < NH_N - 1; j += 2 * distance) {
                uint16_t W = omega[jt++];
                uint16_t temp = a[j];
#if NH_N == 1024
                a[j] = barr((uint16_t)(temp + a[j + distance]));
#else
                a[j] = (uint16_t)((temp + a[j + distance]) % NH_Q);
#endif
                a[j + distance] = mont((uint32_t)W * ((uint32_t)temp + 3 * NH_Q - a[j + distance]));
            }
        }
    }
}

/* poly.cu:519-527 */
void NH_FN(ntt)(uint16_t *r) {
    nh_init();
    mul_coefficients(r, g_gammas);
    ntt_core(r, g_gammas);
}

/* poly.cu:538-547 */
void NH_FN(invntt)(uint16_t *r) {
    nh_init();
    bitrev_vector(r);
    ntt_core(r, g_omegas_inv);
    mul_coefficients(r, g_gammas_inv);
}

/* poly.cu:21-33 */
static uint16_t coeff_freeze(uint16_t x) {
    uint16_t m, r;
    int16_t c;
    r = (uint16_t)(x % NH_Q);
    m = (uint16_t)(r - NH_Q);
    c = (int16_t)m;
    c >>= 15;
    r = (uint16_t)(m ^ ((r ^ m) & c));
    return r;
}

/* poly.cu:44-52 */
static uint16_t flipabs(uint16_t x) {
    int16_t r, m;
    r = (int16_t)coeff_freeze(x);
    r = (int16_t)(r - NH_Q / 2);
    m = (int16_t)(r >> 15);
    return (uint16_t)((r + m) ^ m);
}

/* poly.cu:62-74 */
static void poly_frombytes(uint16_t *r, const uint8_t *a) {
    for (int i = 0; i < NH_N / 4; i++) {
        r[4 * i + 0] = (uint16_t)(a[7 * i + 0] | (((uint16_t)a[7 * i + 1] & 0x3f) << 8));
        r[4 * i + 1] = (uint16_t)((a[7 * i + 1] >> 6) | (((uint16_t)a[7 * i + 2]) << 2) | (((uint16_t)a[7 * i + 3] & 0x0f) << 10));
        r[4 * i + 2] = (uint16_t)((a[7 * i + 3] >> 4) | (((uint16_t)a[7 * i + 4]) << 4) | (((uint16_t)a[7 * i + 5] & 0x03) << 12));
        r[4 * i + 3] = (uint16_t)((a[7 * i + 5] >> 2) | (((uint16_t)a[7 * i + 6]) << 6));
    }
}

/* poly.cu:105-127 */
static void poly_tobytes(uint8_t *r, const uint16_t *p) {
    for (int i = 0; i < NH_N / 4; i++) {
        uint16_t t0 = coeff_freeze(p[4 * i + 0]), t1 = coeff_freeze(p[4 * i + 1]);
        uint16_t t2 = coeff_freeze(p[4 * i + 2]), t3 = coeff_freeze(p[4 * i + 3]);
        r[7 * i + 0] = (uint8_t)(t0 & 0xff);
        r[7 * i + 1] = (uint8_t)((t0 >> 8) | (t1 << 6));
        r[7 * i + 2] = (uint8_t)(t1 >> 2);
        r[7 * i + 3] = (uint8_t)((t1 >> 10) | (t2 << 4));
        r[7 * i + 4] = (uint8_t)(t2 >> 4);
        r[7 * i + 5] = (uint8_t)((t2 >> 12) | (t3 << 2));
        r[7 * i + 6] = (uint8_t)(t3 >> 6);
    }
}

/* poly.cu:165-186 */
static void poly_compress(uint8_t *r, const uint16_t *p) {
    unsigned k = 0;
    uint32_t t[8];
    for (unsigned i = 0; i < NH_N; i += 8) {
        for (unsigned j = 0; j < 8; j++) {
            t[j] = coeff_freeze(p[i + j]);
            t[j] = (((t[j] << 3) + NH_Q / 2) / NH_Q) & 0x7;
        }
        r[k] = (uint8_t)(t[0] | (t[1] << 3) | (t[2] << 6));
        r[k + 1] = (uint8_t)((t[2] >> 2) | (t[3] << 1) | (t[4] << 4) | (t[5] << 7));
        r[k + 2] = (uint8_t)((t[5] >> 1) | (t[6] << 2) | (t[7] << 5));
        k += 3;
    }
}

/* poly.cu:197-216 */
static void poly_decompress(uint16_t *r, const uint8_t *a) {
    for (unsigned i = 0; i < NH_N; i += 8) {
        r[i + 0] = a[0] & 7;
        r[i + 1] = (a[0] >> 3) & 7;
        r[i + 2] = (uint16_t)((a[0] >> 6) | ((a[1] << 2) & 4));
        r[i + 3] = (a[1] >> 1) & 7;
        r[i + 4] = (a[1] >> 4) & 7;
        r[i + 5] = (uint16_t)((a[1] >> 7) | ((a[2] << 1) & 6));
        r[i + 6] = (a[2] >> 2) & 7;
        r[i + 7] = (a[2] >> 5);
        a += 3;
        for (unsigned j = 0; j < 8; j++) r[i + j] = (uint16_t)(((uint32_t)r[i + j] * NH_Q + 4) >> 3);
    }
}

/* poly.cu:226-251 */
static void poly_frommsg(uint16_t *r, const uint8_t *msg) {
    for (unsigned i = 0; i < 32; i++)
        for (unsigned j = 0; j < 8; j++) {
            unsigned mask = -(unsigned)((msg[i] >> j) & 1);
            for (unsigned q = 0; q < NH_N / 256; q++) r[8 * i + j + 256 * q] = (uint16_t)(mask & (NH_Q / 2));
        }
}

/* poly.cu:261-291 */
static void poly_tomsg(uint8_t *msg, const uint16_t *x) {
    memset(msg, 0, 32);
    for (unsigned i = 0; i < 256; i++) {
        uint16_t t = flipabs(x[i]);
        t = (uint16_t)(t + flipabs(x[i + 256]));
#if NH_N == 1024
        t = (uint16_t)(t + flipabs(x[i + 512]));
        t = (uint16_t)(t + flipabs(x[i + 768]));
        t = (uint16_t)(t - NH_Q);
#else
        t = (uint16_t)(t - NH_Q / 2);
#endif
        t >>= 15;
        msg[i >> 3] |= (uint8_t)(t << (i & 7));
    }
}

/* poly.cu:332-366: 16 SHAKE128(seed || i) streams, 16-bit candidates < 5q kept
 * unreduced, 64 coefficients each */
void NH_FN(poly_uniform)(uint16_t *a, const uint8_t seed[32]) {
    uint8_t ext[NH_SYM + 1], buf[168 * 8];
    memcpy(ext, seed, NH_SYM);
    for (unsigned i = 0; i < NH_N / 64; i++) {
        ext[NH_SYM] = (uint8_t)i;
        /* squeeze enough blocks; the stream is consumed block by block */
        unsigned ctr = 0, nblk = 1;
        for (;;) {
            oracle_shake128(buf, 168 * nblk, ext, sizeof ext);
            ctr = 0;
            for (unsigned j = 0; j + 1 < 168 * nblk && ctr < 64; j += 2) {
                uint16_t val = (uint16_t)(buf[j] | ((uint16_t)buf[j + 1] << 8));
                if (val < 5 * NH_Q) a[i * 64 + ctr++] = val;
            }
            if (ctr == 64 || nblk == 8) break;
            nblk++;
        }
    }
}

/* poly.cu:376-444: hw(a) + q - hw(b) over SHAKE256(seed || nonce || i, 128) */
void NH_FN(poly_sample)(uint16_t *r, const uint8_t seed[32], uint8_t nonce) {
    uint8_t ext[NH_SYM + 2], buf[128];
    memcpy(ext, seed, NH_SYM);
    ext[NH_SYM] = nonce;
    for (unsigned i = 0; i < NH_N / 64; i++) {
        ext[NH_SYM + 1] = (uint8_t)i;
        oracle_shake256(buf, 128, ext, sizeof ext);
        for (unsigned j = 0; j < 64; j++) {
            unsigned a = buf[2 * j], b = buf[2 * j + 1], ha = 0, hb = 0;
            for (int k = 0; k < 8; k++) {
                ha += (a >> k) & 1;
                hb += (b >> k) & 1;
            }
            r[64 * i + j] = (uint16_t)(ha + NH_Q - hb);
        }
    }
}

/* poly.cu:455-468 */
static void poly_mul_pointwise(uint16_t *r, const uint16_t *a, const uint16_t *b) {
    for (int i = 0; i < NH_N; i++) {
        uint16_t t = mont(3186u * b[i]);
        r[i] = mont((uint32_t)a[i] * t);
    }
}

static void poly_add(uint16_t *r, const uint16_t *a, const uint16_t *b) {
    for (int i = 0; i < NH_N; i++) r[i] = (uint16_t)((a[i] + b[i]) % NH_Q);
}

static void poly_sub(uint16_t *r, const uint16_t *a, const uint16_t *b) {
    for (int i = 0; i < NH_N; i++) r[i] = (uint16_t)((a[i] + 3 * NH_Q - b[i]) % NH_Q);
}

/* cpapke.cu:265-338 */
int NH_FN(cpa_keypair)(uint8_t *pk, uint8_t *sk, const uint8_t seed[32]) {
    uint16_t ahat[NH_N], ehat[NH_N], ahat_shat[NH_N], bhat[NH_N], shat[NH_N];
    uint8_t z[2 * NH_SYM];
    nh_init();
    oracle_shake256(z, 2 * NH_SYM, seed, NH_SYM);
    const uint8_t *publicseed = z, *noiseseed = z + NH_SYM;
    NH_FN(poly_uniform)(ahat, publicseed);
    NH_FN(poly_sample)(shat, noiseseed, 0);
    NH_FN(ntt)(shat);
    NH_FN(poly_sample)(ehat, noiseseed, 1);
    NH_FN(ntt)(ehat);
    poly_mul_pointwise(ahat_shat, shat, ahat);
    poly_add(bhat, ehat, ahat_shat);
    poly_tobytes(sk, shat);
    poly_tobytes(pk, bhat);
    memcpy(pk + NH_POLYBYTES, publicseed, NH_SYM);
    return 0;
}

/* cpapke.cu:353-429 */
int NH_FN(cpa_enc)(uint8_t *c, const uint8_t m[32], const uint8_t *pk, const uint8_t coin[32]) {
    uint16_t sprime[NH_N], eprime[NH_N], vprime[NH_N], ahat[NH_N], bhat[NH_N], epp[NH_N], uhat[NH_N], v[NH_N];
    uint8_t publicseed[NH_SYM];
    nh_init();
    poly_frommsg(v, m);
    poly_frombytes(bhat, pk);
    memcpy(publicseed, pk + NH_POLYBYTES, NH_SYM);
    NH_FN(poly_uniform)(ahat, publicseed);
    NH_FN(poly_sample)(sprime, coin, 0);
    NH_FN(poly_sample)(eprime, coin, 1);
    NH_FN(poly_sample)(epp, coin, 2);
    NH_FN(ntt)(sprime);
    NH_FN(ntt)(eprime);
    poly_mul_pointwise(uhat, ahat, sprime);
    poly_add(uhat, uhat, eprime);
    poly_mul_pointwise(vprime, bhat, sprime);
    NH_FN(invntt)(vprime);
    poly_add(vprime, vprime, epp);
    poly_add(vprime, vprime, v);
    poly_tobytes(c, uhat);
    poly_compress(c + NH_POLYBYTES, vprime);
    return 0;
}

/* cpapke.cu:443-476 */
int NH_FN(cpa_dec)(uint8_t m[32], const uint8_t *c, const uint8_t *sk) {
    uint16_t vprime[NH_N], uhat[NH_N], tmp[NH_N], shat[NH_N];
    nh_init();
    poly_frombytes(shat, sk);
    poly_frombytes(uhat, c);
    poly_decompress(vprime, c + NH_POLYBYTES);
    poly_mul_pointwise(tmp, shat, uhat);
    NH_FN(invntt)(tmp);
    poly_sub(tmp, tmp, vprime);
    poly_tomsg(m, tmp);
    return 0;
}

/* NewHope CCA-KEM: the Fujisaki-Okamoto transform of the NewHope
 * specification (CCA-KEM Gen / Encaps / Decaps, SHAKE256 for every hash) over the
 * CPA-PKE above.  The reference fixes its sizes (NewHope_NIST_GPU/params.h:37-39:
 * sk = sk_cpa || pk || H(pk) || s, ct = ct_cpa || d) and its verify/cmov
 * (verify.cu), but has no KEM code.  Keypair coins: CPA seed || s. */
#define NH_PK (NH_POLYBYTES + NH_SYM)
#define NH_CT (NH_POLYBYTES + NH_CBYTES)
#define NH_CCA_SK (NH_POLYBYTES + NH_PK + 2 * NH_SYM)
#define NH_CCA_CT (NH_CT + NH_SYM)

int NH_FN(cca_keypair)(uint8_t *pk, uint8_t *sk, const uint8_t coins[64]) {
    NH_FN(cpa_keypair)(pk, sk, coins);
    memcpy(sk + NH_POLYBYTES, pk, NH_PK);
    oracle_shake256(sk + NH_POLYBYTES + NH_PK, NH_SYM, pk, NH_PK);
    memcpy(sk + NH_CCA_SK - NH_SYM, coins + NH_SYM, NH_SYM);
    return 0;
}

int NH_FN(cca_enc)(uint8_t *ct, uint8_t ss[32], const uint8_t *pk, const uint8_t coins[32]) {
    uint8_t buf[2 * NH_SYM], kcd[3 * NH_SYM];
    oracle_shake256(buf, NH_SYM, coins, NH_SYM);        /* mu = H(coins): raw coins are not released */
    oracle_shake256(buf + NH_SYM, NH_SYM, pk, NH_PK);   /* H(pk) */
    oracle_shake256(kcd, 3 * NH_SYM, buf, 2 * NH_SYM);  /* K || coin' || d */
    NH_FN(cpa_enc)(ct, buf, pk, kcd + NH_SYM);
    memcpy(ct + NH_CT, kcd + 2 * NH_SYM, NH_SYM);
    oracle_shake256(kcd + NH_SYM, NH_SYM, ct, NH_CCA_CT); /* H(c || d) */
    oracle_shake256(ss, NH_SYM, kcd, 2 * NH_SYM);
    return 0;
}

int NH_FN(cca_dec)(uint8_t ss[32], const uint8_t *ct, const uint8_t *sk) {
    uint8_t buf[2 * NH_SYM], kcd[3 * NH_SYM], ct2[NH_CCA_CT];
    NH_FN(cpa_dec)(buf, ct, sk);
    memcpy(buf + NH_SYM, sk + NH_POLYBYTES + NH_PK, NH_SYM);
    oracle_shake256(kcd, 3 * NH_SYM, buf, 2 * NH_SYM);
    NH_FN(cpa_enc)(ct2, buf, sk + NH_POLYBYTES, kcd + NH_SYM);
    memcpy(ct2 + NH_CT, kcd + 2 * NH_SYM, NH_SYM);
    const int fail = memcmp(ct, ct2, NH_CCA_CT) != 0;
    oracle_shake256(kcd + NH_SYM, NH_SYM, ct, NH_CCA_CT);
    if (fail) memcpy(kcd, sk + NH_CCA_SK - NH_SYM, NH_SYM); /* implicit rejection: s */
    oracle_shake256(ss, NH_SYM, kcd, 2 * NH_SYM);
    return 0;
}

/* threaded triple driver (the CPU baseline of the NewHope-1024 CPA triple workload) */
typedef struct {
    size_t lo, hi;
    uint8_t *pk, *sk, *ct, *m2;
    const uint8_t *seed, *msg, *coin;
} nh_job;

static void *nh_run(void *arg) {
    nh_job *j = (nh_job *)arg;
    for (size_t e = j->lo; e < j->hi; e++) {
        NH_FN(cpa_keypair)(j->pk + e * NH_PK, j->sk + e * NH_POLYBYTES, j->seed + e * 32);
        NH_FN(cpa_enc)(j->ct + e * NH_CT, j->msg + e * 32, j->pk + e * NH_PK, j->coin + e * 32);
        NH_FN(cpa_dec)(j->m2 + e * 32, j->ct + e * NH_CT, j->sk + e * NH_POLYBYTES);
    }
    return NULL;
}

int NH_FN(triple_batch)(size_t n, uint8_t *pk, uint8_t *sk, uint8_t *ct, uint8_t *m2,
                                const uint8_t *seed, const uint8_t *msg, const uint8_t *coin, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    nh_job jobs[256];
    uint8_t warm[32] = {0};
    nh_init();
    oracle_sha3_256(warm, warm, 0);
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (nh_job){n * t / nthreads, n * (t + 1) / nthreads, pk, sk, ct, m2, seed, msg, coin};
        if (pthread_create(&th[t], NULL, nh_run, &jobs[t])) { nthreads = t; break; }
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    return 0;
}
