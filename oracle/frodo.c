/*
 * FrodoKEM-976-SHAKE restatement for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * KEM flow ........ FrodoKEM_GPU/kem.c:13-214 (keypair, enc, dec)
 * sampler ......... FrodoKEM_GPU/noise.c:10-29 with CDF_TABLE frodo976.cu:21
 * matrix A ........ generateMatA2 (frodo_macrify_reference.cu:164-188):
 *                   row i = SHAKE128(le16(i) || seedA, 2n) read as LE uint16
 * matrix products . frodo_macrify_reference.cu: A*S+E (matMulKernelTile
 *                   :55-99), S'*A+E' (matMulKernelSA :102-159), B'*S
 *                   (:237-258), S'*B+E'' (:261-281), add/sub (:284-309),
 *                   key_encode/decode (:311-361); all mod 2^16
 * packing ......... FrodoKEM_GPU/util.cu:18-113 (MSB-first)
 * parameters ...... FrodoKEM_GPU/params.h:17-26, api.h:11-14
 * "shake" is SHAKE256 (params.h:28).  kem.c calls CPU matrix functions that
 * are not defined anywhere (Q10); they are restated from the GPU kernels'
 * arithmetic.  The S'/E'/E'' stream is standard SHAKE256 as in kem.c (the
 * GPU's non-standard shake256_kernel_single, Q9, is not reproduced) and the
 * final key selection is a constant-time mask instead of kem.c:195's branch.
 */
#include <string.h>
#include <stdlib.h>
#include <pthread.h>
#include "oracle.h"

#define FN 976
#define FNB 8
#define FSS 24                         /* CRYPTO_BYTES */
#define FSEEDA 16
#define FMU 24                         /* 3 * 8 * 8 / 8 */
#define FPK (FSEEDA + 2 * FN * FNB)     /* 15632 */
#define FCT (2 * FN * FNB + 2 * FNB * FNB) /* 15744 */
#define FSK (FSS + FPK + 2 * FN * FNB + FSS) /* 31296 */

static const uint16_t CDF[11] = {5638, 15915, 23689, 28571, 31116, 32217, 32613, 32731, 32760, 32766, 32767};

/* noise.c:10-29 */
void oracle_frodo_sample_n(uint16_t *s, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        uint8_t sample = 0;
        uint16_t prnd = (uint16_t)(s[i] >> 1);
        uint8_t sign = (uint8_t)(s[i] & 0x1);
        for (unsigned j = 0; j < 10; j++) sample = (uint8_t)(sample + ((uint16_t)(CDF[j] - prnd) >> 15));
        s[i] = (uint16_t)(((uint16_t)(-sign) ^ sample) + sign);
    }
}

/* util.cu:18-64: MSB-first packing of `lsb` bits per element */
void oracle_frodo_pack(uint8_t *out, size_t outlen, const uint16_t *in, size_t inlen, unsigned lsb) {
    memset(out, 0, outlen);
    size_t bit = 0;
    for (size_t j = 0; j < inlen && bit < outlen * 8; j++)
        for (int b = (int)lsb - 1; b >= 0 && bit < outlen * 8; b--, bit++)
            if ((in[j] >> b) & 1) out[bit >> 3] |= (uint8_t)(0x80u >> (bit & 7));
}

/* util.cu:67-113 */
void oracle_frodo_unpack(uint16_t *out, size_t outlen, const uint8_t *in, size_t inlen, unsigned lsb) {
    memset(out, 0, outlen * sizeof(uint16_t));
    size_t bit = 0;
    for (size_t i = 0; i < outlen && bit < inlen * 8; i++)
        for (int b = (int)lsb - 1; b >= 0 && bit < inlen * 8; b--, bit++)
            if ((in[bit >> 3] >> (7 - (bit & 7))) & 1) out[i] |= (uint16_t)(1u << b);
}

/* generateMatA2: A[i][j] = LE16 of SHAKE128(le16(i) || seedA, 2n) */
void oracle_frodo_gen_a(uint16_t *A, const uint8_t seedA[16]) {
    uint8_t in[2 + FSEEDA];
    uint8_t row[2 * FN];
    memcpy(in + 2, seedA, FSEEDA);
    for (int i = 0; i < FN; i++) {
        in[0] = (uint8_t)i;
        in[1] = (uint8_t)(i >> 8);
        oracle_shake128(row, sizeof row, in, sizeof in);
        for (int j = 0; j < FN; j++) A[(size_t)i * FN + j] = (uint16_t)(row[2 * j] | (row[2 * j + 1] << 8));
    }
}

/* out (n x nbar) = A S + E with S stored transposed (nbar x n) */
static void mul_add_as_plus_e(uint16_t *out, const uint16_t *A, const uint16_t *St, const uint16_t *E) {
    for (int i = 0; i < FN; i++)
        for (int k = 0; k < FNB; k++) {
            uint16_t sum = E[i * FNB + k];
            for (int j = 0; j < FN; j++) sum = (uint16_t)(sum + A[(size_t)i * FN + j] * St[k * FN + j]);
            out[i * FNB + k] = sum;
        }
}

/* out (nbar x n) = S' A + E' */
void oracle_frodo_mul_add_sa_plus_e(uint16_t *out, const uint16_t *A, const uint16_t *S, const uint16_t *E) {
    for (int k = 0; k < FNB; k++)
        for (int i = 0; i < FN; i++) out[k * FN + i] = E[k * FN + i];
    for (int k = 0; k < FNB; k++)
        for (int j = 0; j < FN; j++) {
            const uint32_t s = S[k * FN + j];
            const uint16_t *arow = A + (size_t)j * FN;
            uint16_t *o = out + k * FN;
            for (int i = 0; i < FN; i++) o[i] = (uint16_t)(o[i] + s * arow[i]);
        }
}

/* W (nbar x nbar) = B' S, S stored transposed (frodo_mul_bs_kernel) */
static void mul_bs(uint16_t *out, const uint16_t *Bp, const uint16_t *St) {
    for (int i = 0; i < FNB; i++)
        for (int j = 0; j < FNB; j++) {
            uint32_t sum = 0;
            for (int k = 0; k < FN; k++) sum += (uint32_t)Bp[i * FN + k] * St[j * FN + k];
            out[i * FNB + j] = (uint16_t)sum;
        }
}

/* V (nbar x nbar) = S' B + E'' (frodo_mul_add_sb_plus_e_kernel) */
static void mul_add_sb_plus_e(uint16_t *out, const uint16_t *B, const uint16_t *S, const uint16_t *E) {
    for (int k = 0; k < FNB; k++)
        for (int i = 0; i < FNB; i++) {
            uint32_t sum = E[k * FNB + i];
            for (int j = 0; j < FN; j++) sum += (uint32_t)S[k * FN + j] * B[j * FNB + i];
            out[k * FNB + i] = (uint16_t)sum;
        }
}

/* frodo_key_encode_kernel: 3 bits per coefficient, shifted to the top */
static void key_encode(uint16_t *out, const uint8_t mu[FMU]) {
    for (int w = 0; w < 8; w++) {
        uint64_t t = 0;
        for (int j = 0; j < 3; j++) t |= (uint64_t)mu[w * 3 + j] << (8 * j);
        for (int j = 0; j < 8; j++) {
            out[w * 8 + j] = (uint16_t)((t & 7) << 13);
            t >>= 3;
        }
    }
}

/* frodo_key_decode_kernel */
static void key_decode(uint8_t mu[FMU], const uint16_t *in) {
    for (int w = 0; w < 8; w++) {
        uint64_t t = 0;
        for (int j = 0; j < 8; j++) {
            uint16_t v = (uint16_t)((in[w * 8 + j] + (1 << 12)) >> 13);
            t |= (uint64_t)(v & 7) << (3 * j);
        }
        for (int j = 0; j < 3; j++) mu[w * 3 + j] = (uint8_t)(t >> (8 * j));
    }
}

static void le_bytes_to_u16(uint16_t *out, const uint8_t *in, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = (uint16_t)(in[2 * i] | (in[2 * i + 1] << 8));
}

/* kem.c:13-61; randomness = s (24) || seedSE (24) || z (16) */
int oracle_frodo_keypair_derand(uint8_t *pk, uint8_t *sk, const uint8_t randomness[64]) {
    uint16_t *A = (uint16_t *)malloc(sizeof(uint16_t) * FN * FN);
    uint16_t B[FN * FNB], S[2 * FN * FNB];
    uint8_t buf[2 * 2 * FN * FNB], in[1 + FSS];
    if (!A) return -2;
    uint16_t *E = S + FN * FNB;
    oracle_shake256(pk, FSEEDA, randomness + 2 * FSS, FSEEDA);
    in[0] = 0x5F;
    memcpy(in + 1, randomness + FSS, FSS);
    oracle_shake256(buf, sizeof buf, in, sizeof in);
    le_bytes_to_u16(S, buf, 2 * FN * FNB);
    oracle_frodo_sample_n(S, FN * FNB);
    oracle_frodo_sample_n(E, FN * FNB);
    oracle_frodo_gen_a(A, pk);
    mul_add_as_plus_e(B, A, S, E);
    oracle_frodo_pack(pk + FSEEDA, FPK - FSEEDA, B, FN * FNB, 16);
    memcpy(sk, randomness, FSS);
    memcpy(sk + FSS, pk, FPK);
    for (int i = 0; i < FN * FNB; i++) {
        sk[FSS + FPK + 2 * i] = (uint8_t)S[i];
        sk[FSS + FPK + 2 * i + 1] = (uint8_t)(S[i] >> 8);
    }
    oracle_shake256(sk + FSS + FPK + 2 * FN * FNB, FSS, pk, FPK);
    free(A);
    return 0;
}

/* re-encryption core shared by enc and dec: Bp, C from (A, B, seedSE, mu) */
static void encrypt_core(uint16_t *Bp, uint16_t *C, const uint16_t *A, const uint16_t *B, const uint8_t seedSE[FSS],
                         const uint8_t mu[FMU]) {
    static const size_t NS = (2 * FN + FNB) * FNB;
    uint16_t *Sp = (uint16_t *)malloc(sizeof(uint16_t) * NS);
    uint8_t *buf = (uint8_t *)malloc(2 * NS);
    uint8_t in[1 + FSS];
    uint16_t V[FNB * FNB], Cm[FNB * FNB];
    in[0] = 0x96;
    memcpy(in + 1, seedSE, FSS);
    oracle_shake256(buf, 2 * NS, in, sizeof in);
    le_bytes_to_u16(Sp, buf, NS);
    uint16_t *Ep = Sp + FN * FNB, *Epp = Sp + 2 * FN * FNB;
    oracle_frodo_sample_n(Sp, FN * FNB);
    oracle_frodo_sample_n(Ep, FN * FNB);
    oracle_frodo_mul_add_sa_plus_e(Bp, A, Sp, Ep);
    oracle_frodo_sample_n(Epp, FNB * FNB);
    mul_add_sb_plus_e(V, B, Sp, Epp);
    key_encode(Cm, mu);
    for (int i = 0; i < FNB * FNB; i++) C[i] = (uint16_t)(V[i] + Cm[i]);
    free(Sp);
    free(buf);
}

/* kem.c:64-127 with an explicit mu; A may be precomputed (NULL: derived from pk) */
int oracle_frodo_enc_derand_a(uint8_t *ct, uint8_t ss[FSS], const uint8_t *pk, const uint8_t mu[FMU],
                              const uint16_t *Apre) {
    uint16_t *A = (uint16_t *)Apre;
    if (!A) {
        A = (uint16_t *)malloc(sizeof(uint16_t) * FN * FN);
        if (!A) return -2;
        oracle_frodo_gen_a(A, pk);
    }
    uint16_t B[FN * FNB], Bp[FN * FNB], C[FNB * FNB];
    uint8_t g2in[FSS + FMU], g2out[2 * FSS];
    uint8_t *fin = (uint8_t *)malloc(FCT + FSS);
    oracle_shake256(g2in, FSS, pk, FPK);
    memcpy(g2in + FSS, mu, FMU);
    oracle_shake256(g2out, sizeof g2out, g2in, sizeof g2in);
    oracle_frodo_unpack(B, FN * FNB, pk + FSEEDA, FPK - FSEEDA, 16);
    encrypt_core(Bp, C, A, B, g2out, mu);
    oracle_frodo_pack(ct, 2 * FN * FNB, Bp, FN * FNB, 16);
    oracle_frodo_pack(ct + 2 * FN * FNB, 2 * FNB * FNB, C, FNB * FNB, 16);
    memcpy(fin, ct, FCT);
    memcpy(fin + FCT, g2out + FSS, FSS);
    oracle_shake256(ss, FSS, fin, FCT + FSS);
    free(fin);
    if (!Apre) free(A);
    return 0;
}

int oracle_frodo_enc_derand(uint8_t *ct, uint8_t ss[FSS], const uint8_t *pk, const uint8_t mu[FMU]) {
    return oracle_frodo_enc_derand_a(ct, ss, pk, mu, NULL);
}

/* kem.c:130-214 with a branch-free selection of k' or s */
int oracle_frodo_dec_a(uint8_t ss[FSS], const uint8_t *ct, const uint8_t *sk, const uint16_t *Apre) {
    const uint8_t *pk = sk + FSS;
    uint16_t *A = (uint16_t *)Apre;
    if (!A) {
        A = (uint16_t *)malloc(sizeof(uint16_t) * FN * FN);
        if (!A) return -2;
        oracle_frodo_gen_a(A, pk);
    }
    uint16_t B[FN * FNB], Bp[FN * FNB], BBp[FN * FNB], C[FNB * FNB], CC[FNB * FNB], W[FNB * FNB], St[FN * FNB];
    uint8_t g2in[FSS + FMU], g2out[2 * FSS];
    uint8_t *fin = (uint8_t *)malloc(FCT + FSS);
    oracle_frodo_unpack(Bp, FN * FNB, ct, 2 * FN * FNB, 16);
    oracle_frodo_unpack(C, FNB * FNB, ct + 2 * FN * FNB, 2 * FNB * FNB, 16);
    le_bytes_to_u16(St, sk + FSS + FPK, FN * FNB);
    mul_bs(W, Bp, St);
    for (int i = 0; i < FNB * FNB; i++) W[i] = (uint16_t)(C[i] - W[i]);
    memcpy(g2in, sk + FSS + FPK + 2 * FN * FNB, FSS);
    key_decode(g2in + FSS, W);
    oracle_shake256(g2out, sizeof g2out, g2in, sizeof g2in);
    oracle_frodo_unpack(B, FN * FNB, pk + FSEEDA, FPK - FSEEDA, 16);
    encrypt_core(BBp, CC, A, B, g2out, g2in + FSS);
    uint16_t diff = 0;
    for (int i = 0; i < FN * FNB; i++) diff |= (uint16_t)(Bp[i] ^ BBp[i]);
    for (int i = 0; i < FNB * FNB; i++) diff |= (uint16_t)(C[i] ^ CC[i]);
    const uint8_t mask = (uint8_t)(0 - (uint8_t)(((uint32_t)diff + 0xFFFFu) >> 16)); /* 0xFF iff diff != 0 */
    memcpy(fin, ct, FCT);
    for (int i = 0; i < FSS; i++)
        fin[FCT + i] = (uint8_t)((g2out[FSS + i] & ~mask) | (sk[i] & mask));
    oracle_shake256(ss, FSS, fin, FCT + FSS);
    free(fin);
    if (!Apre) free(A);
    return 0;
}

int oracle_frodo_dec(uint8_t ss[FSS], const uint8_t *ct, const uint8_t *sk) { return oracle_frodo_dec_a(ss, ct, sk, NULL); }

/* threaded drivers sharing one A (one key pair): the CPU baseline of the Frodo-976 shared-A decaps workload */
typedef struct {
    int op;
    size_t lo, hi;
    uint8_t *o0, *o1;
    const uint8_t *i0, *i1;
    const uint16_t *A;
} fr_job;

static void *fr_run(void *arg) {
    fr_job *j = (fr_job *)arg;
    for (size_t e = j->lo; e < j->hi; e++) {
        if (j->op == 0)
            oracle_frodo_enc_derand_a(j->o0 + e * FCT, j->o1 + e * FSS, j->i0, j->i1 + e * FMU, j->A);
        else
            oracle_frodo_dec_a(j->o0 + e * FSS, j->i0 + e * FCT, j->i1, j->A);
    }
    return NULL;
}

static int fr_batch(int op, size_t n, uint8_t *o0, uint8_t *o1, const uint8_t *i0, const uint8_t *i1,
                    const uint8_t *seedA, int nthreads) {
    uint16_t *A = (uint16_t *)malloc(sizeof(uint16_t) * FN * FN);
    if (!A) return -2;
    oracle_frodo_gen_a(A, seedA);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    fr_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (fr_job){op, n * t / nthreads, n * (t + 1) / nthreads, o0, o1, i0, i1, A};
        if (pthread_create(&th[t], NULL, fr_run, &jobs[t])) { nthreads = t; break; }
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(A);
    return 0;
}

/* n encapsulations to one public key (mu: 24 bytes per element) */
int oracle_frodo_enc_batch(size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu, int nthreads) {
    return fr_batch(0, n, ct, ss, pk, mu, pk, nthreads);
}

/* n decapsulations under one secret key */
int oracle_frodo_dec_batch(size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk, int nthreads) {
    return fr_batch(1, n, ss, NULL, ct, sk, sk + FSS, nthreads);
}
