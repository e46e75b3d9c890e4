/*
 * pqc_api.h - C ABI of the MI355X batched lattice-KEM engine (libpqc_amd.so).
 *
 * Plain pointers and sizes only.  Every entry point returns PQC_OK (0) or a
 * negative PQC_ERR_* code; nothing calls exit() or prints.
 *
 * Reference interfaces replaced (paths relative to the nainag/PQC tree):
 *   crypto_kem_keypair/enc/dec ... Kyber_GPU_Batched/api.h:18-22,
 *                                  FrodoKEM_GPU/api.h:20-24
 *   indcpa_keypair/enc/dec ....... Kyber_GPU_Batched/indcpa.h:16-26 (batched,
 *                                  COUNT + device pointers + stream)
 *   cpapke_keypair/enc/dec ....... NewHope_NIST_GPU/cpapke.h:14-24
 *   sha3_256_n / sha3_512_n ...... Kyber_GPU_Batched/fips202.h:19-38
 *   poly_ntt_n / poly_invntt_n ... Kyber_GPU_Batched/poly.h (poly.cu:159-190)
 *   poly_ntt / poly_invntt,
 *   poly_sample (NewHope) ........ NewHope_NIST_GPU/poly.cu:396-444,519-547
 *   crypto_kem_*_gpu (Frodo) ..... FrodoKEM_GPU/api.h:20-24, frodo976.cu:87-690
 *   poly_set4 scratch ............ Kyber_GPU_Batched/main.h:14-39 (now owned
 *                                  by a pqc_ctx)
 */
#ifndef PQC_API_H
#define PQC_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define PQC_OK 0
#define PQC_ERR_INVALID_ARG (-1)     /* null pointer, misaligned buffer, bad op */
#define PQC_ERR_UNSUPPORTED (-2)     /* algorithm not enabled in this context */
#define PQC_ERR_BATCH_TOO_LARGE (-3) /* n above the context's max_batch */
#define PQC_ERR_NO_MEMORY (-4)       /* HBM budget exceeded or allocation failed */
#define PQC_ERR_DEVICE (-5)          /* HIP runtime error */
#define PQC_ERR_NO_DEVICE (-6)       /* no gfx950 device visible */

/* ---- algorithms ------------------------------------------------------------- */
#define PQC_ALG_KYBER512 1        /* round 2, k=2, du=10, dv=3 */
#define PQC_ALG_KYBER768 2        /* round 2, k=3, du=10, dv=4 */
#define PQC_ALG_KYBER1024 3       /* round 2, k=4, du=11, dv=5 */
#define PQC_ALG_NEWHOPE1024_CPA 4 /* NewHope-1024 CPA-PKE */
#define PQC_ALG_FRODO976 5        /* FrodoKEM-976-SHAKE */
#define PQC_ALG_NEWHOPE1024 6     /* NewHope-1024 CCA-KEM: pk 1824, sk 3680, ct 2208, ss 32
                                     (NewHope_NIST_GPU/params.h:37-39) */
#define PQC_ALG_NEWHOPE512 7      /* NewHope-512 CCA-KEM: pk 928, sk 1888, ct 1120, ss 32 */
#define PQC_ALG_NEWHOPE512_CPA 8  /* NewHope-512 CPA-PKE (the NEWHOPE_N == 512 branches) */
#define PQC_ALG_MASK(a) (1u << (a))

/* ---- call flags ------------------------------------------------------------- */
#define PQC_F_HOST_PTRS 0x1u /* buffers are host memory: staged through the
                                context, and the call returns when done.
                                Pageable memory moves in 16 MiB chunks through
                                two pinned buffers (host copy of one chunk
                                overlaps the DMA of the other); memory from
                                pqc_host_alloc is copied directly. */
#define PQC_F_PLAIN 0x2u     /* run the plain per-op kernels instead of the
                                fused persistent kernels (same outputs) */

typedef struct pqc_ctx pqc_ctx;

/* Context: owns device scratch for up to max_batch elements of every enabled
 * algorithm.  Creation fails with PQC_ERR_NO_MEMORY when the scratch would
 * exceed hbm_budget_bytes (0 = no budget).  One context per host thread. */
int pqc_ctx_create(pqc_ctx **ctx, unsigned alg_mask, size_t max_batch, size_t hbm_budget_bytes);
void pqc_ctx_destroy(pqc_ctx *ctx);
size_t pqc_ctx_device_bytes(const pqc_ctx *ctx);
int pqc_ctx_set_tuning(pqc_ctx *ctx, int fused_waves_per_cu, int fused_grid);

/* Launch parameters the autotuner chooses (replaces the reference's per-GPU
 * block-size tables and its ANALYSIS_MODE sweep, indcpa.cu:27-70,499-562).
 * Every value only changes how the work is spread, never a result.          */
typedef struct pqc_tuning {
  int32_t plain_block;        /* threads per workgroup, plain per-op Kyber kernels: 64..1024, multiple of 64 */
  int32_t newhope_block;      /* same for the NewHope-1024 kernels */
  int32_t fused_waves_per_cu; /* fused persistent grid = CUs x this: 1..4 (LDS holds 4) */
  int32_t mixed_order;        /* mixed-batch tile order: 0 = K 4,3,2; 1 = K 2,3,4; 2 = round-robin over K;
                                 3 = K 4,3,2 with 32-request tiles for K = 4 and K = 2 */
  int32_t frodo_stream_block; /* threads per workgroup of the Frodo noise-stream kernel: 64, 128 or 256 */
  int32_t lowlat_max;         /* Kyber batches of at most this many requests run one four-wave
                                 workgroup per request (latency) instead of 64-request tiles
                                 (throughput): 0 = library defaults (8192 decapsulations, 4096
                                 encapsulations), -1 = never, else 1..65536 for both */
  int32_t frodo_chunk;        /* FrodoKEM-976 batches under one key above this many requests are cut
                                 into chunks of this many, run on four streams: 0 = library default
                                 (no split up to 65536), else a multiple of 32 in 32..65536 */
  int32_t fused_dec_tile;     /* requests per tile of the fused Kyber decapsulation: 32 (two waves per
                                 SIMD, A^T in a global scratch) or 64 (one wave per SIMD, A^T in an LDS
                                 ring); 0 = library default (32) */
  int32_t fused_enc_tile;     /* the same for the fused Kyber encapsulation */
  int32_t fused_key_tile;     /* the same for the fused Kyber key generation */
  int32_t frodo_lane_min;     /* Keccak lane layout of the FrodoKEM-976 noise stream and F prefix in
                                 shared-key decapsulation: calls of at least this many requests run one state
                                 per lane (throughput), smaller ones one state per lane pair (latency);
                                 0 = library default (49152), -1 = lane pairs always, else >= 1 */
} pqc_tuning;
int pqc_ctx_get_tuning(const pqc_ctx *ctx, pqc_tuning *t);
int pqc_ctx_put_tuning(pqc_ctx *ctx, const pqc_tuning *t); /* PQC_ERR_INVALID_ARG outside the ranges */

/* Pinned (page-locked) host memory for PQC_F_HOST_PTRS buffers: the DMA
 * engines read and write it directly, with no staging copy.  Replaces the
 * reference's cudaMemcpy from pageable buffers inside its timed loop
 * (Kyber_GPU_Batched/main.cu:376-423).  Free only after the calls that use
 * the block have returned.                                                    */
int pqc_host_alloc(size_t bytes, void **out);
int pqc_host_free(void *p); /* PQC_ERR_INVALID_ARG for a pointer pqc_host_alloc did not return */

/* Byte sizes of one element: public key, secret key, ciphertext, shared secret. */
int pqc_alg_sizes(int alg, size_t *pk, size_t *sk, size_t *ct, size_t *ss);

/* ---- batched KEM (element-contiguous buffers: element i at i * size) --------
 * Device pointers unless PQC_F_HOST_PTRS; asynchronous on `stream` (a
 * hipStream_t, NULL = default stream) for device pointers.
 * coins: keypair 64 bytes per element (Kyber d || z; Frodo s || seedSE || z),
 * enc 32 bytes per element (Kyber) or 24 (Frodo: mu).
 * FrodoKEM-976 here has one key per element (frodo976.cu:87-690 per call): A is
 * expanded per element, 128 at a time; batches under one key are far faster
 * through pqc_frodo976_load_key + pqc_frodo976_enc/dec_batch.                  */
int crypto_kem_keypair_derand_batch(pqc_ctx *ctx, int alg, size_t n, uint8_t *pk, uint8_t *sk,
                                    const uint8_t *coins, unsigned flags, void *stream);
int crypto_kem_enc_derand_batch(pqc_ctx *ctx, int alg, size_t n, uint8_t *ct, uint8_t *ss,
                                const uint8_t *pk, const uint8_t *coins, unsigned flags, void *stream);
int crypto_kem_dec_batch(pqc_ctx *ctx, int alg, size_t n, uint8_t *ss, const uint8_t *ct,
                         const uint8_t *sk, unsigned flags, void *stream);
/* Same, with coins drawn from the library's randomness source. */
int crypto_kem_keypair_batch(pqc_ctx *ctx, int alg, size_t n, uint8_t *pk, uint8_t *sk,
                             unsigned flags, void *stream);
int crypto_kem_enc_batch(pqc_ctx *ctx, int alg, size_t n, uint8_t *ct, uint8_t *ss,
                         const uint8_t *pk, unsigned flags, void *stream);

/* ---- autotuner ----------------------------------------------------------------
 * Times the candidate configurations of one workload family (the knobs that
 * family uses) with HIP events and keeps every measurement as a sample
 * (14 features, log ms) in the context.  `run` launches the workload once,
 * asynchronously, on the given stream and returns 0 or a PQC_ERR code.
 * use_model = 0: candidates in a fixed order (the configuration in use first)
 * up to `budget` (0 = all).  use_model = 1: once the context's learned cost
 * model has been fitted (after the first use_model call), it ranks all
 * candidates and only the best-ranked `budget` are measured (with a budget of
 * two or more, the configuration in use is always one of them, so the result is
 * never measured slower than it); the model is then refitted on the GPU with
 * every sample.  The best measured configuration is
 * installed in the context and returned.                                       */
#define PQC_TUNE_PLAIN_KYBER 0 /* plain_block */
#define PQC_TUNE_FUSED_KYBER 1 /* fused_waves_per_cu */
#define PQC_TUNE_MIXED_KYBER 2 /* fused_waves_per_cu x tile layout: 32-request tiles both ways, or
                                  64-request tiles (fused_enc_tile = fused_dec_tile = 64) in mixed_order 0, 1, 3 */
#define PQC_TUNE_FRODO 3       /* frodo_stream_block x Keccak lane layout (frodo_lane_min) */
#define PQC_TUNE_NEWHOPE 4     /* newhope_block */
typedef int (*pqc_workload_fn)(void *user, void *stream);
typedef struct pqc_tune_request {
  int32_t family;    /* PQC_TUNE_* */
  int32_t k;         /* Kyber K (2, 3, 4), else 0 */
  int32_t op;        /* 0 key generation, 1 encapsulation, 2 decapsulation */
  int32_t budget;    /* measurements allowed, 0 = every candidate */
  uint64_t n;        /* batch size */
  int32_t use_model; /* 1 = let the learned cost model choose what to measure */
  int32_t reps;      /* timed runs per measurement (the median counts) */
} pqc_tune_request;
typedef struct pqc_tune_result {
  pqc_tuning best;
  float best_ms;      /* measured */
  float predicted_ms; /* the model's prediction for `best` (0 without the model) */
  int32_t measured;   /* configurations timed */
  int32_t reserved;
} pqc_tune_result;
int pqc_autotune(pqc_ctx *ctx, const pqc_tune_request *req, pqc_workload_fn run, void *user, void *stream,
                 pqc_tune_result *out);
/* features of (request, configuration) as the cost model sees them; returns 14 */
int pqc_tune_features(const pqc_tune_request *req, const pqc_tuning *t, float *out, int count);
/* copies up to cap samples (14 features each, log ms) and returns the count */
int pqc_tune_samples(pqc_ctx *ctx, float *x, float *y, int cap);

/* ---- the learned cost model (an MLP trained on the GPU) -------------------------
 * in_dim -> 32 -> 32 -> 1 with ReLU, fp32, loss = mean squared error.  Forward,
 * backward and Adam run in the library's own HIP kernels.  Host arrays in/out;
 * every call synchronises.  pqc_ctx_costmodel returns the context's model
 * (created on first use, owned by the context).                                */
typedef struct pqc_costmodel pqc_costmodel;
int pqc_costmodel_create(pqc_costmodel **m, int in_dim, int max_samples, uint64_t seed);
void pqc_costmodel_destroy(pqc_costmodel *m);
int pqc_costmodel_num_params(const pqc_costmodel *m);
int pqc_costmodel_trained_steps(const pqc_costmodel *m); /* Adam steps taken so far */
int pqc_costmodel_get_params(const pqc_costmodel *m, float *out, size_t count);
int pqc_costmodel_set_params(pqc_costmodel *m, const float *in, size_t count);
int pqc_costmodel_predict(pqc_costmodel *m, int n, const float *x, float *y);
int pqc_costmodel_grad(pqc_costmodel *m, int n, const float *x, const float *t, float *loss, float *grad,
                       size_t count);
int pqc_costmodel_fit(pqc_costmodel *m, int n, const float *x, const float *t, int steps, float lr,
                      float *final_loss);
pqc_costmodel *pqc_ctx_costmodel(pqc_ctx *ctx);

/* ---- mixed Kyber-512/768/1024 batches -----------------------------------------
 * One descriptor per request (the reference has no counterpart: its batches are
 * one parameter set at a time, indcpa.h:16-26).  Buffers are named by byte
 * offsets (multiples of 8) into four device arenas whose sizes are given for
 * bounds checks.  The requests are bucketed by parameter set on the device and
 * run through the fused persistent kernels; each result lands at its
 * descriptor's offsets, identical to the single-parameter-set calls.  A
 * descriptor that fails validation (algorithm not in the context, unaligned or
 * out-of-arena offset) is skipped and reported by pqc_ctx_take_status.
 * Device pointers only, asynchronous on the stream, and capturable into a graph
 * (no allocation, no host synchronisation).                                     */
typedef struct pqc_kem_req {
  int32_t alg;        /* PQC_ALG_KYBER512 / 768 / 1024 */
  uint32_t reserved;  /* 0 */
  uint64_t key_off;   /* pk (encapsulation) or sk (decapsulation) in the key arena */
  uint64_t ct_off;    /* ciphertext: written by encapsulation, read by decapsulation */
  uint64_t ss_off;    /* 32-byte shared secret */
  uint64_t coins_off; /* 32 coin bytes (encapsulation only) */
} pqc_kem_req;
int pqc_kem_enc_mixed(pqc_ctx *ctx, size_t n, const pqc_kem_req *reqs, uint8_t *ct_arena, size_t ct_bytes,
                      const uint8_t *pk_arena, size_t pk_bytes, uint8_t *ss_arena, size_t ss_bytes,
                      const uint8_t *coins_arena, size_t coins_bytes, void *stream);
int pqc_kem_dec_mixed(pqc_ctx *ctx, size_t n, const pqc_kem_req *reqs, const uint8_t *ct_arena, size_t ct_bytes,
                      const uint8_t *sk_arena, size_t sk_bytes, uint8_t *ss_arena, size_t ss_bytes,
                      void *stream);
/* Waits for the stream, then reads and clears the context's device status word:
 * bit 0 = a matrix sampler hit its block cap, bit 1 = a mixed-batch descriptor
 * was rejected.  Host-pointer calls check it themselves.                        */
int pqc_ctx_take_status(pqc_ctx *ctx, void *stream, uint32_t *status);

/* ---- allocation-free decapsulation ------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory or call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation): Kyber-512/768/1024 decapsulation of n
 * ciphertexts with no context, bit-identical to crypto_kem_dec_batch with the
 * default tuning (the same kernels).  The tile kernel's scratch (PRF rows,
 * (2K+1) x 128 bytes per request) lives in a library-owned device array of
 * 256 + PQC_NOALLOC_MAX_BATCH x 1152 bytes (18.9 MB), one per GPU, held until
 * the process ends; nothing is allocated or freed per call.  n above
 * PQC_NOALLOC_MAX_BATCH returns PQC_ERR_BATCH_TOO_LARGE.  Device pointers
 * (8-byte aligned), asynchronous on `stream`.  `status` (device word, 4-byte
 * aligned, may be NULL) collects the call's status bits (bit 0: a matrix
 * sampler hit its block cap) by OR; it is not cleared.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (on different streams or from different host threads) race and must be
 * serialised by the caller.                                                   */
#define PQC_NOALLOC_MAX_BATCH 16384
int pqc_kyber_dec_batch_noalloc(int alg, size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk,
                                uint32_t *status, void *stream);

/* ---- graph capture of device-pointer batch calls --------------------------------
 * Between begin and end, device-pointer calls on `stream` (non-NULL) are recorded
 * instead of run; the graph replays them, on the same buffers, with one launch. */
typedef struct pqc_graph pqc_graph;
int pqc_graph_capture_begin(void *stream);
int pqc_graph_capture_end(void *stream, pqc_graph **graph);
int pqc_graph_launch(pqc_graph *graph, void *stream);
void pqc_graph_destroy(pqc_graph *graph);

/* ---- serving loop ---------------------------------------------------------------
 * A server holds keys_per_alg device-resident key pairs per parameter set (key
 * pair i of set a from coins = the library's generator with key_seed, stream
 * 0x6B6579 + a, index i) and runs a scheduler thread: requests are admitted in
 * arrival order into up to `slots` batches in flight at once (continuous
 * batching: arrivals join the next free slot while others run), each batch at
 * most max_batch requests and arena_bytes of inputs / outputs, and the bytes of
 * all requests in flight at most hbm_budget (0 = no limit).  A batch runs the
 * fused mixed-parameter-set Kyber kernels and the many-key Frodo path, or with
 * PQC_SERVER_PLAIN the plain per-op kernels.  Results are the same bytes as the
 * batch calls give for the request alone.  The reference has no counterpart.  */
#define PQC_SERVER_PLAIN 1u
/* PQC_SERVER_GRAPH: each fused Kyber batch (input copies, the mixed Kyber
 * kernels, output copies) is captured into a hipGraph and launched with one
 * call.  A slot keeps one executable graph per batch shape (which request
 * groups are present); the next batch of that shape updates the executable's
 * parameters in place (hipGraphExecUpdate) instead of re-instantiating it.
 * With Kyber and Frodo both enabled, the slots form two pools: half of them (at
 * least one) take only the Frodo requests, launched eagerly, and the others
 * only the Kyber requests, as graphs, so a Kyber batch never waits for a slot
 * whose Frodo chains still run.  Same bytes as the eager launches.  Without
 * this flag, a batch holds both kinds and its Kyber requests complete as soon
 * as its Kyber kernels finish. */
#define PQC_SERVER_GRAPH 2u
typedef struct pqc_server pqc_server;
typedef struct pqc_server_config {
  uint32_t alg_mask;     /* Kyber sets and/or FrodoKEM-976 */
  uint32_t keys_per_alg; /* device-resident key pairs per parameter set */
  uint64_t key_seed;
  uint32_t max_batch;    /* requests per batch */
  uint32_t slots;        /* batches in flight, 1..8; a fused slot drives 3 streams (transfers +
                          * Kyber, Frodo encaps, Frodo decaps): streams beyond the process's
                          * hardware queues (HIP's GPU_MAX_HW_QUEUES, default 4) share queues
                          * and serialise, so keep 3 * slots within it */
  uint64_t hbm_budget;   /* bytes of in-flight request inputs + outputs, 0 = no limit */
  uint64_t arena_bytes;  /* per-slot input and output arena (pinned + device), >= 1 MiB */
  uint32_t max_wait_us;  /* while other batches run, a batch waits at most this long to fill */
  uint32_t flags;        /* PQC_SERVER_PLAIN or PQC_SERVER_GRAPH */
  pqc_tuning tuning;     /* launch parameters for the slots' contexts (all zero: defaults) */
} pqc_server_config;
typedef struct pqc_server_request {
  int32_t alg;          /* PQC_ALG_* served */
  int32_t op;           /* 1 encapsulate, 2 decapsulate */
  uint32_t key_id;      /* < keys_per_alg */
  uint32_t reserved;
  const uint8_t *in;    /* host: encaps coins (Kyber 32 B, Frodo mu 24 B) or decaps ciphertext */
  uint8_t *out;         /* host: encaps ct || ss, decaps ss; written before the completion */
  uint64_t tag;         /* returned in the completion */
} pqc_server_request;
typedef struct pqc_server_completion {
  uint64_t tag;
  int32_t status;       /* PQC_OK or PQC_ERR_* */
  uint32_t batch;       /* requests in the batch that carried it */
  double submit_us;     /* server clock */
  double done_us;
} pqc_server_completion;
typedef struct pqc_server_stats {
  uint64_t submitted, completed, batches;
  double mean_batch;
  uint64_t inflight_bytes;
  /* per batch, means over the batches so far: host time packing inputs and
   * launching, device time from the first copy to the last, host time handing
   * the outputs back */
  double pack_us, device_us, retire_us;
} pqc_server_stats;
typedef struct pqc_trace_stats {
  uint64_t completed, errors;
  double duration_s, ops_per_s;
  double p50_us, p90_us, p99_us, p999_us, max_us; /* arrival to completion */
  double mean_batch;
} pqc_trace_stats;
int pqc_server_create(pqc_server **srv, const pqc_server_config *cfg);
void pqc_server_destroy(pqc_server *srv);
/* host copies of key pair key_id (either pointer may be NULL) */
int pqc_server_key(pqc_server *srv, int alg, uint32_t key_id, uint8_t *pk, uint8_t *sk);
/* thread-safe, non-blocking; in / out must stay valid until the completion */
int pqc_server_submit(pqc_server *srv, const pqc_server_request *reqs, size_t n);
/* up to cap completions; waits up to wait_us for the first */
int pqc_server_poll(pqc_server *srv, pqc_server_completion *out, size_t cap, size_t *got, uint32_t wait_us);
int pqc_server_drain(pqc_server *srv); /* until every submitted request completed */
int pqc_server_get_stats(pqc_server *srv, pqc_server_stats *st);
double pqc_server_now_us(pqc_server *srv);
/* replays n requests at arrival_us[i] (microseconds from the call) from the calling
 * thread; latency_us[i] (optional) = completion - arrival; tags are overwritten */
int pqc_server_run_trace(pqc_server *srv, size_t n, const double *arrival_us, const pqc_server_request *reqs,
                         double *latency_us, pqc_trace_stats *out);

/* ---- NIST drop-ins: one instance, host pointers, synchronous -----------------
 * The sizes and name of Kyber_GPU_Batched/api.h:10-16 and FrodoKEM_GPU/api.h:11-17
 * (CRYPTO_SECRETKEYBYTES, CRYPTO_PUBLICKEYBYTES, CRYPTO_CIPHERTEXTBYTES,
 * CRYPTO_BYTES, CRYPTO_ALGNAME) describe the parameter set that the unprefixed
 * crypto_kem_keypair / enc / dec mean.  Define PQC_NIST_ALG to one of the
 * PQC_ALG_* KEM ids before including this header to choose it; the default is
 * PQC_ALG_KYBER768, which the exported unprefixed symbols implement.  Any other
 * choice maps the unprefixed names onto the prefixed functions below by macro,
 * so a caller written against the reference's api.h compiles and links
 * unchanged.  Note: the reference's own params.h:13 defaults to KYBER_K 4
 * ("Kyber1024"); a caller that relied on that default defines
 * PQC_NIST_ALG PQC_ALG_KYBER1024.  With PQC_ALG_FRODO976, the reference's
 * crypto_kem_keypair_gpu and crypto_kem_keypair_random (FrodoKEM_GPU/api.h:21-22)
 * also map onto the Frodo key generation.  NewHope: NewHope_NIST_GPU/api.h:10-23
 * names only the CPA-KEM sizes, whose KEM functions the reference never wrote;
 * here PQC_ALG_NEWHOPE1024 / PQC_ALG_NEWHOPE512 select the CCA-KEM drop-ins
 * (sizes of params.h:37-39).
 *
 * Every call runs on a library-owned context of the current device.  Calls
 * from several host threads are safe: each call checks a context out of a
 * per-device pool for its whole duration, so no two calls share staging or
 * pinned buffers or a status word (the pool grows to the largest number of
 * concurrent callers and is kept until the process exits).                  */
#ifndef PQC_NIST_ALG
#define PQC_NIST_ALG PQC_ALG_KYBER768
#endif
#if PQC_NIST_ALG == PQC_ALG_KYBER512
#define CRYPTO_SECRETKEYBYTES 1632
#define CRYPTO_PUBLICKEYBYTES 800
#define CRYPTO_CIPHERTEXTBYTES 736
#define CRYPTO_BYTES 32
#define CRYPTO_ALGNAME "Kyber512"
#define crypto_kem_keypair pqc_kyber512_crypto_kem_keypair
#define crypto_kem_enc pqc_kyber512_crypto_kem_enc
#define crypto_kem_dec pqc_kyber512_crypto_kem_dec
#elif PQC_NIST_ALG == PQC_ALG_KYBER768
#define CRYPTO_SECRETKEYBYTES 2400
#define CRYPTO_PUBLICKEYBYTES 1184
#define CRYPTO_CIPHERTEXTBYTES 1088
#define CRYPTO_BYTES 32
#define CRYPTO_ALGNAME "Kyber768"
#elif PQC_NIST_ALG == PQC_ALG_KYBER1024
#define CRYPTO_SECRETKEYBYTES 3168
#define CRYPTO_PUBLICKEYBYTES 1568
#define CRYPTO_CIPHERTEXTBYTES 1568
#define CRYPTO_BYTES 32
#define CRYPTO_ALGNAME "Kyber1024"
#define crypto_kem_keypair pqc_kyber1024_crypto_kem_keypair
#define crypto_kem_enc pqc_kyber1024_crypto_kem_enc
#define crypto_kem_dec pqc_kyber1024_crypto_kem_dec
#elif PQC_NIST_ALG == PQC_ALG_FRODO976
#define CRYPTO_SECRETKEYBYTES 31296
#define CRYPTO_PUBLICKEYBYTES 15632
#define CRYPTO_CIPHERTEXTBYTES 15744
#define CRYPTO_BYTES 24
#define CRYPTO_ALGNAME "FrodoKEM-976"
#define crypto_kem_keypair pqc_frodo976_crypto_kem_keypair
#define crypto_kem_keypair_gpu pqc_frodo976_crypto_kem_keypair
#define crypto_kem_keypair_random pqc_frodo976_crypto_kem_keypair
#define crypto_kem_enc pqc_frodo976_crypto_kem_enc
#define crypto_kem_dec pqc_frodo976_crypto_kem_dec
#elif PQC_NIST_ALG == PQC_ALG_NEWHOPE1024
#define CRYPTO_SECRETKEYBYTES 3680
#define CRYPTO_PUBLICKEYBYTES 1824
#define CRYPTO_CIPHERTEXTBYTES 2208
#define CRYPTO_BYTES 32
#define CRYPTO_ALGNAME "NewHope1024-CCAKEM"
#define crypto_kem_keypair pqc_newhope1024_crypto_kem_keypair
#define crypto_kem_enc pqc_newhope1024_crypto_kem_enc
#define crypto_kem_dec pqc_newhope1024_crypto_kem_dec
#elif PQC_NIST_ALG == PQC_ALG_NEWHOPE512
#define CRYPTO_SECRETKEYBYTES 1888
#define CRYPTO_PUBLICKEYBYTES 928
#define CRYPTO_CIPHERTEXTBYTES 1120
#define CRYPTO_BYTES 32
#define CRYPTO_ALGNAME "NewHope512-CCAKEM"
#define crypto_kem_keypair pqc_newhope512_crypto_kem_keypair
#define crypto_kem_enc pqc_newhope512_crypto_kem_enc
#define crypto_kem_dec pqc_newhope512_crypto_kem_dec
#else
#error "PQC_NIST_ALG must be a KEM: PQC_ALG_KYBER512/768/1024, PQC_ALG_FRODO976, PQC_ALG_NEWHOPE1024/512"
#endif
int crypto_kem_keypair(unsigned char *pk, unsigned char *sk);
int crypto_kem_enc(unsigned char *ct, unsigned char *ss, const unsigned char *pk);
int crypto_kem_dec(unsigned char *ss, const unsigned char *ct, const unsigned char *sk);
int pqc_kyber512_crypto_kem_keypair(unsigned char *pk, unsigned char *sk);
int pqc_kyber512_crypto_kem_enc(unsigned char *ct, unsigned char *ss, const unsigned char *pk);
int pqc_kyber512_crypto_kem_dec(unsigned char *ss, const unsigned char *ct, const unsigned char *sk);
int pqc_kyber768_crypto_kem_keypair(unsigned char *pk, unsigned char *sk);
int pqc_kyber768_crypto_kem_enc(unsigned char *ct, unsigned char *ss, const unsigned char *pk);
int pqc_kyber768_crypto_kem_dec(unsigned char *ss, const unsigned char *ct, const unsigned char *sk);
int pqc_kyber1024_crypto_kem_keypair(unsigned char *pk, unsigned char *sk);
int pqc_kyber1024_crypto_kem_enc(unsigned char *ct, unsigned char *ss, const unsigned char *pk);
int pqc_kyber1024_crypto_kem_dec(unsigned char *ss, const unsigned char *ct, const unsigned char *sk);
/* FrodoKEM-976-SHAKE: FrodoKEM_GPU/api.h:20-24 (crypto_kem_keypair / enc / dec) */
int pqc_frodo976_crypto_kem_keypair(unsigned char *pk, unsigned char *sk);
int pqc_frodo976_crypto_kem_enc(unsigned char *ct, unsigned char *ss, const unsigned char *pk);
int pqc_frodo976_crypto_kem_dec(unsigned char *ss, const unsigned char *ct, const unsigned char *sk);
/* NewHope-1024 CCA-KEM: the sizes of NewHope_NIST_GPU/params.h:37-39 and api.h:11-14;
 * the reference has no KEM code, the transform is the NewHope specification's */
int pqc_newhope1024_crypto_kem_keypair(unsigned char *pk, unsigned char *sk);
int pqc_newhope1024_crypto_kem_enc(unsigned char *ct, unsigned char *ss, const unsigned char *pk);
int pqc_newhope1024_crypto_kem_dec(unsigned char *ss, const unsigned char *ct, const unsigned char *sk);
/* NewHope-512 CCA-KEM: the same transform at N = 512 */
int pqc_newhope512_crypto_kem_keypair(unsigned char *pk, unsigned char *sk);
int pqc_newhope512_crypto_kem_enc(unsigned char *ct, unsigned char *ss, const unsigned char *pk);
int pqc_newhope512_crypto_kem_dec(unsigned char *ss, const unsigned char *ct, const unsigned char *sk);

/* ---- NewHope-1024 CPA-PKE, plain batched kernels ------------------------------
 * Replaces cpapke_keypair/enc/dec (NewHope_NIST_GPU/cpapke.h:14-24).
 * Element sizes: seed 32, pk 1824, sk 1792, ct 2176, message 32, coins 32.
 * Requires PQC_ALG_NEWHOPE1024_CPA in the context's mask.                       */
int pqc_newhope1024_cpa_keypair_batch(pqc_ctx *ctx, size_t n, uint8_t *pk, uint8_t *sk,
                                      const uint8_t *seeds, unsigned flags, void *stream);
int pqc_newhope1024_cpa_enc_batch(pqc_ctx *ctx, size_t n, uint8_t *ct, const uint8_t *msgs,
                                  const uint8_t *pk, const uint8_t *coins, unsigned flags, void *stream);
int pqc_newhope1024_cpa_dec_batch(pqc_ctx *ctx, size_t n, uint8_t *msgs, const uint8_t *ct,
                                  const uint8_t *sk, unsigned flags, void *stream);
/* NewHope-512 CPA-PKE: the reference's NEWHOPE_N == 512 configuration
 * (params.h:9-39 at N 512; precomp.cu:75-190; ntt.cu:13-37,130-180;
 * poly.cu:244,279).  Element sizes: seed 32, pk 928, sk 896, ct 1088.
 * Requires PQC_ALG_NEWHOPE512_CPA in the context's mask.                        */
int pqc_newhope512_cpa_keypair_batch(pqc_ctx *ctx, size_t n, uint8_t *pk, uint8_t *sk,
                                     const uint8_t *seeds, unsigned flags, void *stream);
int pqc_newhope512_cpa_enc_batch(pqc_ctx *ctx, size_t n, uint8_t *ct, const uint8_t *msgs,
                                 const uint8_t *pk, const uint8_t *coins, unsigned flags, void *stream);
int pqc_newhope512_cpa_dec_batch(pqc_ctx *ctx, size_t n, uint8_t *msgs, const uint8_t *ct,
                                 const uint8_t *sk, unsigned flags, void *stream);

/* ---- FrodoKEM-976-SHAKE: batches under one device-resident key ------------------
 * Replaces crypto_kem_keypair_gpu/enc/dec (FrodoKEM_GPU/api.h:20-24,
 * frodo976.cu:87-690).  Element sizes: randomness 64 (s || seedSE || z),
 * pk 15632, sk 31296, ct 15744, ss 24, mu 24.
 * pqc_frodo976_load_key expands A once (is_sk = 1: a secret key, needed for
 * decapsulation; 0: a public key); enc/dec batches then share it, and their
 * S'.A and S'.B products run as one int8-MFMA GEMM.  pqc_frodo976_products
 * copies the last batch's B'' (8 x 976 uint16 per element) and V (8 x 8).  */
int pqc_frodo976_keypair_batch(pqc_ctx *ctx, size_t n, uint8_t *pk, uint8_t *sk,
                               const uint8_t *randomness, unsigned flags, void *stream);
int pqc_frodo976_load_key(pqc_ctx *ctx, const uint8_t *key, int is_sk, unsigned flags, void *stream);
int pqc_frodo976_enc_batch(pqc_ctx *ctx, size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *mu,
                           unsigned flags, void *stream);
int pqc_frodo976_dec_batch(pqc_ctx *ctx, size_t n, uint8_t *ss, const uint8_t *ct, unsigned flags,
                           void *stream);
int pqc_frodo976_products(pqc_ctx *ctx, size_t n, uint16_t *bpp, uint16_t *v, unsigned flags,
                          void *stream);

/* ---- randomness for the non-derand entry points ------------------------------- */
void pqc_randombytes_seed(uint64_t seed); /* deterministic SHAKE256 DRBG */
void pqc_randombytes_system(void);        /* OS randomness (default) */
int pqc_randombytes(uint8_t *out, size_t n);
/* The NIST PQC known-answer-test source under the names of the reference's
 * rng.h (Kyber_GPU_Batched/rng.h:22-26, whose rng.cu:17-39 body is a counter):
 * AES-256 CTR_DRBG without derivation function (SP 800-90A 10.2.1), seeded
 * with 48 bytes.  After randombytes_init every non-derand entry point draws
 * from it, call by call as the NIST submissions do, so PQCgenKAT_kem.c links
 * against this library unchanged.  pqc_randombytes_seed/_system switch away. */
void randombytes_init(unsigned char *entropy_input, unsigned char *personalization_string,
                      int security_strength);
int randombytes(unsigned char *x, unsigned long long xlen);

/* ---- device plumbing, so that callers need no other GPU binding -------------- */
int pqc_device_count(int *count);
int pqc_set_device(int device);
int pqc_device_info(char *name, size_t name_len, int *cu_count, size_t *hbm_bytes);
int pqc_device_malloc(void **ptr, size_t bytes);
int pqc_device_free(void *ptr);
int pqc_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int pqc_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int pqc_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
int pqc_stream_create(void **stream);
int pqc_stream_destroy(void *stream);
int pqc_stream_synchronize(void *stream);
int pqc_event_create(void **event);
int pqc_event_destroy(void *event);
int pqc_event_record(void *event, void *stream);
int pqc_event_synchronize(void *event);
int pqc_event_elapsed_ms(float *ms, void *start, void *stop);

/* ---- component kernels (tests and microbenchmarks) ---------------------------
 * pqc_keccak_batch: n independent fixed-length messages (element i at
 * in + i*inlen, digest at out + i*outlen); op 0 SHA3-256 (outlen 32),
 * 1 SHA3-512 (outlen 64), 2 SHAKE128, 3 SHAKE256.  inlen must be a multiple
 * of 8.  Replaces sha3_256_n / sha3_512_n (fips202.cu:634-717).
 * pqc_kyber_poly_transform: n polynomials of 256 int16 in place; op 0 =
 * poly_ntt (ntt + Barrett), 1 = poly_invntt.
 * pqc_newhope1024_poly_transform: n polynomials of 1024 uint16 in place
 * (element-contiguous, device memory); op 0 = poly_ntt (mul_coefficients by
 * gammas + ntt), 1 = poly_invntt (bitrev_vector + ntt + mul_coefficients),
 * replacing NewHope_NIST_GPU/poly.cu:519-547 (the reference runs it per
 * polynomial inside its CPA kernels; here one wave per polynomial).
 * pqc_newhope1024_poly_sample: n noise polynomials (1024 uint16 each, device)
 * from the 32-byte seeds at seeds + i*seed_stride (seed_stride a multiple of 8)
 * and one nonce byte: poly_sample (NewHope_NIST_GPU/poly.cu:396-444), one lane
 * per 64-coefficient SHAKE256 stream.
 * pqc_frodo976_gemm: re-runs only the shared-A product
 * [B'' | V] = S'.[A | B] + [E' | E''] of the last Frodo-976 batch of n >= this n
 * elements (frodo_mul_add_sa_plus_e_gpu + frodo_mul_add_sb_plus_e_kernel,
 * frodo_macrify_reference.cu:215-234,261-281), asynchronously on the stream.
 * With ct (device pointer, the last decapsulation batch's ciphertexts) it runs
 * the decapsulation form that compares against ct; with ct = NULL the form
 * that writes B'' and V to the context (read them with pqc_frodo976_products). */
int pqc_keccak_batch(int op, size_t n, uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen,
                     void *stream);
int pqc_kyber_poly_transform(int op, size_t n, int16_t *polys, void *stream);
int pqc_newhope1024_poly_transform(int op, size_t n, uint16_t *polys, void *stream);
int pqc_newhope1024_poly_sample(size_t n, uint16_t *polys, const uint8_t *seeds, size_t seed_stride, int nonce,
                                void *stream);
int pqc_frodo976_gemm(pqc_ctx *ctx, size_t n, const uint8_t *ct, void *stream);
/* VALU issue-rate microbenchmark (no reference counterpart): sustained cycles
 * per wave64 instruction per SIMD of one instruction kind, op 0 = v_xor_b32,
 * 1 = v_add_u32, 2 = v_alignbit_b32, 3 = v_mul_lo_u32, 4 = v_bitop3_b32,
 * 5 = v_lshlrev_b32, 6 = v_mad_u32_u24, 7 = v_sub_u32, 8 = v_and_b32,
 * 9 = v_ashrrev_i32, 10 = v_fma_f32, 11 = v_mul_f32, 12 = v_mul_u32_u24,
 * 13 = v_mad_i32_i24, 14 = v_bfe_u32, 15 = v_perm_b32, eight independent chains
 * per wave, over a full grid at waves_per_simd (1..8) waves per SIMD on the
 * current device (best of three launches, synchronous on the stream).  It checks
 * on the device the 32-bit vector peak the VALU-bound kernels are priced
 * against, and which instruction kinds issue at the full rate.               */
int pqc_valu_issue_rate(int op, int waves_per_simd, double *cycles_per_instr, void *stream);
/* Keccak-f[1600] throughput microbenchmark (no reference counterpart): state
 * permutations per second of the whole device for the three state layouts the
 * kernels use (form 0 one state per lane, 1 one state per lane pair with DPP
 * swaps, 2 the bit-interleaved lane pair) at waves_per_simd (1..8) waves per
 * SIMD.  Synchronous on the stream.                                          */
int pqc_keccak_rate(int form, int waves_per_simd, double *state_perms_per_s, void *stream);

const char *pqc_version(void);
const char *pqc_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif /* PQC_API_H */
