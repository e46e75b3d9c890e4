// KEM serving loop: device-resident key pairs, continuous batching over several
// in-flight slots, admission by HBM budget, latency accounting.
//
// The reference has no server; its benchmark loops whole batches synchronously
// (Kyber_GPU_Batched/main.cu:376-428).  Here a scheduler thread keeps up to
// `slots` batches on the GPU at once: requests that arrive while batches are
// running are admitted into the next free slot as soon as it and the budget
// allow, so their transfers and kernels overlap the running ones.  Each slot
// owns a context (scratch), a stream, pinned staging and device arenas.
//
// A batch is laid out as dense blocks per (operation, parameter set):
//   Kyber encaps: coins -> ct, ss      Kyber decaps: ct -> ss
//   Frodo encaps: mu -> ct, ss         Frodo decaps: ct -> ss
// Fused mode runs all Kyber requests of the batch through the mixed-batch
// kernels (descriptors name each request's key in the key store) and the Frodo
// requests through the many-key path with the cached A of each key.  Plain mode
// gathers each request's key into dense arrays and runs the plain per-op Kyber
// kernels per parameter set, and Frodo with A expanded per request.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "internal.h"
#include "threadpool.h"

namespace {

using clk = std::chrono::steady_clock;

struct Set {
  int alg;
  size_t pk, sk, ct, ss;
};

bool set_of(int alg, Set* s) {
  if (alg != PQC_ALG_KYBER512 && alg != PQC_ALG_KYBER768 && alg != PQC_ALG_KYBER1024 && alg != PQC_ALG_FRODO976)
    return false;
  s->alg = alg;
  return pqc_alg_sizes(alg, &s->pk, &s->sk, &s->ct, &s->ss) == PQC_OK;
}

// groups of a batch: 0..2 Kyber encaps K 2..4, 3..5 Kyber decaps, 6 Frodo encaps, 7 Frodo decaps
constexpr int kGroups = 8;
int group_of(int alg, int op) {
  const bool enc = op == 1;
  if (alg == PQC_ALG_FRODO976) return enc ? 6 : 7;
  return (enc ? 0 : 3) + (alg - PQC_ALG_KYBER512);
}
size_t in_bytes(const Set& s, int op) { return op == 1 ? (s.alg == PQC_ALG_FRODO976 ? 24 : 32) : s.ct; }
size_t out_bytes(const Set& s, int op) { return op == 1 ? s.ct + s.ss : s.ss; }
size_t align8(size_t v) { return (v + 7) & ~(size_t)7; }

struct Pending {
  pqc_server_request r;
  Set set;
  double t_submit;
};

// Slot classes.  With Kyber and Frodo both enabled in graph mode the slots form
// two pools: a Frodo batch is bounded by its two serial SHAKE256 chains (several
// milliseconds whatever its size), a Kyber batch by its much shorter kernels, so
// Kyber batches keep their own slots instead of waiting for a slot whose Frodo
// chains still run.
constexpr int kAnyClass = 0, kKyberClass = 1, kFrodoClass = 2;

struct Slot {
  pqc_ctx* ctx = nullptr;
  pqc_ctx* ctx_fd = nullptr;  // fused mode: Frodo decapsulation, with scratch of its own
  hipStream_t s = nullptr;    // transfers and the Kyber kernels
  hipStream_t s2 = nullptr;   // fused mode: Frodo encapsulation, overlapping the Kyber kernels
  hipStream_t s3 = nullptr;   // fused mode: Frodo decapsulation, overlapping both
  hipEvent_t done = nullptr, in_ready = nullptr, frodo_done = nullptr, fd_done = nullptr;
  hipEvent_t start = nullptr;  // timing pair with `done`
  // eager fused batches with Kyber and Frodo requests: the Kyber outputs (a prefix
  // of the output arena) and a status snapshot come back before the Frodo chains
  // end, and the Kyber requests complete at `kyber_done`
  hipEvent_t kyber_done = nullptr;
  bool kyber_pending = false, kyber_retired = false;
  uint8_t *h_in = nullptr, *h_out = nullptr, *d_in = nullptr, *d_out = nullptr, *d_keys = nullptr;
  pqc_kem_req *h_desc = nullptr, *d_desc = nullptr;  // Kyber encaps first, then decaps
  uint32_t *h_kid = nullptr, *d_kid = nullptr;       // key index per request, in batch order
  std::vector<Pending> reqs;
  std::vector<size_t> out_a, out_b, len_a, len_b;  // where each request's output lies in h_out
  std::vector<size_t> in_at;                       // where each request's input goes in h_in
  uint32_t* h_status = nullptr;  // pinned: status of ctx and ctx_fd after the batch; [2]: ctx after the Kyber part
  std::map<uint32_t, hipGraphExec_t> graphs;  // graph mode: one executable per batch shape (groups present)
  size_t bytes = 0;
  bool busy = false;
  int cls = kAnyClass;  // which requests the slot takes (split pools: Kyber or Frodo only)
};

}  // namespace

struct pqc_server {
  pqc_server_config cfg{};
  unsigned kyber_mask = 0;
  // key store: per parameter set, keys_per_alg key pairs
  uint8_t* d_pk = nullptr;  // all Kyber public keys, sets back to back
  uint8_t* d_sk = nullptr;  // all Kyber secret keys
  size_t pk_base[4] = {0}, sk_base[4] = {0}, pk_bytes = 0, sk_bytes = 0;
  uint8_t* d_fsk = nullptr;  // Frodo secret keys
  uint16_t* d_fa = nullptr;  // Frodo expanded A per key
  std::vector<uint8_t> h_pk[6], h_sk[6];  // host copies (for clients and tests)
  std::vector<Slot> slots;
  size_t arena = 0, key_arena = 0;
  // queues
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> submitted;
  std::deque<Pending> pending;    // every request, or the Kyber ones with split pools
  std::deque<Pending> pending_f;  // split pools: the Frodo requests
  bool split = false;
  std::deque<pqc_server_completion> completions;
  size_t inflight_bytes = 0;
  uint64_t n_submitted = 0, n_completed = 0, n_batches = 0, n_batched_reqs = 0;
  double pack_us = 0, device_us = 0, retire_us = 0;  // totals over the batches
  std::atomic<bool> stop{false};
  pqc::CopyPool copy_pool{8};  // packs batches and delivers results (the worker thread is thread 0)
  std::thread worker;
  clk::time_point t0;
  int device = 0;

  double now_us() const { return std::chrono::duration<double, std::micro>(clk::now() - t0).count(); }
};

namespace {

constexpr size_t kCopyPer = 128;  // requests per copy thread (about 140 KB of Kyber-768 ciphertexts)

void free_slot(Slot& sl) {
  for (auto& g : sl.graphs)
    if (g.second) (void)hipGraphExecDestroy(g.second);
  sl.graphs.clear();
  if (sl.ctx) pqc_ctx_destroy(sl.ctx);
  if (sl.ctx_fd) pqc_ctx_destroy(sl.ctx_fd);
  for (hipStream_t t : {sl.s, sl.s2, sl.s3})
    if (t) (void)hipStreamDestroy(t);
  for (hipEvent_t e : {sl.done, sl.in_ready, sl.frodo_done, sl.fd_done, sl.start, sl.kyber_done})
    if (e) (void)hipEventDestroy(e);
  for (void* p : {(void*)sl.h_in, (void*)sl.h_out, (void*)sl.h_desc, (void*)sl.h_kid, (void*)sl.h_status})
    if (p) (void)hipHostFree(p);
  for (void* p : {(void*)sl.d_in, (void*)sl.d_out, (void*)sl.d_desc, (void*)sl.d_kid, (void*)sl.d_keys})
    if (p) (void)hipFree(p);
}

void free_server(pqc_server* v) {
  for (auto& sl : v->slots) free_slot(sl);
  for (void* p : {(void*)v->d_pk, (void*)v->d_sk, (void*)v->d_fsk, (void*)v->d_fa})
    if (p) (void)hipFree(p);
  delete v;
}

// Graph mode: the slot's executable graph for this batch shape takes the new
// batch's parameters (sizes, counts, grids) by an in-place update; a first batch
// of a shape, or an update the runtime refuses (a different kernel, say), makes
// a new executable.  The slot's previous batch has finished before the slot is
// reused, so no update touches a graph in flight.
int launch_graph(Slot& sl, uint32_t shape, hipGraph_t g, hipStream_t st) {
  hipGraphExec_t& ex = sl.graphs[shape];
  if (ex) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
    if (hipGraphExecUpdate(ex, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(ex);
      ex = nullptr;
    }
  }
  if (!ex && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
    ex = nullptr;
    return PQC_ERR_DEVICE;
  }
  return hipGraphLaunch(ex, st) == hipSuccess ? PQC_OK : PQC_ERR_DEVICE;
}

// Launch one slot's batch: H2D inputs, kernels, D2H outputs, event.
int launch(pqc_server* v, Slot& sl) {
  const bool plain = (v->cfg.flags & PQC_SERVER_PLAIN) != 0;
  // block layout
  size_t cnt[kGroups] = {0};
  for (auto& p : sl.reqs) cnt[group_of(p.r.alg, p.r.op)]++;
  size_t in_base[kGroups], out_base[kGroups], out2_base[kGroups], io = 0, oo = 0;
  size_t grp_first[kGroups], pos = 0;
  int alg_of_group[kGroups];
  Set gs[kGroups];
  for (int g = 0; g < kGroups; ++g) {
    alg_of_group[g] = g < 6 ? PQC_ALG_KYBER512 + g % 3 : PQC_ALG_FRODO976;
    set_of(alg_of_group[g], &gs[g]);
    const int op = (g < 3 || g == 6) ? 1 : 2;
    grp_first[g] = pos;
    pos += cnt[g];
    in_base[g] = io;
    io += align8(cnt[g] * in_bytes(gs[g], op));
    out_base[g] = oo;
    if (op == 1) {  // ct block, then ss block
      oo += align8(cnt[g] * gs[g].ct);
      out2_base[g] = oo;
      oo += align8(cnt[g] * gs[g].ss);
    } else {
      out2_base[g] = oo;
      oo += align8(cnt[g] * gs[g].ss);
    }
  }
  // place every request: inputs into h_in, output spans, key ids, descriptors
  size_t fill[kGroups] = {0};
  sl.out_a.assign(sl.reqs.size(), 0);
  sl.out_b.assign(sl.reqs.size(), 0);
  sl.len_a.assign(sl.reqs.size(), 0);
  sl.len_b.assign(sl.reqs.size(), 0);
  // descriptor slots: Kyber encaps (K = 2, 3, 4) first, then decaps
  size_t desc_first[6], dpos = 0;
  for (int g = 0; g < 6; ++g) {
    desc_first[g] = dpos;
    dpos += cnt[g];
  }
  const size_t n_enc = cnt[0] + cnt[1] + cnt[2];
  sl.in_at.assign(sl.reqs.size(), 0);
  for (size_t i = 0; i < sl.reqs.size(); ++i) {
    const Pending& p = sl.reqs[i];
    const int g = group_of(p.r.alg, p.r.op);
    const size_t j = fill[g]++;
    const Set& s = gs[g];
    const size_t ib = in_bytes(s, p.r.op);
    sl.in_at[i] = in_base[g] + j * ib;
    sl.h_kid[grp_first[g] + j] = p.r.key_id;
    if (p.r.op == 1) {
      sl.out_a[i] = out_base[g] + j * s.ct;
      sl.len_a[i] = s.ct;
      sl.out_b[i] = out2_base[g] + j * s.ss;
      sl.len_b[i] = s.ss;
    } else {
      sl.out_a[i] = out2_base[g] + j * s.ss;
      sl.len_a[i] = s.ss;
    }
    if (g < 6) {  // Kyber descriptor
      const int kk = g % 3;
      pqc_kem_req& d = sl.h_desc[desc_first[g] + j];
      d.alg = p.r.alg;
      d.reserved = 0;
      d.key_off = g < 3 ? v->pk_base[kk] + (size_t)p.r.key_id * s.pk : v->sk_base[kk] + (size_t)p.r.key_id * s.sk;
      if (g < 3) {
        d.coins_off = in_base[g] + j * 32;
        d.ct_off = out_base[g] + j * s.ct;
        d.ss_off = out2_base[g] + j * s.ss;
      } else {
        d.coins_off = 0;
        d.ct_off = in_base[g] + j * s.ct;
        d.ss_off = out2_base[g] + j * s.ss;
      }
    }
  }
  // the inputs themselves, copied by several threads
  v->copy_pool.parallel_for(sl.reqs.size(), kCopyPer, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) memcpy(sl.h_in + sl.in_at[i], sl.reqs[i].r.in, in_bytes(sl.reqs[i].set, sl.reqs[i].r.op));
  });
  hipStream_t st = sl.s;
  const size_t n = sl.reqs.size();
  const size_t n_kyber = n_enc + cnt[3] + cnt[4] + cnt[5];
  if (hipEventRecord(sl.start, st) != hipSuccess) return PQC_ERR_DEVICE;
  sl.h_status[0] = sl.h_status[1] = sl.h_status[2] = 0;
  // graph mode captures the Kyber batches; with split pools the Frodo batches stay
  // eager (their launch sequence is short next to their serial chains, and capture
  // plus update cost more than it saves there)
  const bool graph = !plain && (v->cfg.flags & PQC_SERVER_GRAPH) != 0 && sl.cls != kFrodoClass;
  const bool early = !plain && !graph && n_kyber && (cnt[6] || cnt[7]);
  const size_t kyber_end = out_base[6];  // the Kyber groups' outputs come first
  if (graph && hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return PQC_ERR_DEVICE;
  // everything from here to the output copies is stream-ordered work on st (and
  // the Frodo streams forked off it), recorded into a graph in graph mode
  auto issue = [&]() -> int {
    if (pqc::zero_words(sl.ctx->status, 1, st) || (sl.ctx_fd && pqc::zero_words(sl.ctx_fd->status, 1, st)))
      return PQC_ERR_DEVICE;
    if (hipMemcpyAsync(sl.d_in, sl.h_in, io, hipMemcpyHostToDevice, st) != hipSuccess) return PQC_ERR_DEVICE;
    if (hipMemcpyAsync(sl.d_kid, sl.h_kid, n * 4, hipMemcpyHostToDevice, st) != hipSuccess) return PQC_ERR_DEVICE;
    // fused mode: Frodo encapsulation and decapsulation each run on a stream of their
    // own once the inputs are in
    hipStream_t sf = plain ? st : sl.s2, sd = plain ? st : sl.s3;
    if (!plain && (cnt[6] || cnt[7])) {
      if (hipEventRecord(sl.in_ready, st) != hipSuccess) return PQC_ERR_DEVICE;
      if (cnt[6] && hipStreamWaitEvent(sf, sl.in_ready, 0) != hipSuccess) return PQC_ERR_DEVICE;
      if (cnt[7] && hipStreamWaitEvent(sd, sl.in_ready, 0) != hipSuccess) return PQC_ERR_DEVICE;
    }
    int rc = PQC_OK;
    if (!plain) {
      if (n_kyber &&
          hipMemcpyAsync(sl.d_desc, sl.h_desc, n_kyber * sizeof(pqc_kem_req), hipMemcpyHostToDevice, st) != hipSuccess)
        return PQC_ERR_DEVICE;
      if (n_enc)
        rc = pqc_kem_enc_mixed(sl.ctx, n_enc, sl.d_desc, sl.d_out, v->arena, v->d_pk, v->pk_bytes, sl.d_out, v->arena,
                               sl.d_in, v->arena, st);
      if (!rc && n_kyber > n_enc)
        rc = pqc_kem_dec_mixed(sl.ctx, n_kyber - n_enc, sl.d_desc + n_enc, sl.d_in, v->arena, v->d_sk, v->sk_bytes,
                               sl.d_out, v->arena, st);
    } else {
      for (int g = 0; g < 6 && !rc; ++g) {
        if (!cnt[g]) continue;
        const int kk = g % 3, K = kk + 2;
        const Set& s = gs[g];
        if (g < 3) {
          rc = pqc::gather_records(cnt[g], sl.d_keys, s.pk, v->d_pk + v->pk_base[kk], s.pk, sl.d_kid + grp_first[g], st);
          if (!rc)
            rc = pqc::kyber_plain_enc(sl.ctx, K, cnt[g], sl.d_out + out_base[g], sl.d_out + out2_base[g], sl.d_keys,
                                      sl.d_in + in_base[g], st);
        } else {
          rc = pqc::gather_records(cnt[g], sl.d_keys, s.sk, v->d_sk + v->sk_base[kk], s.sk, sl.d_kid + grp_first[g], st);
          if (!rc)
            rc = pqc::kyber_plain_dec(sl.ctx, K, cnt[g], sl.d_out + out2_base[g], sl.d_in + in_base[g], sl.d_keys, st);
        }
      }
    }
    // eager fused mode: the Kyber outputs and status come back now, ahead of the Frodo chains
    if (!rc && early) {
      if (hipMemcpyAsync(sl.h_out, sl.d_out, kyber_end, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipMemcpyAsync(sl.h_status + 2, sl.ctx->status, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipEventRecord(sl.kyber_done, st) != hipSuccess)
        return PQC_ERR_DEVICE;
    }
    // Frodo groups
    if (!rc && cnt[6]) {
      if (!plain) {
        rc = pqc::frodo_enc_store(sl.ctx, cnt[6], sl.d_out + out_base[6], sl.d_out + out2_base[6], sl.d_in + in_base[6],
                                  v->d_fsk, v->d_fa, sl.d_kid + grp_first[6], sf);
      } else {
        // public keys gathered out of the secret keys (pk at byte 24 of sk); records
        // start 8 bytes early to keep 8-byte alignment, so the pk sits at +8.  A is
        // expanded per request.
        rc = pqc::gather_records(cnt[6], sl.d_keys, 15632 + 8, v->d_fsk + 24 - 8, 31296, sl.d_kid + grp_first[6], st);
        if (!rc)
          rc = pqc::frodo_enc_batch_strided(sl.ctx, cnt[6], sl.d_out + out_base[6], sl.d_out + out2_base[6],
                                            sl.d_keys + 8, 15632 + 8, sl.d_in + in_base[6], st);
      }
    }
    if (!rc && cnt[7]) {
      if (!plain) {
        rc = pqc::frodo_dec_store(sl.ctx_fd, cnt[7], sl.d_out + out2_base[7], sl.d_in + in_base[7], v->d_fsk,
                                  v->d_fa, sl.d_kid + grp_first[7], sd);
      } else {
        rc = pqc::gather_records(cnt[7], sl.d_keys, 31296, v->d_fsk, 31296, sl.d_kid + grp_first[7], st);
        if (!rc) rc = pqc::frodo_dec_batch(sl.ctx, cnt[7], sl.d_out + out2_base[7], sl.d_in + in_base[7], sl.d_keys, st);
      }
    }
    if (!rc && !plain && cnt[6] &&
        (hipEventRecord(sl.frodo_done, sf) != hipSuccess || hipStreamWaitEvent(st, sl.frodo_done, 0) != hipSuccess))
      rc = PQC_ERR_DEVICE;
    if (!rc && !plain && cnt[7] &&
        (hipEventRecord(sl.fd_done, sd) != hipSuccess || hipStreamWaitEvent(st, sl.fd_done, 0) != hipSuccess))
      rc = PQC_ERR_DEVICE;
    if (rc) return rc;
    const size_t o0 = early ? kyber_end : 0;
    if (hipMemcpyAsync(sl.h_out + o0, sl.d_out + o0, oo - o0, hipMemcpyDeviceToHost, st) != hipSuccess)
      return PQC_ERR_DEVICE;
    // the contexts' status words (cleared before the batch) come back with the outputs
    if (hipMemcpyAsync(sl.h_status, sl.ctx->status, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (sl.ctx_fd && hipMemcpyAsync(sl.h_status + 1, sl.ctx_fd->status, 4, hipMemcpyDeviceToHost, st) != hipSuccess))
      return PQC_ERR_DEVICE;
    return PQC_OK;
  };
  int rc = issue();
  if (graph) {
    // the capture always ends, also after an error, so the stream leaves capture mode
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (!rc && (e != hipSuccess || !g)) rc = PQC_ERR_DEVICE;
    uint32_t shape = 0;
    for (int k = 0; k < kGroups; ++k) shape |= cnt[k] ? 1u << k : 0u;
    if (!rc) rc = launch_graph(sl, shape, g, st);
    if (g) (void)hipGraphDestroy(g);
  }
  if (rc) return rc;
  if (hipEventRecord(sl.done, st) != hipSuccess) return PQC_ERR_DEVICE;
  sl.kyber_pending = early;
  sl.kyber_retired = false;
  return PQC_OK;
}

size_t req_bytes(const Pending& p) { return in_bytes(p.set, p.r.op) + out_bytes(p.set, p.r.op); }

// slot key-arena use of one request in plain mode (its gathered key)
size_t key_bytes(const Pending& p) {
  if (p.set.alg == PQC_ALG_FRODO976) return p.r.op == 1 ? 15632 + 8 : 31296;
  return p.r.op == 1 ? p.set.pk : p.set.sk;
}

// Complete a part of the slot's requests (kAll, kKyberOnly at kyber_done, kFrodoOnly
// after the Kyber ones completed).  With status PQC_OK their outputs go to the
// callers' buffers first, copied by the pool without holding the lock.
enum RetirePart { kAll, kKyberOnly, kFrodoOnly };
void retire_part(pqc_server* v, Slot& sl, std::unique_lock<std::mutex>& lk, int status, RetirePart part) {
  auto mine = [&](const Pending& p) {
    const bool frodo = p.r.alg == PQC_ALG_FRODO976;
    return part == kAll || (part == kKyberOnly ? !frodo : frodo);
  };
  if (status == PQC_OK) {
    const double r0 = v->now_us();
    lk.unlock();
    v->copy_pool.parallel_for(sl.reqs.size(), kCopyPer, [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        const Pending& p = sl.reqs[i];
        if (!mine(p)) continue;
        memcpy(p.r.out, sl.h_out + sl.out_a[i], sl.len_a[i]);
        if (sl.len_b[i]) memcpy(p.r.out + sl.len_a[i], sl.h_out + sl.out_b[i], sl.len_b[i]);
      }
    });
    lk.lock();
    v->retire_us += v->now_us() - r0;
  }
  const double t = v->now_us();
  size_t done = 0;
  for (const Pending& p : sl.reqs) {
    if (!mine(p)) continue;
    v->completions.push_back(pqc_server_completion{p.r.tag, status, (uint32_t)sl.reqs.size(), p.t_submit, t});
    ++done;
  }
  v->n_completed += done;
  v->cv_done.notify_all();
}

void worker_loop(pqc_server* v) {
  (void)hipSetDevice(v->device);
  const bool plain = (v->cfg.flags & PQC_SERVER_PLAIN) != 0;
  std::unique_lock<std::mutex> lk(v->mu);
  while (!v->stop) {
    // 1. move submissions in (with split pools, Frodo requests queue apart)
    while (!v->submitted.empty()) {
      const Pending& p = v->submitted.front();
      (v->split && p.r.alg == PQC_ALG_FRODO976 ? v->pending_f : v->pending).push_back(p);
      v->submitted.pop_front();
    }
    bool progress = false;
    // 2. retire finished slots
    for (auto& sl : v->slots) {
      if (!sl.busy) continue;
      if (sl.kyber_pending) {
        const hipError_t qk = hipEventQuery(sl.kyber_done);
        if (qk != hipErrorNotReady) {
          const int kst = qk != hipSuccess ? PQC_ERR_DEVICE : sl.h_status[2] ? PQC_ERR_DEVICE : PQC_OK;
          retire_part(v, sl, lk, kst, kKyberOnly);
          sl.kyber_pending = false;
          sl.kyber_retired = true;
          progress = true;
        }
      }
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      // a status bit raised on the device (a rejected descriptor, a sampler at its
      // block cap) fails the whole batch: its outputs are not trusted
      const int status = q != hipSuccess ? PQC_ERR_DEVICE : (sl.h_status[0] | sl.h_status[1]) ? PQC_ERR_DEVICE : PQC_OK;
      if (status == PQC_OK) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, sl.start, sl.done) == hipSuccess) v->device_us += 1e3 * ms;
      }
      // every request not completed at kyber_done (all of them otherwise; also when
      // kyber_done was still pending at its query above and done is reached now)
      retire_part(v, sl, lk, status, sl.kyber_retired ? kFrodoOnly : kAll);
      sl.kyber_pending = sl.kyber_retired = false;
      v->inflight_bytes -= sl.bytes;
      sl.reqs.clear();
      sl.bytes = 0;
      sl.busy = false;
      progress = true;
      v->cv_done.notify_all();
    }
    // 3. launch into free slots; a class whose next batch is still filling (or
    // does not fit) is held for this pass
    bool hold[3] = {false, false, false};
    for (auto& sl : v->slots) {
      std::deque<Pending>& q = sl.cls == kFrodoClass ? v->pending_f : v->pending;
      if (sl.busy || q.empty() || hold[sl.cls]) continue;
      const double waited = v->now_us() - q.front().t_submit;
      bool any_busy = false, class_busy = false;
      for (auto& o : v->slots) {
        any_busy |= o.busy;
        class_busy |= o.busy && o.cls == sl.cls;
      }
      if (class_busy && q.size() < v->cfg.max_batch && waited < v->cfg.max_wait_us) {
        hold[sl.cls] = true;
        continue;
      }
      size_t in_sz = 0, out_sz = 0, key_sz = 0;
      while (!q.empty() && sl.reqs.size() < v->cfg.max_batch) {
        const Pending& p = q.front();
        const size_t ib = align8(in_bytes(p.set, p.r.op)) + 8, ob = align8(out_bytes(p.set, p.r.op)) + 16;
        const size_t rb = req_bytes(p);
        if (in_sz + ib > v->arena || out_sz + ob > v->arena) break;
        if (plain && key_sz + key_bytes(p) > v->key_arena) break;
        if (v->cfg.hbm_budget && v->inflight_bytes + sl.bytes + rb > v->cfg.hbm_budget && !sl.reqs.empty()) break;
        if (v->cfg.hbm_budget && v->inflight_bytes + rb > v->cfg.hbm_budget && any_busy) break;
        in_sz += ib;
        out_sz += ob;
        key_sz += key_bytes(p);
        sl.bytes += rb;
        sl.reqs.push_back(p);
        q.pop_front();
      }
      if (sl.reqs.empty()) {
        hold[sl.cls] = true;
        continue;
      }
      v->inflight_bytes += sl.bytes;
      lk.unlock();
      const double l0 = v->now_us();
      const int rc = launch(v, sl);
      const double l1 = v->now_us();
      lk.lock();
      v->pack_us += l1 - l0;
      v->n_batches++;
      v->n_batched_reqs += sl.reqs.size();
      if (rc) {  // complete the batch with the error
        const double t = v->now_us();
        for (auto& p : sl.reqs)
          v->completions.push_back(pqc_server_completion{p.r.tag, rc, (uint32_t)sl.reqs.size(), p.t_submit, t});
        v->n_completed += sl.reqs.size();
        v->inflight_bytes -= sl.bytes;
        sl.reqs.clear();
        sl.bytes = 0;
        v->cv_done.notify_all();
      } else {
        sl.busy = true;
      }
      progress = true;
    }
    if (!progress) {
      bool any_busy = false;
      for (auto& o : v->slots) any_busy |= o.busy;
      // poll the GPU every few microseconds while batches run, else sleep until work
      if (any_busy || !v->pending.empty() || !v->pending_f.empty())
        v->cv_work.wait_for(lk, std::chrono::microseconds(5));
      else
        v->cv_work.wait_for(lk, std::chrono::milliseconds(2));
    }
  }
}

}  // namespace

extern "C" {

int pqc_server_create(pqc_server** out, const pqc_server_config* cfg) {
  if (!out || !cfg) return PQC_ERR_INVALID_ARG;
  *out = nullptr;
  const unsigned kyb = PQC_ALG_MASK(PQC_ALG_KYBER512) | PQC_ALG_MASK(PQC_ALG_KYBER768) | PQC_ALG_MASK(PQC_ALG_KYBER1024);
  if (!cfg->alg_mask || (cfg->alg_mask & ~(kyb | PQC_ALG_MASK(PQC_ALG_FRODO976))) || cfg->keys_per_alg == 0 ||
      cfg->max_batch == 0 || cfg->slots == 0 || cfg->slots > 8 || cfg->arena_bytes < (1u << 20))
    return PQC_ERR_INVALID_ARG;
  auto* v = new (std::nothrow) pqc_server;
  if (!v) return PQC_ERR_NO_MEMORY;
  v->cfg = *cfg;
  v->t0 = clk::now();
  v->kyber_mask = cfg->alg_mask & kyb;
  v->arena = cfg->arena_bytes;
  v->key_arena = cfg->arena_bytes;
  if (hipGetDevice(&v->device) != hipSuccess) {
    delete v;
    return PQC_ERR_NO_DEVICE;
  }
  int rc = PQC_OK;
  const size_t K = cfg->keys_per_alg;
  hipStream_t s0 = nullptr;
  if (hipStreamCreate(&s0) != hipSuccess) {
    delete v;
    return PQC_ERR_DEVICE;
  }
  // ---- key store (key pair i of set a from coins DRBG(seed, 0x6B6579 + a, i))
  {
    unsigned gen_mask = cfg->alg_mask;
    pqc_ctx* gctx = nullptr;
    rc = pqc_ctx_create(&gctx, gen_mask, K, 0);
    std::vector<uint8_t> coins(K * 64);
    for (int a = PQC_ALG_KYBER512; !rc && a <= PQC_ALG_KYBER1024; ++a) {
      Set s;
      set_of(a, &s);
      v->pk_base[a - 1] = v->pk_bytes;
      v->sk_base[a - 1] = v->sk_bytes;
      if (!(v->kyber_mask & PQC_ALG_MASK(a))) continue;
      v->pk_bytes += K * s.pk;
      v->sk_bytes += K * s.sk;
    }
    if (!rc && v->kyber_mask) {
      if (hipMalloc(&v->d_pk, v->pk_bytes) != hipSuccess || hipMalloc(&v->d_sk, v->sk_bytes) != hipSuccess)
        rc = PQC_ERR_NO_MEMORY;
    }
    for (int a = PQC_ALG_KYBER512; !rc && a <= PQC_ALG_FRODO976; ++a) {
      if (a == PQC_ALG_NEWHOPE1024_CPA || !(cfg->alg_mask & PQC_ALG_MASK(a))) continue;
      Set s;
      set_of(a, &s);
      for (size_t i = 0; i < K; ++i) pqc::host_drbg(&coins[i * 64], 64, cfg->key_seed, 0x6B6579 + a, i);
      v->h_pk[a].resize(K * s.pk);
      v->h_sk[a].resize(K * s.sk);
      if (a == PQC_ALG_FRODO976) {
        if (hipMalloc(&v->d_fsk, K * s.sk) != hipSuccess || hipMalloc(&v->d_fa, K * 976 * 976 * 2) != hipSuccess) {
          rc = PQC_ERR_NO_MEMORY;
          break;
        }
        rc = crypto_kem_keypair_derand_batch(gctx, a, K, v->h_pk[a].data(), v->h_sk[a].data(), coins.data(),
                                             PQC_F_HOST_PTRS, s0);
        if (!rc && hipMemcpy(v->d_fsk, v->h_sk[a].data(), K * s.sk, hipMemcpyHostToDevice) != hipSuccess)
          rc = PQC_ERR_DEVICE;
        if (!rc) rc = pqc::frodo_expand_a(K, v->d_fa, v->d_fsk, s0);
      } else {
        rc = crypto_kem_keypair_derand_batch(gctx, a, K, v->h_pk[a].data(), v->h_sk[a].data(), coins.data(),
                                             PQC_F_HOST_PTRS, s0);
        if (!rc && (hipMemcpy(v->d_pk + v->pk_base[a - 1], v->h_pk[a].data(), K * s.pk, hipMemcpyHostToDevice) !=
                        hipSuccess ||
                    hipMemcpy(v->d_sk + v->sk_base[a - 1], v->h_sk[a].data(), K * s.sk, hipMemcpyHostToDevice) !=
                        hipSuccess))
          rc = PQC_ERR_DEVICE;
      }
    }
    if (!rc && hipStreamSynchronize(s0) != hipSuccess) rc = PQC_ERR_DEVICE;
    if (gctx) pqc_ctx_destroy(gctx);
  }
  (void)hipStreamDestroy(s0);
  // ---- slots: split pools when Kyber and Frodo share a graph-mode server (half the
  // slots, at least one, take the Frodo requests).  Eager mode keeps one pool: its
  // many small Kyber launches compete with the Frodo pool's on the scheduler
  // thread, which measured a less stable p99 than one pool whose batches complete
  // their Kyber part early.
  v->slots.resize(cfg->slots);
  v->split = !(cfg->flags & PQC_SERVER_PLAIN) && (cfg->flags & PQC_SERVER_GRAPH) && v->kyber_mask &&
             (cfg->alg_mask & PQC_ALG_MASK(PQC_ALG_FRODO976)) && cfg->slots >= 2;
  for (size_t i = 0; i < v->slots.size(); ++i)
    v->slots[i].cls = !v->split ? kAnyClass : i < v->slots.size() / 2 ? kFrodoClass : kKyberClass;
  for (auto& sl : v->slots) {
    if (rc) break;
    rc = pqc_ctx_create(&sl.ctx, cfg->alg_mask, cfg->max_batch, 0);
    if (!rc && !(cfg->flags & PQC_SERVER_PLAIN) && (cfg->alg_mask & PQC_ALG_MASK(PQC_ALG_FRODO976)))
      rc = pqc_ctx_create(&sl.ctx_fd, PQC_ALG_MASK(PQC_ALG_FRODO976), cfg->max_batch, 0);
    if (rc) break;
    if (hipStreamCreate(&sl.s) != hipSuccess || hipStreamCreate(&sl.s2) != hipSuccess ||
        hipStreamCreate(&sl.s3) != hipSuccess || hipEventCreate(&sl.done) != hipSuccess ||
        hipEventCreate(&sl.start) != hipSuccess ||
        hipEventCreateWithFlags(&sl.in_ready, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&sl.frodo_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&sl.fd_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&sl.kyber_done, hipEventDisableTiming) != hipSuccess)
      rc = PQC_ERR_DEVICE;
    const size_t nb = cfg->max_batch;
    if (!rc && (hipHostMalloc(&sl.h_in, v->arena) != hipSuccess || hipHostMalloc(&sl.h_out, v->arena) != hipSuccess ||
                hipHostMalloc(&sl.h_desc, nb * sizeof(pqc_kem_req)) != hipSuccess ||
                hipHostMalloc(&sl.h_kid, nb * 4) != hipSuccess || hipHostMalloc(&sl.h_status, 16) != hipSuccess))
      rc = PQC_ERR_NO_MEMORY;
    if (!rc && (hipMalloc(&sl.d_in, v->arena) != hipSuccess || hipMalloc(&sl.d_out, v->arena) != hipSuccess ||
                hipMalloc(&sl.d_desc, nb * sizeof(pqc_kem_req)) != hipSuccess ||
                hipMalloc(&sl.d_kid, nb * 4) != hipSuccess))
      rc = PQC_ERR_NO_MEMORY;
    if (!rc && (cfg->flags & PQC_SERVER_PLAIN) && hipMalloc(&sl.d_keys, v->key_arena) != hipSuccess)
      rc = PQC_ERR_NO_MEMORY;
    if (!rc) (void)pqc_ctx_put_tuning(sl.ctx, &cfg->tuning);  // an all-zero tuning is rejected: defaults stay
    if (!rc && sl.ctx_fd) (void)pqc_ctx_put_tuning(sl.ctx_fd, &cfg->tuning);
  }
  if (rc) {
    free_server(v);
    return rc;
  }
  v->worker = std::thread(worker_loop, v);
  *out = v;
  return PQC_OK;
}

void pqc_server_destroy(pqc_server* v) {
  if (!v) return;
  {
    std::lock_guard<std::mutex> lk(v->mu);
    v->stop = true;
  }
  v->cv_work.notify_all();
  if (v->worker.joinable()) v->worker.join();
  (void)hipDeviceSynchronize();
  free_server(v);
}

int pqc_server_key(pqc_server* v, int alg, uint32_t key_id, uint8_t* pk, uint8_t* sk) {
  Set s;
  if (!v || !set_of(alg, &s) || !(v->cfg.alg_mask & PQC_ALG_MASK(alg))) return PQC_ERR_INVALID_ARG;
  if (key_id >= v->cfg.keys_per_alg) return PQC_ERR_INVALID_ARG;
  if (pk) memcpy(pk, v->h_pk[alg].data() + (size_t)key_id * s.pk, s.pk);
  if (sk) memcpy(sk, v->h_sk[alg].data() + (size_t)key_id * s.sk, s.sk);
  return PQC_OK;
}

int pqc_server_submit(pqc_server* v, const pqc_server_request* reqs, size_t n) {
  if (!v || (!reqs && n)) return PQC_ERR_INVALID_ARG;
  const double t = v->now_us();
  for (size_t i = 0; i < n; ++i) {
    Set s;
    const pqc_server_request& r = reqs[i];
    if (!set_of(r.alg, &s) || !(v->cfg.alg_mask & PQC_ALG_MASK(r.alg)) || (r.op != 1 && r.op != 2) ||
        r.key_id >= v->cfg.keys_per_alg || !r.in || !r.out)
      return PQC_ERR_INVALID_ARG;
    if (v->cfg.hbm_budget && in_bytes(s, r.op) + out_bytes(s, r.op) > v->cfg.hbm_budget) return PQC_ERR_NO_MEMORY;
  }
  {
    std::lock_guard<std::mutex> lk(v->mu);
    for (size_t i = 0; i < n; ++i) {
      Pending p{reqs[i], {}, t};
      set_of(reqs[i].alg, &p.set);
      v->submitted.push_back(p);
    }
    v->n_submitted += n;
  }
  v->cv_work.notify_one();
  return PQC_OK;
}

int pqc_server_poll(pqc_server* v, pqc_server_completion* out, size_t cap, size_t* got, uint32_t wait_us) {
  if (!v || !got || (!out && cap)) return PQC_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(v->mu);
  if (v->completions.empty() && wait_us)
    v->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !v->completions.empty(); });
  size_t k = 0;
  while (k < cap && !v->completions.empty()) {
    out[k++] = v->completions.front();
    v->completions.pop_front();
  }
  *got = k;
  return PQC_OK;
}

int pqc_server_drain(pqc_server* v) {
  if (!v) return PQC_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(v->mu);
  // completions keep coming while the GPU works; 30 s without one is a failure
  uint64_t seen = v->n_completed;
  while (v->n_completed != v->n_submitted) {
    if (!v->cv_done.wait_for(lk, std::chrono::seconds(30), [&] { return v->n_completed != seen; }))
      return PQC_ERR_DEVICE;
    seen = v->n_completed;
  }
  return PQC_OK;
}

int pqc_server_get_stats(pqc_server* v, pqc_server_stats* st) {
  if (!v || !st) return PQC_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(v->mu);
  st->submitted = v->n_submitted;
  st->completed = v->n_completed;
  st->batches = v->n_batches;
  st->mean_batch = v->n_batches ? (double)v->n_batched_reqs / (double)v->n_batches : 0.0;
  st->inflight_bytes = v->inflight_bytes;
  const double nb = v->n_batches ? (double)v->n_batches : 1.0;
  st->pack_us = v->pack_us / nb;
  st->device_us = v->device_us / nb;
  st->retire_us = v->retire_us / nb;
  return PQC_OK;
}

double pqc_server_now_us(pqc_server* v) { return v ? v->now_us() : 0.0; }

// Replay an arrival trace from the calling thread: request i is submitted at
// arrival_us[i] (server clock, offset to "now" at the call), completions are
// drained as they come.  latency_us[i] (optional) = completion - arrival.
}  // extern "C"

namespace {
// An unfinished trace returns only once the requests it submitted are done
// (their outputs land in the caller's buffers, which must not be freed first)
// or the GPU made no progress for 30 s; their completions are discarded so
// that later polls do not see trace indices as tags.
int abandon_trace(pqc_server* v, int rc) {
  (void)pqc_server_drain(v);
  std::vector<pqc_server_completion> comp(4096);
  size_t got = 0;
  do {
    got = 0;
    pqc_server_poll(v, comp.data(), comp.size(), &got, 0);
  } while (got);
  return rc;
}
}  // namespace

extern "C" {

int pqc_server_run_trace(pqc_server* v, size_t n, const double* arrival_us, const pqc_server_request* reqs,
                         double* latency_us, pqc_trace_stats* out) {
  if (!v || !out || (n && (!arrival_us || !reqs))) return PQC_ERR_INVALID_ARG;
  std::vector<double> lat(n, 0.0);
  std::vector<pqc_server_request> rq(reqs, reqs + n);
  for (size_t i = 0; i < n; ++i) rq[i].tag = i;  // the trace index
  const double base = v->now_us() + 50.0;
  size_t next = 0, done = 0;
  uint64_t errors = 0;
  std::vector<pqc_server_completion> comp(4096);
  double last = base, progress = v->now_us();
  while (done < n) {
    if (v->now_us() - progress > 30e6) return abandon_trace(v, PQC_ERR_DEVICE);  // 30 s without a completion
    const double now = v->now_us();
    size_t end = next;
    while (end < n && base + arrival_us[end] <= now) ++end;
    if (end > next) {
      // one submission for everything that has arrived, each stamped now
      int rc = pqc_server_submit(v, rq.data() + next, end - next);
      if (rc) return abandon_trace(v, rc);
      next = end;
    }
    size_t got = 0;
    pqc_server_poll(v, comp.data(), comp.size(), &got, next < n ? 0 : 200);
    if (got) progress = v->now_us();
    for (size_t k = 0; k < got; ++k) {
      const size_t i = (size_t)comp[k].tag;
      lat[i] = comp[k].done_us - (base + arrival_us[i]);
      errors += comp[k].status != PQC_OK;
      last = std::max(last, comp[k].done_us);
    }
    done += got;
    if (!got && next < n) {
      // spin toward the next arrival without burning the scheduler's core
      const double wait = base + arrival_us[next] - v->now_us();
      if (wait > 20) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait - 10)));
    }
  }
  std::vector<double> sorted = lat;
  std::sort(sorted.begin(), sorted.end());
  auto pct = [&](double q) { return n ? sorted[std::min(n - 1, (size_t)(q * (double)(n - 1) + 0.5))] : 0.0; };
  out->completed = n;
  out->errors = errors;
  out->duration_s = (last - base) / 1e6;
  out->ops_per_s = out->duration_s > 0 ? (double)n / out->duration_s : 0.0;
  out->p50_us = pct(0.50);
  out->p90_us = pct(0.90);
  out->p99_us = pct(0.99);
  out->p999_us = pct(0.999);
  out->max_us = n ? sorted.back() : 0.0;
  pqc_server_stats st;
  pqc_server_get_stats(v, &st);
  out->mean_batch = st.mean_batch;
  if (latency_us) std::copy(lat.begin(), lat.end(), latency_us);
  return PQC_OK;
}

}  // extern "C"
