// Autotuner for the launch parameters in pqc_tuning (replaces the reference's
// per-GPU block-size tables and its ANALYSIS_MODE sweep, which timed 10 block
// sizes per kernel and printed the best: Kyber_GPU_Batched/indcpa.cu:27-70).
//
// A tuning request names a workload family (which knobs apply), its batch size
// and parameter set, and a callback that runs the workload once on a stream.
// Each candidate configuration is timed with HIP events (median of `reps`
// runs after one warm-up) and kept as a sample (features, log ms) in the
// context.  Without the cost model the candidates are measured in a fixed order
// (current configuration first) up to the budget; with it, the model trained on
// every sample so far ranks all candidates and only the best-ranked `budget` are
// measured, after which the model is refitted with the new samples.
#include <algorithm>
#include <cmath>
#include <new>
#include <numeric>
#include <vector>

#include "internal.h"

namespace pqc {
namespace tuner {

constexpr int kFeatures = 14;

struct Samples {
  std::vector<float> x, y;  // kFeatures per sample; log(ms)
  pqc_costmodel* model = nullptr;
};

static float log2f_safe(double v) { return (float)std::log2(v > 1.0 ? v : 1.0); }

// The mixed family's tile layouts: 32-request tiles for both operations (the
// tile order does not apply), or 64-request tiles in tile order 0, 1 or 3.
constexpr int kMixedOrders[3] = {0, 1, 3};
static int mixed_layout(const pqc_tuning& t) {
  if (t.fused_enc_tile != 64 && t.fused_dec_tile != 64) return 0;
  for (int i = 0; i < 3; ++i)
    if (t.mixed_order == kMixedOrders[i]) return 1 + i;
  return 3;  // order 2 is not a candidate: nearest slot
}

// one-hot family (5), K / 4, op / 2, log2(n) / 17, log2(block) / 10, waves / 4,
// mixed tile layout one-hot (4); for the Frodo family, slot 10 is the stream's
// Keccak lane layout at this n (1 = one state per lane)
static void features(const pqc_tune_request& r, const pqc_tuning& t, float* f) {
  for (int i = 0; i < kFeatures; ++i) f[i] = 0.f;
  if (r.family >= 0 && r.family < 5) f[r.family] = 1.f;
  f[5] = (float)r.k / 4.f;
  f[6] = (float)r.op / 2.f;
  f[7] = log2f_safe((double)r.n) / 17.f;
  int block = 0;
  switch (r.family) {
    case PQC_TUNE_PLAIN_KYBER: block = t.plain_block; break;
    case PQC_TUNE_NEWHOPE: block = t.newhope_block; break;
    case PQC_TUNE_FRODO: block = t.frodo_stream_block; break;
    default: break;
  }
  f[8] = block ? log2f_safe(block) / 10.f : 0.f;
  const bool fused = r.family == PQC_TUNE_FUSED_KYBER || r.family == PQC_TUNE_MIXED_KYBER;
  f[9] = fused ? (float)t.fused_waves_per_cu / 4.f : 0.f;
  if (r.family == PQC_TUNE_MIXED_KYBER) f[10 + mixed_layout(t)] = 1.f;
  if (r.family == PQC_TUNE_FRODO) f[10] = frodo_lanes(t, r.n) ? 1.f : 0.f;
}

static std::vector<pqc_tuning> candidates(const pqc_tune_request& r, const pqc_tuning& base) {
  std::vector<pqc_tuning> out;
  auto add = [&](pqc_tuning t) { out.push_back(t); };
  switch (r.family) {
    case PQC_TUNE_PLAIN_KYBER:
      for (int b = 64; b <= 1024; b += 64) {
        pqc_tuning t = base;
        t.plain_block = b;
        add(t);
      }
      break;
    case PQC_TUNE_NEWHOPE:
      for (int b = 64; b <= 1024; b += 64) {
        pqc_tuning t = base;
        t.newhope_block = b;
        add(t);
      }
      break;
    case PQC_TUNE_FUSED_KYBER:
      for (int w = 1; w <= 4; ++w) {
        pqc_tuning t = base;
        t.fused_waves_per_cu = w;
        add(t);
      }
      break;
    case PQC_TUNE_MIXED_KYBER:
      for (int w = 1; w <= 4; ++w)
        for (int l = 0; l < 4; ++l) {
          pqc_tuning t = base;
          t.fused_waves_per_cu = w;
          t.mixed_order = l == 0 ? 0 : kMixedOrders[l - 1];
          t.fused_enc_tile = t.fused_dec_tile = l == 0 ? 32 : 64;
          add(t);
        }
      break;
    case PQC_TUNE_FRODO:
      // block size x Keccak lane layout at this n (lane pairs or one state per lane);
      // the layout in use keeps its threshold, so the configuration in use is a candidate
      for (int lanes = 0; lanes < 2; ++lanes)
        for (int b : {64, 128, 256}) {
          pqc_tuning t = base;
          t.frodo_stream_block = b;
          if (frodo_lanes(base, r.n) != (lanes != 0)) t.frodo_lane_min = lanes ? 1 : -1;
          add(t);
        }
      break;
    default: break;
  }
  return out;
}

// 0 and 32 both mean 32-request tiles
static int tile_of(int v) { return v == 64 ? 64 : 32; }
// 0 means the library's threshold
static int lane_min_of(int v) { return v == 0 ? kFrodoLaneMinDefault : v; }

static bool same(const pqc_tuning& a, const pqc_tuning& b) {
  return a.plain_block == b.plain_block && a.newhope_block == b.newhope_block &&
         a.fused_waves_per_cu == b.fused_waves_per_cu && a.mixed_order == b.mixed_order &&
         a.frodo_stream_block == b.frodo_stream_block && tile_of(a.fused_enc_tile) == tile_of(b.fused_enc_tile) &&
         tile_of(a.fused_dec_tile) == tile_of(b.fused_dec_tile) && tile_of(a.fused_key_tile) == tile_of(b.fused_key_tile) &&
         lane_min_of(a.frodo_lane_min) == lane_min_of(b.frodo_lane_min);
}

// median of reps timed runs (one untimed warm-up first)
static int measure(pqc_ctx* c, const pqc_tuning& t, pqc_workload_fn run, void* user, hipStream_t s, int reps,
                   float* ms) {
  c->tune = t;
  int rc = run(user, (void*)s);
  if (rc) return rc;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return PQC_ERR_DEVICE;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return PQC_ERR_DEVICE;
  }
  std::vector<float> v;
  for (int i = 0; i < reps && rc == PQC_OK; ++i) {
    if (hipEventRecord(a, s) != hipSuccess) rc = PQC_ERR_DEVICE;
    if (!rc) rc = run(user, (void*)s);
    if (!rc && hipEventRecord(b, s) != hipSuccess) rc = PQC_ERR_DEVICE;
    if (!rc && hipEventSynchronize(b) != hipSuccess) rc = PQC_ERR_DEVICE;
    float e = 0.f;
    if (!rc && hipEventElapsedTime(&e, a, b) != hipSuccess) rc = PQC_ERR_DEVICE;
    v.push_back(e);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  std::sort(v.begin(), v.end());
  *ms = v[v.size() / 2];
  return PQC_OK;
}

}  // namespace tuner
}  // namespace pqc

using namespace pqc::tuner;

static Samples* samples_of(pqc_ctx* c) {
  if (!c->tuner) c->tuner = new (std::nothrow) Samples;
  return (Samples*)c->tuner;
}

void pqc::pqc_tuner_free(pqc_ctx* c) {
  auto* S = (Samples*)c->tuner;
  if (!S) return;
  if (S->model) pqc_costmodel_destroy(S->model);
  delete S;
  c->tuner = nullptr;
}

extern "C" {

int pqc_tune_features(const pqc_tune_request* r, const pqc_tuning* t, float* out, int count) {
  if (!r || !t || !out || count < kFeatures) return PQC_ERR_INVALID_ARG;
  features(*r, *t, out);
  return kFeatures;
}

int pqc_tune_samples(pqc_ctx* c, float* x, float* y, int cap) {
  if (!c) return PQC_ERR_INVALID_ARG;
  Samples* S = samples_of(c);
  if (!S) return PQC_ERR_NO_MEMORY;
  const int n = (int)S->y.size();
  if (x && y) {
    const int m = std::min(n, cap);
    std::copy(S->x.begin(), S->x.begin() + (size_t)m * kFeatures, x);
    std::copy(S->y.begin(), S->y.begin() + m, y);
  }
  return n;
}

pqc_costmodel* pqc_ctx_costmodel(pqc_ctx* c) {
  if (!c) return nullptr;
  Samples* S = samples_of(c);
  if (!S) return nullptr;
  if (!S->model && pqc_costmodel_create(&S->model, kFeatures, 8192, 0x5EED) != PQC_OK) S->model = nullptr;
  return S->model;
}

int pqc_autotune(pqc_ctx* c, const pqc_tune_request* r, pqc_workload_fn run, void* user, void* stream,
                 pqc_tune_result* out) {
  if (!c || !r || !run || !out) return PQC_ERR_INVALID_ARG;
  if (r->family < 0 || r->family > 4 || r->reps < 1 || r->budget < 0) return PQC_ERR_INVALID_ARG;
  Samples* S = samples_of(c);
  if (!S) return PQC_ERR_NO_MEMORY;
  hipStream_t s = (hipStream_t)stream;
  const pqc_tuning start = c->tune;
  std::vector<pqc_tuning> cand = candidates(*r, start);
  if (cand.empty()) return PQC_ERR_INVALID_ARG;
  std::vector<float> pred(cand.size(), 0.f);
  std::vector<size_t> order(cand.size());
  std::iota(order.begin(), order.end(), 0);
  const bool model = r->use_model && pqc_ctx_costmodel(c) && pqc_costmodel_trained_steps(S->model) > 0;
  if (model) {
    std::vector<float> fx(cand.size() * kFeatures);
    for (size_t i = 0; i < cand.size(); ++i) features(*r, cand[i], &fx[i * kFeatures]);
    int rc = pqc_costmodel_predict(S->model, (int)cand.size(), fx.data(), pred.data());
    if (rc) return rc;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return pred[a] < pred[b]; });
    // with two or more measurements, the configuration in use is one of them
    // (the last), so a wrong ranking never installs something measured slower
    const size_t b = r->budget == 0 ? cand.size() : std::min<size_t>(r->budget, cand.size());
    if (b >= 2)
      for (size_t i = 0; i < cand.size(); ++i)
        if (same(cand[order[i]], start)) {
          if (i >= b) std::rotate(order.begin() + b - 1, order.begin() + i, order.begin() + i + 1);
          break;
        }
  } else {
    // the configuration in use first, the rest in enumeration order
    for (size_t i = 0; i < cand.size(); ++i)
      if (same(cand[i], start)) {
        std::rotate(order.begin(), order.begin() + i, order.begin() + i + 1);
        break;
      }
  }
  const size_t budget = r->budget == 0 ? cand.size() : std::min<size_t>(r->budget, cand.size());
  float best_ms = 0.f;
  size_t best = order[0];
  int rc = PQC_OK;
  for (size_t j = 0; j < budget; ++j) {
    float ms = 0.f;
    rc = measure(c, cand[order[j]], run, user, s, r->reps, &ms);
    if (rc) break;
    float f[kFeatures];
    features(*r, cand[order[j]], f);
    S->x.insert(S->x.end(), f, f + kFeatures);
    S->y.push_back(std::log(std::max(ms, 1e-4f)));
    if (j == 0 || ms < best_ms) {
      best_ms = ms;
      best = order[j];
    }
  }
  if (rc) {
    c->tune = start;
    return rc;
  }
  c->tune = cand[best];
  out->best = cand[best];
  out->best_ms = best_ms;
  out->measured = (int32_t)budget;
  out->predicted_ms = model ? std::exp(pred[best]) : 0.f;
  // refit the model on everything measured so far (warm start)
  if (r->use_model && pqc_ctx_costmodel(c)) {
    const int n = (int)std::min<size_t>(S->y.size(), 8192);
    const size_t first = S->y.size() - n;
    float loss = 0.f;
    const bool warm = pqc_costmodel_trained_steps(S->model) > 0;
    rc = pqc_costmodel_fit(S->model, n, &S->x[first * kFeatures], &S->y[first], warm ? 200 : 600, 3e-3f, &loss);
    if (rc) return rc;
  }
  return PQC_OK;
}

}  // extern "C"
