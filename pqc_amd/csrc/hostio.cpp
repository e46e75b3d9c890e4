// Host-buffer transfers for batched calls made with PQC_F_HOST_PTRS.
//
// Pageable host memory cannot be read by the DMA engines directly, so a large
// transfer is cut into chunks that a small pool of host threads copies into
// one of two pinned buffers while the DMA engine moves the previous chunk
// (double buffering).  Memory that pqc_host_alloc returned is already pinned
// and is copied directly, asynchronously on the caller's stream.  Transfers
// below kDirectBytes (single-instance calls) go straight to hipMemcpyAsync.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

#include "internal.h"
#include "threadpool.h"

namespace pqc {
namespace {

constexpr size_t kChunkBytes = 16u << 20;   // bytes per pinned staging buffer
constexpr size_t kDirectBytes = 256u << 10; // smaller transfers skip the staging
constexpr int kPoolThreads = 8;             // host threads per context, the caller included

std::mutex g_pin_mu;
std::map<uintptr_t, size_t> g_pinned;  // start -> bytes of every live pqc_host_alloc block

struct HostIo {
  CopyPool pool{kPoolThreads};
  uint8_t* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool busy[2] = {false, false};  // an event recorded after the buffer's last DMA
  int next = 0;
};

int hostio_get(pqc_ctx* c, HostIo** out) {
  if (!c->hostio) {
    HostIo* h = new HostIo;
    for (int b = 0; b < 2; ++b) {
      if (hipHostMalloc((void**)&h->pin[b], kChunkBytes, hipHostMallocDefault) != hipSuccess ||
          hipEventCreateWithFlags(&h->ev[b], hipEventDisableTiming) != hipSuccess) {
        c->hostio = h;
        hostio_free(c);
        return PQC_ERR_NO_MEMORY;
      }
    }
    c->hostio = h;
  }
  *out = (HostIo*)c->hostio;
  return PQC_OK;
}

// dst <- src over `bytes`, split across the pool's threads in 4 KiB-aligned slices.
void par_copy(HostIo* h, uint8_t* dst, const uint8_t* src, size_t bytes) {
  h->pool.run([&](int t, int n) {
    const size_t per = ((bytes + n - 1) / n + 4095) & ~(size_t)4095;
    const size_t b = std::min(bytes, per * t), e = std::min(bytes, b + per);
    if (b < e) memcpy(dst + b, src + b, e - b);
  });
}

// Waits for the buffer's previous DMA, if any.
int settle(HostIo* h, int b) {
  if (!h->busy[b]) return PQC_OK;
  h->busy[b] = false;
  return hip_err(hipEventSynchronize(h->ev[b]));
}

}  // namespace

bool host_is_pinned(const void* p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_pin_mu);
  const uintptr_t a = (uintptr_t)p;
  auto it = g_pinned.upper_bound(a);
  if (it == g_pinned.begin()) return false;
  --it;
  return a >= it->first && a + bytes <= it->first + it->second;
}

int host_to_device(pqc_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) return PQC_OK;
  if (bytes < kDirectBytes || host_is_pinned(src, bytes))
    return hip_err(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  HostIo* h;
  int rc = hostio_get(c, &h);
  for (size_t off = 0; !rc && off < bytes; off += kChunkBytes) {
    const size_t len = std::min(kChunkBytes, bytes - off);
    const int b = h->next;
    h->next ^= 1;
    if ((rc = settle(h, b))) break;
    par_copy(h, h->pin[b], (const uint8_t*)src + off, len);
    if (hipMemcpyAsync((uint8_t*)dst + off, h->pin[b], len, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipEventRecord(h->ev[b], s) != hipSuccess)
      rc = PQC_ERR_DEVICE;
    else
      h->busy[b] = true;
  }
  return rc;
}

int device_to_host(pqc_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) return PQC_OK;
  if (bytes < kDirectBytes || host_is_pinned(dst, bytes))
    return hip_err(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  HostIo* h;
  int rc = hostio_get(c, &h);
  int prev = -1;
  size_t prev_off = 0, prev_len = 0;
  for (size_t off = 0; !rc && off < bytes; off += kChunkBytes) {
    const size_t len = std::min(kChunkBytes, bytes - off);
    const int b = h->next;
    h->next ^= 1;
    if ((rc = settle(h, b))) break;
    if (hipMemcpyAsync(h->pin[b], (const uint8_t*)src + off, len, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipEventRecord(h->ev[b], s) != hipSuccess) {
      rc = PQC_ERR_DEVICE;
      break;
    }
    h->busy[b] = true;
    // the previous chunk has landed once its event has: unpack it while this one moves
    if (prev >= 0) {
      if ((rc = settle(h, prev))) break;
      par_copy(h, (uint8_t*)dst + prev_off, h->pin[prev], prev_len);
    }
    prev = b;
    prev_off = off;
    prev_len = len;
  }
  if (!rc && prev >= 0 && !(rc = settle(h, prev))) par_copy(h, (uint8_t*)dst + prev_off, h->pin[prev], prev_len);
  return rc;
}

void hostio_free(pqc_ctx* c) {
  HostIo* h = (HostIo*)c->hostio;
  if (!h) return;
  for (int b = 0; b < 2; ++b) {
    if (h->ev[b]) {
      if (h->busy[b]) (void)hipEventSynchronize(h->ev[b]);
      (void)hipEventDestroy(h->ev[b]);
    }
    if (h->pin[b]) (void)hipHostFree(h->pin[b]);
  }
  delete h;
  c->hostio = nullptr;
}

}  // namespace pqc

extern "C" {

int pqc_host_alloc(size_t bytes, void** out) {
  if (!out || !bytes) return PQC_ERR_INVALID_ARG;
  *out = nullptr;
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return PQC_ERR_NO_MEMORY;
  {
    std::lock_guard<std::mutex> lk(pqc::g_pin_mu);
    pqc::g_pinned[(uintptr_t)p] = bytes;
  }
  *out = p;
  return PQC_OK;
}

int pqc_host_free(void* p) {
  if (!p) return PQC_OK;
  {
    std::lock_guard<std::mutex> lk(pqc::g_pin_mu);
    auto it = pqc::g_pinned.find((uintptr_t)p);
    if (it == pqc::g_pinned.end()) return PQC_ERR_INVALID_ARG;
    pqc::g_pinned.erase(it);
  }
  return pqc::hip_err(hipHostFree(p));
}

}  // extern "C"
