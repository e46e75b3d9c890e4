// Small kernels of the serving loop: the record gather of the plain-path
// comparison (copies the key of each request out of a key store into a dense
// per-request array), and a word fill.
#include "internal.h"

namespace pqc {

__global__ void k_gather(int n, int words, uint64_t* __restrict__ dst, const uint8_t* __restrict__ src,
                         size_t src_stride, const uint32_t* __restrict__ idx) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * words) return;
  const int r = (int)(t / words), w = (int)(t % words);
  dst[t] = reinterpret_cast<const uint64_t*>(src + (size_t)idx[r] * src_stride)[w];
}

int gather_records(size_t n, uint8_t* dst, size_t bytes, const uint8_t* src, size_t src_stride, const uint32_t* idx,
                   hipStream_t s) {
  if (n == 0) return PQC_OK;
  if (bytes % 8 || src_stride % 8) return PQC_ERR_INVALID_ARG;
  const size_t total = n * (bytes / 8);
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, (int)n, (int)(bytes / 8),
                     reinterpret_cast<uint64_t*>(dst), src, src_stride, idx);
  return hip_err(hipGetLastError());
}

// Zero `n` 32-bit words with a kernel (a memset node inside a captured graph is
// avoided: fill kernels replay reliably).
__global__ void k_zero_words(int n, uint32_t* __restrict__ p) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0u;
}

int zero_words(uint32_t* p, int n, hipStream_t s) {
  if (n <= 0) return PQC_OK;
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, p);
  return hip_err(hipGetLastError());
}

}  // namespace pqc
