// VALU issue-rate microbenchmark: confirms on the device the 32-bit vector peak
// that the VALU-bound kernels are priced against, and how it depends on the
// number of waves per SIMD.  Every wave of a full grid (1 to 8 waves per SIMD)
// runs kIters x 8 groups of eight independent instruction chains of one kind,
// written as the instructions themselves so that the count is exact; the
// elapsed time gives the sustained cycles per wave64 instruction per SIMD.
// Register-only work: no memory is touched besides one word per lane written
// at the end.
#include "internal.h"
#include "keccak.cuh"

namespace pqc {
namespace {

constexpr int kIters = 2000;
constexpr int kBlock = 64;         // one wave per workgroup: waves_per_simd x 4 workgroups per CU
constexpr int kInstrPerIter = 64;  // 8 groups x 8 chains

#define PQC_CHAINS8(ins, tail)                                                                          \
  asm volatile(ins " %0, %0, %8" tail "\n\t" ins " %1, %1, %8" tail "\n\t" ins " %2, %2, %8" tail "\n\t" \
               ins " %3, %3, %8" tail "\n\t" ins " %4, %4, %8" tail "\n\t" ins " %5, %5, %8" tail "\n\t" \
               ins " %6, %6, %8" tail "\n\t" ins " %7, %7, %8" tail                                      \
               : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)           \
               : "v"(b), "v"(c))

template <int OP>
__global__ __launch_bounds__(kBlock) void k_valu_rate(uint32_t* out, uint32_t seed) {
  uint32_t a0 = threadIdx.x ^ seed, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3;
  uint32_t a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
  const uint32_t b = seed | 1u, c = seed * 3u;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      if constexpr (OP == 0) PQC_CHAINS8("v_xor_b32", "");
      if constexpr (OP == 1) PQC_CHAINS8("v_add_u32", "");
      if constexpr (OP == 2) PQC_CHAINS8("v_alignbit_b32", ", 7");
      if constexpr (OP == 3) PQC_CHAINS8("v_mul_lo_u32", "");
      if constexpr (OP == 4) PQC_CHAINS8("v_bitop3_b32", ", %9 bitop3:0x96");
      if constexpr (OP == 5) PQC_CHAINS8("v_lshlrev_b32", "");
      if constexpr (OP == 6) PQC_CHAINS8("v_mad_u32_u24", ", %9");
      if constexpr (OP == 7) PQC_CHAINS8("v_sub_u32", "");
      if constexpr (OP == 8) PQC_CHAINS8("v_and_b32", "");
      if constexpr (OP == 9) PQC_CHAINS8("v_ashrrev_i32", "");
      if constexpr (OP == 10) PQC_CHAINS8("v_fma_f32", ", %9");
      if constexpr (OP == 11) PQC_CHAINS8("v_mul_f32", "");
      if constexpr (OP == 12) PQC_CHAINS8("v_mul_u32_u24", "");
      if constexpr (OP == 13) PQC_CHAINS8("v_mad_i32_i24", ", %9");
      if constexpr (OP == 14) PQC_CHAINS8("v_bfe_u32", ", 16");
      if constexpr (OP == 15) PQC_CHAINS8("v_perm_b32", ", %9");
    }
  }
  out[(size_t)blockIdx.x * kBlock + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}
#undef PQC_CHAINS8

}  // namespace
}  // namespace pqc

extern "C" int pqc_valu_issue_rate(int op, int waves_per_simd, double* cycles_per_instr, void* stream) {
  using namespace pqc;
  if (!cycles_per_instr || op < 0 || op > 15 || waves_per_simd < 1 || waves_per_simd > 8) return PQC_ERR_INVALID_ARG;
  int dev = 0, cus = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, dev) != hipSuccess || cus <= 0 || khz <= 0)
    return PQC_ERR_NO_DEVICE;
  hipStream_t s = (hipStream_t)stream;
  const int blocks = cus * 4 * waves_per_simd;
  uint32_t* out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * kBlock * sizeof(uint32_t)) != hipSuccess) return PQC_ERR_NO_MEMORY;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PQC_OK;
  float best = 0.f;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = PQC_ERR_DEVICE;
  for (int rep = 0; !rc && rep < 3; ++rep) {  // the first launch also warms the clocks up
    if (hipEventRecord(e0, s) != hipSuccess) rc = PQC_ERR_DEVICE;
    if (!rc) {
      auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, s, out, 7u); };
      if (op == 0) launch(k_valu_rate<0>);
      if (op == 1) launch(k_valu_rate<1>);
      if (op == 2) launch(k_valu_rate<2>);
      if (op == 3) launch(k_valu_rate<3>);
      if (op == 4) launch(k_valu_rate<4>);
      if (op == 5) launch(k_valu_rate<5>);
      if (op == 6) launch(k_valu_rate<6>);
      if (op == 7) launch(k_valu_rate<7>);
      if (op == 8) launch(k_valu_rate<8>);
      if (op == 9) launch(k_valu_rate<9>);
      if (op == 10) launch(k_valu_rate<10>);
      if (op == 11) launch(k_valu_rate<11>);
      if (op == 12) launch(k_valu_rate<12>);
      if (op == 13) launch(k_valu_rate<13>);
      if (op == 14) launch(k_valu_rate<14>);
      if (op == 15) launch(k_valu_rate<15>);
      if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
          hipEventSynchronize(e1) != hipSuccess)
        rc = PQC_ERR_DEVICE;
    }
    float ms = 0.f;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PQC_ERR_DEVICE;
    if (!rc && (rep == 0 || ms < best)) best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (rc) return rc;
  // wave64 instructions per SIMD (4 SIMDs per CU) over the elapsed cycles
  const double per_simd = (double)waves_per_simd * kIters * kInstrPerIter;
  *cycles_per_instr = (double)best * 1e-3 * (double)khz * 1e3 / per_simd;
  return PQC_OK;
}

// Keccak-f[1600] throughput of the three state layouts the kernels use, on a
// full grid of `waves_per_simd` waves per SIMD: form 0 one state per lane
// (keccakf1600), 1 one state per lane pair with DPP swaps (keccakf1600_pair),
// 2 the bit-interleaved lane pair (keccakf1600_pair_il).  Reports state
// permutations per second for the whole GPU (a lane pair is one state).
namespace pqc {
namespace {
constexpr int kKIters = 400;

template <int FORM>
__global__ __launch_bounds__(64) void k_keccak_rate(uint32_t* out, uint32_t seed) {
  const int l = threadIdx.x;
  uint32_t x = 0;
  if constexpr (FORM == 0) {
    uint64_t A[25];
#pragma unroll
    for (int i = 0; i < 25; ++i) A[i] = (uint64_t)(l * 7 + i) ^ seed;
    for (int it = 0; it < kKIters; ++it) keccakf1600(A);
#pragma unroll
    for (int i = 0; i < 25; ++i) x ^= (uint32_t)A[i] ^ (uint32_t)(A[i] >> 32);
  } else {
    uint32_t A[25];
#pragma unroll
    for (int i = 0; i < 25; ++i) A[i] = (uint32_t)(l * 7 + i) ^ seed;
    for (int it = 0; it < kKIters; ++it) {
      if constexpr (FORM == 1) keccakf1600_pair(A, (uint32_t)(l & 1));
      else keccakf1600_pair_il(A, (uint32_t)(l & 1));
    }
#pragma unroll
    for (int i = 0; i < 25; ++i) x ^= A[i];
  }
  out[(size_t)blockIdx.x * 64 + l] = x;
}
}  // namespace
}  // namespace pqc

extern "C" int pqc_keccak_rate(int form, int waves_per_simd, double* state_perms_per_s, void* stream) {
  using namespace pqc;
  if (!state_perms_per_s || form < 0 || form > 2 || waves_per_simd < 1 || waves_per_simd > 8)
    return PQC_ERR_INVALID_ARG;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return PQC_ERR_NO_DEVICE;
  hipStream_t s = (hipStream_t)stream;
  const int blocks = cus * 4 * waves_per_simd;
  uint32_t* out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * 64 * sizeof(uint32_t)) != hipSuccess) return PQC_ERR_NO_MEMORY;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PQC_OK;
  float best = 0.f;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = PQC_ERR_DEVICE;
  for (int rep = 0; !rc && rep < 3; ++rep) {
    if (hipEventRecord(e0, s) != hipSuccess) rc = PQC_ERR_DEVICE;
    if (!rc) {
      if (form == 0) hipLaunchKernelGGL(k_keccak_rate<0>, dim3(blocks), dim3(64), 0, s, out, 7u);
      if (form == 1) hipLaunchKernelGGL(k_keccak_rate<1>, dim3(blocks), dim3(64), 0, s, out, 7u);
      if (form == 2) hipLaunchKernelGGL(k_keccak_rate<2>, dim3(blocks), dim3(64), 0, s, out, 7u);
      if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
          hipEventSynchronize(e1) != hipSuccess)
        rc = PQC_ERR_DEVICE;
    }
    float ms = 0.f;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PQC_ERR_DEVICE;
    if (!rc && (rep == 0 || ms < best)) best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (rc) return rc;
  const double states = (double)blocks * 64 / (form == 0 ? 1 : 2);
  *state_perms_per_s = states * kKIters / ((double)best * 1e-3);
  return PQC_OK;
}
