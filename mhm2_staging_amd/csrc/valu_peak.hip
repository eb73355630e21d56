// Sustained packed 16-bit integer VALU rate of one device, the roofline the
// Smith-Waterman kernels are measured against (they are max-plus integer DP,
// so neither MFMA nor HBM bounds them).
//
// Every lane runs 8 independent chains of v_pk_add_u16 / v_pk_max_i16 (the
// two instruction kinds of the DP cell update), 2 int16 lane-ops each, on a
// grid of 32 wavefronts per CU.  The kernel also reads the shader cycle
// counter and the constant-rate wall clock around its loop, so the clock the
// device actually sustained under this load is measured too.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mhmsw.h"

namespace {

constexpr int kChains = 8;
constexpr int kInner = 16;  // unrolled rounds per loop iteration

__global__ __launch_bounds__(256) void valu_peak_kernel(uint32_t seed, int iters, uint32_t* sink,
                                                        unsigned long long* clocks) {
  uint32_t a[kChains];
  const uint32_t b = seed ^ (threadIdx.x * 0x00010001u);
#pragma unroll
  for (int k = 0; k < kChains; k++) a[k] = seed * (uint32_t)(k + 3) + threadIdx.x;
  const unsigned long long c0 = clock64();
  const unsigned long long w0 = wall_clock64();
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int r = 0; r < kInner; r++) {
#pragma unroll
      for (int k = 0; k < kChains; k++) {
        asm volatile("v_pk_add_u16 %0, %0, %1" : "+v"(a[k]) : "v"(b));
        asm volatile("v_pk_max_i16 %0, %0, %1" : "+v"(a[k]) : "v"(b));
      }
    }
  }
  const unsigned long long c1 = clock64();
  const unsigned long long w1 = wall_clock64();
  uint32_t x = 0;
#pragma unroll
  for (int k = 0; k < kChains; k++) x ^= a[k];
  if (x == seed) sink[0] = x;  // keeps the chains live; practically never taken
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    clocks[0] = c1 - c0;
    clocks[1] = w1 - w0;
  }
}

}  // namespace

extern "C" int mhmsw_measure_valu_peak(int device_id, mhmsw_valu_peak* out) {
  if (!out) return MHMSW_E_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MHMSW_E_NODEVICE;
  if (device_id < 0 || device_id >= n) return MHMSW_E_ARG;
  if (hipSetDevice(device_id) != hipSuccess) return MHMSW_E_HIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_id) != hipSuccess) return MHMSW_E_HIP;
  int wall_khz = 0;
  if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device_id) != hipSuccess || wall_khz <= 0)
    return MHMSW_E_HIP;
  const int cus = prop.multiProcessorCount;
  const int blocks = cus * 8;  // 8 blocks of 4 wavefronts per CU: 8 wavefronts per SIMD
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  uint32_t* sink = nullptr;
  unsigned long long* clocks = nullptr;
  int rc = MHMSW_OK;
  unsigned long long hclk[2] = {0, 0};
  float ms = 0.f;
  int iters = 2000;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess || hipMalloc((void**)&sink, 256) != hipSuccess ||
      hipMalloc((void**)&clocks, 2 * sizeof(unsigned long long)) != hipSuccess) {
    rc = MHMSW_E_HIP;
  }
  // a short warm-up run, then one sized to run for about 100 ms
  for (int pass = 0; pass < 2 && rc == MHMSW_OK; pass++) {
    hipLaunchKernelGGL(valu_peak_kernel, dim3(blocks), dim3(256), 0, st, 0x1234567u, iters, sink, clocks);
    if (hipEventRecord(e0, st) != hipSuccess) rc = MHMSW_E_HIP;
    if (rc == MHMSW_OK)
      hipLaunchKernelGGL(valu_peak_kernel, dim3(blocks), dim3(256), 0, st, 0x1234567u, iters, sink, clocks);
    if (rc == MHMSW_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
                           hipStreamSynchronize(st) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = MHMSW_E_HIP;
    if (rc == MHMSW_OK && pass == 0) {
      const double want = 100.0 / (ms > 0.01f ? ms : 0.01f);
      iters = (int)(iters * (want < 1.0 ? 1.0 : (want > 2000.0 ? 2000.0 : want)));
    }
  }
  if (rc == MHMSW_OK &&
      hipMemcpy(hclk, clocks, sizeof(hclk), hipMemcpyDeviceToHost) != hipSuccess)
    rc = MHMSW_E_HIP;
  if (rc == MHMSW_OK) {
    const double lane_ops = (double)blocks * 256.0 * (double)iters * kInner * kChains * 2 /* instr */ * 2 /* halves */;
    out->lane_ops_per_s = lane_ops / (ms * 1e-3);
    out->kernel_ms = ms;
    out->cus = cus;
    const double wall_s = (double)hclk[1] / ((double)wall_khz * 1e3);
    out->clock_mhz = wall_s > 0 ? (double)hclk[0] / wall_s / 1e6 : 0.0;
    out->max_clock_mhz = prop.clockRate / 1e3;
  }
  if (sink) (void)hipFree(sink);
  if (clocks) (void)hipFree(clocks);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}
