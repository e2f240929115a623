// lds_probe.hip -- measures the LDS read bandwidth of the device (used by
// bench.py for the LDS roofline peak; not part of the engine's C ABI).
//
// Every workgroup (256 threads) fills 16 KB of LDS and then reads it back with
// 128-bit loads in a bank-conflict-free pattern (consecutive lanes read
// consecutive 16-byte words), many times over; the XOR of everything read is
// stored so that the loads stay live.
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace {

constexpr int kThreads = 256;
constexpr int kWords = 1024;          // uint4 per workgroup = 16 KB

__global__ __launch_bounds__(kThreads) void lds_read_kernel(uint4* out, uint32_t iters)
{
  __shared__ uint4 buf[kWords];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < kWords; i += kThreads)
    buf[i] = make_uint4(i, i * 3u, i * 5u, i * 7u ^ blockIdx.x);
  __syncthreads();
  uint4 acc = make_uint4(0, 0, 0, 0);
  for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < kWords / kThreads; k++) {
      const uint4 v = buf[(t + k * kThreads + it) & (kWords - 1)];
      acc.x ^= v.x;
      acc.y ^= v.y;
      acc.z ^= v.z;
      acc.w ^= v.w;
    }
  }
  out[blockIdx.x * kThreads + t] = acc;
}

} // namespace

/// Runs the probe on the current device; returns 0 and the measured LDS read
/// bandwidth in GB/s, or a negative value on a HIP error.
extern "C" int pcg_probe_lds_read_gbs(double* gbs)
{
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
  const uint32_t blocks = (uint32_t) cus * 8;          // 8 workgroups (32 waves) per CU
  const uint32_t iters = 2048;
  uint4* out = nullptr;
  if (hipMalloc(&out, (size_t) blocks * kThreads * sizeof(uint4)) != hipSuccess) return -2;
  hipEvent_t e0, e1;
  (void) hipEventCreate(&e0);
  (void) hipEventCreate(&e1);
  hipLaunchKernelGGL(lds_read_kernel, dim3(blocks), dim3(kThreads), 0, 0, out, 16u);   // warm-up
  (void) hipEventRecord(e0, 0);
  hipLaunchKernelGGL(lds_read_kernel, dim3(blocks), dim3(kThreads), 0, 0, out, iters);
  (void) hipEventRecord(e1, 0);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  (void) hipEventDestroy(e0);
  (void) hipEventDestroy(e1);
  (void) hipFree(out);
  if (err != hipSuccess || ms <= 0) return -3;
  const double bytes = (double) blocks * kThreads * iters * (kWords / kThreads) * 16.0;
  *gbs = bytes / (ms * 1e-3) / 1e9;
  return 0;
}
