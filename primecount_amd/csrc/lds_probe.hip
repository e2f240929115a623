// lds_probe.hip -- measures the LDS read bandwidth of the device, the peak of
// the LDS roofline of the sieve kernel (pcg_lds_read_gbs).
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

namespace pcg {

/// Runs the probe on the current device: the measured LDS read bandwidth in
/// GB/s (bytes read by all workgroups / kernel time), or the HIP error of a
/// failed call or launch.
hipError_t lds_read_gbs(double* gbs)
{
  int dev = 0, cus = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err == hipSuccess) err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (err != hipSuccess) return err;
  const uint32_t blocks = (uint32_t) cus * 8;          // 8 workgroups (32 waves) per CU
  const uint32_t iters = 2048;
  uint4* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t st = nullptr;
  float ms = 0;
  if ((err = hipMalloc(&out, (size_t) blocks * kThreads * sizeof(uint4))) != hipSuccess) return err;
  if ((err = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) goto done;
  if ((err = hipEventCreate(&e0)) != hipSuccess) goto done;
  if ((err = hipEventCreate(&e1)) != hipSuccess) goto done;
  hipLaunchKernelGGL(lds_read_kernel, dim3(blocks), dim3(kThreads), 0, st, out, 16u);  // warm-up
  if ((err = hipGetLastError()) != hipSuccess) goto done;
  if ((err = hipEventRecord(e0, st)) != hipSuccess) goto done;
  hipLaunchKernelGGL(lds_read_kernel, dim3(blocks), dim3(kThreads), 0, st, out, iters);
  if ((err = hipGetLastError()) != hipSuccess) goto done;
  if ((err = hipEventRecord(e1, st)) != hipSuccess) goto done;
  if ((err = hipEventSynchronize(e1)) != hipSuccess) goto done;
  if ((err = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) goto done;
  if (ms <= 0) {
    err = hipErrorUnknown;
    goto done;
  }
  *gbs = (double) blocks * kThreads * iters * (kWords / kThreads) * 16.0 / (ms * 1e-3) / 1e9;
done:
  if (e0) (void) hipEventDestroy(e0);
  if (e1) (void) hipEventDestroy(e1);
  if (st) (void) hipStreamDestroy(st);
  (void) hipFree(out);
  return err;
}

} // namespace pcg
