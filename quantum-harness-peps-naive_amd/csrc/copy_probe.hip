// copy_probe.hip -- HBM traffic calibration for rocprofv3 counters.
//
// Copies a buffer of a known size with a plain streaming kernel so that the
// FETCH_SIZE / WRITE_SIZE counters of the contraction kernels can be scaled
// by what the same counters read for exactly `bytes` read and `bytes` written;
// also reports the device's streaming copy and read rates.
// Usage: peps-copy-probe [MiB] [repeats]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const uint4* __restrict__ in, uint4* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

// Read-only stream (sum of all words): what a read pass can reach.
__global__ void k_read8(const uint4* __restrict__ in, size_t n, unsigned* __restrict__ sink) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  unsigned acc = 0;
  for (size_t base = (size_t)blockIdx.x * blockDim.x + threadIdx.x; base < n; base += 8 * stride) {
    uint4 v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = base + j * stride < n ? in[base + j * stride] : make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 8; ++j) acc += v[j].x ^ v[j].y ^ v[j].z ^ v[j].w;
  }
  if (acc == 0x9e3779b9u) sink[0] = acc;  // keeps the loads; practically never stores
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 4096;
  const int reps = argc > 2 ? atoi(argv[2]) : 3;
  const size_t bytes = mib << 20;
  const size_t n = bytes / sizeof(uint4);
  uint4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int r = 0; r < reps; ++r) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    std::printf("copy bytes_read=%zu bytes_written=%zu ms=%.4f GBps=%.1f\n", bytes, bytes, ms,
                2.0 * bytes / (ms * 1e-3) / 1e9);
  }
  unsigned* sink = nullptr;
  if (hipMalloc(&sink, sizeof(unsigned)) != hipSuccess) return 1;
  for (int r = 0; r < reps; ++r) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(k_read8, dim3(4096), dim3(256), 0, 0, a, n, sink);
    (void)hipEventRecord(e1);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    std::printf("read8 bytes_read=%zu ms=%.4f GBps=%.1f\n", bytes, ms, bytes / (ms * 1e-3) / 1e9);
  }
  (void)hipFree(sink);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
