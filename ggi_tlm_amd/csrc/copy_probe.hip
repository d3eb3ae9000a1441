// copy_probe.hip -- calibration program for HBM measurements.
//
// Streams a plain float4 copy of a known number of bytes, so that rocprof's
// FETCH_SIZE / WRITE_SIZE counters can be scaled (bytes the counter reports
// for a copy of known size), and prints the copy bandwidth as the practical
// HBM ceiling of the box.  A second kernel streams twelve float planes in
// and out one cell per lane (the link-plane access pattern without
// neighbour reads), the practical ceiling of a link-plane update.
// Usage: copy_probe [MiB=1024] [reps=20]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

// one float4 per lane, no loop: the practical streaming ceiling
__global__ __launch_bounds__(256) void k_copy1(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// twelve float planes of stride P in, twelve out, one cell per lane
__global__ __launch_bounds__(256) void k_planes12(const float* __restrict__ a, float* __restrict__ b, long long P) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P) return;
  float v[12];
#pragma unroll
  for (int p = 0; p < 12; p++) v[p] = a[p * P + i];
#pragma unroll
  for (int p = 0; p < 12; p++) b[p * P + i] = v[(p + 1) % 12];
}

static double time_ms(hipEvent_t e0, hipEvent_t e1, int reps) {
  float ms = 0.f;
  hipEventSynchronize(e1);
  hipEventElapsedTime(&ms, e0, e1);
  return ms / reps;
}

int main(int argc, char** argv) {
  long long mib = argc > 1 ? atoll(argv[1]) : 1024;
  int reps = argc > 2 ? atoi(argv[2]) : 20;
  if (mib < 1 || reps < 1) {
    fprintf(stderr, "usage: copy_probe [MiB>=1] [reps>=1]\n");
    return 2;
  }
  long long bytes = mib << 20;
  long long n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const int blocks = 256 * 32;
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
  hipDeviceSynchronize();
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
  hipEventRecord(e1, 0);
  double per = time_ms(e0, e1, reps);
  printf("{\"kernel\": \"copy\", \"copy_bytes_read\": %lld, \"copy_bytes_written\": %lld, \"avg_ms\": %.5f, "
         "\"GB_s\": %.1f}\n", bytes, bytes, per, 2.0 * bytes / (per * 1e-3) / 1e9);
  const int g1 = (int)((n + 255) / 256);
  hipLaunchKernelGGL(k_copy1, dim3(g1), dim3(256), 0, 0, a, b, n);
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy1, dim3(g1), dim3(256), 0, 0, a, b, n);
  hipEventRecord(e1, 0);
  per = time_ms(e0, e1, reps);
  printf("{\"kernel\": \"copy1\", \"avg_ms\": %.5f, \"GB_s\": %.1f}\n", per, 2.0 * bytes / (per * 1e-3) / 1e9);
  // twelve planes of 2^24 cells (the 256^3 grid) with the engine's plane
  // padding of 32 cells, when the buffer holds them; else a twelfth of it
  long long P = 12LL * ((1LL << 24) + 32) * 4 <= bytes ? (1LL << 24) + 32 : bytes / 4 / 12 / 64 * 64;
  const int g = (int)((P + 255) / 256);
  hipLaunchKernelGGL(k_planes12, dim3(g), dim3(256), 0, 0, (const float*)a, (float*)b, P);
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_planes12, dim3(g), dim3(256), 0, 0, (const float*)a, (float*)b, P);
  hipEventRecord(e1, 0);
  per = time_ms(e0, e1, reps);
  printf("{\"kernel\": \"planes12\", \"cells\": %lld, \"avg_ms\": %.5f, \"GB_s\": %.1f}\n", P, per,
         2.0 * 48 * P / (per * 1e-3) / 1e9);
  hipFree(a);
  hipFree(b);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}
