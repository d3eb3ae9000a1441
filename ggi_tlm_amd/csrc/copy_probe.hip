// copy_probe.hip -- calibration program for HBM measurements.
//
// Streams a plain float4 copy of a known number of bytes, so that rocprof's
// FETCH_SIZE / WRITE_SIZE counters can be scaled (bytes the counter reports
// for a copy of known size), and prints the copy bandwidth as the practical
// HBM ceiling of the box.  Usage: copy_probe [MiB=1024] [reps=20]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

// one float4 per thread, no loop
__global__ __launch_bounds__(256) void k_copy1(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// twelve float planes in, twelve out, one cell per thread: the access
// pattern of the link planes without neighbour reads
__global__ __launch_bounds__(256) void k_planes12(const float* __restrict__ a, float* __restrict__ b, long long P) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)gridDim.x * blockDim.x || i >= P) return;
  float v[12];
#pragma unroll
  for (int p = 0; p < 12; p++) v[p] = a[p * P + i];
#pragma unroll
  for (int p = 0; p < 12; p++) b[p * P + i] = v[(p + 1) % 12];
}

// as k_planes12, but rows of 256 cells skip their first and last `skip`
// cells (written by another kernel): partial cache lines at both row ends
__global__ __launch_bounds__(256) void k_planes12_skip(const float* __restrict__ a, float* __restrict__ b, long long P,
                                                       long long n, int skip) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x = (int)(i & 255);
  if (x < skip || x >= 256 - skip) return;
  float v[12];
#pragma unroll
  for (int p = 0; p < 12; p++) v[p] = a[p * P + i];
#pragma unroll
  for (int p = 0; p < 12; p++) b[p * P + i] = v[(p + 1) % 12];
}

// only the first and last 32 cells of each 256-cell row, one cell per
// thread (threads packed onto those cells): whole 128-byte lines in
// 256-byte islands 1 KiB apart, the pattern of lines at both ends of a row
__global__ __launch_bounds__(256) void k_planes12_ends(const float* __restrict__ a, float* __restrict__ b, long long P,
                                                       long long rows) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long row = t >> 6;
  if (row >= rows) return;
  int k = (int)(t & 63);
  long long i = row * 256 + (k < 32 ? k : 192 + k);
  float v[12];
#pragma unroll
  for (int p = 0; p < 12; p++) v[p] = a[p * P + i];
#pragma unroll
  for (int p = 0; p < 12; p++) b[p * P + i] = v[(p + 1) % 12];
}

// complement of k_planes12_skip: only the first and last `skip` cells of
// each 256-cell row, over rows [r0, r1)
__global__ __launch_bounds__(256) void k_planes12_edges(const float* __restrict__ a, float* __restrict__ b, long long P,
                                                        long long r0, long long r1, int skip) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long row = r0 + t / (2 * skip);
  if (row >= r1) return;
  int k = (int)(t % (2 * skip));
  long long i = row * 256 + (k < skip ? k : 256 - 2 * skip + k);
  float v[12];
#pragma unroll
  for (int p = 0; p < 12; p++) v[p] = a[p * P + i];
#pragma unroll
  for (int p = 0; p < 12; p++) b[p * P + i] = v[(p + 1) % 12];
}

// k_planes12_skip over rows [r0, r1)
__global__ __launch_bounds__(256) void k_planes12_skip_rows(const float* __restrict__ a, float* __restrict__ b,
                                                            long long P, long long r0, long long r1, int skip) {
  long long i = r0 * 256 + blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= r1 * 256) return;
  int x = (int)(i & 255);
  if (x < skip || x >= 256 - skip) return;
  float v[12];
#pragma unroll
  for (int p = 0; p < 12; p++) v[p] = a[p * P + i];
#pragma unroll
  for (int p = 0; p < 12; p++) b[p * P + i] = v[(p + 1) % 12];
}

int main(int argc, char** argv) {
  long long mib = argc > 1 ? atoll(argv[1]) : 1024;
  int reps = argc > 2 ? atoi(argv[2]) : 20;
  long long bytes = mib << 20;
  long long n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  hipMemset(a, 0, bytes);
  hipMemset(b, 0, bytes);
  int blocks = 256 * 32;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
  hipDeviceSynchronize();
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
  hipEventRecord(e1, 0);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  double per = ms / reps;
  printf("{\"copy_bytes_read\": %lld, \"copy_bytes_written\": %lld, \"avg_ms\": %.5f, \"GB_s\": %.1f}\n", bytes, bytes,
         per, 2.0 * bytes / (per * 1e-3) / 1e9);
  // one float4 per thread
  {
    int g1 = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_copy1, dim3(g1), dim3(256), 0, 0, a, b, n);
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy1, dim3(g1), dim3(256), 0, 0, a, b, n);
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    per = ms / reps;
    printf("{\"kernel\": \"copy1\", \"avg_ms\": %.5f, \"GB_s\": %.1f}\n", per, 2.0 * bytes / (per * 1e-3) / 1e9);
  }
  // twelve planes of 2^24 cells with padded plane strides: power-of-two
  // strides put every plane on the same memory channel at the same time
  {
    const long long N = 1LL << 24;
    const long long pads[] = {0, 64, 256, 1024, 4096, 4288, 65536 + 4288};
    for (long long pad : pads) {
      long long P = N + pad;
      if (12 * P * 4 > bytes) break;
      int g1 = (int)((N + 255) / 256);
      hipLaunchKernelGGL(k_planes12, dim3(g1), dim3(256), 0, 0, (const float*)a, (float*)b, P);
      hipEventRecord(e0, 0);
      for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_planes12, dim3(g1), dim3(256), 0, 0, (const float*)a, (float*)b, P);
      hipEventRecord(e1, 0);
      hipEventSynchronize(e1);
      hipEventElapsedTime(&ms, e0, e1);
      per = ms / reps;
      printf("{\"kernel\": \"planes12\", \"cells\": %lld, \"pad\": %lld, \"avg_ms\": %.5f, \"GB_s\": %.1f}\n", N, pad, per,
             2.0 * 48 * N / (per * 1e-3) / 1e9);
    }
  }
  // partial rows
  if (12 * (1LL << 24) * 4 <= bytes) {
    const long long N = 1LL << 24;
    const int skips[] = {0, 8, 16, 32};
    for (int sk : skips) {
      int g1 = (int)((N + 255) / 256);
      hipLaunchKernelGGL(k_planes12_skip, dim3(g1), dim3(256), 0, 0, (const float*)a, (float*)b, N, N, sk);
      hipEventRecord(e0, 0);
      for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_planes12_skip, dim3(g1), dim3(256), 0, 0, (const float*)a, (float*)b, N, N, sk);
      hipEventRecord(e1, 0);
      hipEventSynchronize(e1);
      hipEventElapsedTime(&ms, e0, e1);
      per = ms / reps;
      double cells = (double)N * (256 - 2 * sk) / 256.0;
      printf("{\"kernel\": \"planes12_skip\", \"skip\": %d, \"avg_ms\": %.5f, \"GB_s\": %.1f}\n", sk, per,
             2.0 * 48 * cells / (per * 1e-3) / 1e9);
    }
  }
  // row ends only
  if (12 * (1LL << 24) * 4 <= bytes) {
    const long long N = 1LL << 24;
    const long long rows = N / 256;
    int g1 = (int)((rows * 64 + 255) / 256);
    hipLaunchKernelGGL(k_planes12_ends, dim3(g1), dim3(256), 0, 0, (const float*)a, (float*)b, N, rows);
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++)
      hipLaunchKernelGGL(k_planes12_ends, dim3(g1), dim3(256), 0, 0, (const float*)a, (float*)b, N, rows);
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    per = ms / reps;
    printf("{\"kernel\": \"planes12_ends\", \"avg_ms\": %.5f, \"GB_s\": %.1f}\n", per,
           2.0 * 48 * rows * 64 / (per * 1e-3) / 1e9);
  }
  // complementary partial writes from two kernels (interior, then row
  // edges), over the whole array or chunk by chunk: do the partial lines of
  // the first kernel merge with the second's while still cached?
  if (12 * (1LL << 24) * 4 <= bytes) {
    const long long N = 1LL << 24;
    const long long rows = N / 256;
    const int skip = 8;
    const long long chunks[] = {rows, rows / 4, rows / 16, rows / 64, rows / 256};
    for (long long ch : chunks) {
      auto pass = [&]() {
        for (long long r0 = 0; r0 < rows; r0 += ch) {
          long long r1 = r0 + ch < rows ? r0 + ch : rows;
          int g1 = (int)(((r1 - r0) * 256 + 255) / 256);
          hipLaunchKernelGGL(k_planes12_skip_rows, dim3(g1), dim3(256), 0, 0, (const float*)a, (float*)b, N, r0, r1,
                             skip);
          int g2 = (int)(((r1 - r0) * 2 * skip + 255) / 256);
          hipLaunchKernelGGL(k_planes12_edges, dim3(g2), dim3(256), 0, 0, (const float*)a, (float*)b, N, r0, r1,
                             skip);
        }
      };
      pass();
      hipEventRecord(e0, 0);
      for (int r = 0; r < reps; r++) pass();
      hipEventRecord(e1, 0);
      hipEventSynchronize(e1);
      hipEventElapsedTime(&ms, e0, e1);
      per = ms / reps;
      printf("{\"kernel\": \"planes12_split\", \"chunk_rows\": %lld, \"chunk_MB\": %.1f, \"avg_ms\": %.5f, "
             "\"GB_s\": %.1f}\n", ch, 96.0 * ch * 256 / 1e6, per, 2.0 * 48 * N / (per * 1e-3) / 1e9);
    }
  }
  // twelve planes, original layout
  {
    long long P = bytes / 4 / 12;
    int g1 = (int)((P + 255) / 256);
    hipLaunchKernelGGL(k_planes12, dim3(g1), dim3(256), 0, 0, (const float*)a, (float*)b, P);
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_planes12, dim3(g1), dim3(256), 0, 0, (const float*)a, (float*)b, P);
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    per = ms / reps;
    printf("{\"kernel\": \"planes12\", \"avg_ms\": %.5f, \"GB_s\": %.1f}\n", per, 2.0 * 48 * P / (per * 1e-3) / 1e9);
  }
  hipFree(a);
  hipFree(b);
  return 0;
}
