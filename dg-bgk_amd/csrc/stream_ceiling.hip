// stream_ceiling.hip -- the HBM bandwidth this device reaches with plain
// streaming kernels, the practical ceiling the update and moments kernels
// are measured against (the published 8 TB/s is the peak).
//
// Kernels over buffers of B bytes (default 8 GiB each), 16-byte vector
// accesses, grid-stride loops, each timed with HIP events as the best of 5:
//   copy_nt     b[i] = a[i], non-temporal stores (as the update's f store)
//   copy_plain  b[i] = a[i]
//   copy4_nt    four loads in flight per thread, non-temporal stores
//   read        sum of a (one read of every byte)
//   write_nt    non-temporal stores of a constant
// Bytes moved: 2B for the copies, B for read and write.
//
//   stream_ceiling [GiB per buffer]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float v4f __attribute__((ext_vector_type(4)));

#define CHECK(x)                                                                   \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                                \
    }                                                                              \
  } while (0)

__global__ void __launch_bounds__(256) k_copy_nt(const v4f* __restrict__ a, v4f* __restrict__ b,
                                                 size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(a[i], b + i);
}

__global__ void __launch_bounds__(256) k_copy_plain(const v4f* __restrict__ a,
                                                    v4f* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

__global__ void __launch_bounds__(256) k_copy4_nt(const v4f* __restrict__ a, v4f* __restrict__ b,
                                                  size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const v4f x0 = a[i], x1 = a[i + stride], x2 = a[i + 2 * stride], x3 = a[i + 3 * stride];
    __builtin_nontemporal_store(x0, b + i);
    __builtin_nontemporal_store(x1, b + i + stride);
    __builtin_nontemporal_store(x2, b + i + 2 * stride);
    __builtin_nontemporal_store(x3, b + i + 3 * stride);
  }
  for (; i < n; i += stride) __builtin_nontemporal_store(a[i], b + i);
}

__global__ void __launch_bounds__(256) k_read(const v4f* __restrict__ a, float* __restrict__ out,
                                              size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  float s = 0.f;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    const v4f x = a[i];
    s += x.x + x.y + x.z + x.w;
  }
  if (s == 12345.f) out[0] = s;  // keeps the loads; never true for the zero buffer
}

__global__ void __launch_bounds__(256) k_write_nt(v4f* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  const v4f c = {1.f, 2.f, 3.f, 4.f};
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(c, b + i);
}

int main(int argc, char** argv) {
  const double gib = argc > 1 ? std::atof(argv[1]) : 8.0;
  if (!(gib > 0.0 && gib <= 64.0)) {
    std::fprintf(stderr, "buffer size must be in (0, 64] GiB\n");
    return 2;
  }
  const size_t n = (size_t)(gib * (double)(1ull << 30)) / 16;
  const size_t bytes = n * 16;
  v4f *a = nullptr, *b = nullptr;
  float* o = nullptr;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMalloc(&o, 64));
  CHECK(hipMemset(a, 0, bytes));
  CHECK(hipMemset(b, 0, bytes));
  CHECK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const char* names[] = {"copy_nt", "copy_plain", "copy4_nt", "read", "write_nt"};
  const int grids[] = {1024, 2048, 4096, 8192, 16384, 65536};
  double best_gbs[5] = {0, 0, 0, 0, 0};
  std::printf("buffers: 2 x %.2f GiB\n", bytes / (double)(1ull << 30));
  for (int g : grids) {
    for (int kind = 0; kind < 5; ++kind) {
      float best = 1e30f;
      for (int rep = 0; rep < 5; ++rep) {
        CHECK(hipEventRecord(e0));
        if (kind == 0) hipLaunchKernelGGL(k_copy_nt, dim3(g), dim3(256), 0, 0, a, b, n);
        if (kind == 1) hipLaunchKernelGGL(k_copy_plain, dim3(g), dim3(256), 0, 0, a, b, n);
        if (kind == 2) hipLaunchKernelGGL(k_copy4_nt, dim3(g), dim3(256), 0, 0, a, b, n);
        if (kind == 3) hipLaunchKernelGGL(k_read, dim3(g), dim3(256), 0, 0, a, o, n);
        if (kind == 4) hipLaunchKernelGGL(k_write_nt, dim3(g), dim3(256), 0, 0, b, n);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
      }
      const double moved = (kind <= 2 ? 2.0 : 1.0) * (double)bytes;
      const double gbs = moved / best / 1e6;
      if (gbs > best_gbs[kind]) best_gbs[kind] = gbs;
      std::printf("grid %6d %-10s %8.3f ms %8.1f GB/s\n", g, names[kind], best, gbs);
      std::fflush(stdout);
    }
  }
  std::printf("best:");
  for (int k = 0; k < 5; ++k) std::printf(" %s %.1f GB/s%s", names[k], best_gbs[k], k < 4 ? "," : "\n");
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  CHECK(hipFree(o));
  return 0;
}
