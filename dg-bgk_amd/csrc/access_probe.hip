// Probe: the update kernel's memory-access structure without its arithmetic
// or halo, on the particle workload's shape (4,907 tiles of 256 slots, 2048
// ordinates, 5 ordinate chunks, FP64 f stored velocity-major: row v holds
// every tile's 256 x 3 values).  Each block (tile, chunk) walks its chunk's
// ordinates: every thread reads its slot's values at ordinate v, adds 1 and
// writes them to the same place in the second buffer, as k_update reads f^k
// and writes f^{k+1}.  Variants: doubles per lane (3, as an element's three
// nodes, or 2), an LDS exchange and barrier per ordinate (as k_update's),
// the next ordinate's loads in flight (as k_update's prefetch),
// non-temporal or plain stores, and, as a layout alternative, f stored
// tile-major (each tile's ordinates contiguous); then the FP32-bulk shape
// (3 floats per lane, 7 chunks of 293 ordinates) in the first half of
// each buffer.  Each variant: one warm-up launch, then the
// mean of 3 launches timed with HIP events; bytes = one read + one write.
//
//   access_probe
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), \
                   __LINE__);                                                  \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

constexpr int kSlots = 256;

template <int kW, bool kBar, bool kPre, bool kNT, bool kTileMajor = false, typename T = double>
__global__ void __launch_bounds__(kSlots) walk(const T* __restrict__ a, T* __restrict__ b,
                                               int ntiles, int V, int vchunk) {
  const int tile = blockIdx.x % ntiles, c = blockIdx.x / ntiles;
  const int v0 = c * vchunk, v1 = min(V, v0 + vchunk);
  __shared__ T sh[kSlots];
  T acc = 0;
  // ordinate stride and this lane's offset: velocity-major rows of all
  // tiles, or each tile's ordinates contiguous
  const size_t row = kTileMajor ? (size_t)kW * kSlots : (size_t)ntiles * kW * kSlots;
  const size_t base = (kTileMajor ? (size_t)tile * V * kW * kSlots : (size_t)tile * kW * kSlots) +
                      kW * threadIdx.x;
  T x[kW], y[kW];
  for (int j = 0; j < kW; ++j) x[j] = a[(size_t)v0 * row + base + j];
  for (int v = v0; v < v1; ++v) {
    if (kPre) {  // clamped, so the load is unconditional
      const int vn = min(v + 1, v1 - 1);
      for (int j = 0; j < kW; ++j) y[j] = a[(size_t)vn * row + base + j];
    } else {
      for (int j = 0; j < kW; ++j) x[j] = a[(size_t)v * row + base + j];
    }
    if (kBar) {
      sh[threadIdx.x] = x[0];
      __syncthreads();
      acc += sh[(threadIdx.x + 1) % kSlots];
    }
    T* q = b + (size_t)v * row + base;
    for (int j = 0; j < kW; ++j) {
      if (kNT)
        __builtin_nontemporal_store(x[j] + T(1), q + j);
      else
        q[j] = x[j] + T(1);
    }
    if (kBar) __syncthreads();
    if (kPre)
      for (int j = 0; j < kW; ++j) x[j] = y[j];
  }
  if (acc == T(-1)) b[0] = acc;  // keeps the LDS exchange
}

typedef void (*Kern)(const double*, double*, int, int, int);
typedef void (*KernF)(const float*, float*, int, int, int);

int main() {
  const int ntiles = 4907, V = 2048, vchunk = 410;
  const int nchunks = (V + vchunk - 1) / vchunk;
  const size_t n = (size_t)ntiles * V * 3 * kSlots;  // doubles per buffer (3 per lane)
  double *a = nullptr, *b = nullptr;
  CK(hipMalloc((void**)&a, n * sizeof(double)));
  CK(hipMalloc((void**)&b, n * sizeof(double)));
  CK(hipMemset(a, 0, n * sizeof(double)));
  CK(hipMemset(b, 0, n * sizeof(double)));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  struct Variant {
    const char* name;
    Kern f;
    int w;
  } vs[] = {
      {"3/lane barrier prefetch nt", walk<3, true, true, true>, 3},
      {"3/lane barrier nt", walk<3, true, false, true>, 3},
      {"3/lane prefetch nt", walk<3, false, true, true>, 3},
      {"3/lane barrier prefetch plain", walk<3, true, true, false>, 3},
      {"2/lane barrier prefetch nt", walk<2, true, true, true>, 2},
      {"2/lane prefetch nt", walk<2, false, true, true>, 2},
      {"3/lane barrier prefetch nt, tile-major", walk<3, true, true, true, true>, 3},
  };
  const dim3 grid(ntiles * nchunks), block(kSlots);
  std::printf("buffers: 2 x %.1f GB; %d tiles x %d slots, %d ordinates in %d chunks\n",
              n * sizeof(double) / 1e9, ntiles, kSlots, V, nchunks);
  for (int rep = 0; rep < 2; ++rep)
    for (const Variant& x : vs) {
      const double bytes = 2.0 * ntiles * V * x.w * kSlots * sizeof(double);
      hipLaunchKernelGGL(x.f, grid, block, 0, 0, a, b, ntiles, V, vchunk);
      CK(hipGetLastError());
      CK(hipEventRecord(e0));
      for (int it = 0; it < 3; ++it) hipLaunchKernelGGL(x.f, grid, block, 0, 0, a, b, ntiles, V, vchunk);
      CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      ms /= 3.f;
      std::printf("rep %d %-40s %8.3f ms %8.1f GB/s\n", rep, x.name, ms, bytes / ms / 1e6);
      std::fflush(stdout);
    }
  // the FP32-bulk shape: 3 floats per lane (f of one element), the FP32
  // update's 7 ordinate chunks of 293
  {
    const int vchunk32 = 293, nch32 = (V + vchunk32 - 1) / vchunk32;
    const dim3 grid32(ntiles * nch32);
    struct VariantF {
      const char* name;
      KernF f;
    } vf[] = {
        {"fp32 3/lane barrier prefetch nt", walk<3, true, true, true, false, float>},
        {"fp32 3/lane barrier prefetch plain", walk<3, true, true, false, false, float>},
    };
    for (int rep = 0; rep < 2; ++rep)
      for (const VariantF& x : vf) {
        const double bytes = 2.0 * ntiles * V * 3 * kSlots * sizeof(float);
        hipLaunchKernelGGL(x.f, grid32, block, 0, 0, (const float*)a, (float*)b, ntiles, V,
                           vchunk32);
        CK(hipGetLastError());
        CK(hipEventRecord(e0));
        for (int it = 0; it < 3; ++it)
          hipLaunchKernelGGL(x.f, grid32, block, 0, 0, (const float*)a, (float*)b, ntiles, V,
                             vchunk32);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        ms /= 3.f;
        std::printf("rep %d %-40s %8.3f ms %8.1f GB/s\n", rep, x.name, ms, bytes / ms / 1e6);
        std::fflush(stdout);
      }
  }
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
