// mlp.hip -- coordinate MLP of the learned warm start, BF16 MFMA on gfx950.
//
// The network maps per-point features (x, y, wall distance, Mach, Kn) to the
// steady state (rho, u, v, T): Linear(5,W) + SiLU, (L-1) x [Linear(W,W) +
// SiLU], Linear(W,4).  Weights and activations are BF16, accumulation and
// master weights FP32, optimiser Adam.  The W x W layers (all but ~1% of the
// work) run on gfx950's v_mfma_f32_32x32x16_bf16 through one tiled GEMM kernel with
// three epilogues: forward (bias + SiLU, keeps the pre-activation), backward
// data (times SiLU' of the previous pre-activation) and backward weight
// (split-K partials over the batch).  The 5-wide input and 4-wide output
// layers are plain FP32 SIMT kernels.  The network only seeds the solver's
// initial Maxwellian (dgbgk_seed_maxwellian); it has no counterpart in the
// reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dgbgk.h"

namespace dgbgk {
namespace mlp {

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef unsigned short bf16;

__device__ __forceinline__ bf16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}

enum Epi { kFwd = 0, kBwdData = 1, kPartial = 2 };

constexpr int kTile = 64, kKC = 32, kPad = 8, kLd = kKC + kPad;
// rows per block of the row-split reductions (bias and small-layer weight
// gradients): short enough that a 65,536-row batch gives 512 blocks per
// output column set, so the loads of many blocks overlap
constexpr int kRowChunk = 128;

// C[M][N] = sum_k A(m,k) B(k,n), 64x64 per block, 32x32 per wave.
// A(m,k) = kAT ? A[k*lda + m] : A[m*lda + k]
// B(k,n) = kBT ? B[k*ldb + n] : B[n*ldb + k]
// Epilogues: kFwd: Z = C + bias (bf16), O = SiLU(Z) (bf16)
//            kBwdData: O = bf16(C * SiLU'(Zprev))
//            kPartial: P[z][m][n] = C (fp32), K split over blockIdx.z
template <bool kAT, bool kBT, int kEpi>
__global__ void __launch_bounds__(256) k_gemm(const bf16* __restrict__ A, int lda,
                                              const bf16* __restrict__ B, int ldb, int K,
                                              const float* __restrict__ bias,
                                              bf16* __restrict__ Z, bf16* __restrict__ O,
                                              const bf16* __restrict__ Zprev,
                                              float* __restrict__ P, int ldc, int Msize) {
  __shared__ bf16 As[kTile * kLd];
  __shared__ bf16 Bs[kTile * kLd];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int n0 = blockIdx.x * kTile, m0 = blockIdx.y * kTile;
  const int kper = (kEpi == kPartial) ? K / gridDim.z : K;
  const int kb = (kEpi == kPartial) ? blockIdx.z * kper : 0;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  floatx16 acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
  for (int k0 = kb; k0 < kb + kper; k0 += kKC) {
    // stage A and B tiles k-contiguous: As[m][k], Bs[n][k]
    if (!kAT) {
      const int m = t >> 2, kq = (t & 3) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(A + (size_t)(m0 + m) * lda + k0 + kq);
      *reinterpret_cast<uint4*>(&As[m * kLd + kq]) = v;
    } else {
      const int k = t >> 3, mq = (t & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(A + (size_t)(k0 + k) * lda + m0 + mq);
      const bf16* h = reinterpret_cast<const bf16*>(&v);
#pragma unroll
      for (int j = 0; j < 8; ++j) As[(mq + j) * kLd + k] = h[j];
    }
    if (!kBT) {
      const int n = t >> 2, kq = (t & 3) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(B + (size_t)(n0 + n) * ldb + k0 + kq);
      *reinterpret_cast<uint4*>(&Bs[n * kLd + kq]) = v;
    } else {
      const int k = t >> 3, nq = (t & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(B + (size_t)(k0 + k) * ldb + n0 + nq);
      const bf16* h = reinterpret_cast<const bf16*>(&v);
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[(nq + j) * kLd + k] = h[j];
    }
    __syncthreads();
    // gfx950's 32x32x16 BF16 MFMA: lane l supplies row (col) l % 32 of A
    // (B) at k = 8 (l / 32) .. 8 (l / 32) + 7 of the 16-wide k step
#pragma unroll
    for (int kk = 0; kk < kKC; kk += 16) {
      const int ko = kk + 8 * (lane >> 5);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[(wm + (lane & 31)) * kLd + ko]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[(wn + (lane & 31)) * kLd + ko]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // accumulator layout of the 32x32 MFMA: lane holds column (lane % 32) and
  // rows 8*(r/4) + 4*(lane/32) + r%4 for r = 0..15
  const int col = n0 + wn + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const size_t idx = (size_t)row * ldc + col;
    if (kEpi == kFwd) {
      const float z = acc[r] + bias[col];
      Z[idx] = f2bf(z);
      O[idx] = f2bf(silu(z));
    } else if (kEpi == kBwdData) {
      O[idx] = f2bf(acc[r] * dsilu(bf2f(Zprev[idx])));
    } else {
      P[(size_t)blockIdx.z * Msize * ldc + idx] = acc[r];
    }
  }
}

// sum over splits of partials: G[i] = sum_z P[z n + i].  A block takes 32
// outputs x 8 split groups (coalesced 128-byte reads per group), the groups
// combine in LDS.  With nin > 0 the outputs are k_wgrad_* partials,
// [nout][nin + 1], scattered into the packed weight [nout][nin] and bias
// [nout] gradients; with nin == 0 G is written directly.
constexpr int kRedCols = 32, kRedGroups = 8;
__global__ void __launch_bounds__(256) k_reduce_splits(const float* __restrict__ P, int splits,
                                                       int n, int nin, float* __restrict__ G,
                                                       float* __restrict__ Gb) {
  __shared__ float s_sum[kRedGroups][kRedCols];
  const int c = threadIdx.x % kRedCols, grp = threadIdx.x / kRedCols;
  const int i = blockIdx.x * kRedCols + c;
  float s = 0.0f;
  if (i < n)
    for (int z = grp; z < splits; z += kRedGroups) s += P[(size_t)z * n + i];
  s_sum[grp][c] = s;
  __syncthreads();
  if (grp != 0 || i >= n) return;
  float t = 0.0f;
#pragma unroll
  for (int g = 0; g < kRedGroups; ++g) t += s_sum[g][c];
  if (nin == 0) {
    G[i] = t;
  } else {
    const int o = i / (nin + 1), k = i % (nin + 1);
    if (k < nin) G[o * nin + k] = t;
    else Gb[o] = t;
  }
}

// column sums of a bf16 [rows][cols] matrix over row chunks (bias
// gradients): P[chunk][c].  Each thread reads 8 columns as one 16-byte
// load; cols / 8 threads cover a row, the block's 256 / (cols / 8) row
// lanes stride the chunk and combine in LDS (cols / 8 must divide 256).
__global__ void __launch_bounds__(256) k_colsum8(const bf16* __restrict__ D, int rows, int cols,
                                                 int rows_per, float* __restrict__ P) {
  __shared__ float s_sum[256 * 8];
  const int gpr = cols / 8, lanes = 256 / gpr;
  const int cg = threadIdx.x % gpr, rl = threadIdx.x / gpr;
  const int r0 = blockIdx.x * rows_per, r1 = min(rows, r0 + rows_per);
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int r = r0 + rl; r < r1; r += lanes) {
    const uint4 v = *reinterpret_cast<const uint4*>(D + (size_t)r * cols + 8 * cg);
    const bf16* h = reinterpret_cast<const bf16*>(&v);
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] += bf2f(h[j]);
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) s_sum[rl * cols + 8 * cg + j] = acc[j];
  __syncthreads();
  for (int c = threadIdx.x; c < cols; c += 256) {
    float t = 0.0f;
    for (int l = 0; l < lanes; ++l) t += s_sum[l * cols + c];
    P[(size_t)blockIdx.x * cols + c] = t;
  }
}

// the same for any cols: one thread per column
__global__ void k_colsum_partial(const bf16* __restrict__ D, int rows, int cols, int rows_per,
                                 float* __restrict__ P) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  const int r0 = blockIdx.y * rows_per, r1 = min(rows, r0 + rows_per);
  float s = 0.0f;
  for (int r = r0; r < r1; ++r) s += bf2f(D[(size_t)r * cols + c]);
  P[(size_t)blockIdx.y * cols + c] = s;
}

// rows per block of the row-blocked layer kernels
constexpr int kRowBlock = 16;
// widest input and output layers the small-layer kernels keep in registers
constexpr int kMaxIO = 8;

// input layer (fin -> W): Z = X W^T + b, A = SiLU(Z); X fp32 (rounded to
// bf16 as the GEMMs' operands are), weights bf16.  A block takes kRowBlock
// rows; a thread keeps its column's fin weights in registers.
__global__ void __launch_bounds__(256) k_in_fwd(const float* __restrict__ X, int fin,
                                                const bf16* __restrict__ W,
                                                const float* __restrict__ b, int width,
                                                int rows, bf16* __restrict__ Z,
                                                bf16* __restrict__ O) {
  const int r0 = blockIdx.x * kRowBlock, r1 = min(rows, r0 + kRowBlock);
  for (int j = threadIdx.x; j < width; j += blockDim.x) {
    float w[kMaxIO];
#pragma unroll
    for (int k = 0; k < kMaxIO; ++k) w[k] = k < fin ? bf2f(W[j * fin + k]) : 0.0f;
    const float bj = b[j];
    for (int r = r0; r < r1; ++r) {
      float z = bj;
#pragma unroll
      for (int k = 0; k < kMaxIO; ++k)
        if (k < fin) z += w[k] * bf2f(f2bf(X[(size_t)r * fin + k]));
      const size_t i = (size_t)r * width + j;
      Z[i] = f2bf(z);
      O[i] = f2bf(silu(z));
    }
  }
}

// output layer (W -> fout) and MSE loss; dY = 2 (Y - T) / (rows * fout).
// Four lanes per row, each a quarter of the row's width in 16-byte loads;
// the quarters combine by lane exchange.  64 rows per block.
__global__ void __launch_bounds__(256) k_out_fwd(const bf16* __restrict__ Ain, int width,
                                                 const bf16* __restrict__ W,
                                                 const float* __restrict__ b, int fout,
                                                 int rows, float* __restrict__ Y,
                                                 const float* __restrict__ Tgt,
                                                 float* __restrict__ dY,
                                                 float* __restrict__ loss_part) {
  __shared__ float s_red[4];
  const int q = threadIdx.x & 3;
  const int r = blockIdx.x * 64 + (threadIdx.x >> 2);
  const int qw = width / 4, k0 = q * qw;
  float acc[kMaxIO];
#pragma unroll
  for (int o = 0; o < kMaxIO; ++o) acc[o] = 0.0f;
  if (r < rows) {
    const bf16* a = Ain + (size_t)r * width + k0;
    for (int k = 0; k < qw; k += 8) {
      const uint4 va = *reinterpret_cast<const uint4*>(a + k);
      const bf16* ha = reinterpret_cast<const bf16*>(&va);
#pragma unroll
      for (int o = 0; o < kMaxIO; ++o) {
        if (o >= fout) break;
        const uint4 vw = *reinterpret_cast<const uint4*>(W + (size_t)o * width + k0 + k);
        const bf16* hw = reinterpret_cast<const bf16*>(&vw);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[o] += bf2f(hw[j]) * bf2f(ha[j]);
      }
    }
  }
#pragma unroll
  for (int o = 0; o < kMaxIO; ++o) {
    acc[o] += __shfl_xor(acc[o], 1, 64);
    acc[o] += __shfl_xor(acc[o], 2, 64);
  }
  float l = 0.0f;
  if (r < rows && q == 0) {
    for (int o = 0; o < fout; ++o) {
      const float y = b[o] + acc[o];
      Y[(size_t)r * fout + o] = y;
      if (Tgt) {
        const float d = y - Tgt[(size_t)r * fout + o];
        l += d * d;
        dY[(size_t)r * fout + o] = 2.0f * d / (float)(rows * fout);
      }
    }
  }
  if (!loss_part) return;
  for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += s_red[i];
    loss_part[blockIdx.x] = s / (float)(rows * fout);
  }
}

// output-layer backward: dZ = (dY W) * SiLU'(Zin), bf16, for the last
// hidden layer.  kRowBlock rows per block, a column's fout weights in
// registers.
__global__ void __launch_bounds__(256) k_out_bwd(const float* __restrict__ dY, int fout,
                                                 const bf16* __restrict__ W, int width,
                                                 int rows, const bf16* __restrict__ Zin,
                                                 bf16* __restrict__ dZ) {
  const int r0 = blockIdx.x * kRowBlock, r1 = min(rows, r0 + kRowBlock);
  for (int k = threadIdx.x; k < width; k += blockDim.x) {
    float w[kMaxIO];
#pragma unroll
    for (int o = 0; o < kMaxIO; ++o) w[o] = o < fout ? bf2f(W[o * width + k]) : 0.0f;
    for (int r = r0; r < r1; ++r) {
      float g = 0.0f;
#pragma unroll
      for (int o = 0; o < kMaxIO; ++o)
        if (o < fout) g += dY[(size_t)r * fout + o] * w[o];
      const size_t i = (size_t)r * width + k;
      dZ[i] = f2bf(g * dsilu(bf2f(Zin[i])));
    }
  }
}

// weight and bias gradient partials of the output layer over one row chunk
// per block: P[chunk][o (width + 1) + k] = sum_r dY[r][o] A[r][k] (k <
// width), sum_r dY[r][o] (k == width).  A thread per column k of A (16-bit
// loads coalesced across the block), all fout outputs in registers.
__global__ void __launch_bounds__(256) k_wgrad_out(const float* __restrict__ dY, int fout,
                                                   const bf16* __restrict__ A, int width,
                                                   int rows, int rows_per,
                                                   float* __restrict__ P) {
  const int r0 = blockIdx.x * rows_per, r1 = min(rows, r0 + rows_per);
  float* Pc = P + (size_t)blockIdx.x * fout * (width + 1);
  for (int k = threadIdx.x; k < width; k += blockDim.x) {
    float acc[kMaxIO];
#pragma unroll
    for (int o = 0; o < kMaxIO; ++o) acc[o] = 0.0f;
    for (int r = r0; r < r1; ++r) {
      const float x = bf2f(A[(size_t)r * width + k]);
#pragma unroll
      for (int o = 0; o < kMaxIO; ++o)
        if (o < fout) acc[o] += dY[(size_t)r * fout + o] * x;
    }
    for (int o = 0; o < fout; ++o) Pc[o * (width + 1) + k] = acc[o];
  }
  for (int o = threadIdx.x; o < fout; o += blockDim.x) {
    float s = 0.0f;
    for (int r = r0; r < r1; ++r) s += dY[(size_t)r * fout + o];
    Pc[o * (width + 1) + width] = s;
  }
}

// weight and bias gradient partials of the input layer over one row chunk
// per block: P[chunk][o (fin + 1) + k] = sum_r dZ[r][o] X[r][k] (k < fin,
// X rounded to bf16 as in the forward pass), sum_r dZ[r][o] (k == fin).  A
// thread per column o of dZ.
__global__ void __launch_bounds__(256) k_wgrad_in(const bf16* __restrict__ dZ, int width,
                                                  const float* __restrict__ X, int fin,
                                                  int rows, int rows_per,
                                                  float* __restrict__ P) {
  const int r0 = blockIdx.x * rows_per, r1 = min(rows, r0 + rows_per);
  float* Pc = P + (size_t)blockIdx.x * width * (fin + 1);
  for (int o = threadIdx.x; o < width; o += blockDim.x) {
    float acc[kMaxIO + 1];
#pragma unroll
    for (int k = 0; k <= kMaxIO; ++k) acc[k] = 0.0f;
    for (int r = r0; r < r1; ++r) {
      const float d = bf2f(dZ[(size_t)r * width + o]);
#pragma unroll
      for (int k = 0; k < kMaxIO; ++k)
        if (k < fin) acc[k] += d * bf2f(f2bf(X[(size_t)r * fin + k]));
      acc[kMaxIO] += d;
    }
    for (int k = 0; k < fin; ++k) Pc[o * (fin + 1) + k] = acc[k];
    Pc[o * (fin + 1) + fin] = acc[kMaxIO];
  }
}

// Adam on FP32 master parameters, then the BF16 copy the kernels read
__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, bf16* __restrict__ pb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  pb[i] = f2bf(pi);
}

__global__ void k_to_bf16(const float* __restrict__ p, size_t n, bf16* __restrict__ pb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pb[i] = f2bf(p[i]);
}

// batch gather with a counter-based hash of (seed, step, row)
__global__ void k_gather(const float* __restrict__ X, const float* __restrict__ Y, long ns,
                         int fin, int fout, int rows, unsigned seed, unsigned step,
                         float* __restrict__ Xb, float* __restrict__ Yb) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  unsigned long long h = ((unsigned long long)seed << 32) ^ ((unsigned long long)step << 20) ^
                         (unsigned long long)r;
  h ^= h >> 33; h *= 0xff51afd7ed558ccdULL; h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ULL;
  h ^= h >> 33;
  const long idx = (long)(h % (unsigned long long)ns);
  for (int k = 0; k < fin; ++k) Xb[(size_t)r * fin + k] = X[(size_t)idx * fin + k];
  for (int k = 0; k < fout; ++k) Yb[(size_t)r * fout + k] = Y[(size_t)idx * fout + k];
}

__global__ void k_sum_loss(const float* __restrict__ part, int n, float* __restrict__ out) {
  __shared__ float s_red[4];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += s_red[i];
    *out = t;
  }
}

}  // namespace mlp
}  // namespace dgbgk

using namespace dgbgk::mlp;

namespace {
thread_local std::string g_mlp_err;
int mfail(int code, const std::string& m) {
  g_mlp_err = m;
  return code;
}
#define MCK(x)                                                                      \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return mfail(DGBGK_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));  \
  } while (0)
}  // namespace

struct dgbgk_mlp {
  int fin = 5, fout = 4, width = 256, hidden = 6, batch = 0;
  int nlayers = 0;                 // hidden + 1 linear layers
  std::vector<int> din, dout;
  std::vector<size_t> woff, boff;  // offsets of W_l [dout][din] and b_l [dout]
  size_t nparams = 0;
  hipStream_t stream = nullptr;
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // master, Adam, gradient
  bf16* pb = nullptr;                                             // BF16 parameters
  std::vector<bf16*> Z, A;         // per hidden layer, [batch][width]
  bf16 *dZa = nullptr, *dZb = nullptr;
  float *Xb = nullptr, *Yb = nullptr, *Yout = nullptr, *dY = nullptr;
  float *part = nullptr, *loss_part = nullptr, *losses = nullptr;
  size_t part_n = 0;
  float *Xall = nullptr, *Yall = nullptr;
  long nall = 0;
  long step = 0;
  std::vector<void*> owned;
};

namespace {

template <typename X>
int dalloc(dgbgk_mlp* h, X** p, size_t n) {
  MCK(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(X)));
  h->owned.push_back(*p);
  return 0;
}

// split of the batch dimension for the W x W weight gradients
int wsplits(int rows) {
  int s = std::min(64, std::max(1, rows / 1024));
  while (s > 1 && rows % (s * kKC) != 0) --s;
  return s;
}

int forward(dgbgk_mlp* h, int rows, bool with_loss) {
  const int W = h->width;
  const dim3 b256(256);
  hipLaunchKernelGGL(k_in_fwd, dim3((rows + kRowBlock - 1) / kRowBlock), b256, 0, h->stream,
                     h->Xb, h->fin, h->pb + h->woff[0], h->p + h->boff[0], W, rows, h->Z[0],
                     h->A[0]);
  for (int l = 1; l < h->hidden; ++l)
    hipLaunchKernelGGL((k_gemm<false, false, kFwd>), dim3(W / kTile, rows / kTile, 1), b256,
                       0, h->stream, h->A[l - 1], W, h->pb + h->woff[l], W, W,
                       h->p + h->boff[l], h->Z[l], h->A[l], nullptr, nullptr, W, rows);
  const int L = h->nlayers - 1;
  hipLaunchKernelGGL(k_out_fwd, dim3((rows + 63) / 64), b256, 0, h->stream,
                     h->A[h->hidden - 1], W, h->pb + h->woff[L], h->p + h->boff[L], h->fout,
                     rows, h->Yout, with_loss ? h->Yb : nullptr, h->dY,
                     with_loss ? h->loss_part : nullptr);
  MCK(hipGetLastError());
  return 0;
}

// gradient of the mean squared error into h->g (packed like the parameters)
int backward(dgbgk_mlp* h, int rows) {
  const int W = h->width, L = h->nlayers - 1, H = h->hidden;
  const dim3 b256(256);
  const int rchunk = kRowChunk, nch = (rows + rchunk - 1) / rchunk;
  // output layer: weights and bias
  {
    const int n = h->fout * (W + 1);
    hipLaunchKernelGGL(k_wgrad_out, dim3(nch), b256, 0, h->stream, h->dY, h->fout, h->A[H - 1],
                       W, rows, rchunk, h->part);
    hipLaunchKernelGGL(k_reduce_splits, dim3((n + kRedCols - 1) / kRedCols), b256, 0, h->stream,
                       h->part, nch, n, W, h->g + h->woff[L], h->g + h->boff[L]);
    hipLaunchKernelGGL(k_out_bwd, dim3((rows + kRowBlock - 1) / kRowBlock), b256, 0, h->stream,
                       h->dY, h->fout, h->pb + h->woff[L], W, rows, h->Z[H - 1], h->dZa);
  }
  bf16* dz = h->dZa;
  bf16* dzn = h->dZb;
  const int sp = wsplits(rows);
  for (int l = H - 1; l >= 1; --l) {
    // dW_l[out][in] = sum_r dZ_l[r][out] A_{l-1}[r][in]   (split over rows)
    hipLaunchKernelGGL((k_gemm<true, true, kPartial>), dim3(W / kTile, W / kTile, sp), b256, 0,
                       h->stream, dz, W, h->A[l - 1], W, rows, nullptr, nullptr, nullptr,
                       nullptr, h->part, W, W);
    const int nw = W * W;
    hipLaunchKernelGGL(k_reduce_splits, dim3((nw + kRedCols - 1) / kRedCols), b256, 0,
                       h->stream, h->part, sp, nw, 0, h->g + h->woff[l], nullptr);
    // db_l = column sums of dZ_l
    if (256 % (W / 8) == 0)
      hipLaunchKernelGGL(k_colsum8, dim3(nch), b256, 0, h->stream, dz, rows, W, rchunk, h->part);
    else
      hipLaunchKernelGGL(k_colsum_partial, dim3((W + 255) / 256, nch), b256, 0, h->stream, dz,
                         rows, W, rchunk, h->part);
    hipLaunchKernelGGL(k_reduce_splits, dim3((W + kRedCols - 1) / kRedCols), b256, 0, h->stream,
                       h->part, nch, W, 0, h->g + h->boff[l], nullptr);
    // dZ_{l-1} = (dZ_l W_l) * SiLU'(Z_{l-1})
    hipLaunchKernelGGL((k_gemm<false, true, kBwdData>), dim3(W / kTile, rows / kTile, 1), b256,
                       0, h->stream, dz, W, h->pb + h->woff[l], W, W, nullptr, nullptr, dzn,
                       h->Z[l - 1], nullptr, W, rows);
    std::swap(dz, dzn);
  }
  // input layer
  {
    const int n = W * (h->fin + 1);
    hipLaunchKernelGGL(k_wgrad_in, dim3(nch), b256, 0, h->stream, dz, W, h->Xb, h->fin, rows,
                       rchunk, h->part);
    hipLaunchKernelGGL(k_reduce_splits, dim3((n + kRedCols - 1) / kRedCols), b256, 0, h->stream,
                       h->part, nch, n, h->fin, h->g + h->woff[0], h->g + h->boff[0]);
  }
  MCK(hipGetLastError());
  return 0;
}

int adam(dgbgk_mlp* h, float lr) {
  h->step += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)h->step), c2 = 1.0f - std::pow(b2, (float)h->step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((h->nparams + 255) / 256)), dim3(256), 0, h->stream,
                     h->p, h->m, h->v, h->g, h->nparams, lr, b1, b2, eps, c1, c2, h->pb);
  MCK(hipGetLastError());
  return 0;
}

int check_rows(dgbgk_mlp* h, int rows) {
  if (rows <= 0 || rows > h->batch || rows % kTile != 0)
    return mfail(DGBGK_ERR_ARG, "rows must be a positive multiple of 64 and <= batch");
  return 0;
}

}  // namespace

extern "C" {

const char* dgbgk_mlp_last_error(void) { return g_mlp_err.c_str(); }

int dgbgk_mlp_create(int fin, int width, int hidden, int fout, int batch, dgbgk_mlp** out) {
  if (!out || fin < 1 || fin > kMaxIO || fout < 1 || fout > kMaxIO || hidden < 2 ||
      width % kTile != 0 || width < kTile ||
      batch < kTile || batch % kTile != 0)
    return mfail(DGBGK_ERR_ARG, "bad MLP shape (1..8 inputs and outputs, width and batch "
                                "multiples of 64, hidden >= 2)");
  *out = nullptr;
  auto* h = new (std::nothrow) dgbgk_mlp;
  if (!h) return mfail(DGBGK_ERR_NOMEM, "out of host memory");
  h->fin = fin; h->width = width; h->hidden = hidden; h->fout = fout; h->batch = batch;
  h->nlayers = hidden + 1;
  for (int l = 0; l < h->nlayers; ++l) {
    const int di = l == 0 ? fin : width, dn = l == h->nlayers - 1 ? fout : width;
    h->din.push_back(di);
    h->dout.push_back(dn);
    h->woff.push_back(h->nparams);
    h->nparams += (size_t)di * dn;
    h->boff.push_back(h->nparams);
    h->nparams += dn;
  }
  // error path of create: release what was made, report the first error
  auto bail = [&](int code) {
    for (void* q : h->owned) (void)hipFree(q);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return code;
  };
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return bail(mfail(DGBGK_ERR_HIP, "no HIP device available"));
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(mfail(DGBGK_ERR_HIP, "hipStreamCreate failed"));
  const size_t act = (size_t)batch * width;
  int st = 0;
  st |= dalloc(h, &h->p, h->nparams);
  st |= dalloc(h, &h->m, h->nparams);
  st |= dalloc(h, &h->v, h->nparams);
  st |= dalloc(h, &h->g, h->nparams);
  st |= dalloc(h, &h->pb, h->nparams);
  h->Z.assign(hidden, nullptr);
  h->A.assign(hidden, nullptr);
  for (int l = 0; l < hidden; ++l) {
    st |= dalloc(h, &h->Z[l], act);
    st |= dalloc(h, &h->A[l], act);
  }
  st |= dalloc(h, &h->dZa, act);
  st |= dalloc(h, &h->dZb, act);
  st |= dalloc(h, &h->Xb, (size_t)batch * fin);
  st |= dalloc(h, &h->Yb, (size_t)batch * fout);
  st |= dalloc(h, &h->Yout, (size_t)batch * fout);
  st |= dalloc(h, &h->dY, (size_t)batch * fout);
  const int nch = (batch + kRowChunk - 1) / kRowChunk;
  h->part_n = std::max({(size_t)64 * width * width, (size_t)nch * width * (fin + 1),
                        (size_t)nch * fout * (width + 1), (size_t)nch * width});
  st |= dalloc(h, &h->part, h->part_n);
  st |= dalloc(h, &h->loss_part, (size_t)(batch + 63) / 64);
  st |= dalloc(h, &h->losses, 1);
  if (st) return bail(st);
  if (hipMemset(h->m, 0, h->nparams * sizeof(float)) != hipSuccess ||
      hipMemset(h->v, 0, h->nparams * sizeof(float)) != hipSuccess)
    return bail(mfail(DGBGK_ERR_HIP, "hipMemset failed"));
  *out = h;
  return 0;
}

void dgbgk_mlp_destroy(dgbgk_mlp* h) {
  if (!h) return;
  // teardown has no caller to report to: every release is attempted
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void* q : h->owned) (void)hipFree(q);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int dgbgk_mlp_nparams(dgbgk_mlp* h, long* n) {
  if (!h || !n) return mfail(DGBGK_ERR_ARG, "null argument");
  *n = (long)h->nparams;
  return 0;
}

int dgbgk_mlp_set_params(dgbgk_mlp* h, const float* params) {
  if (!h || !params) return mfail(DGBGK_ERR_ARG, "null argument");
  MCK(hipMemcpyAsync(h->p, params, h->nparams * sizeof(float), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((h->nparams + 255) / 256)), dim3(256), 0,
                     h->stream, h->p, h->nparams, h->pb);
  MCK(hipMemsetAsync(h->m, 0, h->nparams * sizeof(float), h->stream));
  MCK(hipMemsetAsync(h->v, 0, h->nparams * sizeof(float), h->stream));
  h->step = 0;
  MCK(hipStreamSynchronize(h->stream));
  return 0;
}

int dgbgk_mlp_get_params(dgbgk_mlp* h, float* params) {
  if (!h || !params) return mfail(DGBGK_ERR_ARG, "null argument");
  MCK(hipMemcpyAsync(params, h->p, h->nparams * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  MCK(hipStreamSynchronize(h->stream));
  return 0;
}

// y[n][fout] = network(x[n][fin]) for any n (rows padded to multiples of 64)
int dgbgk_mlp_forward(dgbgk_mlp* h, long n, const float* x, float* y) {
  if (!h || !x || !y || n < 0) return mfail(DGBGK_ERR_ARG, "bad arguments");
  for (long r0 = 0; r0 < n; r0 += h->batch) {
    const int rows = (int)std::min<long>(h->batch, n - r0);
    const int padded = (rows + kTile - 1) / kTile * kTile;
    MCK(hipMemsetAsync(h->Xb, 0, (size_t)padded * h->fin * sizeof(float), h->stream));
    MCK(hipMemcpyAsync(h->Xb, x + (size_t)r0 * h->fin, (size_t)rows * h->fin * sizeof(float),
                       hipMemcpyHostToDevice, h->stream));
    int st = forward(h, padded, false);
    if (st) return st;
    MCK(hipMemcpyAsync(y + (size_t)r0 * h->fout, h->Yout, (size_t)rows * h->fout * sizeof(float),
                       hipMemcpyDeviceToHost, h->stream));
  }
  MCK(hipStreamSynchronize(h->stream));
  return 0;
}

// one batch: loss = mean (net(x) - y)^2 and its gradient (packed like the
// parameters; grad may be NULL)
int dgbgk_mlp_loss_grad(dgbgk_mlp* h, int rows, const float* x, const float* y, float* loss,
                        float* grad) {
  if (!h || !x || !y) return mfail(DGBGK_ERR_ARG, "null argument");
  int st = check_rows(h, rows);
  if (st) return st;
  MCK(hipMemcpyAsync(h->Xb, x, (size_t)rows * h->fin * sizeof(float), hipMemcpyHostToDevice,
                     h->stream));
  MCK(hipMemcpyAsync(h->Yb, y, (size_t)rows * h->fout * sizeof(float), hipMemcpyHostToDevice,
                     h->stream));
  if ((st = forward(h, rows, true))) return st;
  if ((st = backward(h, rows))) return st;
  hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, h->stream, h->loss_part,
                     (rows + 63) / 64, h->losses);
  if (loss)
    MCK(hipMemcpyAsync(loss, h->losses, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (grad)
    MCK(hipMemcpyAsync(grad, h->g, h->nparams * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  MCK(hipStreamSynchronize(h->stream));
  return 0;
}

// training set kept on the device: x [n][fin], y [n][fout]
int dgbgk_mlp_set_data(dgbgk_mlp* h, long n, const float* x, const float* y) {
  if (!h || !x || !y || n < 1) return mfail(DGBGK_ERR_ARG, "bad arguments");
  // the previous training set, replaced
  if (h->Xall) {
    MCK(hipFree(h->Xall));
    h->owned.erase(std::find(h->owned.begin(), h->owned.end(), (void*)h->Xall));
  }
  if (h->Yall) {
    MCK(hipFree(h->Yall));
    h->owned.erase(std::find(h->owned.begin(), h->owned.end(), (void*)h->Yall));
  }
  h->Xall = h->Yall = nullptr;
  int st = dalloc(h, &h->Xall, (size_t)n * h->fin);
  st |= dalloc(h, &h->Yall, (size_t)n * h->fout);
  if (st) return st;
  MCK(hipMemcpy(h->Xall, x, (size_t)n * h->fin * sizeof(float), hipMemcpyHostToDevice));
  MCK(hipMemcpy(h->Yall, y, (size_t)n * h->fout * sizeof(float), hipMemcpyHostToDevice));
  h->nall = n;
  return 0;
}

// `steps` Adam steps on batches of `batch` rows drawn from the training set;
// losses (may be NULL) gets the loss of every step
int dgbgk_mlp_train(dgbgk_mlp* h, int steps, float lr, unsigned seed, float* losses) {
  if (!h || steps < 0 || h->nall < 1) return mfail(DGBGK_ERR_ARG, "no training data");
  const int rows = h->batch;
  float* dl = nullptr;
  if (losses) {
    MCK(hipMalloc((void**)&dl, std::max(steps, 1) * sizeof(float)));
  }
  int st = 0;
  for (int s = 0; s < steps && !st; ++s) {
    hipLaunchKernelGGL(k_gather, dim3((rows + 255) / 256), dim3(256), 0, h->stream, h->Xall,
                       h->Yall, h->nall, h->fin, h->fout, rows, seed, (unsigned)(h->step + 1),
                       h->Xb, h->Yb);
    st = forward(h, rows, true);
    if (!st) st = backward(h, rows);
    if (!st && dl)
      hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, h->stream, h->loss_part,
                         (rows + 63) / 64, dl + s);
    if (!st) st = adam(h, lr);
  }
  if (!st && dl) {
    hipError_t e = hipMemcpyAsync(losses, dl, steps * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) st = mfail(DGBGK_ERR_HIP, "loss copy failed");
  }
  hipError_t e2 = hipStreamSynchronize(h->stream);
  const hipError_t e3 = dl ? hipFree(dl) : hipSuccess;
  if (st) return st;
  if (e2 != hipSuccess) return mfail(DGBGK_ERR_HIP, hipGetErrorString(e2));
  if (e3 != hipSuccess) return mfail(DGBGK_ERR_HIP, hipGetErrorString(e3));
  return 0;
}

}  // extern "C"
