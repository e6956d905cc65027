// dgbgk.hip -- HIP kernels (gfx950) and the C ABI of the DG-BGK time march.
//
// One step (src/advnce.f90:220-449) is a fixed sequence of launches on the
// context's stream, with no host synchronisation:
//   k_eta      FLUCON2 wall mass-flux parameter per type-4 side
//   k_update   all ordinates of all elements: GTEQNF + GETELC + EDGFLXOPT +
//              GETINC + ADTHEM + INFLOW, element tiles staged in LDS
//   k_wall     in-place sweep order: reflected wall fluxes, ordinate by
//              ordinate per (wall element, speed ring)
//   k_moments  MACROS + COMPMU on f^{k+1}: FP64 compensated sums
//   k_points   point averages (GETMAC) when the reference gathers them
//   k_final    residual max, mass, forces, next DTE (ALOTIM2)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dgbgk.h"
#include "kernels.hpp"
#include "prep.hpp"

namespace dgbgk {

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// The hardware hands consecutive workgroup ids to the 8 XCDs in turn.  Map
// them back so each XCD works on one contiguous run of logical blocks: the
// tiles of a run are mesh neighbours and share halo data in that XCD's L2.
__device__ __forceinline__ int xcd_logical(int bid, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7;
  const int xcd = bid & 7, loc = bid >> 3;
  return xcd * q + min(xcd, r) + loc;
}


// Unbiased rounding of an FP32 store f + d: d gains a dither uniform over
// +-1/2 ulp of f before the one rounding, so the stored value rounds up
// with probability equal to the exact sum's fraction of an ulp.  Below
// about an ulp per step a plain rounded store drops the increment, and a
// run near steady state stalls at a state biased by ~ulp / (decay rate per
// step).  The dither fraction r in [0, 1) runs through a Weyl sequence per
// element (element, ordinate and step phases summed mod 1), equidistributed
// over steps.  f = 0 and denormals have no exponent bits and get no dither, so
// clamped nodes stay exactly zero.
// The dither is (r - 1/2) 2^-23 times 2^exponent(f), i.e. (r - 1/2) ulp(f):
// t = fma(r, 2^-23, -2^-24) once per element and ordinate, then per node
// d + t * (f with its mantissa bits cleared).

// The update pass's f store, non-temporal: the written rows take no L2
// lines, so the lines a neighbouring tile still has to read as its halo
// stay (the update's counted reads fall from 1.38x to 1.27x of its
// algorithmic bytes in FP32, 1.44x to 1.34x in FP64), and the moments
// pass that follows starts without dirty lines to write back (15% faster
// in FP32, 10% in FP64).  The wall pass's scattered stores stay plain
// (non-temporal ones made it slower).
template <typename T>
__device__ __forceinline__ void fstore_nt(T* p, T x) {
  __builtin_nontemporal_store(x, p);
}

// The moments pass's f loads, non-temporal: f^{k+1} is read once there
// (3% faster on 1M triangles in both forms).  The update kernel's loads stay
// plain: its halo elements are read again by the neighbouring tile, and
// non-temporal loads made it 5-15% slower.
template <typename T>
__device__ __forceinline__ T fload_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}

// The increment of a node and its residual term, mdt = M DTE.  IMMAT = 1,
// GETINC (src/getinc.f90:34-41): d = M RHS DTE + COLL DTE, residual |d|.
// IMMAT = 0, CMMINC as the reference writes it (src/cmminc.f90:30-39): each
// node adds C(k,j) RHS(k) DTE over j, C the element's CONMAT matrix, whose
// row sum is M = 6/RJ (diagonal 3M, off-diagonal -M), so d has the same
// form; the residual is the largest signed term C(k,j) RHS DTE + COLL DTE,
// i.e. 3 M RHS DTE + COLL DTE or -M RHS DTE + COLL DTE, from 0 (no
// absolute value).  The reference sums the three products one by one; here
// the row sum is taken first (a rounding-level difference).
template <typename T, typename R>
__device__ __forceinline__ T node_increment(bool cmminc, T mdt, T rhs, T cold, R& res) {
  const T d = mdt * rhs + cold;
  if (!cmminc)
    res = fmax(res, (R)fabs(d));
  else
    res = fmax(res, (R)fmax((T(3) * mdt) * rhs + cold, cold - mdt * rhs));
  return d;
}

__device__ __forceinline__ float dexp(float x) { return __expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }

// Maxwellian exponential with beta^2 pre-scaled by bscale<T>(): FP32 takes
// the hardware base-2 exponential (beta^2 log2(e) folded in once per block),
// FP64 the correctly rounded exp.
template <typename T>
__device__ __forceinline__ T bscale();
template <>
__device__ __forceinline__ float bscale<float>() { return 1.4426950408889634f; }
template <>
__device__ __forceinline__ double bscale<double>() { return 1.0; }
__device__ __forceinline__ float mexp(float x) { return __builtin_amdgcn_exp2f(x); }
// FP64: x = n ln2 + r with |r| <= ln2/2 (Cody-Waite reduction, ln2 split
// into a head with its low 21 mantissa bits clear, so n * head is exact,
// and a tail), exp(r) by its Taylor polynomial of degree 13 (truncation
// below 6e-18 relative on that interval), times 2^n by ldexp, which gives
// 0 below the denormals and inf above the range.  No branches: the
// Maxwellian's exponent is never positive, and NaN propagates.
__device__ __forceinline__ double mexp(double x) {
  const double n = rint(x * 1.4426950408889634);
  const double r = fma(-n, 1.9082149292705877e-10, fma(-n, 0.6931471803691238, x));
  double p = 1.6059043836821613e-10;  // 1/13!
  p = fma(p, r, 2.08767569878681e-09);
  p = fma(p, r, 2.505210838544172e-08);
  p = fma(p, r, 2.755731922398589e-07);
  p = fma(p, r, 2.7557319223985893e-06);
  p = fma(p, r, 2.48015873015873e-05);
  p = fma(p, r, 1.984126984126984e-04);
  p = fma(p, r, 1.388888888888889e-03);
  p = fma(p, r, 8.333333333333333e-03);
  p = fma(p, r, 4.1666666666666664e-02);
  p = fma(p, r, 1.6666666666666666e-01);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)n);
}

// Loads from tables the kernel never writes, through the constant address
// space: a uniform index then becomes one scalar load.
template <typename T>
__device__ __forceinline__ T ldc(const T* p, int i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

template <typename T>
__device__ __forceinline__ void edge_rhsi(T ux, T uy, T nx, T ny, T alen, T g1, T g2,
                                          T& r1, T& r2) {
  // normal flux of the upwind values, times the 1-D P1 mass matrix
  // [[2,1],[1,2]] * length/6 (src/edgflxopt.f90:108-120)
  const T fn1 = nx * (ux * g1) + ny * (uy * g1);
  const T fn2 = nx * (ux * g2) + ny * (uy * g2);
  r1 = -(fn1 * T(2) * alen) - fn2 * alen;
  r2 = -(fn1 * alen) - fn2 * T(2) * alen;
}

__device__ __forceinline__ double block_max(double x, double* sh) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_down(x, o, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 1; i < nw; ++i) x = fmax(x, sh[i]);
  }
  return x;  // valid in thread 0
}

__device__ __forceinline__ double block_sum(double x, double* sh) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 1; i < nw; ++i) x += sh[i];
  }
  return x;
}

// Kahan compensated accumulation (FP32-bulk moments); the FP64 path sums
// plainly in ordinate order, as the reference does.
template <bool kComp>
__device__ __forceinline__ void nsum(double& s, double& c, double x) {
  if (kComp) {
    const double y = x - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  } else {
    s += x;
  }
}

// ---------------------------------------------------------------------------
// k_update: one block = one tile of <= 256 elements (owned + halo) for a
// chunk of ordinates.  Per ordinate: phase A computes every slot's predictor
// (GTEQNF + GETELC's UMEAN, src/getelc.f90:54-83) and stages f and U in
// LDS; phase B gives each owned element its volume term, its upwind edge
// fluxes (element-centric EDGFLXOPT: each element recomputes the flux of its
// own sides, so there is no scatter and no atomic), GETINC, ADTHEM and
// INFLOW.  Wall sides are left to k_wall: wall elements hand their partial
// right-hand side to it through wpart.
// ---------------------------------------------------------------------------

// What one slot keeps in registers over all ordinates of its block.
template <typename T>
struct SlotState {
  T gx[3], gy[3], pA[3], pB[3], pU[3], pV[3], pMd[3];  // phase A (all slots)
  T gxa[3], gya[3], mdt, nx[3], ny[3];                 // phase B (owned slots)
  int info[3], mask, wsl, e;
  int nb1[3], nb2[3];  // LDS index (4 slot + node) of the neighbour's values at k, k1
  float dph;           // FP32 store dither: the element's fixed phase
  bool owned, active;
};

template <typename T>
__device__ __forceinline__ void store4(T* p, T a, T b, T c, T d);
template <>
__device__ __forceinline__ void store4<float>(float* p, float a, float b, float c, float d) {
  *reinterpret_cast<float4*>(p) = make_float4(a, b, c, d);
}
template <>
__device__ __forceinline__ void store4<double>(double* p, double a, double b, double c,
                                                   double d) {
  reinterpret_cast<double2*>(p)[0] = make_double2(a, b);
  reinterpret_cast<double2*>(p)[1] = make_double2(c, d);
}

// One ordinate v of one slot.  f: the slot's three node values at v; sf:
// this ordinate's LDS buffer, [slot][4] = g = f - U at the three nodes and
// a zero.  ph: the ordinate's and step's dither phase, dsc0/dsc1 the
// dither's scale (zero without it).
// Contains the block's one barrier per ordinate, so every thread of the
// block calls it for every v.
template <typename T, bool kMom, bool kCM>
__device__ __forceinline__ void update_ordinate(const UpdArgs<T>& a, const SlotState<T>& s,
                                                int v, T ux, T uy, T fin, float ph,
                                                float dsc0, float dsc1,
                                                const T (&f)[3],
                                                T* __restrict__ sf, T dte, T& res, T& chg,
                                                double wv, double wcx, double wcy,
                                                double (&m0)[3], double (&m1)[3],
                                                double (&m2)[3]) {
  const int slot = threadIdx.x;
  const T c13 = T(1) / T(3);
  // collision term times DTE: pMd = MU * DTE (src/getinc.f90:34-41)
  // phase A runs on every slot without a branch: slots without an element
  // hold zero parameters and element 0's f, and nobody reads their result
  T cold[3], h[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T dx = s.pU[k] - ux, dy = s.pV[k] - uy;
    const T feq = s.pA[k] * mexp(-s.pB[k] * (dx * dx + dy * dy));
    cold[k] = s.pMd[k] * (feq - f[k]);
  }
  const T cs = cold[0] + cold[1] + cold[2];
  T rhs[3], U;
  if constexpr (sizeof(T) == 8) {
    // GETELC's predictor: U = DTE/2 (div(c f) - (1/3) sum MU (f_eq - f))
    const T sx = s.gx[0] * f[0] + s.gx[1] * f[1] + s.gx[2] * f[2];
    const T sy = s.gy[0] * f[0] + s.gy[1] * f[1] + s.gy[2] * f[2];
    U = T(0.5) * ((ux * sx + uy * sy) * dte - c13 * cs);
#pragma unroll
    for (int k = 0; k < 3; ++k) h[k] = f[k] - U;
    store4(sf + 4 * slot, h[0], h[1], h[2], T(0));
    __syncthreads();
    if (!s.owned) return;
    // volume term (src/getelc.f90:87-117): area * grad N_a . c (mean of f - U)
    const T sm = c13 * (h[0] + h[1] + h[2]);
    const T fx = ux * sm, fy = uy * sm;
#pragma unroll
    for (int k = 0; k < 3; ++k) rhs[k] = s.gxa[k] * fx + s.gya[k] * fy;
    // edges (src/edgflxopt.f90:72-125, src/getbou.f90:142-187).  Local edge k
    // runs from node k to node k1 = (k+1)%3 with this element's outward
    // normal n; its flux adds -(len/6) (c.n) M2 (g_k, g_k1) with
    // M2 = [[2,1],[1,2]] and g = f - U of the upwind element.  The reference
    // orients each side once (left element) and subtracts the left element's
    // flux from the right element; with the right element's outward normal
    // that is the same expression.  Upwind: own values when c.n > 0 (on a
    // tie the right element's values, as in the reference), always on a
    // boundary side; inflow and wall sides carry a zero (len/6) n here.
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int k1 = (k + 1) % 3;
      const int in = s.info[k];
      // w = -(len/6) c.n from the edge's (len/6) n: c.n > 0 exactly when w < 0
      const T w = -(ux * s.nx[k] + uy * s.ny[k]);
      const bool own = (in & kUpdOwn) || w < T(0) || ((in & kUpdRight) && w == T(0));
      // the neighbour's values are read whichever side is upwind (fixed LDS
      // indices, no address arithmetic); the element's own are in registers
      const T n1 = sf[s.nb1[k]], n2 = sf[s.nb2[k]];
      const T g1 = own ? h[k] : n1, g2 = own ? h[k1] : n2;
      rhs[k] += w * (T(2) * g1 + g2);
      rhs[k1] += w * (g1 + T(2) * g2);
    }
  } else {
    // FP32: the same operator in exact arithmetic, evaluated on differences
    // of nearby values so that FP32 rounds the small terms, not f.
    // Predictor: the gradients sum to zero, so div(c f) takes f - f_0.
    const T d1 = f[1] - f[0], d2 = f[2] - f[0];
    const T sx = s.gx[1] * d1 + s.gx[2] * d2;
    const T sy = s.gy[1] * d1 + s.gy[2] * d2;
    U = T(0.5) * ((ux * sx + uy * sy) * dte - c13 * cs);
    // neighbours read raw f and U: each reader forms differences from its
    // own values before anything is rounded at the size of f
    store4(sf + 4 * slot, f[0], f[1], f[2], U);
    __syncthreads();
    if (!s.owned) return;
    // area grad N_k is half the sum of len n over the element's edges at
    // node k, so the volume term folds into the edges: each edge acts on
    // (upwind g - the element's mean gbar of f - U), the mean about f_0 as
    // gbar = f_0 + db - U with db the mean of (0, d1, d2)
    const T db = c13 * (d1 + d2);
    h[0] = -db;
    h[1] = d1 - db;
    h[2] = d2 - db;
    const T gbar = (f[0] + db) - U;  // an inflow side's upwind values are 0
#pragma unroll
    for (int k = 0; k < 3; ++k) rhs[k] = T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int k1 = (k + 1) % 3;
      const int in = s.info[k];
      const T w = -(ux * s.nx[k] + uy * s.ny[k]);
      const bool own = (in & kUpdOwn) || w < T(0) || ((in & kUpdRight) && w == T(0));
      const T n1 = sf[s.nb1[k]], n2 = sf[s.nb2[k]], un = sf[s.nb1[k] | 3];
      const T du = un - U;
      T e1 = ((n1 - f[0]) - db) - du, e2 = ((n2 - f[0]) - db) - du;
      // every lane forms the neighbour's values (the reads are in bounds
      // for all slots), so the upwind choice below is a select, not a
      // divergent branch around the reads per edge
      asm volatile("" : "+v"(e1), "+v"(e2));
      const bool inflow = (in & kUpdInflow) != 0;
      const T g1 = own ? (inflow ? -gbar : h[k]) : e1;
      const T g2 = own ? (inflow ? -gbar : h[k1]) : e2;
      rhs[k] += w * (T(2) * g1 + g2);
      rhs[k1] += w * (g1 + T(2) * g2);
    }
  }
  if (s.wsl >= 0) {
    T* wp = a.wpart + ((size_t)s.wsl * a.V + v) * 8;
    store4(wp, rhs[0], rhs[1], rhs[2], cold[0]);
    store4(wp + 4, cold[1], cold[2], U, T(0));
    return;
  }
  // GETINC + ADTHEM + INFLOW (src/getinc.f90:30-42, src/adthem.f90:17-26,
  // src/inflow.f90:43-59); mdt = 6/RJ * DTE
  T* fo = a.fnxt + (size_t)v * a.ldf + 3 * (size_t)s.e;
  T d[3], fn[3], cg[3];
  // FP32: the dithered store with one dither fraction for the element's
  // three nodes (a shared fraction leaves each node's rounding unbiased);
  // dsc = (2^-23, -2^-24), or (0, 0) without the dither, which then adds
  // exactly 0.  The change is the step's own increment, not what the
  // rounding adds or drops: |d| where the store was not clamped (clamped:
  // f), and the reset below sets inflow nodes' |f_inf - f|.
  float t = 0.f;
  if constexpr (sizeof(T) == 4) t = fmaf(__builtin_amdgcn_fractf(s.dph + ph), dsc0, dsc1);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    d[k] = node_increment(kCM, s.mdt, rhs[k], cold[k], res);
    T dk = d[k];
    if constexpr (sizeof(T) == 4)
      dk = fmaf(t, __uint_as_float(__float_as_uint(f[k]) & 0x7f800000u), dk);
    fn[k] = f[k] + dk;
    const bool neg = fn[k] < T(0);
    fn[k] = neg ? T(0) : fn[k];  // NaN stays NaN, as in the reference
    if constexpr (sizeof(T) == 4) cg[k] = neg ? f[k] : T(fabs(d[k]));
  }
  if (__ballot(s.mask != 0)) {  // inflow nodes: rare; a wave-uniform branch
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if ((s.mask >> k) & 1) {
        fn[k] = fin;
        if constexpr (sizeof(T) == 4) cg[k] = fabs(fin - f[k]);
      }
  }
  if constexpr (sizeof(T) == 8) {  // FP64: the change as stored
#pragma unroll
    for (int k = 0; k < 3; ++k) cg[k] = fabs(fn[k] - f[k]);
  }
  chg = fmax(chg, fmax(fmax(cg[0], cg[1]), cg[2]));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    fstore_nt(fo + k, fn[k]);
    if (kMom) {
      // MACROS' first pass (src/macros.f90:93-108) on f^{k+1}, this chunk's
      // ordinates, in FP64 (wcx = w cx, wcy = w cy; k_moments sums the same
      // products in the same order)
      const double fd = (double)fn[k];
      m0[k] = fma(wv, fd, m0[k]);
      m1[k] = fma(wcx, fd, m1[k]);
      m2[k] = fma(wcy, fd, m2[k]);
    }
  }
}

// kDepth: ordinates whose f is in flight ahead of the one being computed.
// kDepth + 1 register slots are indexed by compile-time positions of an
// unrolled loop, so no register copy ever waits on a load still in flight.
// kMom: also accumulate the first moment pass of f^{k+1} (n, n u, n v per
// node over the block's ordinates) into msum.
// kS: slots (threads) per block, the tile size the host built for.
// Minimum wavefronts per SIMD the FP32 update kernel is compiled for (1:
// no constraint, the compiler's own register allocation).
#ifndef DGBGK_FP32_MIN_WAVES
#define DGBGK_FP32_MIN_WAVES 1
#endif
template <typename T>
constexpr int kMinWaves = sizeof(T) == 4 ? DGBGK_FP32_MIN_WAVES : 1;

// kCM: IMMAT = 0 (the consistent-mass branch): CMMINC's increment and
// residual (node_increment), launched at prefetch depth 1.
template <typename T, int kDepth, bool kMom, int kS, bool kCM>
__global__ void __launch_bounds__(kS) __attribute__((amdgpu_waves_per_eu(kMinWaves<T>, 8)))
k_update(UpdArgs<T> a) {
  constexpr int kQ = kDepth + 1;
  // one LDS buffer per queue position: ordinate v + kQ reuses v's buffer
  // after kQ - 1 >= 1 barriers, and the buffer is a compile-time offset
  __shared__ __align__(16) T s_f[kQ][kS * 4];
  __shared__ double s_red[8];
  const int logical = xcd_logical(blockIdx.x, gridDim.x);
  const int tile = logical % a.ntiles;
  const int v0 = (logical / a.ntiles) * a.vchunk;
  const int v1 = min(a.V, v0 + a.vchunk);
  const int e0 = a.tiles[4 * tile], nown = a.tiles[4 * tile + 1];
  const int hoff = a.tiles[4 * tile + 2], nh = a.tiles[4 * tile + 3];
  const int slot = threadIdx.x;
  const size_t N3 = 3 * (size_t)a.E;
  const int E = a.E, V = a.V;

  SlotState<T> s;
  s.owned = slot < nown;
  s.active = slot < nown + nh;
  s.e = s.owned ? e0 + slot : (s.active ? a.halo[hoff + slot - nown] : 0);
  const int e = s.e;
  const T dte = (T)a.scal[0];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s.gx[k] = s.gy[k] = s.pA[k] = s.pB[k] = s.pU[k] = s.pV[k] = s.pMd[k] = T(0);
    s.gxa[k] = s.gya[k] = s.nx[k] = s.ny[k] = T(0);
    s.info[k] = kUpdOwn;
    s.nb1[k] = 4 * slot + k;
    s.nb2[k] = 4 * slot + (k + 1) % 3;

  }
  s.mdt = T(0);
  s.dph = 0.f;
  s.mask = 0;
  s.wsl = -1;
  if (s.active) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s.gx[k] = a.geo[k * E + e];
      s.gy[k] = a.geo[(3 + k) * E + e];
      const size_t nd = 3 * (size_t)e + k;
      s.pA[k] = a.np[nd];
      s.pB[k] = a.np[N3 + nd] * bscale<T>();
      s.pU[k] = a.np[2 * N3 + nd];
      s.pV[k] = a.np[3 * N3 + nd];
      s.pMd[k] = a.np[4 * N3 + nd] * dte;
    }
  }
  if (s.owned) {
    const T area = a.geo[6 * E + e];
    s.mdt = a.geo[7 * E + e] * dte;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s.gxa[k] = s.gx[k] * area;
      s.gya[k] = s.gy[k] * area;
      s.info[k] = a.uinfo[k * E + e];
      if (!(s.info[k] & kUpdOwn)) {  // interior side: the neighbour's slot and nodes
        const int nbs = s.info[k] >> kUpdSlot;
        s.nb1[k] = 4 * nbs + ((s.info[k] >> kUpdNbK) & 3);
        s.nb2[k] = 4 * nbs + ((s.info[k] >> kUpdNbK1) & 3);
      }
      s.nx[k] = a.unx[k * E + e];
      s.ny[k] = a.uny[k * E + e];
    }
    s.mask = a.emask[e];
    s.wsl = a.ewslot[e];
    s.dph = a.ephase[e];
  }
  T res = 0, chg = 0;

  // prime the queue: f and the ordinate's velocity and inflow value for
  // v0 .. v0+kDepth-1
  T q[kQ][3], qx[kQ], qy[kQ], qi[kQ];
  float qp[kQ];  // FP32 store dither: the ordinate's phase plus this step's
  double qw[kQ], qcx[kQ], qcy[kQ];
  const double stw = a.scal[6] * 0.6180339887498949;  // Weyl step phase
  const float sph = (float)(stw - floor(stw));
  const float dsc0 = a.dither ? 0x1p-23f : 0.f, dsc1 = a.dither ? -0x1p-24f : 0.f;
  const T* fe = a.fcur + 3 * (size_t)e;
#pragma unroll
  for (int j = 0; j < kQ; ++j) {
    q[j][0] = q[j][1] = q[j][2] = qx[j] = qy[j] = qi[j] = T(0);
    qp[j] = 0.f;
    qw[j] = qcx[j] = qcy[j] = 0.0;
    if (j < kDepth) {  // slots without an element read element 0
      const int v = min(v0 + j, v1 - 1);
      const T* fp = fe + (size_t)v * a.ldf;
      q[j][0] = fp[0];
      q[j][1] = fp[1];
      q[j][2] = fp[2];
      qx[j] = ldc(a.vtab, v);
      qy[j] = ldc(a.vtab, V + v);
      qi[j] = ldc(a.vtab, 2 * V + v);
      qp[j] = sph + (float)ldc(a.vtab, 3 * V + v);
      if (kMom) {
        qw[j] = ldc(a.vtd, 4 * V + v);
        qcx[j] = ldc(a.vtd, 5 * V + v);
        qcy[j] = ldc(a.vtd, 6 * V + v);
      }
    }
  }
  double m0[3] = {0, 0, 0}, m1[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
  for (int vb = v0; vb < v1; vb += kQ) {
#pragma unroll
    for (int j = 0; j < kQ; ++j) {
      const int v = vb + j;
      if (v < v1) {  // uniform over the block
        // unconditional (clamped at the chunk's last ordinate), so the load
        // lands in its slot without a branch that would force a wait
        const int vn = min(v + kDepth, v1 - 1);
        const int jn = (j + kDepth) % kQ;
        {
          const T* fp = fe + (size_t)vn * a.ldf;
          q[jn][0] = fp[0];
          q[jn][1] = fp[1];
          q[jn][2] = fp[2];
          qx[jn] = ldc(a.vtab, vn);
          qy[jn] = ldc(a.vtab, V + vn);
          qi[jn] = ldc(a.vtab, 2 * V + vn);
          qp[jn] = sph + (float)ldc(a.vtab, 3 * V + vn);
          if (kMom) {
            qw[jn] = ldc(a.vtd, 4 * V + vn);
            qcx[jn] = ldc(a.vtd, 5 * V + vn);
            qcy[jn] = ldc(a.vtd, 6 * V + vn);
          }
        }
        update_ordinate<T, kMom, kCM>(a, s, v, qx[j], qy[j], qi[j], qp[j], dsc0, dsc1, q[j], s_f[j], dte,
                                 res, chg, qw[j], qcx[j], qcy[j], m0, m1, m2);
      }
    }
  }
  if (kMom && s.owned && s.wsl < 0) {
    double* ms = a.msum + (size_t)(logical / a.ntiles) * 3 * N3 + 3 * (size_t)e;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ms[k] = m0[k];
      ms[N3 + k] = m1[k];
      ms[2 * N3 + k] = m2[k];
    }
  }
  const double r = block_max((double)res, s_red);
  if (threadIdx.x == 0) a.res_part[logical] = r;
  const double rc = block_max((double)chg, s_red);
  if (threadIdx.x == 0) a.res_part[a.res_stride + logical] = rc;
}

// ---------------------------------------------------------------------------
// k_wall: the wall sides of the wall elements, and their update.  Per
// ordinate, a wall side's flux is the interior form on the element's own
// values when c.n >= 0 (src/getbou.f90:174-187), else the reflected
// distribution (src/getbou.f90:190-255).  In the reference's in-place
// velocity sweep the reflection reads f at ordinates I1, I2 of the same
// element, and those already advanced this step (index < v) are read from
// the new buffer (src/getbou.f90:245-250 inside src/advnce.f90:273-383);
// with jacobi set every read is from the previous step.  One thread per
// (wall element, speed ring) when every reflection stays in its ring, else
// per wall element over all ordinates.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_wall(WallArgs<T> a) {
  __shared__ double s_red[8];
  // per thread, the ring's freshly advanced f (when reflections stay in the
  // ring): the in-place reads of ordinates < v come from LDS
  extern __shared__ unsigned char s_dyn[];
  T* s_ring = reinterpret_cast<T*>(s_dyn) + (size_t)threadIdx.x * 3 * a.T_;
  const bool use_lds = a.lds_ring;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  double res = 0.0, chg = 0.0;
  if (gid < a.nwe * a.ngroups) {
    const int we = gid / a.ngroups, grp = gid % a.ngroups;
    const int e = a.wall_elems[we];
    const int E = a.E, V = a.V;
    const size_t N3 = 3 * (size_t)E;
    const int vb = grp * a.gsize;
    const int ve = vb + a.gsize;
    const T dte = (T)a.scal[0];
    const T mdt = a.geo[7 * E + e] * dte;
    int info[3], nsl[3];
    T enx[3], eny[3], eal[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      info[k] = a.einfo[k * E + e];
      nsl[k] = a.eslot[k * E + e];
      enx[k] = a.enx[k * E + e];
      eny[k] = a.eny[k * E + e];
      eal[k] = a.ealen[k * E + e];
    }
    const int mask = a.emask[e];
    const T alpha = a.alpha;
    // FP32 store dither (as in k_update): node, ordinate and step phases
    const float wph = a.ephase[e];
    const double stw = a.scal[6] * 0.6180339887498949;
    const float sph = (float)(stw - floor(stw));
    for (int v = vb; v < ve; ++v) {
      const T* wp = a.wpart + ((size_t)we * V + v) * 8;
      T rhs[3] = {wp[0], wp[1], wp[2]};
      const T cold[3] = {wp[3], wp[4], wp[5]};  // collision term times DTE
      const T U = wp[6];
      const T ux = a.vtab[v], uy = a.vtab[V + v];
      const T* fc = a.fcur + (size_t)v * a.ldf + 3 * (size_t)e;
      const T f[3] = {fc[0], fc[1], fc[2]};
      // FP32: the edge terms act on (g - gbar), as in k_update's FP32 form
      // (the element's own values as differences from f_0)
      T hw[3] = {T(0), T(0), T(0)}, gbar = T(0);
      if constexpr (sizeof(T) == 4) {
        const T d1 = f[1] - f[0], d2 = f[2] - f[0];
        const T db = (T(1) / T(3)) * (d1 + d2);
        hw[0] = -db;
        hw[1] = d1 - db;
        hw[2] = d2 - db;
        gbar = (f[0] + db) - U;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if ((info[k] & 3) != 3) continue;
        const int a1 = k, a2 = (k + 1) % 3;  // boundary side: (p1, p2) = (k, k+1)
        const int ws = nsl[k];
        const size_t rk = (size_t)ws * V + v;
        const int i1 = a.refl_i[2 * rk];
        T g1, g2;
        if (i1 < 0) {  // molecules leaving through the wall: own values
          g1 = f[a1] - U;
          g2 = f[a2] - U;
        } else if (a.diffuse) {  // alpha = 1: (1 - alpha) * Rf is exactly 0
          const T mf = a.refl_t[3 * rk + 2] * (T)a.eta[a.wall_bside[ws]];
          g1 = alpha * mf;
          g2 = alpha * mf;
        } else {
          const int i2 = a.refl_i[2 * rk + 1];
          const T zd1 = a.refl_t[3 * rk], zd2 = a.refl_t[3 * rk + 1];
          const T mw = a.refl_t[3 * rk + 2];
          const T* s1;
          const T* s2;
          if (use_lds) {
            s1 = i1 < v ? s_ring + 3 * (i1 - vb) : a.fcur + (size_t)i1 * a.ldf + 3 * (size_t)e;
            s2 = i2 < v ? s_ring + 3 * (i2 - vb) : a.fcur + (size_t)i2 * a.ldf + 3 * (size_t)e;
          } else if (a.jacobi) {
            s1 = a.fcur + (size_t)i1 * a.ldf + 3 * (size_t)e;
            s2 = a.fcur + (size_t)i2 * a.ldf + 3 * (size_t)e;
          } else {
            s1 = (i1 < v ? a.fnxt : a.fcur) + (size_t)i1 * a.ldf + 3 * (size_t)e;
            s2 = (i2 < v ? a.fnxt : a.fcur) + (size_t)i2 * a.ldf + 3 * (size_t)e;
          }
          const T rf1 = (zd2 * s1[a1] + zd1 * s2[a1]) / (zd1 + zd2);
          const T rf2 = (zd2 * s1[a2] + zd1 * s2[a2]) / (zd1 + zd2);
          const T mf = mw * (T)a.eta[a.wall_bside[ws]];
          g1 = (T(1) - alpha) * rf1 + alpha * mf;
          g2 = (T(1) - alpha) * rf2 + alpha * mf;
        }
        if constexpr (sizeof(T) == 4) {  // FP32 form: g - gbar
          if (i1 < 0) {
            g1 = hw[a1];
            g2 = hw[a2];
          } else {
            g1 = g1 - gbar;
            g2 = g2 - gbar;
          }
        }
        T r1, r2;
        edge_rhsi(ux, uy, enx[k], eny[k], eal[k], g1, g2, r1, r2);
        rhs[a1] += r1;
        rhs[a2] += r2;
      }
      T* fo = a.fnxt + (size_t)v * a.ldf + 3 * (size_t)e;
      const T fin = a.vtab[2 * V + v];
      const float ph = sph + (float)a.vtab[3 * V + v];
      const float wt = fmaf(__builtin_amdgcn_fractf(wph + ph), a.dither ? 0x1p-23f : 0.f,
                            a.dither ? -0x1p-24f : 0.f);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T d = node_increment(a.cmminc != 0, mdt, rhs[k], cold[k], res);
        T dk = d;
        if constexpr (sizeof(T) == 4) {
          dk = fmaf(wt, __uint_as_float(__float_as_uint(f[k]) & 0x7f800000u), dk);
        }
        T fn = f[k] + dk;
        fn = fn < T(0) ? T(0) : fn;
        if ((mask >> k) & 1) fn = fin;
        T cg = fabs(fn - f[k]);
        if constexpr (sizeof(T) == 4) {  // the increment, as in k_update
          if (fn != T(0) && !((mask >> k) & 1)) cg = fabs(d);
        }
        chg = fmax(chg, (double)cg);
        fo[k] = fn;
        if (use_lds) s_ring[3 * (v - vb) + k] = fn;
      }
    }
  }
  const double r = block_max(res, s_red);
  if (threadIdx.x == 0) a.res_part[blockIdx.x] = r;
  const double rc = block_max(chg, s_red);
  if (threadIdx.x == 0) a.res_part[a.res_stride + blockIdx.x] = rc;
}

// ---------------------------------------------------------------------------
// k_moments: MACROS (src/macros.f90:80-164) and COMPMU
// (src/compmu.f90:36-97) per DG node, FP64 compensated sums over all ordinates.  Pass 1:
// n, n*u, n*v; pass 2: pressure and mean relative speed about u.  Writes the
// FP64 node fields and the next step's GTEQNF/COMPMU parameters in T.
// ---------------------------------------------------------------------------
// k_moments geometry: a block of 256 threads serves 256 / S nodes with S
// wavefronts per node, each summing its share of the ordinates; the shares
// combine through LDS in wavefront order.  Few nodes give few wavefronts
// (one per 64 nodes at S = 1), too few to hide FP64 latency; many nodes
// keep the chip busy at S = 1, where the combine would only cost.
// kMomRing: register sets of four ordinates each wavefront keeps in its
// load ring.
#ifndef DGBGK_MOM_RING32
#define DGBGK_MOM_RING32 2
#endif
#ifndef DGBGK_MOM_RING64
#define DGBGK_MOM_RING64 2
#endif
template <typename T>
constexpr int kMomRing = sizeof(T) == 4 ? DGBGK_MOM_RING32 : DGBGK_MOM_RING64;

template <typename T, int S>
__global__ void __launch_bounds__(256) k_moments(MomArgs<T> a) {
  constexpr bool kComp = sizeof(T) < sizeof(double);
  constexpr int kMomNodes = 256 / S;
  __shared__ double s_red[8];
  __shared__ double s_part[S > 1 ? S : 1][6][S > 1 ? kMomNodes : 1];
  // the wavefront's share index, made a scalar: the ordinate loops then
  // read the velocity tables with scalar loads
  const int lane = threadIdx.x % kMomNodes;
  const int sub = __builtin_amdgcn_readfirstlane(threadIdx.x / kMomNodes);
  const size_t N3 = 3 * (size_t)a.E;
  const size_t k = (size_t)blockIdx.x * kMomNodes + lane;
  const bool live = k < N3;
  const int V = a.V;
  const T* fp = a.f + (live ? k : 0);
  const double* cxa = a.vtd;
  const double* cya = a.vtd + V;
  const double* cxb = a.vtd + 2 * V;
  const double* cyb = a.vtd + 3 * V;
  const double* wv = a.vtd + 4 * V;
  const double* wcx = a.vtd + 5 * V;
  const double* wcy = a.vtd + 6 * V;
  const double* hv = a.vtd + 7 * V;

  // First pass (n, n u, n v), summed per ordinate chunk of the update
  // kernel, then over the chunks: this wavefront's contiguous share of the
  // chunks, then the shares in wavefront order.  Plain sums in the FP64
  // form (one chunk and one share give the reference's ordinate order);
  // Kahan compensation over chunk sums and shares in the FP32-bulk form.
  // Inside a step k_update has left the chunk sums in msum (wall elements
  // excepted), computed in exactly this order.
  double s0 = 0, c0 = 0, s1 = 0, c1 = 0, s2 = 0, c2 = 0;
  if (live) {
    const bool have = a.msum != nullptr && a.ewslot[k / 3] < 0;
    const int cper = (a.nchunks + S - 1) / S;
    const int cb = min(a.nchunks, sub * cper), ce = min(a.nchunks, cb + cper);
    for (int ch = cb; ch < ce; ++ch) {
      double b0 = 0, b1 = 0, b2 = 0;
      if (have) {
        const double* ms = a.msum + (size_t)ch * 3 * N3 + k;
        b0 = ms[0];
        b1 = ms[N3];
        b2 = ms[2 * N3];
      } else {
        const int vb = ch * a.vchunk, ve = min(V, vb + a.vchunk);
        int v = vb;
        for (; v + 4 <= ve; v += 4) {
          double fv[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) fv[j] = (double)fp[(size_t)(v + j) * a.ldf];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            b0 = fma(ldc(wv, v + j), fv[j], b0);
            b1 = fma(ldc(wcx, v + j), fv[j], b1);
            b2 = fma(ldc(wcy, v + j), fv[j], b2);
          }
        }
        for (; v < ve; ++v) {
          const double fd = (double)fp[(size_t)v * a.ldf];
          b0 = fma(ldc(wv, v), fd, b0);
          b1 = fma(ldc(wcx, v), fd, b1);
          b2 = fma(ldc(wcy, v), fd, b2);
        }
      }
      nsum<kComp>(s0, c0, b0);
      nsum<kComp>(s1, c1, b1);
      nsum<kComp>(s2, c2, b2);
    }
  }
  double t0 = s0, e0 = c0, t1 = s1, e1 = c1, t2 = s2, e2s = c2;
  if constexpr (S > 1) {
    s_part[sub][0][lane] = s0;
    s_part[sub][1][lane] = c0;
    s_part[sub][2][lane] = s1;
    s_part[sub][3][lane] = c1;
    s_part[sub][4][lane] = s2;
    s_part[sub][5][lane] = c2;
    __syncthreads();
    // every wavefront combines the shares in the same order, so all hold
    // the same n, u, v
    t0 = e0 = t1 = e1 = t2 = e2s = 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      nsum<kComp>(t0, e0, s_part[i][0][lane] - s_part[i][1][lane]);
      nsum<kComp>(t1, e1, s_part[i][2][lane] - s_part[i][3][lane]);
      nsum<kComp>(t2, e2s, s_part[i][4][lane] - s_part[i][5][lane]);
    }
  }
  const double n = t0 - e0;
  double nd = 0.0, ux = 0.0, uy = 0.0;
  if (n >= 1e-20) {
    nd = n;
    ux = (t1 - e1) / n;
    uy = (t2 - e2s) / n;
  }
  const bool dens = n >= 1e-20;

  // Second pass (pressure and COMPMU's mean relative speed) over this
  // wavefront's contiguous share of the ordinates, four at a time.
  // FP64: the reference's terms, 0.5 |c_b - u|^2 w f (src/macros.f90:144-148)
  // and |c_a - u| w f (COMPMU's mean relative speed).
  // FP32-bulk: the pressure sum expanded about the first pass's mean
  // velocity, sum of f (0.5 w |c_b|^2) here and 0.5 n |u|^2 subtracted once
  // (the two agree in exact arithmetic since n u = sum w c_b f), and the
  // relative speed in single precision (the hardware square root, far
  // below the FP32 rounding of f); both sums in FP64, compensated.
  const float uxf = (float)ux, uyf = (float)uy;
  double sp = 0, cp = 0, st = 0, ct = 0;
  auto term = [&](T raw, int v, double& bp, double& bt) {
    if constexpr (kComp) {
      const float ex = uxf - (float)ldc(a.vtab, v), ey = uyf - (float)ldc(a.vtab, V + v);
      const float rel = __builtin_amdgcn_sqrtf(fmaf(ex, ex, ey * ey));
      bp = fma((double)raw, ldc(hv, v), bp);
      bt = fma((double)(rel * (float)raw), ldc(wv, v), bt);
    } else {
      const double fw = (double)raw * ldc(wv, v);
      const double dx = ldc(cxb, v) - ux, dy = ldc(cyb, v) - uy;
      const double ex = ux - ldc(cxa, v), ey = uy - ldc(cya, v);
      sp += 0.5 * (dx * dx + dy * dy) * fw;
      if (dens) st += sqrt(ex * ex + ey * ey) * fw;
    }
  };
  auto block4 = [&](const T (&raw)[4], int v) {
    double bp = 0, bt = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) term(raw[j], v + j, bp, bt);
    if (kComp) {
      nsum<true>(sp, cp, bp);
      if (dens) nsum<true>(st, ct, bt);
    }
  };
  const int vper = ((V + S - 1) / S + 3) / 4 * 4;
  const int vb = min(V, sub * vper), ve = min(V, vb + vper);
  int v = vb;
  if (live) {
    // A ring of kMomRing register sets of four ordinates: while one set is
    // summed, the loads of the next sets are in flight.  Ring positions are
    // compile-time (unrolled), loads are clamped to the share's last
    // ordinate (none is conditional), and the sums run in ordinate order.
    constexpr int kR = kMomRing<T>;
    if (ve - vb >= 4 * kR) {
      T ring[kR][4];
#pragma unroll
      for (int r = 0; r < kR; ++r)
#pragma unroll
        for (int j = 0; j < 4; ++j) ring[r][j] = fload_nt(fp + (size_t)(vb + 4 * r + j) * a.ldf);
      for (; v + 4 * kR <= ve; v += 4 * kR) {
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          T cur[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            cur[j] = ring[r][j];
            ring[r][j] = fload_nt(fp + (size_t)min(v + 4 * (kR + r) + j, ve - 1) * a.ldf);
          }
          block4(cur, v + 4 * r);
        }
      }
      // fewer than 4 kR ordinates remain; the ring holds them in order
#pragma unroll
      for (int r = 0; r < kR; ++r)
        if (v + 4 <= ve) {
          T cur[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) cur[j] = ring[r][j];
          block4(cur, v);
          v += 4;
        }
    }
    for (; v + 4 <= ve; v += 4) {
      T rc[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) rc[j] = fp[(size_t)(v + j) * a.ldf];
      block4(rc, v);
    }
    for (; v < ve; ++v) {
      double bp = 0, bt = 0;
      term(fp[(size_t)v * a.ldf], v, bp, bt);
      if (kComp) {
        nsum<true>(sp, cp, bp);
        if (dens) nsum<true>(st, ct, bt);
      }
    }
  }
  double tp = sp, ep = cp, tt = st, et = ct;
  if constexpr (S > 1) {
    __syncthreads();  // every wavefront has read the first-pass shares
    s_part[sub][0][lane] = sp;
    s_part[sub][1][lane] = cp;
    s_part[sub][2][lane] = st;
    s_part[sub][3][lane] = ct;
    __syncthreads();
    tp = ep = tt = et = 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      nsum<kComp>(tp, ep, s_part[i][0][lane] - s_part[i][1][lane]);
      nsum<kComp>(tt, et, s_part[i][2][lane] - s_part[i][3][lane]);
    }
  }
  double mu = 0.0, mass = 0.0;
  if (sub == 0 && live) {
    const double ps = ((tp - ep) - (kComp ? 0.5 * nd * (ux * ux + uy * uy) : 0.0)) * a.mm;
    const double ctv = dens ? (tt - et) / n : 0.0;
    mu = sqrt(2.0) * ctv * a.sigt * nd;
    a.mom[k] = nd;
    a.mom[N3 + k] = ux;
    a.mom[2 * N3 + k] = uy;
    a.mom[3 * N3 + k] = ps;
    a.mom[4 * N3 + k] = mu;
    // GTEQNF parameters (src/gteqnf.f90:28-37)
    double pa = 0.0, pb = 0.0;
    if (!(ps < 1e-20)) {
      pb = (a.mm * nd) / (2 * ps);
      pa = nd * (pb / 3.1416);
    }
    a.np[k] = (T)pa;
    a.np[N3 + k] = (T)pb;
    a.np[2 * N3 + k] = (T)ux;
    a.np[3 * N3 + k] = (T)uy;
    a.np[4 * N3 + k] = (T)mu;
    mass = nd * a.rj[k / 3] / 6.0;
  }
  const double m1 = block_max(mu, s_red);
  if (threadIdx.x == 0) a.mu_part[blockIdx.x] = m1;
  const double m2 = block_sum(mass, s_red);
  if (threadIdx.x == 0) a.mass_part[blockIdx.x] = m2;
}

// FLUCON2 (src/flucon2.f90:133-233): one block per type-4 wall side.
template <typename T>
__global__ void __launch_bounds__(256) k_eta(const T* __restrict__ f,
                                             const int* __restrict__ side,
                                             const double* __restrict__ geo,
                                             const double* __restrict__ vtd,
                                             double* __restrict__ eta, int E, int V,
                                             long long ldf) {
  __shared__ double s_red[8];
  const int j = blockIdx.x;
  const int b = side[4 * j], e = side[4 * j + 1], in1 = side[4 * j + 2], in2 = side[4 * j + 3];
  const double anx = geo[3 * j], any = geo[3 * j + 1], bottom = geo[3 * j + 2];
  const size_t N3 = 3 * (size_t)E;
  double s = 0.0;
  for (int v = threadIdx.x; v < V; v += blockDim.x) {
    const double test = vtd[2 * V + v] * anx + vtd[3 * V + v] * any;
    if (test > 0.0) {
      const T* fp = f + (size_t)v * ldf + 3 * (size_t)e;
      const double nf = 0.5 * ((double)fp[in1] + (double)fp[in2]);
      s += fabs(test) * nf * vtd[4 * V + v];
    }
  }
  s = block_sum(s, s_red);
  if (threadIdx.x == 0) eta[b] = s / bottom;
}

// MACROS point averages (src/macros.f90:156-188) of point q over the point
// CSR; gather: also GETMAC's copy of the pressure for GETFOR.
__device__ __forceinline__ void point_average(int q, const double* __restrict__ mom,
                                              const double* __restrict__ rj,
                                              const int* __restrict__ ptr,
                                              const int* __restrict__ node,
                                              double* __restrict__ pts,
                                              double* __restrict__ ps_glob, int E, int P,
                                              double mm, double rgas, int gather) {
  const size_t N3 = 3 * (size_t)E;
  double flag = 0, nd = 0, u = 0, vv = 0, ps = 0;
  for (int i = ptr[q]; i < ptr[q + 1]; ++i) {
    const int k = node[i];
    const double ae = rj[k / 3];
    const double n = mom[k];
    if (n >= 1e-20) {
      u += ae * mom[N3 + k];
      vv += ae * mom[2 * N3 + k];
    }
    flag += ae;
    nd += n * ae;
    ps += mom[3 * N3 + k] * ae;
  }
  double rho = 0, t = 0;
  if (nd < 1e-20) {
    u = vv = ps = 0.0;
  } else {
    nd = nd / flag;
    rho = nd * mm;
    u = u / flag;
    vv = vv / flag;
    ps = ps / flag;
    t = ps / (rgas * rho);
  }
  pts[q] = nd;
  pts[P + q] = rho;
  pts[2 * P + q] = u;
  pts[3 * P + q] = vv;
  pts[4 * P + q] = ps;
  pts[5 * P + q] = t;
  if (gather) ps_glob[q] = ps;
}

__global__ void __launch_bounds__(256) k_points(const double* __restrict__ mom,
                                                const double* __restrict__ rj,
                                                const int* __restrict__ ptr,
                                                const int* __restrict__ node,
                                                double* __restrict__ pts,
                                                double* __restrict__ ps_glob, int E,
                                                int P, double mm, double rgas, int gather) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < P) point_average(q, mom, rj, ptr, node, pts, ps_glob, E, P, mm, rgas, gather);
}

struct PointArgs {
  const double* mom;
  const double* rj;
  const int* ptr;
  const int* node;
  double* pts;
  int E, P;
  double mm, rgas;
};

// Step epilogue: GETRES max, mass, GETFOR (src/getfor.f90:29-49) on the
// gathered point pressure, the residual record, and ALOTIM2's next DTE.
__global__ void __launch_bounds__(1024) k_final(const double* __restrict__ res_part, int nres,
                                               const double* __restrict__ mu_part,
                                               const double* __restrict__ mass_part, int nmom,
                                               const int* __restrict__ fside, int nf,
                                               const double* __restrict__ coord,
                                               double* __restrict__ ps_glob,
                                               double* __restrict__ scal,
                                               StepRec* __restrict__ rec, int rec_cap,
                                               int record, PointArgs pa) {
  __shared__ double s_red[6][16];
  const double steps_done = scal[6];  // in flight with the partials below
  double r = 0.0, rc = 0.0, m = 0.0, ms = 0.0, lift = 0.0, drag = 0.0;
  // each thread's partials, up to kLd per array loaded before any is
  // combined (one memory latency instead of one per partial); combined in
  // index order, so the mass sum keeps its order (out-of-range slots add
  // 0 and all partials are >= 0)
  constexpr int kLd = 8;
  const int nt = blockDim.x;
  for (int b = threadIdx.x; b < nres; b += kLd * nt) {
    double x[kLd], y[kLd];
#pragma unroll
    for (int j = 0; j < kLd; ++j) {
      const int i = b + j * nt;
      x[j] = i < nres ? res_part[i] : 0.0;
      y[j] = i < nres ? res_part[nres + i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kLd; ++j) {
      r = fmax(r, x[j]);
      rc = fmax(rc, y[j]);
    }
  }
  for (int b = threadIdx.x; b < nmom; b += kLd * nt) {
    double x[kLd], y[kLd];
#pragma unroll
    for (int j = 0; j < kLd; ++j) {
      const int i = b + j * nt;
      x[j] = i < nmom ? mu_part[i] : 0.0;
      y[j] = i < nmom ? mass_part[i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kLd; ++j) {
      m = fmax(m, x[j]);
      ms += y[j];
    }
  }
  // In a step, GETMAC on every 1000th step (src/advnce.f90:403-406): the
  // point averages and the gathered pressure that GETFOR below uses.  Rare,
  // so this block does it rather than a launch every step.
  if (record >= 0 && (((long long)steps_done + 1) % 1000) == 0) {
    for (int q = threadIdx.x; q < pa.P; q += blockDim.x)
      point_average(q, pa.mom, pa.rj, pa.ptr, pa.node, pa.pts, ps_glob, pa.E, pa.P, pa.mm,
                    pa.rgas, 1);
    __syncthreads();
  }
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    const int p1 = fside[2 * i], p2 = fside[2 * i + 1];
    const double x1 = coord[2 * p1], y1 = coord[2 * p1 + 1];
    const double x2 = coord[2 * p2], y2 = coord[2 * p2 + 1];
    const double len = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    const double px = -(y2 - y1), py = x2 - x1;
    const double mag = sqrt(px * px + py * py);
    const double psav = 0.5 * (ps_glob[p1] + ps_glob[p2]);
    lift -= psav * len * (py / mag);
    drag -= psav * len * (px / mag);
  }
  // the three maxima and three sums in one LDS round: each wavefront
  // reduces its six values, then thread j combines value j over the
  // wavefronts in wavefront order (the order of block_max / block_sum)
  for (int o = 32; o > 0; o >>= 1) {
    r = fmax(r, __shfl_down(r, o, 64));
    rc = fmax(rc, __shfl_down(rc, o, 64));
    m = fmax(m, __shfl_down(m, o, 64));
    ms += __shfl_down(ms, o, 64);
    lift += __shfl_down(lift, o, 64);
    drag += __shfl_down(drag, o, 64);
  }
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_red[0][w] = r;
    s_red[1][w] = rc;
    s_red[2][w] = m;
    s_red[3][w] = ms;
    s_red[4][w] = lift;
    s_red[5][w] = drag;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int j = threadIdx.x;
    double x = s_red[j][0];
    for (int i = 1; i < nw; ++i) x = j < 3 ? fmax(x, s_red[j][i]) : x + s_red[j][i];
    s_red[j][0] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    r = s_red[0][0];
    rc = s_red[1][0];
    m = s_red[2][0];
    ms = s_red[3][0];
    lift = s_red[4][0];
    drag = s_red[5][0];
    if (record >= 0) {
      // scal[6] counts completed steps (ITIME of src/advnce.f90:220)
      const long long itime = (long long)scal[6] + 1;
      scal[6] = (double)itime;
      const int slot = (int)((itime - 1) % rec_cap);
      if (itime == 1) scal[1] = r;
      rec[slot].sumres = r;
      rec[slot].logres = log(r / scal[1]);
      rec[slot].dte = scal[0];
      rec[slot].mass = ms;
      rec[slot].lift = lift;
      rec[slot].drag = drag;
      rec[slot].change = rc;
    }
    scal[2] = lift;
    scal[3] = drag;
    scal[5] = ms;
    const double dteb = (m > 0.0) ? 0.9 * (1 / m) : 1000000.0;
    scal[0] = dteb < scal[4] ? dteb : scal[4];
  }
}

// Maxwellian seed per DG node from per-point (n, beta^2, u, v): the INICON2
// formula (src/inicon2.f90:149-163) with point-wise state.
template <typename T>
__global__ void __launch_bounds__(256) k_seed(T* __restrict__ f, const int* __restrict__ intma,
                                              const double* __restrict__ ps4,
                                              const double* __restrict__ vtd, int E, int V,
                                              int P, long long ldf) {
  const size_t N3 = 3 * (size_t)E;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (k >= N3) return;
  const int q = intma[k] - 1;
  const double n1 = ps4[q], b2 = ps4[P + q], u0 = ps4[2 * P + q], v0 = ps4[3 * P + q];
  const double ux = vtd[2 * V + v], uy = vtd[3 * V + v];
  const double tmp = (ux - u0) * (ux - u0) + (uy - v0) * (uy - v0);
  const double sp = sqrt(tmp);
  const double f0 = (b2 / 3.1416) * exp(-(b2 * (sp * sp)));
  f[(size_t)v * ldf + k] = (T)(n1 * f0);
}

// Minimum, maximum and count of non-finite values of f, per block: rows v
// of blockIdx.y, grid-stride over the 3E node values of the row (the row
// padding is not read).  The host combines the blocks' partials.
template <typename T>
__global__ void __launch_bounds__(256) k_fstats(const T* __restrict__ f, int E, long long ldf,
                                                double* __restrict__ part) {
  __shared__ double sh[8];
  const size_t N3 = 3 * (size_t)E;
  const T* row = f + (size_t)blockIdx.y * ldf;
  double lo = INFINITY, hi = -INFINITY, bad = 0.0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < N3;
       k += (size_t)gridDim.x * blockDim.x) {
    const double x = (double)row[k];
    if (isfinite(x)) {
      lo = fmin(lo, x);
      hi = fmax(hi, x);
    } else {
      bad += 1.0;
    }
  }
  const double mlo = -block_max(-lo, sh);
  const double mhi = block_max(hi, sh);
  const double nb = block_sum(bad, sh);
  if (threadIdx.x == 0) {
    double* o = part + 3 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x);
    o[0] = mlo;
    o[1] = mhi;
    o[2] = nb;
  }
}

// Combines n per-block (min, max, non-finite count) triples of k_fstats
// into one triple (one block): min and max are exact in any order, and the
// counts are integers summed exactly in double.
__global__ void __launch_bounds__(256) k_fstats_combine(const double* __restrict__ part, int n,
                                                        double* __restrict__ out) {
  __shared__ double sh[8];
  double lo = INFINITY, hi = -INFINITY, bad = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) {
    lo = fmin(lo, part[3 * (size_t)b]);
    hi = fmax(hi, part[3 * (size_t)b + 1]);
    bad += part[3 * (size_t)b + 2];
  }
  const double mlo = -block_max(-lo, sh);
  const double mhi = block_max(hi, sh);
  const double nb = block_sum(bad, sh);
  if (threadIdx.x == 0) {
    out[0] = mlo;
    out[1] = mhi;
    out[2] = nb;
  }
}

// scratch: the allocation-free statistics entry point's work array
// (dgbgk_f_stats_noalloc): k_fstats's per-block triples, then the combined
// triple.  Each device holds its own copy for the life of the process; host
// code reaches it only through the address hipGetSymbolAddress returns on
// that device.  Room for kFstatsMaxBlocks blocks (16 per ordinate row at
// up to 4096 ordinates): 1.5 MiB.
constexpr int kFstatsMaxBlocks = 16 * 4096;
static __device__ __align__(256) double scratch[3 * kFstatsMaxBlocks + 4];

// Layout conversions between the reference's DISNF(3,V,E) (double, input
// element numbering) and the device's velocity-major T[V][ldf] (storage
// order; eorder maps a storage position to its input element, or is null
// when the two coincide).
template <typename T>
__global__ void k_to_device(const double* __restrict__ in, T* __restrict__ out,
                            const int* __restrict__ eorder, int E, int V, long long ldf) {
  const size_t N3 = 3 * (size_t)E;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (k >= N3) return;
  const size_t e = k / 3, a = k % 3;
  const size_t ei = eorder ? (size_t)eorder[e] : e;
  out[(size_t)v * ldf + k] = (T)in[(ei * V + v) * 3 + a];
}

template <typename T>
__global__ void k_to_host(const T* __restrict__ in, double* __restrict__ out,
                          const int* __restrict__ eorder, int E, int V, long long ldf) {
  const size_t N3 = 3 * (size_t)E;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (k >= N3) return;
  const size_t e = k / 3, a = k % 3;
  const size_t ei = eorder ? (size_t)eorder[e] : e;
  out[(ei * V + v) * 3 + a] = (double)in[(size_t)v * ldf + k];
}

// ---------------------------------------------------------------------------
// host context
// ---------------------------------------------------------------------------
namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define HIPCK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess)                                                     \
      return fail(kErrHip, std::string(#x) + ": " + hipGetErrorString(e_));   \
  } while (0)

template <typename X>
int upload(X** d, const std::vector<X>& h, std::vector<void*>& owned) {
  size_t n = std::max<size_t>(h.size(), 1);
  HIPCK(hipMalloc((void**)d, n * sizeof(X)));
  owned.push_back(*d);
  if (!h.empty()) HIPCK(hipMemcpy(*d, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice));
  return kOk;
}

template <typename X>
int alloc(X** d, size_t n, std::vector<void*>& owned) {
  HIPCK(hipMalloc((void**)d, std::max<size_t>(n, 1) * sizeof(X)));
  owned.push_back(*d);
  return kOk;
}

}  // namespace

struct Ctx {
  int dev = 0;  // the device current at dgbgk_create, which holds every buffer
  Prep p;  // built on the mesh in storage order
  // storage position -> input element (0-based); empty when the input
  // numbering is kept.  With a renumbering, the input mesh and setup
  // arguments are kept for the reference tables.
  std::vector<int> eorder;
  int* d_eorder = nullptr;
  struct Input {
    std::vector<int> intma, bsido;
    std::vector<double> coord, eta, w;
    dgbgk_params prm{};
    int rorder = 0, torder = 0;
  } in;
  int prec = 0;   // 0 fp64, 1 fp32
  int gs = 1;     // in-place sweep order
  size_t tsz = 8;
  size_t ldf = 0;  // row stride of f (3E padded to 64 values)
  hipStream_t stream = nullptr;
  std::vector<void*> owned;
  void* f[2] = {nullptr, nullptr};
  int cur = 0;
  void *geo = nullptr, *enx = nullptr, *eny = nullptr, *ealen = nullptr, *np = nullptr,
       *unx = nullptr, *uny = nullptr,
       *vtab = nullptr, *refl_t = nullptr, *wpart = nullptr;
  int *einfo = nullptr, *eslot = nullptr, *emask = nullptr, *ewslot = nullptr, *uinfo = nullptr,
      *refl_i = nullptr, *wall_bside = nullptr, *tiles = nullptr, *halo = nullptr,
      *wall_elems = nullptr, *eta_side = nullptr, *fside = nullptr, *pt_ptr = nullptr,
      *pt_node = nullptr, *intma = nullptr;
  double *vtd = nullptr, *eta_geo = nullptr, *eta = nullptr, *rj = nullptr, *mom = nullptr,
         *pts = nullptr, *ps_glob = nullptr, *coord = nullptr, *res_part = nullptr,
         *mu_part = nullptr, *mass_part = nullptr, *scal = nullptr, *ps4 = nullptr,
         *msum = nullptr;
  float* ephase = nullptr;  // [E] FP32 store dither phase per element
  StepRec* rec = nullptr;
  int rec_cap = 0;
  int wall_tpb = 256;
  int gsize = 1, wall_diffuse = 0, wall_lds = 0;  // k_wall ordinates per thread, modes
  int slots = 256;   // update kernel: tile slots (threads per block)
  int depth = 1;     // update kernel: ordinates of f in flight (1 or 2)
  int fuse_mom = 1;  // update kernel also does the first moment pass
  int dither = 1;    // FP32 bulk: unbiased (dithered) rounding of the f store
  int mom_split = 1; // k_moments wavefronts per node (1, 2 or 4)
  int ntiles = 0, vchunk = 1, nvchunks = 1, nblk_upd = 0, nblk_wall = 0, ngroups = 1,
      nblk_mom = 0, neta = 0, nforce = 0;
  int itime = 0;
  bool use_graph = true;
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};  // by current f buffer
  // timing: events recorded around launches, read after a synchronisation
  bool timing = false;
  struct Ev {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<Ev> ev_pool;   // all created events
  size_t ev_used = 0;        // recorded, not yet read
  double t_ms[8] = {};
  long t_cnt[8] = {};
};

namespace {

// partials: k_update of this step left the first moment pass in msum
template <typename T>
int launch_moments(Ctx& c, bool partials = false) {
  MomArgs<T> m;
  m.f = (const T*)c.f[c.cur];
  m.vtd = c.vtd;
  m.vtab = (const T*)c.vtab;
  m.rj = c.rj;
  m.mom = c.mom;
  m.np = (T*)c.np;
  m.mu_part = c.mu_part;
  m.mass_part = c.mass_part;
  m.mm = c.p.prm.m / kNa;
  m.sigt = kPi * c.p.prm.d * c.p.prm.d;
  m.msum = partials ? c.msum : nullptr;
  m.ewslot = c.ewslot;
  m.nchunks = c.nvchunks;
  m.vchunk = c.vchunk;
  m.E = c.p.E;
  m.ldf = (long long)c.ldf;
  m.V = c.p.V;
  if (c.mom_split == 4)
    hipLaunchKernelGGL((k_moments<T, 4>), dim3(c.nblk_mom), dim3(256), 0, c.stream, m);
  else if (c.mom_split == 2)
    hipLaunchKernelGGL((k_moments<T, 2>), dim3(c.nblk_mom), dim3(256), 0, c.stream, m);
  else
    hipLaunchKernelGGL((k_moments<T, 1>), dim3(c.nblk_mom), dim3(256), 0, c.stream, m);
  HIPCK(hipGetLastError());
  return kOk;
}

int launch_final(Ctx& c, int record) {
  PointArgs pa{c.mom, c.rj, c.pt_ptr, c.pt_node, c.pts, c.p.E, c.p.P, c.p.prm.m / kNa,
               c.p.prm.rgas};
  hipLaunchKernelGGL(k_final, dim3(1), dim3(1024), 0, c.stream, c.res_part,
                     c.nblk_upd + c.nblk_wall, c.mu_part, c.mass_part, c.nblk_mom, c.fside,
                     c.nforce, c.coord, c.ps_glob, c.scal, c.rec, c.rec_cap, record, pa);
  HIPCK(hipGetLastError());
  return kOk;
}

// in_step: gate on the device step counter (GETMAC every 1000 steps)
int launch_points(Ctx& c, int gather) {
  const int P = c.p.P;
  hipLaunchKernelGGL(k_points, dim3((P + 255) / 256), dim3(256), 0, c.stream, c.mom, c.rj,
                     c.pt_ptr, c.pt_node, c.pts, c.ps_glob, c.p.E, P, c.p.prm.m / kNa,
                     c.p.prm.rgas, gather);
  HIPCK(hipGetLastError());
  return kOk;
}

// Reads every recorded event pair (the stream must be synchronised).
void tharvest(Ctx& c) {
  for (size_t i = 0; i < c.ev_used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c.ev_pool[i].a, c.ev_pool[i].b) == hipSuccess) {
      c.t_ms[c.ev_pool[i].kernel] += ms;
      c.t_cnt[c.ev_pool[i].kernel] += 1;
    }
  }
  c.ev_used = 0;
}
void tbegin(Ctx& c, int k) {
  if (!c.timing) return;
  if (c.ev_used == c.ev_pool.size()) {
    Ctx::Ev e{k, nullptr, nullptr};
    if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) {
      c.timing = false;
      return;
    }
    c.ev_pool.push_back(e);
  }
  c.ev_pool[c.ev_used].kernel = k;
  if (hipEventRecord(c.ev_pool[c.ev_used].a, c.stream) != hipSuccess) c.timing = false;
}
void tend(Ctx& c, int k) {
  if (!c.timing) return;
  (void)k;
  // a failed record leaves the pair unused: timing stops, the step goes on
  if (hipEventRecord(c.ev_pool[c.ev_used].b, c.stream) != hipSuccess) {
    c.timing = false;
    return;
  }
  c.ev_used += 1;
}

template <typename T, int kS>
void launch_update_s(Ctx& c, const UpdArgs<T>& u) {
  const dim3 g(c.nblk_upd), b(kS);
  if (u.cmminc) {  // IMMAT = 0 runs at prefetch depth 1 (set at create)
    if (c.fuse_mom)
      hipLaunchKernelGGL((k_update<T, 1, true, kS, true>), g, b, 0, c.stream, u);
    else
      hipLaunchKernelGGL((k_update<T, 1, false, kS, true>), g, b, 0, c.stream, u);
    return;
  }
  if (c.fuse_mom) {
    if (c.depth == 2)
      hipLaunchKernelGGL((k_update<T, 2, true, kS, false>), g, b, 0, c.stream, u);
    else
      hipLaunchKernelGGL((k_update<T, 1, true, kS, false>), g, b, 0, c.stream, u);
  } else {
    if (c.depth == 2)
      hipLaunchKernelGGL((k_update<T, 2, false, kS, false>), g, b, 0, c.stream, u);
    else
      hipLaunchKernelGGL((k_update<T, 1, false, kS, false>), g, b, 0, c.stream, u);
  }
}

// Resident blocks per CU of the update instantiation c will launch.
template <typename T, int kS>
int update_blocks_per_cu_s(const Ctx& c) {
  int n = 0;
  hipError_t e;
  if (c.p.prm.immat == 0)
    e = c.fuse_mom
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_update<T, 1, true, kS, true>, kS, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_update<T, 1, false, kS, true>, kS,
                                                           0);
  else if (c.fuse_mom)
    e = c.depth == 2
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_update<T, 2, true, kS, false>, kS, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_update<T, 1, true, kS, false>, kS, 0);
  else
    e = c.depth == 2
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_update<T, 2, false, kS, false>, kS, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_update<T, 1, false, kS, false>, kS, 0);
  return e == hipSuccess && n > 0 ? n : 1;
}

template <typename T>
int update_blocks_per_cu(const Ctx& c) {
  return c.slots == 512 ? update_blocks_per_cu_s<T, 512>(c) : update_blocks_per_cu_s<T, 256>(c);
}

template <typename T>
void launch_update(Ctx& c, const UpdArgs<T>& u) {
  if (c.slots == 512)
    launch_update_s<T, 512>(c, u);
  else
    launch_update_s<T, 256>(c, u);
}

template <typename T>
int step_once(Ctx& c) {
  const Prep& p = c.p;
  const T* fc = (const T*)c.f[c.cur];
  T* fn = (T*)c.f[1 - c.cur];
  if (c.neta) {
    tbegin(c, 3);
    hipLaunchKernelGGL(k_eta<T>, dim3(c.neta), dim3(256), 0, c.stream, fc, c.eta_side,
                       c.eta_geo, c.vtd, c.eta, p.E, p.V, (long long)c.ldf);
    HIPCK(hipGetLastError());
    tend(c, 3);
  }
  UpdArgs<T> u;
  u.fcur = fc; u.fnxt = fn; u.geo = (const T*)c.geo; u.uinfo = c.uinfo;
  u.unx = (const T*)c.unx; u.uny = (const T*)c.uny;
  u.emask = c.emask; u.ewslot = c.ewslot; u.np = (const T*)c.np; u.vtab = (const T*)c.vtab;
  u.tiles = c.tiles; u.halo = c.halo; u.wpart = (T*)c.wpart; u.vtd = c.vtd; u.msum = c.msum;
  u.scal = c.scal; u.res_part = c.res_part; u.res_stride = c.nblk_upd + c.nblk_wall;
  u.E = p.E; u.V = p.V; u.ntiles = c.ntiles; u.vchunk = c.vchunk;
  u.ephase = c.ephase;
  u.ldf = (long long)c.ldf;
  u.dither = c.dither;
  u.cmminc = c.p.prm.immat == 0;
  tbegin(c, 0);
  launch_update<T>(c, u);
  HIPCK(hipGetLastError());
  tend(c, 0);
  if (c.nblk_wall) {
    WallArgs<T> w;
    w.fcur = fc; w.fnxt = fn; w.geo = (const T*)c.geo; w.einfo = c.einfo; w.eslot = c.eslot;
    w.enx = (const T*)c.enx; w.eny = (const T*)c.eny; w.ealen = (const T*)c.ealen;
    w.emask = c.emask; w.vtab = (const T*)c.vtab; w.refl_i = c.refl_i;
    w.refl_t = (const T*)c.refl_t; w.wall_bside = c.wall_bside; w.eta = c.eta;
    w.wall_elems = c.wall_elems; w.wpart = (const T*)c.wpart; w.scal = c.scal;
    w.ephase = c.ephase;
    w.res_part = c.res_part + c.nblk_upd; w.res_stride = c.nblk_upd + c.nblk_wall;
    w.alpha = (T)p.prm.alpha; w.E = p.E; w.V = p.V;
    w.ldf = (long long)c.ldf;
    w.dither = c.dither;
    w.cmminc = c.p.prm.immat == 0;
    w.T_ = p.T; w.ngroups = c.ngroups; w.gsize = c.gsize; w.nwe = (int)p.wall_elems.size();
    w.jacobi = c.gs ? 0 : 1;
    w.diffuse = c.wall_diffuse;
    w.lds_ring = c.wall_lds;
    tbegin(c, 1);
    const size_t lds = c.wall_lds ? (size_t)c.wall_tpb * 3 * p.T * sizeof(T) : 0;
    hipLaunchKernelGGL(k_wall<T>, dim3(c.nblk_wall), dim3(c.wall_tpb), lds, c.stream, w);
    HIPCK(hipGetLastError());
    tend(c, 1);
  }
  c.cur = 1 - c.cur;
  tbegin(c, 2);
  int st = launch_moments<T>(c, c.fuse_mom != 0);
  if (st) return st;
  tend(c, 2);
  tbegin(c, 4);
  st = launch_final(c, 0);  // with GETMAC's point gather every 1000th step
  tend(c, 4);
  return st;
}

// A captured sequence of kGraphSteps steps (an even count, so the two f
// buffers end where they started) replayed as one HIP graph.
constexpr int kGraphSteps = 64;

template <typename T>
int capture_graph(Ctx& c) {
  const int slot = c.cur;
  if (c.graph_exec[slot]) return kOk;
  HIPCK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  int st = kOk;
  for (int s = 0; s < kGraphSteps && st == kOk; ++s) st = step_once<T>(c);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c.stream, &g);
  if (st) {
    if (g) (void)hipGraphDestroy(g);  // the step's own error is the one reported
    return st;
  }
  if (e != hipSuccess) return fail(kErrHip, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&c.graph_exec[slot], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);  // the executable graph no longer needs it
  if (e != hipSuccess) {
    c.graph_exec[slot] = nullptr;
    return fail(kErrHip, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  return kOk;
}

template <typename T>
int run_steps(Ctx& c, int n) {
  int done = 0;
  if (c.use_graph && !c.timing && n >= kGraphSteps) {
    int st = capture_graph<T>(c);
    if (st) return st;
    while (n - done >= kGraphSteps) {
      HIPCK(hipGraphLaunch(c.graph_exec[c.cur], c.stream));
      done += kGraphSteps;
    }
  }
  for (; done < n; ++done) {
    int st = step_once<T>(c);
    if (st) return st;
  }
  return kOk;
}

// An element's fixed dither phase in [0, 1) for the FP32 store: an integer
// hash (lowbias32 mixing) of the bits of its centroid, so the same on every
// device and run, whatever the element numbering or storage order.
uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

float element_phase(const Prep& p, int e) {
  const int* t = &p.intma[3 * (size_t)e];
  double cx = 0.0, cy = 0.0;
  for (int a = 0; a < 3; ++a) {
    cx += p.coord[2 * (size_t)(t[a] - 1)];
    cy += p.coord[2 * (size_t)(t[a] - 1) + 1];
  }
  uint64_t bx, by;
  std::memcpy(&bx, &cx, 8);
  std::memcpy(&by, &cy, 8);
  uint32_t h = mix32((uint32_t)by ^ 0x9e3779b9u);
  h = mix32(h ^ (uint32_t)(by >> 32));
  h = mix32(h ^ (uint32_t)bx);
  h = mix32(h ^ (uint32_t)(bx >> 32));
  return (float)(h >> 8) * 0x1p-24f;
}

template <typename T>
int upload_tables(Ctx& c) {
  const Prep& p = c.p;
  const int E = p.E, V = p.V;
  const size_t N3 = 3 * (size_t)E;
  {
    std::vector<float> ph(E);
    for (int e = 0; e < E; ++e) ph[e] = element_phase(p, e);
    int st = upload(&c.ephase, ph, c.owned);
    if (st) return st;
  }
  std::vector<T> geo(8 * (size_t)E), enx(3 * (size_t)E), eny(3 * (size_t)E), eal(3 * (size_t)E);
  std::vector<int> einfo(3 * (size_t)E), eslot(3 * (size_t)E), emask(E), ewslot(E);
  std::vector<T> unx(3 * (size_t)E), uny(3 * (size_t)E);
  std::vector<int> uinfo(3 * (size_t)E);
  for (int e = 0; e < E; ++e) {
    const double* g = &p.geome[7 * (size_t)e];
    for (int k = 0; k < 6; ++k) geo[(size_t)k * E + e] = (T)g[k];
    geo[6 * (size_t)E + e] = (T)(0.5 * g[6]);
    geo[7 * (size_t)E + e] = (T)(p.prm.immat == 0 ? -p.cm12[e] : p.mmat[3 * (size_t)e]);
    for (int k = 0; k < 3; ++k) {
      const ElemEdge& ed = p.edges[3 * (size_t)e + k];
      const size_t i = (size_t)k * E + e;
      einfo[i] = einfo_pack(ed.kind, ed.a1, ed.a2, ed.b1 < 0 ? 0 : ed.b1,
                            ed.b2 < 0 ? 0 : ed.b2, ed.sign > 0 ? 1 : 0);
      eslot[i] = ed.kind == kInterior ? p.nbr_slot[3 * (size_t)e + k]
                                      : (ed.kind == kWall ? ed.wside : -1);
      enx[i] = (T)ed.nx;
      eny[i] = (T)ed.ny;
      eal[i] = (T)ed.alen;
      // update kernel: (len/6) times the outward normal (the right
      // element's is the side's normal negated), rounded once; zero on
      // wall sides (k_wall's), and on inflow sides in FP64 (no flux); the
      // FP32 form's edge terms carry the volume term's share, so an inflow
      // side keeps its normal there and takes zero upwind values.  The
      // neighbour's nodes at this edge's nodes k and k+1.
      const double sg = ed.sign > 0 ? 1.0 : -1.0;
      const bool flux = ed.kind == kInterior || ed.kind == kOutflow ||
                        (sizeof(T) == 4 && ed.kind == kInflow);
      unx[i] = flux ? (T)(sg * ed.alen * ed.nx) : T(0);
      uny[i] = flux ? (T)(sg * ed.alen * ed.ny) : T(0);
      if (ed.kind == kInterior) {
        const int nbk = ed.sign > 0 ? ed.b1 : ed.b2, nbk1 = ed.sign > 0 ? ed.b2 : ed.b1;
        uinfo[i] = (nbk << kUpdNbK) | (nbk1 << kUpdNbK1) | (ed.sign > 0 ? 0 : kUpdRight) |
                   (p.nbr_slot[3 * (size_t)e + k] << kUpdSlot);
      } else {
        uinfo[i] = kUpdOwn | (ed.kind == kInflow ? kUpdInflow : 0);
      }
    }
    emask[e] = p.inflow_mask[e];
    ewslot[e] = p.wall_elem_slot[e];
  }
  int st;
  if ((st = upload((T**)&c.geo, geo, c.owned))) return st;
  if ((st = upload((T**)&c.enx, enx, c.owned))) return st;
  if ((st = upload((T**)&c.eny, eny, c.owned))) return st;
  if ((st = upload((T**)&c.ealen, eal, c.owned))) return st;
  if ((st = upload(&c.einfo, einfo, c.owned))) return st;
  if ((st = upload(&c.eslot, eslot, c.owned))) return st;
  if ((st = upload(&c.emask, emask, c.owned))) return st;
  if ((st = upload(&c.ewslot, ewslot, c.owned))) return st;
  if ((st = upload((T**)&c.unx, unx, c.owned))) return st;
  if ((st = upload((T**)&c.uny, uny, c.owned))) return st;
  if ((st = upload(&c.uinfo, uinfo, c.owned))) return st;
  std::vector<T> vtab(4 * (size_t)V);
  std::vector<double> vtd(8 * (size_t)V);
  for (int v = 0; v < V; ++v) {
    vtab[v] = (T)p.cxa[v];
    vtab[V + v] = (T)p.cya[v];
    vtab[2 * V + v] = (T)p.finf[v];
    // the FP32 store dither's ordinate phase: a Weyl sequence over v
    const double ow = 0.7548776662466927 * v;
    vtab[3 * V + v] = (T)(ow - std::floor(ow));
    vtd[v] = p.cxa[v];
    vtd[V + v] = p.cya[v];
    vtd[2 * V + v] = p.cxb[v];
    vtd[3 * V + v] = p.cyb[v];
    vtd[4 * V + v] = p.wv[v];
    // the first moment pass's weights: n u = sum (w cx) f, n v = sum (w cy) f
    vtd[5 * V + v] = p.wv[v] * p.cxb[v];
    vtd[6 * V + v] = p.wv[v] * p.cyb[v];
    // the FP32-bulk pressure sum's weight: 0.5 w |c_b|^2
    vtd[7 * V + v] = 0.5 * p.wv[v] * (p.cxb[v] * p.cxb[v] + p.cyb[v] * p.cyb[v]);
  }
  if ((st = upload((T**)&c.vtab, vtab, c.owned))) return st;
  if ((st = upload(&c.vtd, vtd, c.owned))) return st;
  const size_t NW = p.wall_sides.size();
  std::vector<int> ri(2 * NW * V);
  std::vector<T> rt(3 * NW * V);
  for (size_t i = 0; i < NW * V; ++i) {
    ri[2 * i] = p.refl_i1[i];
    ri[2 * i + 1] = p.refl_i2[i];
    rt[3 * i] = (T)p.refl_zd1[i];
    rt[3 * i + 1] = (T)p.refl_zd2[i];
    rt[3 * i + 2] = (T)p.refl_mw[i];
  }
  if ((st = upload(&c.refl_i, ri, c.owned))) return st;
  if ((st = upload((T**)&c.refl_t, rt, c.owned))) return st;
  if ((st = upload(&c.wall_bside, p.wall_sides, c.owned))) return st;
  std::vector<int> tiles(4 * p.tile_e0.size());
  for (size_t t = 0; t < p.tile_e0.size(); ++t) {
    tiles[4 * t] = p.tile_e0[t];
    tiles[4 * t + 1] = p.tile_nown[t];
    tiles[4 * t + 2] = p.tile_hoff[t];
    tiles[4 * t + 3] = p.tile_nhalo[t];
  }
  if ((st = upload(&c.tiles, tiles, c.owned))) return st;
  if ((st = upload(&c.halo, p.halo_elems, c.owned))) return st;
  if ((st = upload(&c.wall_elems, p.wall_elems, c.owned))) return st;
  std::vector<int> es(4 * p.eta_side.size());
  std::vector<double> eg(3 * p.eta_side.size());
  for (size_t j = 0; j < p.eta_side.size(); ++j) {
    es[4 * j] = p.eta_side[j];
    es[4 * j + 1] = p.eta_elem[j];
    es[4 * j + 2] = p.eta_in1[j];
    es[4 * j + 3] = p.eta_in2[j];
    eg[3 * j] = p.eta_nx[j];
    eg[3 * j + 1] = p.eta_ny[j];
    eg[3 * j + 2] = p.eta_bottom[j];
  }
  if ((st = upload(&c.eta_side, es, c.owned))) return st;
  if ((st = upload(&c.eta_geo, eg, c.owned))) return st;
  std::vector<int> fs(2 * p.force_p1.size());
  for (size_t j = 0; j < p.force_p1.size(); ++j) {
    fs[2 * j] = p.force_p1[j];
    fs[2 * j + 1] = p.force_p2[j];
  }
  if ((st = upload(&c.fside, fs, c.owned))) return st;
  std::vector<double> rj(E);
  for (int e = 0; e < E; ++e) rj[e] = p.geome[7 * (size_t)e + 6];
  if ((st = upload(&c.rj, rj, c.owned))) return st;
  if ((st = upload(&c.pt_ptr, p.pt_ptr, c.owned))) return st;
  if ((st = upload(&c.pt_node, p.pt_node, c.owned))) return st;
  if ((st = upload(&c.coord, p.coord, c.owned))) return st;
  if ((st = upload(&c.intma, p.intma, c.owned))) return st;
  std::vector<double> eta0(std::max(p.B, 1), 0.0);
  if ((st = upload(&c.eta, eta0, c.owned))) return st;
  if ((st = alloc((T**)&c.np, 5 * N3, c.owned))) return st;
  if ((st = alloc(&c.mom, 5 * N3, c.owned))) return st;
  if ((st = alloc(&c.pts, 6 * (size_t)p.P, c.owned))) return st;
  if ((st = alloc(&c.ps_glob, (size_t)p.P, c.owned))) return st;
  if ((st = alloc(&c.ps4, 4 * (size_t)p.P, c.owned))) return st;
  if ((st = alloc((T**)&c.wpart, std::max<size_t>(1, p.wall_elems.size() * (size_t)V * 8), c.owned)))
    return st;
  // f rows padded to a multiple of 64 values: each row starts on a
  // 256-byte boundary, so a wavefront's 64 consecutive nodes of one
  // ordinate never straddle more cache lines than they fill
  c.ldf = (N3 + 63) / 64 * 64;
  for (int b = 0; b < 2; ++b)
    if ((st = alloc((T**)&c.f[b], (size_t)V * c.ldf, c.owned))) return st;
  return kOk;
}

template <typename T>
int seed_points(Ctx& c, const std::vector<double>& ps4) {
  HIPCK(hipMemcpyAsync(c.ps4, ps4.data(), ps4.size() * sizeof(double), hipMemcpyHostToDevice,
                       c.stream));
  const size_t N3 = 3 * (size_t)c.p.E;
  hipLaunchKernelGGL(k_seed<T>, dim3((unsigned)((N3 + 255) / 256), c.p.V), dim3(256), 0,
                     c.stream, (T*)c.f[c.cur], c.intma, c.ps4, c.vtd, c.p.E, c.p.V, c.p.P,
                     (long long)c.ldf);
  HIPCK(hipGetLastError());
  return kOk;
}

// after f changes outside a step: initial MACROS + GETMAC + first DTE
template <typename T>
int refresh(Ctx& c) {
  int st = launch_moments<T>(c);
  if (st) return st;
  if ((st = launch_points(c, 1))) return st;
  // DTE for the next step, no record
  HIPCK(hipMemsetAsync(c.res_part, 0, 2 * sizeof(double) * (c.nblk_upd + c.nblk_wall), c.stream));
  return launch_final(c, -1);
}

int do_refresh(Ctx& c) { return c.prec ? refresh<float>(c) : refresh<double>(c); }

}  // namespace
}  // namespace dgbgk

using namespace dgbgk;

struct dgbgk_ctx {
  Ctx c;
};
struct dgbgk_setup {
  Prep p;          // input numbering: the reference tables
  int ntiles = 0;  // update tiles in the device's storage order
};

// Elements are renumbered along a Hilbert curve unless DGBGK_REORDER=0.
static bool reorder_enabled() {
  const char* r = std::getenv("DGBGK_REORDER");
  return !(r && r[0] == '0');
}

static Params to_params(const dgbgk_params* prm) { return params_from_api(*prm); }

extern "C" {

const char* dgbgk_last_error(void) { return g_err.c_str(); }

int dgbgk_setup_create(int nelem, int npoin, int nboun, const int* intma, const double* coord,
                       const int* bsido, int rorder, int torder, const double* eta,
                       const double* w, const dgbgk_params* prm, dgbgk_setup** out) {
  if (!out || !prm || !intma || !coord || (!bsido && nboun) || !eta || !w)
    return fail(kErrArg, "null argument");
  *out = nullptr;
  auto* s = new (std::nothrow) dgbgk_setup;
  if (!s) return fail(kErrNoMem, "out of host memory");
  std::string msg;
  int st = prepare(nelem, npoin, nboun, intma, coord, bsido, rorder, torder, eta, w,
                   to_params(prm), kSlots, kSlots, s->p, msg);
  if (st) {
    delete s;
    return fail(st, msg);
  }
  // the tiles a device context builds, on its storage order
  std::vector<int> order;
  Prep q;
  const char* bo = std::getenv("DGBGK_BORDER_ORDER");
  st = prepare_storage(nelem, npoin, nboun, intma, coord, bsido, rorder, torder, eta, w,
                       to_params(prm), kSlots, kSlots, reorder_enabled(),
                       reorder_enabled() && !(bo && bo[0] == '0'), q, order, msg);
  if (st) {
    delete s;
    return fail(st, msg);
  }
  s->ntiles = (int)q.tile_e0.size();
  *out = s;
  return kOk;
}

int dgbgk_setup_sizes(dgbgk_setup* s, int* nside, int* nvel, int* ntiles, int* nwall_sides,
                      int* ring_separable) {
  if (!s) return fail(kErrArg, "null setup");
  if (nside) *nside = s->p.S;
  if (nvel) *nvel = s->p.V;
  if (ntiles) *ntiles = s->ntiles;
  if (nwall_sides) *nwall_sides = (int)s->p.wall_sides.size();
  if (ring_separable) *ring_separable = s->p.ring_separable ? 1 : 0;
  return kOk;
}

static void copy_tables(const Prep& p, int* iside, double* nx, double* ny, double* el,
                        double* geome, double* relen, double* mmat, double* rsido,
                        double* vcord, double* sumweight, double* dtec) {
  auto cp = [](auto* dst, const auto& v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  cp(iside, p.iside); cp(nx, p.nx); cp(ny, p.ny); cp(el, p.el); cp(geome, p.geome);
  cp(relen, p.relen); cp(mmat, p.mmat); cp(rsido, p.rsido); cp(vcord, p.vcord);
  if (sumweight) *sumweight = p.sumweight;
  if (dtec) *dtec = p.dtec;
}

int dgbgk_setup_tables_host(dgbgk_setup* s, int* iside, double* nx, double* ny, double* el,
                            double* geome, double* relen, double* mmat, double* rsido,
                            double* vcord, double* sumweight, double* dtec) {
  if (!s) return fail(kErrArg, "null setup");
  copy_tables(s->p, iside, nx, ny, el, geome, relen, mmat, rsido, vcord, sumweight, dtec);
  return kOk;
}

int dgbgk_setup_reflections(dgbgk_setup* s, int* side, int* i1, int* i2, double* zd1,
                            double* zd2) {
  if (!s) return fail(kErrArg, "null setup");
  const Prep& p = s->p;
  const size_t n = p.refl_i1.size();
  for (size_t k = 0; side && k < p.wall_sides.size(); ++k) side[k] = p.wall_sides[k] + 1;
  for (size_t k = 0; k < n; ++k) {
    if (i1) i1[k] = p.refl_i1[k] + 1;   // -1 (no reflection) becomes 0
    if (i2) i2[k] = p.refl_i2[k] + 1;
    if (zd1) zd1[k] = p.refl_zd1[k];
    if (zd2) zd2[k] = p.refl_zd2[k];
  }
  return kOk;
}

void dgbgk_setup_destroy(dgbgk_setup* s) { delete s; }

int dgbgk_create(int nelem, int npoin, int nboun, const int* intma, const double* coord,
                 const int* bsido, int rorder, int torder, const double* eta, const double* w,
                 const dgbgk_params* prm, dgbgk_ctx** out) {
  if (!out || !prm || !intma || !coord || (!bsido && nboun) || !eta || !w)
    return fail(kErrArg, "null argument");
  *out = nullptr;
  if (prm->precision != DGBGK_FP64 && prm->precision != DGBGK_FP32)
    return fail(kErrArg, "precision must be DGBGK_FP64 or DGBGK_FP32");
  auto* h = new (std::nothrow) dgbgk_ctx;
  if (!h) return fail(kErrNoMem, "out of host memory");
  Ctx& c = h->c;
  std::string msg;
  // update tiles: 256 slots, owned elements up to the slots left by the
  // halo.  512 slots, or an owned cap (a multiple of 64 keeps the halo in
  // its own wavefront, which then skips phase B), on request; both
  // measured slower on gfx950.
  c.slots = kSlots;
  if (const char* ps = std::getenv("DGBGK_SLOTS")) c.slots = std::atoi(ps) == 512 ? 512 : 256;
  int owned_cap = c.slots;
  if (const char* po = std::getenv("DGBGK_TILE_OWNED")) owned_cap = std::atoi(po);
  Params dprm = to_params(prm);
  if (const char* ta = std::getenv("DGBGK_TILE_ALIGN")) dprm.tile_align = std::max(1, std::atoi(ta));
  const char* bo = std::getenv("DGBGK_BORDER_ORDER");
  int st = prepare_storage(nelem, npoin, nboun, intma, coord, bsido, rorder, torder, eta, w,
                           dprm, c.slots, owned_cap, reorder_enabled(),
                           reorder_enabled() && !(bo && bo[0] == '0'), c.p, c.eorder, msg);
  if (st) {
    delete h;
    return fail(st, msg);
  }
  if (!c.eorder.empty()) {
    c.in.intma.assign(intma, intma + 3 * (size_t)nelem);
    c.in.bsido.assign(bsido, bsido + 6 * (size_t)std::max(nboun, 0));
    c.in.coord.assign(coord, coord + 2 * (size_t)npoin);
    c.in.eta.assign(eta, eta + rorder);
    c.in.w.assign(w, w + rorder);
    c.in.prm = *prm;
    c.in.rorder = rorder;
    c.in.torder = torder;
  }
  c.prec = prm->precision;
  c.gs = prm->order == DGBGK_ORDER_REFERENCE ? 1 : 0;
  if (const char* ng = std::getenv("DGBGK_NO_GRAPH")) c.use_graph = ng[0] == '0';
  c.tsz = c.prec ? 4 : 8;
  // ordinates of f in flight per update thread beyond the one computed:
  // one (a second costs a wavefront of occupancy in both forms: FP64 3 -> 2,
  // FP32 5 -> 4 per SIMD; measured 6% slower in FP32 at 1M triangles), and
  // whether the update kernel also sums the first moment pass
  c.depth = 1;
  if (const char* pd = std::getenv("DGBGK_PREFETCH")) {
    const int d = std::atoi(pd);
    if (d >= 1 && d <= 2 && prm->immat != 0) c.depth = d;  // IMMAT = 0: depth 1 only
  }
  if (const char* fm = std::getenv("DGBGK_FUSE_MOMENTS")) c.fuse_mom = fm[0] != '0';
  if (const char* fd = std::getenv("DGBGK_FP32_DITHER")) c.dither = fd[0] != '0';
  const Prep& p = c.p;
  // error path of create: release what was made, report the first error
  auto cleanup = [&](int code) {
    for (void* q : c.owned) (void)hipFree(q);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete h;
    return code;
  };
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return cleanup(fail(kErrHip, "no HIP device available"));
  if (hipGetDevice(&c.dev) != hipSuccess)
    return cleanup(fail(kErrHip, "hipGetDevice failed"));
  if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(kErrHip, "hipStreamCreate failed"));
  // launch geometry
  c.ntiles = (int)p.tile_e0.size();
  int vch = prm->vchunk;
  if (vch <= 0) {
    // Ordinate chunks per tile.  Each chunk re-reads its slots' tables
    // (Maxwellian parameters, geometry, edges: ~160 B in FP32, ~300 B in
    // FP64 per slot), so fewer is better, as long as the blocks fill the
    // chip's resident capacity and do not leave a short last wave: cost =
    // (capacity rounds used / rounds of work) x (1 + table share x chunks).
    // The capacity takes 256 CUs (MI355X) whatever the device, so the
    // chunking, and with it the FP64 summation order, depends on the mesh
    // only.
    const int bpc = c.prec ? update_blocks_per_cu<float>(c) : update_blocks_per_cu<double>(c);
    const double cap = 256.0 * bpc;
    const double tab = (c.prec ? 160.0 : 300.0) / (2.0 * (c.prec ? 4 : 8) * 3.0 * p.V);
    double best = 1e300;
    int bnv = 1;
    for (int nv = 1; nv <= std::min(p.V, 64); ++nv) {
      const double B = (double)c.ntiles * nv;
      const double eff = std::max(cap / B, std::ceil(B / cap) * cap / B);
      const double cost = eff * (1.0 + tab * nv);
      if (cost < best * (1.0 - 1e-9)) {
        best = cost;
        bnv = nv;
      }
    }
    vch = (p.V + bnv - 1) / bnv;
  }
  c.vchunk = std::max(1, std::min(vch, p.V));
  c.nvchunks = (p.V + c.vchunk - 1) / c.vchunk;
  c.nblk_upd = c.ntiles * c.nvchunks;
  const int nwe = (int)p.wall_elems.size();
  // k_wall: ordinates that depend on each other through in-place reflected
  // reads are walked by one thread in order.  Without in-place reads
  // (Jacobi order, or a fully diffuse wall, whose reflected part has weight
  // 1 - alpha = 0) every (wall element, ordinate) is independent.
  c.wall_tpb = 256;
  c.wall_diffuse = prm->alpha == 1.0 ? 1 : 0;
  c.wall_lds = 0;
  if (!c.gs || c.wall_diffuse) {
    c.ngroups = p.V;
    c.gsize = 1;
  } else if (p.ring_separable) {
    // one thread per speed ring; its freshly advanced f is staged in LDS
    c.ngroups = p.R;
    c.gsize = p.T;
    const size_t per = 3 * (size_t)p.T * c.tsz;
    while (c.wall_tpb > 64 && (size_t)c.wall_tpb * per > 65536) c.wall_tpb >>= 1;
    if ((size_t)c.wall_tpb * per <= 65536) c.wall_lds = 1;
    else c.wall_tpb = 256;
  } else {
    c.ngroups = 1;
    c.gsize = p.V;
  }
  c.nblk_wall = nwe ? (nwe * c.ngroups + c.wall_tpb - 1) / c.wall_tpb : 0;
  // k_moments: one wavefront per 64 nodes leaves fewer than ~8 per SIMD
  // below 512k nodes, and then up to four share each node's ordinates.  The
  // choice depends on the node count only, so a mesh gives the same sums
  // (and results) on any device.
  {
    const size_t n3 = 3 * (size_t)p.E;
    c.mom_split = n3 < (size_t)1 << 19 ? 4 : (n3 < (size_t)1 << 20 ? 2 : 1);
    if (const char* ms = std::getenv("DGBGK_MOM_SPLIT")) {
      const int v = std::atoi(ms);
      if (v == 1 || v == 2 || v == 4) c.mom_split = v;
    }
    const size_t per = 256 / c.mom_split;
    c.nblk_mom = (int)((n3 + per - 1) / per);
  }
  c.neta = (int)p.eta_side.size();
  c.nforce = (int)p.force_p1.size();
  st = c.prec ? upload_tables<float>(c) : upload_tables<double>(c);
  if (st) return cleanup(st);
  if (!c.eorder.empty() && (st = upload(&c.d_eorder, c.eorder, c.owned))) return cleanup(st);
  if ((st = alloc(&c.res_part, 2 * ((size_t)c.nblk_upd + c.nblk_wall), c.owned)))
    return cleanup(st);
  if ((st = alloc(&c.msum, (size_t)c.nvchunks * 3 * 3 * (size_t)p.E, c.owned)))
    return cleanup(st);
  if ((st = alloc(&c.mu_part, (size_t)c.nblk_mom, c.owned))) return cleanup(st);
  if ((st = alloc(&c.mass_part, (size_t)c.nblk_mom, c.owned))) return cleanup(st);
  std::vector<double> scal(8, 0.0);
  scal[4] = p.dtec;
  if ((st = upload(&c.scal, scal, c.owned))) return cleanup(st);
  c.rec_cap = 1024;
  if ((st = alloc(&c.rec, (size_t)c.rec_cap, c.owned))) return cleanup(st);
  // INICON2 (src/inicon2.f90:120-173): uniform Maxwellian from T1, P1, U0, V0
  {
    const double t1 = prm->t1, p1 = prm->p1, gc = prm->rgas, m = prm->m;
    const double rho1 = (p1 * 100000.0) / (gc * t1);
    const double n1 = rho1 * kNa / m;
    const double tmp1 = rho1 / (2 * p1 * 100000.0);
    const double beta1 = std::sqrt(tmp1);
    std::vector<double> ps4(4 * (size_t)p.P);
    for (int q = 0; q < p.P; ++q) {
      ps4[q] = n1 >= 1e-20 ? n1 : 0.0;
      ps4[p.P + q] = beta1 * beta1;
      ps4[2 * (size_t)p.P + q] = prm->u0;
      ps4[3 * (size_t)p.P + q] = prm->v0;
    }
    st = c.prec ? seed_points<float>(c, ps4) : seed_points<double>(c, ps4);
    if (st) return cleanup(st);
  }
  if ((st = do_refresh(c))) return cleanup(st);
  if (hipStreamSynchronize(c.stream) != hipSuccess)
    return cleanup(fail(kErrHip, "initialisation failed on the device"));
  *out = h;
  return kOk;
}

void dgbgk_destroy(dgbgk_ctx* h) {
  if (!h) return;
  Ctx& c = h->c;
  // teardown has no caller to report to: every release is attempted
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  for (auto& e : c.ev_pool) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  for (auto& g : c.graph_exec)
    if (g) (void)hipGraphExecDestroy(g);
  for (void* q : c.owned) (void)hipFree(q);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete h;
}

int dgbgk_step(dgbgk_ctx* h, int nsteps, dgbgk_step_info* info) {
  if (!h || nsteps < 0) return fail(kErrArg, "bad arguments");
  Ctx& c = h->c;
  int done = 0;
  while (done < nsteps) {
    const int batch = std::min(nsteps - done, c.rec_cap);
    const int first = c.itime % c.rec_cap;  // slot of this batch's first record
    int st = c.prec ? run_steps<float>(c, batch) : run_steps<double>(c, batch);
    if (st) return st;
    c.itime += batch;
    if (info) {
      const int n1 = std::min(batch, c.rec_cap - first);
      HIPCK(hipMemcpyAsync(info + done, c.rec + first, n1 * sizeof(StepRec),
                           hipMemcpyDeviceToHost, c.stream));
      if (batch > n1)
        HIPCK(hipMemcpyAsync(info + done + n1, c.rec, (batch - n1) * sizeof(StepRec),
                             hipMemcpyDeviceToHost, c.stream));
    }
    HIPCK(hipStreamSynchronize(c.stream));
    tharvest(c);
    done += batch;
  }
  return kOk;
}

int dgbgk_get_f(dgbgk_ctx* h, double* f) {
  if (!h || !f) return fail(kErrArg, "null argument");
  Ctx& c = h->c;
  const size_t n = 3 * (size_t)c.p.E * c.p.V;
  double* d = nullptr;
  HIPCK(hipMalloc((void**)&d, n * sizeof(double)));
  const size_t N3 = 3 * (size_t)c.p.E;
  dim3 g((unsigned)((N3 + 255) / 256), c.p.V);
  if (c.prec)
    hipLaunchKernelGGL(k_to_host<float>, g, dim3(256), 0, c.stream, (const float*)c.f[c.cur], d,
                       (const int*)c.d_eorder, c.p.E, c.p.V, (long long)c.ldf);
  else
    hipLaunchKernelGGL(k_to_host<double>, g, dim3(256), 0, c.stream,
                       (const double*)c.f[c.cur], d, (const int*)c.d_eorder, c.p.E, c.p.V,
                       (long long)c.ldf);
  hipError_t e1 = hipMemcpyAsync(f, d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream);
  hipError_t e2 = hipStreamSynchronize(c.stream);
  const hipError_t e3 = hipFree(d);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(kErrHip, "get_f failed");
  return kOk;
}

int dgbgk_set_f(dgbgk_ctx* h, const double* f) {
  if (!h || !f) return fail(kErrArg, "null argument");
  Ctx& c = h->c;
  const size_t n = 3 * (size_t)c.p.E * c.p.V;
  double* d = nullptr;
  HIPCK(hipMalloc((void**)&d, n * sizeof(double)));
  hipError_t e1 = hipMemcpyAsync(d, f, n * sizeof(double), hipMemcpyHostToDevice, c.stream);
  const size_t N3 = 3 * (size_t)c.p.E;
  dim3 g((unsigned)((N3 + 255) / 256), c.p.V);
  if (c.prec)
    hipLaunchKernelGGL(k_to_device<float>, g, dim3(256), 0, c.stream, d, (float*)c.f[c.cur],
                       (const int*)c.d_eorder, c.p.E, c.p.V, (long long)c.ldf);
  else
    hipLaunchKernelGGL(k_to_device<double>, g, dim3(256), 0, c.stream, d,
                       (double*)c.f[c.cur], (const int*)c.d_eorder, c.p.E, c.p.V,
                       (long long)c.ldf);
  int st = do_refresh(c);
  hipError_t e2 = hipStreamSynchronize(c.stream);
  const hipError_t e3 = hipFree(d);
  if (st) return st;
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(kErrHip, "set_f failed");
  return kOk;
}

int dgbgk_seed_maxwellian(dgbgk_ctx* h, const double* rho, const double* u, const double* v,
                          const double* temp) {
  if (!h || !rho || !u || !v || !temp) return fail(kErrArg, "null argument");
  Ctx& c = h->c;
  const Prep& p = c.p;
  std::vector<double> ps4(4 * (size_t)p.P);
  for (int q = 0; q < p.P; ++q) {
    // INICON2 with point state: n = rho*NA/M, beta^2 = rho/(2p), p = rho*R*T
    const double pr = rho[q] * p.prm.rgas * temp[q];
    const double n1 = rho[q] * kNa / p.prm.m;
    ps4[q] = (n1 >= 1e-20 && pr > 0.0) ? n1 : 0.0;
    ps4[p.P + q] = pr > 0.0 ? rho[q] / (2 * pr) : 0.0;
    ps4[2 * (size_t)p.P + q] = u[q];
    ps4[3 * (size_t)p.P + q] = v[q];
  }
  int st = c.prec ? seed_points<float>(c, ps4) : seed_points<double>(c, ps4);
  if (st) return st;
  if ((st = do_refresh(c))) return st;
  HIPCK(hipStreamSynchronize(c.stream));
  return kOk;
}

int dgbgk_node_fields(dgbgk_ctx* h, double* disnd, double* disux, double* disuy,
                      double* disps, double* mu) {
  if (!h) return fail(kErrArg, "null context");
  Ctx& c = h->c;
  const size_t N3 = 3 * (size_t)c.p.E;
  double* dst[5] = {disnd, disux, disuy, disps, mu};
  std::vector<double> tmp(c.eorder.empty() ? 0 : N3);
  for (int i = 0; i < 5; ++i) {
    if (!dst[i]) continue;
    double* to = c.eorder.empty() ? dst[i] : tmp.data();
    HIPCK(hipMemcpyAsync(to, c.mom + i * N3, N3 * sizeof(double), hipMemcpyDeviceToHost,
                         c.stream));
    HIPCK(hipStreamSynchronize(c.stream));
    if (!c.eorder.empty())
      for (size_t e = 0; e < (size_t)c.p.E; ++e)
        for (int a = 0; a < 3; ++a) dst[i][3 * (size_t)c.eorder[e] + a] = tmp[3 * e + a];
  }
  return kOk;
}

int dgbgk_point_fields(dgbgk_ctx* h, double* nd, double* rho, double* uvel, double* vvel,
                       double* ps, double* temp) {
  if (!h) return fail(kErrArg, "null context");
  Ctx& c = h->c;
  int st = launch_points(c, 0);
  if (st) return st;
  const size_t P = c.p.P;
  double* dst[6] = {nd, rho, uvel, vvel, ps, temp};
  for (int i = 0; i < 6; ++i)
    if (dst[i])
      HIPCK(hipMemcpyAsync(dst[i], c.pts + i * P, P * sizeof(double), hipMemcpyDeviceToHost,
                           c.stream));
  HIPCK(hipStreamSynchronize(c.stream));
  return kOk;
}

int dgbgk_forces(dgbgk_ctx* h, double* lift, double* drag) {
  if (!h) return fail(kErrArg, "null context");
  Ctx& c = h->c;
  const Prep& p = c.p;
  std::vector<double> ps(p.P);
  int st = dgbgk_point_fields(h, nullptr, nullptr, nullptr, nullptr, ps.data(), nullptr);
  if (st) return st;
  double l = 0.0, d = 0.0;
  for (size_t j = 0; j < p.force_p1.size(); ++j) {
    const int p1 = p.force_p1[j], p2 = p.force_p2[j];
    const double x1 = p.coord[2 * (size_t)p1], y1 = p.coord[2 * (size_t)p1 + 1];
    const double x2 = p.coord[2 * (size_t)p2], y2 = p.coord[2 * (size_t)p2 + 1];
    const double len = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    const double px = -(y2 - y1), py = x2 - x1, mag = std::sqrt(px * px + py * py);
    const double psav = 0.5 * (ps[p1] + ps[p2]);
    l -= psav * len * (py / mag);
    d -= psav * len * (px / mag);
  }
  if (lift) *lift = l;
  if (drag) *drag = d;
  return kOk;
}

int dgbgk_wall_eta(dgbgk_ctx* h, double* eta) {
  if (!h || !eta) return fail(kErrArg, "null argument");
  Ctx& c = h->c;
  HIPCK(hipMemcpyAsync(eta, c.eta, std::max(c.p.B, 0) * sizeof(double), hipMemcpyDeviceToHost,
                       c.stream));
  HIPCK(hipStreamSynchronize(c.stream));
  return kOk;
}

int dgbgk_sizes(dgbgk_ctx* h, int* nside, int* nvel, int* nelem, int* npoin, int* nboun,
                int* itime) {
  if (!h) return fail(kErrArg, "null context");
  const Ctx& c = h->c;
  if (nside) *nside = c.p.S;
  if (nvel) *nvel = c.p.V;
  if (nelem) *nelem = c.p.E;
  if (npoin) *npoin = c.p.P;
  if (nboun) *nboun = c.p.B;
  if (itime) *itime = c.itime;
  return kOk;
}

int dgbgk_setup_tables(dgbgk_ctx* h, int* iside, double* nx, double* ny, double* el,
                       double* geome, double* relen, double* mmat, double* rsido, double* vcord,
                       double* sumweight, double* dtec) {
  if (!h) return fail(kErrArg, "null context");
  const Ctx& c = h->c;
  if (c.eorder.empty()) {
    copy_tables(c.p, iside, nx, ny, el, geome, relen, mmat, rsido, vcord, sumweight, dtec);
    return kOk;
  }
  // the reference's tables are those of the input numbering
  Prep q;
  std::string msg;
  const Ctx::Input& in = c.in;
  int st = prepare(c.p.E, c.p.P, c.p.B, in.intma.data(), in.coord.data(), in.bsido.data(),
                   in.rorder, in.torder, in.eta.data(), in.w.data(), to_params(&in.prm), kSlots,
                   kSlots, q, msg);
  if (st) return fail(st, msg);
  copy_tables(q, iside, nx, ny, el, geome, relen, mmat, rsido, vcord, sumweight, dtec);
  return kOk;
}

int dgbgk_timing(dgbgk_ctx* h, int enable) {
  if (!h) return fail(kErrArg, "null context");
  Ctx& c = h->c;
  HIPCK(hipStreamSynchronize(c.stream));
  tharvest(c);
  c.timing = enable != 0;
  for (int k = 0; k < 8; ++k) {
    c.t_ms[k] = 0;
    c.t_cnt[k] = 0;
  }
  return kOk;
}

int dgbgk_kernel_time(dgbgk_ctx* h, int kernel, double* ms_total, long* count) {
  if (!h || kernel < 0 || kernel >= 8) return fail(kErrArg, "bad arguments");
  HIPCK(hipStreamSynchronize(h->c.stream));
  tharvest(h->c);
  if (ms_total) *ms_total = h->c.t_ms[kernel];
  if (count) *count = h->c.t_cnt[kernel];
  return kOk;
}

int dgbgk_sync(dgbgk_ctx* h) {
  if (!h) return fail(kErrArg, "null context");
  HIPCK(hipStreamSynchronize(h->c.stream));
  return kOk;
}

int dgbgk_f_stats(dgbgk_ctx* h, double* fmin, double* fmax, long long* nonfinite) {
  if (!h) return fail(kErrArg, "null context");
  Ctx& c = h->c;
  const size_t N3 = 3 * (size_t)c.p.E;
  const unsigned gx = (unsigned)std::min<size_t>(16, (N3 + 255) / 256);
  const size_t nb = (size_t)gx * c.p.V;
  double* d = nullptr;
  HIPCK(hipMalloc((void**)&d, 3 * nb * sizeof(double)));
  dim3 g(gx, c.p.V);
  if (c.prec)
    hipLaunchKernelGGL(k_fstats<float>, g, dim3(256), 0, c.stream, (const float*)c.f[c.cur],
                       c.p.E, (long long)c.ldf, d);
  else
    hipLaunchKernelGGL(k_fstats<double>, g, dim3(256), 0, c.stream, (const double*)c.f[c.cur],
                       c.p.E, (long long)c.ldf, d);
  std::vector<double> part(3 * nb);
  hipError_t e1 = hipMemcpyAsync(part.data(), d, 3 * nb * sizeof(double), hipMemcpyDeviceToHost,
                                 c.stream);
  hipError_t e2 = hipStreamSynchronize(c.stream);
  const hipError_t e3 = hipFree(d);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
    return fail(kErrHip, "f_stats failed");
  double lo = INFINITY, hi = -INFINITY, bad = 0.0;
  for (size_t b = 0; b < nb; ++b) {
    lo = std::min(lo, part[3 * b]);
    hi = std::max(hi, part[3 * b + 1]);
    bad += part[3 * b + 2];
  }
  if (fmin) *fmin = lo;
  if (fmax) *fmax = hi;
  if (nonfinite) *nonfinite = (long long)bad;
  return kOk;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};  // scratch's address on each device

// The scratch array's device address on the current device, looked up once
// per device with that device current.
// The address of device dev's copy of scratch, looked up once per device
// with that device current.
int fstats_scratch_ptr(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(kErrArg, "device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = 0;
    HIPCK(hipGetDevice(&cur));
    if (cur != dev) HIPCK(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCK(hipSetDevice(cur));
    HIPCK(e);
    g_scratch_dev[dev] = (double*)ptr;
  }
  *out = g_scratch_dev[dev];
  return kOk;
}
}  // namespace

int dgbgk_f_stats_noalloc(dgbgk_ctx* h, double* fmin, double* fmax, long long* nonfinite) {
  if (!h) return fail(kErrArg, "null context");
  Ctx& c = h->c;
  const size_t N3 = 3 * (size_t)c.p.E;
  const unsigned gx = (unsigned)std::min<size_t>(16, (N3 + 255) / 256);
  const size_t nb = (size_t)gx * c.p.V;  // <= 16 * 4096 = kFstatsMaxBlocks
  if (c.p.V > kFstatsMaxBlocks / 16 || nb > (size_t)kFstatsMaxBlocks)
    return fail(kErrArg, "dgbgk_f_stats_noalloc: " + std::to_string(c.p.V) +
                             " ordinates exceed its scratch (at most 4096)");
  double* sp = nullptr;  // the context's device's copy of scratch
  int st = fstats_scratch_ptr(c.dev, &sp);
  if (st) return st;
  double* part = sp;
  double* res = sp + 3 * (size_t)kFstatsMaxBlocks;
  dim3 g(gx, c.p.V);
  if (c.prec)
    hipLaunchKernelGGL(k_fstats<float>, g, dim3(256), 0, c.stream, (const float*)c.f[c.cur],
                       c.p.E, (long long)c.ldf, part);
  else
    hipLaunchKernelGGL(k_fstats<double>, g, dim3(256), 0, c.stream, (const double*)c.f[c.cur],
                       c.p.E, (long long)c.ldf, part);
  HIPCK(hipGetLastError());
  hipLaunchKernelGGL(k_fstats_combine, dim3(1), dim3(256), 0, c.stream, part, (int)nb, res);
  HIPCK(hipGetLastError());
  double r[3];
  HIPCK(hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, c.stream));
  HIPCK(hipStreamSynchronize(c.stream));
  if (fmin) *fmin = r[0];
  if (fmax) *fmax = r[1];
  if (nonfinite) *nonfinite = (long long)r[2];
  return kOk;
}

int dgbgk_dof_bytes(dgbgk_ctx* h, int* bytes_per_value) {
  if (!h) return fail(kErrArg, "null context");
  if (bytes_per_value) *bytes_per_value = (int)h->c.tsz;
  return kOk;
}

int dgbgk_launch_config(dgbgk_ctx* h, int* slots, int* depth, int* vchunk, int* nvchunks,
                        int* ntiles, int* mom_split) {
  if (!h) return fail(kErrArg, "null context");
  const Ctx& c = h->c;
  if (slots) *slots = c.slots;
  if (depth) *depth = c.depth;
  if (vchunk) *vchunk = c.vchunk;
  if (nvchunks) *nvchunks = c.nvchunks;
  if (ntiles) *ntiles = c.ntiles;
  if (mom_split) *mom_split = c.mom_split;
  return kOk;
}

}  // extern "C"
