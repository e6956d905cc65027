// kernels.hpp -- device kernels of the DG-BGK time march (gfx950).
//
// Data layout in HBM (one device, no partitioning):
//   f        T[V][ldf]  velocity-major; node (e,a) at 3e+a of each ordinate;
//                       rows padded to ldf = 3E rounded up to 64 values, so
//                       every row starts on a 256-byte boundary
//   geo      T[8][E]    dN/dx (3), dN/dy (3), area, 1/lumped mass  (SoA)
//                       (IMMAT = 0: row 7 is minus CONMAT's off-diagonal entry,
//                       its row sum)
//   einfo    int[3][E]  per element edge: kind | a1 | a2 | b1 | b2 | left
//   eslot    int[3][E]  interior: neighbour's LDS slot; wall: wall-side slot
//   enx/eny/ealen T[3][E] side normal and length/6 of each element edge
//   uinfo/unx/uny        the same edges for the update kernel: (len/6)
//                        times the outward normal (zero on wall sides, and
//                        on inflow sides in FP64), neighbour slot and nodes
//                        (see kUpd*)
//   np       T[5][3E]   per node: n*beta2/pi, beta2, u, v, mu (GTEQNF/COMPMU)
//   vtab     T[4][V]    ordinate velocity (ADVNCE form), inflow value, and
//                       the FP32 store dither's ordinate phase
//   mom      double[5][3E] DISND, DISUX, DISUY, DISPS, MU
#pragma once

#include <hip/hip_runtime.h>

namespace dgbgk {

constexpr int kSlots = 256;  // LDS slots (= threads) per update block

// einfo bit fields
__host__ __device__ constexpr int einfo_pack(int kind, int a1, int a2, int b1, int b2,
                                             int left) {
  return kind | (a1 << 2) | (a2 << 4) | (b1 << 6) | (b2 << 8) | (left << 10);
}

struct StepRec {  // same layout as dgbgk_step_info
  double sumres, logres, dte, mass, lift, drag, change;
};

// Update-kernel edge word (uinfo), one per element edge k (local nodes k,
// k1 = (k+1)%3): the flux source when the neighbour is upwind and the
// tie rule of the side's orientation.
constexpr int kUpdNbK = 0;      // bits 0-1: neighbour's local node at node k
constexpr int kUpdNbK1 = 2;     // bits 2-3: neighbour's local node at node k1
constexpr int kUpdRight = 1 << 4;  // this element is the side's right element
constexpr int kUpdOwn = 1 << 5;    // boundary side: own values are always upwind
constexpr int kUpdInflow = 1 << 6; // inflow side: upwind values are zero (FP32 form)
constexpr int kUpdSlot = 8;     // bits 8+: neighbour's LDS slot

template <typename T>
struct UpdArgs {
  const T* __restrict__ fcur;
  T* __restrict__ fnxt;
  const T* __restrict__ geo;
  const int* __restrict__ uinfo;   // [3][E] edge words
  const T* __restrict__ unx;       // [3][E] (len/6) x outward normal; 0 on inflow
  const T* __restrict__ uny;       //        and wall sides
  const int* __restrict__ emask;   // inflow bits 0..2
  const int* __restrict__ ewslot;  // wall element slot or -1
  const T* __restrict__ np;
  const T* __restrict__ vtab;
  const int* __restrict__ tiles;   // [ntiles][4] e0, nown, hoff, nhalo
  const int* __restrict__ halo;
  T* __restrict__ wpart;           // [NWE][V][8] partial RHS (3), COLL (3), U
  const double* __restrict__ vtd;  // [8][V]: cxa, cya, cxb, cyb, wv, wv cxb, wv cyb, ...
  double* __restrict__ msum;       // [nvchunks][3][3E] moment partials of f^{k+1}
  const double* __restrict__ scal; // [0] = DTE
  double* __restrict__ res_part;   // per block: max|increment|; +res_stride: max|change|
  const float* __restrict__ ephase; // [E] FP32 store dither phase per element
  int res_stride;
  int E, V, ntiles, vchunk;
  long long ldf;  // row stride of f
  int dither;     // FP32: dithered (unbiased) rounding of the f store
  int cmminc;     // IMMAT = 0: CMMINC's increment and residual (geo row 7 its row sum)
};

template <typename T>
struct WallArgs {
  const T* __restrict__ fcur;
  T* __restrict__ fnxt;
  const T* __restrict__ geo;
  const int* __restrict__ einfo;
  const int* __restrict__ eslot;
  const T* __restrict__ enx;
  const T* __restrict__ eny;
  const T* __restrict__ ealen;
  const int* __restrict__ emask;
  const T* __restrict__ vtab;
  const int* __restrict__ refl_i;
  const T* __restrict__ refl_t;
  const int* __restrict__ wall_bside;
  const double* __restrict__ eta;
  const int* __restrict__ wall_elems;
  const T* __restrict__ wpart;
  const double* __restrict__ scal;
  double* __restrict__ res_part;
  const float* __restrict__ ephase; // [E] FP32 store dither phase per element
  int res_stride;
  T alpha;
  int E, V, T_, ngroups, gsize, nwe;  // ngroups threads per wall element, gsize ordinates each
  int jacobi;    // reflections read f of the previous step only
  int diffuse;   // alpha = 1: the reflected part has weight 0 and is not read
  int lds_ring;  // one thread per speed ring, in-place reads staged in LDS
  long long ldf;  // row stride of f
  int dither;     // FP32: dithered (unbiased) rounding of the f store
  int cmminc;     // IMMAT = 0: as in UpdArgs
};

template <typename T>
struct MomArgs {
  const T* __restrict__ f;
  const double* __restrict__ vtd;  // [8][V]: cxa, cya, cxb, cyb, wv, wv cxb, wv cyb,
                                   //        0.5 wv |c_b|^2
  const T* __restrict__ vtab;      // [4][V]: cxa, cya as T (rows 0, 1)
  const double* __restrict__ rj;   // GEOME(7,e)
  double* __restrict__ mom;        // [5][3E]
  T* __restrict__ np;              // [5][3E]
  double* __restrict__ mu_part;
  double* __restrict__ mass_part;
  const double* __restrict__ msum;  // k_update's partials, or null: two passes
  const int* __restrict__ ewslot;   // wall elements have no partials
  double mm, sigt;
  int E, V, nchunks, vchunk;
  long long ldf;  // row stride of f
};

}  // namespace dgbgk
