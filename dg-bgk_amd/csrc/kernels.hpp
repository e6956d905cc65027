// kernels.hpp -- device kernels of the DG-BGK time march (gfx950).
//
// Data layout in HBM (one device, no partitioning):
//   f        T[V][3E]   velocity-major; node (e,a) at 3e+a of each ordinate
//   geo      T[8][E]    dN/dx (3), dN/dy (3), area, 1/lumped mass  (SoA)
//   einfo    int[3][E]  per element edge: kind | a1 | a2 | b1 | b2 | left
//   eslot    int[3][E]  interior: neighbour's LDS slot; wall: wall-side slot
//   enx/eny/ealen T[3][E] flux normal and length/6 of each element edge
//   np       T[5][3E]   per node: n*beta2/pi, beta2, u, v, mu (GTEQNF/COMPMU)
//   vtab     T[3][V]    ordinate velocity (ADVNCE form) and inflow value
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

struct StepRec {
  double sumres, logres, dte, mass, lift, drag;
};

template <typename T>
struct UpdArgs {
  const T* __restrict__ fcur;
  T* __restrict__ fnxt;
  const T* __restrict__ geo;
  const int* __restrict__ einfo;
  const int* __restrict__ eslot;
  const T* __restrict__ enx;
  const T* __restrict__ eny;
  const T* __restrict__ ealen;
  const int* __restrict__ emask;   // inflow bits 0..2
  const int* __restrict__ ewslot;  // wall element slot or -1
  const T* __restrict__ np;
  const T* __restrict__ vtab;
  const int* __restrict__ refl_i;  // [NW][V][2]
  const T* __restrict__ refl_t;    // [NW][V][3] zd1, zd2, mw
  const int* __restrict__ wall_bside;
  const double* __restrict__ eta;  // per boundary side
  const int* __restrict__ tiles;   // [ntiles][4] e0, nown, hoff, nhalo
  const int* __restrict__ halo;
  T* __restrict__ wpart;           // [NWE][V][6] partial RHS and COLL
  const double* __restrict__ scal; // [0] = DTE
  double* __restrict__ res_part;
  T alpha;
  int E, V, ntiles, vchunk;
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
  T alpha;
  int E, V, T_, ngroups, nwe;
};

template <typename T>
struct MomArgs {
  const T* __restrict__ f;
  const double* __restrict__ vtd;  // [5][V]: cxa, cya, cxb, cyb, wv
  const double* __restrict__ rj;   // GEOME(7,e)
  double* __restrict__ mom;        // [5][3E]
  T* __restrict__ np;              // [5][3E]
  double* __restrict__ mu_part;
  double* __restrict__ mass_part;
  double mm, sigt;
  int E, V;
};

}  // namespace dgbgk
