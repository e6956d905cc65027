// prep.hpp -- host-side setup for the DG-BGK time march: the reference's mesh
// preprocessing (bit-exact side table and geometry) and the device tables
// derived from it (element edges, tiles, wall reflection table).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dgbgk.h"

namespace dgbgk {

// Reference constants (e.g. src/advnce.f90:94, src/gteqnf.f90:18).
constexpr double kPi = 3.1416;
constexpr double kNa = 6.022e26;

enum Status {
  kOk = 0,
  kErrArg = -1,
  kErrMxsid = -2,          // src/getsid.f90:137-140
  kErrBoundarySearch = -3, // src/getbou.f90:48-53 search runs past NBOUN
  kErrWallSide = -4,       // src/flucon2.f90:125-128
  kErrSumbottom = -5,      // src/flucon2.f90:227-231
  kErrReflectLookup = -6,  // src/getbou.f90:223-250 reads outside VCORD/DISNF
  kErrMesh = -7,           // ids out of range, degenerate element
  kErrUnsupported = -8,    // a configuration the library does not build
  kErrHip = -9,
  kErrNoMem = -10,
  kErrTile = -11,
  kErrState = -12,
};

struct Params {
  double rv, csafm, alpha, rgas, d, m;
  double t1, p1, u0, v0;
  double cinf[4];
  int inf, immat, precision, lookup, order, vchunk;
  int tile_align = 1;  // update tiles: owned elements a multiple of this
};

enum EdgeKind : int { kInterior = 0, kInflow = 1, kOutflow = 2, kWall = 3 };

struct ElemEdge {
  int kind;     // EdgeKind
  int nbr;      // 0-based neighbour element (interior) or -1
  int bside;    // 0-based boundary side (BSIDO column) or -1
  int wside;    // 0-based wall-side slot (kind == kWall) or -1
  int a1, a2;   // own local node (0..2) at side point 1 and point 2
  int b1, b2;   // neighbour local node at side point 1 and point 2
  int sign;     // +1: own element is ISIDE(3) (left), -1: it is ISIDE(4)
  double nx, ny, alen;  // side normal and length/6 used for this flux
};

struct Prep {
  int E = 0, P = 0, B = 0, S = 0, R = 0, T = 0, V = 0;
  Params prm{};
  // reference tables, 1-based ids, Fortran column-major
  std::vector<int> intma, bsido, iside;
  std::vector<double> coord, nx, ny, el, geome, relen, mmat, rsido, vcord;
  std::vector<double> cm12;  // IMMAT = 0: CONMAT's off-diagonal entry per element
  double sumweight = 0.0, dtec = 0.0;
  // per ordinate: velocity in the two radius forms the reference uses
  // (A: (rv/2)*(eta+1) in ADVNCE/COMPMU, B: eta*(rv/2)+rv/2 in MACROS,
  // FLUCON2, INICON2), the Jacobian and the moment weight w*jac*frac
  std::vector<double> cxa, cya, cxb, cyb, jac, wv, finf;
  // element edges (3 per element, local edge k runs from node k to k+1)
  std::vector<ElemEdge> edges;
  std::vector<uint8_t> inflow_mask;  // bit a set: node a reset by INFLOW
  // walls
  std::vector<int> wall_sides;       // wall slot -> boundary side (0-based)
  std::vector<int> wall_elem_slot;   // element -> wall element slot or -1
  std::vector<int> wall_elems;       // wall element slot -> element
  // reflection table, (wall slot, ordinate): i1/i2 0-based ordinates or -1
  // when the ordinate does not reflect (TEST >= 0); zd1/zd2 the reference's
  // interpolation distances; mw the wall Maxwellian factor
  std::vector<int> refl_i1, refl_i2;
  std::vector<double> refl_zd1, refl_zd2, refl_mw;
  bool ring_separable = true;  // every reflection reads its own speed ring
  // FLUCON2 sides (type 4): boundary side, element, local nodes, 1/bottom
  std::vector<int> eta_side, eta_elem, eta_in1, eta_in2;
  std::vector<double> eta_nx, eta_ny, eta_bottom;
  // forces: type-4 sides (point ids 0-based)
  std::vector<int> force_p1, force_p2;
  // point CSR: nodes (3*e+a) touching each point, in element order
  std::vector<int> pt_ptr, pt_node;
  // tiles of the update kernel
  std::vector<int> tile_e0, tile_nown, tile_hoff, tile_nhalo, halo_elems;
  std::vector<int> nbr_slot;  // 3 per element: LDS slot of the neighbour
};

// The C-ABI parameter block (include/dgbgk.h) as the host setup's Params.
Params params_from_api(const dgbgk_params& a);

// Builds everything; returns a Status and fills msg on error.
int prepare(int nelem, int npoin, int nboun, const int* intma, const double* coord,
            const int* bsido, int rorder, int torder, const double* eta,
            const double* w, const Params& prm, int tile_slots, int tile_owned, Prep& out,
            std::string& msg, const std::vector<int>* tile_bounds = nullptr);

// The reference's preprocessing steps, exposed for bit-exact checks.
int getsid(int nelem, int npoin, const int* intma, std::vector<int>& iside);

// Device storage order of the elements: the input elements sorted along a
// Hilbert curve through their centroids (stable in the input order).  The
// update kernel's tiles are runs of consecutive elements, so in this order
// they are compact patches whatever the input numbering.  Returns 1 and
// fills order (storage position -> input element, 0-based) when the order
// differs from the input's; 0 (order empty) when it is the input order or a
// point id is out of range (prepare then reports the mesh error).
int hilbert_order(int nelem, int npoin, const int* intma, const double* coord,
                  std::vector<int>& order);

// The mesh renumbered into a storage order: INTMA columns moved, BSIDO(3)
// (the side's element) renumbered.
void renumber_elements(int nelem, int nboun, const int* intma, const int* bsido,
                       const std::vector<int>& order, std::vector<int>& intma_out,
                       std::vector<int>& bsido_out);

// Refinement of a storage order whose tiles are built (p): inside each
// tile's owned range, the elements no other tile reads first, then those
// that other tiles read as halo, grouped by the lowest such tile, each group
// in storage order.  A tile's halo then reads runs of consecutive elements
// rather than elements scattered over its neighbours' ranges.  perm maps a
// refined position to a position of p's order; bounds are the tile starts
// and E, for prepare to rebuild the same tiles.
void border_order(const Prep& p, std::vector<int>& perm, std::vector<int>& bounds);

// Point CSR of a renumbered mesh back in input element order, the order in
// which MACROS accumulates the point averages (src/macros.f90:156-170).
void csr_input_order(Prep& p, const std::vector<int>& order);

// The whole host setup of a device context: the storage order (the Hilbert
// order when reorder, refined by border_order when border), the tables
// built on it, and the point CSR back in input order.  eorder: storage
// position -> input element, 0-based, empty when the input order is kept.
int prepare_storage(int nelem, int npoin, int nboun, const int* intma, const double* coord,
                    const int* bsido, int rorder, int torder, const double* eta,
                    const double* w, const Params& prm, int tile_slots, int tile_owned,
                    bool reorder, bool border, Prep& p, std::vector<int>& eorder,
                    std::string& msg);

}  // namespace dgbgk
