// prep.cpp -- host setup: the reference's mesh preprocessing restated with
// bit-identical results (compiled with -ffp-contract=off so every product and
// sum rounds as the Fortran's does), plus the tables the HIP kernels read.
#include "prep.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace dgbgk {

// Side table ISIDE(8,NSIDE) (src/getsid.f90:1-216).  Points are visited in
// order; around each point the elements are visited in element order and
// each element contributes the sides to its next (left side) and previous
// (right side) vertex with a larger point id.  A side seen only from its
// right is flipped so its element becomes the left one.
int getsid(int nelem, int npoin, const int* intma, std::vector<int>& iside) {
  std::vector<int> deg(npoin + 1, 0), start(npoin + 2, 0), cone(3 * (size_t)nelem);
  for (int k = 0; k < 3 * nelem; ++k) {
    int p = intma[k];
    if (p < 1 || p > npoin) return kErrMesh;
    deg[p]++;
  }
  for (int p = 1; p <= npoin; ++p) start[p + 1] = start[p] + deg[p];
  std::vector<int> fill(npoin + 1, 0);
  for (int e = 1; e <= nelem; ++e)
    for (int a = 0; a < 3; ++a) {
      int p = intma[3 * (e - 1) + a];
      cone[start[p] + fill[p]++] = e;
    }
  iside.clear();
  iside.reserve(8 * (size_t)(2 * nelem + 16));
  auto S = [&](size_t s, int k) -> int& { return iside[8 * s + k]; };
  for (int ip = 1; ip <= npoin; ++ip) {
    size_t first = iside.size() / 8;
    for (int c = start[ip]; c < start[ip + 1]; ++c) {
      int e = cone[c];
      const int* el = &intma[3 * (size_t)(e - 1)];
      int in1 = 0;
      for (int a = 0; a < 3; ++a) { in1 = a; if (el[a] == ip) break; }
      for (int j = 1; j <= 2; ++j) {
        int ip2 = el[(in1 + j) % 3];
        if (ip2 < ip) continue;
        size_t nside = iside.size() / 8, found = SIZE_MAX;
        for (size_t s = first; s < nside; ++s)
          if (S(s, 1) == ip2) { found = s; break; }
        if (found == SIZE_MAX) {
          iside.insert(iside.end(), 8, 0);
          S(nside, 0) = ip;
          S(nside, 1) = ip2;
          S(nside, 1 + j) = e;
        } else {
          S(found, 1 + j) = e;
        }
      }
    }
    size_t nside = iside.size() / 8;
    for (size_t s = first; s < nside; ++s) {
      if (S(s, 2) != 0) continue;
      S(s, 2) = S(s, 3);
      S(s, 3) = 0;
      S(s, 0) = S(s, 1);
      S(s, 1) = ip;
    }
  }
  size_t nside = iside.size() / 8;
  for (size_t s = 0; s < nside; ++s) {
    int ip1 = S(s, 0), ip2 = S(s, 1), iel = S(s, 2), ier = S(s, 3);
    const int* L = &intma[3 * (size_t)(iel - 1)];
    int iloc = -2;
    for (int a = 0; a < 3; ++a)
      if (L[a] == ip1) { iloc = a + 1; break; }
    S(s, 4) = iloc;
    S(s, 5) = (iloc == 3) ? 1 : iloc + 1;
    if (ier == 0) {
      S(s, 6) = 0;
      S(s, 7) = 0;
    } else {
      const int* Rr = &intma[3 * (size_t)(ier - 1)];
      for (int a = 0; a < 3; ++a)
        if (Rr[a] == ip2) { iloc = a + 1; break; }
      if (iloc == 3) { S(s, 6) = 1; S(s, 7) = 3; }
      else { S(s, 7) = iloc; S(s, 6) = iloc + 1; }
    }
  }
  return (int)nside;
}

namespace {

// Reflected ordinate of GETBOU (src/getbou.f90:190-250).  i1/i2 0-based.
int reflect_lookup(const Prep& p, double ux, double uy, double anx, double any,
                   int& i1, int& i2, double& zd1, double& zd2) {
  const int R = p.R, T = p.T, V = p.V;
  auto VC = [&](int k, int v1) { return p.vcord[3 * (size_t)(v1 - 1) + k]; };
  double test = ux * anx + uy * any;
  double cxr = ux + std::fabs(2 * test) * anx;
  double cyr = uy + std::fabs(2 * test) * any;
  double rt = std::sqrt(cxr * cxr + cyr * cyr);
  double quo = std::fabs(cxr / rt);
  double theta = std::fabs(std::acos(quo));
  if (cxr < 0.0) theta = kPi - theta;
  if (cyr < 0.0) theta = -theta;
  double etaref = ((2 * rt) / p.prm.rv) - 1;
  if (etaref > 1.0) etaref = 1.0;
  double zetaref = theta / kPi;
  int ring;
  if (p.prm.lookup == 0) {
    // reference rule: ETATEST = VCORD(1, 1+(IETA-1)*(RORDER-1)), ring IETA-1
    int ieta;
    for (ieta = 1; ieta <= R; ++ieta) {
      int idx = 1 + (ieta - 1) * (R - 1);
      if (idx < 1 || idx > V) return kErrReflectLookup;
      if (std::fabs(etaref - VC(0, idx)) < 0.0001) break;
    }
    ring = ieta - 1;
  } else {
    // corrected rule: the ring whose own eta matches
    int j;
    for (j = 1; j <= R; ++j)
      if (std::fabs(etaref - VC(0, (j - 1) * T + 1)) < 0.0001) break;
    if (j > R) return kErrReflectLookup;
    ring = j;
  }
  int iz;
  double d1 = 0.0, d2 = 0.0;
  for (iz = 1; iz <= T - 1; ++iz) {
    double z1 = VC(1, iz), z2 = VC(1, iz + 1);
    d1 = z2 - z1;
    d2 = zetaref - z1;
    if (zetaref >= z1 && zetaref <= z2) break;
  }
  int a, b;
  if (iz == T) {
    d1 = VC(1, 1) - (-1);
    d2 = zetaref - (-1);
    a = ring * T;
    b = (ring - 1) * T + 1;
  } else {
    a = (ring - 1) * T + iz;
    b = (ring - 1) * T + iz + 1;
  }
  if (a < 1 || a > V || b < 1 || b > V) return kErrReflectLookup;
  i1 = a - 1;
  i2 = b - 1;
  zd1 = d1;
  zd2 = d2;
  return kOk;
}

uint64_t pair_key(int a, int b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

}  // namespace

Params params_from_api(const dgbgk_params& a) {
  Params q;
  q.rv = a.rv; q.csafm = a.csafm; q.alpha = a.alpha; q.rgas = a.rgas;
  q.d = a.d; q.m = a.m; q.t1 = a.t1; q.p1 = a.p1; q.u0 = a.u0; q.v0 = a.v0;
  for (int k = 0; k < 4; ++k) q.cinf[k] = a.cinf[k];
  q.inf = a.inf; q.immat = a.immat; q.precision = a.precision;
  q.lookup = a.lookup; q.order = a.order; q.vchunk = a.vchunk;
  return q;
}

int prepare(int nelem, int npoin, int nboun, const int* intma, const double* coord,
            const int* bsido, int rorder, int torder, const double* eta,
            const double* w, const Params& prm, int tile_slots, int tile_owned, Prep& p,
            std::string& msg, const std::vector<int>* tile_bounds) {
  if (nelem < 1 || npoin < 3 || nboun < 0 || rorder < 1 || torder < 1) {
    msg = "invalid sizes";
    return kErrArg;
  }
  if (prm.immat != 0 && prm.immat != 1) {
    msg = "IMMAT must be 1 (lumped mass) or 0 (the consistent-mass branch)";
    return kErrArg;
  }
  p = Prep();
  p.E = nelem; p.P = npoin; p.B = nboun; p.R = rorder; p.T = torder;
  p.V = rorder * torder;
  p.prm = prm;
  p.intma.assign(intma, intma + 3 * (size_t)nelem);
  p.bsido.assign(bsido, bsido + 6 * (size_t)nboun);
  p.coord.assign(coord, coord + 2 * (size_t)npoin);
  for (int b = 0; b < nboun; ++b) {
    const int* bs = &p.bsido[6 * (size_t)b];
    if (bs[0] < 1 || bs[0] > npoin || bs[1] < 1 || bs[1] > npoin || bs[2] < 1 ||
        bs[2] > nelem) {
      msg = "boundary side " + std::to_string(b + 1) + " has an id out of range";
      return kErrMesh;
    }
  }
  int ns = getsid(nelem, npoin, p.intma.data(), p.iside);
  if (ns < 0) {
    msg = ns == kErrMxsid ? "MXSID is not big enough" : "connectivity out of range";
    return ns;
  }
  if (ns > 2 * (nelem + nboun)) {
    msg = "MXSID = 2*(NELEM+NBOUN) is not big enough";
    return kErrMxsid;
  }
  p.S = ns;
  const int E = nelem, B = nboun, S = ns, V = p.V, R = rorder, T = torder;
  auto IS = [&](int s1, int k1) { return p.iside[8 * (size_t)(s1 - 1) + (k1 - 1)]; };
  auto CO = [&](int k1, int p1) { return p.coord[2 * (size_t)(p1 - 1) + (k1 - 1)]; };
  auto BS = [&](int k1, int b1) { return p.bsido[6 * (size_t)(b1 - 1) + (k1 - 1)]; };
  // NORMAL (src/normal.f90:24-33)
  p.nx.resize(S); p.ny.resize(S); p.el.resize(S);
  for (int s = 1; s <= S; ++s) {
    int ip1 = IS(s, 1), ip2 = IS(s, 2);
    double dx = CO(1, ip2) - CO(1, ip1);
    double dy = CO(2, ip2) - CO(2, ip1);
    p.el[s - 1] = std::sqrt((dx * dx) + (dy * dy));
    p.nx[s - 1] = dy / p.el[s - 1];
    p.ny[s - 1] = -dx / p.el[s - 1];
  }
  // GETGEO (src/getgeo.f90:22-59)
  p.geome.resize(7 * (size_t)E);
  static const double nxi[3] = {-1.0, 1.0, 0.0}, net[3] = {-1.0, 0.0, 1.0};
  for (int e = 0; e < E; ++e) {
    const int* t = &p.intma[3 * (size_t)e];
    double x[3], y[3];
    for (int a = 0; a < 3; ++a) { x[a] = CO(1, t[a]); y[a] = CO(2, t[a]); }
    double x21 = x[1] - x[0], x31 = x[2] - x[0], y21 = y[1] - y[0], y31 = y[2] - y[0];
    double rj = x21 * y31 - x31 * y21;
    if (!(rj > 0.0)) {
      msg = "element " + std::to_string(e + 1) + " is not counter-clockwise (2A=" +
            std::to_string(rj) + ")";
      return kErrMesh;
    }
    double rj1 = 1. / rj;
    double xix = y31 * rj1, xiy = -x31 * rj1, etx = -y21 * rj1, ety = x21 * rj1;
    double* g = &p.geome[7 * (size_t)e];
    for (int a = 0; a < 3; ++a) {
      g[a] = xix * nxi[a] + etx * net[a];
      g[a + 3] = xiy * nxi[a] + ety * net[a];
    }
    g[6] = rj;
  }
  // GETLEN (src/getlen.f90:17-26); GETMAT (src/getmat.f90:8-40) for IMMAT =
  // 1, else CONMAT (src/conmat.f90:20-37, src/dg_lin_convec.f90:113-117),
  // kept as its off-diagonal entry -0.25/(RJ/24) (the diagonal is -3 times
  // it, the same double as the reference's 0.75/(RJ/24)); MMAT stays zero
  p.relen.resize(E);
  p.mmat.assign(3 * (size_t)E, 0.0);
  if (prm.immat == 0) p.cm12.resize(E);
  for (int e = 0; e < E; ++e) {
    const double* g = &p.geome[7 * (size_t)e];
    double dist[3];
    for (int a = 0; a < 3; ++a) dist[a] = 1.0 / std::sqrt(g[a] * g[a] + g[a + 3] * g[a + 3]);
    p.relen[e] = std::min(dist[0], std::min(dist[1], dist[2]));
    double rj24 = g[6] * (1. / 24.);
    if (prm.immat == 0) {
      p.cm12[e] = -0.25 * (1 / rj24);
      continue;
    }
    for (int a = 0; a < 3; ++a) {
      double m = 0.0;
      m = m + 4.0 * rj24;
      p.mmat[3 * (size_t)e + a] = 1. / m;
    }
  }
  // GETNOR (src/getnor.f90:21-31)
  p.rsido.resize(3 * (size_t)B);
  for (int b = 1; b <= B; ++b) {
    int p0 = BS(1, b), p1 = BS(2, b);
    double dx = CO(1, p1) - CO(1, p0), dy = CO(2, p1) - CO(2, p0);
    double rl = std::sqrt(dx * dx + dy * dy), rt = 1.0 / rl;
    p.rsido[3 * (size_t)(b - 1) + 0] = dy * rt;
    p.rsido[3 * (size_t)(b - 1) + 1] = -dx * rt;
    p.rsido[3 * (size_t)(b - 1) + 2] = rl;
  }
  // VSPACE (src/vspace.f90:16-49)
  p.vcord.resize(3 * (size_t)V);
  {
    double gap = 2.0 / (double)T, s = 0.0;
    int i = 0;
    for (int ie = 0; ie < R; ++ie)
      for (int iz = 1; iz <= T; ++iz) {
        p.vcord[3 * (size_t)i + 0] = eta[ie];
        p.vcord[3 * (size_t)i + 1] = -1 + iz * gap;
        p.vcord[3 * (size_t)i + 2] = w[ie] * 1.0;
        s = s + p.vcord[3 * (size_t)i + 2];
        ++i;
      }
    p.sumweight = s;
  }
  // ALOTIM with MXSPD = rv (src/alotim.f90:12-18, src/advnce.f90:133)
  for (int e = 0; e < E; ++e) {
    double d = (prm.csafm * p.relen[e]) / (prm.rv);
    if (e == 0 || d < p.dtec) p.dtec = d;
  }
  // per-ordinate tables
  const double frac = 4.0 / p.sumweight;
  p.cxa.resize(V); p.cya.resize(V); p.cxb.resize(V); p.cyb.resize(V);
  p.jac.resize(V); p.wv.resize(V); p.finf.resize(V);
  double cinf[4];
  for (int k = 0; k < 4; ++k) cinf[k] = prm.inf == 1 ? prm.cinf[k] : 0.0;
  for (int v = 0; v < V; ++v) {
    double et = p.vcord[3 * (size_t)v], ze = p.vcord[3 * (size_t)v + 1];
    double th = ze * kPi;
    double ra = (prm.rv / 2) * (et + 1), rb = et * (prm.rv / 2) + (prm.rv / 2);
    p.cxa[v] = ra * std::cos(th); p.cya[v] = ra * std::sin(th);
    p.cxb[v] = rb * std::cos(th); p.cyb[v] = rb * std::sin(th);
    p.jac[v] = kPi * prm.rv * prm.rv * 0.25 * (et + 1);
    p.wv[v] = p.vcord[3 * (size_t)v + 2] * p.jac[v] * frac;
    // INFLOW value (src/inflow.f90:28-39)
    double tin = cinf[0], pin = cinf[1] * 100000.0;
    double rho = pin / (prm.rgas * tin), beta = std::sqrt(rho / (2 * pin));
    double c0 = (rho * kNa) / prm.m, c1 = (beta * beta) / kPi;
    double ud = p.cxa[v] - cinf[2], vd = p.cya[v] - cinf[3];
    p.finf[v] = c0 * c1 * std::exp(-(beta * beta) * (ud * ud + vd * vd));
  }
  // GETBOU's boundary-side search: first ISB with BSIDO(3)=IEL, BSIDO(1)=IP1
  std::unordered_map<uint64_t, int> bou_first;
  bou_first.reserve(2 * (size_t)B + 1);
  for (int b = B; b >= 1; --b)
    bou_first[((uint64_t)(uint32_t)BS(3, b) << 32) | (uint32_t)BS(1, b)] = b;
  // element edges
  p.edges.assign(3 * (size_t)E, ElemEdge{});
  std::vector<uint8_t> seen(3 * (size_t)E, 0);
  std::vector<int> side_isb(S, 0);
  for (int s = 1; s <= S; ++s) {
    int iel = IS(s, 3), ier = IS(s, 4);
    int inl1 = IS(s, 5), inl2 = IS(s, 6), inr1 = IS(s, 7), inr2 = IS(s, 8);
    if (inl1 < 1 || inl1 > 3) { msg = "side table corrupt"; return kErrMesh; }
    int kl = inl1 - 1;  // local edge of the left element starts at INL1
    ElemEdge& L = p.edges[3 * (size_t)(iel - 1) + kl];
    if (seen[3 * (size_t)(iel - 1) + kl]++) { msg = "element edge used twice"; return kErrMesh; }
    L.a1 = inl1 - 1; L.a2 = inl2 - 1; L.sign = +1; L.wside = -1;
    if (ier == 0) {
      auto it = bou_first.find(((uint64_t)(uint32_t)iel << 32) | (uint32_t)IS(s, 1));
      if (it == bou_first.end()) {
        msg = "boundary side (" + std::to_string(IS(s, 1)) + "," + std::to_string(IS(s, 2)) +
              ") of element " + std::to_string(iel) + " is not in BSIDO";
        return kErrBoundarySearch;
      }
      int isb = it->second;
      side_isb[s - 1] = isb;
      int type = BS(4, isb);
      L.kind = type == 1 ? kInflow : (type == 2 ? kOutflow : kWall);
      L.nbr = -1; L.bside = isb - 1; L.b1 = L.b2 = -1;
      L.nx = p.rsido[3 * (size_t)(isb - 1)];
      L.ny = p.rsido[3 * (size_t)(isb - 1) + 1];
      L.alen = (1. / 6.) * p.rsido[3 * (size_t)(isb - 1) + 2];
    } else {
      int kr = inr2 - 1;  // right element's edge runs INR2 -> INR1
      ElemEdge& Rr = p.edges[3 * (size_t)(ier - 1) + kr];
      if (seen[3 * (size_t)(ier - 1) + kr]++) { msg = "element edge used twice"; return kErrMesh; }
      double alen = (1.0 / 6.0) * p.el[s - 1];
      L.kind = kInterior; L.nbr = ier - 1; L.bside = -1;
      L.b1 = inr1 - 1; L.b2 = inr2 - 1;
      L.nx = p.nx[s - 1]; L.ny = p.ny[s - 1]; L.alen = alen;
      Rr.kind = kInterior; Rr.nbr = iel - 1; Rr.bside = -1; Rr.wside = -1;
      Rr.a1 = inr1 - 1; Rr.a2 = inr2 - 1; Rr.b1 = inl1 - 1; Rr.b2 = inl2 - 1;
      Rr.sign = -1; Rr.nx = p.nx[s - 1]; Rr.ny = p.ny[s - 1]; Rr.alen = alen;
    }
  }
  for (size_t k = 0; k < seen.size(); ++k)
    if (!seen[k]) { msg = "element edge without a side"; return kErrMesh; }
  // the update kernel relies on: local edge k runs from node k to node
  // (k+1)%3, and the side's points (p1, p2) are those nodes in that order
  // for the left element and reversed for the right element
  for (int e = 0; e < E; ++e)
    for (int k = 0; k < 3; ++k) {
      const ElemEdge& ed = p.edges[3 * (size_t)e + k];
      const int k1 = (k + 1) % 3;
      const bool ok = ed.sign > 0 ? (ed.a1 == k && ed.a2 == k1) : (ed.a1 == k1 && ed.a2 == k);
      if (!ok) { msg = "inconsistent element edge orientation"; return kErrMesh; }
    }
  // inflow node mask: every DG node on a point of a type-1 side
  {
    std::vector<uint8_t> ptin(npoin + 1, 0);
    for (int b = 1; b <= B; ++b)
      if (BS(4, b) == 1) { ptin[BS(1, b)] = 1; ptin[BS(2, b)] = 1; }
    p.inflow_mask.assign(E, 0);
    for (int e = 0; e < E; ++e)
      for (int a = 0; a < 3; ++a)
        if (ptin[p.intma[3 * (size_t)e + a]]) p.inflow_mask[e] |= (uint8_t)(1u << a);
    bool any_inflow = false;
    for (int b = 1; b <= B; ++b) any_inflow |= BS(4, b) == 1;
    if (any_inflow && !(cinf[0] * cinf[1] > 0.0)) {
      msg = "inflow sides present but the inflow state (CINF/INF) is empty";
      return kErrArg;
    }
  }
  // wall sides and wall elements; reflection table
  {
    std::unordered_map<int, int> wslot;
    p.wall_elem_slot.assign(E, -1);
    for (int e = 0; e < E; ++e)
      for (int k = 0; k < 3; ++k) {
        ElemEdge& ed = p.edges[3 * (size_t)e + k];
        if (ed.kind != kWall) continue;
        auto it = wslot.find(ed.bside);
        if (it == wslot.end()) {
          it = wslot.emplace(ed.bside, (int)p.wall_sides.size()).first;
          p.wall_sides.push_back(ed.bside);
        }
        ed.wside = it->second;
        if (p.wall_elem_slot[e] < 0) {
          p.wall_elem_slot[e] = (int)p.wall_elems.size();
          p.wall_elems.push_back(e);
        }
      }
    const size_t NW = p.wall_sides.size();
    p.refl_i1.assign(NW * V, -1); p.refl_i2.assign(NW * V, -1);
    p.refl_zd1.assign(NW * V, 0.0); p.refl_zd2.assign(NW * V, 0.0);
    p.refl_mw.assign(NW * V, 0.0);
    for (size_t ws = 0; ws < NW; ++ws) {
      int b = p.wall_sides[ws];
      double anx = p.rsido[3 * (size_t)b], any = p.rsido[3 * (size_t)b + 1];
      double temp = (double)p.bsido[6 * (size_t)b + 4];
      for (int v = 0; v < V; ++v) {
        double ux = p.cxa[v], uy = p.cya[v];
        double test = ux * anx + uy * any;
        if (test >= 0.0) continue;
        int i1, i2;
        double zd1, zd2;
        int st = reflect_lookup(p, ux, uy, anx, any, i1, i2, zd1, zd2);
        if (st != kOk) {
          msg = "wall reflection of ordinate " + std::to_string(v + 1) +
                " falls outside the velocity grid (the reference ring lookup "
                "needs RORDER == TORDER and a first radial node at -1)";
          return st;
        }
        size_t k = ws * V + v;
        p.refl_i1[k] = i1; p.refl_i2[k] = i2;
        p.refl_zd1[k] = zd1; p.refl_zd2[k] = zd2;
        double speed = std::sqrt(ux * ux + uy * uy);
        p.refl_mw[k] = std::exp(-(speed * speed) / (2 * prm.rgas * temp));
        if (i1 / T != v / T || i2 / T != v / T) p.ring_separable = false;
      }
    }
  }
  // FLUCON2 sides: first side in ISIDE with the same two points
  {
    std::unordered_map<uint64_t, int> side_of;
    side_of.reserve(2 * (size_t)S + 1);
    for (int s = S; s >= 1; --s) side_of[pair_key(IS(s, 1), IS(s, 2))] = s;
    for (int b = 1; b <= B; ++b) {
      if (BS(4, b) != 4) continue;
      int ip1 = BS(1, b), ip2 = BS(2, b);
      auto it = side_of.find(pair_key(ip1, ip2));
      if (it == side_of.end()) {
        msg = "wall side " + std::to_string(b) + " not found in the side table";
        return kErrWallSide;
      }
      int s = it->second, in1, in2;
      if (IS(s, 1) == ip1 && IS(s, 2) == ip2) { in1 = IS(s, 5); in2 = IS(s, 6); }
      else { in1 = IS(s, 6); in2 = IS(s, 5); }
      int tw = BS(5, b);
      double r = prm.rgas;
      double bottom = std::sqrt(2 * kPi * (r * r * r) * (double)(tw * tw * tw));
      if (bottom < 0.1) {
        msg = "error calculating flux conservation parameter (SUMBOTTOM < 0.1)";
        return kErrSumbottom;
      }
      p.eta_side.push_back(b - 1);
      p.eta_elem.push_back(BS(3, b) - 1);
      p.eta_in1.push_back(in1 - 1);
      p.eta_in2.push_back(in2 - 1);
      p.eta_nx.push_back(p.rsido[3 * (size_t)(b - 1)]);
      p.eta_ny.push_back(p.rsido[3 * (size_t)(b - 1) + 1]);
      p.eta_bottom.push_back(bottom);
      p.force_p1.push_back(ip1 - 1);
      p.force_p2.push_back(ip2 - 1);
    }
  }
  // point CSR in element order (MACROS accumulation order)
  p.pt_ptr.assign(npoin + 1, 0);
  for (int k = 0; k < 3 * E; ++k) p.pt_ptr[p.intma[k]]++;
  for (int q = 0; q < npoin; ++q) p.pt_ptr[q + 1] += p.pt_ptr[q];
  p.pt_node.assign(3 * (size_t)E, 0);
  {
    std::vector<int> pos(p.pt_ptr.begin(), p.pt_ptr.end() - 1);
    for (int k = 0; k < 3 * E; ++k) p.pt_node[pos[p.intma[k] - 1]++] = k;
  }
  // tiles: contiguous owned ranges of at most tile_owned elements whose
  // owned + halo elements fit the slots (tile_owned a multiple of the
  // wavefront size keeps halo slots out of the owned wavefronts), or the
  // ranges given by tile_bounds (tile starts, then E)
  {
    if (tile_owned <= 0 || tile_owned > tile_slots) tile_owned = tile_slots;
    const bool forced = tile_bounds && !tile_bounds->empty();
    if (forced && (tile_bounds->front() != 0 || tile_bounds->back() != E)) {
      msg = "tile bounds do not cover the mesh";
      return kErrTile;
    }
    p.nbr_slot.assign(3 * (size_t)E, -1);
    std::vector<int> mark(E, -1), hidx(E, -1);
    int e0 = 0;
    int tile = 0;
    while (e0 < E) {
      int nh = 0, e1 = e0;
      if (forced) {
        e1 = tile + 1 < (int)tile_bounds->size() ? (*tile_bounds)[tile + 1] : e0;
        if (e1 <= e0 || e1 > E) { msg = "tile bounds out of order"; return kErrTile; }
      }
      // grow [e0, e1)
      while (!forced && e1 < E) {
        // tentative halo after adding e1
        int add = 0;
        for (int k = 0; k < 3; ++k) {
          int nb = p.edges[3 * (size_t)e1 + k].nbr;
          if (nb >= 0 && (nb < e0 || nb > e1) && mark[nb] != tile) ++add;
        }
        int rem = (mark[e1] == tile) ? 1 : 0;
        int nown = e1 - e0 + 1;
        if (nown > tile_owned || nown + nh + add - rem > tile_slots) break;
        if (rem) { mark[e1] = -2; --nh; }
        for (int k = 0; k < 3; ++k) {
          int nb = p.edges[3 * (size_t)e1 + k].nbr;
          if (nb >= 0 && (nb < e0 || nb > e1) && mark[nb] != tile) { mark[nb] = tile; ++nh; }
        }
        ++e1;
      }
      if (e1 == e0) { msg = "tile construction failed"; return kErrTile; }
      // owned counts a multiple of tile_align (all tiles but the last), so
      // tile boundaries fall on cache-line boundaries of every f row; the
      // elements given back can only replace halo entries, so the slots
      // still suffice
      if (!forced && prm.tile_align > 1 && e1 < E && e1 - e0 >= prm.tile_align)
        e1 = e0 + (e1 - e0) / prm.tile_align * prm.tile_align;
      int nown = e1 - e0;
      // collect halo elements (neighbours outside the owned range), then
      // give them slots in storage order: halo slots of one wavefront then
      // load runs of consecutive elements
      int hoff = (int)p.halo_elems.size(), nhal = 0;
      for (int e = e0; e < e1; ++e)
        for (int k = 0; k < 3; ++k) {
          int nb = p.edges[3 * (size_t)e + k].nbr;
          if (nb < 0 || (nb >= e0 && nb < e1)) continue;
          if (hidx[nb] < hoff || hidx[nb] >= hoff + nhal || p.halo_elems[hidx[nb]] != nb) {
            hidx[nb] = hoff + nhal;
            p.halo_elems.push_back(nb);
            ++nhal;
          }
        }
      std::sort(p.halo_elems.begin() + hoff, p.halo_elems.end());
      for (int j = 0; j < nhal; ++j) hidx[p.halo_elems[hoff + j]] = hoff + j;
      for (int e = e0; e < e1; ++e)
        for (int k = 0; k < 3; ++k) {
          int nb = p.edges[3 * (size_t)e + k].nbr;
          if (nb < 0) continue;
          p.nbr_slot[3 * (size_t)e + k] = (nb >= e0 && nb < e1) ? nb - e0 : nown + (hidx[nb] - hoff);
        }
      if (nown + nhal > tile_slots) { msg = "tile overflow"; return kErrTile; }
      p.tile_e0.push_back(e0);
      p.tile_nown.push_back(nown);
      p.tile_hoff.push_back(hoff);
      p.tile_nhalo.push_back(nhal);
      e0 = e1;
      ++tile;
    }
  }
  return kOk;
}

namespace {

// Index along the Hilbert curve of order `bits` of the cell (xi, yi), by the
// usual quadrant-and-rotate recursion from the top bit down.
uint64_t hilbert_d(int64_t xi, int64_t yi, int bits) {
  const int64_t n = int64_t(1) << bits;
  uint64_t d = 0;
  for (int64_t s = n >> 1; s > 0; s >>= 1) {
    const int64_t rx = (xi & s) ? 1 : 0, ry = (yi & s) ? 1 : 0;
    d += (uint64_t)(s * s) * (uint64_t)((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        xi = n - 1 - xi;
        yi = n - 1 - yi;
      }
      std::swap(xi, yi);
    }
  }
  return d;
}

}  // namespace

int hilbert_order(int nelem, int npoin, const int* intma, const double* coord,
                  std::vector<int>& order) {
  order.clear();
  if (nelem < 2) return 0;
  const size_t E = (size_t)nelem;
  for (size_t k = 0; k < 3 * E; ++k)
    if (intma[k] < 1 || intma[k] > npoin) return 0;
  // centroids ((x1 + x2) + x3) / 3 over the input's node order, scaled by
  // the larger side of their bounding box into 2^16 cells a side
  std::vector<double> cx(E), cy(E);
  double lox = 0.0, loy = 0.0, hix = 0.0, hiy = 0.0;
  for (size_t e = 0; e < E; ++e) {
    const int* t = &intma[3 * e];
    const double* p0 = &coord[2 * (size_t)(t[0] - 1)];
    const double* p1 = &coord[2 * (size_t)(t[1] - 1)];
    const double* p2 = &coord[2 * (size_t)(t[2] - 1)];
    cx[e] = (p0[0] + p1[0] + p2[0]) / 3.0;
    cy[e] = (p0[1] + p1[1] + p2[1]) / 3.0;
    if (e == 0 || cx[e] < lox) lox = cx[e];
    if (e == 0 || cy[e] < loy) loy = cy[e];
    if (e == 0 || cx[e] > hix) hix = cx[e];
    if (e == 0 || cy[e] > hiy) hiy = cy[e];
  }
  if (!(std::isfinite(lox) && std::isfinite(loy) && std::isfinite(hix) && std::isfinite(hiy)))
    return 0;
  const int bits = 16;
  const int64_t n = int64_t(1) << bits;
  const double span = std::max(std::max(hix - lox, hiy - loy), 1e-300);
  std::vector<uint64_t> key(E);
  for (size_t e = 0; e < E; ++e) {
    const int64_t xi = std::min((int64_t)((cx[e] - lox) / span * (double)(n - 1)), n - 1);
    const int64_t yi = std::min((int64_t)((cy[e] - loy) / span * (double)(n - 1)), n - 1);
    key[e] = hilbert_d(xi, yi, bits);
  }
  order.resize(E);
  for (size_t e = 0; e < E; ++e) order[e] = (int)e;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  for (size_t e = 0; e < E; ++e)
    if (order[e] != (int)e) return 1;
  order.clear();
  return 0;
}

void renumber_elements(int nelem, int nboun, const int* intma, const int* bsido,
                       const std::vector<int>& order, std::vector<int>& intma_out,
                       std::vector<int>& bsido_out) {
  const size_t E = (size_t)nelem;
  std::vector<int> inv(E);
  for (size_t s = 0; s < E; ++s) inv[order[s]] = (int)s;
  intma_out.resize(3 * E);
  for (size_t s = 0; s < E; ++s)
    for (int a = 0; a < 3; ++a) intma_out[3 * s + a] = intma[3 * (size_t)order[s] + a];
  bsido_out.assign(bsido, bsido + 6 * (size_t)std::max(nboun, 0));
  for (int b = 0; b < nboun; ++b) {
    int& e = bsido_out[6 * (size_t)b + 2];
    if (e >= 1 && e <= nelem) e = inv[e - 1] + 1;  // out of range: prepare reports it
  }
}

void border_order(const Prep& p, std::vector<int>& perm, std::vector<int>& bounds) {
  const int E = p.E, nt = (int)p.tile_e0.size();
  // the lowest tile that reads each element as halo, or -1
  std::vector<int> reader(E, -1);
  for (int t = nt - 1; t >= 0; --t)
    for (int j = 0; j < p.tile_nhalo[t]; ++j) reader[p.halo_elems[p.tile_hoff[t] + j]] = t;
  perm.resize(E);
  for (int e = 0; e < E; ++e) perm[e] = e;
  for (int t = 0; t < nt; ++t) {
    const int e0 = p.tile_e0[t], e1 = e0 + p.tile_nown[t];
    std::stable_sort(perm.begin() + e0, perm.begin() + e1,
                     [&](int a, int b) { return reader[a] < reader[b]; });
  }
  bounds.assign(p.tile_e0.begin(), p.tile_e0.end());
  bounds.push_back(E);
}

void csr_input_order(Prep& p, const std::vector<int>& order) {
  for (int q = 0; q < p.P; ++q)
    std::stable_sort(p.pt_node.begin() + p.pt_ptr[q], p.pt_node.begin() + p.pt_ptr[q + 1],
                     [&](int a, int b) { return order[a / 3] < order[b / 3]; });
}

int prepare_storage(int nelem, int npoin, int nboun, const int* intma, const double* coord,
                    const int* bsido, int rorder, int torder, const double* eta,
                    const double* w, const Params& prm, int tile_slots, int tile_owned,
                    bool reorder, bool border, Prep& p, std::vector<int>& eorder,
                    std::string& msg) {
  // the device works on the mesh renumbered into storage order (compact
  // tiles); the ABI's per-element arrays stay in the input numbering
  eorder.clear();
  const int* im = intma;
  const int* bs = bsido;
  std::vector<int> im_r, bs_r;
  if (reorder && hilbert_order(nelem, npoin, intma, coord, eorder)) {
    renumber_elements(nelem, nboun, intma, bsido, eorder, im_r, bs_r);
    im = im_r.data();
    bs = bs_r.data();
  }
  int st = prepare(nelem, npoin, nboun, im, coord, bs, rorder, torder, eta, w, prm, tile_slots,
                   tile_owned, p, msg);
  if (st) return st;
  // inside each tile, the elements that other tiles read as halo last,
  // grouped by reader; the same tiles are rebuilt on that order
  if (border) {
    std::vector<int> perm, bounds;
    border_order(p, perm, bounds);
    bool moved = false;
    for (int i = 0; i < nelem && !moved; ++i) moved = perm[i] != i;
    if (moved) {
      std::vector<int> fin(nelem);
      for (int i = 0; i < nelem; ++i) fin[i] = eorder.empty() ? perm[i] : eorder[perm[i]];
      eorder.swap(fin);
      renumber_elements(nelem, nboun, intma, bsido, eorder, im_r, bs_r);
      st = prepare(nelem, npoin, nboun, im_r.data(), coord, bs_r.data(), rorder, torder, eta,
                   w, prm, tile_slots, tile_owned, p, msg, &bounds);
      if (st) return st;
    }
  }
  if (!eorder.empty()) csr_input_order(p, eorder);
  return kOk;
}

}  // namespace dgbgk
