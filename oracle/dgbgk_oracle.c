/*
 * dgbgk_oracle.c -- FP64 CPU restatement of the DG-BGK explicit time march
 * (reference: DrBenEvans/DG-BGK, Fortran 90 + MPI, run at one rank).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and the
 * "port" CPU baseline.  Nothing in the product library calls this code.
 *
 * The restatement keeps the reference's order of operations: the ordinate
 * loop updates f in place (so a wall reflection reads ordinates that were
 * already advanced in the same step), sums run in the reference order, and
 * the reference's literal constants (PI = 3.1416, NA = 6.022e26) are kept.
 * Evaluation order inside each expression follows Fortran's left-to-right
 * rule for operators of equal precedence; compile with -ffp-contract=off.
 *
 * Deliberate differences from the Fortran, none of which changes a value:
 *  - per-ordinate COS/SIN/JAC are tabulated once instead of recomputed per
 *    DG node (same arguments, same libm, so the same bits);
 *  - unless faithful_scans is set, the reference's linear searches (GETBOU
 *    boundary-side search, FLUCON2 side search, INFLOW element scan) are
 *    replaced by tables built once with the same first-match rule;
 *  - conditions on which the reference STOPs, or reads out of bounds, return
 *    a negative status instead.
 * Vacuum nodes (DISND < 1e-20) leave COMPMU's CT uninitialised in the
 * reference (src/compmu.f90:55-76); here CT = 0 there (parity unpinned).
 */
#include "dgbgk_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PI_R 3.1416     /* PARAMETER(PI=3.1416), e.g. src/advnce.f90:94 */
#define NA_R 6.022e26   /* PARAMETER(NA=6.022E+26), src/gteqnf.f90:18 */

#define INTMA(j, e) intma[3 * ((size_t)(e)-1) + (j)-1]
#define ISIDE(k, s) iside[8 * ((size_t)(s)-1) + (k)-1]
#define BSIDO(k, b) bsido[6 * ((size_t)(b)-1) + (k)-1]
#define GEOME(k, e) geome[7 * ((size_t)(e)-1) + (k)-1]
#define VCORD(k, v) vcord[3 * ((size_t)(v)-1) + (k)-1]
#define RSIDO(k, b) rsido[3 * ((size_t)(b)-1) + (k)-1]
#define COORD(k, p) coord[2 * ((size_t)(p)-1) + (k)-1]
#define NE(a, e) (3 * ((size_t)(e)-1) + (a)-1) /* (NNODE,NELEM) arrays */

/* FP32 emulation of chosen stages of the step (ora_set_emulation; tests and
 * precision studies only, serial step only).  A(x) rounds one arithmetic
 * result to FP32 when the step's arithmetic is emulated. */
static int g_emu = 0;
static inline double emu_round(int flag, double x) { return (g_emu & flag) ? (double)(float)x : x; }
#define A(x) emu_round(ORA_EMU_ARITH, (x))
#define AP(x) emu_round(ORA_EMU_ARITH | ORA_EMU_PRED, (x))
#define AE(x) emu_round(ORA_EMU_ARITH | ORA_EMU_EDGE, (x))
#define AI(x) emu_round(ORA_EMU_ARITH | ORA_EMU_INC, (x))
#define P(x) emu_round(ORA_EMU_PARAMS, (x))

/* ------------------------------------------------------------------------ */
/* GETSID: side table ISIDE(8,NSIDE) (src/getsid.f90:1-216).                 */
/* ------------------------------------------------------------------------ */
int ora_getsid(int nelem, int npoin, const int *intma, int *iside, int mxsid)
{
  int *lhowm = calloc((size_t)npoin + 1, sizeof(int));
  int *lwher = calloc((size_t)npoin + 1, sizeof(int));
  int *icone = calloc(3 * (size_t)nelem + 1, sizeof(int));
  int iloca = 0, status = 0;
  if (!lhowm || !lwher || !icone) { status = -100; goto done; }
  /* LHOWM: elements per point (src/getsid.f90:14-24) */
  for (int ie = 1; ie <= nelem; ie++)
    for (int in = 1; in <= 3; in++) lhowm[INTMA(in, ie)] += 1;
  /* LWHER: start of each point in ICONE (:28-31) */
  lwher[1] = 0;
  for (int ip = 2; ip <= npoin; ip++) lwher[ip] = lwher[ip - 1] + lhowm[ip - 1];
  /* ICONE: elements around each point, in element order (:35-46) */
  for (int ip = 1; ip <= npoin; ip++) lhowm[ip] = 0;
  for (int ie = 1; ie <= nelem; ie++)
    for (int in = 1; in <= 3; in++) {
      int ip = INTMA(in, ie);
      lhowm[ip] += 1;
      icone[lwher[ip] + lhowm[ip]] = ie;
    }
  /* loop over points (:52-133) */
  for (int ip = 1; ip <= npoin; ip++) {
    int iloc1 = iloca, iele = lhowm[ip];
    if (iele == 0) continue;
    for (int is = 1; is <= iele + 2; is++) {
      if (is + iloc1 > mxsid) { status = -1; goto done; }
      ISIDE(3, is + iloc1) = 0;
      ISIDE(4, is + iloc1) = 0;
    }
    int iwher = lwher[ip], ip1 = ip;
    for (int iel = 1; iel <= iele; iel++) {
      int ie = icone[iwher + iel];
      int in1 = 1;
      for (int in = 1; in <= 3; in++) { in1 = in; if (INTMA(in, ie) == ip) break; }
      for (int j = 1; j <= 2; j++) {
        int in2 = in1 + j;
        if (in2 > 3) in2 -= 3;
        int ip2 = INTMA(in2, ie);
        if (ip2 < ip1) continue;
        int jloca = 0, found = 0;
        for (int is = iloc1 + 1; is <= iloca; is++) {
          jloca = is;
          if (ISIDE(2, is) == ip2) { found = 1; break; }
        }
        if (!found) { /* new side (:104-108) */
          iloca += 1;
          if (iloca > mxsid) { status = -1; goto done; }
          ISIDE(1, iloca) = ip1;
          ISIDE(2, iloca) = ip2;
          ISIDE(2 + j, iloca) = ie;
        } else {      /* old side (:112-113) */
          ISIDE(2 + j, jloca) = ie;
        }
      }
    }
    /* sides whose only element is on the right are flipped (:122-129) */
    for (int is = iloc1 + 1; is <= iloca; is++) {
      if (ISIDE(3, is) != 0) continue;
      ISIDE(3, is) = ISIDE(4, is);
      ISIDE(4, is) = 0;
      ISIDE(1, is) = ISIDE(2, is);
      ISIDE(2, is) = ip1;
    }
  }
  /* local node numbers of each side (:143-213) */
  for (int is = 1; is <= iloca; is++) {
    int ip1 = ISIDE(1, is), ip2 = ISIDE(2, is);
    int iel = ISIDE(3, is), ier = ISIDE(4, is);
    int iloc = -2;
    for (int in = 1; in <= 3; in++)
      if (INTMA(in, iel) == ip1) { iloc = in; break; }
    int ip1l, ip2l;
    if (iloc == 3) { ip1l = 3; ip2l = 1; } else { ip1l = iloc; ip2l = iloc + 1; }
    ISIDE(5, is) = ip1l;
    ISIDE(6, is) = ip2l;
    if (ier == 0) {
      ISIDE(7, is) = 0;
      ISIDE(8, is) = 0;
    } else {
      for (int in = 1; in <= 3; in++)
        if (INTMA(in, ier) == ip2) { iloc = in; break; }
      int ip1r, ip2r;
      if (iloc == 3) { ip1r = 1; ip2r = 3; } else { ip2r = iloc; ip1r = iloc + 1; }
      ISIDE(7, is) = ip1r;
      ISIDE(8, is) = ip2r;
    }
  }
  status = iloca;
done:
  free(lhowm); free(lwher); free(icone);
  return status;
}

/* NORMAL: side normals and lengths (src/normal.f90:24-33). */
void ora_normal(int nside, const int *iside, const double *coord,
                double *nx, double *ny, double *el)
{
  for (int is = 1; is <= nside; is++) {
    int ip1 = ISIDE(1, is), ip2 = ISIDE(2, is);
    double dx = COORD(1, ip2) - COORD(1, ip1);
    double dy = COORD(2, ip2) - COORD(2, ip1);
    el[is - 1] = sqrt((dx * dx) + (dy * dy));
    nx[is - 1] = dy / el[is - 1];
    ny[is - 1] = -dx / el[is - 1];
  }
}

/* GETGEO: shape-function gradients and RJ = 2*area (src/getgeo.f90:22-59). */
void ora_getgeo(int nelem, const int *intma, const double *coord, double *geome)
{
  static const double nxi[3] = {-1.0, 1.0, 0.0}, net[3] = {-1.0, 0.0, 1.0};
  for (int ie = 1; ie <= nelem; ie++) {
    double x[3], y[3];
    for (int in = 1; in <= 3; in++) {
      x[in - 1] = COORD(1, INTMA(in, ie));
      y[in - 1] = COORD(2, INTMA(in, ie));
    }
    double x21 = x[1] - x[0], x31 = x[2] - x[0];
    double y21 = y[1] - y[0], y31 = y[2] - y[0];
    double rj = x21 * y31 - x31 * y21;
    double rj1 = 1. / rj;
    double xix = y31 * rj1, xiy = -x31 * rj1;
    double etx = -y21 * rj1, ety = x21 * rj1;
    for (int in = 1; in <= 3; in++) {
      GEOME(in, ie) = xix * nxi[in - 1] + etx * net[in - 1];
      GEOME(in + 3, ie) = xiy * nxi[in - 1] + ety * net[in - 1];
    }
    GEOME(7, ie) = rj;
  }
}

/* GETLEN: minimum element altitude (src/getlen.f90:17-26). */
void ora_getlen(int nelem, const double *geome, double *relen)
{
  for (int ie = 1; ie <= nelem; ie++) {
    double dist[3];
    for (int in = 1; in <= 3; in++) {
      double anx = GEOME(in, ie), any = GEOME(in + 3, ie);
      dist[in - 1] = 1.0 / sqrt(anx * anx + any * any);
    }
    double m = dist[0];
    if (dist[1] < m) m = dist[1];
    if (dist[2] < m) m = dist[2];
    relen[ie - 1] = m;
  }
}

/* CONMAT: the inverse consistent mass matrix as the reference forms it,
 * 0.75/(RJ/24) on the diagonal and -0.25/(RJ/24) off it (src/conmat.f90:1-44);
 * CMMAT(IN,JN,IE) at (IN-1) + 3 (JN-1) + 9 (IE-1). */
void ora_conmat(int nelem, const double *geome, double *cmmat)
{
  const double c24 = 1. / 24.;
  for (int ie = 1; ie <= nelem; ie++) {
    double rj = GEOME(7, ie);
    double rj24 = rj * c24;
    double rj1 = (1 / rj24);
    double rj11 = 0.75 * rj1, rj12 = -0.25 * rj1;
    for (int in = 1; in <= 3; in++)
      for (int jn = 1; jn <= 3; jn++)
        cmmat[(in - 1) + 3 * (jn - 1) + 9 * (size_t)(ie - 1)] = in == jn ? rj11 : rj12;
  }
}

/* GETMAT: inverse lumped mass, 1/(4*RJ/24) (src/getmat.f90:8-40). */
void ora_getmat(int nelem, const double *geome, double *mmat)
{
  const double c24 = 1. / 24., c4 = 4.0;
  for (int ie = 1; ie <= nelem; ie++) {
    double rj24 = GEOME(7, ie) * c24;
    for (int in = 1; in <= 3; in++) {
      double m = 0.0;
      m = m + c4 * rj24;
      mmat[NE(in, ie)] = 1. / m;
    }
  }
}

/* GETNOR: boundary normals and lengths RSIDO(3,NBOUN) (src/getnor.f90:21-31). */
void ora_getnor(int nboun, const int *bsido, const double *coord, double *rsido)
{
  for (int ib = 1; ib <= nboun; ib++) {
    int p0 = BSIDO(1, ib), p1 = BSIDO(2, ib);
    double dx = COORD(1, p1) - COORD(1, p0);
    double dy = COORD(2, p1) - COORD(2, p0);
    double rl = sqrt(dx * dx + dy * dy);
    double rt = 1.0 / rl;
    RSIDO(1, ib) = dy * rt;
    RSIDO(2, ib) = -dx * rt;
    RSIDO(3, ib) = rl;
  }
}

/* VSPACE: ring-major ordinate table VCORD(3,VNPNT) (src/vspace.f90:16-49). */
void ora_vspace(int rorder, int torder, const double *eta, const double *w,
                double *vcord, double *sumweight)
{
  double gap = 2.0 / (double)torder;
  int i = 0;
  double s = 0.0;
  for (int ieta = 1; ieta <= rorder; ieta++)
    for (int izeta = 1; izeta <= torder; izeta++) {
      i += 1;
      VCORD(1, i) = eta[ieta - 1];
      VCORD(2, i) = -1 + izeta * gap;
      VCORD(3, i) = w[ieta - 1] * 1.0;
      s = s + VCORD(3, i);
    }
  *sumweight = s;
}

/* ------------------------------------------------------------------------ */
/* run context                                                               */
/* ------------------------------------------------------------------------ */
struct ora_ctx {
  int E, P, B, S, R, T, V;
  int *intma, *bsido, *iside;
  double *coord, *nx, *ny, *el, *geome, *relen, *mmat, *rsido, *vcord;
  double *cmmat; /* CONMAT's matrices (immat = 0 only) */
  double sumweight, dtec;
  ora_params prm;
  double cinf[4];
  /* per-ordinate tables: radius forms A = (rv/2)*(eta+1), B = eta*(rv/2)+rv/2 */
  double *cxa, *cya, *cxb, *cyb, *jac;
  /* state */
  double *f, *fsnap;
  double *disnd, *disux, *disuy, *disps, *mu, *eta_b;
  double *nd_pp, *rho_pp, *uvel_pp, *vvel_pp, *ps_pp, *temp_pp, *flag;
  double *nd, *rho, *uvel, *vvel, *ps, *temp; /* GETMAC globals (stale) */
  /* work */
  double *nfo, *fluxp, *fluyp, *umean, *rhs, *delun, *residual;
  /* search tables */
  int *bou_of_side;  /* GETBOU's ISB for each boundary side in ISIDE */
  int *flu_in;       /* FLUCON2's (IN1,IN2) for each boundary side */
  unsigned char *inflow_node; /* 1 if DG node is reset by INFLOW */
  int itime;
  double sritm1;
  /* threaded stepping (ora_set_threads): each element's sides in increasing
   * side number, +is for the left element, -is for the right one */
  int nthreads;
  int emul; /* ORA_EMU_* stages emulated in FP32 (precision studies) */
  int *elem_sides;
  double *node_si, *node_su, *node_sv, *node_sp;
};

static void *zal(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

int ora_nside(const ora_ctx *c) { return c->S; }
int ora_nvel(const ora_ctx *c) { return c->V; }
const int *ora_iside(const ora_ctx *c) { return c->iside; }
const double *ora_vcord(const ora_ctx *c) { return c->vcord; }
double ora_sumweight(const ora_ctx *c) { return c->sumweight; }

#define F(c, in, iv, ie) (c)->f[(((size_t)(ie)-1) * (c)->V + ((iv)-1)) * 3 + ((in)-1)]
#define FS(c, in, iv, ie) (c)->fsnap[(((size_t)(ie)-1) * (c)->V + ((iv)-1)) * 3 + ((in)-1)]

/* GTEQNF: local Maxwellian for one ordinate (src/gteqnf.f90:26-40). */
static void gteqnf_n(int nthreads, int nelem, const double *disnd, const double *disux,
                     const double *disuy, const double *disps, double ux, double uy,
                     double m, double *nfo)
{
#pragma omp parallel for schedule(static) if (nthreads > 1) num_threads(nthreads)
  for (int ie = 1; ie <= nelem; ie++)
    for (int in = 1; in <= 3; in++) {
      size_t k = NE(in, ie);
      nfo[k] = 0.0;
      double nd = disnd[k];
      double mm = m / NA_R;
      double rho = mm * nd;
      double ps = disps[k];
      if (ps < 1e-20) continue;
      double cx = disux[k], cy = disuy[k];
      double beta2 = rho / (2 * ps);
      if (g_emu & ORA_EMU_PARAMS) { /* FP32 parameters, FP32 exponential */
        const float pa = (float)(nd * (beta2 / PI_R)), pb = (float)beta2;
        const float dx = (float)cx - (float)ux, dy = (float)cy - (float)uy;
        nfo[k] = (double)(pa * expf(-pb * (dx * dx + dy * dy)));
        continue;
      }
      double tmp = -beta2 * ((cx - ux) * (cx - ux) + (cy - uy) * (cy - uy));
      nfo[k] = nd * (beta2 / PI_R) * exp(tmp);
    }
}

void ora_gteqnf(int nelem, const double *disnd, const double *disux,
                const double *disuy, const double *disps, double ux, double uy,
                double m, double *nfo)
{
  gteqnf_n(1, nelem, disnd, disux, disuy, disps, ux, uy, m, nfo);
}

/* Inflow Maxwellian f_inf(c) (src/inflow.f90:28-39). */
double ora_inflow_value(const double *cinf, double rgas, double m, double ux,
                        double uy)
{
  double tin = cinf[0], pin = cinf[1] * 100000.0, u0 = cinf[2], v0 = cinf[3];
  double rho = pin / (rgas * tin);
  double beta = sqrt(rho / (2 * pin));
  double c0 = (rho * NA_R) / m;
  double c1 = (beta * beta) / PI_R;
  double ud = ux - u0, vd = uy - v0;
  double c2 = (ud * ud + vd * vd);
  return c0 * c1 * exp(-(beta * beta) * c2);
}

/* GETFOR: pressure lift and drag over type-4 sides (src/getfor.f90:27-49). */
void ora_getfor(int nboun, const int *bsido, const double *ps,
                const double *coord, double *lift, double *drag)
{
  double l = 0.0, d = 0.0;
  for (int ib = 1; ib <= nboun; ib++) {
    if (BSIDO(4, ib) != 4) continue;
    int ip1 = BSIDO(1, ib), ip2 = BSIDO(2, ib);
    double x1 = COORD(1, ip1), y1 = COORD(2, ip1);
    double x2 = COORD(1, ip2), y2 = COORD(2, ip2);
    double len = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    double vec1 = x2 - x1, vec2 = y2 - y1;
    /* integer ROTMAT = [[0,-1],[1,0]] applied in real arithmetic */
    double perp1 = 0.0 * vec1 + (-1.0) * vec2;
    double perp2 = 1.0 * vec1 + 0.0 * vec2;
    double mag = sqrt(perp1 * perp1 + perp2 * perp2);
    perp1 = perp1 / mag;
    perp2 = perp2 / mag;
    double psav = 0.5 * (ps[ip1 - 1] + ps[ip2 - 1]);
    l = l - psav * len * perp2;
    d = d - psav * len * perp1;
  }
  *lift = l;
  *drag = d;
}

/* MACROS with the per-node velocity sums split over threads (each node's
 * sums keep the ordinate order) and the point accumulation kept serial in
 * element order, so every value equals the serial loop's bit for bit.  Used
 * by the threaded CPU baseline only. */
static void macros_threaded(ora_ctx *c, double frac, double mm)
{
  const int E = c->E, V = c->V;
  const int *intma = c->intma;
  const double *geome = c->geome, *vcord = c->vcord;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
  for (int ie = 1; ie <= E; ie++)
    for (int in = 1; in <= 3; in++) {
      size_t k = NE(in, ie);
      c->disnd[k] = c->disux[k] = c->disuy[k] = c->disps[k] = 0.0;
      double si = 0.0, su = 0.0, sv = 0.0, sp = 0.0;
      for (int iv = 1; iv <= V; iv++) {
        double w = VCORD(3, iv), jac = c->jac[iv - 1];
        double fv = F(c, in, iv, ie);
        si = si + fv * w * jac * frac;
        su = su + c->cxb[iv - 1] * fv * w * jac * frac;
        sv = sv + c->cyb[iv - 1] * fv * w * jac * frac;
      }
      if (si >= 1e-20) {
        c->disnd[k] = si;
        c->disux[k] = su / si;
        c->disuy[k] = sv / si;
      }
      for (int iv = 1; iv <= V; iv++) {
        double w = VCORD(3, iv), jac = c->jac[iv - 1];
        double cxd = c->cxb[iv - 1] - c->disux[k];
        double cyd = c->cyb[iv - 1] - c->disuy[k];
        double tmp = 0.5 * (cxd * cxd + cyd * cyd);
        sp = sp + tmp * F(c, in, iv, ie) * w * jac * frac;
      }
      c->node_si[k] = si; c->node_su[k] = su; c->node_sv[k] = sv; c->node_sp[k] = sp;
      c->disps[k] = sp * mm;
    }
  for (int ie = 1; ie <= E; ie++) {
    double ae = GEOME(7, ie);
    for (int in = 1; in <= 3; in++) {
      size_t k = NE(in, ie);
      int ip = INTMA(in, ie);
      double si = c->node_si[k];
      if (si >= 1e-20) {
        c->uvel_pp[ip - 1] = c->uvel_pp[ip - 1] + ae * c->node_su[k] / si;
        c->vvel_pp[ip - 1] = c->vvel_pp[ip - 1] + ae * c->node_sv[k] / si;
      }
      c->flag[ip - 1] = c->flag[ip - 1] + ae;
      c->nd_pp[ip - 1] = c->nd_pp[ip - 1] + si * ae;
      c->ps_pp[ip - 1] = c->ps_pp[ip - 1] + mm * c->node_sp[k] * ae;
    }
  }
}

/* MACROS: node moments and area-weighted point averages (src/macros.f90:55-188). */
static void macros(ora_ctx *c)
{
  const int E = c->E, V = c->V, P = c->P;
  const int *intma = c->intma;
  const double *geome = c->geome, *vcord = c->vcord;
  double frac = 4.0 / c->sumweight;
  double mm = c->prm.m / 6.022E26;
  for (int ip = 0; ip < P; ip++) {
    c->rho_pp[ip] = c->uvel_pp[ip] = c->vvel_pp[ip] = 0.0;
    c->temp_pp[ip] = c->ps_pp[ip] = c->nd_pp[ip] = c->flag[ip] = 0.0;
  }
  if (c->nthreads > 1) {
    macros_threaded(c, frac, mm);
    goto points;
  }
  for (int ie = 1; ie <= E; ie++) {
    double ae = GEOME(7, ie);
    for (int in = 1; in <= 3; in++) {
      size_t k = NE(in, ie);
      int ip = INTMA(in, ie);
      c->disnd[k] = c->disux[k] = c->disuy[k] = c->disps[k] = 0.0;
      double si = 0.0, su = 0.0, sv = 0.0, sp = 0.0;
      for (int iv = 1; iv <= V; iv++) {
        double w = VCORD(3, iv), jac = c->jac[iv - 1];
        double fv = F(c, in, iv, ie);
        si = si + fv * w * jac * frac;
        su = su + c->cxb[iv - 1] * fv * w * jac * frac;
        sv = sv + c->cyb[iv - 1] * fv * w * jac * frac;
      }
      if (si >= 1e-20) {
        c->uvel_pp[ip - 1] = c->uvel_pp[ip - 1] + ae * su / si;
        c->vvel_pp[ip - 1] = c->vvel_pp[ip - 1] + ae * sv / si;
        c->disnd[k] = si;
        c->disux[k] = su / si;
        c->disuy[k] = sv / si;
      }
      for (int iv = 1; iv <= V; iv++) {
        double w = VCORD(3, iv), jac = c->jac[iv - 1];
        double cxd = c->cxb[iv - 1] - c->disux[k];
        double cyd = c->cyb[iv - 1] - c->disuy[k];
        double tmp = 0.5 * (cxd * cxd + cyd * cyd);
        sp = sp + tmp * F(c, in, iv, ie) * w * jac * frac;
      }
      /* point accumulation (:156-167) -- same element/node order */
      c->flag[ip - 1] = c->flag[ip - 1] + ae;
      c->nd_pp[ip - 1] = c->nd_pp[ip - 1] + si * ae;
      c->ps_pp[ip - 1] = c->ps_pp[ip - 1] + mm * sp * ae;
      c->disps[k] = sp * mm;
    }
  }
points:
  for (int ip = 0; ip < P; ip++) {
    if (c->nd_pp[ip] < 1e-20) {
      c->rho_pp[ip] = c->uvel_pp[ip] = c->vvel_pp[ip] = 0.0;
      c->ps_pp[ip] = c->temp_pp[ip] = 0.0;
    } else {
      c->nd_pp[ip] = c->nd_pp[ip] / c->flag[ip];
      c->rho_pp[ip] = c->nd_pp[ip] * mm;
      c->uvel_pp[ip] = c->uvel_pp[ip] / c->flag[ip];
      c->vvel_pp[ip] = c->vvel_pp[ip] / c->flag[ip];
      c->ps_pp[ip] = c->ps_pp[ip] / c->flag[ip];
      c->temp_pp[ip] = c->ps_pp[ip] / (c->prm.rgas * c->rho_pp[ip]);
    }
  }
}

/* The np=1 MACROS point arrays are the globals GETMAC would assemble. */
static void getmac(ora_ctx *c)
{
  size_t n = (size_t)c->P * sizeof(double);
  memcpy(c->nd, c->nd_pp, n);
  memcpy(c->rho, c->rho_pp, n);
  memcpy(c->uvel, c->uvel_pp, n);
  memcpy(c->vvel, c->vvel_pp, n);
  memcpy(c->ps, c->ps_pp, n);
  memcpy(c->temp, c->temp_pp, n);
}

/* COMPMU: collision frequency per DG node (src/compmu.f90:36-97). */
static void compmu(ora_ctx *c)
{
  const int E = c->E, V = c->V;
  const double *vcord = c->vcord;
  double frac = 4.0 / c->sumweight;
  double sigt = PI_R * c->prm.d * c->prm.d;
#pragma omp parallel for schedule(static) if (c->nthreads > 1) num_threads(c->nthreads)
  for (int ie = 1; ie <= E; ie++)
    for (int in = 1; in <= 3; in++) {
      size_t k = NE(in, ie);
      double nd = c->disnd[k], ct = 0.0;
      if (nd >= 1e-20) {
        double cx0 = c->disux[k], cy0 = c->disuy[k];
        for (int iv = 1; iv <= V; iv++) {
          double cx = c->cxa[iv - 1], cy = c->cya[iv - 1];
          double w = VCORD(3, iv), jac = c->jac[iv - 1];
          double c1 = (cx0 - cx) * (cx0 - cx) + (cy0 - cy) * (cy0 - cy);
          double cdash = sqrt(c1);
          ct = ct + cdash * (F(c, in, iv, ie) / nd) * w * jac * frac;
        }
      }
      double cr = sqrt(2.0) * ct;
      c->mu[k] = cr * sigt * c->disnd[k];
    }
}

/* FLUCON2: wall mass-flux parameter per type-4 side (src/flucon2.f90:43-241). */
static int flucon2(ora_ctx *c)
{
  const int B = c->B, V = c->V, S = c->S;
  const int *bsido = c->bsido, *iside = c->iside;
  const double *rsido = c->rsido, *vcord = c->vcord;
  double frac = 4.0 / c->sumweight;
  for (int ib = 1; ib <= B; ib++) {
    c->eta_b[ib - 1] = 0.0;
    if (BSIDO(4, ib) != 4) continue;
    int ip1 = BSIDO(1, ib), ip2 = BSIDO(2, ib), iel = BSIDO(3, ib);
    double anx = RSIDO(1, ib), any = RSIDO(2, ib);
    int in1 = -1, in2 = -1;
    if (c->prm.faithful_scans) {
      int is;
      for (is = 1; is <= S; is++) {
        if (ip1 == ISIDE(1, is) && ip2 == ISIDE(2, is)) {
          in1 = ISIDE(5, is); in2 = ISIDE(6, is); break;
        } else if (ip2 == ISIDE(1, is) && ip1 == ISIDE(2, is)) {
          in1 = ISIDE(6, is); in2 = ISIDE(5, is); break;
        }
      }
      if (is == S + 1) return -3;
    } else {
      in1 = c->flu_in[2 * (ib - 1)];
      in2 = c->flu_in[2 * (ib - 1) + 1];
      if (in1 < 0) return -3;
    }
    double top = 0.0;
    for (int iv = 1; iv <= V; iv++) {
      double cx = c->cxb[iv - 1], cy = c->cyb[iv - 1];
      double test = cx * anx + cy * any;
      double vnorm = fabs(test);
      double jac = c->jac[iv - 1];
      double w = VCORD(3, iv);
      double nf = 0.5 * (F(c, in1, iv, iel) + F(c, in2, iv, iel));
      if (test > 0.0) top = top + vnorm * nf * jac * w * frac;
    }
    c->eta_b[ib - 1] = top; /* MPI_ALLREDUCE over one V-rank is a copy */
  }
  for (int ib = 1; ib <= B; ib++) {
    if (BSIDO(4, ib) != 4) { c->eta_b[ib - 1] = 0.0; continue; }
    int tw = BSIDO(5, ib);
    double r = c->prm.rgas;
    double bottom = sqrt(2 * PI_R * (r * r * r) * (double)(tw * tw * tw));
    if (bottom < 0.1) return -4;
    c->eta_b[ib - 1] = c->eta_b[ib - 1] / bottom;
  }
  return 0;
}

/* Reflected-ordinate lookup of GETBOU (src/getbou.f90:196-250), both the
 * reference ring rule and the corrected one.  Returns 0 or a negative status. */
static int reflect_lookup(const ora_ctx *c, double ux, double uy, double anx,
                          double any, int *i1, int *i2, double *zd1, double *zd2)
{
  const int R = c->R, T = c->T, V = c->V;
  const double *vcord = c->vcord;
  double test = ux * anx + uy * any;
  double cxr = ux + fabs(2 * test) * anx;
  double cyr = uy + fabs(2 * test) * any;
  double rt = sqrt(cxr * cxr + cyr * cyr);
  double quo = fabs(cxr / rt);
  double theta = fabs(acos(quo));
  if (cxr < 0.0) theta = PI_R - theta;
  if (cyr < 0.0) theta = -theta;
  double etaref = ((2 * rt) / c->prm.rv) - 1;
  if (etaref > 1.0) etaref = 1.0;
  double zetaref = theta / PI_R;
  int ring; /* 1-based ring whose ordinates the reflection reads */
  if (c->prm.lookup == 0) {
    int ieta;
    for (ieta = 1; ieta <= R; ieta++) {
      int idx = 1 + (ieta - 1) * (R - 1);
      if (idx < 1 || idx > V) return -5;
      if (fabs(etaref - VCORD(1, idx)) < 0.0001) break;
    }
    ring = ieta - 1; /* I1 = (IETA-2)*TORDER + IZETA */
  } else {
    int j;
    for (j = 1; j <= R; j++)
      if (fabs(etaref - VCORD(1, (j - 1) * T + 1)) < 0.0001) break;
    if (j > R) return -6;
    ring = j;
  }
  int izeta;
  double d1 = 0.0, d2 = 0.0;
  for (izeta = 1; izeta <= T - 1; izeta++) {
    double zt1 = VCORD(2, izeta), zt2 = VCORD(2, izeta + 1);
    d1 = zt2 - zt1;
    d2 = zetaref - zt1;
    if (zetaref >= zt1 && zetaref <= zt2) break;
  }
  if (izeta == T) { /* wrap between the last and first angle (:236-240) */
    d1 = VCORD(2, 1) - (-1);
    d2 = zetaref - (-1);
    *i1 = ring * T;
    *i2 = (ring - 1) * T + 1;
  } else {
    *i1 = (ring - 1) * T + izeta;
    *i2 = (ring - 1) * T + izeta + 1;
  }
  if (*i1 < 1 || *i1 > V || *i2 < 1 || *i2 > V) return -5;
  *zd1 = d1;
  *zd2 = d2;
  return 0;
}

/* GETFLU on a 2-vector (src/getflu.f90:7-18). */
static inline void getflu2(const double u[2], double ux, double uy, double fx[2],
                           double fy[2])
{
  for (int i = 0; i < 2; i++) {
    if (fabs(u[i]) == 0.0) { fx[i] = 0.0; fy[i] = 0.0; }
    else { fx[i] = AE(ux * u[i]); fy[i] = AE(uy * u[i]); }
  }
}

/* 1-D P1 edge mass matrix times normal flux (src/edgflxopt.f90:114-120). */
static inline void edge_rhsi(const double fn[2], double alen, double rhsi[2])
{
  static const double mm[2][2] = {{2.0, 1.0}, {1.0, 2.0}};
  rhsi[0] = 0.0;
  rhsi[1] = 0.0;
  for (int in = 0; in < 2; in++)
    for (int jn = 0; jn < 2; jn++) {
      double cm = AE(AE(fn[jn] * mm[in][jn]) * alen);
      rhsi[in] = AE(rhsi[in] - cm);
    }
}

/* GETBOU: one boundary side for ordinate iv (src/getbou.f90:48-283). */
static int getbou(ora_ctx *c, int is, int iv, double ux, double uy)
{
  const int *iside = c->iside, *bsido = c->bsido;
  const double *rsido = c->rsido;
  int ip1 = ISIDE(1, is), iel = ISIDE(3, is);
  int in1 = ISIDE(5, is), in2 = ISIDE(6, is);
  int isb;
  if (c->prm.faithful_scans) {
    for (isb = 1; isb <= c->B; isb++)
      if (iel == BSIDO(3, isb) && ip1 == BSIDO(1, isb)) break;
    if (isb > c->B) return -2;
  } else {
    isb = c->bou_of_side[is - 1];
    if (isb < 1) return -2;
  }
  double alen = AE((1. / 6.) * RSIDO(3, isb));
  double anx = AE(RSIDO(1, isb)), any = AE(RSIDO(2, isb));
  int index = BSIDO(4, isb);
  double temp = (double)BSIDO(5, isb);
  double *rhs = c->rhs;
  double unk1 = F(c, in1, iv, iel), unk2 = F(c, in2, iv, iel);
  double unsid[2] = {AE(unk1 - c->umean[iel - 1]), AE(unk2 - c->umean[iel - 1])};
  double fx[2], fy[2], fn[2], rhsi[2];
  getflu2(unsid, ux, uy, fx, fy);
  fn[0] = AE(AE(anx * fx[0]) + AE(any * fy[0]));
  fn[1] = AE(AE(anx * fx[1]) + AE(any * fy[1]));
  /* the branch and the reflection lookup on the FP64 values, also when the
   * arithmetic is emulated in FP32: the reference's lookup is discontinuous
   * at its bracket ends, and the library builds it on the host in FP64 */
  const double uxd = c->cxa[iv - 1], uyd = c->cya[iv - 1];
  const double anxd = RSIDO(1, isb), anyd = RSIDO(2, isb);
  double test = uxd * anxd + uyd * anyd;
  if (index == 1) return 0; /* inflow side: no flux (:88-138) */
  if (index == 2 || test >= 0.0) {
    edge_rhsi(fn, alen, rhsi);
    rhs[NE(in1, iel)] = AE(rhs[NE(in1, iel)] + rhsi[0]);
    rhs[NE(in2, iel)] = AE(rhs[NE(in2, iel)] + rhsi[1]);
    return 0;
  }
  /* wall, molecules leaving the wall (:189-280) */
  double speed = sqrt(uxd * uxd + uyd * uyd);
  double mw = exp(-(speed * speed) / (2 * c->prm.rgas * temp));
  double mf = mw * c->eta_b[isb - 1];
  int i1, i2;
  double zd1, zd2;
  int st = reflect_lookup(c, uxd, uyd, anxd, anyd, &i1, &i2, &zd1, &zd2);
  if (st) return st;
  double rf11, rf12, rf21, rf22;
  if (c->prm.jacobi) {
    rf11 = FS(c, in1, i1, iel); rf12 = FS(c, in1, i2, iel);
    rf21 = FS(c, in2, i1, iel); rf22 = FS(c, in2, i2, iel);
  } else {
    rf11 = F(c, in1, i1, iel); rf12 = F(c, in1, i2, iel);
    rf21 = F(c, in2, i1, iel); rf22 = F(c, in2, i2, iel);
  }
  mf = AE(mf);
  double rf1 = AE(AE(AE(zd2 * rf11) + AE(zd1 * rf12)) / AE(zd1 + zd2));
  double rf2 = AE(AE(AE(zd2 * rf21) + AE(zd1 * rf22)) / AE(zd1 + zd2));
  double a = c->prm.alpha;
  unsid[0] = AE(AE((1 - a) * rf1) + AE(a * mf));
  unsid[1] = AE(AE((1 - a) * rf2) + AE(a * mf));
  getflu2(unsid, ux, uy, fx, fy);
  fn[0] = AE(AE(anx * fx[0]) + AE(any * fy[0]));
  fn[1] = AE(AE(anx * fx[1]) + AE(any * fy[1]));
  edge_rhsi(fn, alen, rhsi);
  rhs[NE(in1, iel)] = AE(rhs[NE(in1, iel)] + rhsi[0]);
  rhs[NE(in2, iel)] = AE(rhs[NE(in2, iel)] + rhsi[1]);
  return 0;
}

/* Optional stage snapshots of one ordinate (test hook). */
typedef struct ora_stage {
  double *nfo, *umean, *rhs_elc, *rhs_edge, *delun;
} ora_stage;
static ora_stage *g_stage = NULL;

/* EDGFLXOPT's side loop, element by element for the threaded baseline: each
 * element visits its own sides in increasing side number, computes each
 * side's flux with the serial loop's arithmetic and adds it to its own two
 * nodes, so every node receives the same terms in the same order as in the
 * serial loop (an interior side's flux is computed by both its elements).
 * Boundary sides call GETBOU as the serial loop does; it writes only that
 * element's nodes. */
static int edges_threaded(ora_ctx *c, int iv, double ux, double uy)
{
  const int E = c->E;
  const int *iside = c->iside;
  int status = 0;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
  for (int ie = 1; ie <= E; ie++) {
    for (int j = 0; j < 3; j++) {
      int code = c->elem_sides[3 * (size_t)(ie - 1) + j];
      if (code == 0) continue;
      int is = code > 0 ? code : -code;
      int iel = ISIDE(3, is), ier = ISIDE(4, is);
      int inl1 = ISIDE(5, is), inl2 = ISIDE(6, is);
      int inr1 = ISIDE(7, is), inr2 = ISIDE(8, is);
      if (ier == 0) {
        int st = getbou(c, is, iv, ux, uy);
        if (st) {
#pragma omp critical
          status = st;
        }
        continue;
      }
      double scpr = ux * c->nx[is - 1] + uy * c->ny[is - 1];
      double edgun[2];
      if (scpr > 0.0) {
        edgun[0] = F(c, inl1, iv, iel) - c->umean[iel - 1];
        edgun[1] = F(c, inl2, iv, iel) - c->umean[iel - 1];
      } else {
        edgun[0] = F(c, inr1, iv, ier) - c->umean[ier - 1];
        edgun[1] = F(c, inr2, iv, ier) - c->umean[ier - 1];
      }
      double fx[2], fy[2], fn[2], rhsi[2];
      getflu2(edgun, ux, uy, fx, fy);
      double alen = (1.0 / 6.0) * c->el[is - 1];
      fn[0] = c->nx[is - 1] * fx[0] + c->ny[is - 1] * fy[0];
      fn[1] = c->nx[is - 1] * fx[1] + c->ny[is - 1] * fy[1];
      edge_rhsi(fn, alen, rhsi);
      if (code < 0) {
        c->rhs[NE(inr1, ier)] = c->rhs[NE(inr1, ier)] - rhsi[0];
        c->rhs[NE(inr2, ier)] = c->rhs[NE(inr2, ier)] - rhsi[1];
      } else {
        c->rhs[NE(inl1, iel)] = c->rhs[NE(inl1, iel)] + rhsi[0];
        c->rhs[NE(inl2, iel)] = c->rhs[NE(inl2, iel)] + rhsi[1];
      }
    }
  }
  return status;
}

/* Threads for ora_step (1 = the serial restatement).  The threaded step
 * computes the same values bit for bit; it is the all-core CPU baseline.
 * Returns 0, or -100 when a table cannot be allocated. */
int ora_set_threads(ora_ctx *c, int nthreads)
{
  c->nthreads = nthreads > 1 ? nthreads : 1;
  if (c->nthreads == 1 || c->elem_sides) return 0;
  const size_t E = c->E;
  const int *iside = c->iside;
  c->elem_sides = calloc(3 * E, sizeof(int));
  c->node_si = calloc(3 * E, sizeof(double));
  c->node_su = calloc(3 * E, sizeof(double));
  c->node_sv = calloc(3 * E, sizeof(double));
  c->node_sp = calloc(3 * E, sizeof(double));
  if (!c->elem_sides || !c->node_si || !c->node_su || !c->node_sv || !c->node_sp) {
    c->nthreads = 1;
    return -100;
  }
  int *cnt = calloc(E + 1, sizeof(int));
  if (!cnt) { c->nthreads = 1; return -100; }
  for (int is = 1; is <= c->S; is++) { /* increasing side number */
    int iel = ISIDE(3, is), ier = ISIDE(4, is);
    if (cnt[iel] < 3) c->elem_sides[3 * (size_t)(iel - 1) + cnt[iel]++] = is;
    if (ier > 0 && cnt[ier] < 3) c->elem_sides[3 * (size_t)(ier - 1) + cnt[ier]++] = -is;
  }
  free(cnt);
  return 0;
}

/* An FP32 store of f + d with a dither uniform over +-1/2 ulp of f added to
 * d before the one rounding (stochastic rounding at f's exponent): the
 * dither fraction is a hash of (element, ordinate, step), shared by the
 * element's three nodes.  The caller clamps and rounds the result. */
static double emu_store_sr(double f, double d, int ie, int iv, int itime)
{
  if (f == 0.0) return f + d;
  uint32_t h = (uint32_t)ie * 2654435761u ^ (uint32_t)iv * 2246822519u ^
               (uint32_t)itime * 3266489917u;
  h ^= h >> 15; h *= 2246822519u; h ^= h >> 13; h *= 3266489917u; h ^= h >> 16;
  const double r = (double)h / 4294967296.0;
  int ex;
  frexp((double)(float)f, &ex);                 /* f = m 2^ex, 0.5 <= m < 1 */
  const double ulp = ldexp(1.0, ex - 24);        /* FP32 ulp of f */
  return f + (d + (r - 0.5) * ulp);
}

/* The upwind values GETBOU takes on a boundary side (precision studies of
 * the mean form): 0 on an inflow side (no flux), the element's own values
 * on an outflow side or where c.n >= 0, else the wall's reflected
 * distribution; the same branches and arithmetic as getbou(). */
static int getbou_values(ora_ctx *c, int is, int iv, double ux, double uy, double g[2])
{
  const int *iside = c->iside, *bsido = c->bsido;
  const double *rsido = c->rsido;
  int iel = ISIDE(3, is), in1 = ISIDE(5, is), in2 = ISIDE(6, is);
  int isb = c->bou_of_side[is - 1];
  if (isb < 1) return -2;
  int index = BSIDO(4, isb);
  double temp = (double)BSIDO(5, isb);
  const double uxd = c->cxa[iv - 1], uyd = c->cya[iv - 1];
  const double anxd = RSIDO(1, isb), anyd = RSIDO(2, isb);
  double test = uxd * anxd + uyd * anyd;
  (void)ux; (void)uy;
  if (index == 1) { g[0] = g[1] = 0.0; return 0; }
  if (index == 2 || test >= 0.0) {
    g[0] = AE(F(c, in1, iv, iel) - c->umean[iel - 1]);
    g[1] = AE(F(c, in2, iv, iel) - c->umean[iel - 1]);
    return 0;
  }
  double speed = sqrt(uxd * uxd + uyd * uyd);
  double mw = exp(-(speed * speed) / (2 * c->prm.rgas * temp));
  double mf = AE(mw * c->eta_b[isb - 1]);
  int i1, i2;
  double zd1, zd2;
  int st = reflect_lookup(c, uxd, uyd, anxd, anyd, &i1, &i2, &zd1, &zd2);
  if (st) return st;
  double rf11 = F(c, in1, i1, iel), rf12 = F(c, in1, i2, iel);
  double rf21 = F(c, in2, i1, iel), rf22 = F(c, in2, i2, iel);
  double rf1 = AE(AE(AE(zd2 * rf11) + AE(zd1 * rf12)) / AE(zd1 + zd2));
  double rf2 = AE(AE(AE(zd2 * rf21) + AE(zd1 * rf22)) / AE(zd1 + zd2));
  double a = c->prm.alpha;
  g[0] = AE(AE((1 - a) * rf1) + AE(a * mf));
  g[1] = AE(AE((1 - a) * rf2) + AE(a * mf));
  return 0;
}

/* The volume term folded into the edge terms (precision studies): since
 * area grad N_k is half the sum of len n over the element's edges at node
 * k, each edge of an element acts on (upwind values - the element's mean
 * of f - U), an inflow side on (0 - mean).  Equal to the reference's
 * operator in exact arithmetic; its terms are differences, so their
 * rounding is relative to the variation of f instead of f. */
static int edges_meanform(ora_ctx *c, int iv, double ux, double uy)
{
  const int E = c->E;
  const int *iside = c->iside;
  const double c13 = AE(1. / 3.);
  if (!c->elem_sides && ora_set_threads(c, 2)) return -100;
  c->nthreads = 1;
  for (int ie = 1; ie <= E; ie++) {
    /* differences from the element's first node: r = f_1, dl_k = f_k - r,
     * the mean of g = f - U is gb = r + db - U with db the mean of dl */
    const double r = F(c, 1, iv, ie), uo = c->umean[ie - 1];
    const double dl[3] = {0.0, AE(F(c, 2, iv, ie) - r), AE(F(c, 3, iv, ie) - r)};
    const double db = AE(c13 * AE(dl[1] + dl[2]));
    const double gb = AE(AE(r + db) - uo);
    for (int j = 0; j < 3; j++) {
      int code = c->elem_sides[3 * (size_t)(ie - 1) + j];
      if (code == 0) continue;
      int is = code > 0 ? code : -code;
      int iel = ISIDE(3, is), ier = ISIDE(4, is);
      int inl1 = ISIDE(5, is), inl2 = ISIDE(6, is);
      int inr1 = ISIDE(7, is), inr2 = ISIDE(8, is);
      /* h = upwind g - gb, formed from differences of nearby values */
      const int mine1 = code > 0 ? inl1 : inr1, mine2 = code > 0 ? inl2 : inr2;
      double h[2];
      int own = 0;
      if (ier == 0) {
        double g[2];
        int st = getbou_values(c, is, iv, ux, uy, g);
        if (st) return st;
        int isb = c->bou_of_side[is - 1];
        const double tst = c->cxa[iv - 1] * c->rsido[3 * (isb - 1)] +
                           c->cya[iv - 1] * c->rsido[3 * (isb - 1) + 1];
        const int index = c->bsido[6 * (isb - 1) + 3];
        if (index != 1 && (index == 2 || tst >= 0.0)) own = 1;
        else { h[0] = AE(g[0] - gb); h[1] = AE(g[1] - gb); }
      } else {
        const int left_up = ux * c->nx[is - 1] + uy * c->ny[is - 1] > 0.0;
        if (left_up == (code > 0)) own = 1;
        else {
          const int nb = code > 0 ? ier : iel;
          const int n1 = code > 0 ? inr1 : inl1, n2 = code > 0 ? inr2 : inl2;
          const double du = AE(c->umean[nb - 1] - uo);
          h[0] = AE(AE(AE(F(c, n1, iv, nb) - r) - db) - du);
          h[1] = AE(AE(AE(F(c, n2, iv, nb) - r) - db) - du);
        }
      }
      if (own) { h[0] = AE(dl[mine1 - 1] - db); h[1] = AE(dl[mine2 - 1] - db); }
      double fx[2], fy[2], fn[2], rhsi[2];
      getflu2(h, ux, uy, fx, fy);
      const double snx = ier == 0 ? AE(c->rsido[3 * (c->bou_of_side[is - 1] - 1)]) : AE(c->nx[is - 1]);
      const double sny = ier == 0 ? AE(c->rsido[3 * (c->bou_of_side[is - 1] - 1) + 1]) : AE(c->ny[is - 1]);
      const double len = ier == 0 ? c->rsido[3 * (c->bou_of_side[is - 1] - 1) + 2] : c->el[is - 1];
      double alen = AE((1.0 / 6.0) * len);
      fn[0] = AE(AE(snx * fx[0]) + AE(sny * fy[0]));
      fn[1] = AE(AE(snx * fx[1]) + AE(sny * fy[1]));
      edge_rhsi(fn, alen, rhsi);
      if (code > 0) {
        c->rhs[NE(inl1, iel)] = AE(c->rhs[NE(inl1, iel)] + rhsi[0]);
        c->rhs[NE(inl2, iel)] = AE(c->rhs[NE(inl2, iel)] + rhsi[1]);
      } else {
        c->rhs[NE(inr1, ier)] = AE(c->rhs[NE(inr1, ier)] - rhsi[0]);
        c->rhs[NE(inr2, ier)] = AE(c->rhs[NE(inr2, ier)] - rhsi[1]);
      }
    }
  }
  return 0;
}

/* FP32 emulation of stages of the step (precision studies; serial steps
 * only: the threaded step is switched off while it is set). */
void ora_set_emulation(ora_ctx *c, int flags)
{
  c->emul = flags;
  if (flags & (ORA_EMU_STORE_RN | ORA_EMU_STORE_SR))
    for (size_t k = 0; k < 3 * (size_t)c->V * c->E; k++) c->f[k] = (double)(float)c->f[k];
}

/* One ordinate of the velocity loop (src/advnce.f90:286-377). */
static int ordinate(ora_ctx *c, int iv, double dte)
{
  const int E = c->E, S = c->S;
  const int *iside = c->iside, *intma = c->intma, *bsido = c->bsido;
  const double *geome = c->geome;
  double ux = emu_round(ORA_EMU_ARITH | ORA_EMU_PRED | ORA_EMU_EDGE | ORA_EMU_INC, c->cxa[iv - 1]),
         uy = emu_round(ORA_EMU_ARITH | ORA_EMU_PRED | ORA_EMU_EDGE | ORA_EMU_INC, c->cya[iv - 1]);
  const double c13 = AP(1. / 3.);
  const int nt = c->nthreads > 1 && !g_emu ? c->nthreads : 1;
  /* GTEQNF */
  gteqnf_n(nt, E, c->disnd, c->disux, c->disuy, c->disps, ux, uy, c->prm.m, c->nfo);
  /* GETFLA (src/getfla.f90:16-30) */
#pragma omp parallel for schedule(static) if (nt > 1) num_threads(nt)
  for (int ie = 1; ie <= E; ie++)
    for (int in = 1; in <= 3; in++) {
      double fv = F(c, in, iv, ie);
      size_t k = NE(in, ie);
      if (fabs(fv) == 0.0) { c->fluxp[k] = 0.0; c->fluyp[k] = 0.0; }
      else { c->fluxp[k] = AP(ux * fv); c->fluyp[k] = AP(uy * fv); }
    }
  /* GETELC (src/getelc.f90:40-121) */
#pragma omp parallel for schedule(static) if (nt > 1) num_threads(nt)
  for (int ie = 1; ie <= E; ie++) {
    double area = AP(0.5 * GEOME(7, ie));
    double unkel[3];
    double dive = 0.0, coll = 0.0;
    for (int in = 1; in <= 3; in++) unkel[in - 1] = F(c, in, iv, ie);
    if (g_emu & ORA_EMU_FORM) { /* difference form: the gradients sum to zero */
      const double d2 = AP(F(c, 2, iv, ie) - F(c, 1, iv, ie));
      const double d3 = AP(F(c, 3, iv, ie) - F(c, 1, iv, ie));
      const double sx = AP(AP(AP(GEOME(2, ie)) * d2) + AP(AP(GEOME(3, ie)) * d3));
      const double sy = AP(AP(AP(GEOME(5, ie)) * d2) + AP(AP(GEOME(6, ie)) * d3));
      dive = AP(AP(ux * sx) + AP(uy * sy));
    } else
    for (int in = 1; in <= 3; in++) {
      size_t k = NE(in, ie);
      dive = AP(AP(dive + AP(AP(GEOME(in, ie)) * c->fluxp[k])) + AP(AP(GEOME(in + 3, ie)) * c->fluyp[k]));
    }
    for (int in = 1; in <= 3; in++) {
      size_t k = NE(in, ie);
      coll = AP(coll + AP(P(c->mu[k]) * AP(c->nfo[k] - F(c, in, iv, ie))));
    }
    coll = AP(c13 * coll);
    dive = AP(dive - coll);
    dive = AP(AP(dive * dte) * 0.5);
    c->umean[ie - 1] = dive;
    for (int in = 0; in < 3; in++) unkel[in] = AP(unkel[in] - dive);
    double fxme = 0.0, fyme = 0.0;
    for (int in = 0; in < 3; in++) {
      double fx = 0.0, fy = 0.0;
      if (fabs(unkel[in]) != 0.0) { fx = AP(ux * unkel[in]); fy = AP(uy * unkel[in]); }
      fxme = AP(fxme + fx);
      fyme = AP(fyme + fy);
    }
    fxme = AP(c13 * fxme);
    fyme = AP(c13 * fyme);
    for (int in = 1; in <= 3; in++) {
      double anx = AP(AP(GEOME(in, ie)) * area), any = AP(AP(GEOME(in + 3, ie)) * area);
      size_t k = NE(in, ie);
      c->rhs[k] = 0.0;
      if (g_emu & ORA_EMU_FORM) continue; /* folded into the edge terms */
      c->rhs[k] = AP(AP(c->rhs[k] + AP(anx * fxme)) + AP(any * fyme));
    }
  }
  if (g_stage) {
    memcpy(g_stage->nfo, c->nfo, 3 * (size_t)E * sizeof(double));
    memcpy(g_stage->umean, c->umean, (size_t)E * sizeof(double));
    memcpy(g_stage->rhs_elc, c->rhs, 3 * (size_t)E * sizeof(double));
  }
  /* RFILLA(DELUN) */
  for (size_t k = 0; k < 3 * (size_t)E; k++) c->delun[k] = 0.0;
  if (nt > 1) {
    int st = edges_threaded(c, iv, ux, uy);
    if (st) return st;
    goto edges_done;
  }
  if (g_emu & ORA_EMU_FORM) {
    int st = edges_meanform(c, iv, ux, uy);
    if (st) return st;
    goto edges_done;
  }
  /* EDGFLXOPT (src/edgflxopt.f90:72-125) */
  for (int is = 1; is <= S; is++) {
    int iel = ISIDE(3, is), ier = ISIDE(4, is);
    int inl1 = ISIDE(5, is), inl2 = ISIDE(6, is);
    int inr1 = ISIDE(7, is), inr2 = ISIDE(8, is);
    if (ier == 0) {
      int st = getbou(c, is, iv, ux, uy);
      if (st) return st;
      continue;
    }
    double scpr = ux * c->nx[is - 1] + uy * c->ny[is - 1];
    double edgun[2];
    if (c->prm.mutant == ORA_MUTANT_SWAP_NODES) { int t = inl1; inl1 = inl2; inl2 = t; }
    int upwind_left = scpr > 0.0;
    if (c->prm.mutant == ORA_MUTANT_DOWNWIND) upwind_left = !upwind_left;
    if (upwind_left) {
      edgun[0] = AE(F(c, inl1, iv, iel) - c->umean[iel - 1]);
      edgun[1] = AE(F(c, inl2, iv, iel) - c->umean[iel - 1]);
    } else {
      edgun[0] = AE(F(c, inr1, iv, ier) - c->umean[ier - 1]);
      edgun[1] = AE(F(c, inr2, iv, ier) - c->umean[ier - 1]);
    }
    double fx[2], fy[2], fn[2], rhsi[2];
    getflu2(edgun, ux, uy, fx, fy);
    double alen = AE((1.0 / 6.0) * c->el[is - 1]);
    const double snx = AE(c->nx[is - 1]), sny = AE(c->ny[is - 1]);
    fn[0] = AE(AE(snx * fx[0]) + AE(sny * fy[0]));
    fn[1] = AE(AE(snx * fx[1]) + AE(sny * fy[1]));
    edge_rhsi(fn, alen, rhsi);
    c->rhs[NE(inr1, ier)] = AE(c->rhs[NE(inr1, ier)] - rhsi[0]);
    c->rhs[NE(inr2, ier)] = AE(c->rhs[NE(inr2, ier)] - rhsi[1]);
    c->rhs[NE(inl1, iel)] = AE(c->rhs[NE(inl1, iel)] + rhsi[0]);
    c->rhs[NE(inl2, iel)] = AE(c->rhs[NE(inl2, iel)] + rhsi[1]);
  }
edges_done:
  if (g_stage) memcpy(g_stage->rhs_edge, c->rhs, 3 * (size_t)E * sizeof(double));
  /* RHS *= DTE (src/advnce.f90:342-346) */
#pragma omp parallel for schedule(static) if (nt > 1) num_threads(nt)
  for (size_t k = 0; k < 3 * (size_t)E; k++) c->rhs[k] = AI(dte * c->rhs[k]);
  double res = 0.0;
  if (c->prm.immat == 0) {
    /* CMMINC (src/cmminc.f90:20-41), as written: each node's increment adds
     * CMMAT(IN,JN)*RHS(IN) over JN -- RHS(IN), not RHS(JN), so the sum is the
     * matrix row sum (6/RJ, the lumped value) times RHS(IN) -- after the
     * collision term; the residual is the largest signed CMMAT(IN,JN)*RHS(IN)
     * + COLL*DTE, from 0 (no ABS) */
#pragma omp parallel for schedule(static) reduction(max : res) if (nt > 1) num_threads(nt)
    for (int ie = 1; ie <= E; ie++)
      for (int in = 1; in <= 3; in++) {
        size_t k = NE(in, ie);
        double nf = F(c, in, iv, ie);
        double coll = AI(P(c->mu[k]) * AI(c->nfo[k] - nf));
        double d = AI(c->delun[k] + AI(coll * dte));
        for (int jn = 1; jn <= 3; jn++) {
          double cm = c->cmmat[(in - 1) + 3 * (jn - 1) + 9 * (size_t)(ie - 1)];
          double rh = c->prm.mutant == ORA_MUTANT_CMMINC_RHS_JN ? c->rhs[NE(jn, ie)] : c->rhs[k];
          double r1 = AI(AI(AI(cm) * rh) + AI(coll * dte));
          if (r1 > res) res = r1;
          d = AI(d + AI(AI(cm) * rh));
        }
        c->delun[k] = d;
      }
    c->residual[iv - 1] = res;
    goto increments_done;
  }
  /* GETINC (src/getinc.f90:22-43); the largest |increment| (a max, exact in
   * any order) */
#pragma omp parallel for schedule(static) reduction(max : res) if (nt > 1) num_threads(nt)
  for (int ie = 1; ie <= E; ie++)
    for (int in = 1; in <= 3; in++) {
      size_t k = NE(in, ie);
      double nf = F(c, in, iv, ie);
      double coll = AI(P(c->mu[k]) * AI(c->nfo[k] - nf));
      double r1 = AI(AI(AI(c->mmat[k]) * c->rhs[k]) + AI(coll * dte));
      if (c->prm.mutant == ORA_MUTANT_RES_AFTER && c->inflow_node[k]) r1 = 0.0;
      if (fabs(r1) > res) res = fabs(r1);
      c->delun[k] = AI(AI(c->delun[k] + AI(AI(c->mmat[k]) * c->rhs[k])) + AI(coll * dte));
    }
  c->residual[iv - 1] = res;
increments_done:
  if (g_stage) memcpy(g_stage->delun, c->delun, 3 * (size_t)E * sizeof(double));
  /* ADTHEM (src/adthem.f90:17-26) */
#pragma omp parallel for schedule(static) if (nt > 1) num_threads(nt)
  for (int ie = 1; ie <= E; ie++)
    for (int in = 1; in <= 3; in++) {
      double v;
      if (g_emu & ORA_EMU_STORE_SR)
        v = emu_store_sr(F(c, in, iv, ie), c->delun[NE(in, ie)], ie, iv, c->itime);
      else
        v = F(c, in, iv, ie) + c->delun[NE(in, ie)];
      if (v < 0.0) v = 0.0;
      if (g_emu & (ORA_EMU_STORE_RN | ORA_EMU_STORE_SR)) v = (double)(float)v;
      F(c, in, iv, ie) = v;
    }
  /* INFLOW (src/inflow.f90:28-59) */
  int have_inflow = 0;
  for (int ib = 1; ib <= c->B; ib++) if (BSIDO(4, ib) == 1) { have_inflow = 1; break; }
  if (have_inflow) {
    double nf = ora_inflow_value(c->cinf, c->prm.rgas, c->prm.m, ux, uy);
    if (g_emu & (ORA_EMU_STORE_RN | ORA_EMU_STORE_SR)) nf = (double)(float)nf;
    if (c->prm.faithful_scans) {
      for (int ib = 1; ib <= c->B; ib++) {
        if (BSIDO(4, ib) != 1) continue;
        int ip1 = BSIDO(1, ib), ip2 = BSIDO(2, ib);
        for (int ie = 1; ie <= E; ie++)
          for (int in = 1; in <= 3; in++) {
            int ipt = INTMA(in, ie);
            if (ipt == ip1 || ipt == ip2) F(c, in, iv, ie) = nf;
          }
      }
    } else {
#pragma omp parallel for schedule(static) if (nt > 1) num_threads(nt)
      for (int ie = 1; ie <= E; ie++)
        for (int in = 1; in <= 3; in++)
          if (c->inflow_node[NE(in, ie)]) F(c, in, iv, ie) = nf;
    }
  }
  return 0;
}

static double total_mass(const ora_ctx *c)
{
  const double *geome = c->geome;
  double s = 0.0;
  for (int ie = 1; ie <= c->E; ie++)
    for (int in = 1; in <= 3; in++)
      s += c->disnd[NE(in, ie)] * GEOME(7, ie) / 6.0;
  return s;
}

int ora_step(ora_ctx *c, int nsteps, ora_step_info *info)
{
  for (int s = 0; s < nsteps; s++) {
    c->itime += 1;
    int st = flucon2(c);
    if (st) return st;
    compmu(c);
    /* ALOTIM2 (src/alotim2.f90:10-15) and the min with the Courant step */
    double maxmu = c->mu[0];
    for (size_t k = 1; k < 3 * (size_t)c->E; k++) if (c->mu[k] > maxmu) maxmu = c->mu[k];
    double dteb = (maxmu > 0.0) ? 0.9 * (1 / maxmu) : 1000000.0;
    double dte = (dteb < c->dtec) ? dteb : c->dtec;
    if (c->prm.jacobi)
      memcpy(c->fsnap, c->f, 3 * (size_t)c->V * c->E * sizeof(double));
    g_emu = c->emul;
    if (g_emu & ORA_EMU_DTE) dte = (double)(float)dte;
    for (int iv = 1; iv <= c->V; iv++) {
      st = ordinate(c, iv, dte);
      if (st) { g_emu = 0; return st; }
    }
    g_emu = 0;
    double sumres = 0.0;
    for (int iv = 0; iv < c->V; iv++)
      if (fabs(c->residual[iv]) > sumres) sumres = fabs(c->residual[iv]);
    macros(c);
    if (c->itime % 1000 == 0 || c->prm.mutant == ORA_MUTANT_GETMAC_EVERY) getmac(c);
    if (c->itime == 1) c->sritm1 = sumres;
    if (info) {
      info[s].sumres = sumres;
      info[s].logres = log(sumres / c->sritm1);
      info[s].dte = dte;
      info[s].mass = total_mass(c);
      ora_getfor(c->B, c->bsido, c->ps, c->coord, &info[s].lift, &info[s].drag);
    }
  }
  return 0;
}

/* INICON2 Maxwellian branch (src/inicon2.f90:120-173) and the CINF copy
 * (:185-193). */
static void inicon2(ora_ctx *c)
{
  const double *vcord = c->vcord;
  double t1 = c->prm.t1, p1 = c->prm.p1, u0 = c->prm.u0, v0 = c->prm.v0;
  double gc = c->prm.rgas, m = c->prm.m;
  double rho1 = (p1 * 100000.0) / (gc * t1);
  double n1 = rho1 * NA_R / m;
  double tmp1 = rho1 / (2 * p1 * 100000.0);
  double beta1 = sqrt(tmp1);
  memset(c->f, 0, 3 * (size_t)c->V * c->E * sizeof(double));
  if (n1 >= 1e-20) {
    double co = (beta1 * beta1) / (PI_R);
    for (int ie = 1; ie <= c->E; ie++)
      for (int in = 1; in <= 3; in++)
        for (int inv = 1; inv <= c->V; inv++) {
          double ux = c->cxb[inv - 1], uy = c->cyb[inv - 1];
          double tmp = (ux - u0) * (ux - u0) + (uy - v0) * (uy - v0);
          double speed = sqrt(tmp);
          tmp = -((beta1 * beta1) * (speed * speed));
          double f0 = co * exp(tmp);
          F(c, in, inv, ie) = n1 * f0;
        }
  }
  (void)vcord;
  for (int k = 0; k < 4; k++) c->cinf[k] = (c->prm.inf == 1) ? c->prm.cinf[k] : 0.0;
}

int ora_create(int nelem, int npoin, int nboun, const int *intma_in,
               const double *coord_in, const int *bsido_in, int rorder,
               int torder, const double *eta, const double *w,
               const ora_params *prm, ora_ctx **out)
{
  *out = NULL;
  ora_ctx *c = zal(1, sizeof(ora_ctx));
  if (!c) return -100;
  c->E = nelem; c->P = npoin; c->B = nboun; c->R = rorder; c->T = torder;
  c->V = rorder * torder;
  c->prm = *prm;
  const size_t E = nelem, P = npoin, B = nboun, V = c->V;
  c->intma = zal(3 * E, sizeof(int));
  c->bsido = zal(6 * B, sizeof(int));
  c->coord = zal(2 * P, sizeof(double));
  memcpy(c->intma, intma_in, 3 * E * sizeof(int));
  memcpy(c->bsido, bsido_in, 6 * B * sizeof(int));
  memcpy(c->coord, coord_in, 2 * P * sizeof(double));
  for (size_t k = 0; k < 3 * E; k++)
    if (c->intma[k] < 1 || c->intma[k] > npoin) { ora_destroy(c); return -7; }
  for (size_t b = 0; b < B; b++)
    if (c->bsido[6 * b] < 1 || c->bsido[6 * b] > npoin || c->bsido[6 * b + 1] < 1 ||
        c->bsido[6 * b + 1] > npoin || c->bsido[6 * b + 2] < 1 ||
        c->bsido[6 * b + 2] > nelem) { ora_destroy(c); return -7; }
  int mxsid = 2 * (nelem + nboun);
  c->iside = zal(8 * (size_t)mxsid, sizeof(int));
  int ns = ora_getsid(nelem, npoin, c->intma, c->iside, mxsid);
  if (ns < 0) { ora_destroy(c); return ns; }
  c->S = ns;
  const size_t S = ns;
  c->nx = zal(S, sizeof(double)); c->ny = zal(S, sizeof(double)); c->el = zal(S, sizeof(double));
  c->geome = zal(7 * E, sizeof(double)); c->relen = zal(E, sizeof(double));
  c->mmat = zal(3 * E, sizeof(double)); c->rsido = zal(3 * B, sizeof(double));
  c->vcord = zal(3 * V, sizeof(double));
  ora_normal(ns, c->iside, c->coord, c->nx, c->ny, c->el);
  ora_getgeo(nelem, c->intma, c->coord, c->geome);
  ora_getlen(nelem, c->geome, c->relen);
  /* dg_lin_convec: GETMAT for IMMAT = 1, else CONMAT (src/dg_lin_convec.f90:113-117) */
  if (prm->immat == 0) {
    c->cmmat = zal(9 * E, sizeof(double));
    ora_conmat(nelem, c->geome, c->cmmat);
  } else {
    ora_getmat(nelem, c->geome, c->mmat);
  }
  ora_getnor(nboun, c->bsido, c->coord, c->rsido);
  ora_vspace(rorder, torder, eta, w, c->vcord, &c->sumweight);
  /* ALOTIM (src/alotim.f90:12-18) with MXSPD = rv */
  c->dtec = 0.0;
  for (size_t e = 0; e < E; e++) {
    double d = (prm->csafm * c->relen[e]) / (prm->rv);
    if (e == 0 || d < c->dtec) c->dtec = d;
  }
  /* per-ordinate tables */
  c->cxa = zal(V, sizeof(double)); c->cya = zal(V, sizeof(double));
  c->cxb = zal(V, sizeof(double)); c->cyb = zal(V, sizeof(double));
  c->jac = zal(V, sizeof(double));
  {
    const double *vcord = c->vcord;
    double rv = prm->rv;
    for (size_t v = 1; v <= V; v++) {
      double et = VCORD(1, v), ze = VCORD(2, v);
      double th = ze * PI_R;
      double ra = (rv / 2) * (et + 1);
      double rb = et * (rv / 2) + (rv / 2);
      c->cxa[v - 1] = ra * cos(th); c->cya[v - 1] = ra * sin(th);
      c->cxb[v - 1] = rb * cos(th); c->cyb[v - 1] = rb * sin(th);
      c->jac[v - 1] = PI_R * rv * rv * 0.25 * (et + 1);
    }
  }
  c->f = zal(3 * V * E, sizeof(double));
  if (!c->f) { ora_destroy(c); return -100; }
  if (prm->jacobi) c->fsnap = zal(3 * V * E, sizeof(double));
  c->disnd = zal(3 * E, sizeof(double)); c->disux = zal(3 * E, sizeof(double));
  c->disuy = zal(3 * E, sizeof(double)); c->disps = zal(3 * E, sizeof(double));
  c->mu = zal(3 * E, sizeof(double)); c->eta_b = zal(B, sizeof(double));
  c->nd_pp = zal(P, sizeof(double)); c->rho_pp = zal(P, sizeof(double));
  c->uvel_pp = zal(P, sizeof(double)); c->vvel_pp = zal(P, sizeof(double));
  c->ps_pp = zal(P, sizeof(double)); c->temp_pp = zal(P, sizeof(double));
  c->flag = zal(P, sizeof(double));
  c->nd = zal(P, sizeof(double)); c->rho = zal(P, sizeof(double));
  c->uvel = zal(P, sizeof(double)); c->vvel = zal(P, sizeof(double));
  c->ps = zal(P, sizeof(double)); c->temp = zal(P, sizeof(double));
  c->nfo = zal(3 * E, sizeof(double)); c->fluxp = zal(3 * E, sizeof(double));
  c->fluyp = zal(3 * E, sizeof(double)); c->umean = zal(E, sizeof(double));
  c->rhs = zal(3 * E, sizeof(double)); c->delun = zal(3 * E, sizeof(double));
  c->residual = zal(V, sizeof(double));
  /* search tables with the reference's first-match rules */
  {
    const int *iside = c->iside, *bsido = c->bsido, *intma = c->intma;
    c->bou_of_side = zal(S, sizeof(int));
    for (size_t is = 1; is <= S; is++) {
      if (ISIDE(4, is) != 0) continue;
      int found = 0;
      for (int isb = 1; isb <= nboun; isb++)
        if (ISIDE(3, is) == BSIDO(3, isb) && ISIDE(1, is) == BSIDO(1, isb)) {
          found = isb; break;
        }
      c->bou_of_side[is - 1] = found;
    }
    c->flu_in = zal(2 * B, sizeof(int));
    for (size_t ib = 1; ib <= B; ib++) {
      int ip1 = BSIDO(1, ib), ip2 = BSIDO(2, ib), in1 = -1, in2 = -1;
      for (size_t is = 1; is <= S; is++) {
        if (ip1 == ISIDE(1, is) && ip2 == ISIDE(2, is)) { in1 = ISIDE(5, is); in2 = ISIDE(6, is); break; }
        if (ip2 == ISIDE(1, is) && ip1 == ISIDE(2, is)) { in1 = ISIDE(6, is); in2 = ISIDE(5, is); break; }
      }
      c->flu_in[2 * (ib - 1)] = in1;
      c->flu_in[2 * (ib - 1) + 1] = in2;
    }
    c->inflow_node = zal(3 * E, 1);
    unsigned char *ptin = zal(P + 1, 1);
    for (size_t ib = 1; ib <= B; ib++)
      if (BSIDO(4, ib) == 1) { ptin[BSIDO(1, ib)] = 1; ptin[BSIDO(2, ib)] = 1; }
    for (size_t ie = 1; ie <= E; ie++)
      for (int in = 1; in <= 3; in++)
        c->inflow_node[NE(in, ie)] = ptin[INTMA(in, ie)];
    free(ptin);
  }
  inicon2(c);
  /* initial MACROS + GETMAC (src/advnce.f90:139-160) */
  macros(c);
  getmac(c);
  c->itime = 0;
  c->sritm1 = 0.0;
  *out = c;
  return 0;
}

void ora_destroy(ora_ctx *c)
{
  if (!c) return;
  void *p[] = {c->intma, c->bsido, c->iside, c->coord, c->nx, c->ny, c->el,
               c->geome, c->relen, c->mmat, c->cmmat, c->rsido, c->vcord, c->cxa, c->cya,
               c->cxb, c->cyb, c->jac, c->f, c->fsnap, c->disnd, c->disux,
               c->disuy, c->disps, c->mu, c->eta_b, c->nd_pp, c->rho_pp,
               c->uvel_pp, c->vvel_pp, c->ps_pp, c->temp_pp, c->flag, c->nd,
               c->rho, c->uvel, c->vvel, c->ps, c->temp, c->nfo, c->fluxp,
               c->fluyp, c->umean, c->rhs, c->delun, c->residual,
               c->bou_of_side, c->flu_in, c->inflow_node, c->elem_sides,
               c->node_si, c->node_su, c->node_sv, c->node_sp};
  for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++) free(p[i]);
  free(c);
}

void ora_cmmat(const ora_ctx *c, double *cmmat)
{
  if (c->cmmat) memcpy(cmmat, c->cmmat, 9 * (size_t)c->E * sizeof(double));
  else memset(cmmat, 0, 9 * (size_t)c->E * sizeof(double));
}

void ora_get_f(const ora_ctx *c, double *f)
{
  memcpy(f, c->f, 3 * (size_t)c->V * c->E * sizeof(double));
}

void ora_set_f(ora_ctx *c, const double *f)
{
  memcpy(c->f, f, 3 * (size_t)c->V * c->E * sizeof(double));
  macros(c);
  getmac(c);
}

void ora_get_node_fields(const ora_ctx *c, double *disnd, double *disux,
                         double *disuy, double *disps, double *mu)
{
  size_t n = 3 * (size_t)c->E * sizeof(double);
  if (disnd) memcpy(disnd, c->disnd, n);
  if (disux) memcpy(disux, c->disux, n);
  if (disuy) memcpy(disuy, c->disuy, n);
  if (disps) memcpy(disps, c->disps, n);
  if (mu) memcpy(mu, c->mu, n);
}

void ora_get_point_fields(const ora_ctx *c, double *nd, double *rho,
                          double *uvel, double *vvel, double *ps, double *temp)
{
  size_t n = (size_t)c->P * sizeof(double);
  if (nd) memcpy(nd, c->nd_pp, n);
  if (rho) memcpy(rho, c->rho_pp, n);
  if (uvel) memcpy(uvel, c->uvel_pp, n);
  if (vvel) memcpy(vvel, c->vvel_pp, n);
  if (ps) memcpy(ps, c->ps_pp, n);
  if (temp) memcpy(temp, c->temp_pp, n);
}

void ora_get_eta(const ora_ctx *c, double *eta)
{
  memcpy(eta, c->eta_b, (size_t)c->B * sizeof(double));
}

/* ---- test hooks: one ordinate with stage snapshots ---- */

/* FLUCON2 + COMPMU + ALOTIM2 for the current f, as at the top of a step;
 * returns DTE through *dte. */
int ora_prepare_step(ora_ctx *c, double *dte)
{
  int st = flucon2(c);
  if (st) return st;
  compmu(c);
  double maxmu = c->mu[0];
  for (size_t k = 1; k < 3 * (size_t)c->E; k++) if (c->mu[k] > maxmu) maxmu = c->mu[k];
  double dteb = (maxmu > 0.0) ? 0.9 * (1 / maxmu) : 1000000.0;
  *dte = (dteb < c->dtec) ? dteb : c->dtec;
  if (c->prm.jacobi) memcpy(c->fsnap, c->f, 3 * (size_t)c->V * c->E * sizeof(double));
  return 0;
}

int ora_debug_ordinate(ora_ctx *c, int iv, double dte, double *nfo, double *umean,
                       double *rhs_elc, double *rhs_edge, double *delun, double *residual)
{
  ora_stage st = {nfo, umean, rhs_elc, rhs_edge, delun};
  g_stage = &st;
  int r = ordinate(c, iv, dte);
  g_stage = NULL;
  *residual = c->residual[iv - 1];
  return r;
}

void ora_velocity(const ora_ctx *c, int iv, double *ux, double *uy)
{
  *ux = c->cxa[iv - 1];
  *uy = c->cya[iv - 1];
}

void ora_tables(const ora_ctx *c, double *nx, double *ny, double *el, double *geome,
                double *relen, double *mmat, double *rsido, double *dtec)
{
  if (nx) memcpy(nx, c->nx, (size_t)c->S * sizeof(double));
  if (ny) memcpy(ny, c->ny, (size_t)c->S * sizeof(double));
  if (el) memcpy(el, c->el, (size_t)c->S * sizeof(double));
  if (geome) memcpy(geome, c->geome, 7 * (size_t)c->E * sizeof(double));
  if (relen) memcpy(relen, c->relen, (size_t)c->E * sizeof(double));
  if (mmat) memcpy(mmat, c->mmat, 3 * (size_t)c->E * sizeof(double));
  if (rsido) memcpy(rsido, c->rsido, 3 * (size_t)c->B * sizeof(double));
  if (dtec) *dtec = c->dtec;
}

void ora_cinf(const ora_ctx *c, double *cinf)
{
  memcpy(cinf, c->cinf, 4 * sizeof(double));
}
