/*
 * dgbgk_oracle.h -- FP64 CPU restatement of the DG-BGK explicit time march.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the "port" CPU baseline; it is never linked into, loaded by or
 * called from the product library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may use it.
 *
 * Conventions follow the reference (DrBenEvans/DG-BGK, Fortran 90):
 *   - integer tables hold 1-based ids, stored column-major exactly like the
 *     Fortran arrays (INTMA(3,E) -> intma[3*(e-1)+(j-1)]);
 *   - the distribution function uses the reference layout DISNF(3,V,E):
 *     f[((e-1)*V + (v-1))*3 + (a-1)];
 *   - every routine names the reference file:line it restates.
 */
#ifndef DGBGK_ORACLE_H
#define DGBGK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mesh preprocessing (one-time setup) ---- */
int ora_getsid(int nelem, int npoin, const int *intma, int *iside, int mxsid);
void ora_normal(int nside, const int *iside, const double *coord,
                double *nx, double *ny, double *el);
void ora_getgeo(int nelem, const int *intma, const double *coord, double *geome);
void ora_getlen(int nelem, const double *geome, double *relen);
void ora_getmat(int nelem, const double *geome, double *mmat);
void ora_conmat(int nelem, const double *geome, double *cmmat); /* CMMAT(3,3,NELEM) */
void ora_getnor(int nboun, const int *bsido, const double *coord, double *rsido);
void ora_vspace(int rorder, int torder, const double *eta, const double *w,
                double *vcord, double *sumweight);

/* ---- run context ---- */
typedef struct ora_params {
  double rv, csafm, alpha, rgas, d, m;
  double t1, p1, u0, v0;  /* INICON2 Maxwellian seed */
  double cinf[4];         /* inflow state (T, P[bar], U, V) */
  int inf;                /* IVD%INF */
  int lookup;             /* 0 = reference ring lookup, 1 = corrected */
  int jacobi;             /* 0 = in-place velocity sweep (reference), 1 = Jacobi */
  int faithful_scans;     /* 1 = literal O(N^2) searches of the reference */
  int mutant;             /* 0 = the restatement; otherwise a planted mistake,
                             used only by the tests that show a pin catches it:
                             ORA_MUTANT_* below */
  int immat;              /* IVD%IMMAT: 1 lumped mass (GETMAT, GETINC), 0 the
                             consistent-mass branch (CONMAT, CMMINC) */
} ora_params;

/* Planted mistakes (tests only).  Each is a plausible misreading of the
 * reference; the pins must reject every one of them. */
#define ORA_MUTANT_DOWNWIND 1     /* EDGFLXOPT takes the downwind element's values */
#define ORA_MUTANT_SWAP_NODES 2   /* EDGFLXOPT reads ISIDE(5:6) swapped */
#define ORA_MUTANT_GETMAC_EVERY 3 /* GETMAC's point gather every step, not every 1000th */
#define ORA_MUTANT_RES_AFTER 4    /* residual taken after INFLOW's reset */
#define ORA_MUTANT_CMMINC_RHS_JN 5 /* CMMINC multiplies CMMAT(IN,JN) by RHS(JN) */

typedef struct ora_ctx ora_ctx;

/* Creates a context from the raw mesh (INTMA, COORD, BSIDO) and the radial
 * quadrature (eta, w); status 0 on success, negative on a reference STOP
 * condition.  f is seeded with the INICON2 Maxwellian. */
int ora_create(int nelem, int npoin, int nboun, const int *intma,
               const double *coord, const int *bsido, int rorder, int torder,
               const double *eta, const double *w, const ora_params *prm,
               ora_ctx **out);
void ora_destroy(ora_ctx *c);

int ora_nside(const ora_ctx *c);
int ora_nvel(const ora_ctx *c);
const int *ora_iside(const ora_ctx *c);
const double *ora_vcord(const ora_ctx *c);
double ora_sumweight(const ora_ctx *c);

/* f in the reference layout (3,V,E). */
void ora_get_f(const ora_ctx *c, double *f);
/* CMMAT(3,3,NELEM) of a context created with immat = 0 (zeros otherwise). */
void ora_cmmat(const ora_ctx *c, double *cmmat);
void ora_set_f(ora_ctx *c, const double *f); /* re-runs the initial MACROS */

/* Per-step record. */
typedef struct ora_step_info {
  double sumres;   /* MAXVAL(ABS(RESIDUAL)) */
  double logres;   /* LOG(SUMRES/SRITM1) */
  double dte;      /* time step used */
  double mass;     /* sum over DG nodes of DISND * RJ/6 after the step */
  double lift, drag; /* GETFOR on the point pressure last gathered (stale) */
} ora_step_info;

int ora_step(ora_ctx *c, int nsteps, ora_step_info *info);

/* Threads for ora_step (default 1, the serial restatement).  More threads
 * give the same values bit for bit (per-node sums and each element's edge
 * terms keep the serial order); used for the all-core CPU baseline. */
int ora_set_threads(ora_ctx *c, int nthreads);

/* FP32 emulation of stages of the step, for precision studies of the
 * FP32-bulk form (the restatement itself is FP64; flags 0 restores it).
 * Storing f rounds the current f once. */
#define ORA_EMU_STORE_RN 1 /* f stored in FP32, rounded to nearest */
#define ORA_EMU_STORE_SR 2 /* f stored in FP32 with a +-1/2 ulp dither (stochastic) */
#define ORA_EMU_PARAMS 4   /* Maxwellian parameters and mu in FP32, FP32 exponential */
#define ORA_EMU_ARITH 8    /* predictor, volume, edge, wall and increment arithmetic in FP32 */
#define ORA_EMU_DTE 16     /* the time step rounded to FP32 */
#define ORA_EMU_PRED 32    /* only the predictor and volume term (GETFLA, GETELC) in FP32 */
#define ORA_EMU_EDGE 64    /* only the edge and wall fluxes (EDGFLXOPT, GETBOU) in FP32 */
#define ORA_EMU_INC 128    /* only the increment (RHS*DTE, GETINC) in FP32 */
#define ORA_EMU_FORM 256   /* difference-form predictor and the volume term folded
                              into the edges (not the reference's order; studies) */
void ora_set_emulation(ora_ctx *c, int flags);

/* DIS* node fields (3,E) and point fields (P) from the most recent MACROS. */
void ora_get_node_fields(const ora_ctx *c, double *disnd, double *disux,
                         double *disuy, double *disps, double *mu);
void ora_get_point_fields(const ora_ctx *c, double *nd, double *rho,
                          double *uvel, double *vvel, double *ps, double *temp);
void ora_get_eta(const ora_ctx *c, double *eta);

/* ---- individual per-step routines (for unit checks) ---- */
void ora_gteqnf(int nelem, const double *disnd, const double *disux,
                const double *disuy, const double *disps, double ux, double uy,
                double m, double *nfo);
void ora_getfor(int nboun, const int *bsido, const double *ps,
                const double *coord, double *lift, double *drag);
double ora_inflow_value(const double *cinf, double rgas, double m, double ux,
                        double uy);

/* ---- test hooks ---- */
int ora_prepare_step(ora_ctx *c, double *dte);
int ora_debug_ordinate(ora_ctx *c, int iv, double dte, double *nfo, double *umean,
                       double *rhs_elc, double *rhs_edge, double *delun, double *residual);
void ora_velocity(const ora_ctx *c, int iv, double *ux, double *uy);
void ora_tables(const ora_ctx *c, double *nx, double *ny, double *el, double *geome,
                double *relen, double *mmat, double *rsido, double *dtec);
void ora_cinf(const ora_ctx *c, double *cinf);

#ifdef __cplusplus
}
#endif
#endif
