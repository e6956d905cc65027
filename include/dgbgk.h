/*
 * dgbgk.h -- C ABI of the MI355X DG-BGK explicit time-march library.
 *
 * The library replaces the reference's time march ADVNCE
 * (src/advnce.f90:3-492, called once from src/dg_lin_convec.f90:185-206)
 * and the one-time setup that feeds it (src/dg_lin_convec.f90:96-177).
 * Arguments are plain pointers and sizes in the reference's conventions:
 * integer tables hold 1-based ids stored column-major exactly as the
 * Fortran arrays (INTMA(3,NELEM), BSIDO(6,NBOUN), COORD(2,NPOIN)), and the
 * distribution function crosses the boundary in the reference layout
 * DISNF(3,VNPNT,NELEM) as double.  Inside, f stays resident in HBM,
 * velocity-major, in FP64 or FP32.
 *
 * Every call returns 0 on success or a negative DGBGK_ERR_* code; the
 * reference's STOP conditions map to codes (see below) and
 * dgbgk_last_error() gives the message.  One context owns one HIP stream;
 * calls on a context must come from one host thread at a time.
 */
#ifndef DGBGK_H
#define DGBGK_H

#ifdef __cplusplus
extern "C" {
#endif

#define DGBGK_OK 0
#define DGBGK_ERR_ARG -1
#define DGBGK_ERR_MXSID -2          /* src/getsid.f90:137-140 */
#define DGBGK_ERR_BOUNDARY_SEARCH -3 /* src/getbou.f90:48-53 (search overruns) */
#define DGBGK_ERR_WALL_SIDE -4      /* src/flucon2.f90:125-128 */
#define DGBGK_ERR_SUMBOTTOM -5      /* src/flucon2.f90:227-231 */
#define DGBGK_ERR_REFLECT_LOOKUP -6 /* src/getbou.f90:223-250 (out of VCORD) */
#define DGBGK_ERR_MESH -7
#define DGBGK_ERR_UNSUPPORTED -8    /* a configuration the library does not build */
#define DGBGK_ERR_HIP -9
#define DGBGK_ERR_NOMEM -10
#define DGBGK_ERR_TILE -11
#define DGBGK_ERR_STATE -12

#define DGBGK_FP64 0 /* f, flux and relaxation in FP64 */
#define DGBGK_FP32 1 /* f, flux and relaxation in FP32; moments in FP64 */

#define DGBGK_LOOKUP_REFERENCE 0 /* ring rule of src/getbou.f90:223-243 */
#define DGBGK_LOOKUP_CORRECTED 1 /* ring found by its own eta */

#define DGBGK_ORDER_REFERENCE 0 /* in-place velocity sweep (src/advnce.f90:273-383) */
#define DGBGK_ORDER_JACOBI 1    /* reflections read f of the previous step */

/* Deck keys of the /Input/ namelist (src/input_variable_module.f90:5-33)
 * that the time march reads, plus build choices. */
typedef struct dgbgk_params {
  double rv, csafm, alpha, rgas, d, m; /* IVD%rv, CSAFM, ALPHA, R, d, M */
  double t1, p1, u0, v0;               /* IVD%T1, P1 [bar], U0, V0 */
  double cinf[4];                      /* IVD%CINF: T, P [bar], U, V */
  int inf;                             /* IVD%INF */
  int immat;                           /* IVD%IMMAT: 1 lumped mass (GETINC), 0 the
                                          consistent-mass branch (CONMAT, CMMINC,
                                          src/advnce.f90:349-361); else an error */
  int precision;                       /* DGBGK_FP64 or DGBGK_FP32 */
  int lookup;                          /* DGBGK_LOOKUP_* */
  int order;                           /* DGBGK_ORDER_* */
  int vchunk;                          /* ordinates per update block, 0 = auto */
} dgbgk_params;

/* One time step's record: the values ADVNCE prints or writes per step. */
typedef struct dgbgk_step_info {
  double sumres; /* MAXVAL(ABS(RESIDUAL)), src/advnce.f90:422 */
  double logres; /* LOG(SUMRES/SRITM1), src/advnce.f90:427-429 */
  double dte;    /* time step, src/advnce.f90:251-268 */
  double mass;   /* sum over DG nodes of DISND*RJ/6 after the step */
  double lift;   /* GETFOR on the last gathered point pressure, */
  double drag;   /* src/advnce.f90:431 (refreshed every 1000 steps) */
  double change; /* max |f^{k+1} - f^k| over all nodes and ordinates, after
                    the clamp and the inflow reset (no reference counterpart:
                    SUMRES counts the increment before the reset, so at
                    nodes held at the inflow state it never vanishes; this
                    is the steady-state measure).  DGBGK_FP32: the step's
                    increment before the store's rounding, |d| (f where
                    clamped, |f_inf - f| where reset) */
} dgbgk_step_info;

typedef struct dgbgk_ctx dgbgk_ctx;

/* Setup: the reference's GETSID, NORMAL, GETGEO, GETLEN, GETMAT, GETNOR,
 * VSPACE, ALOTIM (src/dg_lin_convec.f90:96-123, src/spectralbgk.f90:205,
 * src/advnce.f90:133), the INICON2 Maxwellian seed (src/inicon2.f90:120-193)
 * and the initial MACROS (src/advnce.f90:139-160).  eta/w are the radial
 * quadrature (the Lobatto file's two columns, src/vspace.f90:16-25).
 * Every per-element array of this interface is in the caller's element
 * numbering; the device stores elements in Hilbert order of their
 * centroids (DGBGK_REORDER=0 in the environment keeps the input order). */
int dgbgk_create(int nelem, int npoin, int nboun, const int *intma,
                 const double *coord, const int *bsido, int rorder, int torder,
                 const double *eta, const double *w, const dgbgk_params *prm,
                 dgbgk_ctx **out);
void dgbgk_destroy(dgbgk_ctx *ctx);

/* Message of the last failing call on this thread. */
const char *dgbgk_last_error(void);

/* NTIME steps of ADVNCE's loop body (src/advnce.f90:220-449); info gets
 * nsteps records (or pass NULL). */
int dgbgk_step(dgbgk_ctx *ctx, int nsteps, dgbgk_step_info *info);

/* DISNF_PP in the reference layout (3,VNPNT,NELEM), column-major. */
int dgbgk_get_f(dgbgk_ctx *ctx, double *f);
int dgbgk_set_f(dgbgk_ctx *ctx, const double *f);

/* Maxwellian seed from per-point primitives (rho, u, v, T), the INICON2
 * formula applied point by point; the warm start's entry into the method. */
int dgbgk_seed_maxwellian(dgbgk_ctx *ctx, const double *rho, const double *u,
                          const double *v, const double *temp);

/* MACROS node fields (3,NELEM): DISND, DISUX, DISUY, DISPS and COMPMU's MU
 * (src/macros.f90:126-164, src/compmu.f90:86-97) for the current f. */
int dgbgk_node_fields(dgbgk_ctx *ctx, double *disnd, double *disux,
                      double *disuy, double *disps, double *mu);

/* MACROS point averages ND, RHO, UVEL, VVEL, PS, TEMP (src/macros.f90:156-188;
 * GETMAC is the identity on one device) for the current f. */
int dgbgk_point_fields(dgbgk_ctx *ctx, double *nd, double *rho, double *uvel,
                       double *vvel, double *ps, double *temp);

/* GETFOR (src/getfor.f90:27-49) on the current point pressure. */
int dgbgk_forces(dgbgk_ctx *ctx, double *lift, double *drag);

/* FLUCON2's ETA per boundary side (src/flucon2.f90:218-241), last step. */
int dgbgk_wall_eta(dgbgk_ctx *ctx, double *eta);

/* Smallest and largest finite value of the current f over every node and
 * ordinate, and the count of non-finite values (no counterpart in the
 * reference: a device-side check of f >= 0 and finiteness that needs no
 * host copy of f). */
int dgbgk_f_stats(dgbgk_ctx *ctx, double *fmin, double *fmax, long long *nonfinite);

/* dgbgk_f_stats without any allocation (no counterpart in the reference;
 * for callers that cannot allocate device memory: tight serving or stepping
 * loops, code that forbids hipMalloc).  Same values as dgbgk_f_stats.  Its
 * scratch is a fixed 1.5 MiB device array per GPU, held until the process
 * ends, so at most 4096 ordinates (DGBGK_ERR_ARG beyond).  Calls that
 * overlap on one GPU share that array and race; one call at a time per
 * GPU. */
int dgbgk_f_stats_noalloc(dgbgk_ctx *ctx, double *fmin, double *fmax, long long *nonfinite);

/* Sizes: NSIDE, VNPNT, NELEM, NPOIN, NBOUN, steps taken. */
int dgbgk_sizes(dgbgk_ctx *ctx, int *nside, int *nvel, int *nelem, int *npoin,
                int *nboun, int *itime);

/* Host-side setup tables, for bit-exact checks against the reference:
 * ISIDE(8,NSIDE), NX/NY/EL(NSIDE), GEOME(7,NELEM), RELEN(NELEM),
 * MMAT(3,NELEM), RSIDO(3,NBOUN), VCORD(3,VNPNT), SUMWEIGHT, DTEc.  Any
 * pointer may be NULL. */
int dgbgk_setup_tables(dgbgk_ctx *ctx, int *iside, double *nx, double *ny,
                       double *el, double *geome, double *relen, double *mmat,
                       double *rsido, double *vcord, double *sumweight,
                       double *dtec);

/* The same setup without a GPU (no HIP call): for host checks and tools. */
typedef struct dgbgk_setup dgbgk_setup;
int dgbgk_setup_create(int nelem, int npoin, int nboun, const int *intma,
                       const double *coord, const int *bsido, int rorder,
                       int torder, const double *eta, const double *w,
                       const dgbgk_params *prm, dgbgk_setup **out);
int dgbgk_setup_sizes(dgbgk_setup *s, int *nside, int *nvel, int *ntiles,
                      int *nwall_sides, int *ring_separable);
int dgbgk_setup_tables_host(dgbgk_setup *s, int *iside, double *nx, double *ny,
                            double *el, double *geome, double *relen,
                            double *mmat, double *rsido, double *vcord,
                            double *sumweight, double *dtec);
/* GETBOU's reflection table of the same setup (src/getbou.f90:190-250):
 * side[k] is wall slot k's boundary side (1-based row of BSIDO), k <
 * nwall_sides; for slot k and ordinate v (arrays nwall_sides x nvel,
 * slot-major) i1/i2 are the 1-based ordinates the reflected velocity is
 * interpolated between (0 where v leaves through the wall and is not
 * reflected), zd1 the ZETA width of their bracket and zd2 the reflected
 * ZETA's offset from i1's.  Any pointer may be NULL. */
int dgbgk_setup_reflections(dgbgk_setup *s, int *side, int *i1, int *i2,
                            double *zd1, double *zd2);
void dgbgk_setup_destroy(dgbgk_setup *s);

/* Kernel timing with HIP events on the context's stream.  Kernel ids:
 * 0 update, 1 wall pass, 2 moments, 3 eta, 4 finalize+forces, 5 points. */
int dgbgk_timing(dgbgk_ctx *ctx, int enable);
int dgbgk_kernel_time(dgbgk_ctx *ctx, int kernel, double *ms_total, long *count);
int dgbgk_sync(dgbgk_ctx *ctx);

/* Bytes of f per ordinate-node (8 or 4) and algorithmic bytes per DOF-v
 * update of the update kernel (one read and one write of f). */
int dgbgk_dof_bytes(dgbgk_ctx *ctx, int *bytes_per_value);

/* Launch configuration the library chose for this mesh (no counterpart in
 * the reference, which has no device launch): update tile slots, f
 * ordinates in flight, ordinates per update block and chunks, tiles, and
 * the moments kernel's wavefronts per node. */
int dgbgk_launch_config(dgbgk_ctx *ctx, int *slots, int *depth, int *vchunk, int *nvchunks,
                        int *ntiles, int *mom_split);

/* ---- learned warm start: coordinate MLP (no counterpart in the reference) ----
 * fin -> width (SiLU), (hidden-1) x width -> width (SiLU), width -> fout.
 * BF16 weights and activations, FP32 accumulation (MFMA) and FP32 master
 * weights, Adam.  Parameters are packed per layer as W [dout][din] then
 * b [dout].  width and batch must be multiples of 64, fin and fout at
 * most 8. */
typedef struct dgbgk_mlp dgbgk_mlp;
int dgbgk_mlp_create(int fin, int width, int hidden, int fout, int batch, dgbgk_mlp **out);
void dgbgk_mlp_destroy(dgbgk_mlp *m);
const char *dgbgk_mlp_last_error(void);
int dgbgk_mlp_nparams(dgbgk_mlp *m, long *n);
int dgbgk_mlp_set_params(dgbgk_mlp *m, const float *params);
int dgbgk_mlp_get_params(dgbgk_mlp *m, float *params);
/* y[n][fout] = net(x[n][fin]) */
int dgbgk_mlp_forward(dgbgk_mlp *m, long n, const float *x, float *y);
/* one batch of rows (multiple of 64, <= batch): MSE loss and its gradient */
int dgbgk_mlp_loss_grad(dgbgk_mlp *m, int rows, const float *x, const float *y, float *loss,
                        float *grad);
/* device-resident training set and Adam steps on random batches */
int dgbgk_mlp_set_data(dgbgk_mlp *m, long n, const float *x, const float *y);
int dgbgk_mlp_train(dgbgk_mlp *m, int steps, float lr, unsigned seed, float *losses);

#ifdef __cplusplus
}
#endif
#endif
