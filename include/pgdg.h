/*
 * pgdg.h - C ABI of the MI355X-native PG-DG time step.
 *
 * The reference (gaotr3306/PGDG_p2) has no plugin or FFI boundary: its hot
 * path is the set of C++ member functions of `class MPM` and
 * `class Particle_base` (S3 = 3D_generalized/PGDG_3D_spheric,
 * T2 = 2D_cases/Triple_point_2phase).  Each entry point below replaces one of
 * those member functions; the comment on each names the reference routine
 * (file:line) it stands in for.  The C++ host class `pgdg::MPM`
 * (include/pgdg_mpm.hpp) keeps the reference method names and argument
 * lists and forwards to these functions.
 *
 * Conventions
 *   - every function returns an int status: PGDG_OK (0) or a negative code;
 *     nothing aborts, nothing throws across the boundary;
 *   - pointers documented as "device" are HIP device pointers, "host" are
 *     ordinary host memory; `stream` is a hipStream_t passed as void*
 *     (NULL = the null stream);
 *   - all floating-point state is IEEE binary64 unless stated otherwise.
 */
#ifndef PGDG_H
#define PGDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PGDG_OK               0
#define PGDG_ERR_ARG         -1  /* invalid argument / shape               */
#define PGDG_ERR_HIP         -2  /* a HIP runtime call failed              */
#define PGDG_ERR_CAPACITY    -3  /* particle / list capacity exceeded      */
#define PGDG_ERR_NONFINITE   -4  /* non-finite value detected (checked op) */
#define PGDG_ERR_STATE       -5  /* call out of order / missing setup      */
#define PGDG_ERR_NODEVICE    -6  /* no HIP device visible                  */

/* ---- library identity -------------------------------------------------- */
/* Returns a static string "pgdg-mi355x <version> gfx950". */
const char* pgdg_version(void);

/* Number of visible HIP devices (0 when none).  Makes no allocation. */
int pgdg_device_count(int* count);

/* Device memory helpers, so that callers without a GPU framework of their
 * own can stage inputs (no reference counterpart).  kind: 0 host->device,
 * 1 device->host, 2 device->device.  pgdg_device_sync waits for all work. */
int pgdg_malloc(void** ptr, int64_t bytes);
int pgdg_free(void* ptr);
int pgdg_memcpy(void* dst, const void* src, int64_t bytes, int kind);
int pgdg_device_sync(void);

/* Streaming device-to-device copy by a plain kernel (16-byte aligned
 * buffers, bytes a multiple of 16): known HBM traffic for calibrating
 * hardware byte counters against algorithmic bytes.                       */
int pgdg_copy_kernel(void* dst, const void* src, int64_t bytes, void* stream);

/* Stream-ordered timing (HIP events) for callers without a GPU framework:
 * record on the stream the work is launched on; elapsed waits for `stop`. */
int pgdg_event_create(void** ev);
int pgdg_event_record(void* ev, void* stream);
int pgdg_event_elapsed_ms(void* start, void* stop, float* ms);
int pgdg_event_destroy(void* ev);

/* ======================================================================== *
 *  MGFM two-medium Riemann solve (A9) and stiffened-gas EOS (A10)          *
 * ======================================================================== */

/* Reference variants of MPM::MGFM.
 *   PGDG_MGFM_S3 : S3:mpm.cpp:882-926  (clamps p,rho >= 1e-10,
 *                  p*_0 = max(pL,pR), returns rho*_L and u*_L at the
 *                  pre-update p*)
 *   PGDG_MGFM_T2 : T2:mpm.cpp:1373-1416 (clamps p >= 0, p*_0 = (pL+pR)/2,
 *                  u*_L uses the updated p* with the last W_L)           */
#define PGDG_MGFM_S3 0
#define PGDG_MGFM_T2 1

/* Stopping rule.
 *   PGDG_TOL_REFERENCE : stop when |f(p*)| <= max(1e-8, 1e-3 |uL-uR|)
 *                        (f evaluated before the update) or 50 iterations
 *                        (S3:mpm.cpp:892, 921);
 *   PGDG_TOL_RELATIVE  : stop when |p*_new - p*_old| <= rtol * p*_new or
 *                        max_iter; the returned (rho*_L, u*_L) are then
 *                        evaluated at the returned p* (project extension,
 *                        no reference counterpart).                         */
#define PGDG_TOL_REFERENCE 0
#define PGDG_TOL_RELATIVE  1

typedef struct pgdg_eos_pair {
  double gamma_l, pinf_l;   /* EOS of the left  medium (reference `s`)  */
  double gamma_r, pinf_r;   /* EOS of the right medium (reference `s1`) */
} pgdg_eos_pair;

typedef struct pgdg_mgfm_opts {
  int    variant;    /* PGDG_MGFM_S3 or PGDG_MGFM_T2                     */
  int    tol_mode;   /* PGDG_TOL_REFERENCE or PGDG_TOL_RELATIVE          */
  double rtol;       /* relative tolerance for PGDG_TOL_RELATIVE         */
  int    max_iter;   /* iteration cap (the reference uses 50)            */
} pgdg_mgfm_opts;

/* Batched MGFM, replaces `Vector<Real,3> MPM::MGFM(ul, ur, pl, pr, rhol,
 * rhor, s, s1)` (S3:mpm.cpp:882; T2:mpm.cpp:1373) called once per face.
 *   in     device, SoA [6][n] doubles: ul, ur, pl, pr, rhol, rhor
 *   out    device, SoA [3][n] doubles: rho*_L, p*, u*_L
 *   iters  device, [n] int32 Newton iterations taken (may be NULL)
 *   pstar0 device, [n] initial p* guesses (NULL = the variant's own guess);
 *          used by the learned-guess path.                                 */
int pgdg_mgfm_batch(const double* in, double* out, int32_t* iters, int64_t n,
                    const pgdg_eos_pair* eos, const pgdg_mgfm_opts* opts,
                    const double* pstar0, void* stream);

/* Elementwise y = x^(1./dim) rounded to nearest, the particle spacing
 * pow(vol, 1./DIM) of S3:mpm.cpp:1089 / S3:particle_flow.cpp:473
 * (dim 2: sqrt; dim 3: correctly rounded x^(1./3.)).  x, y device.       */
int pgdg_root_dim(const double* x, double* y, int64_t n, int dim, void* stream);

/* ======================================================================== *
 *  Learned initial guess for the MGFM Newton solve (no reference          *
 *  counterpart: the reference starts from max(pL, pR), S3:mpm.cpp:891)    *
 * ======================================================================== */

/* A 4-layer MLP 12 -> 32 -> 32 -> 32 -> 2 (ReLU) mapping the standardised
 * face features (ln pL, ln pR, ln rhoL, ln rhoR, uL, uR, (pL - pR) / (pL +
 * pR), uL - uR, gammaL, ln(1 + PinfL), gammaR, ln(1 + PinfR); the two
 * differences formed in FP64 so that small jumps survive the BF16 inputs)
 * to the
 * standardised targets
 * (ln p* - ln max(pL, pR), u*).  Matrix products run on MFMA with BF16
 * operands and FP32 accumulation; the master parameters and the optimiser
 * (Adam) are FP32.  Parameter vector layout (FP32, PGDG_MLP_NPARAM):
 * W0[16][32] (rows 12..15 zero), b0[32], W1[32][32], b1[32], W2[32][32],
 * b2[32], W3[32][16] (columns 2..15 zero), b3[16] (entries 2..15 zero).  */
#define PGDG_MLP_NIN     12
#define PGDG_MLP_NPAD    16
#define PGDG_MLP_WIDTH   32
#define PGDG_MLP_NOUT    2
#define PGDG_MLP_NPARAM  3184

typedef struct pgdg_mlp pgdg_mlp;

typedef struct pgdg_mlp_norm {
  float feat_mu[PGDG_MLP_NIN], feat_sd[PGDG_MLP_NIN];
  float out_mu[PGDG_MLP_NOUT], out_sd[PGDG_MLP_NOUT];
  /* A predicted |ln(p*0 / max(pL, pR))| below keep leaves the initial p* at
   * max(pL, pR): the network cannot place p* closer than its own residual
   * there.  Training sets keep to twice the residual RMS of that output on
   * the training faces; 0 always takes the network's p*0.               */
  float keep;
} pgdg_mlp_norm;

typedef struct pgdg_mlp_train_opts {
  int      epochs;       /* passes over the training faces                  */
  int      batch;        /* samples per optimiser step (multiple of 64)     */
  float    lr;           /* Adam step size (cosine-decayed to lr/20)        */
  uint64_t seed;         /* parameter initialisation                        */
  int      variant;      /* MGFM variant that labels the faces              */
} pgdg_mlp_train_opts;

int pgdg_mlp_create(pgdg_mlp** out);
int pgdg_mlp_destroy(pgdg_mlp* m);
/* Labels the n device faces in[6][n] (ul, ur, pl, pr, rhol, rhor) with the
 * exact MGFM solve (relative tolerance 1e-12), sets the normalisation from
 * them and trains.  loss_hist (host, may be NULL) receives the mean loss of
 * each optimiser step, up to hist_cap entries; *steps the number of steps. */
int pgdg_mlp_train(pgdg_mlp* m, const double* in, int64_t n, const pgdg_eos_pair* eos,
                   const pgdg_mlp_train_opts* opts, float* loss_hist, int hist_cap, int* steps,
                   void* stream);
/* p*0 for each face (device in[6][n] -> device pstar0[n]); a non-finite or
 * non-positive prediction falls back to max(pl, pr).                      */
int pgdg_mlp_guess(pgdg_mlp* m, const double* in, int64_t n, const pgdg_eos_pair* eos,
                   double* pstar0, void* stream);
/* pgdg_mgfm_batch with the learned p*0 (the Newton iteration still runs to
 * the requested tolerance).                                                */
int pgdg_mgfm_batch_learned(pgdg_mlp* m, const double* in, double* out, int32_t* iters, int64_t n,
                            const pgdg_eos_pair* eos, const pgdg_mgfm_opts* opts, void* stream);
/* Faces whose media differ from face to face: pair[i] (device, int32, may be
 * NULL when npairs == 1) selects eos[pair[i]] for face i, as the reference's
 * MGFM(ul, ur, pl, pr, rhol, rhor, s, s1) (S3:mpm.cpp:882; UX:mpm.cpp:739)
 * takes the two media with every face.  pgdg_mlp_train_pairs: as
 * pgdg_mlp_train over such faces (npairs <= PGDG_MAX_PHASE^2; the EOS
 * features of each face enter the network).  pgdg_mgfm_batch_pairs: the
 * batched solve over such faces, from the network's p*0 when m is given
 * (trained), else from the variant's own.                                 */
int pgdg_mlp_train_pairs(pgdg_mlp* m, const double* in, const int32_t* pair, int64_t n,
                         const pgdg_eos_pair* eos, int npairs, const pgdg_mlp_train_opts* opts,
                         float* loss_hist, int hist_cap, int* steps, void* stream);
int pgdg_mgfm_batch_pairs(pgdg_mlp* m, const double* in, const int32_t* pair, double* out,
                          int32_t* iters, int64_t n, const pgdg_eos_pair* eos, int npairs,
                          const pgdg_mgfm_opts* opts, void* stream);
/* Parameters and normalisation (host buffers).                             */
int pgdg_mlp_get_params(pgdg_mlp* m, float* params, pgdg_mlp_norm* norm);
int pgdg_mlp_set_params(pgdg_mlp* m, const float* params, const pgdg_mlp_norm* norm);
/* One forward (y[n][2], standardised) or forward + backward on standardised
 * device inputs x[n][16] and targets t[n][2] (n a multiple of 64): mean loss
 * 0.5 * sum_o (y - t)^2 / n and its gradient grad[PGDG_MLP_NPARAM] (device).  */
int pgdg_mlp_forward(pgdg_mlp* m, const float* x, float* y, int64_t n, void* stream);
int pgdg_mlp_loss_grad(pgdg_mlp* m, const float* x, const float* t, int64_t n, float* grad,
                       float* loss, void* stream);

/* ======================================================================== *
 *  MLS particle-to-grid moment assembly with a per-cell least-squares      *
 *  solve (extension of MPM::MLS_P2G, S3:mpm.cpp:274-382)                   *
 * ======================================================================== */

/* Precision of the assembly.  The per-cell factorisation and solve are
 * FP64 in both.
 *   PGDG_PREC_FP64  : FP64 inputs (absolute positions), FP64 sums;
 *   PGDG_PREC_MIXED : FP32 inputs (cell-local coordinates) and FP32 sums. */
#define PGDG_PREC_FP64  0
#define PGDG_PREC_MIXED 1

typedef struct pgdg_mls_problem {
  int     n[3];          /* cells per axis on [0, n*dx]^3, index (i*n1+j)*n2+k */
  double  dx;            /* cell size                                        */
  double  cut;           /* particle cut-off distance (reference 1.5*p_sz)   */
  double  h;             /* Wendland kernel width (reference 1.6*p_g)        */
  int     nfields;       /* 1..6                                             */
  int     order;         /* 2: P2 (10 zero-mean modes), 1: P1 (4), 0: P0     */
  int64_t np;            /* particles, sorted by cell                        */
  const int64_t* off;    /* device [nc+1] first particle of each cell        */
  const double*  x;      /* device [3][np] positions        (FP64 path)      */
  const double*  vol;    /* device [np] volumes             (FP64 path)      */
  const double*  val;    /* device [nf][np] field values    (FP64 path)      */
  const float*   u;      /* device [3][np] cell-local coords in [0,1) (mixed)*/
  const float*   volf;   /* device [np]                     (mixed)          */
  const float*   valf;   /* device [nf][np]                 (mixed)          */
} pgdg_mls_problem;

/* Per-cell weighted least-squares fit of the fields of the particles of the
 * 3x3x3 neighbour cells within `cut` (weights W(r, h) * vol, the reference's
 * MLS_P2G weights) in the zero-mean P2 basis of the cell-scaled coordinate.
 * coef: device [nf][10][nc]; order_used: device [nc] (order actually
 * solved: a singular moment matrix falls back to P1, then P0) or NULL.    */
int pgdg_mls_ls_fit(const pgdg_mls_problem* p, int precision, double* coef, int32_t* order_used,
                    void* stream);
/* ======================================================================== *
 *  PG-DG time step: case description and host state layout                *
 * ======================================================================== */

/* Boundary-face tags, as in the reference Cell_dg::bound
 * (S3:mpm_ini.cpp:447-549): 0 interior, 1 wall, -1 inlet, -2 outlet,
 * -3 symmetry.                                                           */
#define PGDG_BC_INTERIOR  0
#define PGDG_BC_WALL      1
#define PGDG_BC_INLET    -1
#define PGDG_BC_OUTLET   -2
#define PGDG_BC_SYMMETRY -3

/* Side order of pgdg_case.bc, the reference face order of a cell:
 * yL, xR, yR, xL, zL, zR (S3:mpm_ini.cpp:447-549; T2:mpm_ini.cpp:329-375). */
enum { PGDG_SIDE_YL = 0, PGDG_SIDE_XR, PGDG_SIDE_YR, PGDG_SIDE_XL, PGDG_SIDE_ZL, PGDG_SIDE_ZR };

/* Semantics of the reference source the case follows.
 *   PGDG_SEM_T2 : 2-D triple point (T2/T3): every boundary face is a wall in
 *                 the DG flux, T2 MGFM, per-phase v in the cell acceleration,
 *                 no inlet/outlet and no mark_in;
 *   PGDG_SEM_S3 : 3-D generalized code (S3): inlet/outlet/symmetry faces,
 *                 S3 MGFM, inlet particle copies and outlet removal;
 *   PGDG_SEM_UX : 2-D underwater explosion (UX), two phases, FP64 form
 *                 only: HLL flux at interior face points
 *                 (UX:particle_flow.cpp:280-294), HLL against a ghost copy
 *                 of the inner state at domain-boundary face points, with
 *                 the y velocity reflected on PGDG_BC_WALL sides and kept
 *                 on PGDG_BC_OUTLET sides (UX:particle_flow.cpp:314-366;
 *                 the reference tags them 2 and -1, walls on y sides
 *                 only); the T2 MGFM; Calculate_limiter with a zero wall
 *                 velocity in the x columns of G_v, boundary neighbours
 *                 equal to the cell's own means with the momentum normal
 *                 to a domain edge set to 0, and the map
 *                 2 ds / (ds + 1) in every cell (UX:mpm.cpp:352-587);
 *                 MLS_P2G keeps an absent phase's density and momentum
 *                 (UX:mpm.cpp:299-334).                                   */
#define PGDG_SEM_T2 2
#define PGDG_SEM_S3 3
#define PGDG_SEM_UX 4

#define PGDG_MAX_PHASE 3

typedef struct pgdg_case {
  int    dim;            /* D: 2 or 3                                      */
  int    nphase;         /* F: 2 or 3 (reference FldNum)                   */
  int    order;          /* 1: P1 basis (1,x,y[,z]) of the reference;
                            2: zero-mean P2 basis (UX:glbfunc.h:210-222;
                            3D: 1,x,y,z,xy,xz,yz,x2,y2,z2) with 3-point
                            Gauss rules (no reference counterpart)        */
  int    rk_stages;      /* 2: the reference's two-stage Heun update;
                            3: SSP-RK3 (no reference counterpart)          */
  int    semantics;      /* PGDG_SEM_T2, PGDG_SEM_S3 or PGDG_SEM_UX        */
  int    n[3];           /* cells per axis (n[2] = 1 when dim == 2)         */
  int    bc[6];          /* tag of each domain side, PGDG_SIDE_* order      */
  double box_l[3], box_r[3];
  double p_sz;           /* particle spacing (reference p_sz)               */
  double p_g;            /* grid spacing used for kernel widths (p_g)       */
  double gamma[PGDG_MAX_PHASE], pinf[PGDG_MAX_PHASE], ini_rho[PGDG_MAX_PHASE];
  double inlet_rho, inlet_u[3], inlet_p;   /* inlet ghost state (S3)       */
  double p2g_cut;        /* P2G particle cut-off distance; the reference
                            uses 1.5*p_sz (S3:mpm.cpp:307), which covers a
                            whole cell only at one particle per cell       */
  double output_dt;      /* dt is truncated to land on multiples of this    */
  double max_v0;         /* initial max_v (reference 1e-7)                  */
  int    precision;      /* PGDG_PREC_FP64: the reference's FP64 arithmetic
                            and order (bit-identical to the CPU restatement);
                            PGDG_PREC_MIXED: FP32 DG quadrature and fluxes
                            (MGFM Newton, mass division and the RK update in
                            FP64), FP32 limiter slope ratios and positivity
                            probe, FP32 P2G kernel weights, FP32 particle-
                            shifting pair terms in cell-local coordinates;
                            the state, its sums and the topology stay FP64 */
  double cfl;            /* time-step factor; 0 selects the reference 0.3
                            (S3:particle.cpp:560)                          */
  int    ppc_min, ppc_max; /* per-cell particle-count bounds of split and
                            merge (0, 0: the reference's volume criteria
                            only).  A split (merge) is admitted only if no
                            cell it adds particles to (removes them from)
                            passes ppc_max (ppc_min), counting the cell's
                            fluid particles when the pass starts plus the
                            net changes of every earlier candidate of the
                            pass, in particle order; merge candidates also
                            need a cell holding more than ppc_min          */
} pgdg_case;

/* Host SoA views.  Arrays are sized by capacity `cap` (particles) or by the
 * cell count nc = n0*n1*n2 (cells, index (i*n1 + j)*n2 + k); vector fields
 * are stored component-major ([3][cap]); coef is [U*M][cap] with entry
 * u*M + m.  U = dim+2; M = dim+1 for order 1, 6 (2D) or 10 (3D) for
 * order 2.                                                                 */
typedef struct pgdg_particles {
  int64_t n, cap;
  double *x, *v, *tv, *rhov;                 /* [3][cap]                 */
  double *vol, *mass, *h, *P, *rho, *E, *c;  /* [cap]                    */
  double *coef;                              /* [U*M][cap]               */
  int32_t *phase, *mark_in;                  /* [cap]                    */
} pgdg_particles;

typedef struct pgdg_walls {                  /* boundary particles (phase 10) */
  int64_t n;
  double *x, *norm;                          /* [3][n]                   */
  double *vol;                               /* [n]                      */
} pgdg_walls;

typedef struct pgdg_cells {
  double *coef, *coef0, *dcoef;   /* [F][U][M][nc]                         */
  double *lim;                    /* [F][U][nc]                            */
  double *alph, *rho, *E, *c, *ac;/* [F][nc]                               */
  double *rhov;                   /* [F][3][nc]                            */
  int32_t *ph_state, *ph_state_b; /* [F][nc]                               */
  double *v, *a;                  /* [3][nc]  mixture velocity, acceleration */
  double *P;                      /* [nc]                                  */
  double *G_v;                    /* [9][nc]  velocity gradient, row-major  */
} pgdg_cells;

typedef struct pgdg_scalars {
  double  dt;             /* glbl_timestep                                 */
  double  max_v;          /* max |v| of the last Set_timestep              */
  double  integral_time;  /* time since the last output instant            */
  double  run_time;
  int64_t iterate_num;    /* steps taken                                   */
  int64_t n_split, n_inlet, n_outlet, n_merge;  /* event counts, last step */
} pgdg_scalars;

/* ======================================================================== *
 *  PG-DG time step: context and entry points                               *
 *  One context owns one case's state on the current GPU; every call is     *
 *  ordered on the context's stream.                                        *
 * ======================================================================== */
typedef struct pgdg_ctx pgdg_ctx;

/* Allocates the device state for `c` with room for particle_cap fluid
 * particles and n_wall wall particles (replaces the state ownership of
 * MPM::MPM / MPM::Initialize_case, S3:mpm.h:160-194).  Supported: order 1
 * or 2, rk_stages 2 or 3, dim 2 (PGDG_SEM_T2) or 3 (PGDG_SEM_S3), 2..3
 * phases; PGDG_SEM_UX with dim 2, 2 phases, PGDG_PREC_FP64, PGDG_BC_OUTLET
 * on the x sides and PGDG_BC_WALL or PGDG_BC_OUTLET on the y sides
 * (anything else is PGDG_ERR_ARG).                                        */
int pgdg_create(const pgdg_case* c, int64_t particle_cap, int64_t n_wall, pgdg_ctx** out);
int pgdg_destroy(pgdg_ctx* ctx);

/* Host <-> device copies of the whole state (the reference's in-memory
 * particle_fluid / particle_bound / cell arrays, S3:mpm.h:160-194).       */
int pgdg_upload_state(pgdg_ctx* ctx, const pgdg_particles* p, const pgdg_walls* w,
                      const pgdg_cells* c, const pgdg_scalars* s);
int pgdg_download_state(pgdg_ctx* ctx, pgdg_particles* p, pgdg_cells* c, pgdg_scalars* s);

/* nsteps iterations of the Solver::MPM_Solver loop body (S3:solver.cpp:33-51):
 * Set_timestep, output-instant truncation of dt, MPM_run_simulation.     */
int pgdg_step(pgdg_ctx* ctx, int nsteps);

/* Reference-named stages; each replaces the MPM member of the same name.  */
int pgdg_set_timestep(pgdg_ctx* ctx);              /* MPM::Set_timestep, S3:mpm.cpp:19          */
int pgdg_used_state_update(pgdg_ctx* ctx);         /* MPM::Used_state_update, S3:mpm.cpp:184    */
int pgdg_map_particles_to_cells(pgdg_ctx* ctx);    /* Reset_cell_list_info + Map_the_particle_to_tree,
                                                      S3:mpm.cpp:65-96                           */
int pgdg_split_particle_fast(pgdg_ctx* ctx);       /* MPM::Split_particle_fast, S3:mpm.cpp:1036 */
int pgdg_merge_particle_fast(pgdg_ctx* ctx);       /* MPM::Merge_particle_fast, S3:mpm.cpp:1536 */
int pgdg_refresh_cell_membership(pgdg_ctx* ctx);   /* Refresh_neighbor_info(4), S3:mpm.cpp:137-161 */
int pgdg_refresh_neighbor_info(pgdg_ctx* ctx, int stage); /* Refresh_neighbor_info(2),
                                                      S3:level_infor.cpp:153                     */
int pgdg_mls_p2g(pgdg_ctx* ctx);                   /* MPM::MLS_P2G, S3:mpm.cpp:274             */
int pgdg_calculate_limiter(pgdg_ctx* ctx);         /* MPM::Calculate_limiter, S3:mpm.cpp:392   */
int pgdg_grid_acceleration(pgdg_ctx* ctx);         /* MPM::Grid_acceleration, S3:mpm.cpp:747   */
int pgdg_update_grid(pgdg_ctx* ctx, int md);       /* MPM::Update_grid, S3:mpm.cpp:768         */
int pgdg_update_particle(pgdg_ctx* ctx, int md);   /* MPM::Update_particle, S3:mpm.cpp:839     */
int pgdg_run_simulation(pgdg_ctx* ctx, int64_t n); /* MPM::MPM_run_simulation, S3:mpm_ini.cpp:682 */

/* Stage by number (PGDG_FN_*), as used by the entries above.              */
#define PGDG_FN_SET_TIMESTEP    0
#define PGDG_FN_USED            1
#define PGDG_FN_BIN             2
#define PGDG_FN_SPLIT           3
#define PGDG_FN_MERGE           4
#define PGDG_FN_MEMBERSHIP      5
#define PGDG_FN_NEIGHBORS       6
#define PGDG_FN_P2G             7
#define PGDG_FN_LIMITER         8
#define PGDG_FN_RHS             9
#define PGDG_FN_UPDATE_GRID    10
#define PGDG_FN_UPDATE_PARTICLE 11
#define PGDG_FN_RUN            12
int pgdg_stage(pgdg_ctx* ctx, int fn, int arg);

/* Utilities: wait for the context's stream, its HIP stream handle, the
 * live fluid particle count, device bytes held, and per-stage HIP-event
 * timing (slot order: the profile slots of pgdg_p2_amd.step).             */
int pgdg_sync(pgdg_ctx* ctx);
int pgdg_get_stream(pgdg_ctx* ctx, void** stream);
int pgdg_particle_count(pgdg_ctx* ctx, int64_t* n);
int pgdg_memory_bytes(pgdg_ctx* ctx, int64_t* bytes);
int pgdg_profile_enable(pgdg_ctx* ctx, int on);
/* RK stages of pgdg_step / pgdg_run_simulation as HIP graphs (default on;
 * profiling runs the stages one launch at a time).                       */
int pgdg_graphs_enable(pgdg_ctx* ctx, int on);
/* Autotuner (default on): during the first step after it is enabled, the
 * workgroup sizes (64, 128 or 256) of the particle-shifting, RHS and
 * limiter kernels are chosen by timing each candidate in place (these
 * kernels overwrite their outputs, so the extra launches change nothing).
 * pgdg_tuning_read returns {shifting, RHS, limiter}; PGDG_ERR_STATE before
 * the tuning step has run.  pgdg_tuning_set fixes the three sizes (each 64,
 * 128 or 256, else PGDG_ERR_ARG) and switches the autotuner off.  The
 * step's results do not depend on the sizes.                              */
int pgdg_autotune_enable(pgdg_ctx* ctx, int on);
int pgdg_tuning_read(pgdg_ctx* ctx, int* blocks);
int pgdg_tuning_set(pgdg_ctx* ctx, const int* blocks);
/* Mixed form, 3-D: stage the particle-shifting candidates in LDS (default
 * on) or let every lane read its candidates from global memory.  Both give
 * the same sums bit for bit; the switch exists to compare them.           */
int pgdg_pst_staging_enable(pgdg_ctx* ctx, int on);
/* Mixed form, 3-D P2: the DG right-hand side as 16x16 matrix products on
 * the matrix cores (k_rhs_mfma) or one lane per cell on the vector ALUs
 * (default).  The two sum the same point terms in different orders.  No
 * counterpart in the reference (whose right-hand side is FP64,
 * S3:particle_flow.cpp:16-405).  Off by default: both kernels move about
 * 3.5 KB per cell through HBM (face fluxes, coefficients, the fused stage
 * update), and the matrix-core kernel keeps three wavefronts per SIMD to
 * hide it against the vector kernel's more (6.0 against 4.5 ms per stage
 * at 256x128x128 on MI355X).                                             */
int pgdg_rhs_mfma_enable(pgdg_ctx* ctx, int on);
/* Mixed form: how the step solves MGFM at its interface face points (the
 * reference calls MPM::MGFM once per face point inside the face integral,
 * S3:particle_flow.cpp:192-204).
 *   PGDG_IFACE_INLINE  (default) each point's Newton solve inside the face
 *                      kernel, from the reference's initial p*;
 *   PGDG_IFACE_BATCH   the stage's points collected as records and solved
 *                      as one lane-refilling batch, from the reference's
 *                      initial p* (the same solutions bit for bit);
 *   PGDG_IFACE_LEARNED as BATCH, from the trained network m's p*0 (no
 *                      reference counterpart); the Newton iteration runs to
 *                      the reference's tolerance, so the solutions differ
 *                      only within it.  The context reads m's parameters
 *                      when a stage runs: set the mode again after
 *                      retraining m (its normalisation is copied here).
 * PGDG_ERR_ARG for the FP64 form (which keeps the reference's solve).      */
#define PGDG_IFACE_INLINE  0
#define PGDG_IFACE_BATCH   1
#define PGDG_IFACE_LEARNED 2
int pgdg_iface_mgfm_mode(pgdg_ctx* ctx, int mode, pgdg_mlp* m);
/* The EOS pair of each record pair code a * PGDG_MAX_PHASE + b (left medium
 * phase a, right medium phase b): eos[PGDG_MAX_PHASE^2], host.             */
int pgdg_iface_pairs(pgdg_ctx* ctx, pgdg_eos_pair* eos);
/* The interface-point records of the last stage solved in a batch mode:
 * *n records; the first min(*n, cap) are copied (any memory) as compact SoA
 * in[6][k] (ul, ur, pl, pr, rhol, rhor), out[3][k] (rho*_L, p*, u*_L),
 * pair[k], iters[k] (each may be NULL) - the training set for
 * pgdg_mlp_train_pairs and the iteration counts of the step's solves.     */
int pgdg_iface_records(pgdg_ctx* ctx, double* in, double* out, int32_t* pair, int32_t* iters, int64_t cap,
                       int64_t* n);
int pgdg_profile_read(pgdg_ctx* ctx, double* ms, int64_t* launches, int nslot);

/* ======================================================================== *
 *  Output and diagnostics (the reference's Output_plt_file and             *
 *  Output_point)                                                           *
 * ======================================================================== */
/* Neighbour-list sizes of the fluid particles (particle order), as the
 * reference's particle->neighbor.size() after a step: the status-2 list of
 * the last RK stage (entries, walls included, within h * 2 of the stage-start
 * position; S3:level_infor.cpp:153-195).  Zero for every particle until
 * the first RK stage after pgdg_upload_state or pgdg_checkpoint_read (the
 * reference's lists are empty after its set-up).  n_neighbor: host [n].  */
int pgdg_neighbor_counts(pgdg_ctx* ctx, int32_t* n_neighbor);
/* pgdg_neighbor_counts without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * counts are staged in a fixed device array of 2^21 int32 (8 MiB) per GPU,
 * held until the process ends; PGDG_ERR_CAPACITY when the context holds
 * more than 2^21 fluid particles.  Calls that overlap on one GPU share the
 * array and race.                                                         */
int pgdg_neighbor_counts_noalloc(pgdg_ctx* ctx, int32_t* n_neighbor);
/* Cells in use (Used_Ele_cen, S3:mpm.cpp:184-263): host [nc], 1 or 0.      */
int pgdg_used_cells(pgdg_ctx* ctx, int32_t* used);
/* Fluid particles in each cell's list (the cell lists of the last binning
 * and split/merge edits; the reference sums them as particles per cell in
 * Output_point, T2:mpm_ini.cpp:599-654): host [nc].  PGDG_ERR_STATE before
 * the first binning after an upload.                                      */
int pgdg_cell_particle_counts(pgdg_ctx* ctx, int32_t* counts);

/* Checkpoint / restart.  The reference has none: it creates a restart
 * directory and never writes it (S3:Makefile:95, 115).  A checkpoint file
 * holds the particle, wall and cell arrays, the time-loop scalars, the
 * used-cell flags and the neighbour lists the next step's split and merge
 * read, so that pgdg_checkpoint_read into a context created for the same
 * case continues bit for bit as the written context would have.  The file
 * is in the host's byte order; read rejects a file of another case
 * (PGDG_ERR_ARG) or one that exceeds the context's capacity
 * (PGDG_ERR_CAPACITY).                                                   */
int pgdg_checkpoint_write(pgdg_ctx* ctx, const char* path);
int pgdg_checkpoint_read(pgdg_ctx* ctx, const char* path);

/* Host-only writers (no GPU), from host state:
 * pgdg_plt_write_particles  the particle file of MPM::Output_plt_file
 *   (S3:mpm_ini.cpp:832-946; T2:mpm_ini.cpp:662-805): Tecplot binary
 *   #!TDV112, variables x, y[, z], phase, vx, vy, vol, rho, P, |tv|, |v|,
 *   n_neighbor in block order.
 * pgdg_plt_write_cells      the cell file (S3:mpm_ini.cpp:946-978;
 *   T2:mpm_ini.cpp:807-838; UX:mpm_ini.cpp:766-795, with its extra "ca"
 *   column): ASCII, one line per cell in i, j[, k] order.
 * pgdg_point_diagnostics    MPM::Output_point's sums (S3:mpm_ini.cpp:774-827:
 *   phase-2 front positions and mass; T2:mpm_ini.cpp:599-654: particle and
 *   grid mass and energy, particles per cell in use), in the reference's
 *   summation order; pgdg_point_append writes its lines (pointa.txt and
 *   mass.txt for S3, point_p.txt and point_g.txt for T2) into dir,
 *   truncating the files when first != 0.                                   */
typedef struct pgdg_point_diag {
  double  time;
  double  dw, jet, ups, mass2;      /* S3                                  */
  int64_t n_fluid;
  double  p_mass[2], p_energy;      /* T2: particle sums                   */
  double  g_mass[2], g_energy;      /* T2: grid sums over the cells in use */
  double  p2g;                      /* fluid particles per cell in use     */
} pgdg_point_diag;
int pgdg_plt_write_particles(const char* path, const pgdg_case* c, const pgdg_particles* p,
                             const int32_t* n_neighbor, double run_time);
int pgdg_plt_write_cells(const char* path, const pgdg_case* c, const pgdg_cells* cl, double run_time);
int pgdg_point_diagnostics(const pgdg_case* c, const pgdg_particles* p, const pgdg_cells* cl,
                           const int32_t* used, double time, pgdg_point_diag* out);
int pgdg_point_append(const char* dir, const pgdg_case* c, const pgdg_point_diag* d, int first);

#ifdef __cplusplus
}
#endif
#endif /* PGDG_H */
