/*
 * pgdg.h - C ABI of the MI355X-native PG-DG time step.
 *
 * The reference (gaotr3306/PGDG_p2) has no plugin or FFI boundary: its hot
 * path is the set of C++ member functions of `class MPM` and
 * `class Particle_base` (S3 = 3D_generalized/PGDG_3D_spheric,
 * T2 = 2D_cases/Triple_point_2phase).  Each entry point below replaces one of
 * those member functions; the comment on each names the reference routine
 * (file:line) it stands in for.  The C++ host class `pgdg::MPM`
 * (pgdg_p2_amd/csrc/mpm_host.hpp) keeps the reference method names and
 * forwards to these functions.
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

#ifdef __cplusplus
}
#endif
#endif /* PGDG_H */
