// riemann.hpp - stiffened-gas EOS and the MGFM two-shock Newton solve,
// written once for host and device (__host__ __device__).
//
// Every expression keeps the evaluation order of the reference text
// (S3:mpm.cpp:859-926, S3:glbfunc.h:314-337, T2:mpm.cpp:1373-1416) so that a
// build with -ffp-contract=off reproduces the FP64 results bit for bit.
// pow(W,2) of S3:glbfunc.h:334 is written W*W (one rounding, the correctly
// rounded square).
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define PGDG_HD __host__ __device__ inline
#else
#define PGDG_HD inline
#endif

namespace pgdg {

struct Eos {
  double gamma;
  double pinf;
};

// S3:mpm.cpp:859-865  p = (rho e)(gamma-1) - gamma P_inf
PGDG_HD double eos_p(double rhoe, Eos m) { return rhoe * (m.gamma - 1) - m.gamma * m.pinf; }

// S3:mpm.cpp:870-876  rho e = (p + gamma P_inf)/(gamma-1)
PGDG_HD double eos_rhoe(double p, Eos m) { return (p + m.gamma * m.pinf) / (m.gamma - 1); }

// S3:glbfunc.h:314-321  post-shock density across the wave at pressure p*
PGDG_HD double shock_rho(double rho, double p, double ps, Eos m) {
  double num = (m.gamma + 1) * ps + (m.gamma - 1) * p + 2 * m.gamma * m.pinf;
  double den = (m.gamma - 1) * ps + (m.gamma + 1) * p + 2 * m.gamma * m.pinf;
  return rho * num / den;
}

// S3:glbfunc.h:323-328  mass flux W of the shock
PGDG_HD double shock_w(double rho, double p, double ps, Eos m) {
  return sqrt(0.5 * rho * ((m.gamma + 1.0) * ps + (m.gamma - 1.0) * p + 2.0 * m.gamma * m.pinf));
}

// S3:glbfunc.h:331-337  d/dp* of (p*-p)/W
PGDG_HD double shock_dk(double w, double p, double rho, double ps, Eos m) {
  return 1. / w - (ps - p) / (w * w) * rho * (m.gamma + 1) / 4. / w;
}

struct MgfmOut {
  double rho_l;   // rho*_L
  double pstar;   // p*
  double u_l;     // u*_L
  int iters;
};

// Tolerance / variant selectors (mirror PGDG_MGFM_* and PGDG_TOL_* of pgdg.h).
enum { MGFM_S3 = 0, MGFM_T2 = 1 };
enum { TOL_REF = 0, TOL_REL = 1 };

// One MGFM solve.  `guess` < 0 selects the variant's own initial p*.
// S3 variant: S3:mpm.cpp:882-926.  T2 variant: T2:mpm.cpp:1373-1416.
PGDG_HD MgfmOut mgfm_solve(double ul, double ur, double pl, double pr, double rhol, double rhor,
                           Eos ml, Eos mr, int variant, int tol_mode, double rtol, int max_iter,
                           double guess) {
  MgfmOut o;
  if (variant == MGFM_S3) {
    pl = pl >= 1e-10 ? pl : 1e-10;   // AMAX1(pl, 1e-10)
    pr = pr >= 1e-10 ? pr : 1e-10;
    rhol = rhol >= 1e-10 ? rhol : 1e-10;
    rhor = rhor >= 1e-10 ? rhor : 1e-10;
  } else {
    if (pl < 0) pl = 0;
    if (pr < 0) pr = 0;
  }
  double du = fabs(ul - ur);
  double tol = 1e-8 >= 1e-3 * du ? 1e-8 : 1e-3 * du;
  double ps;
  if (guess >= 0)
    ps = guess;
  else if (variant == MGFM_S3)
    ps = pl >= pr ? pl : pr;
  else
    ps = (pl + pr) * 0.5;

  double rls = 0, wl = 0, wr = 0, usl = 0, f = 1.;
  int nn = 0;
  if (tol_mode == TOL_REF) {
    do {
      if (variant == MGFM_S3) {
        wl = shock_w(rhol, pl, ps, ml);
        wr = shock_w(rhor, pr, ps, mr);
        usl = ul - (ps - pl) / wl;
        rls = shock_rho(rhol, pl, ps, ml);
      } else {
        rls = shock_rho(rhol, pl, ps, ml);
        wl = shock_w(rhol, pl, ps, ml);
        wr = shock_w(rhor, pr, ps, mr);
      }
      f = ul - ur - (ps - pl) / wl - (ps - pr) / wr;
      double jac = -shock_dk(wl, pl, rhol, ps, ml) - shock_dk(wr, pr, rhor, ps, mr);
      double dp = -f / jac;
      if (fabs(dp) > 0.6 * ps) dp = copysign(0.6 * ps, dp);
      ps += dp;
      ps = ps >= 1e-8 ? ps : 1e-8;
      nn++;
    } while (fabs(f) > tol && nn < max_iter);
    if (variant == MGFM_T2) usl = ul - (ps - pl) / wl;
  } else {
    // Relative-tolerance extension: iterate the same Newton map until the
    // update is below rtol*p*, then evaluate rho*_L and u*_L at the result.
    double ps_old;
    do {
      wl = shock_w(rhol, pl, ps, ml);
      wr = shock_w(rhor, pr, ps, mr);
      f = ul - ur - (ps - pl) / wl - (ps - pr) / wr;
      double jac = -shock_dk(wl, pl, rhol, ps, ml) - shock_dk(wr, pr, rhor, ps, mr);
      double dp = -f / jac;
      if (fabs(dp) > 0.6 * ps) dp = copysign(0.6 * ps, dp);
      ps_old = ps;
      ps += dp;
      ps = ps >= 1e-8 ? ps : 1e-8;
      nn++;
    } while (fabs(ps - ps_old) > rtol * ps && nn < max_iter);
    wl = shock_w(rhol, pl, ps, ml);
    rls = shock_rho(rhol, pl, ps, ml);
    usl = ul - (ps - pl) / wl;
  }
  o.rho_l = rls;
  o.pstar = ps;
  o.u_l = usl;
  o.iters = nn;
  return o;
}

}  // namespace pgdg
