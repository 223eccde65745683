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

// One MGFM solve, split into begin / iterate / end so that a batch kernel
// can advance many faces one Newton iteration at a time; mgfm_solve runs the
// three in sequence.  S3 variant: S3:mpm.cpp:882-926.  T2 variant:
// T2:mpm.cpp:1373-1416.
struct MgfmState {
  double ul, ur, pl, pr, rhol, rhor;
  double tol, ps, rls, wl, wr, usl, f;
  int nn;
};

// `guess` < 0 selects the variant's own initial p*.
PGDG_HD void mgfm_begin(MgfmState& S, double ul, double ur, double pl, double pr, double rhol,
                        double rhor, int variant, double guess) {
  if (variant == MGFM_S3) {
    pl = pl >= 1e-10 ? pl : 1e-10;   // AMAX1(pl, 1e-10)
    pr = pr >= 1e-10 ? pr : 1e-10;
    rhol = rhol >= 1e-10 ? rhol : 1e-10;
    rhor = rhor >= 1e-10 ? rhor : 1e-10;
  } else {
    if (pl < 0) pl = 0;
    if (pr < 0) pr = 0;
  }
  S.ul = ul;
  S.ur = ur;
  S.pl = pl;
  S.pr = pr;
  S.rhol = rhol;
  S.rhor = rhor;
  double du = fabs(ul - ur);
  S.tol = 1e-8 >= 1e-3 * du ? 1e-8 : 1e-3 * du;
  if (guess >= 0)
    S.ps = guess;
  else if (variant == MGFM_S3)
    S.ps = pl >= pr ? pl : pr;
  else
    S.ps = (pl + pr) * 0.5;
  S.rls = S.wl = S.wr = S.usl = 0;
  S.f = 1.;
  S.nn = 0;
}

// One Newton iteration; true while another is needed.  TOL_REF: stop when
// |f| (evaluated before the update) is within the reference tolerance;
// TOL_REL (extension): stop when the update is below rtol * p*.
PGDG_HD bool mgfm_iterate(MgfmState& S, Eos ml, Eos mr, int variant, int tol_mode, double rtol,
                          int max_iter) {
  const double ps = S.ps;
  if (tol_mode == TOL_REF) {
    if (variant == MGFM_S3) {
      S.wl = shock_w(S.rhol, S.pl, ps, ml);
      S.wr = shock_w(S.rhor, S.pr, ps, mr);
      S.usl = S.ul - (ps - S.pl) / S.wl;
      S.rls = shock_rho(S.rhol, S.pl, ps, ml);
    } else {
      S.rls = shock_rho(S.rhol, S.pl, ps, ml);
      S.wl = shock_w(S.rhol, S.pl, ps, ml);
      S.wr = shock_w(S.rhor, S.pr, ps, mr);
    }
  } else {
    S.wl = shock_w(S.rhol, S.pl, ps, ml);
    S.wr = shock_w(S.rhor, S.pr, ps, mr);
  }
  S.f = S.ul - S.ur - (ps - S.pl) / S.wl - (ps - S.pr) / S.wr;
  double jac = -shock_dk(S.wl, S.pl, S.rhol, ps, ml) - shock_dk(S.wr, S.pr, S.rhor, ps, mr);
  double dp = -S.f / jac;
  if (fabs(dp) > 0.6 * ps) dp = copysign(0.6 * ps, dp);
  double pn = ps + dp;
  pn = pn >= 1e-8 ? pn : 1e-8;
  S.ps = pn;
  S.nn++;
  if (tol_mode == TOL_REF) return fabs(S.f) > S.tol && S.nn < max_iter;
  return fabs(pn - ps) > rtol * pn && S.nn < max_iter;
}

PGDG_HD MgfmOut mgfm_end(MgfmState& S, Eos ml, int variant, int tol_mode) {
  if (tol_mode == TOL_REF) {
    if (variant == MGFM_T2) S.usl = S.ul - (S.ps - S.pl) / S.wl;
  } else {
    S.wl = shock_w(S.rhol, S.pl, S.ps, ml);
    S.rls = shock_rho(S.rhol, S.pl, S.ps, ml);
    S.usl = S.ul - (S.ps - S.pl) / S.wl;
  }
  MgfmOut o;
  o.rho_l = S.rls;
  o.pstar = S.ps;
  o.u_l = S.usl;
  o.iters = S.nn;
  return o;
}

PGDG_HD MgfmOut mgfm_solve(double ul, double ur, double pl, double pr, double rhol, double rhor,
                           Eos ml, Eos mr, int variant, int tol_mode, double rtol, int max_iter,
                           double guess) {
  MgfmState S;
  mgfm_begin(S, ul, ur, pl, pr, rhol, rhor, variant, guess);
  while (mgfm_iterate(S, ml, mr, variant, tol_mode, rtol, max_iter)) {
  }
  return mgfm_end(S, ml, variant, tol_mode);
}

}  // namespace pgdg
