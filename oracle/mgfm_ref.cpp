// mgfm_ref.cpp - CPU FP64 restatement of the reference MGFM interface solver.
//
// TEST INFRASTRUCTURE ONLY.  This file is part of the oracle: it is linked
// into liboracle.so, which only the tests, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load, and only as the checker.
//
// Follows, line by line:
//   S3:mpm.cpp:882-926     MPM::MGFM (3D cases)
//   T2:mpm.cpp:1373-1416   MPM::MGFM (2D triple point)
//   S3:glbfunc.h:314-337   RhoLs_cal, Wl_cal, Jaco_cal
// pow(W,2) (S3:glbfunc.h:334) is written W*W.  Pinned by the closed-form
// checks in tests/test_oracle_mgfm.py (residual, Jacobian vs finite
// difference, equal-state and symmetric cases, two-shock Toro reference).
#include <cmath>
#include <cstdint>

namespace oracle {

static inline double AMAX1(double a, double b) { return a >= b ? a : b; }

static inline double RhoLs_cal(double rhol, double pl, double pstar, double gam, double Pp0) {
  double valN = (gam + 1) * pstar + (gam - 1) * pl + 2 * gam * Pp0;
  double valD = (gam - 1) * pstar + (gam + 1) * pl + 2 * gam * Pp0;
  return rhol * valN / valD;
}

static inline double Wl_cal(double rhol, double pl, double pstar, double gam, double Pp0) {
  return std::sqrt(0.5 * rhol * ((gam + 1.0) * pstar + (gam - 1.0) * pl + 2.0 * gam * Pp0));
}

static inline double Jaco_cal(double W, double pl, double rhol, double pstar, double gam,
                              double Pp0) {
  return 1. / W - (pstar - pl) / (W * W) * rhol * (gam + 1) / 4. / W;
}

// S3:mpm.cpp:882-926.  gamma/Pp0 of the two media are passed explicitly.
void MGFM_S3(double ul, double ur, double pl, double pr, double rhol, double rhor, double gs,
             double ps0, double gs1, double ps1, int max_iter, double guess, double out[3],
             int* iters) {
  pl = AMAX1(pl, 1e-10);
  pr = AMAX1(pr, 1e-10);
  rhol = AMAX1(rhol, 1e-10);
  rhor = AMAX1(rhor, 1e-10);
  double tol = AMAX1(1e-8, 1e-3 * std::abs(ul - ur));
  double pstar = AMAX1(pl, pr);
  if (guess >= 0) pstar = guess;
  double f1 = 1.;
  double rhols = 0, usl = 0, Wl = 0, Wr = 0;
  int nn = 0;
  do {
    Wl = Wl_cal(rhol, pl, pstar, gs, ps0);
    Wr = Wl_cal(rhor, pr, pstar, gs1, ps1);
    usl = ul - (pstar - pl) / Wl;
    rhols = RhoLs_cal(rhol, pl, pstar, gs, ps0);
    f1 = ul - ur - (pstar - pl) / Wl - (pstar - pr) / Wr;
    double Jacobi = -Jaco_cal(Wl, pl, rhol, pstar, gs, ps0) - Jaco_cal(Wr, pr, rhor, pstar, gs1, ps1);
    double dp = -f1 / Jacobi;
    if (std::abs(dp) > 0.6 * pstar) dp = std::copysign(0.6 * pstar, dp);
    pstar += dp;
    pstar = AMAX1(pstar, 1e-8);
    nn++;
  } while (std::abs(f1) > tol && nn < max_iter);
  out[0] = rhols;
  out[1] = pstar;
  out[2] = usl;
  if (iters) *iters = nn;
}

// T2:mpm.cpp:1373-1416.
void MGFM_T2(double ul, double ur, double pl, double pr, double rhol, double rhor, double gs,
             double ps0, double gs1, double ps1, int max_iter, double guess, double out[3],
             int* iters) {
  if (pl < 0) pl = 0;
  if (pr < 0) pr = 0;
  double tol = AMAX1(1e-8, 1e-3 * std::abs(ul - ur));
  double pstar = (pl + pr) * 0.5;
  if (guess >= 0) pstar = guess;
  double f1 = 1.;
  double rhols = 0, Wl = 0, Wr = 0;
  int nn = 0;
  do {
    rhols = RhoLs_cal(rhol, pl, pstar, gs, ps0);
    Wl = Wl_cal(rhol, pl, pstar, gs, ps0);
    Wr = Wl_cal(rhor, pr, pstar, gs1, ps1);
    f1 = ul - ur - (pstar - pl) / Wl - (pstar - pr) / Wr;
    double Jacobi = -Jaco_cal(Wl, pl, rhol, pstar, gs, ps0) - Jaco_cal(Wr, pr, rhor, pstar, gs1, ps1);
    double dp = -f1 / Jacobi;
    if (std::abs(dp) > 0.6 * pstar) dp = std::copysign(0.6 * pstar, dp);
    pstar += dp;
    pstar = AMAX1(pstar, 1e-8);
    nn++;
  } while (std::abs(f1) > tol && nn < max_iter);
  double usl = ul - (pstar - pl) / Wl;
  out[0] = rhols;
  out[1] = pstar;
  out[2] = usl;
  if (iters) *iters = nn;
}

// Relative-tolerance extension (no reference counterpart): the same Newton
// map with the S3 (variant 0) or T2 (variant 1) clamps and initial guess,
// stopping when |dp| <= rtol * p*; rho*_L and u*_L are evaluated at the
// returned p*.
void MGFM_rel(int variant, double ul, double ur, double pl, double pr, double rhol, double rhor,
              double gs, double ps0, double gs1, double ps1, double rtol, int max_iter,
              double guess, double out[3], int* iters) {
  double pstar;
  if (variant == 0) {
    pl = AMAX1(pl, 1e-10);
    pr = AMAX1(pr, 1e-10);
    rhol = AMAX1(rhol, 1e-10);
    rhor = AMAX1(rhor, 1e-10);
    pstar = AMAX1(pl, pr);
  } else {
    if (pl < 0) pl = 0;
    if (pr < 0) pr = 0;
    pstar = (pl + pr) * 0.5;
  }
  if (guess >= 0) pstar = guess;
  int nn = 0;
  double old;
  do {
    double Wl = Wl_cal(rhol, pl, pstar, gs, ps0);
    double Wr = Wl_cal(rhor, pr, pstar, gs1, ps1);
    double f1 = ul - ur - (pstar - pl) / Wl - (pstar - pr) / Wr;
    double Jacobi = -Jaco_cal(Wl, pl, rhol, pstar, gs, ps0) - Jaco_cal(Wr, pr, rhor, pstar, gs1, ps1);
    double dp = -f1 / Jacobi;
    if (std::abs(dp) > 0.6 * pstar) dp = std::copysign(0.6 * pstar, dp);
    old = pstar;
    pstar += dp;
    pstar = AMAX1(pstar, 1e-8);
    nn++;
  } while (std::abs(pstar - old) > rtol * pstar && nn < max_iter);
  double Wl = Wl_cal(rhol, pl, pstar, gs, ps0);
  out[0] = RhoLs_cal(rhol, pl, pstar, gs, ps0);
  out[1] = pstar;
  out[2] = ul - (pstar - pl) / Wl;
  if (iters) *iters = nn;
}

}  // namespace oracle

extern "C" {

// Batched driver over SoA [6][n] inputs (ul, ur, pl, pr, rhol, rhor); the
// oracle counterpart of pgdg_mgfm_batch.  mode: 0 reference tolerance,
// 1 relative tolerance.  guess may be NULL.
int oracle_mgfm_batch(const double* in, double* out, int32_t* iters, int64_t n, double gl,
                      double pinfl, double gr, double pinfr, int variant, int mode, double rtol,
                      int max_iter, const double* guess) {
  for (int64_t i = 0; i < n; i++) {
    double o[3];
    int it = 0;
    double g = guess ? guess[i] : -1.0;
    if (mode == 0) {
      if (variant == 0)
        oracle::MGFM_S3(in[i], in[n + i], in[2 * n + i], in[3 * n + i], in[4 * n + i],
                        in[5 * n + i], gl, pinfl, gr, pinfr, max_iter, g, o, &it);
      else
        oracle::MGFM_T2(in[i], in[n + i], in[2 * n + i], in[3 * n + i], in[4 * n + i],
                        in[5 * n + i], gl, pinfl, gr, pinfr, max_iter, g, o, &it);
    } else {
      oracle::MGFM_rel(variant, in[i], in[n + i], in[2 * n + i], in[3 * n + i], in[4 * n + i],
                       in[5 * n + i], gl, pinfl, gr, pinfr, rtol, max_iter, g, o, &it);
    }
    out[i] = o[0];
    out[n + i] = o[1];
    out[2 * n + i] = o[2];
    if (iters) iters[i] = it;
  }
  return 0;
}

}  // extern "C"
