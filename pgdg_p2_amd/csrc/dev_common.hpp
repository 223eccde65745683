// dev_common.hpp - device-side views, geometry constants and scalar math of
// the PG-DG step kernels.
//
// Data layout in HBM (all FP64 unless noted, SoA so a wavefront touching 64
// consecutive particles or cells reads 512 contiguous bytes per field):
//   particles  x[3][cap], v[3][cap], tv[3][cap], rhov[3][cap],
//              vol, mass, h, P, rho, E, c [cap], coef[U*M][cap],
//              phase, mark_in (int32) [cap]
//   cells      coef/coef0/dcoef[F][U][M][nc], lim[F][U][nc],
//              alph/rho/E/c/ac[F][nc], rhov[F][3][nc], ph_state(_b)[F][nc],
//              v/a[3][nc], P[nc], G[9][nc]; cell index (i*n1 + j)*n2 + k
//   bins       CSR over cells: bin_off[nc+1], bin_ent[] (particle refs:
//              >= 0 fluid particle, <= -2 wall -(w+2), -1 removed)
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "pgdg.h"
#include "riemann.hpp"

namespace pgdg {

constexpr double kCellRatio = 3.2;
constexpr double kCutOff = 2.;
constexpr double kPiRef = 3.1415926;  // S3:glbparam.h:36

__host__ __device__ inline bool ref_is_wall(int r) { return r <= -2; }
__host__ __device__ inline int ref_wall(int r) { return -r - 2; }

// Geometry and constants of one case, passed by value to every kernel.
struct Geo {
  int N[3];
  int nc;
  int F;
  int sem;
  int bc[6];
  int radius;  // status-2 neighbour search radius in cells
  double ps[3];
  double box_l[3], box_r[3];
  double p_sz, p_g, p2g_cut;
  double gamma[PGDG_MAX_PHASE], pinf[PGDG_MAX_PHASE], ini_rho[PGDG_MAX_PHASE];
  double inlet_rho, inlet_u[3], inlet_p;
  double Gp[3], Gw[3];  // Gauss rule per axis: 2 points (order 1), 3 (order 2)
  double q12[3];        // ps[a]^2 / 12, the constants of the P2 square modes
  double cfl;           // time-step factor
  double probe_c[4], probe_s[4], sqrt2;
  double vol_max, vol_min, dss0;
  int ord;              // 1: P1 basis, 2: P2 basis
  int prec;             // PGDG_PREC_FP64 or PGDG_PREC_MIXED
  int ppc_min, ppc_max; // per-cell particle-count bounds of split/merge (0: off)
  int mfma_rhs;         // mixed form, 3-D P2: RHS on the matrix cores (k_rhs_mfma), 0 by default
};

// Modes per cell and Gauss points per axis.
template <int D, int ORD>
struct Basis {
  static constexpr int M = ORD == 1 ? D + 1 : (D == 2 ? 6 : 10);
  static constexpr int NQ = ORD + 1;
};

// Modes 1..M-1 of the cell polynomial at offset r from the cell centre.
// Order 2 follows the zero-mean P2 form of UX:glbfunc.h:210-222 (2D: x, y,
// xy, x^2 - dx^2/12, y^2 - dy^2/12); 3D: x, y, z, xy, xz, yz, x^2 - dx^2/12,
// y^2 - dy^2/12, z^2 - dz^2/12.
template <int D, int ORD, typename T>
__device__ inline void basis_tail(const T* r, const double* q12, T* phi) {
#pragma unroll
  for (int a = 0; a < D; a++) phi[a] = r[a];
  if constexpr (ORD == 2) {
    if constexpr (D == 2) {
      phi[2] = r[0] * r[1];
      phi[3] = r[0] * r[0] - (T)q12[0];
      phi[4] = r[1] * r[1] - (T)q12[1];
    } else {
      phi[3] = r[0] * r[1];
      phi[4] = r[0] * r[2];
      phi[5] = r[1] * r[2];
      phi[6] = r[0] * r[0] - (T)q12[0];
      phi[7] = r[1] * r[1] - (T)q12[1];
      phi[8] = r[2] * r[2] - (T)q12[2];
    }
  }
}

// The same modes without their constants, psi_m(r) = phi_m(r) - phi_m(0):
// a particle's copy is a polynomial about the particle's own position.
template <int D, int ORD>
__device__ inline void incr_tail(const double* r, double* psi) {
  const double z[3] = {0, 0, 0};
  basis_tail<D, ORD>(r, z, psi);
}

// Gradient at offset r of a polynomial whose tail coefficients are c[1..].
template <int D, int ORD>
__device__ inline void grad_at(const double* c, const double* r, double* g) {
  if constexpr (ORD == 1) {
#pragma unroll
    for (int a = 0; a < D; a++) g[a] = c[1 + a];
  } else if constexpr (D == 2) {
    g[0] = c[1] + c[3] * r[1] + 2. * c[4] * r[0];
    g[1] = c[2] + c[3] * r[0] + 2. * c[5] * r[1];
  } else {
    g[0] = c[1] + c[4] * r[1] + c[5] * r[2] + 2. * c[7] * r[0];
    g[1] = c[2] + c[4] * r[0] + c[6] * r[2] + 2. * c[8] * r[1];
    g[2] = c[3] + c[5] * r[0] + c[6] * r[1] + 2. * c[9] * r[2];
  }
}

struct PtlDev {
  int64_t cap;
  double *x, *v, *tv, *rhov;                // [3][cap]
  double *vol, *mass, *h, *P, *rho, *E, *c; // [cap]
  double *coef;                             // [U*M][cap]
  int32_t *phase, *mark_in;                 // [cap]
};

struct CellDev {
  double *coef, *coef0, *dcoef, *lim, *alph, *rho, *E, *c, *ac, *rhov, *v, *a, *P, *G;
  int32_t *ph_state, *ph_state_b;
  // mixed form only: interior-face point fluxes, the faces left to the
  // general face kernel and their count; fkind[ax][c] is 1 where the face
  // (c, +ax) holds one flux per point for both sides, unweighted (the side-0
  // slots; each side multiplies by the other cell's volume fraction), 0 where
  // each side has its own, weighted, slots (the listed faces)
  float* fflux;
  int32_t *flist, *fcount;
  int8_t* fkind;
  // mixed form, 3-D: domain-boundary face point fluxes [F][6][NQF][U][plane]
  float* bflux;
  // mixed form, interface points solved as a batch: the MGFM inputs of every
  // (listed face, side, phase pair with different phases, face point)
  // ([6][mg_cap]: ul, ur, pl, pr, rhol, rhor), their pair codes
  // (ss * PGDG_MAX_PHASE + s1), solutions ([3][mg_cap]) and Newton
  // iterations; the record count, the largest count any stage reached, and
  // each listed face's first record (-1: the face kernel solves that face's
  // points itself)
  double *mg_in, *mg_out;
  int32_t *mg_pair, *mg_it, *mg_count, *mg_peak, *mg_base;
  int64_t mg_cap;
};

struct WallDev {
  int64_t n;
  double *x, *norm, *vol;  // x,norm [3][n]
};

// Device-resident scalars of the time loop (kept on the device so stage
// launches can be replayed without host round trips).
struct DevScalars {
  double dt;
  double max_v;
  double integral_time;
  double run_time;
  long long iterate_num;
  unsigned long long red_dt;    // min-reduction slot (IEEE bits of a positive double)
  unsigned long long red_maxv;  // max-reduction slot
  long long np_dev;             // live fluid particles (read by the stage kernels,
  long long nent_dev;           //  bin entries          so a captured stage replays)
};

// get_cell_id (S3:glbfunc.h:161-173) of coord - box_l
template <int D>
__device__ inline int cell_of(const Geo& g, const double* x) {
  int p[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) {
    int q = int(floor((x[a] - g.box_l[a]) / g.ps[a]));
    q = q <= g.N[a] - 1 ? q : g.N[a] - 1;
    p[a] = q >= 0 ? q : 0;
  }
  return (p[0] * g.N[1] + p[1]) * g.N[2] + p[2];
}

__device__ inline void cell_ijk(const Geo& g, int c, int& i, int& j, int& k) {
  k = c % g.N[2];
  int t = c / g.N[2];
  j = t % g.N[1];
  i = t / g.N[1];
}

// Correctly rounded x^(1./3.) (pow(x, 1./3.) of the reference rounded to
// nearest): Newton for the cube root, an fma-exact residual correction and
// the first-order factor (1 - eps ln x) for 1./3. = 1/3 - eps.
__host__ __device__ inline double root13(double x) {
  if (!(x > 0) || isinf(x)) {
    if (x == 0) return 0.0;
    if (isinf(x) && x > 0) return x;
    return NAN;
  }
  int e;
  double m = frexp(x, &e);
  int q = e / 3;
  int r = e - 3 * q;
  if (r < 0) {
    r += 3;
    q -= 1;
  }
  double y0 = ldexp(m, r);
  double t = 0.6 + 0.4 * y0 - 0.02 * y0 * y0;
  for (int it = 0; it < 6; it++) t = t - (t * t * t - y0) / (3.0 * t * t);
  double p = t * t;
  double pe = fma(t, t, -p);
  double c3 = p * t;
  double ce = fma(p, t, -c3);
  double res = ((y0 - c3) - ce) - pe * t;
  double delta = res / (3.0 * p);
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double lnm = 2.0 * s * (1.0 + s2 * (1.0 / 3.0 + s2 * (0.2 + s2 * (1.0 / 7.0 + s2 * (1.0 / 9.0)))));
  double lnx = (double)e * 0.6931471805599453 + lnm;
  const double eps = 1.850371707708594e-17;
  double corr = delta - eps * lnx * t;
  return ldexp(t + corr, q);
}

template <int D>
__host__ __device__ inline double root_d(double x) {
  if (D == 2) return sqrt(x);
  return root13(x);
}

// Wendland kernel W(r, h) and dW/dr / r (S3:particle.cpp:131-149, 297-315)
template <int D>
__device__ inline double kern_C(double h) {
  if (D == 2) return 1.0 / kPiRef / (h * h) * 7. / 4.;
  return 1.0 / kPiRef / ((h * h) * h) * 21. / 16.;
}
template <int D>
__device__ inline double kern_W(double dist, double h) {
  double C = kern_C<D>(h);
  double s = dist / h;
  if (s < 2) {
    double q = 1. - 0.5 * s;
    return C * (((q * q) * q) * q) * (1. + 2. * s);
  }
  return 0.0;
}
template <int D>
__device__ inline double kern_dW(double dist, double h) {
  double C = kern_C<D>(h);
  double s = dist / h;
  if (s < 2) {
    double q = 1. - 0.5 * s;
    return -5. * C * ((q * q) * q) * s / h / dist;
  }
  return 0.0;
}

__device__ inline double amax1(double a, double b) { return a >= b ? a : b; }
__device__ inline double amin1(double a, double b) { return a <= b ? a : b; }
__device__ inline float amax1(float a, float b) { return a >= b ? a : b; }
__device__ inline float amin1(float a, float b) { return a <= b ? a : b; }

}  // namespace pgdg
