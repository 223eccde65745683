// step_cells.hip - per-cell kernels of the PG-DG step: MLS P2G, limiter,
// DG right-hand side (volume + face quadrature with HLLC, MGFM ghost states
// and boundary fluxes) and the RK stage update.
//
// One lane per cell, FP64.  Each kernel keeps the operation order of the
// reference routine it replaces so that a -ffp-contract=off build matches
// the FP64 oracle bit for bit:
//   k_p2g          MPM::MLS_P2G            S3:mpm.cpp:274-382
//   k_limiter      MPM::Calculate_limiter  S3:mpm.cpp:392-738, T2:mpm.cpp:887-1225
//   k_rhs          Particle_base::Grid_acceleration S3:particle_flow.cpp:16-405
//   k_update_grid  MPM::Update_grid        S3:mpm.cpp:768-830, T2:mpm.cpp:1254-1320
// Face fluxes are accumulated only along the face normal: the reference
// forms the full U x D flux and contracts it with an axis-aligned normal, so
// the tangential columns contribute exact zeros.
#include "cell_flux.hpp"
#include "dev_common.hpp"
#include "hip_util.hpp"
#include "step_kernels.hpp"

namespace pgdg {

#define CIDX(f, u, m) ((((size_t)(f) * U + (u)) * M + (m)) * nc + c)

template <int D>
__device__ inline void cell_center(const Geo& g, int i, int j, int k, double* x) {
  x[0] = g.ps[0] * (i + 0.5);
  x[1] = g.ps[1] * (j + 0.5);
  x[2] = D == 3 ? g.ps[2] * (k + 0.5) : 0.0;
}

// ---------------------------------------------------------------------------
// MLS P2G (S3:mpm.cpp:274-382)
// ---------------------------------------------------------------------------
template <int D>
__global__ void k_p2g_gather(PtlDev P, const int* __restrict__ bin_ent, const int* __restrict__ bin_total,
                             P2gBuf pb) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pb.ne || e >= *bin_total) return;
  const int64_t cap = P.cap;
  double* l = pb.light + e * kP2gLight;
  const int p = bin_ent[e];
  if (p < 0) {
    *reinterpret_cast<double2*>(l + 4) = make_double2(-1.0, 0.0);
    return;
  }
  double r[kP2gLight] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) r[a] = P.x[a * cap + p];
  r[3] = P.vol[p];
  r[4] = (double)(P.phase[p] - 1);
#pragma unroll
  for (int q = 0; q < kP2gLight; q += 2) *reinterpret_cast<double2*>(l + q) = make_double2(r[q], r[q + 1]);
}

// Kernel weight of a particle at squared distance sq from the cell centre
// (0 beyond the cut-off).  MIX (mixed form): the cut-off test compares
// squared distances in FP64 and the weight is evaluated in FP32; the FP64
// form takes the square root and the FP64 weight, as the reference.
template <int D, bool MIX>
__device__ inline double p2g_weight(const Geo& g, double sq, double vol) {
  const double hi = g.p_g * 1 * 3.2 / kCutOff;
  if constexpr (MIX) {
    if (!(sq < g.p2g_cut * g.p2g_cut)) return 0.0;
    const float sk = __builtin_amdgcn_sqrtf((float)sq) * (float)(1.0 / hi);
    const float q = 1.f - 0.5f * sk;
    return (double)((float)kern_C<D>(hi) * (((q * q) * q) * q) * (1.f + 2.f * sk) * (float)vol);
  } else {
    double dist = sqrt(sq);
    if (!(dist < g.p2g_cut)) return 0.0;
    return kern_W<D>(dist, hi) * vol;
  }
}

// The end of MLS_P2G for cell c: phase states and fractions, the appearing
// phases' coefficients from the weighted averages (pav [phase][U], cf
// [phase][U][B]; read only for appearing phases), absent phases cleared,
// velocity and pressure.  APPEAR false: the caller knows no phase appears.
// Order 2: qav [phase][U] are the weighted averages of the particles' own
// values; an appearing mean with non-positive density or pressure below
// -0.05 Pp0 (the bound of Update_state_onestep's fallback,
// S3:particle_flow.cpp:460-469) is replaced by them, with zero slopes.
template <int D, int ORD, bool APPEAR>
__device__ inline void p2g_finish(const Geo& g, const CellDev& C, int c, const double* mass_t,
                                  const double* pav, const double* cf, const double* qav = nullptr) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M, B = M - 1;
  const int nc = g.nc;
  double mass_total = 0;
#pragma unroll
  for (int s = 0; s < PGDG_MAX_PHASE; s++)
    if (s < g.F) mass_total += mass_t[s];
  double cv[3] = {C.v[c], C.v[nc + c], C.v[2 * nc + c]};
  double cP = C.P[c];
#pragma unroll
  for (int ss = 0; ss < PGDG_MAX_PHASE; ss++) {
    if (ss >= g.F) break;
    int st = mass_t[ss] > 0 ? 1 : 0;
    C.ph_state[ss * nc + c] = st;
    C.alph[ss * nc + c] = mass_t[ss] / (mass_total + 1e-20);
    int stb = C.ph_state_b[ss * nc + c];
    if (APPEAR && st == 1 && stb == 0) {
      const double* pv = pav + ss * U;
      const double* cs = cf + ss * U * B;
      double pa[U];
#pragma unroll
      for (int u = 0; u < U; u++) pa[u] = pv[u] / (mass_t[ss] + 1e-20);
      bool keep_tail = true;
      if constexpr (ORD == 2) {  // centre value -> mean of the zero-mean basis
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int a = 0; a < D; a++) pa[u] += cs[u * B + (D == 2 ? 3 : 6) + a] / (mass_t[ss] + 1e-20) * g.q12[a];
        double sq = 0;
#pragma unroll
        for (int a = 0; a < D; a++) sq += pa[1 + a] * pa[1 + a];
        const double e_m = pa[U - 1] - 0.5 * sq / pa[0];
        const double P_m = eos_p(e_m, Eos{g.gamma[ss], g.pinf[ss]});
        if (!(pa[0] > 0) || P_m < -0.05 * g.pinf[ss]) {
          keep_tail = false;
#pragma unroll
          for (int u = 0; u < U; u++) pa[u] = qav[ss * U + u] / (mass_t[ss] + 1e-20);
        }
      }
      double rho = pa[0];
      C.coef[CIDX(ss, 0, 0)] = rho;
      C.rho[ss * nc + c] = rho;
      double rv[3] = {0, 0, 0};
#pragma unroll
      for (int a = 0; a < D; a++) {
        rv[a] = pa[1 + a];
        C.coef[CIDX(ss, 1 + a, 0)] = rv[a];
        C.rhov[(ss * 3 + a) * nc + c] = rv[a];
      }
      if (D == 2) rv[2] = C.rhov[(ss * 3 + 2) * nc + c];
      double E = pa[U - 1];
      C.coef[CIDX(ss, U - 1, 0)] = E;
      C.E[ss * nc + c] = E;
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int b = 0; b < B; b++) C.coef[CIDX(ss, u, 1 + b)] = keep_tail ? cs[u * B + b] / (mass_t[ss] + 1e-20) : 0.0;
#pragma unroll
      for (int a = 0; a < 3; a++) cv[a] = rv[a] / rho;
      double sq = 0;
#pragma unroll
      for (int a = 0; a < D; a++) sq += rv[a] * rv[a];
      double e_j = E - 0.5 * sq / rho;
      cP = eos_p(e_j, Eos{g.gamma[ss], g.pinf[ss]});
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int m = 0; m < M; m++) C.coef0[CIDX(ss, u, m)] = C.coef[CIDX(ss, u, m)];
    }
    C.ph_state_b[ss * nc + c] = st;
    if (st == 0) {
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int m = 0; m < M; m++) {
          C.coef[CIDX(ss, u, m)] = 0;
          C.coef0[CIDX(ss, u, m)] = 0;
        }
      // T2/S3 reset the absent phase's density and momentum
      // (T2:mpm.cpp:346-352); UX keeps them (UX:mpm.cpp:330-334)
      if (g.sem != PGDG_SEM_UX) {
        C.rho[ss * nc + c] = g.ini_rho[ss];
#pragma unroll
        for (int a = 0; a < 3; a++) C.rhov[(ss * 3 + a) * nc + c] = 0;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++) C.v[a * nc + c] = cv[a];
  C.P[c] = cP;
}

// The 3^D neighbour bins of cell (i, j, k) (clipped to the domain).
struct P2gBox {
  int t0, t1, s0, s1, q0, q1;
};
template <int D>
__device__ inline P2gBox p2g_box(const Geo& g, int i, int j, int k) {
  const int kc = D == 3 ? 1 : 0;
  return {max(i - 1, 0), min(i + 1, g.N[0] - 1), max(j - 1, 0), min(j + 1, g.N[1] - 1), max(k - kc, 0),
          min(k + kc, g.N[2] - 1)};
}

// MLS_P2G, phase masses: one lane per cell, cells taken in XCD-local runs
// (xcd_block); the walk over the neighbour bins reads only the gathered
// records.  A cell where no phase appears (ph_state 0 -> 1) is finished
// here; a cell where one does (a thin layer at moving interfaces) keeps its
// masses and is flagged for k_p2g_appear.  The pass is bound by L2
// capacity: each cell reads its 3^D neighbour bins' records, and with more
// than one wavefront per SIMD the cells in flight on an XCD need more
// neighbour records than its L2 holds (measured 2x slower).  The workgroup
// therefore reserves LDS so that one wavefront per SIMD runs.
constexpr int kP2gMassLds = 80 * 1024;
template <int D, int ORD, bool MIX>
__global__ void __launch_bounds__(128) k_p2g(Geo g, CellDev C, P2gBuf pb, const int* __restrict__ bin_off,
                                             const int32_t* __restrict__ used) {
  __shared__ int s_occupancy[kP2gMassLds / 4];  // occupancy control only
  if (threadIdx.x == 0) s_occupancy[0] = 0;
  const int nc = g.nc;
  const int c = xcd_block() * blockDim.x + threadIdx.x;
  if (c >= nc || !used[c]) return;
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  double xc[3];
  cell_center<D>(g, i, j, k, xc);
  const P2gBox bx = p2g_box<D>(g, i, j, k);
  double mass_t[PGDG_MAX_PHASE] = {0, 0, 0};
  for (int t = bx.t0; t <= bx.t1; t++)
    for (int s_ = bx.s0; s_ <= bx.s1; s_++)
      for (int q = bx.q0; q <= bx.q1; q++) {
        int cc = (t * g.N[1] + s_) * g.N[2] + q;
        int e0 = bin_off[cc], e1 = bin_off[cc + 1];
        for (int e = e0; e < e1; e++) {
          const double* __restrict__ lp = pb.light + (int64_t)e * kP2gLight;
          const double2 l01 = *reinterpret_cast<const double2*>(lp);
          const double2 l23 = *reinterpret_cast<const double2*>(lp + 2);
          const double2 l45 = *reinterpret_cast<const double2*>(lp + 4);
          const int ph = (int)l45.x;
          if (ph < 0) continue;  // wall particles are not cell members
          const double xr[3] = {l01.x, l01.y, l23.x};
          double sq = 0;
#pragma unroll
          for (int a = 0; a < D; a++) {
            const double xa = xr[a] - xc[a];
            sq += xa * xa;
          }
          const double ww = p2g_weight<D, MIX>(g, sq, l23.y);  // > 0 exactly within the cut-off
          if (ww != 0.0) {
#pragma unroll
            for (int s = 0; s < PGDG_MAX_PHASE; s++)
              if (s < g.F && ph == s) mass_t[s] += ww;
          }
        }
      }
  bool appear = false;
#pragma unroll
  for (int s = 0; s < PGDG_MAX_PHASE; s++)
    if (s < g.F && mass_t[s] > 0 && C.ph_state_b[s * nc + c] == 0) appear = true;
  pb.flag[c] = appear ? 1 : 0;
  if (appear) {
#pragma unroll
    for (int s = 0; s < PGDG_MAX_PHASE; s++) pb.mass[(size_t)s * nc + c] = mass_t[s];
    return;
  }
  p2g_finish<D, ORD, false>(g, C, c, mass_t, nullptr, nullptr);
}

// MLS_P2G, flagged cells: one lane per cell (the others exit at once), so
// that the appearing cells spread over the whole grid; the neighbour bins
// are walked again in the same order, forming the weighted state and slope
// averages of the appearing phases from the particle arrays.
template <int D, int ORD, bool MIX>
__global__ void __launch_bounds__(64) k_p2g_appear(Geo g, CellDev C, PtlDev P, P2gBuf pb,
                                                   const int* __restrict__ bin_off,
                                                   const int* __restrict__ bin_ent,
                                                   const int32_t* __restrict__ used) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M, B = M - 1;
  const int nc = g.nc;
  const int c = xcd_block() * blockDim.x + threadIdx.x;
  if (c >= nc || !used[c] || !pb.flag[c]) return;
  const int64_t cap = P.cap;
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  double xc[3];
  cell_center<D>(g, i, j, k, xc);
  double mass_t[PGDG_MAX_PHASE];
  unsigned appear = 0;
#pragma unroll
  for (int s = 0; s < PGDG_MAX_PHASE; s++) {
    mass_t[s] = pb.mass[(size_t)s * nc + c];
    if (s < g.F && mass_t[s] > 0 && C.ph_state_b[s * nc + c] == 0) appear |= 1u << s;
  }
  double pav[PGDG_MAX_PHASE][U];
  double cf[PGDG_MAX_PHASE][U][B];
  double qav[PGDG_MAX_PHASE][ORD == 2 ? U : 1];
#pragma unroll
  for (int s = 0; s < PGDG_MAX_PHASE; s++) {
#pragma unroll
    for (int u = 0; u < (ORD == 2 ? U : 1); u++) qav[s][u] = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      pav[s][u] = 0;
#pragma unroll
      for (int b = 0; b < B; b++) cf[s][u][b] = 0;
    }
  }
  const P2gBox bx = p2g_box<D>(g, i, j, k);
  for (int t = bx.t0; t <= bx.t1; t++)
    for (int s_ = bx.s0; s_ <= bx.s1; s_++)
      for (int q = bx.q0; q <= bx.q1; q++) {
        int cc = (t * g.N[1] + s_) * g.N[2] + q;
        int e0 = bin_off[cc], e1 = bin_off[cc + 1];
        for (int e = e0; e < e1; e++) {
          const double* __restrict__ lp = pb.light + (int64_t)e * kP2gLight;
          const double2 l01 = *reinterpret_cast<const double2*>(lp);
          const double2 l23 = *reinterpret_cast<const double2*>(lp + 2);
          const double2 l45 = *reinterpret_cast<const double2*>(lp + 4);
          const int ph = (int)l45.x;
          if (ph < 0 || !((appear >> ph) & 1u)) continue;
          const double xr[3] = {l01.x, l01.y, l23.x};
          double xij[3];
#pragma unroll
          for (int a = 0; a < D; a++) xij[a] = xr[a] - xc[a];
          double sq = 0;
#pragma unroll
          for (int a = 0; a < D; a++) sq += xij[a] * xij[a];
          const double ww = p2g_weight<D, MIX>(g, sq, l23.y);
          if (ww == 0.0) continue;
          const int64_t p = bin_ent[e];
          double pj[U];
          pj[0] = P.rho[p];
#pragma unroll
          for (int a = 0; a < D; a++) pj[1 + a] = P.rhov[a * cap + p];
          pj[U - 1] = P.E[p];
          double pq[ORD == 2 ? U : 1];
#pragma unroll
          for (int u = 0; u < (ORD == 2 ? U : 1); u++) pq[u] = pj[u];
          // the particle's polynomial at the cell centre
          const double nx[3] = {-xij[0], -xij[1], D == 3 ? -xij[2] : 0.0};
          double psi[B];
          incr_tail<D, ORD>(nx, psi);
          double sl[U][B];
#pragma unroll
          for (int u = 0; u < U; u++) {
            double acc = 0;
#pragma unroll
            for (int b = 0; b < B; b++) {
              sl[u][b] = P.coef[(size_t)(u * M + 1 + b) * cap + p];
              acc += sl[u][b] * psi[b];
            }
            pj[u] += acc;
          }
#pragma unroll
          for (int s = 0; s < PGDG_MAX_PHASE; s++) {
            if (s < g.F && ph == s) {
#pragma unroll
              for (int u = 0; u < U; u++) pav[s][u] += pj[u] * ww;
              if constexpr (ORD == 2)
#pragma unroll
                for (int u = 0; u < U; u++) qav[s][u] += pq[u] * ww;
#pragma unroll
              for (int u = 0; u < U; u++)
#pragma unroll
                for (int b = 0; b < B; b++) cf[s][u][b] += sl[u][b] * ww;
            }
          }
        }
      }
  p2g_finish<D, ORD, true>(g, C, c, mass_t, &pav[0][0], &cf[0][0][0], &qav[0][0]);
}

// ---------------------------------------------------------------------------
// Calculate_limiter (S3:mpm.cpp:392-738; T2:mpm.cpp:887-1225)
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T limiter_fn(T ds, int bd_fg) {  // S3:glbfunc.h:283-299
  T val = 1;
  if (ds > 0 && ds < T(0.5)) val = 2 * ds;
  else if (ds >= T(0.5)) val = 1;
  if (bd_fg == 1) val = (fabs(ds) + ds) / (fabs(ds) + 1);
  return val;
}

// Face geometry of side f of cell (i,j,k): centre, unit normal (axis and
// sign), tangents, area, neighbour cell (-1 outside) and boundary tag.
template <int D>
struct FaceGeo {
  double cen[3];
  int axis;
  double sgn;
  double tan0[3], tan1[3];
  double S;
  int nbr;
  int b;
};

template <int D>
__device__ inline void face_geo(const Geo& g, int i, int j, int k, const double* xc, int f,
                                FaceGeo<D>& fg) {
  const double px = g.ps[0], py = g.ps[1], pz = g.ps[2];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    fg.cen[a] = xc[a];
    fg.tan0[a] = 0;
    fg.tan1[a] = 0;
  }
  int ni = i, nj = j, nk = k;
  switch (f) {
    case 0: fg.axis = 1; fg.sgn = -1; fg.tan0[0] = 0.5 * px; fg.cen[1] = xc[1] - 0.5 * py; fg.S = D == 3 ? px * pz : px; nj = j - 1; break;
    case 1: fg.axis = 0; fg.sgn = 1; fg.tan0[1] = 0.5 * py; fg.cen[0] = xc[0] + 0.5 * px; fg.S = D == 3 ? py * pz : py; ni = i + 1; break;
    case 2: fg.axis = 1; fg.sgn = 1; fg.tan0[0] = -0.5 * px; fg.cen[1] = xc[1] + 0.5 * py; fg.S = D == 3 ? px * pz : px; nj = j + 1; break;
    case 3: fg.axis = 0; fg.sgn = -1; fg.tan0[1] = -0.5 * py; fg.cen[0] = xc[0] - 0.5 * px; fg.S = D == 3 ? py * pz : py; ni = i - 1; break;
    case 4: fg.axis = 2; fg.sgn = -1; fg.tan0[0] = 0.5 * px; fg.tan1[1] = 0.5 * py; fg.cen[2] = xc[2] - 0.5 * pz; fg.S = py * px; nk = k - 1; break;
    default: fg.axis = 2; fg.sgn = 1; fg.tan0[0] = 0.5 * px; fg.tan1[1] = 0.5 * py; fg.cen[2] = xc[2] + 0.5 * pz; fg.S = py * px; nk = k + 1; break;
  }
  if (D == 3 && f < 4) fg.tan1[2] = 0.5 * pz;
  bool in = ni >= 0 && ni < g.N[0] && nj >= 0 && nj < g.N[1] && nk >= 0 && nk < g.N[2];
  fg.nbr = in ? (ni * g.N[1] + nj) * g.N[2] + nk : -1;
  fg.b = in ? 0 : g.bc[f];
}

// T: the arithmetic of the slope ratios and the positivity probe (double:
// the reference's FP64 order; float: the mixed form).  The neighbour
// extrema and the velocity gradient stay FP64.
template <int D, int ORD, int BS, typename T>
__global__ void __launch_bounds__(BS) k_limiter(Geo g, CellDev C, const int32_t* __restrict__ used) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M, NF = 2 * D;
  const int nc = g.nc;
  int c = xcd_block() * blockDim.x + threadIdx.x;
  if (c >= nc || !used[c]) return;
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  double xc[3];
  cell_center<D>(g, i, j, k, xc);
  const int F = g.F;
  // float form: phases absent from the cell get no limiter (nothing in the
  // mixed step reads an absent phase's limiter); the FP64 form computes all,
  // as the reference
  unsigned need = 0;
  for (int s = 0; s < PGDG_MAX_PHASE; s++)
    if (s < F && (sizeof(T) == 8 || C.ph_state[s * nc + c] == 1)) need |= 1u << s;
  double umax[PGDG_MAX_PHASE][U], umin[PGDG_MAX_PHASE][U], pi[PGDG_MAX_PHASE][U];
  double crho[PGDG_MAX_PHASE], calph[PGDG_MAX_PHASE], crv[PGDG_MAX_PHASE][3];
  for (int s = 0; s < PGDG_MAX_PHASE; s++) {
    if (s >= F) break;
#pragma unroll
    for (int u = 0; u < U; u++) {
      umax[s][u] = -1e20;
      umin[s][u] = 1e20;
    }
    crho[s] = C.rho[s * nc + c];
    calph[s] = C.alph[s * nc + c];
#pragma unroll
    for (int a = 0; a < 3; a++) crv[s][a] = C.rhov[(s * 3 + a) * nc + c];
    pi[s][0] = crho[s];
#pragma unroll
    for (int a = 0; a < D; a++) pi[s][1 + a] = crv[s][a];
    pi[s][U - 1] = C.E[s * nc + c];
  }
  // velocity gradient G_v (S3:mpm.cpp:471-533)
  double G[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) G[a][b] = 0.;
  int bd_fg = 0;
  double vc[3] = {C.v[c], C.v[nc + c], C.v[2 * nc + c]};
  // the underwater-explosion semantics run the FP64 form only (the mixed
  // form refuses them): no UX code in the FP32 kernels
  const bool ux = sizeof(T) == 8 && g.sem == PGDG_SEM_UX;
  {
    // UX takes a zero wall velocity on the x sides (UX:mpm.cpp:388-399);
    // T2/S3 the cell's own (a zero difference)
    if (i == g.N[0] - 1) {
#pragma unroll
      for (int a = 0; a < D; a++) G[a][0] = ux ? (0.0 - vc[a]) / (0.5 * g.ps[0]) : (vc[a] - vc[a]) / (0.5 * g.ps[0]);
      bd_fg = 1;
    } else if (i == 0) {
#pragma unroll
      for (int a = 0; a < D; a++) G[a][0] = ux ? (vc[a] - 0.0) / (0.5 * g.ps[0]) : (vc[a] - vc[a]) / (0.5 * g.ps[0]);
      bd_fg = 1;
    } else {
      int cr = c + g.N[1] * g.N[2], cl = c - g.N[1] * g.N[2];
#pragma unroll
      for (int a = 0; a < D; a++) G[a][0] = (C.v[a * nc + cr] - C.v[a * nc + cl]) / (2 * g.ps[0]);
    }
    if (j == g.N[1] - 1) {
#pragma unroll
      for (int a = 0; a < D; a++) G[a][1] = (0.0 - vc[a]) / (0.5 * g.ps[1]);
      bd_fg = 1;
    } else if (j == 0) {
#pragma unroll
      for (int a = 0; a < D; a++) G[a][1] = (vc[a] - 0.0) / (0.5 * g.ps[1]);
      bd_fg = 1;
    } else {
      int cr = c + g.N[2], cl = c - g.N[2];
#pragma unroll
      for (int a = 0; a < D; a++) G[a][1] = (C.v[a * nc + cr] - C.v[a * nc + cl]) / (2 * g.ps[1]);
    }
    if (D == 3) {
      if (k == g.N[2] - 1) {
#pragma unroll
        for (int a = 0; a < D; a++) G[a][2] = (0.0 - vc[a]) / (0.5 * g.ps[2]);
        bd_fg = 1;
      } else if (k == 0) {
#pragma unroll
        for (int a = 0; a < D; a++) G[a][2] = (vc[a] - 0.0) / (0.5 * g.ps[2]);
        bd_fg = 1;
      } else {
#pragma unroll
        for (int a = 0; a < D; a++) G[a][2] = (C.v[a * nc + c + 1] - C.v[a * nc + c - 1]) / (2 * g.ps[2]);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) C.G[(a * 3 + b) * nc + c] = G[a][b];
  const bool s3 = g.sem == PGDG_SEM_S3;
  // neighbour extrema
  for (int f = 0; f < NF; f++) {
    FaceGeo<D> fg;
    face_geo<D>(g, i, j, k, xc, f, fg);
    if (fg.b == 0) {
      int nb = fg.nbr;
      for (int s = 0; s < PGDG_MAX_PHASE; s++) {
        if (s >= F) break;
        if (!((need >> s) & 1u)) continue;
        double pj[U];
#pragma unroll
        for (int u = 0; u < U; u++) pj[u] = 0;
        for (int s1 = 0; s1 < PGDG_MAX_PHASE; s1++) {
          if (s1 >= F) break;
          double al = C.alph[s1 * nc + nb];
          if (al == 0) continue;
          double rhoj = C.rho[s1 * nc + nb];
          double rvj[3];
#pragma unroll
          for (int a = 0; a < D; a++) rvj[a] = C.rhov[(s1 * 3 + a) * nc + nb];
          double Ej = C.E[s1 * nc + nb];
          if (s1 != s) {
            double vj[3];
#pragma unroll
            for (int a = 0; a < D; a++) vj[a] = rvj[a] / rhoj;
            double sq = 0;
#pragma unroll
            for (int a = 0; a < D; a++) sq += vj[a] * vj[a];
            double rhoe_j = Ej - 0.5 * rhoj * sq;
            double P_j = eos_p(rhoe_j, Eos{g.gamma[s1], g.pinf[s1]});
            rhoj = crho[s];
#pragma unroll
            for (int a = 0; a < D; a++) rvj[a] = rhoj * vj[a];
            Ej = eos_rhoe(P_j, Eos{g.gamma[s], g.pinf[s]}) + 0.5 * rhoj * sq;
          }
          pj[0] += rhoj * al;
#pragma unroll
          for (int a = 0; a < D; a++) pj[1 + a] += rvj[a] * al;
          pj[U - 1] += Ej * al;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          double d = pj[u] - pi[s][u];
          umax[s][u] = amax1(umax[s][u], d);
          umin[s][u] = amin1(umin[s][u], d);
        }
      }
    } else {
      double pj[U];
#pragma unroll
      for (int u = 0; u < U; u++) pj[u] = 0;
      for (int s = 0; s < PGDG_MAX_PHASE; s++) {
        if (s >= F) break;
        if (s3) {
          pj[0] = crho[s];
#pragma unroll
          for (int a = 0; a < D; a++) pj[1 + a] = crv[s][a];
          pj[U - 1] = pi[s][U - 1];
          if (fg.b == -1 || fg.b == -2)
#pragma unroll
            for (int u = 0; u < U; u++) pj[u] = pi[s][u];
          if (j == g.N[1] - 1 || j == 0) pj[2] = -crv[s][1];
          if (D == 3 && (k == g.N[2] - 1 || k == 0)) pj[3] = -crv[s][2];
        } else if (ux) {
          // the cell's own means, the momentum normal to a domain edge set
          // to 0, whatever the side's tag (UX:mpm.cpp:478-487)
          pj[0] = crho[s];
#pragma unroll
          for (int a = 0; a < D; a++) pj[1 + a] = crv[s][a];
          pj[U - 1] = pi[s][U - 1];
          if (i == g.N[0] - 1 || i == 0) pj[1] = 0;
          if (j == g.N[1] - 1 || j == 0) pj[2] = 0;
        } else {
          if (fg.b == 1) {
            pj[0] = crho[s];
            pj[1] = crv[s][0];
            pj[2] = -crv[s][1];
            pj[U - 1] = pi[s][U - 1];
          } else if (fg.b == -1 || fg.b == -2) {
#pragma unroll
            for (int u = 0; u < U; u++) pj[u] = pi[s][u];
          }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          double d = pj[u] - pi[s][u];
          umax[s][u] = amax1(umax[s][u], d);
          umin[s][u] = amin1(umin[s][u], d);
        }
      }
    }
  }
  for (int s = 0; s < PGDG_MAX_PHASE; s++) {
    if (s >= F) break;
#pragma unroll
    for (int u = 0; u < U; u++) {
      umax[s][u] = amax1(umax[s][u], 0.);
      umin[s][u] = amin1(umin[s][u], 0.);
    }
  }
  // ratio of the allowed range to the slope increment at neighbour centres
  for (int s = 0; s < PGDG_MAX_PHASE; s++) {
    if (s >= F) break;
    if (!((need >> s) & 1u)) continue;
    T sl[U][M - 1];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int b = 0; b < M - 1; b++) sl[u][b] = (T)C.coef[CIDX(s, u, 1 + b)];
    T lim[U];
#pragma unroll
    for (int u = 0; u < U; u++) lim[u] = 1.;
    // UX's Limiter is 2 ds / (ds + 1) in every cell (UX:glbfunc.h:291-306)
    int lim_fg = ux ? 1 : bd_fg;
    if (calph[s] != 1) lim_fg = 1;
    for (int f = 0; f < NF; f++) {
      FaceGeo<D> fg;
      face_geo<D>(g, i, j, k, xc, f, fg);
      double dr[3];
      if (fg.b == 0) {
        int ni, nj, nk;
        cell_ijk(g, fg.nbr, ni, nj, nk);
        double xn[3];
        cell_center<D>(g, ni, nj, nk, xn);
#pragma unroll
        for (int a = 0; a < 3; a++) dr[a] = xn[a] - xc[a];
      } else {
#pragma unroll
        for (int a = 0; a < 3; a++) dr[a] = fg.cen[a] - xc[a];
      }
      const T drT[3] = {(T)dr[0], (T)dr[1], (T)dr[2]};
      T phi[M - 1];
      basis_tail<D, ORD>(drT, g.q12, phi);
#pragma unroll
      for (int u = 0; u < U; u++) {
        T dU = 0;
#pragma unroll
        for (int b = 0; b < M - 1; b++) dU += sl[u][b] * phi[b];
        T sd = 1;
        if (dU >= 0) {
          sd = qdiv((T)umax[s][u], dU + T(1e-20));
          sd = limiter_fn(fabs(sd), lim_fg);
        } else if (dU < 0) {
          sd = qdiv((T)umin[s][u], dU + T(1e-20));
          sd = limiter_fn(fabs(sd), lim_fg);
        }
        lim[u] = amin1(lim[u], sd);
      }
    }
    // positivity probe
    int fgp = 0;
    T c0[U];
#pragma unroll
    for (int u = 0; u < U; u++) c0[u] = (T)C.coef[CIDX(s, u, 0)];
    const EosT<T> mt(Eos{g.gamma[s], g.pinf[s]});
    for (int jj = 0; jj < 4 && !fgp; jj++) {
      for (int k1 = 0; k1 < (D == 3 ? 2 : 1) && !fgp; k1++) {
        double dr[3] = {0, 0, 0};
        dr[0] = g.sqrt2 * g.p_g * g.probe_c[jj];
        dr[1] = g.sqrt2 * g.p_g * g.probe_s[jj];
        if (D == 3) dr[2] = (-0.5 + k1) * g.p_g;
        const T drT[3] = {(T)dr[0], (T)dr[1], (T)dr[2]};
        T phi[M - 1];
        basis_tail<D, ORD>(drT, g.q12, phi);
        T Ui[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          T acc = 0;
#pragma unroll
          for (int b = 0; b < M - 1; b++) acc += (lim[u] * sl[u][b]) * phi[b];
          Ui[u] = c0[u] + acc;
        }
        T sq = 0;
#pragma unroll
        for (int a = 0; a < D; a++) sq += Ui[1 + a] * Ui[1 + a];
        T e_p = Ui[U - 1] - qdiv(T(0.5) * sq, Ui[0]);
        T P_p = mt.p(e_p);
        if (P_p < T(-0.05) * mt.pinf || Ui[0] < 0) fgp = 1;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) C.lim[(s * U + u) * nc + c] = fgp ? 0. : lim[u];
  }
}

// ---------------------------------------------------------------------------
// DG right-hand side (S3:particle_flow.cpp:16-405)
// ---------------------------------------------------------------------------
// T: the arithmetic of the quadrature (double: the reference's FP64 order,
// bit-identical to the oracle; float: the mixed form, FP32 evaluation, fluxes
// and sums with the MGFM Newton solve in FP64 and the mass division and
// stored rates in FP64).  The float form takes the interior-face point
// fluxes from the face kernels (each face's HLLC evaluated once for both of
// its cells) and evaluates only the volume integral and the domain-boundary
// faces itself.
// md >= 0 (float form only): also apply stage md's coefficient update.
template <int D, int ORD, int BS, typename T>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
k_rhs(Geo g, CellDev C, const int32_t* __restrict__ used, const DevScalars* __restrict__ sc, int md) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M, B = M - 1, NF = 2 * D, NQ = Basis<D, ORD>::NQ;
  constexpr bool kFaceBuf = sizeof(T) == 4;
  constexpr int NQF = D == 3 ? NQ * NQ : NQ;
  // FP64 form: per-lane LDS slot for the neighbour's coefficients and
  // limiter of its first present phase, loaded once per face instead of once
  // per face point
  __shared__ T s_nb[kFaceBuf ? 1 : U * M + U][BS];
  const int ln = threadIdx.x;
  const int nc = g.nc;
  int c = xcd_block() * blockDim.x + threadIdx.x;
  if (c >= nc || !used[c]) return;
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  // geometry (offsets, weights) stays FP64; the state arithmetic is T
  double xc[3];
  cell_center<D>(g, i, j, k, xc);
  const bool s3 = g.sem == PGDG_SEM_S3;
  const bool ux = sizeof(T) == 8 && g.sem == PGDG_SEM_UX;  // FP64 form only (see k_limiter)
  const double hp[3] = {g.ps[0] / 2, g.ps[1] / 2, g.ps[2] / 2};
  const double vol64 = D == 3 ? g.ps[0] * g.ps[1] * g.ps[2] : g.ps[0] * g.ps[1];
  const T voli = (T)vol64;
  for (int ss = 0; ss < PGDG_MAX_PHASE; ss++) {
    if (ss >= g.F) break;
    if (C.ph_state[ss * nc + c] != 1) continue;
    const Eos me{g.gamma[ss], g.pinf[ss]};
    const EosT<T> mt(me);
    T cf[U][M], lm[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      lm[u] = C.lim[(ss * U + u) * nc + c];
#pragma unroll
      for (int m = 0; m < M; m++) cf[u][m] = C.coef[CIDX(ss, u, m)];
    }
    T dc[U][M];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int m = 0; m < M; m++) dc[u][m] = 0;
    // volume integral, Gauss points in g1, g2[, g3] order
    auto vol_point = [&](int g1, int g2, int g3) {
          double dp[3] = {0, 0, 0};
          dp[0] = g.Gp[g1] * hp[0];
          dp[1] = g.Gp[g2] * hp[1];
          T gw = g.Gw[g1] * g.Gw[g2];
          if (D == 3) {
            dp[2] = g.Gp[g3] * hp[2];
            gw = g.Gw[g1] * g.Gw[g2] * g.Gw[g3];
          }
          const T dpT[3] = {(T)dp[0], (T)dp[1], (T)dp[2]};
          T phv[B];
          basis_tail<D, ORD>(dpT, g.q12, phv);
          T Up[U];
#pragma unroll
          for (int u = 0; u < U; u++) {
            T acc = 0;
#pragma unroll
            for (int b = 0; b < B; b++) {
              if constexpr (kFaceBuf) acc = fmaf(cf[u][1 + b], phv[b], acc);
              else acc += cf[u][1 + b] * phv[b];
            }
            Up[u] = cf[u][0] + lm[u] * acc;
          }
          T rv[3] = {0, 0, 0}, v[3] = {0, 0, 0};
#pragma unroll
          for (int a = 0; a < D; a++) rv[a] = Up[1 + a];
#pragma unroll
          for (int a = 0; a < D; a++) v[a] = qdiv(rv[a], Up[0]);
          T sq = 0;
#pragma unroll
          for (int a = 0; a < D; a++) sq += rv[a] * rv[a];
          T e_p = Up[U - 1] - qdiv(T(0.5) * sq, Up[0]);
          T P_p = mt.p(e_p);
          T col[D][U];
#pragma unroll
          for (int b = 0; b < D; b++) flux_col<D>(Up[0], rv, Up[U - 1], v, P_p, b, col[b]);
          if constexpr (kFaceBuf) {
            // float form: the point weight once, fused products
            const T wv = gw * voli;
#pragma unroll
            for (int bi = 1; bi <= D; bi++)
#pragma unroll
              for (int u = 0; u < U; u++) dc[u][bi] = fmaf(col[bi - 1][u], wv, dc[u][bi]);
            if constexpr (ORD == 2) {
              const T x2[3] = {T(2.) * dpT[0], T(2.) * dpT[1], T(2.) * dpT[2]};
#pragma unroll
              for (int u = 0; u < U; u++) {
                T c0 = col[0][u] * wv, c1 = col[1][u] * wv;
                if constexpr (D == 2) {
                  dc[u][3] = fmaf(c0, dpT[1], fmaf(c1, dpT[0], dc[u][3]));
                  dc[u][4] = fmaf(c0, x2[0], dc[u][4]);
                  dc[u][5] = fmaf(c1, x2[1], dc[u][5]);
                } else {
                  T c2 = col[2][u] * wv;
                  dc[u][4] = fmaf(c0, dpT[1], fmaf(c1, dpT[0], dc[u][4]));
                  dc[u][5] = fmaf(c0, dpT[2], fmaf(c2, dpT[0], dc[u][5]));
                  dc[u][6] = fmaf(c1, dpT[2], fmaf(c2, dpT[1], dc[u][6]));
                  dc[u][7] = fmaf(c0, x2[0], dc[u][7]);
                  dc[u][8] = fmaf(c1, x2[1], dc[u][8]);
                  dc[u][9] = fmaf(c2, x2[2], dc[u][9]);
                }
              }
            }
            return;
          }
#pragma unroll
          for (int bi = 1; bi <= D; bi++)
#pragma unroll
            for (int u = 0; u < U; u++) dc[u][bi] += col[bi - 1][u] * gw * voli * 1;
          if constexpr (ORD == 2) {
            // second-order modes: sum_b col[b] d(phi_m)/dx_b at dp
#pragma unroll
            for (int u = 0; u < U; u++) {
              if constexpr (D == 2) {
                dc[u][3] += (col[0][u] * dpT[1] + col[1][u] * dpT[0]) * gw * voli;
                dc[u][4] += (col[0][u] * (T(2.) * dpT[0])) * gw * voli;
                dc[u][5] += (col[1][u] * (T(2.) * dpT[1])) * gw * voli;
              } else {
                dc[u][4] += (col[0][u] * dpT[1] + col[1][u] * dpT[0]) * gw * voli;
                dc[u][5] += (col[0][u] * dpT[2] + col[2][u] * dpT[0]) * gw * voli;
                dc[u][6] += (col[1][u] * dpT[2] + col[2][u] * dpT[1]) * gw * voli;
                dc[u][7] += (col[0][u] * (T(2.) * dpT[0])) * gw * voli;
                dc[u][8] += (col[1][u] * (T(2.) * dpT[1])) * gw * voli;
                dc[u][9] += (col[2][u] * (T(2.) * dpT[2])) * gw * voli;
              }
            }
          }
    };
    if constexpr (kFaceBuf) {
      // float form: one rolled loop over the points (unrolled, the point
      // bodies hold more values live than the register file has)
      constexpr int NZ = D == 3 ? NQ : 1;
#pragma unroll 1
      for (int p = 0; p < NQ * NQ * NZ; p++) vol_point(p / (NQ * NZ), (p / NZ) % NQ, p % NZ);
    } else {
      for (int g1 = 0; g1 < NQ; g1++)
        for (int g2 = 0; g2 < NQ; g2++)
          for (int g3 = 0; g3 < (D == 3 ? NQ : 1); g3++) vol_point(g1, g2, g3);
    }
    // face integral
    for (int f = 0; f < NF; f++) {
      FaceGeo<D> fg;
      face_geo<D>(g, i, j, k, xc, f, fg);
      const int ax = fg.axis;
      const T sgn = fg.sgn;
      T nrm[3] = {0, 0, 0};
#pragma unroll
      for (int a = 0; a < 3; a++)
        if (a == ax) nrm[a] = sgn;
      // FP32 form: face terms go straight into the rates (scaled by -n S);
      // the FP64 form sums each face first, in the oracle's order
      constexpr bool kDirect = sizeof(T) == 4;
      const T fs = (T)(-(double)fg.sgn * fg.S);
      T DFn[kDirect ? 1 : M][U];
#pragma unroll
      for (int m = 0; m < (kDirect ? 1 : M); m++)
#pragma unroll
        for (int u = 0; u < U; u++) DFn[m][u] = 0;
      if constexpr (kFaceBuf) {
        if (fg.b == 0) {  // interior face: the face kernels' point fluxes
          const int fc = fg.sgn > 0 ? c : fg.nbr;
          // one flux for both sides (weighted here by the other cell's
          // volume fraction) or this side's own, weighted, slots
          const bool one = C.fkind[ax * nc + fc] != 0;
          const int side = one ? 0 : (fg.sgn > 0 ? 0 : 1);
          const T wo = one ? (T)C.alph[ss * nc + fg.nbr] : T(1);
#pragma unroll 1
          for (int q = 0; q < NQF; q++) {
            double off[3], gwd;
            face_point<D, ORD>(g, ax, q, hp, off, gwd);
            T dri[3];
#pragma unroll
            for (int a = 0; a < 3; a++) dri[a] = (T)(off[a] + fg.cen[a] - xc[a]);
            T Nb[M];
            Nb[0] = T(1.);
            basis_tail<D, ORD>(dri, g.q12, Nb + 1);
            const T wq = (T)gwd * fs;
#pragma unroll
            for (int u = 0; u < U; u++) {
              const T Fw = (C.fflux[face_slot<D, ORD>(g, side, ss, q, u, ax, fc)] * wo) * wq;
#pragma unroll
              for (int bi = 0; bi < M; bi++) dc[u][bi] = fmaf(Fw, Nb[bi], dc[u][bi]);
            }
          }
          continue;
        }
      }
      // T2: every boundary face is a wall; S3: tag 1; UX: none (HLL ghost)
      const bool wall_like = s3 ? (fg.b == 1) : (!ux && fg.b != 0);
      int s_first = -1;
      if (!kFaceBuf && fg.b == 0) {
        for (int s1 = 0; s1 < PGDG_MAX_PHASE && s_first < 0; s1++)
          if (s1 < g.F && C.alph[s1 * nc + fg.nbr] != 0) s_first = s1;
        if (s_first >= 0) {
#pragma unroll
          for (int v = 0; v < U * M; v++) s_nb[v][ln] = (T)C.coef[((size_t)s_first * U * M + v) * nc + fg.nbr];
#pragma unroll
          for (int u = 0; u < U; u++) s_nb[U * M + u][ln] = (T)C.lim[(s_first * U + u) * nc + fg.nbr];
        }
      }
      for (int g1 = 0; g1 < NQ; g1++)
        for (int g2 = 0; g2 < (D == 3 ? NQ : 1); g2++) {
          double dp[3];
          T gw;
          if (D == 3) {
#pragma unroll
            for (int a = 0; a < 3; a++) dp[a] = g.Gp[g1] * fg.tan0[a] + g.Gp[g2] * fg.tan1[a];
            gw = g.Gw[g1] * g.Gw[g2];
          } else {
#pragma unroll
            for (int a = 0; a < 3; a++) dp[a] = g.Gp[g1] * fg.tan0[a];
            gw = g.Gw[g1];
          }
          double dri[3];
#pragma unroll
          for (int a = 0; a < 3; a++) dri[a] = dp[a] + fg.cen[a] - xc[a];
          double nb64[B];
          basis_tail<D, ORD>(dri, g.q12, nb64);
          T Nb[M];
          Nb[0] = T(1.);
#pragma unroll
          for (int b = 0; b < B; b++) Nb[1 + b] = (T)nb64[b];
          T Ui[U];
#pragma unroll
          for (int u = 0; u < U; u++) {
            T acc = 0;
#pragma unroll
            for (int b = 0; b < B; b++) acc += cf[u][1 + b] * Nb[1 + b];
            Ui[u] = cf[u][0] + lm[u] * acc;
          }
          const T rho_i = Ui[0], E_i = Ui[U - 1];
          T rvi[3] = {0, 0, 0}, vi[3] = {0, 0, 0};
#pragma unroll
          for (int a = 0; a < D; a++) rvi[a] = Ui[1 + a];
#pragma unroll
          for (int a = 0; a < D; a++) vi[a] = qdiv(rvi[a], rho_i);
          T sq = 0;
#pragma unroll
          for (int a = 0; a < D; a++) sq += rvi[a] * rvi[a];
          T e_i = E_i - qdiv(T(0.5) * sq, rho_i);
          T P_i = mt.p(e_i);
          T ci = qsqrt(qdiv(mt.g * (mt.pinf + P_i), rho_i));
          T vni = 0;
#pragma unroll
          for (int a = 0; a < D; a++) vni += vi[a] * nrm[a];
          T DFi[U];
          flux_col<D>(rho_i, rvi, E_i, vi, P_i, ax, DFi);
          if (!kFaceBuf && fg.b == 0) {
            const int nb = fg.nbr;
            int ni, nj, nk;
            cell_ijk(g, nb, ni, nj, nk);
            double xn[3];
            cell_center<D>(g, ni, nj, nk, xn);
            double drj[3];
#pragma unroll
            for (int a = 0; a < 3; a++) drj[a] = dp[a] + fg.cen[a] - xn[a];
            double pj64[B];
            basis_tail<D, ORD>(drj, g.q12, pj64);
            T phj[B];
#pragma unroll
            for (int b = 0; b < B; b++) phj[b] = (T)pj64[b];
            for (int s1 = 0; s1 < PGDG_MAX_PHASE; s1++) {
              if (s1 >= g.F) break;
              T alph_j = C.alph[s1 * nc + nb];
              if (alph_j == 0) continue;
              T Uj[U];
              if (s1 == s_first) {
#pragma unroll
                for (int u = 0; u < U; u++) {
                  T acc = 0;
#pragma unroll
                  for (int b = 0; b < B; b++) acc += s_nb[u * M + 1 + b][ln] * phj[b];
                  Uj[u] = s_nb[u * M][ln] + s_nb[U * M + u][ln] * acc;
                }
              } else {
#pragma unroll
                for (int u = 0; u < U; u++) {
                  T acc = 0;
#pragma unroll
                  for (int b = 0; b < B; b++) acc += (T)C.coef[(((size_t)s1 * U + u) * M + 1 + b) * nc + nb] * phj[b];
                  Uj[u] = (T)C.coef[(((size_t)s1 * U + u) * M) * nc + nb] + (T)C.lim[(s1 * U + u) * nc + nb] * acc;
                }
              }
              T rho_j = Uj[0], E_j = Uj[U - 1];
              T rvj[3] = {0, 0, 0}, vj[3] = {0, 0, 0};
#pragma unroll
              for (int a = 0; a < D; a++) rvj[a] = Uj[1 + a];
#pragma unroll
              for (int a = 0; a < D; a++) vj[a] = rvj[a] / rho_j;
              T sqj = 0;
#pragma unroll
              for (int a = 0; a < D; a++) sqj += vj[a] * vj[a];
              T rhoe_j = E_j - T(0.5) * sqj * rho_j;
              T P_j = EosT<T>(Eos{g.gamma[s1], g.pinf[s1]}).p(rhoe_j);
              T vnj = 0;
#pragma unroll
              for (int a = 0; a < D; a++) vnj += vj[a] * nrm[a];
              if (ss != s1) {
                MgfmOut o = mgfm_solve(vni, vnj, P_i, P_j, rho_i, rho_j, me,
                                       Eos{g.gamma[s1], g.pinf[s1]}, s3 ? MGFM_S3 : MGFM_T2,
                                       TOL_REF, 0.0, 50, -1.0);
                rho_j = o.rho_l;
                P_j = o.pstar;
                vnj = o.u_l;
#pragma unroll
                for (int a = 0; a < D; a++) vj[a] = vi[a] - vni * nrm[a] + vnj * nrm[a];
#pragma unroll
                for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
                rhoe_j = mt.rhoe(P_j);
                T sq2 = 0;
#pragma unroll
                for (int a = 0; a < D; a++) sq2 += vj[a] * vj[a];
                E_j = rhoe_j + T(0.5) * rho_j * sq2;
              }
              T cj = sqrt(mt.g * (mt.pinf + P_j) / rho_j);
              T Ujs[U];
              Ujs[0] = rho_j;
#pragma unroll
              for (int a = 0; a < D; a++) Ujs[1 + a] = rvj[a];
              Ujs[U - 1] = E_j;
              T DFj[U];
              flux_col<D>(rho_j, rvj, E_j, vj, P_j, ax, DFj);
              vni = 0;
#pragma unroll
              for (int a = 0; a < D; a++) vni += vi[a] * nrm[a];
              vnj = 0;
#pragma unroll
              for (int a = 0; a < D; a++) vnj += vj[a] * nrm[a];
              T SL = amin1(vni - ci, vnj - cj);
              T SR = amax1(vni + ci, vnj + cj);
              T Fh[U];
              if (ux) {  // HLL (UX:particle_flow.cpp:280-294)
                if (SL > 0) {
#pragma unroll
                  for (int u = 0; u < U; u++) Fh[u] = DFi[u];
                } else if (SL <= 0 && SR >= 0) {
#pragma unroll
                  for (int u = 0; u < U; u++)
                    Fh[u] = qdiv(SR * DFi[u] - SL * DFj[u] + SL * SR * (Ujs[u] - Ui[u]) * sgn, SR - SL);
                } else if (SR < 0) {
#pragma unroll
                  for (int u = 0; u < U; u++) Fh[u] = DFj[u];
                } else {
#pragma unroll
                  for (int u = 0; u < U; u++) Fh[u] = 0;
                }
              } else {
              T S_s = ((P_j - P_i) + rho_i * vni * (SL - vni) - rho_j * vnj * (SR - vnj)) /
                           (rho_i * (SL - vni) - rho_j * (SR - vnj));
              if (SL > 0) {
#pragma unroll
                for (int u = 0; u < U; u++) Fh[u] = DFi[u];
              } else if (SL <= 0 && S_s >= 0) {
                T rho_sL = rho_i * (SL - vni) / (SL - S_s);
                T UsL[U];
                UsL[0] = rho_sL;
#pragma unroll
                for (int a = 0; a < D; a++) UsL[1 + a] = rho_sL * (vi[a] + (S_s - vni) * nrm[a]);
                UsL[U - 1] = rho_sL * (E_i / rho_i + (S_s - vni) * (S_s + (P_i) / (rho_i * (SL - vni))));
#pragma unroll
                for (int u = 0; u < U; u++) Fh[u] = DFi[u] + SL * (UsL[u] - Ui[u]) * sgn;
              } else if (S_s <= 0 && SR >= 0) {
                T rho_sR = rho_j * (SR - vnj) / (SR - S_s);
                T UsR[U];
                UsR[0] = rho_sR;
#pragma unroll
                for (int a = 0; a < D; a++) UsR[1 + a] = rho_sR * (vj[a] + (S_s - vnj) * nrm[a]);
                UsR[U - 1] = rho_sR * (E_j / rho_j + (S_s - vnj) * (S_s + (P_j) / (rho_j * (SR - vnj))));
#pragma unroll
                for (int u = 0; u < U; u++) Fh[u] = DFj[u] + SR * (UsR[u] - Ujs[u]) * sgn;
              } else if (SR <= 0) {
#pragma unroll
                for (int u = 0; u < U; u++) Fh[u] = DFj[u];
              } else {
#pragma unroll
                for (int u = 0; u < U; u++) Fh[u] = 0;
              }
              }
#pragma unroll
              for (int bi = 0; bi < M; bi++)
#pragma unroll
                for (int u = 0; u < U; u++) {
                  if constexpr (kDirect) dc[u][bi] += Fh[u] * Nb[bi] * gw * alph_j * fs;
                  else DFn[bi][u] += Fh[u] * Nb[bi] * gw * alph_j;
                }
            }
          } else if (wall_like) {
#pragma unroll
            for (int bi = 0; bi < M; bi++) {
              const T t = P_i * Nb[bi] * gw;
#pragma unroll
              for (int a = 0; a < D; a++)
                if (a == ax) {
                  if constexpr (kDirect) dc[1 + a][bi] += t * fs;
                  else DFn[bi][1 + a] += t;
                }
            }
          } else {  // S3 inlet / outlet / symmetry and every UX boundary face: HLL
                    // against a ghost state (UX:particle_flow.cpp:314-366)
            T rho_j = rho_i;
            T vj[3] = {0, 0, 0}, rvj[3] = {0, 0, 0};
#pragma unroll
            for (int a = 0; a < D; a++) vj[a] = vi[a];
#pragma unroll
            for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
            T P_j = P_i, E_j = E_i;
            if (fg.b == -1 && s3) {
              rho_j = g.inlet_rho;
#pragma unroll
              for (int a = 0; a < D; a++) vj[a] = g.inlet_u[a];
#pragma unroll
              for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
              P_j = g.inlet_p;
              T sqv = 0;
#pragma unroll
              for (int a = 0; a < D; a++) sqv += vj[a] * vj[a];
              E_j = mt.rhoe(P_j) + T(0.5) * sqv * rho_j;
            }
            // S3 reflects v_y in the first and last y rows; UX on wall-tagged
            // sides (its tag 2)
            if (ux ? fg.b == PGDG_BC_WALL : (j == g.N[1] - 1 || j == 0)) {
              vj[1] = -vi[1];
#pragma unroll
              for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
            }
            if (D == 3 && (k == g.N[2] - 1 || k == 0)) {
              vj[2] = -vi[2];
#pragma unroll
              for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
            }
            T Ujs[U];
            Ujs[0] = rho_j;
#pragma unroll
            for (int a = 0; a < D; a++) Ujs[1 + a] = rvj[a];
            Ujs[U - 1] = E_j;
            T cj = qsqrt(qdiv(mt.g * (mt.pinf + P_j), rho_j));
            T vnj = 0;
#pragma unroll
            for (int a = 0; a < D; a++) vnj += vj[a] * nrm[a];
            T DFj[U];
            flux_col<D>(rho_j, rvj, E_j, vj, P_j, ax, DFj);
            T SL = amin1(vni - ci, vnj - cj);
            T SR = amax1(vni + ci, vnj + cj);
            T Fh[U];
            if (SL > 0) {
#pragma unroll
              for (int u = 0; u < U; u++) Fh[u] = DFi[u];
            } else if (SL <= 0 && SR >= 0) {
#pragma unroll
              for (int u = 0; u < U; u++)
                Fh[u] = qdiv(SR * DFi[u] - SL * DFj[u] + SL * SR * (Ujs[u] - Ui[u]) * sgn, SR - SL);
            } else if (SR < 0) {
#pragma unroll
              for (int u = 0; u < U; u++) Fh[u] = DFj[u];
            } else {
#pragma unroll
              for (int u = 0; u < U; u++) Fh[u] = 0;
            }
#pragma unroll
            for (int bi = 0; bi < M; bi++)
#pragma unroll
              for (int u = 0; u < U; u++) {
                if constexpr (kDirect) dc[u][bi] += Fh[u] * Nb[bi] * gw * fs;
                else DFn[bi][u] += Fh[u] * Nb[bi] * gw;
              }
          }
        }
#pragma unroll
      for (int bi = 0; bi < M; bi++)
#pragma unroll
        for (int u = 0; u < U; u++)
          if constexpr (!kDirect) dc[u][bi] += (-(DFn[bi][u]) * sgn) * fg.S;
    }
    // diagonal mass matrix Vol (1, dx^2/12, ...)
    double mij[M];
    mij[0] = 1.;
#pragma unroll
    for (int a = 0; a < D; a++) mij[1 + a] = hp[a] * hp[a] / 3.;
    if constexpr (ORD == 2) {  // UX:particle_flow.cpp:28-29: (hx hy)^2/9 and 4 hx^4/45
#pragma unroll
      for (int q = 0; q < (D == 2 ? 1 : 3); q++) {
        const double t = q == 0 ? hp[0] * hp[1] : (q == 1 ? hp[0] * hp[2] : hp[1] * hp[2]);
        mij[1 + D + q] = t * t / 9.;
      }
#pragma unroll
      for (int a = 0; a < D; a++) {
        const double t = hp[a] * hp[a];
        mij[M - D + a] = t * t * 4. / 45.;
      }
    }
#pragma unroll
    for (int m = 0; m < M; m++) mij[m] *= vol64;
#pragma unroll
    for (int m = 0; m < M; m++) {
      double inv = 1. / (mij[m] + 1e-20);
#pragma unroll
      for (int u = 0; u < U; u++) C.dcoef[CIDX(ss, u, m)] = dc[u][m] * inv;
    }
    if (kFaceBuf && md >= 0) {
      // float form inside a time step: the RK coefficient update of this
      // present phase (slopes of the coefficients and rates scaled by
      // |Lim_u|, then the stage combination a0 U0 + a1 (U + dt L)) in place
      // of k_update_coef's sweep.  The rates just stored are read back from
      // the cache one variable at a time (holding them in registers
      // through the update spills).  The stored rates stay unscaled:
      // Update_grid reads only the means' rates, which the scaling leaves
      // unchanged.
      const double dt = sc->dt;
      const double a0 = md == 0 ? 0. : md == 1 ? 0.5 : md == 2 ? 0.75 : 1. / 3.;
      const double a1 = md == 0 ? 1. : md == 1 ? 0.5 : md == 2 ? 0.25 : 2. / 3.;
#pragma unroll 1
      for (int u = 0; u < U; u++) {
        const double l = fabs(C.lim[(ss * U + u) * nc + c]);
#pragma unroll
        for (int m = 0; m < M; m++) {
          const size_t o = (((size_t)ss * U + u) * M + m) * nc + c;
          double dcf = C.dcoef[o], cf = C.coef[o];
          if (m >= 1) {
            dcf *= l;
            cf *= l;
          }
          double nw = a1 * (cf + dcf * dt);
          if (md == 0) C.coef0[o] = cf;
          else nw += a0 * C.coef0[o];
          C.coef[o] = nw;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Mixed-form DG right-hand side on the matrix cores (3-D, P2 basis)
// ---------------------------------------------------------------------------
typedef float rhs_f32x4 __attribute__((ext_vector_type(4)));

// One wavefront takes 64 consecutive cells as four blocks of 16; within a
// block every step is a 16x16 matrix product over the block's cells
// (v_mfma_f32_16x16x4_f32, lane l holding row l % 16 / column l % 16 and
// reduction index l / 16 of the operands, and entries (4 (l / 16) + r,
// l % 16) of the result):
//   states  U(point, cell) = sum_m Phi(point, m) Cf(m, cell), Cf the
//           coefficients with the limiter applied to the slopes, Phi the
//           basis at the 27 volume Gauss points;
//   volume  R(m, cell) += sum_b sum_point w dPhi_m/dx_b(point) F_b(U(point));
//   faces   R(m, cell) += sum_point w S (-n) Phi_m(point) Fh(point) over the
//           54 face points, Fh the point fluxes of the face kernels
//           (interior faces: k_face_simple / k_face_general; domain-boundary
//           faces: k_face_boundary3), loaded one block of 16 points ahead.
// The volume products take the flux columns straight from the registers the
// state product left them in: the reduction index of reduction step t in
// lane l is point 4 (l / 16) + t, which is the point of result entry r = t
// of that lane, and the constant operand is laid out in that order.  The
// rates are then divided by the mass matrix and stored, and the stage's
// coefficient update applied, from the result layout (16 consecutive
// cells per mode).  The point fluxes are those of k_rhs<float>; the sums
// run in another order.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_rhs_mfma(Geo g, CellDev C, const int32_t* __restrict__ used,
                                                 const DevScalars* __restrict__ sc, int md) {
  constexpr int D = 3, U = 5, M = 10, NQ = 3;
  const int nc = g.nc;
  // constant operands of every lane position, shared by the workgroup's
  // wavefronts: [0, 6) states at the volume points (pb * 3 + s), [6, 30)
  // volume weights ((b * 2 + pb) * 4 + t), [30, 46) face weights (pb * 4 + t)
  __shared__ float s_A[46][64];
  const int lane = threadIdx.x & 63, grp = lane >> 4, col = lane & 15;
  const int cbase = ((int)xcd_block() * 4 + (int)(threadIdx.x >> 6)) * 64;
  const double hp[3] = {g.ps[0] / 2, g.ps[1] / 2, g.ps[2] / 2};
  const double vol64 = g.ps[0] * g.ps[1] * g.ps[2];
  const float voli = (float)vol64;
  // basis (1, tail) at offset r from the cell centre
  auto phi_at = [&](const float* r, int m) -> float {
    float t[M - 1];
    basis_tail<D, 2>(r, g.q12, t);
    float v = 1.f;
#pragma unroll
    for (int q = 0; q < M - 1; q++)
      if (m == q + 1) v = t[q];
    return m < M ? v : 0.f;
  };
  auto vol_point = [&](int p, float* r, float& gw) {  // p = (g1 * 3 + g2) * 3 + g3
    const int g1 = p / 9, g2 = (p / 3) % 3, g3 = p % 3;
    r[0] = (float)(g.Gp[g1] * hp[0]);
    r[1] = (float)(g.Gp[g2] * hp[1]);
    r[2] = (float)(g.Gp[g3] * hp[2]);
    gw = (float)(g.Gw[g1] * g.Gw[g2] * g.Gw[g3]);
  };
  // derivative of basis mode m along b at offset r
  auto dphi = [&](int m, int b, const float* r) -> float {
    if (m >= 1 && m <= 3) return m - 1 == b ? 1.f : 0.f;
    switch (m) {
      case 4: return b == 0 ? r[1] : (b == 1 ? r[0] : 0.f);
      case 5: return b == 0 ? r[2] : (b == 2 ? r[0] : 0.f);
      case 6: return b == 1 ? r[2] : (b == 2 ? r[1] : 0.f);
      case 7: return b == 0 ? 2.f * r[0] : 0.f;
      case 8: return b == 1 ? 2.f * r[1] : 0.f;
      case 9: return b == 2 ? 2.f * r[2] : 0.f;
      default: return 0.f;
    }
  };
  // constant operands: Phi(point pb * 16 + col, mode 4 s + grp); the
  // volume weight w dPhi_col/dx_b(point pb * 16 + 4 grp + t) vol; the face
  // weight w S (-n) Phi_col(face point pb * 16 + 4 grp + t)
  if (threadIdx.x < 64) {
#pragma unroll 1
    for (int pb = 0; pb < 2; pb++) {
      const int p = pb * 16 + col;
      float r[3], gw;
      vol_point(p < 27 ? p : 0, r, gw);
#pragma unroll
      for (int s = 0; s < 3; s++) s_A[pb * 3 + s][lane] = p < 27 ? phi_at(r, 4 * s + grp) : 0.f;
#pragma unroll 1
      for (int t = 0; t < 4; t++) {
        const int pp = pb * 16 + 4 * grp + t;
        float rr[3], w;
        vol_point(pp < 27 ? pp : 0, rr, w);
#pragma unroll
        for (int b = 0; b < 3; b++) s_A[6 + (b * 2 + pb) * 4 + t][lane] = pp < 27 ? (dphi(col, b, rr) * w) * voli : 0.f;
      }
    }
#pragma unroll 1
    for (int q = 0; q < 16; q++) {
      const int fp = (q >> 2) * 16 + 4 * grp + (q & 3);
      float r[3], wq;
      face_point_cell3<2>(g, fp < 54 ? fp : 0, r, wq);
      s_A[30 + q][lane] = fp < 54 ? phi_at(r, col) * wq : 0.f;
    }
  }
  __syncthreads();
  if (cbase >= nc) return;  // the whole wavefront (no barrier follows)
  // inverse diagonal mass matrix of the result rows m = 4 grp + r
  double minv[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int m = 4 * grp + r;
    double mij = 1.;
    if (m >= 1 && m <= 3) mij = hp[m - 1] * hp[m - 1] / 3.;
    else if (m >= 4 && m <= 6) {
      const double t = m == 4 ? hp[0] * hp[1] : (m == 5 ? hp[0] * hp[2] : hp[1] * hp[2]);
      mij = t * t / 9.;
    } else if (m >= 7 && m <= 9) {
      const double t = hp[m - 7] * hp[m - 7];
      mij = t * t * 4. / 45.;
    }
    minv[r] = 1. / (mij * vol64 + 1e-20);
  }
  const double dt = sc->dt;
  const double a0 = md == 0 ? 0. : md == 1 ? 0.5 : md == 2 ? 0.75 : 1. / 3.;
  const double a1 = md == 0 ? 1. : md == 1 ? 0.5 : md == 2 ? 0.25 : 2. / 3.;

  for (int blk = 0; blk < 4; blk++) {
    const int c0 = cbase + blk * 16;
    if (c0 >= nc) break;  // uniform
    const int c = c0 + col;
    const bool in_range = c < nc;
    const bool use_c = in_range && used[c] != 0;
    int ci, cj, ck;
    cell_ijk(g, in_range ? c : c0, ci, cj, ck);
    for (int ss = 0; ss < PGDG_MAX_PHASE; ss++) {
      if (ss >= g.F) break;
      const bool valid = use_c && C.ph_state[ss * nc + c] == 1;
      if (__ballot(valid) == 0) continue;  // no cell of the block holds the phase
      const EosT<float> mt(Eos{g.gamma[ss], g.pinf[ss]});
      // coefficients: mode 4s + grp of cell col (a placeholder state, unit
      // density and energy, where the cell does not hold the phase)
      float Bv[U][3];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const float lm = valid ? (float)C.lim[(ss * U + u) * nc + c] : 0.f;
#pragma unroll
        for (int s = 0; s < 3; s++) {
          const int m = 4 * s + grp;
          float v = (m == 0 && (u == 0 || u == U - 1)) ? 1.f : 0.f;
          if (valid && m < M) {
            const float cf = (float)C.coef[(((size_t)ss * U + u) * M + m) * nc + c];
            v = m == 0 ? cf : lm * cf;
          }
          Bv[u][s] = v;
        }
      }
      rhs_f32x4 R[U];
#pragma unroll
      for (int u = 0; u < U; u++) R[u] = rhs_f32x4{0.f, 0.f, 0.f, 0.f};
      // face integral: the point fluxes of face block pb (points
      // pb * 16 + 4 grp + t), interior faces from the face kernels' C.fflux,
      // domain-boundary faces from k_face_boundary3's C.bflux
      auto load_faces = [&](int pb, float (&Fo)[4][U]) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int fp = pb * 16 + 4 * grp + t;
#pragma unroll
          for (int u = 0; u < U; u++) Fo[t][u] = 0.f;
          if (fp < 54 && valid) {
            const int f = fp / (NQ * NQ), q = fp % (NQ * NQ);
            int ax;
            float sgn;
            side_of(f, ax, sgn);
            const int st = ax == 0 ? g.N[1] * g.N[2] : (ax == 1 ? g.N[2] : 1);
            const int ia = ax == 0 ? ci : (ax == 1 ? cj : ck);
            const int na = ia + (sgn > 0 ? 1 : -1);
            if (na >= 0 && na < g.N[ax]) {
              const int fc = sgn > 0 ? c : c - st;
              const bool one = C.fkind[ax * nc + fc] != 0;
              const int side = one ? 0 : (sgn > 0 ? 0 : 1);
              const float wo = one ? (float)C.alph[ss * nc + (sgn > 0 ? c + st : c - st)] : 1.f;
#pragma unroll
              for (int u = 0; u < U; u++) Fo[t][u] = C.fflux[face_slot<D, 2>(g, side, ss, q, u, ax, fc)] * wo;
            } else if (g.bc[f] != 0) {
              const int pidx = bplane_index(g, f, ci, cj, ck);
#pragma unroll
              for (int u = 0; u < U; u++) Fo[t][u] = C.bflux[bflux_slot<2>(g, ss, f, q, u, pidx)];
            }
          }
        }
      };
      // volume integral
#pragma unroll 1
      for (int pb = 0; pb < 2; pb++) {
        rhs_f32x4 Up[U];
        const float Ae0 = s_A[pb * 3][lane], Ae1 = s_A[pb * 3 + 1][lane], Ae2 = s_A[pb * 3 + 2][lane];
#pragma unroll
        for (int u = 0; u < U; u++) {
          rhs_f32x4 a = {0.f, 0.f, 0.f, 0.f};
          a = __builtin_amdgcn_mfma_f32_16x16x4f32(Ae0, Bv[u][0], a, 0, 0, 0);
          a = __builtin_amdgcn_mfma_f32_16x16x4f32(Ae1, Bv[u][1], a, 0, 0, 0);
          a = __builtin_amdgcn_mfma_f32_16x16x4f32(Ae2, Bv[u][2], a, 0, 0, 0);
          Up[u] = a;
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int p = pb * 16 + 4 * grp + t;
          float colv[3][U];
          if (p < 27) {
            const float rho = Up[0][t], E = Up[U - 1][t];
            float rv[3], v[3];
#pragma unroll
            for (int a = 0; a < D; a++) rv[a] = Up[1 + a][t];
#pragma unroll
            for (int a = 0; a < D; a++) v[a] = qdiv(rv[a], rho);
            float sq = 0;
#pragma unroll
            for (int a = 0; a < D; a++) sq += rv[a] * rv[a];
            const float P = mt.p(E - qdiv(0.5f * sq, rho));
#pragma unroll
            for (int b = 0; b < D; b++) flux_col<D>(rho, rv, E, v, P, b, colv[b]);
          } else {
#pragma unroll
            for (int b = 0; b < D; b++)
#pragma unroll
              for (int u = 0; u < U; u++) colv[b][u] = 0.f;
          }
#pragma unroll
          for (int b = 0; b < D; b++) {
            const float Ap = s_A[6 + (b * 2 + pb) * 4 + t][lane];
#pragma unroll
            for (int u = 0; u < U; u++) R[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ap, colv[b][u], R[u], 0, 0, 0);
          }
        }
      }
#pragma unroll 1
      for (int pb = 0; pb < 4; pb++) {
        float Fc[4][U];
        load_faces(pb, Fc);
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const float Af = s_A[30 + pb * 4 + t][lane];
#pragma unroll
          for (int u = 0; u < U; u++) R[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af, Fc[t][u], R[u], 0, 0, 0);
        }
      }
      // rates of modes 4 grp + r of cell col; then the stage's update
      if (valid) {
#pragma unroll 1
        for (int u = 0; u < U; u++) {
          const double l = md >= 0 ? fabs(C.lim[(ss * U + u) * nc + c]) : 0.;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int m = 4 * grp + r;
            if (m >= M) continue;
            const size_t o = (((size_t)ss * U + u) * M + m) * nc + c;
            double dcf = (double)R[u][r] * minv[r];
            C.dcoef[o] = dcf;
            if (md >= 0) {
              double cf = C.coef[o];
              if (m >= 1) {
                dcf *= l;
                cf *= l;
              }
              double nw = a1 * (cf + dcf * dt);
              if (md == 0) C.coef0[o] = cf;
              else nw += a0 * C.coef0[o];
              C.coef[o] = nw;
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Update_grid (S3:mpm.cpp:768-830; T2:mpm.cpp:1254-1320)
// ---------------------------------------------------------------------------
// md: 0 first stage (Euler), 1 Heun second stage (reference), 2 and 3 the
// second and third SSP-RK3 stages.
//
// Two kernels: k_update_coef streams every coefficient (one lane per phase,
// variable, mode and cell: slopes of coef and dcoef scaled by |Lim_u|, then
// the stage combination for present phases), k_update_grid then derives the
// cell state from the updated mode-0 coefficients.  Each value goes through
// the same operations as in the reference's single loop.
template <int D, int ORD>
__global__ void __launch_bounds__(256) k_update_coef(Geo g, CellDev C, const int32_t* __restrict__ used,
                                                     const DevScalars* __restrict__ sc, int md) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M;
  const int nc = g.nc;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.F * U * M * nc) return;
  const int c = (int)(t % nc);
  const int r = (int)(t / nc);  // (s * U + u) * M + m
  const int m = r % M, su = r / M, s = su / U;
  if (!used[c]) return;
  // mixed form: an absent phase's coefficients are zero (P2G clears them)
  // and its rates are read by nothing before the RHS rewrites them, so the
  // scaling by |Lim_u| is skipped there (the FP64 form keeps the
  // reference's full sweep)
  if (g.prec == PGDG_PREC_MIXED && C.ph_state[s * nc + c] != 1) return;
  double cf = C.coef[t], dcf = C.dcoef[t];
  if (m >= 1) {
    const double l = fabs(C.lim[(size_t)su * nc + c]);
    dcf *= l;
    cf *= l;
  }
  if (C.ph_state[s * nc + c] == 1) {
    const double dt = sc->dt;
    if (md == 0) {
      C.coef0[t] = cf;
      cf += dcf * dt;
    } else if (md == 1) {
      cf = 0.5 * (dcf * dt + cf + C.coef0[t]);
    } else if (md == 2) {
      cf = 0.75 * C.coef0[t] + 0.25 * (cf + dcf * dt);
    } else {
      cf = (1. / 3.) * C.coef0[t] + (2. / 3.) * (cf + dcf * dt);
    }
  }
  C.coef[t] = cf;
  C.dcoef[t] = dcf;
}

template <int D, int ORD>
__global__ void __launch_bounds__(256) k_update_grid(Geo g, CellDev C, const int32_t* __restrict__ used,
                                                     const DevScalars* __restrict__ sc, int md) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M;
  (void)sc;
  (void)md;
  const int nc = g.nc;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc || !used[c]) return;
  const bool s3 = g.sem == PGDG_SEM_S3;
  double rva[3] = {0, 0, 0};
  double rho_a = 0, P_a = 0;
  double cP = C.P[c];
  double cv[3] = {C.v[c], C.v[nc + c], C.v[2 * nc + c]};
  double ca[3] = {C.a[c], C.a[nc + c], C.a[2 * nc + c]};
  for (int s = 0; s < PGDG_MAX_PHASE; s++) {
    if (s >= g.F) break;
    C.c[s * nc + c] = 0;
    C.ac[s * nc + c] = 0;
    double alph = C.alph[s * nc + c];
    double rho, rv[3], E;
    if (C.ph_state[s * nc + c] == 1) {
      rho = C.coef[CIDX(s, 0, 0)];
      rv[2] = C.rhov[(s * 3 + 2) * nc + c];
#pragma unroll
      for (int a = 0; a < D; a++) rv[a] = C.coef[CIDX(s, 1 + a, 0)];
      E = C.coef[CIDX(s, U - 1, 0)];
      C.rho[s * nc + c] = rho;
#pragma unroll
      for (int a = 0; a < D; a++) C.rhov[(s * 3 + a) * nc + c] = rv[a];
      C.E[s * nc + c] = E;
      double sq = 0;
#pragma unroll
      for (int a = 0; a < D; a++) sq += rv[a] * rv[a];
      double e_j = E - 0.5 * sq / rho;
      if (!s3)
#pragma unroll
        for (int a = 0; a < 3; a++) cv[a] = rv[a] / rho;
      cP = eos_p(e_j, Eos{g.gamma[s], g.pinf[s]});
      double cs = sqrt(g.gamma[s] * (g.pinf[s] + cP) / rho);
      C.c[s * nc + c] = cs;
      const double d0 = C.dcoef[CIDX(s, 0, 0)];
#pragma unroll
      for (int a = 0; a < D; a++) ca[a] = (C.dcoef[CIDX(s, 1 + a, 0)] - d0 * cv[a]) / rho;
      double sqa = 0;
#pragma unroll
      for (int a = 0; a < D; a++) sqa += ca[a] * ca[a];
      C.ac[s * nc + c] = sqrt(sqa);
    } else {
      rho = C.rho[s * nc + c];
#pragma unroll
      for (int a = 0; a < 3; a++) rv[a] = C.rhov[(s * 3 + a) * nc + c];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) rva[a] += rv[a] * alph;
    rho_a += rho * alph;
    P_a += cP * alph;
  }
#pragma unroll
  for (int a = 0; a < 3; a++) C.v[a * nc + c] = rva[a] / rho_a;
#pragma unroll
  for (int a = 0; a < 3; a++) C.a[a * nc + c] = ca[a];
  C.P[c] = P_a;
}

#undef CIDX

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int D, int ORD>
static int launch_cells_o(int which, const Geo& g, const CellDev& C, const int32_t* used,
                          const DevScalars* sc, int arg, hipStream_t st, int block, const IfaceBatch* ib) {
  const int nc = g.nc;
  switch (which) {
    case CELL_LIMITER:
      if (g.prec == PGDG_PREC_MIXED) {
        switch (block) {
          case 64: hipLaunchKernelGGL((k_limiter<D, ORD, 64, float>), dim3(grid_xcd(nc, 64)), dim3(64), 0, st, g, C, used); break;
          case 128: hipLaunchKernelGGL((k_limiter<D, ORD, 128, float>), dim3(grid_xcd(nc, 128)), dim3(128), 0, st, g, C, used); break;
          case 256: hipLaunchKernelGGL((k_limiter<D, ORD, 256, float>), dim3(grid_xcd(nc, 256)), dim3(256), 0, st, g, C, used); break;
          default: return PGDG_ERR_ARG;
        }
      } else {
        switch (block) {
          case 64: hipLaunchKernelGGL((k_limiter<D, ORD, 64, double>), dim3(grid_xcd(nc, 64)), dim3(64), 0, st, g, C, used); break;
          case 128: hipLaunchKernelGGL((k_limiter<D, ORD, 128, double>), dim3(grid_xcd(nc, 128)), dim3(128), 0, st, g, C, used); break;
          case 256: hipLaunchKernelGGL((k_limiter<D, ORD, 256, double>), dim3(grid_xcd(nc, 256)), dim3(256), 0, st, g, C, used); break;
          default: return PGDG_ERR_ARG;
        }
      }
      break;
    case CELL_RHS:
      if (g.prec == PGDG_PREC_MIXED) {
        PGDG_TRY(launch_faces<D>(g, C, used, st, ib));
        if (D == 3 && ORD == 2 && g.mfma_rhs) {
          hipLaunchKernelGGL(k_rhs_mfma, dim3(grid_xcd(((long long)nc + 63) / 64, 4)), dim3(256), 0, st, g, C,
                             used, sc, arg);
          break;
        }
        switch (block) {
          case 64: hipLaunchKernelGGL((k_rhs<D, ORD, 64, float>), dim3(grid_xcd(nc, 64)), dim3(64), 0, st, g, C, used, sc, arg); break;
          case 128: hipLaunchKernelGGL((k_rhs<D, ORD, 128, float>), dim3(grid_xcd(nc, 128)), dim3(128), 0, st, g, C, used, sc, arg); break;
          case 256: hipLaunchKernelGGL((k_rhs<D, ORD, 256, float>), dim3(grid_xcd(nc, 256)), dim3(256), 0, st, g, C, used, sc, arg); break;
          default: return PGDG_ERR_ARG;
        }
      } else {
        switch (block) {
          case 64: hipLaunchKernelGGL((k_rhs<D, ORD, 64, double>), dim3(grid_xcd(nc, 64)), dim3(64), 0, st, g, C, used, sc, -1); break;
          case 128: hipLaunchKernelGGL((k_rhs<D, ORD, 128, double>), dim3(grid_xcd(nc, 128)), dim3(128), 0, st, g, C, used, sc, -1); break;
          case 256: hipLaunchKernelGGL((k_rhs<D, ORD, 256, double>), dim3(grid_xcd(nc, 256)), dim3(256), 0, st, g, C, used, sc, -1); break;
          default: return PGDG_ERR_ARG;
        }
      }
      break;
    case CELL_UPDATE:
      // arg: the stage code, plus kCellSkipCoef when the RHS kernel has
      // already applied the coefficient update
      if (!(arg & kCellSkipCoef))
        hipLaunchKernelGGL((k_update_coef<D, ORD>),
                           dim3((unsigned)(((int64_t)g.F * (D + 2) * Basis<D, ORD>::M * nc + 255) / 256)),
                           dim3(256), 0, st, g, C, used, sc, arg & 3);
      hipLaunchKernelGGL((k_update_grid<D, ORD>), dim3((nc + 255) / 256), dim3(256), 0, st, g, C, used,
                         sc, arg & 3);
      break;
    default:
      return PGDG_ERR_ARG;  // P2G: launch_p2g
  }
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int launch_cells(int which, const Geo& g, const CellDev& C, const PtlDev& P, const int* bin_off,
                 const int* bin_ent, const int32_t* used, const DevScalars* sc, int arg,
                 hipStream_t st, int block, const IfaceBatch* ib) {
  (void)P;
  (void)bin_off;
  (void)bin_ent;
  if (g.ord == 2) return launch_cells_o<D, 2>(which, g, C, used, sc, arg, st, block, ib);
  return launch_cells_o<D, 1>(which, g, C, used, sc, arg, st, block, ib);
}

template <int D>
int launch_p2g(const Geo& g, const CellDev& C, const PtlDev& P, const int* bin_off, const int* bin_ent,
               int64_t n_ent, const int32_t* used, const P2gBuf& pb, hipStream_t st) {
  // every bin entry (the device-side total bin_off[nc]) is gathered; the
  // grid covers the buffer's capacity
  (void)n_ent;
  hipLaunchKernelGGL((k_p2g_gather<D>), dim3((unsigned)((pb.ne + 255) / 256)), dim3(256), 0, st, P, bin_ent,
                     bin_off + g.nc, pb);
  const dim3 gr(grid_xcd(g.nc, 128)), bl(128), ga(grid_xcd(g.nc, 64)), ba(64);
  const bool mix = g.prec == PGDG_PREC_MIXED;
#define PGDG_P2G(O, X)                                                                                 \
  do {                                                                                                \
    hipLaunchKernelGGL((k_p2g<D, O, X>), gr, bl, 0, st, g, C, pb, bin_off, used);                     \
    hipLaunchKernelGGL((k_p2g_appear<D, O, X>), ga, ba, 0, st, g, C, P, pb, bin_off, bin_ent, used);  \
  } while (0)
  if (g.ord == 2) {
    if (mix) PGDG_P2G(2, true);
    else PGDG_P2G(2, false);
  } else {
    if (mix) PGDG_P2G(1, true);
    else PGDG_P2G(1, false);
  }
#undef PGDG_P2G
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}
template int launch_p2g<2>(const Geo&, const CellDev&, const PtlDev&, const int*, const int*, int64_t,
                           const int32_t*, const P2gBuf&, hipStream_t);
template int launch_p2g<3>(const Geo&, const CellDev&, const PtlDev&, const int*, const int*, int64_t,
                           const int32_t*, const P2gBuf&, hipStream_t);

template int launch_cells<2>(int, const Geo&, const CellDev&, const PtlDev&, const int*, const int*,
                             const int32_t*, const DevScalars*, int, hipStream_t, int, const IfaceBatch*);
template int launch_cells<3>(int, const Geo&, const CellDev&, const PtlDev&, const int*, const int*,
                             const int32_t*, const DevScalars*, int, hipStream_t, int, const IfaceBatch*);

}  // namespace pgdg
