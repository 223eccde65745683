// cell_flux.hpp - point-wise pieces of the DG face and volume integrals
// (Particle_base::Grid_acceleration, S3:particle_flow.cpp:16-405): the
// stiffened-gas EOS in arithmetic T, the axis flux column, the HLLC flux of
// an interior face point with the MGFM ghost state at material interfaces,
// and the face quadrature points in a fixed (canonical) order.
#pragma once

#include "dev_common.hpp"

namespace pgdg {

// Stiffened-gas EOS in arithmetic T (the same operations as eos_p and
// eos_rhoe for T = double).
template <typename T>
struct EosT {
  T g, pinf, gm1, gpinf;
  __device__ explicit EosT(const Eos& m)
      : g((T)m.gamma), pinf((T)m.pinf), gm1((T)(m.gamma - 1)), gpinf((T)(m.gamma * m.pinf)) {}
  __device__ T p(T rhoe) const { return rhoe * gm1 - gpinf; }
  __device__ T rhoe(T p_) const { return (p_ + gpinf) / gm1; }
};

// Quotient and square root in arithmetic T: correctly rounded for double
// (the FP64 form's operations), the hardware reciprocal and square root for
// float (the mixed form; within 2 ulp).
template <typename T>
__device__ inline T qdiv(T a, T b) {
  if constexpr (sizeof(T) == 4) return a * __builtin_amdgcn_rcpf(b);
  else return a / b;
}
template <typename T>
__device__ inline T qrcp(T b) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(b);
  else return T(1) / b;
}
template <typename T>
__device__ inline T qsqrt(T a) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_sqrtf(a);
  else return sqrt(a);
}

// Flux column b of DF = U v^T + P I (+ v_b P in the energy row).  b is a
// runtime axis: it selects through comparisons rather than array indexing,
// so the operands stay in registers (a runtime index into a local array
// would place it in scratch memory).
template <int D, typename T>
__device__ inline T pick_axis(const T* v, int b) {
  return b == 0 ? v[0] : (b == 1 || D == 2 ? v[1] : v[2]);
}

template <int D, typename T>
__device__ inline void flux_col(T rho, const T* rv, T E, const T* v, T P, int b, T* out) {
  constexpr int U = D + 2;
  const T vb = pick_axis<D>(v, b);
  out[0] = rho * vb;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const T f = rv[a] * vb;
    out[1 + a] = a == b ? f + P : f;
  }
  out[U - 1] = E * vb + vb * P;
}

// Face quadrature point q of a face normal to axis `ax`, in the canonical
// order shared by the face and cell kernels: the tangential axes t0 < t1,
// q = g1 * NQ + g2 (2-D: q = g1).  off: offset from the face centre; gw: the
// product of the 1-D weights.
template <int D, int ORD>
__device__ inline void face_point(const Geo& g, int ax, int q, const double* hp, double* off, double& gw) {
  constexpr int NQ = Basis<D, ORD>::NQ;
  const int t0 = ax == 0 ? 1 : 0;
  const int t1 = ax == 2 ? 1 : 2;
  off[0] = off[1] = off[2] = 0;
  if (D == 3) {
    const int g1 = q / NQ, g2 = q % NQ;
    off[t0] = g.Gp[g1] * hp[t0];
    off[t1] = g.Gp[g2] * hp[t1];
    gw = g.Gw[g1] * g.Gw[g2];
  } else {
    off[t0] = g.Gp[q] * hp[t0];
    gw = g.Gw[q];
  }
}

// Slot of an interior-face point flux (mixed form).  Face (ax, c) lies
// between cell c and its +ax neighbour; side 0 holds the flux seen from c
// (its +ax face), side 1 the flux seen from the neighbour (its -ax face);
// ss is the phase of that side's cell, q the canonical point, u the
// variable.  Consecutive c are contiguous.
template <int D, int ORD>
__device__ inline size_t face_slot(const Geo& g, int side, int ss, int q, int u, int ax, int c) {
  constexpr int U = D + 2, NQ = Basis<D, ORD>::NQ, NQF = D == 3 ? NQ * NQ : NQ;
  return ((((size_t)side * g.F + ss) * NQF + q) * U + u) * ((size_t)D * g.nc) + (size_t)ax * g.nc + c;
}

// HLLC flux at one interior face point (the per-point body of the
// reference's face integral): cell i's phase ss (state Ui) against the
// neighbour's phase s1 (state Uj), normal sgn * e_ax.  For ss != s1 the
// neighbour state is replaced by the MGFM ghost state in phase ss's EOS
// (Newton solve in FP64); kSamePhase: the caller guarantees ss == s1 (no
// ghost state code).  Fh: the flux the reference multiplies by the basis,
// the weights and alpha_j.  Used by the mixed form only (T = float, with the
// hardware reciprocal and square root); the FP64 form keeps the reference's
// operations inside its cell kernel.
// The MGFM solve of a ghost state, in place (the reference's call and
// initial guess).
struct MgfmInline {
  static constexpr bool kInputsOnly = false;
  __device__ MgfmOut operator()(double ul, double ur, double pl, double pr, double rhol, double rhor, Eos ml,
                                Eos mr, int variant) const {
    return mgfm_solve(ul, ur, pl, pr, rhol, rhor, ml, mr, variant, TOL_REF, 0.0, 50, -1.0);
  }
};

// mg: the MGFM solve of the ghost state (MgfmInline, or a batched solve's
// record); with MG::kInputsOnly the flux stops after handing mg the solve's
// inputs (Fh is left unset).
template <int D, typename T, bool kSamePhase = false, typename MG = MgfmInline>
__device__ inline void iface_flux(const Geo& g, const T* Ui, const T* Uj, int ss, int s1, int ax, T sgn,
                                  T* Fh, MG mg = MG{}) {
  constexpr int U = D + 2;
  const bool s3 = g.sem == PGDG_SEM_S3;
  const Eos me{g.gamma[ss], g.pinf[ss]};
  const EosT<T> mt(me);
  T nrm[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 3; a++)
    if (a == ax) nrm[a] = sgn;
  const T rho_i = Ui[0], E_i = Ui[U - 1];
  const T ir_i = qrcp(rho_i);
  T rvi[3] = {0, 0, 0}, vi[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) rvi[a] = Ui[1 + a];
#pragma unroll
  for (int a = 0; a < D; a++) vi[a] = rvi[a] * ir_i;
  T sq = 0;
#pragma unroll
  for (int a = 0; a < D; a++) sq += rvi[a] * rvi[a];
  const T e_i = E_i - T(0.5) * sq * ir_i;
  const T P_i = mt.p(e_i);
  const T ci = qsqrt(mt.g * (mt.pinf + P_i) * ir_i);
  T vni = 0;
#pragma unroll
  for (int a = 0; a < D; a++) vni += vi[a] * nrm[a];
  T DFi[U];
  flux_col<D>(rho_i, rvi, E_i, vi, P_i, ax, DFi);
  T rho_j = Uj[0], E_j = Uj[U - 1];
  T rvj[3] = {0, 0, 0}, vj[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) rvj[a] = Uj[1 + a];
  const T ir_j = qrcp(rho_j);
#pragma unroll
  for (int a = 0; a < D; a++) vj[a] = rvj[a] * ir_j;
  T sqj = 0;
#pragma unroll
  for (int a = 0; a < D; a++) sqj += vj[a] * vj[a];
  T rhoe_j = E_j - T(0.5) * sqj * rho_j;
  T P_j = EosT<T>(Eos{g.gamma[s1], g.pinf[s1]}).p(rhoe_j);
  T vnj = 0;
#pragma unroll
  for (int a = 0; a < D; a++) vnj += vj[a] * nrm[a];
  if (!kSamePhase && ss != s1) {
    MgfmOut o = mg(vni, vnj, P_i, P_j, rho_i, rho_j, me, Eos{g.gamma[s1], g.pinf[s1]}, s3 ? MGFM_S3 : MGFM_T2);
    if constexpr (MG::kInputsOnly) return;
    rho_j = o.rho_l;
    P_j = o.pstar;
    vnj = o.u_l;
#pragma unroll
    for (int a = 0; a < D; a++) vj[a] = vi[a] - vni * nrm[a] + vnj * nrm[a];
#pragma unroll
    for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
    rhoe_j = qdiv(P_j + mt.gpinf, mt.gm1);
    T sq2 = 0;
#pragma unroll
    for (int a = 0; a < D; a++) sq2 += vj[a] * vj[a];
    E_j = rhoe_j + T(0.5) * rho_j * sq2;
  }
  const T cj = qsqrt(qdiv(mt.g * (mt.pinf + P_j), rho_j));
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
  const T SL = amin1(vni - ci, vnj - cj);
  const T SR = amax1(vni + ci, vnj + cj);
  const T S_s = qdiv((P_j - P_i) + rho_i * vni * (SL - vni) - rho_j * vnj * (SR - vnj),
                     rho_i * (SL - vni) - rho_j * (SR - vnj));
  if (SL > 0) {
#pragma unroll
    for (int u = 0; u < U; u++) Fh[u] = DFi[u];
  } else if (SL <= 0 && S_s >= 0) {
    const T rho_sL = qdiv(rho_i * (SL - vni), SL - S_s);
    T UsL[U];
    UsL[0] = rho_sL;
#pragma unroll
    for (int a = 0; a < D; a++) UsL[1 + a] = rho_sL * (vi[a] + (S_s - vni) * nrm[a]);
    UsL[U - 1] = rho_sL * (E_i * ir_i + (S_s - vni) * (S_s + qdiv(P_i, rho_i * (SL - vni))));
#pragma unroll
    for (int u = 0; u < U; u++) Fh[u] = DFi[u] + SL * (UsL[u] - Ui[u]) * sgn;
  } else if (S_s <= 0 && SR >= 0) {
    const T rho_sR = qdiv(rho_j * (SR - vnj), SR - S_s);
    T UsR[U];
    UsR[0] = rho_sR;
#pragma unroll
    for (int a = 0; a < D; a++) UsR[1 + a] = rho_sR * (vj[a] + (S_s - vnj) * nrm[a]);
    UsR[U - 1] = rho_sR * (qdiv(E_j, rho_j) + (S_s - vnj) * (S_s + qdiv(P_j, rho_j * (SR - vnj))));
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

// Flux of a domain-boundary face point in the float form: the wall
// pressure, or HLL against the inlet / outlet / symmetry ghost state (the
// boundary branches of the face integral, S3:particle_flow.cpp:290-394).
// (j, k): the cell's indices, for the symmetry reflections.
__device__ inline void bface_flux3(const Geo& g, const EosT<float>& mt, int fb, int j, int k, int ax, float sgn,
                                   const float* Ui, float* Fh) {
  constexpr int D = 3, U = 5;
  const bool wall_like = g.sem == PGDG_SEM_S3 ? (fb == 1) : (fb != 0);
  float nrm[3];
#pragma unroll
  for (int a = 0; a < 3; a++) nrm[a] = a == ax ? sgn : 0.f;
  const float rho_i = Ui[0], E_i = Ui[U - 1];
  float rvi[3], vi[3];
#pragma unroll
  for (int a = 0; a < D; a++) rvi[a] = Ui[1 + a];
#pragma unroll
  for (int a = 0; a < D; a++) vi[a] = qdiv(rvi[a], rho_i);
  float sq = 0;
#pragma unroll
  for (int a = 0; a < D; a++) sq += rvi[a] * rvi[a];
  const float e_i = E_i - qdiv(0.5f * sq, rho_i);
  const float P_i = mt.p(e_i);
  if (wall_like) {
    Fh[0] = Fh[U - 1] = 0.f;
#pragma unroll
    for (int a = 0; a < D; a++) Fh[1 + a] = a == ax ? P_i : 0.f;
    return;
  }
  const float ci = qsqrt(qdiv(mt.g * (mt.pinf + P_i), rho_i));
  float vni = 0;
#pragma unroll
  for (int a = 0; a < D; a++) vni += vi[a] * nrm[a];
  float DFi[U];
  flux_col<D>(rho_i, rvi, E_i, vi, P_i, ax, DFi);
  float rho_j = rho_i, P_j = P_i, E_j = E_i;
  float vj[3], rvj[3];
#pragma unroll
  for (int a = 0; a < D; a++) {
    vj[a] = vi[a];
    rvj[a] = rho_j * vj[a];
  }
  if (fb == -1) {
    rho_j = (float)g.inlet_rho;
#pragma unroll
    for (int a = 0; a < D; a++) vj[a] = (float)g.inlet_u[a];
#pragma unroll
    for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
    P_j = (float)g.inlet_p;
    float sqv = 0;
#pragma unroll
    for (int a = 0; a < D; a++) sqv += vj[a] * vj[a];
    E_j = mt.rhoe(P_j) + 0.5f * sqv * rho_j;
  }
  if (j == g.N[1] - 1 || j == 0) {
    vj[1] = -vi[1];
#pragma unroll
    for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
  }
  if (k == g.N[2] - 1 || k == 0) {
    vj[2] = -vi[2];
#pragma unroll
    for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
  }
  float Ujs[U];
  Ujs[0] = rho_j;
#pragma unroll
  for (int a = 0; a < D; a++) Ujs[1 + a] = rvj[a];
  Ujs[U - 1] = E_j;
  const float cj = qsqrt(qdiv(mt.g * (mt.pinf + P_j), rho_j));
  float vnj = 0;
#pragma unroll
  for (int a = 0; a < D; a++) vnj += vj[a] * nrm[a];
  float DFj[U];
  flux_col<D>(rho_j, rvj, E_j, vj, P_j, ax, DFj);
  const float SL = amin1(vni - ci, vnj - cj);
  const float SR = amax1(vni + ci, vnj + cj);
  if (SL > 0) {
#pragma unroll
    for (int u = 0; u < U; u++) Fh[u] = DFi[u];
  } else if (SL <= 0 && SR >= 0) {
#pragma unroll
    for (int u = 0; u < U; u++) Fh[u] = qdiv(SR * DFi[u] - SL * DFj[u] + SL * SR * (Ujs[u] - Ui[u]) * sgn, SR - SL);
  } else if (SR < 0) {
#pragma unroll
    for (int u = 0; u < U; u++) Fh[u] = DFj[u];
  } else {
#pragma unroll
    for (int u = 0; u < U; u++) Fh[u] = 0;
  }
}

// Face side f of a cell (the reference order yL, xR, yR, xL, zL, zR):
// normal axis and sign, and the neighbour's index offsets.
__device__ inline void side_of(int f, int& ax, float& sgn) {
  ax = (f == 1 || f == 3) ? 0 : (f < 4 ? 1 : 2);
  sgn = (f == 1 || f == 2 || f == 5) ? 1.f : -1.f;
}

// Face point fp = f * NQ^2 + q of a 3-D cell (side f, canonical point q):
// offset r from the cell centre and the weight w S (-n) of the point in the
// face integral.
template <int ORD>
__device__ inline void face_point_cell3(const Geo& g, int fp, float* r, float& wq) {
  constexpr int NQ = ORD + 1;
  const int f = fp / (NQ * NQ), q = fp % (NQ * NQ);
  int ax;
  float sgn;
  side_of(f, ax, sgn);
  const double hp[3] = {g.ps[0] / 2, g.ps[1] / 2, g.ps[2] / 2};
  double off[3], gwd;
  face_point<3, ORD>(g, ax, q, hp, off, gwd);
  off[ax] += (double)sgn * hp[ax];
#pragma unroll
  for (int a = 0; a < 3; a++) r[a] = (float)off[a];
  const double S = ax == 0 ? g.ps[1] * g.ps[2] : (ax == 1 ? g.ps[0] * g.ps[2] : g.ps[0] * g.ps[1]);
  wq = (float)gwd * (float)(-(double)sgn * S);
}

// Boundary-face point fluxes of the mixed form, 3-D: plane index of side f
// of cell (i, j, k) (the two tangential cell indices), the plane stride
// (the largest of the three plane sizes) and the slot of phase ss, side f,
// point q, variable u.
__device__ inline int bplane_index(const Geo& g, int f, int i, int j, int k) {
  return (f == 1 || f == 3) ? j * g.N[2] + k : (f < 4 ? i * g.N[2] + k : i * g.N[1] + j);
}
__device__ inline int bplane_stride(const Geo& g) {
  const int a = g.N[1] * g.N[2], b = g.N[0] * g.N[2], c = g.N[0] * g.N[1];
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}
template <int ORD>
__device__ inline size_t bflux_slot(const Geo& g, int ss, int f, int q, int u, int pidx) {
  constexpr int NQF = (ORD + 1) * (ORD + 1);
  return ((((size_t)ss * 6 + f) * NQF + q) * 5 + u) * (size_t)bplane_stride(g) + pidx;
}

}  // namespace pgdg
