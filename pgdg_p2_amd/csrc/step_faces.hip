// step_faces.hip - interior-face fluxes of the mixed-precision DG
// right-hand side (the face integral of Particle_base::Grid_acceleration,
// S3:particle_flow.cpp:16-405, evaluated once per face).
//
// The reference (and the FP64 form) evaluates every interior face twice,
// once from each of its cells.  Where both cells hold the same single phase
// the two HLLC evaluations are the same flux (HLLC is conservative: with
// the roles of the two states and the normal exchanged, SL, SR and S* swap
// and change sign and the same branch formula results), so the mixed form
// evaluates it once per face point and stores it for both cells, each
// weighted by the other cell's volume fraction.  Faces with several phases
// or a material interface (MGFM ghost states, one-sided) are listed and
// evaluated per side by a second kernel.  The cell kernel (k_rhs, float)
// projects the stored point fluxes onto its modes.
#include <algorithm>

#include "cell_flux.hpp"
#include "dev_common.hpp"
#include "hip_util.hpp"
#include "step_kernels.hpp"

namespace pgdg {

__global__ void k_face_clear(int32_t* count, int32_t* mg_count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *count = 0;
    if (mg_count) *mg_count = 0;
  }
}

// Phase masks of a cell: present (ph_state 1) and nonzero volume fraction.
__device__ inline void phase_masks(const Geo& g, const CellDev& C, int c, unsigned& pm, unsigned& am) {
  pm = am = 0;
  for (int s = 0; s < PGDG_MAX_PHASE; s++) {
    if (s >= g.F) break;
    pm |= (C.ph_state[s * g.nc + c] == 1 ? 1u : 0u) << s;
    am |= (C.alph[s * g.nc + c] != 0 ? 1u : 0u) << s;
  }
}

// One lane per cell c, its faces towards +x, +y(, +z).  Same-single-phase
// faces: one HLLC per point, stored once for both sides (fkind 1; each
// cell's kernel weights it by the other cell's volume fraction); other
// faces: appended to the general list (fkind 0).  Faces between two unused
// cells are skipped.
template <int D, int ORD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_face_simple(Geo g, CellDev C, const int32_t* __restrict__ used) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M, B = M - 1, NQ = Basis<D, ORD>::NQ;
  constexpr int NQF = D == 3 ? NQ * NQ : NQ;
  const int nc = g.nc;
  const int c = xcd_block() * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int ijk[3];
  cell_ijk(g, c, ijk[0], ijk[1], ijk[2]);
  const int stride[3] = {g.N[1] * g.N[2], g.N[2], 1};
  const bool uc = used[c] != 0;
  const double hp[3] = {g.ps[0] / 2, g.ps[1] / 2, g.ps[2] / 2};
  unsigned pc, ac;
  phase_masks(g, C, c, pc, ac);
  // c's coefficients of its single phase, loaded once for its three faces
  // (the phase's rows start at base pointers, so the row offsets are
  // uniform multiples of nc)
  const bool one = pc == ac && __popc(pc) == 1;
  const int s = one ? __ffs(pc) - 1 : 0;
  const size_t rows = (size_t)s * U * M;
  const size_t ncz = (size_t)nc;
  float cl[U][M], ll[U];
  if (one) {
    const double* pl = C.coef + rows * ncz + c;
#pragma unroll
    for (int u = 0; u < U; u++) {
      ll[u] = (float)C.lim[((size_t)s * U + u) * ncz + c];
#pragma unroll
      for (int m = 0; m < M; m++) cl[u][m] = (float)pl[(size_t)(u * M + m) * ncz];
    }
  }
#pragma unroll 1
  for (int ax = 0; ax < D; ax++) {
    if (ijk[ax] == g.N[ax] - 1) continue;  // domain boundary: the cell kernel's boundary flux
    const int nb = c + stride[ax];
    const bool un = used[nb] != 0;
    if (!uc && !un) continue;
    unsigned pn, an;
    phase_masks(g, C, nb, pn, an);
    if (!(one && pn == an && pn == pc)) {
      C.fkind[ax * nc + c] = 0;
      C.flist[atomicAdd(C.fcount, 1)] = ax * nc + c;
      continue;
    }
    C.fkind[ax * nc + c] = 1;
    float cr[U][M], lr[U];
    const double* pr = C.coef + rows * ncz + nb;
#pragma unroll
    for (int u = 0; u < U; u++) {
      lr[u] = (float)C.lim[((size_t)s * U + u) * ncz + nb];
#pragma unroll
      for (int m = 0; m < M; m++) cr[u][m] = (float)pr[(size_t)(u * M + m) * ncz];
    }
#pragma unroll 1
    for (int q = 0; q < NQF; q++) {
      double off[3], gwd;
      face_point<D, ORD>(g, ax, q, hp, off, gwd);
      float dl[3] = {(float)off[0], (float)off[1], (float)off[2]}, dr[3] = {dl[0], dl[1], dl[2]};
      dl[ax] += (float)hp[ax];
      dr[ax] -= (float)hp[ax];
      float pl64[B], pr64[B];
      basis_tail<D, ORD>(dl, g.q12, pl64);
      basis_tail<D, ORD>(dr, g.q12, pr64);
      float UL[U], UR[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        float al = 0, ar = 0;
#pragma unroll
        for (int b = 0; b < B; b++) {
          al += cl[u][1 + b] * (float)pl64[b];
          ar += cr[u][1 + b] * (float)pr64[b];
        }
        UL[u] = cl[u][0] + ll[u] * al;
        UR[u] = cr[u][0] + lr[u] * ar;
      }
      float Fh[U];
      iface_flux<D, float, true>(g, UL, UR, s, s, ax, 1.f, Fh);
#pragma unroll
      for (int u = 0; u < U; u++) C.fflux[face_slot<D, ORD>(g, 0, s, q, u, ax, c)] = Fh[u];
    }
  }
}

// The MGFM inputs of one interface point into record r.
struct MgfmCollect {
  static constexpr bool kInputsOnly = true;
  double* in;
  int32_t* pair;
  int64_t ld, r;
  int code;
  __device__ MgfmOut operator()(double ul, double ur, double pl, double pr, double rl, double rr, Eos, Eos,
                                int) const {
    in[r] = ul;
    in[ld + r] = ur;
    in[2 * ld + r] = pl;
    in[3 * ld + r] = pr;
    in[4 * ld + r] = rl;
    in[5 * ld + r] = rr;
    pair[r] = code;
    return MgfmOut{rl, pl, ul, 0};
  }
};

// The batched solve's solution of record r.
struct MgfmRecord {
  static constexpr bool kInputsOnly = false;
  const double* out;
  const int32_t* it;
  int64_t ld, r;
  __device__ MgfmOut operator()(double, double, double, double, double, double, Eos, Eos, int) const {
    return MgfmOut{out[r], out[ld + r], out[2 * ld + r], it[r]};
  }
};

// How k_face_general meets the MGFM solves of its interface points:
// FG_INLINE solves each in place; FG_COLLECT reserves the face's records
// and writes the solves' inputs there (no fluxes); FG_APPLY takes the
// solutions from the records (after the batch kernel has run).
enum { FG_INLINE = 0, FG_COLLECT = 1, FG_APPLY = 2 };

// The listed faces, each side as its cell sees it: for every present phase
// ss of the side's cell, the sum over the other cell's phases s1 (alpha != 0)
// of alpha_s1 * flux(ss, s1), with the MGFM ghost state where s1 != ss.  The
// two cells' coefficient sets are loaded once per phase into registers; the
// sum over s1 accumulates in the output slots (this lane's own, in s1 order).
// A face's MGFM records follow its traversal order (side, ss, s1, point).
template <int D, int ORD, int MODE>
__global__ void __launch_bounds__(128) k_face_general(Geo g, CellDev C, const int32_t* __restrict__ used) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M, B = M - 1, NQ = Basis<D, ORD>::NQ;
  constexpr int NQF = D == 3 ? NQ * NQ : NQ;
  const int nc = g.nc;
  const size_t ncz = (size_t)nc;
  const int n = *C.fcount;
  const int stride[3] = {g.N[1] * g.N[2], g.N[2], 1};
  const double hp[3] = {g.ps[0] / 2, g.ps[1] / 2, g.ps[2] / 2};
  // the phase-s coefficient set of cell c (FP32)
  auto load_set = [&](int c, int s, float (*cf)[M], float* lm) {
    const double* pc = C.coef + (size_t)s * U * M * ncz + c;
#pragma unroll
    for (int u = 0; u < U; u++) {
      lm[u] = (float)C.lim[((size_t)s * U + u) * ncz + c];
#pragma unroll
      for (int m = 0; m < M; m++) cf[u][m] = (float)pc[(size_t)(u * M + m) * ncz];
    }
  };
  auto eval = [&](const float (*cf)[M], const float* lm, const float* ph, float* Uo) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      float acc = 0;
#pragma unroll
      for (int b = 0; b < B; b++) acc += cf[u][1 + b] * ph[b];
      Uo[u] = cf[u][0] + lm[u] * acc;
    }
  };
  const int64_t ld = C.mg_cap;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int id = C.flist[t];
    const int ax = id / nc, c = id % nc;
    const int nb = c + stride[ax];
    int64_t rec = -1;  // this face's next MGFM record (-1: solved in place)
    if constexpr (MODE == FG_COLLECT) {
      int nrec = 0;
      for (int side = 0; side < 2; side++) {
        const int ci = side == 0 ? c : nb, cj = side == 0 ? nb : c;
        if (!used[ci]) continue;
        for (int ss = 0; ss < PGDG_MAX_PHASE && ss < g.F; ss++) {
          if (C.ph_state[ss * nc + ci] != 1) continue;
          for (int s1 = 0; s1 < PGDG_MAX_PHASE && s1 < g.F; s1++)
            if (s1 != ss && C.alph[s1 * nc + cj] != 0) nrec += NQF;
        }
      }
      if (nrec > 0) {
        const int64_t base = atomicAdd(C.mg_count, nrec);
        atomicMax(C.mg_peak, (int32_t)(base + nrec));
        if (base + nrec <= ld) {
          rec = base;
        } else {
          // past the record capacity: this face solves its points in place;
          // its reserved records below the capacity hold a trivial solve
          for (int64_t r = base; r < ld; r++) {
            for (int q = 0; q < 6; q++) C.mg_in[q * ld + r] = q < 2 ? 0.0 : 1.0;
            C.mg_pair[r] = 0;
          }
        }
      }
      C.mg_base[t] = (int32_t)rec;
      if (rec < 0) continue;
    } else if constexpr (MODE == FG_APPLY) {
      rec = C.mg_base[t];
    }
    for (int side = 0; side < 2; side++) {
      const int ci = side == 0 ? c : nb, cj = side == 0 ? nb : c;
      if (!used[ci]) continue;
      const float sgn = side == 0 ? 1.f : -1.f;
      for (int ss = 0; ss < PGDG_MAX_PHASE; ss++) {
        if (ss >= g.F) break;
        if (C.ph_state[ss * nc + ci] != 1) continue;
        float cfi[U][M], lmi[U];
        load_set(ci, ss, cfi, lmi);
        bool first = true;
        for (int s1 = 0; s1 < PGDG_MAX_PHASE; s1++) {
          if (s1 >= g.F) break;
          const float alph_j = (float)C.alph[s1 * nc + cj];
          if (alph_j == 0) continue;
          if (MODE == FG_COLLECT && s1 == ss) continue;  // a same-phase pair holds no MGFM record
          float cfj[U][M], lmj[U];
          load_set(cj, s1, cfj, lmj);
#pragma unroll 1
          for (int q = 0; q < NQF; q++) {
            double off[3], gwd;
            face_point<D, ORD>(g, ax, q, hp, off, gwd);
            float di[3] = {(float)off[0], (float)off[1], (float)off[2]}, dj[3] = {di[0], di[1], di[2]};
            di[ax] += (float)(side == 0 ? hp[ax] : -hp[ax]);
            dj[ax] -= (float)(side == 0 ? hp[ax] : -hp[ax]);
            float phi[B], phj[B], Ui[U], Uj[U], Fh[U];
            basis_tail<D, ORD>(di, g.q12, phi);
            basis_tail<D, ORD>(dj, g.q12, phj);
            eval(cfi, lmi, phi, Ui);
            eval(cfj, lmj, phj, Uj);
            if constexpr (MODE == FG_COLLECT) {
              if (s1 != ss)
                iface_flux<D, float>(g, Ui, Uj, ss, s1, ax, sgn, Fh,
                                     MgfmCollect{C.mg_in, C.mg_pair, ld, rec++, ss * PGDG_MAX_PHASE + s1});
              continue;  // no fluxes in this pass
            } else if constexpr (MODE == FG_APPLY) {
              if (rec >= 0 && s1 != ss) {
                iface_flux<D, float>(g, Ui, Uj, ss, s1, ax, sgn, Fh, MgfmRecord{C.mg_out, C.mg_it, ld, rec++});
              } else {
                iface_flux<D, float>(g, Ui, Uj, ss, s1, ax, sgn, Fh);
              }
            } else {
              iface_flux<D, float>(g, Ui, Uj, ss, s1, ax, sgn, Fh);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
              float* o = &C.fflux[face_slot<D, ORD>(g, side, ss, q, u, ax, c)];
              *o = first ? Fh[u] * alph_j : *o + Fh[u] * alph_j;
            }
          }
          first = false;
        }
        if (MODE != FG_COLLECT && first)  // no phase of the other cell has a volume fraction
#pragma unroll 1
          for (int q = 0; q < NQF; q++)
#pragma unroll
            for (int u = 0; u < U; u++) C.fflux[face_slot<D, ORD>(g, side, ss, q, u, ax, c)] = 0.f;
      }
    }
  }
}

// Domain-boundary faces of the 3-D mixed form (for the matrix-core cell
// kernel): one lane per boundary face (side f, plane index), every present
// phase of its cell; the state at each face point from the cell's
// coefficients (slopes scaled by the limiter), the wall pressure or the HLL
// flux against the inlet / outlet / symmetry ghost state (bface_flux3), into
// C.bflux.
template <int ORD>
__global__ void __launch_bounds__(256) k_face_boundary3(Geo g, CellDev C, const int32_t* __restrict__ used) {
  constexpr int D = 3, U = 5, M = Basis<3, ORD>::M, NQF = (ORD + 1) * (ORD + 1);
  const int nc = g.nc;
  const int PL = bplane_stride(g);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)6 * PL) return;
  const int f = (int)(t / PL), pidx = (int)(t % PL);
  if (g.bc[f] == 0) return;  // this side of the domain has no boundary condition
  int i, j, k;
  if (f == 1 || f == 3) {
    if (pidx >= g.N[1] * g.N[2]) return;
    i = f == 1 ? g.N[0] - 1 : 0;
    j = pidx / g.N[2];
    k = pidx % g.N[2];
  } else if (f < 4) {
    if (pidx >= g.N[0] * g.N[2]) return;
    i = pidx / g.N[2];
    j = f == 2 ? g.N[1] - 1 : 0;
    k = pidx % g.N[2];
  } else {
    if (pidx >= g.N[0] * g.N[1]) return;
    i = pidx / g.N[1];
    j = pidx % g.N[1];
    k = f == 5 ? g.N[2] - 1 : 0;
  }
  const int c = (i * g.N[1] + j) * g.N[2] + k;
  if (!used[c]) return;
  int ax;
  float sgn;
  side_of(f, ax, sgn);
  for (int ss = 0; ss < PGDG_MAX_PHASE; ss++) {
    if (ss >= g.F) break;
    if (C.ph_state[ss * nc + c] != 1) continue;
    const EosT<float> mt(Eos{g.gamma[ss], g.pinf[ss]});
    float cf[U][M];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const float lm = (float)C.lim[(ss * U + u) * nc + c];
#pragma unroll
      for (int m = 0; m < M; m++) {
        const float v = (float)C.coef[(((size_t)ss * U + u) * M + m) * nc + c];
        cf[u][m] = m == 0 ? v : lm * v;
      }
    }
#pragma unroll 1
    for (int q = 0; q < NQF; q++) {
      float r[3], wq;
      face_point_cell3<ORD>(g, f * NQF + q, r, wq);
      float ph[M - 1];
      basis_tail<3, ORD>(r, g.q12, ph);
      float Ui[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        float a = cf[u][0];
#pragma unroll
        for (int b = 0; b < M - 1; b++) a = fmaf(cf[u][1 + b], ph[b], a);
        Ui[u] = a;
      }
      float Fh[U];
      bface_flux3(g, mt, g.bc[f], j, k, ax, sgn, Ui, Fh);
#pragma unroll
      for (int u = 0; u < U; u++) C.bflux[bflux_slot<ORD>(g, ss, f, q, u, pidx)] = Fh[u];
    }
  }
}

// The listed faces with their interface points' MGFM solved in place, or
// collected, solved as one batch (the reference's initial p* or the
// network's) and applied.
template <int D, int ORD>
static int launch_general(const Geo& g, const CellDev& C, const int32_t* used, hipStream_t st,
                          const IfaceBatch* ib) {
  if (!ib || ib->mode == IFACE_INLINE) {
    hipLaunchKernelGGL((k_face_general<D, ORD, FG_INLINE>), dim3(1024), dim3(128), 0, st, g, C, used);
    return PGDG_OK;
  }
  if (!C.mg_in || !C.mg_count || C.mg_cap <= 0) return PGDG_ERR_STATE;
  if (ib->mode == IFACE_LEARNED && !ib->P) return PGDG_ERR_STATE;
  hipLaunchKernelGGL((k_face_general<D, ORD, FG_COLLECT>), dim3(1024), dim3(128), 0, st, g, C, used);
  PGDG_TRY(launch_mgfm_pairs(ib->mode == IFACE_LEARNED ? ib->P : nullptr, ib->nm, ib->tab, C.mg_in, C.mg_pair,
                             C.mg_out, C.mg_it, 0, C.mg_cap, C.mg_count, g.sem == PGDG_SEM_S3 ? MGFM_S3 : MGFM_T2,
                             TOL_REF, 0.0, 50, st));
  hipLaunchKernelGGL((k_face_general<D, ORD, FG_APPLY>), dim3(1024), dim3(128), 0, st, g, C, used);
  return PGDG_OK;
}

template <int D>
int launch_faces(const Geo& g, const CellDev& C, const int32_t* used, hipStream_t st, const IfaceBatch* ib) {
  if (!C.fflux || !C.flist || !C.fcount) return PGDG_ERR_STATE;
  hipLaunchKernelGGL(k_face_clear, dim3(1), dim3(64), 0, st, C.fcount, C.mg_count);
  if (g.ord == 2) {
    hipLaunchKernelGGL((k_face_simple<D, 2>), dim3(grid_xcd(g.nc, 256)), dim3(256), 0, st, g, C, used);
    PGDG_TRY((launch_general<D, 2>(g, C, used, st, ib)));
    if (D == 3 && g.mfma_rhs) {
      if (!C.bflux) return PGDG_ERR_STATE;
      const long long n = 6LL * std::max(g.N[1] * g.N[2], std::max(g.N[0] * g.N[2], g.N[0] * g.N[1]));
      hipLaunchKernelGGL(k_face_boundary3<2>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, C, used);
    }
  } else {
    hipLaunchKernelGGL((k_face_simple<D, 1>), dim3(grid_xcd(g.nc, 256)), dim3(256), 0, st, g, C, used);
    PGDG_TRY((launch_general<D, 1>(g, C, used, st, ib)));
  }
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}
template int launch_faces<2>(const Geo&, const CellDev&, const int32_t*, hipStream_t, const IfaceBatch*);
template int launch_faces<3>(const Geo&, const CellDev&, const int32_t*, hipStream_t, const IfaceBatch*);

}  // namespace pgdg
