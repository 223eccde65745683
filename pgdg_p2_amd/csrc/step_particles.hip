// step_particles.hip - per-particle kernels of the PG-DG step: G2P state
// update, particle shifting (PST) with on-the-fly neighbour search, boundary
// penalty, and the time-step reduction.
//
//   k_g2p       Particle_base::Update_state_onestep S3:particle_flow.cpp:413-477
//   k_pst_binned Particle_base::Update_PST / PST / Bound_pen S3:particle_flow.cpp:485-622
//               with the neighbour list of Level_info::Refresh_neighbor_info
//               (status 2, S3:level_infor.cpp:153-195) generated in place
//   k_timestep  MPM::Set_timestep S3:mpm.cpp:19-57, Particle_base::Set_timestep
//               S3:particle.cpp:554-569; Solver::MPM_Solver dt truncation
//               S3:solver.cpp:33-48
//
// Neighbour lists are not stored for the stage update: PST walks the same
// cells (i, j, k ascending) and bin entries in the same order as the
// reference list builder, and applies the builder's distance test to the
// positions and smoothing length saved when the stage started (xb, hb), so
// it visits exactly the reference list in the reference order.
#include "dev_common.hpp"
#include "hip_util.hpp"
#include "step_kernels.hpp"

namespace pgdg {

// The stage kernels take their particle / bin-entry counts from the device
// scalars, so their launches can be captured once and replayed while the
// counts change from step to step (the grid then covers the capacity).
template <int D>
__global__ void k_stage_save(PtlDev P, StageBufs S, const DevScalars* __restrict__ sc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc->np_dev) return;
  const int64_t cap = P.cap;
#pragma unroll
  for (int a = 0; a < D; a++) S.xb[a * cap + i] = P.x[a * cap + i];
  S.hb[i] = P.h[i];
}

template <int D>
__global__ void k_in_ele(Geo g, PtlDev P, StageBufs S, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[3];
#pragma unroll
  for (int a = 0; a < D; a++) x[a] = P.x[a * P.cap + i];
  S.in_ele[i] = cell_of<D>(g, x);
}

// ---------------------------------------------------------------------------
// Update_state_onestep (S3:particle_flow.cpp:413-477)
// ---------------------------------------------------------------------------
// md: 0 first stage, 1 Heun second stage, 2 and 3 the SSP-RK3 stages (the
// cell update's stage codes).  Order 2: the particle's copy carries the
// gradient of the cell polynomial at the particle and its second-order
// coefficients.  COPY == false skips the coefficient copy and the
// conserved state (momentum, energy, pressure, mass): within a time step
// only P2G (first stage) and split/merge (between steps) read them, so the
// values written by stages before the last are never read.  Without the
// copy a lane holds no coefficient set (the sums take each coefficient as
// it arrives): the copy kernel k_g2p_copy (last stage) holds ~150
// registers, k_g2p (the other stages) fits kG2pWaves wavefronts per SIMD
// for the coefficient gathers (6: 9.0 against 10.4 ms per step at 8, the
// same at 5).  The copy's stores are non-temporal: nothing reads them
// before the next step, and they would evict the cell coefficients the
// gathers share.
constexpr int kG2pWaves = 6;
template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}
template <int D, int ORD, bool COPY>
__device__ __forceinline__ void g2p_body(const Geo& g, const PtlDev& P, const CellDev& C, const StageBufs& S,
                                         const DevScalars* __restrict__ sc, int md) {
  constexpr bool copy = COPY;
  constexpr int U = D + 2, M = Basis<D, ORD>::M;
  // XCD-chunked numbering: the particles of one cell sit a lattice row or
  // plane apart in particle order, so runs of neighbouring workgroups on one
  // XCD read each cell's coefficients into one L2
  int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (i >= sc->np_dev) return;
  const int64_t cap = P.cap;
  const int nc = g.nc;
  const double dt = sc->dt;
  const int c = S.in_ele[i];
  const int ss = P.phase[i] - 1;
  double divu = 0;
#pragma unroll
  for (int a = 0; a < D; a++) divu += C.G[(a * 3 + a) * nc + c];
  double x[3];
#pragma unroll
  for (int a = 0; a < D; a++) x[a] = P.x[a * cap + i];
  double vol = P.vol[i];
  if (md == 0) {
#pragma unroll
    for (int a = 0; a < D; a++) {
      S.x0[a * cap + i] = x[a];
      x[a] += P.tv[a * cap + i] * dt;
    }
    S.vol0[i] = vol;
    vol = (1. + divu * dt) * vol;
  } else if (md == 1) {
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = 0.5 * ((S.x0[a * cap + i] + x[a]) + P.tv[a * cap + i] * dt);
    vol = 0.5 * (S.vol0[i] + vol + divu * dt * vol);
  } else if (md == 2) {
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = 0.75 * S.x0[a * cap + i] + 0.25 * (x[a] + P.tv[a * cap + i] * dt);
    vol = 0.75 * S.vol0[i] + 0.25 * (vol + divu * dt * vol);
  } else {
#pragma unroll
    for (int a = 0; a < D; a++)
      x[a] = (1. / 3.) * S.x0[a * cap + i] + (2. / 3.) * (x[a] + P.tv[a * cap + i] * dt);
    vol = (1. / 3.) * S.vol0[i] + (2. / 3.) * (vol + divu * dt * vol);
  }
  int ci, cj, ck;
  cell_ijk(g, c, ci, cj, ck);
  double xc[3] = {g.ps[0] * (ci + 0.5), g.ps[1] * (cj + 0.5), D == 3 ? g.ps[2] * (ck + 0.5) : 0.0};
  double dri[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) dri[a] = x[a] - xc[a];
  double phi[M - 1];
  basis_tail<D, ORD>(dri, g.q12, phi);
  double Ui[U];
  if constexpr (!COPY) {
#pragma unroll
    for (int t = 0; t < U; t++) {
      const double* cf = C.coef + (((size_t)ss * U + t) * M) * nc + c;
      double acc = 0;
#pragma unroll
      for (int b = 0; b < M - 1; b++) acc += cf[(size_t)(1 + b) * nc] * phi[b];
      Ui[t] = cf[0] + acc;
    }
  } else {
#pragma unroll
  for (int t = 0; t < U; t++) {
    double acc = 0;
    double cw[M];
    cw[0] = C.coef[(((size_t)ss * U + t) * M) * nc + c];
#pragma unroll
    for (int b = 0; b < M - 1; b++) {
      cw[1 + b] = C.coef[(((size_t)ss * U + t) * M + 1 + b) * nc + c];
      acc += cw[1 + b] * phi[b];
    }
    Ui[t] = cw[0] + acc;
    if constexpr (ORD == 2) {
      double gr[3];
      grad_at<D, ORD>(cw, dri, gr);
#pragma unroll
      for (int a = 0; a < D; a++) cw[1 + a] = gr[a];
    }
#pragma unroll
    for (int m = 0; m < M; m++) st_stream(&P.coef[(size_t)(t * M + m) * cap + i], cw[m]);
  }
  }
  double rho = Ui[0], E = Ui[U - 1];
  double rv[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) rv[a] = Ui[1 + a];
  double v[3];
#pragma unroll
  for (int a = 0; a < 3; a++) v[a] = rv[a] / rho;
  double sq = 0;
#pragma unroll
  for (int a = 0; a < D; a++) sq += v[a] * v[a];
  double ei = E - 0.5 * rho * sq;
  double Pp = eos_p(ei, Eos{g.gamma[ss], g.pinf[ss]});
  double cs = sqrt(g.gamma[ss] * (g.pinf[ss] + Pp) / rho);
  if (Pp < -0.05 * g.pinf[ss] || rho < 0) {
    if (copy)
#pragma unroll
      for (int t = 0; t < U; t++)
#pragma unroll
        for (int b = 1; b < M; b++) st_stream(&P.coef[(size_t)(t * M + b) * cap + i], 0.0);
    rho = C.rho[ss * nc + c];
#pragma unroll
    for (int a = 0; a < 3; a++) rv[a] = C.rhov[(ss * 3 + a) * nc + c];
    E = C.E[ss * nc + c];
    cs = C.c[ss * nc + c];
  }
#pragma unroll
  for (int a = 0; a < D; a++) P.x[a * cap + i] = x[a];
  P.vol[i] = vol;
  P.rho[i] = rho;
#pragma unroll
  for (int a = 0; a < 3; a++) P.v[a * cap + i] = v[a];
  P.c[i] = cs;
  P.h[i] = root_d<D>(vol) * kCellRatio / kCutOff;
  if (copy) {  // read only by P2G and split/merge, as the copy
#pragma unroll
    for (int a = 0; a < 3; a++) st_stream(&P.rhov[a * cap + i], rv[a]);
    st_stream(&P.E[i], E);
    st_stream(&P.P[i], Pp);
    st_stream(&P.mass[i], vol * rho);
  }
}

template <int D, int ORD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kG2pWaves)))
    k_g2p(Geo g, PtlDev P, CellDev C, StageBufs S, const DevScalars* __restrict__ sc, int md) {
  g2p_body<D, ORD, false>(g, P, C, S, sc, md);
}

template <int D, int ORD>
__global__ void __launch_bounds__(256) k_g2p_copy(Geo g, PtlDev P, CellDev C, StageBufs S,
                                                  const DevScalars* __restrict__ sc, int md) {
  g2p_body<D, ORD, true>(g, P, C, S, sc, md);
}

// ---------------------------------------------------------------------------
// Update_PST + PST + Bound_pen (S3:particle_flow.cpp:485-622) over
// bin-ordered neighbour data
// ---------------------------------------------------------------------------
// Gather the data PST reads from its neighbours into bin order.
template <int D>
__global__ void k_bin_gather(PtlDev P, WallDev W, StageBufs S, BinData B, const int* __restrict__ bin_ent,
                             const DevScalars* __restrict__ sc) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= sc->nent_dev) return;
  const int ref = bin_ent[e];
  const int64_t cap = P.cap;
  B.ref[e] = ref;
  if (ref_is_wall(ref)) {
    const int w = ref_wall(ref);
#pragma unroll
    for (int a = 0; a < D; a++) {
      double xa = W.x[a * W.n + w];
      B.x[a * B.ne_cap + e] = xa;
      B.xb[a * B.ne_cap + e] = xa;
    }
    B.vol[e] = W.vol[w];
    B.rho[e] = 0.0;
    B.c[e] = 0.0;
    B.ph[e] = 10;
  } else {
#pragma unroll
    for (int a = 0; a < D; a++) {
      B.x[a * B.ne_cap + e] = P.x[a * cap + ref];
      B.xb[a * B.ne_cap + e] = S.xb[a * cap + ref];
    }
    B.vol[e] = P.vol[ref];
    B.rho[e] = P.rho[ref];
    B.c[e] = P.c[ref];
    B.ph[e] = P.phase[ref];
  }
}

// Bounding box of the stage-start positions of each cell's entries (empty
// cells get an empty box).
template <int D>
__global__ void k_cell_bbox(Geo g, BinData B, const int* __restrict__ bin_off) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int e = bin_off[c]; e < bin_off[c + 1]; e++) {
#pragma unroll
    for (int a = 0; a < D; a++) {
      double v = B.xb[a * B.ne_cap + e];
      lo[a] = v < lo[a] ? v : lo[a];
      hi[a] = v > hi[a] ? v : hi[a];
    }
  }
#pragma unroll
  for (int a = 0; a < D; a++) {
    B.bmin[a * g.nc + c] = lo[a];
    B.bmax[a * g.nc + c] = hi[a];
  }
}

// The list of the particle is the reference's status-2 list: cells in i, j,
// k order, bin order within a cell, entries within h*CUT_OFF of the
// stage-start position.  Threads run over bin
// entries, so a wavefront holds particles of the same or adjacent cells and
// reads the same candidate cells.  A cell is skipped when the distance from
// the particle's stage-start position to the cell's box exceeds the list
// radius by a relative margin far above rounding; a candidate's squared
// distance is compared with the squared radius with the same margin, and
// only candidates inside the margin band take the exact sqrt test.
template <int D, int BS>
__global__ void __launch_bounds__(BS) k_pst_binned(Geo g, PtlDev P, WallDev W, StageBufs S, BinData B,
                                                    const int* __restrict__ bin_off,
                                                    const DevScalars* __restrict__ sc) {
  const int64_t E0 = (int64_t)xcd_block() * BS;
  const int64_t nent = sc->nent_dev;
  if (E0 >= nent) return;  // the whole workgroup is past the entries
  const int64_t nb = B.ne_cap;
  int64_t e = E0 + threadIdx.x;
  if constexpr (BS >= 128) {
    // lanes dealt to wavefronts by the target's quadrant (x, y halves of
    // its cell): a wavefront's lanes then sit at the same lateral offset
    // from each candidate line, accept about as many candidates of each
    // candidate cell and evaluate their pair terms together (each lane sums
    // its own target's pairs in list order, whatever lane it runs on)
    __shared__ int s_gcnt[5];
    __shared__ int s_perm[BS];
    if (threadIdx.x < 5) s_gcnt[threadIdx.x] = 0;
    __syncthreads();
    int key = 4;  // walls and past-the-end entries last
    if (e < nent && B.ref[e] >= 0) {
      key = 0;
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const double u = (B.xb[a * nb + e] - g.box_l[a]) / g.ps[a];
        key = 2 * key + (u - floor(u) >= 0.5 ? 1 : 0);
      }
    }
    const int rank = atomicAdd(&s_gcnt[key], 1);
    __syncthreads();
    int at = rank;
    for (int u = 0; u < key; u++) at += s_gcnt[u];
    s_perm[at] = threadIdx.x;
    __syncthreads();
    e = E0 + s_perm[threadIdx.x];
  }
  if (e >= nent) return;
  const int i = B.ref[e];
  if (i < 0) return;  // wall entry
  const int64_t cap = P.cap;
  const double dt = sc->dt;
  const double max_v = sc->max_v;
  double xi[3], xbi[3];
#pragma unroll
  for (int a = 0; a < D; a++) {
    xi[a] = B.x[a * nb + e];
    xbi[a] = B.xb[a * nb + e];
  }
  const double hi = P.h[i], hbi = S.hb[i];
  const double rhoi = B.rho[e], ci = B.c[e];
  const int phi = B.ph[e];
  const double ca = max_v * 5.;
  const double P_b = rhoi * ca * ca;
  const double dxk = hi / kCellRatio * kCutOff;
  const double W_dp = kern_W<D>(dxk, hi);
  const double rlim = hbi * kCutOff;
  const double r2 = rlim * rlim;
  const double r2_in = r2 * (1.0 - 1e-12), r2_out = r2 * (1.0 + 1e-12);
  const double tmp1_base = -1. * P_b * (1. / rhoi);
  double pst[3] = {0, 0, 0}, sharp[3] = {0, 0, 0};
  const int pos = cell_of<D>(g, xbi);
  int pi, pj, pk;
  cell_ijk(g, pos, pi, pj, pk);
  const int r = g.radius;
  const int rk = D == 3 ? r : 0;
  for (int t = max(pi - r, 0); t <= min(pi + r, g.N[0] - 1); t++)
    for (int s = max(pj - r, 0); s <= min(pj + r, g.N[1] - 1); s++)
      for (int m = max(pk - rk, 0); m <= min(pk + rk, g.N[2] - 1); m++) {
        const int cc = (t * g.N[1] + s) * g.N[2] + m;
        double bd = 0;
#pragma unroll
        for (int a = 0; a < D; a++) {
          double lo = B.bmin[a * g.nc + cc], hb_ = B.bmax[a * g.nc + cc];
          double d = xbi[a] < lo ? lo - xbi[a] : (xbi[a] > hb_ ? xbi[a] - hb_ : 0.0);
          bd += d * d;
        }
        if (!(bd <= r2_out)) continue;  // empty cell or wholly outside
        const int e1 = bin_off[cc + 1];
        for (int f0 = bin_off[cc]; f0 < e1; f0 += 32) {
          // list membership of the cell's candidates as a bit mask (loads
          // issued back to back), then the pair terms once per set bit in
          // entry order
          const int nf = min(32, e1 - f0);
          uint32_t msk = 0;
          // four candidates at a time: their loads are issued before any of
          // them is tested, so the load latency is paid once per four
          for (int k = 0; k < nf; k += 4) {
            double sqb[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
              const int f = min(f0 + k + u, e1 - 1);  // a repeat of the last candidate past the end
              double d[3];
#pragma unroll
              for (int a = 0; a < D; a++) d[a] = xbi[a] - B.xb[a * nb + f];
              double q = 0;
#pragma unroll
              for (int a = 0; a < D; a++) q += d[a] * d[a];
              sqb[u] = q;
            }
            bool band = false;
#pragma unroll
            for (int u = 0; u < 4; u++) {
              if (k + u < nf && sqb[u] <= r2_in) msk |= 1u << (k + u);
              band |= k + u < nf && sqb[u] > r2_in && sqb[u] <= r2_out;
            }
            // inside the rounding band the exact test takes the square root
            // (rare: a wave-level branch)
            if (__ballot(band) != 0 && band)
#pragma unroll
              for (int u = 0; u < 4; u++)
                if (k + u < nf && sqb[u] > r2_in && sqb[u] <= r2_out && sqrt(sqb[u]) <= rlim)
                  msk |= 1u << (k + u);
          }
          // the next accepted candidate's position, volume and phase are
          // loaded while the current one's pair terms are evaluated
          int fn = msk ? f0 + __builtin_ctz(msk) : -1;
          double xn[3] = {0, 0, 0}, voln = 0;
          int phn = 0;
          if (fn >= 0) {
#pragma unroll
            for (int a = 0; a < D; a++) xn[a] = B.x[a * nb + fn];
            voln = B.vol[fn];
            phn = B.ph[fn];
          }
          while (msk) {
            const int f = fn;
            msk &= msk - 1;
            double xj[3] = {xn[0], xn[1], xn[2]};
            const double vol_j = voln;
            const int phj = phn;
            fn = msk ? f0 + __builtin_ctz(msk) : -1;
            if (fn >= 0) {
#pragma unroll
              for (int a = 0; a < D; a++) xn[a] = B.x[a * nb + fn];
              voln = B.vol[fn];
              phn = B.ph[fn];
            }
            double dr[3];
#pragma unroll
            for (int a = 0; a < D; a++) dr[a] = xi[a] - xj[a];
            double sq = 0;
#pragma unroll
            for (int a = 0; a < D; a++) sq += dr[a] * dr[a];
            double dist = sqrt(sq);
            if (dist != 0) {
              const bool wl = phj == 10;
              double aw = kern_dW<D>(dist, hi);
              double Wij = kern_W<D>(dist, hi);
              double rr = Wij / W_dp;
              double q2 = rr * rr;
              double R = 0.2 * (q2 * q2);
              double tmp1 = tmp1_base * (1 + R);
              if (!wl) {
#pragma unroll
                for (int a = 0; a < D; a++) pst[a] += dr[a] * aw * tmp1 * vol_j;
                if (phj != phi) {
                  double rhoj = B.rho[f], cj = B.c[f];
                  double tmp2 = 1. / (rhoi + rhoj);
                  double kk = -0.1 * tmp2 * (ci * ci * rhoi + cj * cj * rhoj);
#pragma unroll
                  for (int a = 0; a < D; a++) sharp[a] += kk * (dr[a] * aw) * vol_j;
                }
              } else {
                const int w = ref_wall(B.ref[f]);
#pragma unroll
                for (int a = 0; a < D; a++) pst[a] += W.norm[a * W.n + w] * Wij * tmp1 * vol_j;
              }
            }
          }
        }
      }
  double pen[3] = {0, 0, 0};
  {
    int c = S.in_ele[i];
    int ci_, cj_, ck_;
    cell_ijk(g, c, ci_, cj_, ck_);
    const double xc[3] = {g.ps[0] * (ci_ + 0.5), g.ps[1] * (cj_ + 0.5),
                          D == 3 ? g.ps[2] * (ck_ + 0.5) : 0.0};
    for (int f = 0; f < 2 * D; f++) {
      int ax, ni = ci_, nj = cj_, nk = ck_;
      double sgn, cen[3] = {xc[0], xc[1], xc[2]};
      switch (f) {
        case 0: ax = 1; sgn = -1; cen[1] = xc[1] - 0.5 * g.ps[1]; nj--; break;
        case 1: ax = 0; sgn = 1; cen[0] = xc[0] + 0.5 * g.ps[0]; ni++; break;
        case 2: ax = 1; sgn = 1; cen[1] = xc[1] + 0.5 * g.ps[1]; nj++; break;
        case 3: ax = 0; sgn = -1; cen[0] = xc[0] - 0.5 * g.ps[0]; ni--; break;
        case 4: ax = 2; sgn = -1; cen[2] = xc[2] - 0.5 * g.ps[2]; nk--; break;
        default: ax = 2; sgn = 1; cen[2] = xc[2] + 0.5 * g.ps[2]; nk++; break;
      }
      bool in = ni >= 0 && ni < g.N[0] && nj >= 0 && nj < g.N[1] && nk >= 0 && nk < g.N[2];
      if (in || g.bc[f] == 0) continue;
      double nrm[3] = {0, 0, 0};
      nrm[ax] = sgn;
      double ds = 0;
#pragma unroll
      for (int a = 0; a < D; a++) ds += (xi[a] - cen[a]) * nrm[a];
      if (ds < 0)
#pragma unroll
        for (int a = 0; a < 3; a++) nrm[a] = -1 * nrm[a];
      double dis = fabs(ds);
      if (dis < 0.4 * g.p_sz) {
        double kk = 0.02 / dis * (ci * ci);
#pragma unroll
        for (int a = 0; a < D; a++) pen[a] += kk * nrm[a];
      }
    }
  }
  const double px = g.p_sz;
  double dx[3], dx1[3], dx2[3];
#pragma unroll
  for (int a = 0; a < D; a++) {
    dx[a] = pst[a] * dt * dt / 2.;
    dx1[a] = sharp[a] * dt * dt / 2.;
    dx2[a] = pen[a] * dt * dt / 2.;
  }
  auto clampv = [&](double* d, double cap_) {
    double s2 = 0;
#pragma unroll
    for (int a = 0; a < D; a++) s2 += d[a] * d[a];
    double dis = sqrt(s2);
    double lim = amin1(cap_, dis) / (dis + 1e-20);
#pragma unroll
    for (int a = 0; a < D; a++) d[a] *= lim;
  };
  clampv(dx, px / 10);
  clampv(dx1, px / 10);
  clampv(dx2, px / 2);
  const bool frozen = g.sem == PGDG_SEM_S3 && P.mark_in[i] == 1;
#pragma unroll
  for (int a = 0; a < D; a++) {
    double dv = frozen ? 0.0 : (dx[a] + dx1[a] + dx2[a]) / dt;
    P.tv[a * cap + i] = P.v[a * cap + i] + dv;
  }
}

// the particle-shifting kernel alone, at a workgroup size of 64, 128 or 256
template <int D>
int launch_pst_kernel(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                      const BinData& B, int64_t n_ent, const int* bin_off, const DevScalars* sc,
                      hipStream_t st, int block) {
  switch (block) {
    case 64:
      hipLaunchKernelGGL((k_pst_binned<D, 64>), dim3(grid_xcd(n_ent, 64)), dim3(64), 0, st, g, P, W, S,
                         B, bin_off, sc);
      break;
    case 128:
      hipLaunchKernelGGL((k_pst_binned<D, 128>), dim3(grid_xcd(n_ent, 128)), dim3(128), 0, st, g, P, W,
                         S, B, bin_off, sc);
      break;
    case 256:
      hipLaunchKernelGGL((k_pst_binned<D, 256>), dim3(grid_xcd(n_ent, 256)), dim3(256), 0, st, g, P, W,
                         S, B, bin_off, sc);
      break;
    default:
      return PGDG_ERR_ARG;
  }
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int launch_pst_binned(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                      const BinData& B, int64_t n_ent, const int* bin_off, const int* bin_ent,
                      const DevScalars* sc, hipStream_t st, int block, int parts) {
  if (n_ent <= 0) return PGDG_OK;
  if (parts & 1) {
    hipLaunchKernelGGL(k_bin_gather<D>, dim3(grid_for(n_ent, 256, 1LL << 30)), dim3(256), 0, st, P, W, S,
                       B, bin_ent, sc);
    hipLaunchKernelGGL(k_cell_bbox<D>, dim3((g.nc + 255) / 256), dim3(256), 0, st, g, B, bin_off);
  }
  if (parts & 2) PGDG_TRY(launch_pst_kernel<D>(g, P, W, S, B, n_ent, bin_off, sc, st, block));
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

// Neighbour-list sizes (Refresh_neighbor_info, status 2, S3:level_infor.cpp:
// 153-195) over the bin-ordered stage-start positions: the list test of
// k_pst_binned, counted instead of summed.  cnt is indexed by particle.
template <int D>
__global__ void k_count_neighbors(Geo g, PtlDev P, StageBufs S, BinData B, const int* __restrict__ bin_off,
                                  const DevScalars* __restrict__ sc, int32_t* __restrict__ cnt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= sc->nent_dev) return;
  const int i = B.ref[e];
  if (i < 0) return;  // wall or removed entry
  const int64_t nb = B.ne_cap;
  double xbi[3];
#pragma unroll
  for (int a = 0; a < D; a++) xbi[a] = B.xb[a * nb + e];
  const double rlim = S.hb[i] * kCutOff;
  const double r2 = rlim * rlim;
  const double r2_in = r2 * (1.0 - 1e-12), r2_out = r2 * (1.0 + 1e-12);
  const int pos = cell_of<D>(g, xbi);
  int pi, pj, pk;
  cell_ijk(g, pos, pi, pj, pk);
  const int r = g.radius, rk = D == 3 ? r : 0;
  int n = 0;
  for (int t = max(pi - r, 0); t <= min(pi + r, g.N[0] - 1); t++)
    for (int s = max(pj - r, 0); s <= min(pj + r, g.N[1] - 1); s++)
      for (int m = max(pk - rk, 0); m <= min(pk + rk, g.N[2] - 1); m++) {
        const int cc = (t * g.N[1] + s) * g.N[2] + m;
        for (int f = bin_off[cc]; f < bin_off[cc + 1]; f++) {
          double sqb = 0;
#pragma unroll
          for (int a = 0; a < D; a++) {
            const double d = xbi[a] - B.xb[a * nb + f];
            sqb += d * d;
          }
          if (sqb <= r2_in || (sqb <= r2_out && sqrt(sqb) <= rlim)) n++;
        }
      }
  cnt[i] = n;
}

template <int D>
int launch_neighbor_counts(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                           const BinData& B, int64_t n_ent, const int* bin_off, const int* bin_ent,
                           const DevScalars* sc, hipStream_t st, int32_t* cnt) {
  if (n_ent <= 0) return PGDG_OK;
  // the bin-ordered stage-start positions (the records of the FP64 form)
  hipLaunchKernelGGL(k_bin_gather<D>, dim3(grid_for(n_ent, 256, 1LL << 30)), dim3(256), 0, st, P, W, S, B,
                     bin_ent, sc);
  hipLaunchKernelGGL(k_count_neighbors<D>, dim3(grid_for(n_ent, 256, 1LL << 30)), dim3(256), 0, st, g, P,
                     S, B, bin_off, sc, cnt);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

// ---------------------------------------------------------------------------
// Set_timestep reduction + Solver dt truncation
// ---------------------------------------------------------------------------
__device__ inline unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

__global__ void k_ts_init(DevScalars* sc, double output_dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (!(sc->integral_time < output_dt - 1.0e-10)) sc->integral_time = 0.0;
    sc->red_dt = dbits(1.e20);
    sc->red_maxv = dbits(0.0);
  }
}

template <int D, bool CELLS>
__global__ void __launch_bounds__(256) k_timestep(Geo g, PtlDev P, CellDev C,
                                                  const int32_t* __restrict__ used, int64_t n,
                                                  DevScalars* sc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double ts = 1.e20, vm = 0.0;
  if (i < n && (!CELLS || used[i])) {
    double v[3], h, c12, a12;
    if (CELLS) {
      const int nc = g.nc;
#pragma unroll
      for (int a = 0; a < D; a++) v[a] = C.v[a * nc + i];
      h = kCellRatio / kCutOff * g.p_sz;
      c12 = amax1(C.c[i], C.c[nc + i]);
      a12 = amax1(C.ac[i], C.ac[nc + i]);
    } else {
#pragma unroll
      for (int a = 0; a < D; a++) v[a] = P.v[a * P.cap + i];
      h = P.h[i];
      c12 = amax1(P.c[i], 0.0);
      a12 = amax1(0.0, 0.0);
    }
    double sq = 0;
#pragma unroll
    for (int a = 0; a < D; a++) sq += v[a] * v[a];
    double Uv = sqrt(sq);
    const double cfl = g.cfl;
    double ds = h / kCellRatio * kCutOff;
    double t1 = cfl * ds / (c12 + Uv);
    double t2 = cfl * sqrt(ds / a12);
    ts = amin1(t1, t2);
    vm = Uv;
  }
  // wavefront, then workgroup reductions (min/max are order independent),
  // one pair of integer atomics per workgroup
  for (int off = 32; off > 0; off >>= 1) {
    double o1 = __shfl_down(ts, off, 64);
    double o2 = __shfl_down(vm, off, 64);
    ts = amin1(ts, o1);
    vm = amax1(vm, o2);
  }
  __shared__ double red_ts[4], red_vm[4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red_ts[w] = ts;
    red_vm[w] = vm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
      ts = amin1(ts, red_ts[q]);
      vm = amax1(vm, red_vm[q]);
    }
    if (ts > 0) atomicMin(&sc->red_dt, dbits(ts));
    if (vm >= 0) atomicMax(&sc->red_maxv, dbits(vm));
  }
}

__global__ void k_ts_finalize(DevScalars* sc, double output_dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double dt = __longlong_as_double((long long)sc->red_dt);
    double mv = __longlong_as_double((long long)sc->red_maxv);
    sc->max_v = mv;
    double timestep = dt;
    if (sc->integral_time >= (output_dt - timestep)) timestep = output_dt - sc->integral_time;
    sc->dt = timestep;
    sc->integral_time += timestep;
    sc->run_time += timestep;
    sc->iterate_num += 1;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int D>
int launch_stage_save(const Geo& g, const PtlDev& P, const StageBufs& S, int64_t n, int md,
                      const DevScalars* sc, hipStream_t st) {
  (void)g;
  (void)md;
  if (n <= 0) return PGDG_OK;
  hipLaunchKernelGGL(k_stage_save<D>, dim3(grid_for(n, 256, 1LL << 30)), dim3(256), 0, st, P, S, sc);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int launch_in_ele(const Geo& g, const PtlDev& P, const StageBufs& S, int64_t n, hipStream_t st) {
  if (n <= 0) return PGDG_OK;
  hipLaunchKernelGGL(k_in_ele<D>, dim3(grid_for(n, 256, 1LL << 30)), dim3(256), 0, st, g, P, S, n);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int launch_g2p(const Geo& g, const PtlDev& P, const CellDev& C, const StageBufs& S, int64_t n,
               const DevScalars* sc, int md, bool copy, hipStream_t st) {
  if (n <= 0) return PGDG_OK;
  const dim3 grid(grid_xcd(n, 256));
  if (g.ord == 2) {
    if (copy) hipLaunchKernelGGL((k_g2p_copy<D, 2>), grid, dim3(256), 0, st, g, P, C, S, sc, md);
    else hipLaunchKernelGGL((k_g2p<D, 2>), grid, dim3(256), 0, st, g, P, C, S, sc, md);
  } else {
    if (copy) hipLaunchKernelGGL((k_g2p_copy<D, 1>), grid, dim3(256), 0, st, g, P, C, S, sc, md);
    else hipLaunchKernelGGL((k_g2p<D, 1>), grid, dim3(256), 0, st, g, P, C, S, sc, md);
  }
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int launch_timestep(const Geo& g, const PtlDev& P, const CellDev& C, const int32_t* used,
                    int64_t n, bool cells, DevScalars* sc, double output_dt, hipStream_t st) {
  hipLaunchKernelGGL(k_ts_init, dim3(1), dim3(64), 0, st, sc, output_dt);
  int64_t cnt = cells ? g.nc : n;
  if (cnt > 0) {
    unsigned grid = (unsigned)((cnt + 255) / 256);
    if (cells)
      hipLaunchKernelGGL((k_timestep<D, true>), dim3(grid), dim3(256), 0, st, g, P, C, used, cnt, sc);
    else
      hipLaunchKernelGGL((k_timestep<D, false>), dim3(grid), dim3(256), 0, st, g, P, C, used, cnt, sc);
  }
  hipLaunchKernelGGL(k_ts_finalize, dim3(1), dim3(64), 0, st, sc, output_dt);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

#define PGDG_INST(D)                                                                              \
  template int launch_stage_save<D>(const Geo&, const PtlDev&, const StageBufs&, int64_t, int,     \
                                    const DevScalars*, hipStream_t);                              \
  template int launch_in_ele<D>(const Geo&, const PtlDev&, const StageBufs&, int64_t, hipStream_t); \
  template int launch_neighbor_counts<D>(const Geo&, const PtlDev&, const WallDev&, const StageBufs&, \
                                         const BinData&, int64_t, const int*, const int*,           \
                                         const DevScalars*, hipStream_t, int32_t*);                 \
  template int launch_g2p<D>(const Geo&, const PtlDev&, const CellDev&, const StageBufs&, int64_t, \
                             const DevScalars*, int, bool, hipStream_t);                          \
  template int launch_pst_binned<D>(const Geo&, const PtlDev&, const WallDev&, const StageBufs&,   \
                                    const BinData&, int64_t, const int*, const int*,              \
                                    const DevScalars*, hipStream_t, int, int);                    \
  template int launch_pst_kernel<D>(const Geo&, const PtlDev&, const WallDev&, const StageBufs&,   \
                                    const BinData&, int64_t, const int*, const DevScalars*,       \
                                    hipStream_t, int);                                            \
  template int launch_timestep<D>(const Geo&, const PtlDev&, const CellDev&, const int32_t*,       \
                                  int64_t, bool, DevScalars*, double, hipStream_t);
PGDG_INST(2)
PGDG_INST(3)
#undef PGDG_INST

}  // namespace pgdg
