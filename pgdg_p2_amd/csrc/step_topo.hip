// step_topo.hip - particle topology of the PG-DG step: cell binning, cell-list
// edits, neighbour lists, Split_particle_fast, Merge_particle_fast and the
// stable compaction that removes particles.
//
// Reference routines (single-thread order, see the oracle for the literal
// restatement):
//   binning        MPM::Reset_cell_list_info / Map_the_particle_to_tree
//                  S3:mpm.cpp:65-96, S3:level_infor.cpp:117-147
//   lists          Level_info::Refresh_neighbor_info S3:level_infor.cpp:153-195
//   split          MPM::Split_particle_fast S3:mpm.cpp:1036-1417,
//                  Set_split_ptcl / Copy_split_state / Set_ptcl
//                  S3:particle.cpp:575-609, 667-722
//   merge          MPM::Merge_particle_fast S3:mpm.cpp:1536-1706,
//                  Set_merge_ptcl S3:particle.cpp:614-661
//
// Determinism: every ordering decision is a stable radix sort or an exclusive
// scan; no atomic decides a position.  A cell list edit is "keep the old
// entries that were not removed, in order, then append the new entries in
// the reference's append order", computed per cell from a scan.
//
// Neighbour lists are materialised only where the reference reads a list
// that is not rebuilt first: the split/merge passes read the lists of split
// parents, outlet particles and particles below the merge volume.  At the
// end of a step the lists of exactly those particles are built (from the
// positions and smoothing lengths saved at the start of the last stage, which
// is when the reference built them); each split/merge pass rebuilds the
// lists of the affected particles that are below the merge volume.
#include <hipcub/hipcub.hpp>

#include "dev_common.hpp"
#include "hip_util.hpp"
#include "step_topo.hpp"

namespace pgdg {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__global__ void k_widen(const int* in, int64_t n, int64_t* out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = in[e];
}

__global__ void k_keep(const uint8_t* removed, int64_t n, int* keep) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) keep[e] = removed[e] ? 0 : 1;
}

__global__ void k_split_src(const uint8_t* type, int64_t n, uint8_t* src) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) src[e] = (type[e] == 1 || type[e] == 3) ? 1 : 0;  // split parent or outlet
}

__global__ void k_mark_refs(const int* refs, int64_t k, uint8_t* flag) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < k && refs[e] >= 0) flag[refs[e]] = 1;
}

__global__ void k_set_i32(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Fluid entries (refs >= 0) of every cell's bin.
__global__ void k_fluid_count(const int* __restrict__ bin_off, const int* __restrict__ bin_ent, int nc,
                              int* fcnt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int k = 0;
  for (int e = bin_off[c]; e < bin_off[c + 1]; e++) k += bin_ent[e] >= 0 ? 1 : 0;
  fcnt[c] = k;
}

// Particles whose list the next step's split/merge passes read before any
// rebuild: split candidates, outlet candidates, merge candidates.
template <int D>
__global__ void k_flag_lists(Geo g, PtlDev P, int64_t n, uint8_t* need) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = P.vol[i];
  bool f = v < g.vol_min || v > g.vol_max || (g.sem == PGDG_SEM_S3 && P.x[i] > g.box_r[0]);
  need[i] = f ? 1 : 0;
}

__global__ void k_gather_slab(const double* __restrict__ src, double* __restrict__ dst, int rows,
                              const int32_t* __restrict__ isrc, int32_t* __restrict__ idst,
                              int irows, int64_t n, int64_t cap, const int* __restrict__ newidx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int j = newidx[i];
  if (j < 0) return;
  for (int r = 0; r < rows; r++) dst[(size_t)r * cap + j] = src[(size_t)r * cap + i];
  for (int r = 0; r < irows; r++) idst[(size_t)r * cap + j] = isrc[(size_t)r * cap + i];
}

int ParticleArrays::gather(const int* newidx, int64_t n, hipStream_t st) {
  if (n == 0) return PGDG_OK;
  hipLaunchKernelGGL(k_gather_slab, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f64[0],
                     f64[1], rows64, i32[0], i32[1], 2, n, cap, newidx);
  PGDG_HIP_TRY(hipGetLastError());
  double* t = f64[0];
  f64[0] = f64[1];
  f64[1] = t;
  int32_t* u = i32[0];
  i32[0] = i32[1];
  i32[1] = u;
  return PGDG_OK;
}

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
template <int D>
__global__ void k_bin_keys(Geo g, PtlDev P, WallDev W, int64_t n, uint32_t* keys, int* vals) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n + W.n) return;
  double x[3];
  if (e < n) {
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = P.x[a * P.cap + e];
    vals[e] = (int)e;
  } else {
    int64_t w = e - n;
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = W.x[a * W.n + w];
    vals[e] = -(int)w - 2;
  }
  keys[e] = (uint32_t)cell_of<D>(g, x);
}

__global__ void k_hist(const uint32_t* keys, int64_t n, int* cnt) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) atomicAdd(&cnt[keys[e]], 1);
}

__global__ void k_fill_i32(int* a, int64_t n, int v) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = v;
}

__global__ void k_fill_u8(uint8_t* a, int64_t n, uint8_t v) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = v;
}

// ---------------------------------------------------------------------------
// cell-list edit: keep old entries whose particle is not flagged in `rm`,
// then append `app_ref` (sorted by cell, append order kept)
// ---------------------------------------------------------------------------
__global__ void k_edit_count(const int* off, const int* ent, int nc, const uint8_t* rm, int* cnt,
                             const int* app_cnt) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int k = 0;
  for (int e = off[c]; e < off[c + 1]; e++) {
    int r = ent[e];
    if (r >= 0 && rm[r]) continue;
    k++;
  }
  cnt[c] = k + app_cnt[c];
}

__global__ void k_edit_fill(const int* off, const int* ent, int nc, const uint8_t* rm,
                            const int* noff, int* nent, const int* app_off, const int* app_cnt,
                            const int* app_sorted) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int o = noff[c];
  for (int e = off[c]; e < off[c + 1]; e++) {
    int r = ent[e];
    if (r >= 0 && rm[r]) continue;
    nent[o++] = r;
  }
  int a0 = app_off[c];
  for (int t = 0; t < app_cnt[c]; t++) nent[o++] = app_sorted[a0 + t];
}

// ---------------------------------------------------------------------------
// neighbour lists (status 2: radius cells, dist <= h*CUT_OFF; status 3:
// 2 cells, dist <= 2 p_sz)
// ---------------------------------------------------------------------------
template <int D, bool FILL>
__global__ void __launch_bounds__(128) k_lists(Geo g, const double* __restrict__ xq, int64_t xcap,
                                               const double* __restrict__ hq,
                                               const double* __restrict__ xall, WallDev W,
                                               int64_t n, const uint8_t* __restrict__ need,
                                               int status, const int* __restrict__ bin_off,
                                               const int* __restrict__ bin_ent, int* cnt,
                                               const int64_t* __restrict__ loff, int* pool) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!need[i]) {
    if (!FILL) cnt[i] = 0;
    return;
  }
  double xi[3];
#pragma unroll
  for (int a = 0; a < D; a++) xi[a] = xq[a * xcap + i];
  const double lim = status == 2 ? hq[i] * kCutOff : g.p_sz * 2.;
  const int r = status == 2 ? g.radius : 2;
  const int rk = D == 3 ? r : 0;
  const int pos = cell_of<D>(g, xi);
  int pi, pj, pk;
  cell_ijk(g, pos, pi, pj, pk);
  int k = 0;
  int64_t o = FILL ? loff[i] : 0;
  for (int t = max(pi - r, 0); t <= min(pi + r, g.N[0] - 1); t++)
    for (int s = max(pj - r, 0); s <= min(pj + r, g.N[1] - 1); s++)
      for (int m = max(pk - rk, 0); m <= min(pk + rk, g.N[2] - 1); m++) {
        int cc = (t * g.N[1] + s) * g.N[2] + m;
        for (int e = bin_off[cc]; e < bin_off[cc + 1]; e++) {
          int ref = bin_ent[e];
          if (ref == -1) continue;
          double sq = 0;
#pragma unroll
          for (int a = 0; a < D; a++) {
            double xe = ref_is_wall(ref) ? W.x[a * W.n + ref_wall(ref)] : xall[a * xcap + ref];
            double d = xi[a] - xe;
            sq += d * d;
          }
          if (sqrt(sq) <= lim) {
            if (FILL) pool[o + k] = ref;
            k++;
          }
        }
      }
  if (!FILL) cnt[i] = k;
}

__global__ void k_list_commit(int64_t n, const uint8_t* need, const int* cnt, const int64_t* off,
                              int64_t base, int* lcnt, int64_t* loff) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !need[i]) return;
  lcnt[i] = cnt[i];
  loff[i] = base + off[i];
}

__global__ void k_list_offsets(int64_t n, const uint8_t* need, const int64_t* off, int64_t base,
                               int64_t* loff) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !need[i]) return;
  loff[i] = base + off[i];
}

// ---------------------------------------------------------------------------
// compaction: remap references, gather particle arrays
// ---------------------------------------------------------------------------
__global__ void k_remap_refs(int* a, int64_t n, const int* newidx) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int r = a[e];
  if (r >= 0) a[e] = newidx[r];
}

__global__ void k_newidx(const uint8_t* removed, const int* scan, int64_t n, int* newidx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  newidx[i] = removed[i] ? -1 : scan[i];
}

template <typename T>
__global__ void k_gather(const T* __restrict__ src, T* __restrict__ dst, int64_t n, int64_t cap,
                         int rows, const int* __restrict__ newidx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int j = newidx[i];
  if (j < 0) return;
  for (int r = 0; r < rows; r++) dst[r * cap + j] = src[r * cap + i];
}

// ---------------------------------------------------------------------------
// mark_in (S3:mpm.cpp:1398-1399, 1698-1699)
// ---------------------------------------------------------------------------
template <int D>
__global__ void k_mark_in(Geo g, PtlDev P, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  P.mark_in[i] = (P.x[i] - g.box_l[0] < root_d<D>(P.vol[i])) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// split
// ---------------------------------------------------------------------------
enum { T_NONE = 0, T_SPLIT = 1, T_INLET = 2, T_OUTLET = 3 };

template <int D>
__global__ void k_split_classify(Geo g, PtlDev P, int64_t n, uint8_t* type, int* newcnt,
                                 int* seqcnt, int* incnt, int* outcnt, const int* __restrict__ fcnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool io = g.sem == PGDG_SEM_S3;
  double vol = P.vol[i];
  double dp = root_d<D>(vol);
  double x0 = P.x[i];
  bool f1 = vol > g.vol_max;
  if (f1 && g.ppc_max > 0) {  // no split that would take the cell above ppc_max
    double x[3];
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = P.x[a * P.cap + i];
    f1 = fcnt[cell_of<D>(g, x)] + ((1 << D) - 1) <= g.ppc_max;
  }
  bool f2 = io && x0 - g.box_l[0] >= dp && P.mark_in[i] == 1;
  bool f3 = io && x0 > g.box_r[0];
  uint8_t t = f1 ? T_SPLIT : f2 ? T_INLET : f3 ? T_OUTLET : T_NONE;
  type[i] = t;
  const int nch = D == 3 ? 8 : 4;
  newcnt[i] = t == T_SPLIT ? nch - 1 : t == T_INLET ? 1 : 0;
  seqcnt[i] = t == T_SPLIT ? nch : 0;
  incnt[i] = t == T_INLET ? 1 : 0;
  outcnt[i] = t == T_OUTLET ? 1 : 0;
}

// Increments of the parent's polynomial at offset dr (order 1: slopes * dr;
// order 2: the particle-centred P2 form) and, for order 2, the copy moved to
// the new position (gradient re-evaluated at dr).
template <int D, int ORD>
__device__ inline void split_increment(const double* pcoef, const double* dr, double* sl) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M;
  double psi[M - 1];
  incr_tail<D, ORD>(dr, psi);
#pragma unroll
  for (int u = 0; u < U; u++) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < M - 1; k++) s += pcoef[u * M + 1 + k] * psi[k];
    sl[u] = s;
  }
}
template <int D, int ORD>
__device__ inline void write_moved_copy(const PtlDev& P, int64_t dst, const double* pcoef,
                                        const double* dr) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M;
  const int64_t cap = P.cap;
  for (int q = 0; q < U * M; q++) P.coef[(size_t)q * cap + dst] = pcoef[q];
  if constexpr (ORD == 2)
    for (int u = 0; u < U; u++) {
      double gr[3];
      grad_at<D, ORD>(pcoef + u * M, dr, gr);
#pragma unroll
      for (int a = 0; a < D; a++) P.coef[(size_t)(u * M + 1 + a) * cap + dst] = gr[a];
    }
}

// Child of parent p at coordinate xc (Set_split_ptcl, S3:particle.cpp:575-609)
template <int D, int ORD>
__device__ void write_split_child(const Geo& g, const PtlDev& P, int64_t p, int64_t dst,
                                  const double* xc, const double* pcoef) {
  constexpr int U = D + 2;
  const int64_t cap = P.cap;
  double dr[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) dr[a] = xc[a] - P.x[a * cap + p];
  const int ph = P.phase[p];
  const double div = D == 3 ? 8.0 : 4.0;
  double sl[U];
  split_increment<D, ORD>(pcoef, dr, sl);
  double rho = P.rho[p] + sl[0];
  double rv[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) rv[a] = P.rhov[a * cap + p] + sl[1 + a];
  double E = P.E[p] + sl[U - 1];
  double v[3];
#pragma unroll
  for (int a = 0; a < 3; a++) v[a] = rv[a] / rho;
  double sq = 0;
#pragma unroll
  for (int a = 0; a < D; a++) sq += v[a] * v[a];
  double ei = E - 0.5 * rho * sq;
  const int s = ph - 1;
  double Pp = eos_p(ei, Eos{g.gamma[s], g.pinf[s]});
  double cs = sqrt(g.gamma[s] * (g.pinf[s] + Pp) / rho);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    P.x[a * cap + dst] = a < D ? xc[a] : 0.0;
    P.rhov[a * cap + dst] = rv[a];
    P.v[a * cap + dst] = v[a];
    P.tv[a * cap + dst] = v[a];
  }
  P.phase[dst] = ph;
  P.vol[dst] = P.vol[p] / div;
  P.mass[dst] = P.mass[p] / div;
  P.h[dst] = P.h[p] / 2.0;
  write_moved_copy<D, ORD>(P, dst, pcoef, dr);
  P.rho[dst] = rho;
  P.E[dst] = E;
  P.P[dst] = Pp;
  P.c[dst] = cs;
  P.mark_in[dst] = 0;
}

template <int D, int ORD>
__global__ void k_split_generate(Geo g, PtlDev P, int64_t n, const uint8_t* __restrict__ type,
                                 const int* __restrict__ newbase, const int* __restrict__ seqbase,
                                 const int* __restrict__ inbase, int s_total, uint8_t* rm,
                                 uint8_t* removed, int* app_ref, uint32_t* app_cell, int* lcnt) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t t = type[i];
  const int64_t cap = P.cap;
  if (t == T_OUTLET) {
    rm[i] = 1;
    removed[i] = 1;
    return;
  }
  if (t != T_SPLIT && t != T_INLET) return;
  double dp = root_d<D>(P.vol[i]);
  double pcoef[U * M];
  for (int q = 0; q < U * M; q++) pcoef[q] = P.coef[(size_t)q * cap + i];
  if (t == T_INLET) {
    // Set_ptcl (S3:particle.cpp:696-722) of a copy at x - dp
    int64_t dst = n + newbase[i];
    double rv[3] = {0, 0, 0};
    double rho = P.rho[i], E = P.E[i];
#pragma unroll
    for (int a = 0; a < D; a++) rv[a] = P.rhov[a * cap + i];
    double v[3];
#pragma unroll
    for (int a = 0; a < 3; a++) v[a] = rv[a] / rho;
    double sq = 0;
#pragma unroll
    for (int a = 0; a < D; a++) sq += v[a] * v[a];
    double ei = E - 0.5 * rho * sq;
    const int s = P.phase[i] - 1;
    double Pp = eos_p(ei, Eos{g.gamma[s], g.pinf[s]});
    double cs = sqrt(g.gamma[s] * (g.pinf[s] + Pp) / rho);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      double xa = a < D ? P.x[a * cap + i] : 0.0;
      P.x[a * cap + dst] = a == 0 ? xa - dp : xa;
      P.rhov[a * cap + dst] = rv[a];
      P.v[a * cap + dst] = v[a];
      P.tv[a * cap + dst] = v[a];
    }
    P.vol[dst] = P.vol[i];
    P.mass[dst] = P.mass[i];
    P.h[dst] = P.h[i];
    P.phase[dst] = P.phase[i];
    for (int q = 0; q < U * M; q++) P.coef[(size_t)q * cap + dst] = pcoef[q];
    P.rho[dst] = rho;
    P.E[dst] = E;
    P.P[dst] = Pp;
    P.c[dst] = cs;
    P.mark_in[dst] = 0;
    lcnt[dst] = 0;
    double xd[3];
#pragma unroll
    for (int a = 0; a < D; a++) xd[a] = P.x[a * cap + dst];
    int seq = s_total + inbase[i];
    app_ref[seq] = (int)dst;
    app_cell[seq] = (uint32_t)cell_of<D>(g, xd);
    return;
  }
  // split into 2^D children at (x - dp/4) + dp/2 (ci, cj[, ck])
  const int nch = D == 3 ? 8 : 4;
  double pref[3];
#pragma unroll
  for (int a = 0; a < D; a++) pref[a] = P.x[a * cap + i] - dp / 4.;
  int64_t base = n + newbase[i];
  int seq = seqbase[i];
  double xlast[3];
  for (int ch = 0; ch < nch; ch++) {
    int ca = D == 3 ? ch >> 2 : ch >> 1;
    int cb = D == 3 ? (ch >> 1) & 1 : ch & 1;
    int cc = D == 3 ? ch & 1 : 0;
    double xc[3];
    xc[0] = pref[0] + dp / 2. * ca;
    xc[1] = pref[1] + dp / 2. * cb;
    if (D == 3) xc[2] = pref[2] + dp / 2. * cc;
    if (ch < nch - 1) {
      int64_t dst = base + ch;
      write_split_child<D, ORD>(g, P, i, dst, xc, pcoef);
      lcnt[dst] = 0;
      app_ref[seq + 1 + ch] = (int)dst;
      app_cell[seq + 1 + ch] = (uint32_t)cell_of<D>(g, xc);
    } else {
#pragma unroll
      for (int a = 0; a < D; a++) xlast[a] = xc[a];
    }
  }
  // parent takes the last child's state (Copy_split_state, v not copied)
  {
    double dr[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; a++) dr[a] = xlast[a] - P.x[a * cap + i];
    const int ph = P.phase[i];
    const double div = D == 3 ? 8.0 : 4.0;
    double sl[U];
    split_increment<D, ORD>(pcoef, dr, sl);
    if constexpr (ORD == 2) write_moved_copy<D, ORD>(P, i, pcoef, dr);
    double rho = P.rho[i] + sl[0];
    double rv[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; a++) rv[a] = P.rhov[a * cap + i] + sl[1 + a];
    double E = P.E[i] + sl[U - 1];
    double v[3];
#pragma unroll
    for (int a = 0; a < 3; a++) v[a] = rv[a] / rho;
    double sq = 0;
#pragma unroll
    for (int a = 0; a < D; a++) sq += v[a] * v[a];
    double ei = E - 0.5 * rho * sq;
    const int s = ph - 1;
    double Pp = eos_p(ei, Eos{g.gamma[s], g.pinf[s]});
    double cs = sqrt(g.gamma[s] * (g.pinf[s] + Pp) / rho);
#pragma unroll
    for (int a = 0; a < D; a++) {
      P.x[a * cap + i] = xlast[a];
      P.rhov[a * cap + i] = rv[a];
      P.tv[a * cap + i] = v[a];
    }
    P.vol[i] = P.vol[i] / div;
    P.mass[i] = P.mass[i] / div;
    P.h[i] = P.h[i] / 2.0;
    P.rho[i] = rho;
    P.E[i] = E;
    P.P[i] = Pp;
    P.c[i] = cs;
  }
  rm[i] = 1;
  app_ref[seq] = (int)i;
  app_cell[seq] = (uint32_t)cell_of<D>(g, xlast);
}

// ---------------------------------------------------------------------------
// affected particles of a split or merge pass
// ---------------------------------------------------------------------------
// Mark every cell within +-2 (clamped) of the centre cells.
template <int D>
__global__ void k_window(Geo g, const uint32_t* centers, int64_t k, uint8_t* win) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k) return;
  int c = (int)centers[t];
  int i, j, kk;
  cell_ijk(g, c, i, j, kk);
  const int rk = D == 3 ? 2 : 0;
  for (int a = max(i - 2, 0); a <= min(i + 2, g.N[0] - 1); a++)
    for (int b = max(j - 2, 0); b <= min(j + 2, g.N[1] - 1); b++)
      for (int q = max(kk - rk, 0); q <= min(kk + rk, g.N[2] - 1); q++)
        win[(a * g.N[1] + b) * g.N[2] + q] = 1;
}

// aff[e] = 1 for every fluid entry of the lists of particles with src[i].
__global__ void k_scatter_lists(int64_t n, const uint8_t* src, const int* lcnt, const int64_t* loff,
                                const int* pool, uint8_t* aff) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !src[i]) return;
  int64_t o = loff[i];
  for (int t = 0; t < lcnt[i]; t++) {
    int r = pool[o + t];
    if (r >= 0) aff[r] = 1;
  }
}

// Final affected flag, restricted to particles whose list can be read again
// in this step (below the merge volume), excluding removed particles.
template <int D>
__global__ void k_affected(Geo g, PtlDev P, int64_t n, const uint8_t* win, uint8_t* aff,
                           const uint8_t* removed, uint8_t* need) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[3];
#pragma unroll
  for (int a = 0; a < D; a++) x[a] = P.x[a * P.cap + i];
  bool a_ = aff[i] || win[cell_of<D>(g, x)];
  need[i] = (a_ && !removed[i] && P.vol[i] < g.vol_min) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// merge
// ---------------------------------------------------------------------------
// Merge candidate: volume below vol_min and, with per-cell bounds, a cell
// (of the current position) holding more than ppc_min fluid particles.
template <int D>
__device__ inline bool merge_small(const Geo& g, const PtlDev& P, int64_t i, const int* fcnt) {
  if (!(P.vol[i] < g.vol_min)) return false;
  if (g.ppc_min <= 0) return true;
  double x[3];
#pragma unroll
  for (int a = 0; a < D; a++) x[a] = P.x[a * P.cap + i];
  return fcnt[cell_of<D>(g, x)] > g.ppc_min;
}

template <int D>
__global__ void __launch_bounds__(128) k_merge_candidate(Geo g, PtlDev P, int64_t n, const int* lcnt,
                                                        const int64_t* loff, const int* pool,
                                                        int* cand, const int* __restrict__ fcnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t cap = P.cap;
  int best = -1;
  if (merge_small<D>(g, P, i, fcnt)) {
    double dss = g.dss0;
    double xi[3];
#pragma unroll
    for (int a = 0; a < D; a++) xi[a] = P.x[a * cap + i];
    const int ph = P.phase[i];
    int64_t o = loff[i];
    for (int t = 0; t < lcnt[i]; t++) {
      int r = pool[o + t];
      if (r < 0 || r == (int)i) continue;
      double sq = 0;
#pragma unroll
      for (int a = 0; a < D; a++) {
        double d = xi[a] - P.x[a * cap + r];
        sq += d * d;
      }
      double dist = sqrt(sq);
      if (P.vol[r] < g.vol_min && dist < dss && ph == P.phase[r] && merge_small<D>(g, P, r, fcnt)) {
        dss = dist;
        best = r;
      }
    }
  }
  cand[i] = best;
}

__global__ void k_merge_mark(int64_t n, const int* cand, int* surv) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cand[i];
  surv[i] = (c >= 0 && cand[c] == (int)i && (int)i < c) ? 1 : 0;
}

template <int D>
__global__ void k_merge_apply(Geo g, PtlDev P, int64_t n, const int* cand, const int* surv,
                              const int* sbase, uint8_t* rm, uint8_t* removed, uint8_t* src,
                              int* app_ref, uint32_t* app_cell) {
  constexpr int U = D + 2;
  const int M = g.ord == 2 ? (D == 2 ? 6 : 10) : D + 1;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !surv[i]) return;
  const int64_t cap = P.cap;
  const int p = cand[i];
  double vol0 = P.vol[i], vol1 = P.vol[p];
  double vol = vol0 + vol1;
  double xn[3];
#pragma unroll
  for (int a = 0; a < D; a++) xn[a] = (P.x[a * cap + i] * vol0 + P.x[a * cap + p] * vol1) / vol;
  P.mass[i] = P.mass[i] + P.mass[p];
  P.vol[i] = vol;
  P.h[i] = root_d<D>(vol) * kCellRatio / kCutOff;
#pragma unroll
  for (int a = 0; a < D; a++) P.x[a * cap + i] = xn[a];
  for (int q = 0; q < U * M; q++)
    P.coef[(size_t)q * cap + i] = (P.coef[(size_t)q * cap + i] * vol0 + P.coef[(size_t)q * cap + p] * vol1) / vol;
  double rho = (P.rho[i] * vol0 + P.rho[p] * vol1) / vol;
  double rv[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) rv[a] = (P.rhov[a * cap + i] * vol0 + P.rhov[a * cap + p] * vol1) / vol;
  if (D == 2) rv[2] = P.rhov[2 * cap + i];
  double E = (P.E[i] * vol0 + P.E[p] * vol1) / vol;
  double v[3];
#pragma unroll
  for (int a = 0; a < 3; a++) v[a] = rv[a] / rho;
  double sq = 0;
#pragma unroll
  for (int a = 0; a < D; a++) sq += v[a] * v[a];
  double ei = E - 0.5 * rho * sq;
  const int s = P.phase[i] - 1;
  double Pp = eos_p(ei, Eos{g.gamma[s], g.pinf[s]});
  double cs = sqrt(g.gamma[s] * (g.pinf[s] + Pp) / rho);
  P.rho[i] = rho;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    P.rhov[a * cap + i] = rv[a];
    P.v[a * cap + i] = v[a];
    P.tv[a * cap + i] = v[a];
  }
  P.E[i] = E;
  P.P[i] = Pp;
  P.c[i] = cs;
  rm[i] = 1;
  rm[p] = 1;
  removed[p] = 1;
  src[i] = 1;
  src[p] = 1;
  int seq = sbase[i];
  app_ref[seq] = (int)i;
  app_cell[seq] = (uint32_t)cell_of<D>(g, xn);
}

// ---------------------------------------------------------------------------
// the topology object
// ---------------------------------------------------------------------------
static inline unsigned g1(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1); }

int Topo::init(int64_t cap_, int64_t nwall_, int nc_, hipStream_t st) {
  cap = cap_;
  nwall = nwall_;
  nc = nc_;
  stream = st;
  int64_t ne = cap + nwall;
  size_t need = 0;
  auto A = [&](void** p, size_t bytes) -> int {
    PGDG_HIP_TRY(hipMalloc(p, bytes > 0 ? bytes : 16));
    need += bytes;
    return PGDG_OK;
  };
  PGDG_TRY(A((void**)&bin_off, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&bin_ent, sizeof(int) * ne));
  PGDG_TRY(A((void**)&bin_ent2, sizeof(int) * ne));
  PGDG_TRY(A((void**)&keys, sizeof(uint32_t) * ne));
  PGDG_TRY(A((void**)&keys2, sizeof(uint32_t) * ne));
  PGDG_TRY(A((void**)&vals, sizeof(int) * ne));
  PGDG_TRY(A((void**)&cnt, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&cnt2, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&off2, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&fcnt, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&win, nc));
  PGDG_TRY(A((void**)&rm, cap));
  PGDG_TRY(A((void**)&removed, cap));
  PGDG_TRY(A((void**)&src, cap));
  PGDG_TRY(A((void**)&aff, cap));
  PGDG_TRY(A((void**)&need_, cap));
  PGDG_TRY(A((void**)&type, cap));
  PGDG_TRY(A((void**)&ia, sizeof(int) * (cap + 1)));
  PGDG_TRY(A((void**)&ib, sizeof(int) * (cap + 1)));
  PGDG_TRY(A((void**)&ic, sizeof(int) * (cap + 1)));
  PGDG_TRY(A((void**)&id, sizeof(int) * (cap + 1)));
  PGDG_TRY(A((void**)&sa, sizeof(int) * (cap + 1)));
  PGDG_TRY(A((void**)&sb, sizeof(int) * (cap + 1)));
  PGDG_TRY(A((void**)&sc_, sizeof(int) * (cap + 1)));
  PGDG_TRY(A((void**)&sd, sizeof(int) * (cap + 1)));
  PGDG_TRY(A((void**)&app_ref, sizeof(int) * cap));
  PGDG_TRY(A((void**)&app_ref2, sizeof(int) * cap));
  PGDG_TRY(A((void**)&app_cell, sizeof(uint32_t) * cap));
  PGDG_TRY(A((void**)&app_cell2, sizeof(uint32_t) * cap));
  PGDG_TRY(A((void**)&lcnt, sizeof(int) * cap));
  PGDG_TRY(A((void**)&lcnt2, sizeof(int) * cap));
  PGDG_TRY(A((void**)&loff, sizeof(int64_t) * cap));
  PGDG_TRY(A((void**)&loff2, sizeof(int64_t) * cap));
  PGDG_TRY(A((void**)&lscan, sizeof(int64_t) * (cap + 1)));
  PGDG_TRY(A((void**)&lwide, sizeof(int64_t) * (cap + 1)));
  PGDG_TRY(A((void**)&newidx, sizeof(int) * cap));
  PGDG_TRY(A((void**)&host_small, sizeof(int64_t) * 8));
  pool_cap = cap * 40;
  PGDG_TRY(A((void**)&pool, sizeof(int) * pool_cap));
  key_bits = 1;
  while ((1LL << key_bits) < (long long)nc) key_bits++;
  // CUB temporary storage: max over the uses
  size_t t1 = 0, t2 = 0, t3 = 0;
  PGDG_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, keys, keys2, vals, bin_ent, (int)ne,
                                                  0, key_bits, stream));
  PGDG_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, ia, ib,
                                                (int)(cap > nc ? cap + 1 : nc + 1), stream));
  PGDG_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, t3, lwide, lscan, (int)(cap + 1), stream));
  tmp_bytes = t1 > t2 ? t1 : t2;
  if (t3 > tmp_bytes) tmp_bytes = t3;
  PGDG_TRY(A((void**)&tmp, tmp_bytes));
  bytes_total = need;
  // no lists yet (first step: every list is empty)
  hipLaunchKernelGGL(k_fill_i32, dim3(g1(cap)), dim3(256), 0, stream, lcnt, cap, 0);
  PGDG_HIP_TRY(hipMemsetAsync(loff, 0, sizeof(int64_t) * cap, stream));
  pool_used = 0;
  return PGDG_OK;
}

void Topo::release() {
  void* ptrs[] = {bin_off, bin_ent, bin_ent2, keys, keys2, vals, cnt, cnt2, off2, fcnt, win, rm, removed,
                  src, aff, need_, type, ia, ib, ic, id, sa, sb, sc_, sd, app_ref, app_ref2,
                  app_cell, app_cell2, lcnt, lcnt2, loff, loff2, lscan, lwide, newidx, host_small, pool, tmp};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

int Topo::scan_i32(const int* in, int* out, int64_t n) {
  size_t b = tmp_bytes;
  PGDG_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, b, in, out, (int)n, stream));
  return PGDG_OK;
}

// total = out[n-1] + in[n-1] read back to the host
int Topo::scan_total(const int* in, int* out, int64_t n, int64_t* total) {
  if (n == 0) {
    *total = 0;
    return PGDG_OK;
  }
  PGDG_TRY(scan_i32(in, out, n));
  int a = 0, b = 0;
  PGDG_HIP_TRY(hipMemcpyAsync(&a, out + n - 1, sizeof(int), hipMemcpyDeviceToHost, stream));
  PGDG_HIP_TRY(hipMemcpyAsync(&b, in + n - 1, sizeof(int), hipMemcpyDeviceToHost, stream));
  PGDG_HIP_TRY(hipStreamSynchronize(stream));
  *total = (int64_t)a + b;
  return PGDG_OK;
}

template <int D>
int Topo::build_bins(const Geo& g, const PtlDev& P, const WallDev& W, int64_t n) {
  int64_t ne = n + W.n;
  hipLaunchKernelGGL(k_bin_keys<D>, dim3(g1(ne)), dim3(256), 0, stream, g, P, W, n, keys, vals);
  size_t b = tmp_bytes;
  PGDG_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, b, keys, keys2, vals, bin_ent, (int)ne, 0,
                                                  key_bits, stream));
  PGDG_HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(int) * (nc + 1), stream));
  hipLaunchKernelGGL(k_hist, dim3(g1(ne)), dim3(256), 0, stream, keys2, ne, cnt);
  PGDG_TRY(scan_i32(cnt, bin_off, nc + 1));
  n_ent = ne;
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int Topo::edit_bins(int64_t k_app) {
  // sort the appended entries by cell, keeping append order within a cell
  PGDG_HIP_TRY(hipMemsetAsync(cnt2, 0, sizeof(int) * (nc + 1), stream));
  if (k_app > 0) {
    size_t b = tmp_bytes;
    PGDG_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, b, app_cell, app_cell2, app_ref, app_ref2,
                                                    (int)k_app, 0, key_bits, stream));
    hipLaunchKernelGGL(k_hist, dim3(g1(k_app)), dim3(256), 0, stream, app_cell2, k_app, cnt2);
  }
  PGDG_TRY(scan_i32(cnt2, off2, nc + 1));  // off2: appended offsets per cell
  hipLaunchKernelGGL(k_edit_count, dim3(g1(nc)), dim3(256), 0, stream, bin_off, bin_ent, nc, rm, cnt,
                     cnt2);
  int64_t tot = 0;
  PGDG_TRY(scan_total(cnt, ia, nc, &tot));
  // new offsets: ia[0..nc-1] plus the total at nc
  hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(64), 0, stream, ia + nc, (int)tot);
  hipLaunchKernelGGL(k_edit_fill, dim3(g1(nc)), dim3(256), 0, stream, bin_off, bin_ent, nc, rm, ia,
                     bin_ent2, off2, cnt2, app_ref2);
  PGDG_HIP_TRY(hipMemcpyAsync(bin_off, ia, sizeof(int) * (nc + 1), hipMemcpyDeviceToDevice, stream));
  int* t = bin_ent;
  bin_ent = bin_ent2;
  bin_ent2 = t;
  n_ent = tot;
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int Topo::build_lists(const Geo& g, const double* xq, const double* hq, const double* xall,
                      const WallDev& W, int64_t n, int status, bool reset_pool) {
  if (n == 0) return PGDG_OK;
  // counts
  hipLaunchKernelGGL((k_lists<D, false>), dim3(g1(n, 128)), dim3(128), 0, stream, g, xq, cap, hq,
                     xall, W, n, need_, status, bin_off, bin_ent, ia, loff, pool);
  // int64 scan of the counts
  hipLaunchKernelGGL(k_widen, dim3(g1(n)), dim3(256), 0, stream, ia, n, lwide);
  size_t b = tmp_bytes;
  PGDG_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, b, lwide, lscan, (int)n, stream));
  int64_t last_off = 0;
  int last_cnt = 0;
  PGDG_HIP_TRY(hipMemcpyAsync(&last_off, lscan + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, stream));
  PGDG_HIP_TRY(hipMemcpyAsync(&last_cnt, ia + n - 1, sizeof(int), hipMemcpyDeviceToHost, stream));
  PGDG_HIP_TRY(hipStreamSynchronize(stream));
  int64_t total = last_off + last_cnt;
  if (reset_pool) pool_used = 0;
  if (pool_used + total > pool_cap) return PGDG_ERR_CAPACITY;
  int64_t base = pool_used;
  if (reset_pool) {
    // particles not rebuilt keep no list
    hipLaunchKernelGGL(k_fill_i32, dim3(g1(n)), dim3(256), 0, stream, lcnt, n, 0);
  }
  hipLaunchKernelGGL(k_list_offsets, dim3(g1(n)), dim3(256), 0, stream, n, need_, lscan, base, loff);
  hipLaunchKernelGGL((k_lists<D, true>), dim3(g1(n, 128)), dim3(128), 0, stream, g, xq, cap, hq,
                     xall, W, n, need_, status, bin_off, bin_ent, ia, loff, pool);
  hipLaunchKernelGGL(k_list_commit, dim3(g1(n)), dim3(256), 0, stream, n, need_, ia, lscan, base,
                     lcnt, loff);
  pool_used = base + total;
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int Topo::compact(ParticleArrays& PA, int64_t n, int64_t* n_out) {
  // newidx = exclusive scan of keep
  hipLaunchKernelGGL(k_keep, dim3(g1(n)), dim3(256), 0, stream, removed, n, ia);
  int64_t kept = 0;
  PGDG_TRY(scan_total(ia, ib, n, &kept));
  hipLaunchKernelGGL(k_newidx, dim3(g1(n)), dim3(256), 0, stream, removed, ib, n, newidx);
  PGDG_TRY(PA.gather(newidx, n, stream));
  hipLaunchKernelGGL(k_gather<int>, dim3(g1(n)), dim3(256), 0, stream, lcnt, lcnt2, n, cap, 1, newidx);
  hipLaunchKernelGGL(k_gather<int64_t>, dim3(g1(n)), dim3(256), 0, stream, loff, loff2, n, cap, 1,
                     newidx);
  {
    int* t = lcnt;
    lcnt = lcnt2;
    lcnt2 = t;
    int64_t* u = loff;
    loff = loff2;
    loff2 = u;
  }
  hipLaunchKernelGGL(k_remap_refs, dim3(g1(n_ent)), dim3(256), 0, stream, bin_ent, n_ent, newidx);
  if (pool_used > 0)
    hipLaunchKernelGGL(k_remap_refs, dim3(g1(pool_used)), dim3(256), 0, stream, pool, pool_used, newidx);
  *n_out = kept;
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int Topo::mark_in(const Geo& g, const PtlDev& P, int64_t n) {
  if (g.sem != PGDG_SEM_S3 || n == 0) return PGDG_OK;
  hipLaunchKernelGGL(k_mark_in<D>, dim3(g1(n)), dim3(256), 0, stream, g, P, n);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int Topo::split(const Geo& g, ParticleArrays& PA, const WallDev& W, int64_t* n_io, int64_t counts[3]) {
  int64_t n = *n_io;
  counts[0] = counts[1] = counts[2] = 0;
  if (n == 0) return PGDG_OK;
  PtlDev P = PA.view();
  if (g.ppc_max > 0)
    hipLaunchKernelGGL(k_fluid_count, dim3(g1(nc)), dim3(256), 0, stream, bin_off, bin_ent, nc, fcnt);
  hipLaunchKernelGGL(k_split_classify<D>, dim3(g1(n)), dim3(256), 0, stream, g, P, n, type, ia, ib,
                     ic, id, fcnt);
  int64_t n_new = 0, n_seq = 0, n_in = 0, n_out = 0;
  PGDG_TRY(scan_total(ia, sa, n, &n_new));
  PGDG_TRY(scan_total(ib, sb, n, &n_seq));
  PGDG_TRY(scan_total(ic, sc_, n, &n_in));
  PGDG_TRY(scan_total(id, sd, n, &n_out));
  counts[0] = n_seq / (D == 3 ? 8 : 4);
  counts[1] = n_in;
  counts[2] = n_out;
  if (n_new == 0 && n_out == 0) return mark_in<D>(g, P, n);
  if (n + n_new > cap) return PGDG_ERR_CAPACITY;
  hipLaunchKernelGGL(k_fill_u8, dim3(g1(n + n_new)), dim3(256), 0, stream, rm, n + n_new, (uint8_t)0);
  hipLaunchKernelGGL(k_fill_u8, dim3(g1(n + n_new)), dim3(256), 0, stream, removed, n + n_new, (uint8_t)0);
  hipLaunchKernelGGL(k_fill_u8, dim3(g1(n + n_new)), dim3(256), 0, stream, aff, n + n_new, (uint8_t)0);
  hipLaunchKernelGGL(k_fill_u8, dim3(g1(nc)), dim3(256), 0, stream, win, (int64_t)nc, (uint8_t)0);
  // old lists of split parents and outlet particles -> affected
  hipLaunchKernelGGL(k_split_src, dim3(g1(n)), dim3(256), 0, stream, type, n, src);
  hipLaunchKernelGGL(k_scatter_lists, dim3(g1(n)), dim3(256), 0, stream, n, src, lcnt, loff, pool, aff);
  if (g.ord == 2)
    hipLaunchKernelGGL((k_split_generate<D, 2>), dim3(g1(n)), dim3(256), 0, stream, g, P, n, type, sa, sb,
                       sc_, (int)n_seq, rm, removed, app_ref, app_cell, lcnt);
  else
    hipLaunchKernelGGL((k_split_generate<D, 1>), dim3(g1(n)), dim3(256), 0, stream, g, P, n, type, sa, sb,
                       sc_, (int)n_seq, rm, removed, app_ref, app_cell, lcnt);
  int64_t k_app = n_seq + n_in;
  // centre cells of the windows are the cells of every appended entry
  if (k_app > 0) {
    hipLaunchKernelGGL(k_window<D>, dim3(g1(k_app)), dim3(256), 0, stream, g, app_cell, k_app, win);
    hipLaunchKernelGGL(k_mark_refs, dim3(g1(k_app)), dim3(256), 0, stream, app_ref, k_app, aff);
  }
  PGDG_TRY(edit_bins(k_app));
  int64_t n2 = n + n_new;
  hipLaunchKernelGGL(k_affected<D>, dim3(g1(n2)), dim3(256), 0, stream, g, P, n2, win, aff, removed,
                     need_);
  PGDG_TRY(build_lists<D>(g, P.x, nullptr, P.x, W, n2, 3, false));
  if (n_out > 0) {
    int64_t kept = 0;
    PGDG_TRY(compact(PA, n2, &kept));
    n2 = kept;
  }
  *n_io = n2;
  return mark_in<D>(g, PA.view(), n2);
}

template <int D>
int Topo::merge(const Geo& g, ParticleArrays& PA, const WallDev& W, int64_t* n_io, int64_t* merged) {
  int64_t n = *n_io;
  *merged = 0;
  if (n == 0) return PGDG_OK;
  PtlDev P = PA.view();
  if (g.ppc_min > 0)
    hipLaunchKernelGGL(k_fluid_count, dim3(g1(nc)), dim3(256), 0, stream, bin_off, bin_ent, nc, fcnt);
  hipLaunchKernelGGL(k_merge_candidate<D>, dim3(g1(n, 128)), dim3(128), 0, stream, g, P, n, lcnt, loff,
                     pool, ia, fcnt);
  hipLaunchKernelGGL(k_merge_mark, dim3(g1(n)), dim3(256), 0, stream, n, ia, ib);
  int64_t ns = 0;
  PGDG_TRY(scan_total(ib, sa, n, &ns));
  *merged = ns;
  if (ns == 0) return mark_in<D>(g, P, n);
  hipLaunchKernelGGL(k_fill_u8, dim3(g1(n)), dim3(256), 0, stream, rm, n, (uint8_t)0);
  hipLaunchKernelGGL(k_fill_u8, dim3(g1(n)), dim3(256), 0, stream, removed, n, (uint8_t)0);
  hipLaunchKernelGGL(k_fill_u8, dim3(g1(n)), dim3(256), 0, stream, aff, n, (uint8_t)0);
  hipLaunchKernelGGL(k_fill_u8, dim3(g1(n)), dim3(256), 0, stream, src, n, (uint8_t)0);
  hipLaunchKernelGGL(k_fill_u8, dim3(g1(nc)), dim3(256), 0, stream, win, (int64_t)nc, (uint8_t)0);
  hipLaunchKernelGGL(k_merge_apply<D>, dim3(g1(n)), dim3(256), 0, stream, g, P, n, ia, ib, sa, rm,
                     removed, src, app_ref, app_cell);
  // old lists of survivors and partners, survivors themselves, windows
  hipLaunchKernelGGL(k_scatter_lists, dim3(g1(n)), dim3(256), 0, stream, n, src, lcnt, loff, pool, aff);
  hipLaunchKernelGGL(k_window<D>, dim3(g1(ns)), dim3(256), 0, stream, g, app_cell, ns, win);
  hipLaunchKernelGGL(k_mark_refs, dim3(g1(ns)), dim3(256), 0, stream, app_ref, ns, aff);
  PGDG_TRY(edit_bins(ns));
  hipLaunchKernelGGL(k_affected<D>, dim3(g1(n)), dim3(256), 0, stream, g, P, n, win, aff, removed,
                     need_);
  PGDG_TRY(build_lists<D>(g, P.x, nullptr, P.x, W, n, 3, false));
  int64_t kept = 0;
  PGDG_TRY(compact(PA, n, &kept));
  *n_io = kept;
  return mark_in<D>(g, PA.view(), kept);
}

template <int D>
int Topo::end_of_step_lists(const Geo& g, const PtlDev& P, const double* xb, const double* hb,
                            const WallDev& W, int64_t n) {
  if (n == 0) return PGDG_OK;
  hipLaunchKernelGGL(k_flag_lists<D>, dim3(g1(n)), dim3(256), 0, stream, g, P, n, need_);
  return build_lists<D>(g, xb, hb, xb, W, n, 2, true);
}

#define PGDG_INST(D)                                                                              \
  template int Topo::build_bins<D>(const Geo&, const PtlDev&, const WallDev&, int64_t);           \
  template int Topo::split<D>(const Geo&, ParticleArrays&, const WallDev&, int64_t*, int64_t*);   \
  template int Topo::merge<D>(const Geo&, ParticleArrays&, const WallDev&, int64_t*, int64_t*);   \
  template int Topo::end_of_step_lists<D>(const Geo&, const PtlDev&, const double*, const double*, \
                                          const WallDev&, int64_t);                               \
  template int Topo::mark_in<D>(const Geo&, const PtlDev&, int64_t);
PGDG_INST(2)
PGDG_INST(3)
#undef PGDG_INST

}  // namespace pgdg
