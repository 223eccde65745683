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
// Determinism: every ordering decision is an exclusive scan or a stable
// bucket sort (dev_scan.hip); no atomic decides a position.  A cell list
// edit is "keep the old entries that were not removed, in order, then append
// the new entries in the reference's append order", computed per cell.
//
// No host round trips inside the passes: counts (particles, entries, events,
// pool use) stay in a TopoCounters block on the device, every particle loop
// is a grid-stride loop over the device count, and a pass with nothing to do
// is gated off by a device flag.  Removed particles (outlet, merge partners)
// are flagged dead and skipped; finish() compacts them once after the last
// pass.  Removal is stable, so compacting once leaves the same order as
// compacting after every pass.
//
// Neighbour lists are materialised only where the reference reads a list
// that is not rebuilt first: the split/merge passes read the lists of split
// parents, outlet particles and particles below the merge volume.  At the
// end of a step the lists of exactly those particles are built (from the
// positions and smoothing lengths saved at the start of the last stage, which
// is when the reference built them); each split/merge pass rebuilds the
// lists of the affected particles that are below the merge volume.
#include "dev_common.hpp"
#include "dev_scan.hpp"
#include "hip_util.hpp"
#include "step_topo.hpp"

namespace pgdg {

constexpr int kLoopBlocks = 2048;  // grid of the particle loops
#define PGDG_LOOP(i, n) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

static inline unsigned g1(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1); }

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__global__ void k_mark_refs(const int* refs, Len len, uint8_t* flag) {
  const long long k = len.get();
  PGDG_LOOP(e, k) {
    if (refs[e] >= 0) flag[refs[e]] = 1;
  }
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
__global__ void k_bin_keys(Geo g, PtlDev P, WallDev W, int64_t n, uint32_t* keys) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n + W.n) return;
  double x[3];
  if (e < n) {
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = P.x[a * P.cap + e];
  } else {
    int64_t w = e - n;
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = W.x[a * W.n + w];
  }
  keys[e] = (uint32_t)cell_of<D>(g, x);
}

// Entries in bucket order: fluid particle e (e < n) or wall -(w + 2).
__global__ void k_bin_fill(const int* __restrict__ order, int64_t n, int64_t k, int* ent) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k) return;
  const int e = order[q];
  ent[q] = e < n ? e : -(int)(e - n) - 2;
}

__global__ void k_set_nent(TopoCounters* tc, long long k) {
  if (threadIdx.x == 0 && blockIdx.x == 0) tc->nent = k;
}

// ---------------------------------------------------------------------------
// cell-list edit: keep old entries whose particle is not flagged in `rm`,
// then append the pass's new entries (bucketed by cell, append order kept)
// ---------------------------------------------------------------------------
__global__ void k_edit_count(const int* off, const int* ent, int nc, const uint8_t* rm, int* cnt,
                             const int* aoff, const long long* gate) {
  if (*gate == 0) return;
  PGDG_LOOP(c, nc) {
    int k = 0;
    for (int e = off[c]; e < off[c + 1]; e++) {
      int r = ent[e];
      if (r >= 0 && rm[r]) continue;
      k++;
    }
    cnt[c] = k + aoff[c + 1] - aoff[c];
  }
}

__global__ void k_edit_fill(const int* off, const int* ent, int nc, const uint8_t* rm, const int* noff,
                            int* nent, const int* aoff, const int* aorder, const int* app_ref,
                            const long long* gate, long long ent_cap, TopoCounters* tc) {
  if (*gate == 0) return;
  PGDG_LOOP(c, nc) {
    int o = noff[c];
    if (o + (off[c + 1] - off[c]) + (aoff[c + 1] - aoff[c]) > ent_cap) {  // inconsistent counts
      tc->err = PGDG_ERR_STATE;
      continue;
    }
    for (int e = off[c]; e < off[c + 1]; e++) {
      int r = ent[e];
      if (r >= 0 && rm[r]) continue;
      nent[o++] = r;
    }
    for (int t = aoff[c]; t < aoff[c + 1]; t++) nent[o++] = app_ref[aorder[t]];
  }
}

__global__ void k_copy_len(const int* __restrict__ src, int* dst, Len len) {
  const long long n = len.get();
  PGDG_LOOP(i, n) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// neighbour lists (status 2: radius cells, dist <= h*CUT_OFF; status 3:
// 2 cells, dist <= 2 p_sz)
// ---------------------------------------------------------------------------
template <int D, bool FILL>
__global__ void __launch_bounds__(128) k_lists(Geo g, const double* __restrict__ xq, int64_t xcap,
                                               const double* __restrict__ hq,
                                               const double* __restrict__ xall, WallDev W, Len len,
                                               const uint8_t* __restrict__ need, int status,
                                               const int* __restrict__ bin_off,
                                               const int* __restrict__ bin_ent, int* cnt,
                                               const int64_t* __restrict__ loff, int* pool,
                                               const long long* fill_ok) {
  const long long n = len.get();
  if (FILL && *fill_ok == 0) return;
  PGDG_LOOP(i, n) {
    if (!need[i]) {
      if (!FILL) cnt[i] = 0;
      continue;
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
}

__global__ void k_widen(const int* in, Len len, long long* out) {
  const long long n = len.get();
  PGDG_LOOP(e, n) out[e] = in[e];
}

// Reserve the pool entries of the lists being built (or flag the overflow).
__global__ void k_list_alloc(TopoCounters* tc, int gated) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (gated && tc->active == 0) {
    tc->list_ok = 0;
    return;
  }
  const long long base = tc->pool_used, tot = tc->lists_total;
  const bool ok = base + tot <= tc->pool_cap;
  tc->list_base = base;
  tc->list_ok = ok ? 1 : 0;
  if (ok)
    tc->pool_used = base + tot;
  else
    tc->err = PGDG_ERR_CAPACITY;
}

__global__ void k_list_offsets(Len len, const uint8_t* need, const long long* scan, const TopoCounters* tc,
                               int64_t* loff) {
  const long long n = len.get();
  if (tc->list_ok == 0) return;
  const long long base = tc->list_base;
  PGDG_LOOP(i, n) {
    if (need[i]) loff[i] = base + scan[i];
  }
}

// Commit the counts (zero when the pool overflowed: no stale offsets).
__global__ void k_list_commit(Len len, const uint8_t* need, const int* cnt, const TopoCounters* tc, int* lcnt) {
  const long long n = len.get();
  const bool ok = tc->list_ok != 0;
  PGDG_LOOP(i, n) {
    if (need[i]) lcnt[i] = ok ? cnt[i] : 0;
  }
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

__global__ void k_keep(const uint8_t* dead, int64_t n, int* keep) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keep[i] = dead[i] ? 0 : 1;
}

__global__ void k_newidx(const uint8_t* dead, const int* scan, int64_t n, int* newidx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  newidx[i] = dead[i] ? -1 : scan[i];
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

__global__ void k_compacted(TopoCounters* tc, long long n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    tc->n_total = n;
    tc->n_dead = 0;
  }
}

// ---------------------------------------------------------------------------
// mark_in (S3:mpm.cpp:1398-1399, 1698-1699)
// ---------------------------------------------------------------------------
template <int D>
__global__ void k_mark_in(Geo g, PtlDev P, Len len, const uint8_t* dead) {
  const long long n = len.get();
  PGDG_LOOP(i, n) {
    if (!dead[i]) P.mark_in[i] = (P.x[i] - g.box_l[0] < root_d<D>(P.vol[i])) ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// per-cell bounds (no reference counterpart): a candidate's effect on a cell
// is the net change of its fluid count; candidates are taken in particle
// order and each counts the changes of every earlier candidate of the pass,
// admitted or not, so that admissions are independent of one another.
// ---------------------------------------------------------------------------
struct NetList {
  int cell[9], net[9], k;
  __device__ void add(int c, int d) {
    for (int t = 0; t < k; t++)
      if (cell[t] == c) {
        net[t] += d;
        return;
      }
    cell[k] = c;
    net[k++] = d;
  }
};

// Split of the particle at xp with spacing dp: the 2^D children (the parent
// takes the last one's place) against the parent's cell.
template <int D>
__device__ inline void split_net(const Geo& g, const double* xp, double dp, NetList& L) {
  L.k = 0;
  L.add(cell_of<D>(g, xp), -1);
  double pref[3] = {xp[0] - dp / 4., xp[1] - dp / 4., D == 3 ? xp[2] - dp / 4. : 0.};
  const int nch = D == 3 ? 8 : 4;
  for (int ch = 0; ch < nch; ch++) {
    const int ca = D == 3 ? ch >> 2 : ch >> 1;
    const int cb = D == 3 ? (ch >> 1) & 1 : ch & 1;
    const int cc = D == 3 ? ch & 1 : 0;
    double xc[3];
    xc[0] = pref[0] + dp / 2. * ca;
    xc[1] = pref[1] + dp / 2. * cb;
    xc[2] = D == 3 ? pref[2] + dp / 2. * cc : 0.;
    L.add(cell_of<D>(g, xc), 1);
  }
}

// Merge of survivor i and partner p: both leave their cells, the merged
// particle enters the cell of the volume-weighted position.
template <int D>
__device__ inline void merge_net(const Geo& g, const PtlDev& P, int64_t i, int64_t p, NetList& L) {
  const int64_t cap = P.cap;
  const double vol0 = P.vol[i], vol1 = P.vol[p], vol = vol0 + vol1;
  double xi[3], xp[3], xn[3];
#pragma unroll
  for (int a = 0; a < D; a++) {
    xi[a] = P.x[a * cap + i];
    xp[a] = P.x[a * cap + p];
    xn[a] = (xi[a] * vol0 + xp[a] * vol1) / vol;
  }
  L.k = 0;
  L.add(cell_of<D>(g, xi), -1);
  L.add(cell_of<D>(g, xp), -1);
  L.add(cell_of<D>(g, xn), 1);
}

// Per cell: walk the contributions in candidate order; a candidate whose
// running total would take the cell past its bound is refused.
__global__ void k_bound_walk(const int* __restrict__ coff, const int* __restrict__ corder,
                             const int* __restrict__ ccand, const int* __restrict__ ccost,
                             const int* __restrict__ fcnt, int nc, int lim, int upper, uint8_t* reject,
                             const long long* gate) {
  if (*gate == 0) return;
  PGDG_LOOP(c, nc) {
    int run = 0;
    for (int t = coff[c]; t < coff[c + 1]; t++) {
      const int e = corder[t];
      run += ccost[e];
      const bool bad = upper ? fcnt[c] + run > lim : fcnt[c] - run < lim;
      if (bad) reject[ccand[e]] = 1;
    }
  }
}

__global__ void k_contrib_check(TopoCounters* tc, long long ccap) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && tc->n_contrib > ccap) {
    tc->err = PGDG_ERR_CAPACITY;
    tc->n_contrib = 0;
  }
}

// ---------------------------------------------------------------------------
// split
// ---------------------------------------------------------------------------
enum { T_NONE = 0, T_SPLIT = 1, T_INLET = 2, T_OUTLET = 3 };

// Event totals of a pass: per-lane counts added once per lane that has any
// (sums, so the totals do not depend on the order of the atomics).
__device__ inline void add_counts(TopoCounters* tc, long long cs, long long ci, long long co) {
  if (cs) atomicAdd((unsigned long long*)&tc->tot_split, (unsigned long long)cs);
  if (ci) atomicAdd((unsigned long long*)&tc->tot_inlet, (unsigned long long)ci);
  if (co) atomicAdd((unsigned long long*)&tc->tot_outlet, (unsigned long long)co);
}

__global__ void k_pass_zero(TopoCounters* tc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  tc->tot_split = tc->tot_inlet = tc->tot_outlet = 0;
  tc->n_pairs = 0;
}
enum { F_SPLIT = 1, F_INLET = 2, F_OUTLET = 4 };

__device__ inline uint8_t split_type(int flags) {
  return (flags & F_SPLIT) ? T_SPLIT : (flags & F_INLET) ? T_INLET : (flags & F_OUTLET) ? T_OUTLET : T_NONE;
}

// Classify every particle (split / inlet copy / outlet, in the reference's
// precedence) and clear the pass flags.  With per-cell bounds the split
// flag is provisional: flags are kept in `type` and lc32 counts the cells a
// split candidate adds to.
template <int D>
__global__ void k_split_classify(Geo g, PtlDev P, TopoCounters* tc, uint8_t* type, int* is_s,
                                 int* is_i, int* is_o, int* ncontrib, int bounds, uint8_t* src,
                                 uint8_t* reject) {
  const long long n = tc->n_total;
  const bool io = g.sem == PGDG_SEM_S3;
  long long cs = 0, ci = 0, co = 0;
  PGDG_LOOP(i, n) {
    const double vol = P.vol[i];
    const double dp = root_d<D>(vol);
    const double x0 = P.x[i];
    const bool f1 = vol > g.vol_max;
    const bool f2 = io && x0 - g.box_l[0] >= dp && P.mark_in[i] == 1;
    const bool f3 = io && x0 > g.box_r[0];
    const int flags = (f1 ? F_SPLIT : 0) | (f2 ? F_INLET : 0) | (f3 ? F_OUTLET : 0);
    if (bounds) {
      type[i] = (uint8_t)flags;
      reject[i] = 0;
      int k = 0;
      if (f1) {
        double x[3];
#pragma unroll
        for (int a = 0; a < D; a++) x[a] = P.x[a * P.cap + i];
        NetList L;
        split_net<D>(g, x, dp, L);
        for (int t = 0; t < L.k; t++) k += L.net[t] > 0 ? 1 : 0;
      }
      ncontrib[i] = k;
    } else {
      const uint8_t t = split_type(flags);
      type[i] = t;
      is_s[i] = t == T_SPLIT;
      is_i[i] = t == T_INLET;
      is_o[i] = t == T_OUTLET;
      src[i] = (t == T_SPLIT || t == T_OUTLET) ? 1 : 0;
      cs += t == T_SPLIT;
      ci += t == T_INLET;
      co += t == T_OUTLET;
    }
  }
  add_counts(tc, cs, ci, co);
}

template <int D>
__global__ void k_split_contrib(Geo g, PtlDev P, const TopoCounters* tc, const uint8_t* type, const int* cbase,
                                long long ccap, uint32_t* ckey, int* ccand, int* ccost) {
  const long long n = tc->n_total;
  PGDG_LOOP(i, n) {
    if (!(type[i] & F_SPLIT)) continue;
    double x[3];
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = P.x[a * P.cap + i];
    NetList L;
    split_net<D>(g, x, root_d<D>(P.vol[i]), L);
    long long o = cbase[i];
    for (int t = 0; t < L.k; t++)
      if (L.net[t] > 0) {
        if (o < ccap) {
          ckey[o] = (uint32_t)L.cell[t];
          ccand[o] = (int)i;
          ccost[o] = L.net[t];
        }
        o++;
      }
  }
}

__global__ void k_split_final(TopoCounters* tc, uint8_t* type, const uint8_t* reject, int* is_s, int* is_i,
                              int* is_o, uint8_t* src) {
  const long long n = tc->n_total;
  long long cs = 0, ci = 0, co = 0;
  PGDG_LOOP(i, n) {
    int flags = type[i];
    if (reject[i]) flags &= ~F_SPLIT;
    const uint8_t t = split_type(flags);
    type[i] = t;
    is_s[i] = t == T_SPLIT;
    is_i[i] = t == T_INLET;
    is_o[i] = t == T_OUTLET;
    src[i] = (t == T_SPLIT || t == T_OUTLET) ? 1 : 0;
    cs += t == T_SPLIT;
    ci += t == T_INLET;
    co += t == T_OUTLET;
  }
  add_counts(tc, cs, ci, co);
}

// Totals of the pass: new particles, appended entries, capacity.
__global__ void k_split_begin(TopoCounters* tc, long long cap, int nch, int nc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const long long ns = tc->tot_split, ni = tc->tot_inlet, no = tc->tot_outlet;
  const long long n_new = (nch - 1) * ns + ni;
  long long active = (n_new + no) > 0 ? 1 : 0;
  if (tc->n_total + n_new > cap) {
    tc->err = PGDG_ERR_CAPACITY;
    active = 0;
  }
  tc->active = active;
  tc->win_len = active ? nc : 0;
  tc->k_app = active ? nch * ns + ni : 0;
  tc->n_split = ns;
  tc->n_inlet = ni;
  tc->n_outlet = no;
}

__global__ void k_split_end(TopoCounters* tc, int nch) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || tc->active == 0) return;
  tc->n_total += (nch - 1) * tc->tot_split + tc->tot_inlet;
  tc->n_dead += tc->tot_outlet;
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

__device__ inline void clear_new(int64_t dst, uint8_t* rm, uint8_t* dead, uint8_t* aff, uint8_t* src, int* lcnt) {
  rm[dst] = 0;
  dead[dst] = 0;
  aff[dst] = 0;
  src[dst] = 0;
  lcnt[dst] = 0;
}

// Children of split parents and inlet copies, appended after the n_total
// particles in the reference's order (parents in particle order; per
// parent its 2^D - 1 children, or its inlet copy); outlet particles are
// flagged dead.  sbase/ibase: exclusive scans of the split and inlet
// indicators.
template <int D, int ORD>
__global__ void k_split_generate(Geo g, PtlDev P, TopoCounters* tc, const uint8_t* __restrict__ type,
                                 const int* __restrict__ sbase, const int* __restrict__ ibase, uint8_t* rm,
                                 uint8_t* dead, uint8_t* aff, uint8_t* src, int* app_ref, uint32_t* app_cell,
                                 int* lcnt) {
  constexpr int U = D + 2, M = Basis<D, ORD>::M;
  if (tc->active == 0) return;
  const long long n = tc->n_total;
  const int nch = D == 3 ? 8 : 4;
  const long long s_total = (long long)nch * tc->tot_split;
  const int64_t cap = P.cap;
  PGDG_LOOP(i, n) {
    const uint8_t t = type[i];
    if (t == T_OUTLET) {
      rm[i] = 1;
      dead[i] = 1;
      continue;
    }
    if (t != T_SPLIT && t != T_INLET) continue;
    const int64_t newbase = (int64_t)(nch - 1) * sbase[i] + ibase[i];
    if (n + newbase + (t == T_SPLIT ? nch - 1 : 1) > cap || (long long)nch * sbase[i] + nch > cap) {
      tc->err = PGDG_ERR_STATE;
      continue;
    }
    double dp = root_d<D>(P.vol[i]);
    double pcoef[U * M];
    for (int q = 0; q < U * M; q++) pcoef[q] = P.coef[(size_t)q * cap + i];
    if (t == T_INLET) {
      // Set_ptcl (S3:particle.cpp:696-722) of a copy at x - dp
      int64_t dst = n + newbase;
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
      clear_new(dst, rm, dead, aff, src, lcnt);
      double xd[3];
#pragma unroll
      for (int a = 0; a < D; a++) xd[a] = P.x[a * cap + dst];
      const long long seq = s_total + ibase[i];
      app_ref[seq] = (int)dst;
      app_cell[seq] = (uint32_t)cell_of<D>(g, xd);
      continue;
    }
    // split into 2^D children at (x - dp/4) + dp/2 (ci, cj[, ck])
    double pref[3];
#pragma unroll
    for (int a = 0; a < D; a++) pref[a] = P.x[a * cap + i] - dp / 4.;
    const int64_t base = n + newbase;
    const long long seq = (long long)nch * sbase[i];
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
        clear_new(dst, rm, dead, aff, src, lcnt);
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
}

// ---------------------------------------------------------------------------
// affected particles of a split or merge pass
// ---------------------------------------------------------------------------
// Mark every cell within +-2 (clamped) of the centre cells.
template <int D>
__global__ void k_window(Geo g, const uint32_t* centers, Len len, uint8_t* win) {
  const long long k = len.get();
  PGDG_LOOP(t, k) {
    int c = (int)centers[t];
    int i, j, kk;
    cell_ijk(g, c, i, j, kk);
    const int rk = D == 3 ? 2 : 0;
    for (int a = max(i - 2, 0); a <= min(i + 2, g.N[0] - 1); a++)
      for (int b = max(j - 2, 0); b <= min(j + 2, g.N[1] - 1); b++)
        for (int q = max(kk - rk, 0); q <= min(kk + rk, g.N[2] - 1); q++)
          win[(a * g.N[1] + b) * g.N[2] + q] = 1;
  }
}

// aff[e] = 1 for every fluid entry of the lists of particles with src[i].
__global__ void k_scatter_lists(Len len, const uint8_t* src, const int* lcnt, const int64_t* loff,
                                const int* pool, uint8_t* aff) {
  const long long n = len.get();
  PGDG_LOOP(i, n) {
    if (!src[i]) continue;
    int64_t o = loff[i];
    for (int t = 0; t < lcnt[i]; t++) {
      int r = pool[o + t];
      if (r >= 0) aff[r] = 1;
    }
  }
}

// Final affected flag, restricted to particles whose list can be read again
// in this step (below the merge volume), excluding removed particles.
template <int D>
__global__ void k_affected(Geo g, PtlDev P, Len len, const uint8_t* win, const uint8_t* aff,
                           const uint8_t* dead, uint8_t* need) {
  const long long n = len.get();
  PGDG_LOOP(i, n) {
    double x[3];
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = P.x[a * P.cap + i];
    bool a_ = aff[i] || win[cell_of<D>(g, x)];
    need[i] = (a_ && !dead[i] && P.vol[i] < g.vol_min) ? 1 : 0;
  }
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
__global__ void __launch_bounds__(128) k_merge_candidate(Geo g, PtlDev P, const TopoCounters* tc, const int* lcnt,
                                                        const int64_t* loff, const int* pool, int* cand,
                                                        const int* __restrict__ fcnt, const uint8_t* dead,
                                                        uint8_t* rm, uint8_t* aff, uint8_t* src,
                                                        uint8_t* reject) {
  const long long n = tc->n_total;
  const int64_t cap = P.cap;
  PGDG_LOOP(i, n) {
    if (g.ppc_min > 0) reject[i] = 0;
    int best = -1;
    if (!dead[i] && merge_small<D>(g, P, i, fcnt)) {
      double dss = g.dss0;
      double xi[3];
#pragma unroll
      for (int a = 0; a < D; a++) xi[a] = P.x[a * cap + i];
      const int ph = P.phase[i];
      int64_t o = loff[i];
      for (int t = 0; t < lcnt[i]; t++) {
        int r = pool[o + t];
        if (r < 0 || r == (int)i || dead[r]) continue;
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
}

// Survivors: mutual candidates, the lower index survives.  With per-cell
// bounds, ncontrib counts the cells a pair removes from.
template <int D>
__global__ void k_merge_mark(Geo g, PtlDev P, TopoCounters* tc, const int* cand, int* surv, int bounds,
                             int* ncontrib) {
  const long long n = tc->n_total;
  long long cnt = 0;
  PGDG_LOOP(i, n) {
    int c = cand[i];
    const int s = (c >= 0 && cand[c] == (int)i && (int)i < c) ? 1 : 0;
    surv[i] = s;
    cnt += s;
    if (bounds) {
      int k = 0;
      if (s) {
        NetList L;
        merge_net<D>(g, P, i, c, L);
        for (int t = 0; t < L.k; t++) k += L.net[t] < 0 ? 1 : 0;
      }
      ncontrib[i] = k;
    }
  }
  if (!bounds && cnt) atomicAdd((unsigned long long*)&tc->n_pairs, (unsigned long long)cnt);
}

template <int D>
__global__ void k_merge_contrib(Geo g, PtlDev P, const TopoCounters* tc, const int* cand, const int* surv,
                                const int* cbase, long long ccap, uint32_t* ckey, int* ccand, int* ccost) {
  const long long n = tc->n_total;
  PGDG_LOOP(i, n) {
    if (!surv[i]) continue;
    NetList L;
    merge_net<D>(g, P, i, cand[i], L);
    long long o = cbase[i];
    for (int t = 0; t < L.k; t++)
      if (L.net[t] < 0) {
        if (o < ccap) {
          ckey[o] = (uint32_t)L.cell[t];
          ccand[o] = (int)i;
          ccost[o] = -L.net[t];
        }
        o++;
      }
  }
}

__global__ void k_merge_final(TopoCounters* tc, const uint8_t* reject, int* surv) {
  const long long n = tc->n_total;
  long long cnt = 0;
  PGDG_LOOP(i, n) {
    if (reject[i]) surv[i] = 0;
    cnt += surv[i];
  }
  if (cnt) atomicAdd((unsigned long long*)&tc->n_pairs, (unsigned long long)cnt);
}

__global__ void k_merge_begin(TopoCounters* tc, int nc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const long long np_ = tc->n_pairs;
  tc->active = np_ > 0 ? 1 : 0;
  tc->win_len = np_ > 0 ? nc : 0;
  tc->k_app = np_;
  tc->n_merge += np_;
}

__global__ void k_merge_end(TopoCounters* tc) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || tc->active == 0) return;
  tc->n_dead += tc->n_pairs;
}

template <int D>
__global__ void k_merge_apply(Geo g, PtlDev P, const TopoCounters* tc, const int* cand, const int* surv,
                              const int* sbase, uint8_t* rm, uint8_t* dead, uint8_t* src, int* app_ref,
                              uint32_t* app_cell) {
  constexpr int U = D + 2;
  if (tc->active == 0) return;
  const long long n = tc->n_total;
  const int M = g.ord == 2 ? (D == 2 ? 6 : 10) : D + 1;
  const int64_t cap = P.cap;
  PGDG_LOOP(i, n) {
    if (!surv[i]) continue;
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
    dead[p] = 1;
    src[i] = 1;
    src[p] = 1;
    const int seq = sbase[i];
    if (seq >= cap) continue;
    app_ref[seq] = (int)i;
    app_cell[seq] = (uint32_t)cell_of<D>(g, xn);
  }
}

// ---------------------------------------------------------------------------
// the topology object
// ---------------------------------------------------------------------------
__global__ void k_tc_reset(TopoCounters* tc, long long n, long long pool_cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  TopoCounters z = {};
  z.n_total = n;
  z.pool_cap = pool_cap;
  *tc = z;
}

__global__ void k_tc_pool(TopoCounters* tc, long long used, long long cap_) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  tc->pool_used = used;
  tc->pool_cap = cap_;
}

__global__ void k_tc_step(TopoCounters* tc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  tc->n_merge = 0;
}

__global__ void k_tc_lists_full(TopoCounters* tc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  tc->list_base = 0;
  tc->list_ok = 1;
  tc->pool_used = tc->lists_total;
}

int Topo::init(int64_t cap_, int64_t nwall_, int nc_, hipStream_t st) {
  cap = cap_;
  nwall = nwall_;
  nc = nc_;
  stream = st;
  const int64_t ne = cap + nwall;
  ccap = cap;
  size_t need = 0;
  auto A = [&](void** p, size_t bytes) -> int {
    PGDG_HIP_TRY(hipMalloc(p, bytes > 0 ? bytes : 16));
    need += bytes;
    return PGDG_OK;
  };
  PGDG_TRY(A((void**)&tc, sizeof(TopoCounters)));
  PGDG_TRY(A((void**)&bin_off, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&off_tmp, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&bin_ent, sizeof(int) * ne));
  PGDG_TRY(A((void**)&ent_tmp, sizeof(int) * ne));
  PGDG_TRY(A((void**)&keys, sizeof(uint32_t) * ne));
  PGDG_TRY(A((void**)&order, sizeof(int) * (ne > ccap ? ne : ccap)));
  PGDG_TRY(A((void**)&cursor, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&cnt, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&coff, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&fcnt, sizeof(int) * (nc + 1)));
  PGDG_TRY(A((void**)&win, nc));
  uint8_t** u8s[] = {&rm, &dead, &src, &aff, &need_, &type, &reject};
  for (uint8_t** p : u8s) PGDG_TRY(A((void**)p, cap));
  int** i32s[] = {&is_a, &is_b, &is_c, &cand, &app_ref, &lcnt, &lcnt2, &lc32, &newidx, &cbase};
  for (int** p : i32s) PGDG_TRY(A((void**)p, sizeof(int) * (cap + 1)));
  PGDG_TRY(A((void**)&app_cell, sizeof(uint32_t) * cap));
  PGDG_TRY(A((void**)&ckey, sizeof(uint32_t) * ccap));
  PGDG_TRY(A((void**)&ccand, sizeof(int) * ccap));
  PGDG_TRY(A((void**)&ccost, sizeof(int) * ccap));
  PGDG_TRY(A((void**)&loff, sizeof(int64_t) * cap));
  PGDG_TRY(A((void**)&loff2, sizeof(int64_t) * cap));
  PGDG_TRY(A((void**)&lscan, sizeof(int64_t) * (cap + 1)));
  PGDG_TRY(A((void**)&part32, sizeof(int) * kScanBlocks));
  PGDG_TRY(A((void**)&part64, sizeof(long long) * kScanBlocks));
  pool_cap = cap * 40;
  PGDG_TRY(A((void**)&pool, sizeof(int) * pool_cap));
  bytes_total = need;
  PGDG_TRY(fill_u8(dead, len_host(cap), 0, stream));
  return reset(0);
}

void Topo::release() {
  void* ptrs[] = {tc, bin_off, off_tmp, bin_ent, ent_tmp, keys, order, cursor, cnt, coff, fcnt, win, rm, dead,
                  src, aff, need_, type, reject, is_a, is_b, is_c, cand, app_ref, lcnt, lcnt2, lc32, newidx,
                  cbase, app_cell, ckey, ccand, ccost, loff, loff2, lscan, part32, part64, pool};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

// After an upload: n live particles, no lists (the reference's lists are
// empty after its set-up), no dead particles.
int Topo::reset(int64_t n) {
  hipLaunchKernelGGL(k_tc_reset, dim3(1), dim3(64), 0, stream, tc, (long long)n, (long long)pool_cap);
  PGDG_TRY(fill_i32(lcnt, len_host(cap), 0, stream));
  PGDG_HIP_TRY(hipMemsetAsync(loff, 0, sizeof(int64_t) * cap, stream));
  PGDG_TRY(fill_u8(dead, len_host(cap), 0, stream));
  pool_used = 0;
  n_ent = 0;
  hc = TopoCounters{};
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int Topo::set_pool_used(int64_t used) {
  pool_used = used;
  hipLaunchKernelGGL(k_tc_pool, dim3(1), dim3(64), 0, stream, tc, (long long)used, (long long)pool_cap);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int Topo::fluid_counts() {
  hipLaunchKernelGGL(k_fluid_count, dim3(g1(nc)), dim3(256), 0, stream, (const int*)bin_off, (const int*)bin_ent,
                     nc, fcnt);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int Topo::reset_step_counts() {
  hipLaunchKernelGGL(k_tc_step, dim3(1), dim3(64), 0, stream, tc);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int Topo::build_bins(const Geo& g, const PtlDev& P, const WallDev& W, int64_t n) {
  const int64_t ne = n + W.n;
  hipLaunchKernelGGL(k_bin_keys<D>, dim3(g1(ne)), dim3(256), 0, stream, g, P, W, n, keys);
  PGDG_TRY(bucket_stable(keys, len_host(ne), nc, bin_off, cursor, order, part32, stream));
  hipLaunchKernelGGL(k_bin_fill, dim3(g1(ne)), dim3(256), 0, stream, (const int*)order, n, ne, bin_ent);
  hipLaunchKernelGGL(k_set_nent, dim3(1), dim3(64), 0, stream, tc, (long long)ne);
  n_ent = ne;
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

// Keep/append edit of every cell list for the pass's removed (rm) and
// appended (app_ref, app_cell; tc->k_app of them) entries; nothing when the
// pass is inactive.
int Topo::edit_bins() {
  const long long* gate = &tc->active;
  PGDG_TRY(bucket_stable(app_cell, len_dev(&tc->k_app), nc, coff, cursor, order, part32, stream, gate));
  hipLaunchKernelGGL(k_edit_count, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)bin_off,
                     (const int*)bin_ent, nc, (const uint8_t*)rm, cnt, (const int*)coff, gate);
  PGDG_TRY(scan_excl<int>(cnt, off_tmp, len_host(nc, gate), &tc->nent, true, part32, stream));
  hipLaunchKernelGGL(k_edit_fill, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)bin_off,
                     (const int*)bin_ent, nc, (const uint8_t*)rm, (const int*)off_tmp, ent_tmp,
                     (const int*)coff, (const int*)order, (const int*)app_ref, gate, (long long)(cap + nwall), tc);
  hipLaunchKernelGGL(k_copy_len, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)off_tmp, bin_off,
                     len_host(nc + 1, gate));
  hipLaunchKernelGGL(k_copy_len, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)ent_tmp, bin_ent,
                     len_dev(&tc->nent, gate));
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

// Status-3 lists of the flagged (need_) particles of a split or merge pass,
// allocated in the pool on the device; only when the pass is active.
template <int D>
int Topo::build_lists(const Geo& g, const double* xq, const double* hq, const double* xall,
                      const WallDev& W, int status) {
  const long long* gate = &tc->active;
  const Len N = len_dev(&tc->n_total, gate);
  hipLaunchKernelGGL((k_lists<D, false>), dim3(kLoopBlocks), dim3(128), 0, stream, g, xq, cap, hq, xall, W, N,
                     (const uint8_t*)need_, status, (const int*)bin_off, (const int*)bin_ent, lc32,
                     (const int64_t*)loff, pool, (const long long*)nullptr);
  hipLaunchKernelGGL(k_widen, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)lc32, N, lscan);
  PGDG_TRY(scan_excl<long long>(lscan, lscan, N, &tc->lists_total, false, part64, stream));
  hipLaunchKernelGGL(k_list_alloc, dim3(1), dim3(64), 0, stream, tc, 1);
  hipLaunchKernelGGL(k_list_offsets, dim3(kLoopBlocks), dim3(256), 0, stream, N, (const uint8_t*)need_,
                     (const long long*)lscan, (const TopoCounters*)tc, loff);
  hipLaunchKernelGGL((k_lists<D, true>), dim3(kLoopBlocks), dim3(128), 0, stream, g, xq, cap, hq, xall, W, N,
                     (const uint8_t*)need_, status, (const int*)bin_off, (const int*)bin_ent, lc32,
                     (const int64_t*)loff, pool, (const long long*)&tc->list_ok);
  hipLaunchKernelGGL(k_list_commit, dim3(kLoopBlocks), dim3(256), 0, stream, N, (const uint8_t*)need_,
                     (const int*)lc32, (const TopoCounters*)tc, lcnt);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int Topo::mark_in(const Geo& g, const PtlDev& P, const long long* gate) {
  if (g.sem != PGDG_SEM_S3) return PGDG_OK;
  hipLaunchKernelGGL(k_mark_in<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, P, len_dev(&tc->n_total, gate),
                     (const uint8_t*)dead);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int Topo::split(const Geo& g, ParticleArrays& PA, const WallDev& W) {
  PtlDev P = PA.view();
  const bool bounds = g.ppc_max > 0;
  const int nch = D == 3 ? 8 : 4;
  const Len N = len_dev(&tc->n_total);
  if (bounds)
    hipLaunchKernelGGL(k_fluid_count, dim3(g1(nc)), dim3(256), 0, stream, (const int*)bin_off,
                       (const int*)bin_ent, nc, fcnt);
  hipLaunchKernelGGL(k_pass_zero, dim3(1), dim3(64), 0, stream, tc);
  hipLaunchKernelGGL(k_split_classify<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, P, tc, type, is_a, is_b,
                     is_c, lc32, bounds ? 1 : 0, src, reject);
  if (bounds) {
    // admissions against ppc_max, per cell in particle order
    PGDG_TRY(scan_excl<int>(lc32, cbase, N, &tc->n_contrib, false, part32, stream));
    hipLaunchKernelGGL(k_split_contrib<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, P,
                       (const TopoCounters*)tc, (const uint8_t*)type, (const int*)cbase, (long long)ccap, ckey,
                       ccand, ccost);
    hipLaunchKernelGGL(k_contrib_check, dim3(1), dim3(64), 0, stream, tc, (long long)ccap);
    PGDG_TRY(bucket_stable(ckey, len_dev(&tc->n_contrib), nc, coff, cursor, order, part32, stream,
                           &tc->n_contrib));
    hipLaunchKernelGGL(k_bound_walk, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)coff, (const int*)order,
                       (const int*)ccand, (const int*)ccost, (const int*)fcnt, nc, g.ppc_max, 1, reject,
                       (const long long*)&tc->n_contrib);
    hipLaunchKernelGGL(k_split_final, dim3(kLoopBlocks), dim3(256), 0, stream, tc, type,
                       (const uint8_t*)reject, is_a, is_b, is_c, src);
  }
  // split and inlet bases (the outlet total is all the pass needs)
  PGDG_TRY(scan_excl<int>(is_a, is_a, len_dev(&tc->n_total, &tc->tot_split), nullptr, false, part32, stream));
  PGDG_TRY(scan_excl<int>(is_b, is_b, len_dev(&tc->n_total, &tc->tot_inlet), nullptr, false, part32, stream));
  hipLaunchKernelGGL(k_split_begin, dim3(1), dim3(64), 0, stream, tc, (long long)cap, nch, nc);
  const long long* gate = &tc->active;
  PGDG_TRY(fill_u8(win, len_dev(&tc->win_len), 0, stream));
  PGDG_TRY(fill_u8(rm, len_dev(&tc->n_total, gate), 0, stream));
  PGDG_TRY(fill_u8(aff, len_dev(&tc->n_total, gate), 0, stream));
  // old lists of split parents and outlet particles -> affected
  hipLaunchKernelGGL(k_scatter_lists, dim3(kLoopBlocks), dim3(256), 0, stream, len_dev(&tc->n_total, gate),
                     (const uint8_t*)src, (const int*)lcnt, (const int64_t*)loff, (const int*)pool, aff);
  if (g.ord == 2)
    hipLaunchKernelGGL((k_split_generate<D, 2>), dim3(kLoopBlocks), dim3(256), 0, stream, g, P, tc, (const uint8_t*)type, (const int*)is_a, (const int*)is_b, rm,
                       dead, aff, src, app_ref, app_cell, lcnt);
  else
    hipLaunchKernelGGL((k_split_generate<D, 1>), dim3(kLoopBlocks), dim3(256), 0, stream, g, P, tc, (const uint8_t*)type, (const int*)is_a, (const int*)is_b, rm,
                       dead, aff, src, app_ref, app_cell, lcnt);
  // centre cells of the windows are the cells of every appended entry
  hipLaunchKernelGGL(k_window<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, (const uint32_t*)app_cell,
                     len_dev(&tc->k_app, gate), win);
  hipLaunchKernelGGL(k_mark_refs, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)app_ref,
                     len_dev(&tc->k_app, gate), aff);
  PGDG_TRY(edit_bins());
  hipLaunchKernelGGL(k_split_end, dim3(1), dim3(64), 0, stream, tc, nch);
  hipLaunchKernelGGL(k_affected<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, P, len_dev(&tc->n_total, gate),
                     (const uint8_t*)win, (const uint8_t*)aff, (const uint8_t*)dead, need_);
  PGDG_TRY(build_lists<D>(g, P.x, nullptr, P.x, W, 3));
  PGDG_HIP_TRY(hipGetLastError());
  return mark_in<D>(g, P);
}

template <int D>
int Topo::merge(const Geo& g, ParticleArrays& PA, const WallDev& W) {
  PtlDev P = PA.view();
  const bool bounds = g.ppc_min > 0;
  const Len N = len_dev(&tc->n_total);
  if (bounds)
    hipLaunchKernelGGL(k_fluid_count, dim3(g1(nc)), dim3(256), 0, stream, (const int*)bin_off,
                       (const int*)bin_ent, nc, fcnt);
  hipLaunchKernelGGL(k_pass_zero, dim3(1), dim3(64), 0, stream, tc);
  hipLaunchKernelGGL(k_merge_candidate<D>, dim3(kLoopBlocks), dim3(128), 0, stream, g, P,
                     (const TopoCounters*)tc, (const int*)lcnt, (const int64_t*)loff, (const int*)pool, cand,
                     (const int*)fcnt, (const uint8_t*)dead, rm, aff, src, reject);
  // survivor indicator in is_b, its exclusive scan in is_a
  hipLaunchKernelGGL(k_merge_mark<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, P, tc, (const int*)cand, is_b,
                     bounds ? 1 : 0, lc32);
  if (bounds) {
    // admissions against ppc_min, per cell in survivor order
    PGDG_TRY(scan_excl<int>(lc32, cbase, N, &tc->n_contrib, false, part32, stream));
    hipLaunchKernelGGL(k_merge_contrib<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, P,
                       (const TopoCounters*)tc, (const int*)cand, (const int*)is_b, (const int*)cbase,
                       (long long)ccap, ckey, ccand, ccost);
    hipLaunchKernelGGL(k_contrib_check, dim3(1), dim3(64), 0, stream, tc, (long long)ccap);
    PGDG_TRY(bucket_stable(ckey, len_dev(&tc->n_contrib), nc, coff, cursor, order, part32, stream,
                           &tc->n_contrib));
    hipLaunchKernelGGL(k_bound_walk, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)coff, (const int*)order,
                       (const int*)ccand, (const int*)ccost, (const int*)fcnt, nc, g.ppc_min, 0, reject,
                       (const long long*)&tc->n_contrib);
    hipLaunchKernelGGL(k_merge_final, dim3(kLoopBlocks), dim3(256), 0, stream, tc, (const uint8_t*)reject, is_b);
  }
  PGDG_TRY(scan_excl<int>(is_b, is_a, len_dev(&tc->n_total, &tc->n_pairs), nullptr, false, part32, stream));
  hipLaunchKernelGGL(k_merge_begin, dim3(1), dim3(64), 0, stream, tc, nc);
  const long long* gate = &tc->active;
  PGDG_TRY(fill_u8(win, len_dev(&tc->win_len), 0, stream));
  PGDG_TRY(fill_u8(rm, len_dev(&tc->n_total, gate), 0, stream));
  PGDG_TRY(fill_u8(aff, len_dev(&tc->n_total, gate), 0, stream));
  PGDG_TRY(fill_u8(src, len_dev(&tc->n_total, gate), 0, stream));
  hipLaunchKernelGGL(k_merge_apply<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, P, (const TopoCounters*)tc,
                     (const int*)cand, (const int*)is_b, (const int*)is_a, rm, dead, src, app_ref, app_cell);
  // old lists of survivors and partners, survivors themselves, windows
  hipLaunchKernelGGL(k_scatter_lists, dim3(kLoopBlocks), dim3(256), 0, stream, len_dev(&tc->n_total, gate),
                     (const uint8_t*)src, (const int*)lcnt, (const int64_t*)loff, (const int*)pool, aff);
  hipLaunchKernelGGL(k_window<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, (const uint32_t*)app_cell,
                     len_dev(&tc->k_app, gate), win);
  hipLaunchKernelGGL(k_mark_refs, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)app_ref,
                     len_dev(&tc->k_app, gate), aff);
  PGDG_TRY(edit_bins());
  hipLaunchKernelGGL(k_merge_end, dim3(1), dim3(64), 0, stream, tc);
  hipLaunchKernelGGL(k_affected<D>, dim3(kLoopBlocks), dim3(256), 0, stream, g, P, len_dev(&tc->n_total, gate),
                     (const uint8_t*)win, (const uint8_t*)aff, (const uint8_t*)dead, need_);
  PGDG_TRY(build_lists<D>(g, P.x, nullptr, P.x, W, 3));
  PGDG_HIP_TRY(hipGetLastError());
  // positions and volumes change only for merged particles: an inactive
  // pass leaves mark_in as the last pass computed it
  return mark_in<D>(g, P, gate);
}

int Topo::finish(ParticleArrays& PA, int64_t* n_out) {
  PGDG_HIP_TRY(hipMemcpyAsync(&hc, tc, sizeof(TopoCounters), hipMemcpyDeviceToHost, stream));
  PGDG_HIP_TRY(hipStreamSynchronize(stream));
  if (hc.err != 0) {
    const int e = (int)hc.err;
    // the state of the failed passes is not usable; clear the flag
    TopoCounters z = hc;
    z.err = 0;
    PGDG_HIP_TRY(hipMemcpyAsync(tc, &z, sizeof(TopoCounters), hipMemcpyHostToDevice, stream));
    PGDG_HIP_TRY(hipStreamSynchronize(stream));
    return e;
  }
  n_ent = hc.nent;
  pool_used = hc.pool_used;
  const int64_t n = hc.n_total;
  if (hc.n_dead > 0) {
    hipLaunchKernelGGL(k_keep, dim3(g1(n)), dim3(256), 0, stream, (const uint8_t*)dead, n, is_c);
    PGDG_TRY(scan_excl<int>(is_c, is_c, len_host(n), nullptr, false, part32, stream));
    hipLaunchKernelGGL(k_newidx, dim3(g1(n)), dim3(256), 0, stream, (const uint8_t*)dead, (const int*)is_c, n,
                       newidx);
    PGDG_TRY(PA.gather(newidx, n, stream));
    hipLaunchKernelGGL(k_gather<int>, dim3(g1(n)), dim3(256), 0, stream, (const int*)lcnt, lcnt2, n, cap, 1,
                       (const int*)newidx);
    hipLaunchKernelGGL(k_gather<int64_t>, dim3(g1(n)), dim3(256), 0, stream, (const int64_t*)loff, loff2, n,
                       cap, 1, (const int*)newidx);
    int* t = lcnt;
    lcnt = lcnt2;
    lcnt2 = t;
    int64_t* u = loff;
    loff = loff2;
    loff2 = u;
    hipLaunchKernelGGL(k_remap_refs, dim3(g1(n_ent)), dim3(256), 0, stream, bin_ent, n_ent, (const int*)newidx);
    if (pool_used > 0)
      hipLaunchKernelGGL(k_remap_refs, dim3(g1(pool_used)), dim3(256), 0, stream, pool, pool_used,
                         (const int*)newidx);
    PGDG_TRY(fill_u8(dead, len_host(n), 0, stream));
  }
  const int64_t live = n - hc.n_dead;
  hipLaunchKernelGGL(k_compacted, dim3(1), dim3(64), 0, stream, tc, (long long)live);
  PGDG_HIP_TRY(hipGetLastError());
  *n_out = live;
  return PGDG_OK;
}

template <int D>
int Topo::end_of_step_lists(const Geo& g, const PtlDev& P, const double* xb, const double* hb,
                            const WallDev& W, int64_t n) {
  if (n == 0) return PGDG_OK;
  // particles not rebuilt keep no list
  PGDG_TRY(fill_i32(lcnt, len_host(n), 0, stream));
  hipLaunchKernelGGL(k_flag_lists<D>, dim3(g1(n)), dim3(256), 0, stream, g, P, n, need_);
  const Len N = len_host(n);
  hipLaunchKernelGGL((k_lists<D, false>), dim3(kLoopBlocks), dim3(128), 0, stream, g, xb, cap, hb, xb, W, N,
                     (const uint8_t*)need_, 2, (const int*)bin_off, (const int*)bin_ent, lc32,
                     (const int64_t*)loff, pool, (const long long*)nullptr);
  hipLaunchKernelGGL(k_widen, dim3(kLoopBlocks), dim3(256), 0, stream, (const int*)lc32, N, lscan);
  PGDG_TRY(scan_excl<long long>(lscan, lscan, N, &tc->lists_total, false, part64, stream));
  long long total = 0;
  PGDG_HIP_TRY(hipMemcpyAsync(&total, &tc->lists_total, sizeof(long long), hipMemcpyDeviceToHost, stream));
  PGDG_HIP_TRY(hipStreamSynchronize(stream));
  if (total > pool_cap) {  // grow the pool (no graph holds it)
    const int64_t ncap = total + total / 4 + 1024;
    int* np_ = nullptr;
    PGDG_HIP_TRY(hipMalloc((void**)&np_, sizeof(int) * ncap));
    (void)hipFree(pool);
    pool = np_;
    bytes_total += sizeof(int) * (size_t)(ncap - pool_cap);
    pool_cap = ncap;
  }
  PGDG_TRY(set_pool_used(0));
  hipLaunchKernelGGL(k_tc_lists_full, dim3(1), dim3(64), 0, stream, tc);
  hipLaunchKernelGGL(k_list_offsets, dim3(kLoopBlocks), dim3(256), 0, stream, N, (const uint8_t*)need_,
                     (const long long*)lscan, (const TopoCounters*)tc, loff);
  hipLaunchKernelGGL((k_lists<D, true>), dim3(kLoopBlocks), dim3(128), 0, stream, g, xb, cap, hb, xb, W, N,
                     (const uint8_t*)need_, 2, (const int*)bin_off, (const int*)bin_ent, lc32,
                     (const int64_t*)loff, pool, (const long long*)&tc->list_ok);
  hipLaunchKernelGGL(k_list_commit, dim3(kLoopBlocks), dim3(256), 0, stream, N, (const uint8_t*)need_,
                     (const int*)lc32, (const TopoCounters*)tc, lcnt);
  pool_used = total;
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

#define PGDG_INST(D)                                                                              \
  template int Topo::build_bins<D>(const Geo&, const PtlDev&, const WallDev&, int64_t);           \
  template int Topo::split<D>(const Geo&, ParticleArrays&, const WallDev&);                       \
  template int Topo::merge<D>(const Geo&, ParticleArrays&, const WallDev&);                       \
  template int Topo::end_of_step_lists<D>(const Geo&, const PtlDev&, const double*, const double*, \
                                          const WallDev&, int64_t);                               \
  template int Topo::mark_in<D>(const Geo&, const PtlDev&, const long long*);
PGDG_INST(2)
PGDG_INST(3)
#undef PGDG_INST

}  // namespace pgdg
