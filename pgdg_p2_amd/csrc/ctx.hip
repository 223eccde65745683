// ctx.hip - the step context (device state of one case on one GPU) and the
// C ABI of the time step.
//
// pgdg_step follows Solver::MPM_Solver's loop body (S3:solver.cpp:33-51):
// Set_timestep, the output-time truncation of dt, then
// MPM::MPM_run_simulation (S3:mpm_ini.cpp:682-766):
//   Used_state_update (first step), binning, Split_particle_fast,
//   Merge_particle_fast x4, cell membership, and two RK stages of
//   [neighbour lists, MLS_P2G (stage 0), Calculate_limiter,
//    Grid_acceleration, Update_grid, Update_particle].
// All work is enqueued on the context's stream; dt and the loop scalars stay
// on the device.  Binning, split and merge keep their counts on the device
// too (Topo); the host reads them back once per step, after the last merge
// pass, when the removed particles are compacted, and once more at the end
// of the step to size the neighbour-list pool.
#include <algorithm>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <mutex>
#include <new>
#include <vector>

#include "dev_common.hpp"
#include "hip_util.hpp"
#include "step_kernels.hpp"
#include "step_topo.hpp"

namespace pgdg {

// counts read by the stage kernels (see step_particles.hip)
__global__ void k_set_counts(DevScalars* sc, int64_t np, int64_t nent) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sc->np_dev = np;
    sc->nent_dev = nent;
  }
}

__global__ void k_used_clear(int32_t* used, int nc) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nc) used[c] = 0;
}

// Used_state_update (S3:mpm.cpp:184-263): cells within +-5 of a cell whose
// phase 0 or 1 is present.
__global__ void k_used_mark(Geo g, const int32_t* ph_state, int32_t* used, int D) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  if (!(ph_state[c] == 1 || ph_state[g.nc + c] == 1)) return;
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  const int rk = D == 3 ? 5 : 0;
  for (int t = max(i - 5, 0); t <= min(i + 5, g.N[0] - 1); t++)
    for (int s = max(j - 5, 0); s <= min(j + 5, g.N[1] - 1); s++)
      for (int q = max(k - rk, 0); q <= min(k + rk, g.N[2] - 1); q++) used[(t * g.N[1] + s) * g.N[2] + q] = 1;
}

// Stage timer: when enabled, every stage is bracketed by two HIP events on
// the context stream; pgdg_profile_read sums the elapsed times per slot.
enum {
  PS_TIMESTEP = 0, PS_USED, PS_BIN, PS_SPLIT, PS_MERGE, PS_MEMBERSHIP, PS_STAGE_SAVE, PS_P2G,
  PS_LIMITER, PS_RHS, PS_UPDATE_GRID, PS_G2P, PS_PST, PS_LISTS, PS_PST_PREP, PS_COMPACT, PS_NSLOT
};

struct Profiler {
  bool on = false;
  struct Rec {
    int slot;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  int begin(int slot, hipStream_t st, Rec& r) {
    r.slot = slot;
    PGDG_HIP_TRY(hipEventCreate(&r.a));
    PGDG_HIP_TRY(hipEventCreate(&r.b));
    PGDG_HIP_TRY(hipEventRecord(r.a, st));
    return PGDG_OK;
  }
  int end(Rec& r, hipStream_t st) {
    PGDG_HIP_TRY(hipEventRecord(r.b, st));
    recs.push_back(r);
    return PGDG_OK;
  }
  int read(double* ms, int64_t* cnt, int nslot) {
    for (int k = 0; k < nslot; k++) {
      ms[k] = 0;
      cnt[k] = 0;
    }
    for (auto& r : recs) {
      PGDG_HIP_TRY(hipEventSynchronize(r.b));
      float t = 0;
      PGDG_HIP_TRY(hipEventElapsedTime(&t, r.a, r.b));
      if (r.slot < nslot) {
        ms[r.slot] += t;
        cnt[r.slot] += 1;
      }
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    recs.clear();
    return PGDG_OK;
  }
};

struct Ctx {
  pgdg_case cs;
  Profiler prof;
  Geo g;
  int D, U, M, F, nc;
  hipStream_t st = nullptr;
  bool own_stream = false;
  ParticleArrays PA;
  int64_t np = 0;
  WallDev W{};
  double* w_slab = nullptr;
  CellDev C{};
  double* c_slab = nullptr;
  int32_t* ci_slab = nullptr;
  int32_t* used = nullptr;
  StageBufs S{};
  double* s_slab = nullptr;
  BinData B{};
  double* b_slab = nullptr;
  P2gBuf PB{};
  MixedBins MB{};
  void* mb_slab = nullptr;
  void* ff_slab = nullptr;
  void* mg_slab = nullptr;
  size_t mg_bytes = 0;
  // mixed form: how the step's interface points solve MGFM (in the face
  // kernel, or as a batch from the reference's or the network's p*)
  IfaceBatch ib;
  int32_t* bi_slab = nullptr;
  Topo T;
  DevScalars* sc = nullptr;
  int64_t iterate_num = 0;
  int64_t n_split = 0, n_inlet = 0, n_outlet = 0, n_merge = 0;
  bool loaded = false;
  size_t bytes = 0;
  // RK-stage graphs: one per stage index and per pair of buffers the stage
  // reads through pointers that alternate (particle slab, bin entries)
  struct StageGraph {
    int md;
    const void *k1, *k2;
    hipGraphExec_t exec;
  };
  static constexpr int kMaxGraphs = 8;
  StageGraph graphs[kMaxGraphs];
  int ngraphs = 0;
  bool use_graphs = true;
  bool capturing = false;
  // autotuner: workgroup sizes of the particle-shifting, RHS and limiter
  // kernels, chosen by timing the candidates in place during the first step
  Tuning tune;
  bool autotune = true, tuned = false, tuning_now = false;
  bool rhs_fused = false;  // the last RHS launch applied the coefficient update
  // mixed form: candidate entries of the particle-shifting sum staged in LDS
  // (3-D); off = every lane walks its candidates in global memory
  int pst_staging = 1;
  // a status-2 list refresh (the stage-start save) has run since the last
  // upload: before it every particle's list is empty, as after the
  // reference's set-up
  bool lists_built = false;
  // the cell bins have been built since the last upload (the stages that
  // read them refuse to run before)
  bool bins_valid = false;
  template <typename F>
  int pick_block(F launch, int* best) {
    static const int cand[3] = {64, 128, 256};
    hipEvent_t a = nullptr, b = nullptr;
    PGDG_HIP_TRY(hipEventCreate(&a));
    PGDG_HIP_TRY(hipEventCreate(&b));
    float best_ms = 1e30f;
    int rc = PGDG_OK;
    for (int q = 0; q < 3 && rc == PGDG_OK; q++) {
      rc = launch(cand[q]);  // warm
      if (rc != PGDG_OK) break;
      (void)hipEventRecord(a, st);
      rc = launch(cand[q]);
      (void)hipEventRecord(b, st);
      if (rc != PGDG_OK || hipEventSynchronize(b) != hipSuccess) {
        rc = rc != PGDG_OK ? rc : PGDG_ERR_HIP;
        break;
      }
      float ms = 0;
      (void)hipEventElapsedTime(&ms, a, b);
      if (ms < best_ms) {
        best_ms = ms;
        *best = cand[q];
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
  }
  int sync_counts() {
    hipLaunchKernelGGL(k_set_counts, dim3(1), dim3(64), 0, st, sc, np, T.n_ent);
    PGDG_HIP_TRY(hipGetLastError());
    return PGDG_OK;
  }
  void drop_graphs() {
    for (int q = 0; q < ngraphs; q++) (void)hipGraphExecDestroy(graphs[q].exec);
    ngraphs = 0;
  }
  template <int D_>
  int rk_stage(int md);

  int init(const pgdg_case* c, int64_t cap, int64_t nwall);
  void release();
  int upload(const pgdg_particles* p, const pgdg_walls* w, const pgdg_cells* c,
             const pgdg_scalars* s);
  int download(pgdg_particles* p, pgdg_cells* c, pgdg_scalars* s);
  template <int D_>
  int set_timestep();
  template <int D_>
  int run_simulation(int64_t n);
  int fit_iface_records();
  int alloc_iface_records(int64_t cap_rec);
  template <int D_>
  int stage(int fn, int arg);
  template <int D_>
  int stage_impl(int fn, int arg);
  int step(int nsteps);
};

static void fill_geo(const pgdg_case& cs, Geo& g) {
  memset(&g, 0, sizeof(g));
  const int D = cs.dim;
  for (int a = 0; a < 3; a++) {
    g.N[a] = a < D ? cs.n[a] : 1;
    g.ps[a] = a < D ? (cs.box_r[a] - cs.box_l[a]) / double(g.N[a]) : 1.0;
    g.box_l[a] = cs.box_l[a];
    g.box_r[a] = cs.box_r[a];
  }
  g.nc = g.N[0] * g.N[1] * g.N[2];
  g.F = cs.nphase;
  g.sem = cs.semantics;
  for (int f = 0; f < 6; f++) g.bc[f] = cs.bc[f];
  g.p_sz = cs.p_sz;
  g.p_g = cs.p_g;
  g.p2g_cut = cs.p2g_cut;
  for (int s = 0; s < PGDG_MAX_PHASE; s++) {
    g.gamma[s] = cs.gamma[s];
    g.pinf[s] = cs.pinf[s];
    g.ini_rho[s] = cs.ini_rho[s];
  }
  g.inlet_rho = cs.inlet_rho;
  for (int a = 0; a < 3; a++) g.inlet_u[a] = cs.inlet_u[a];
  g.inlet_p = cs.inlet_p;
  g.radius = (int)std::ceil(kCellRatio / (cs.p_g / cs.p_sz));  // S3:level_infor.cpp:155
  g.ord = cs.order;
  g.prec = cs.precision;
  g.mfma_rhs = 0;
  if (cs.order == 1) {
    g.Gp[0] = -1. / std::sqrt(3.);  // S3:mpm_ini.cpp:37-41
    g.Gp[1] = 1. / std::sqrt(3.);
    g.Gw[0] = g.Gw[1] = 0.5;
  } else {  // 3-point Gauss-Legendre, weights normalised to sum 1
    g.Gp[0] = -std::sqrt(3. / 5.);
    g.Gp[1] = 0.;
    g.Gp[2] = std::sqrt(3. / 5.);
    g.Gw[0] = g.Gw[2] = 5. / 18.;
    g.Gw[1] = 8. / 18.;
  }
  for (int a = 0; a < 3; a++) g.q12[a] = g.ps[a] * g.ps[a] / 12.;
  g.cfl = cs.cfl > 0 ? cs.cfl : 0.3;  // S3:particle.cpp:558
  g.ppc_min = cs.ppc_min;
  g.ppc_max = cs.ppc_max;
  for (int jj = 0; jj < 4; jj++) {  // limiter probe angles, S3:mpm.cpp:697-698
    double theta = (jj + 0.5) * kPiRef * 0.5;
    g.probe_c[jj] = std::cos(theta);
    g.probe_s[jj] = std::sin(theta);
  }
  g.sqrt2 = std::sqrt(2.);
  double pd = D == 3 ? (cs.p_sz * cs.p_sz) * cs.p_sz : cs.p_sz * cs.p_sz;
  g.vol_max = pd * 1.6;          // S3:mpm.cpp:1068
  g.vol_min = pd * 2. / 3.;      // S3:mpm.cpp:1540
  g.dss0 = D == 3 ? root_d<3>(g.vol_min) : root_d<2>(g.vol_min);
}

int Ctx::init(const pgdg_case* c, int64_t cap, int64_t nwall) {
  cs = *c;
  D = cs.dim;
  U = D + 2;
  M = cs.order == 1 ? D + 1 : (D == 2 ? 6 : 10);
  F = cs.nphase;
  fill_geo(cs, g);
  nc = g.nc;
  PGDG_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  own_stream = true;
  PGDG_TRY(PA.init(U, M, cap));
  // walls
  W.n = nwall;
  PGDG_HIP_TRY(hipMalloc((void**)&w_slab, sizeof(double) * 7 * (nwall > 0 ? nwall : 1)));
  W.x = w_slab;
  W.norm = w_slab + 3 * nwall;
  W.vol = w_slab + 6 * nwall;
  // cells
  size_t fum = (size_t)F * U * M * nc, fu = (size_t)F * U * nc, fn = (size_t)F * nc;
  size_t nd = 3 * fum + fu + 5 * fn + 3 * fn + 3 * nc + 3 * nc + nc + 9 * (size_t)nc;
  PGDG_HIP_TRY(hipMalloc((void**)&c_slab, sizeof(double) * nd));
  PGDG_HIP_TRY(hipMemset(c_slab, 0, sizeof(double) * nd));
  double* b = c_slab;
  C.coef = b; b += fum;
  C.coef0 = b; b += fum;
  C.dcoef = b; b += fum;
  C.lim = b; b += fu;
  C.alph = b; b += fn;
  C.rho = b; b += fn;
  C.E = b; b += fn;
  C.c = b; b += fn;
  C.ac = b; b += fn;
  C.rhov = b; b += 3 * fn;
  C.v = b; b += 3 * (size_t)nc;
  C.a = b; b += 3 * (size_t)nc;
  C.P = b; b += nc;
  C.G = b; b += 9 * (size_t)nc;
  PGDG_HIP_TRY(hipMalloc((void**)&ci_slab, sizeof(int32_t) * 2 * fn));
  C.ph_state = ci_slab;
  C.ph_state_b = ci_slab + fn;
  PGDG_HIP_TRY(hipMalloc((void**)&used, sizeof(int32_t) * nc));
  size_t ff_bytes = 0;
  if (cs.precision == PGDG_PREC_MIXED) {
    // interior-face point fluxes [2][F][NQF][U][D][nc] (FP32), the general
    // face list [D][nc] and its count
    const int nq = cs.order + 1, nqf = D == 3 ? nq * nq : nq;
    const size_t nflux = (size_t)2 * F * nqf * U * D * nc;
    // 3-D: boundary-face point fluxes [F][6][NQF][U][plane], plane the
    // largest of the three face planes of the grid
    const size_t plane = D == 3 ? std::max((size_t)cs.n[1] * cs.n[2],
                                           std::max((size_t)cs.n[0] * cs.n[2], (size_t)cs.n[0] * cs.n[1]))
                                : 0;
    const size_t nbflux = (size_t)F * 6 * nqf * U * plane;
    ff_bytes = sizeof(float) * (nflux + nbflux) + sizeof(int32_t) * ((size_t)D * nc + 64) + (size_t)D * nc;
    PGDG_HIP_TRY(hipMalloc(&ff_slab, ff_bytes));
    C.fflux = (float*)ff_slab;
    C.bflux = D == 3 ? C.fflux + nflux : nullptr;
    C.flist = (int32_t*)(C.fflux + nflux + nbflux);
    C.fcount = C.flist + (size_t)D * nc;
    C.fkind = (int8_t*)(C.fcount + 64);
    // interface-point MGFM records for the batched solves: inputs [6] and
    // solutions [3] (FP64), pair code and iterations per record, the first
    // record of each listed face [D][nc], the record count and its peak
    PGDG_TRY(alloc_iface_records(std::max<int64_t>(65536, nc)));
  }
  // stage buffers: xb[3], hb, x0[3], vol0
  PGDG_HIP_TRY(hipMalloc((void**)&s_slab, sizeof(double) * 8 * cap));
  PGDG_HIP_TRY(hipMemset(s_slab, 0, sizeof(double) * 8 * cap));
  S.xb = s_slab;
  S.hb = s_slab + 3 * cap;
  S.x0 = s_slab + 4 * cap;
  S.vol0 = s_slab + 7 * cap;
  PGDG_HIP_TRY(hipMalloc((void**)&S.in_ele, sizeof(int32_t) * cap));
  // bin-ordered neighbour data: x[3], xb[3], vol, rho, c per entry; boxes
  const int64_t ne = cap + nwall;
  B.ne_cap = ne;
  PGDG_HIP_TRY(hipMalloc((void**)&b_slab, sizeof(double) * (9 * ne + 6 * (size_t)nc)));
  PGDG_HIP_TRY(hipMalloc((void**)&bi_slab, sizeof(int32_t) * 2 * ne));
  B.x = b_slab;
  B.xb = b_slab + 3 * ne;
  B.vol = b_slab + 6 * ne;
  B.rho = b_slab + 7 * ne;
  B.c = b_slab + 8 * ne;
  B.bmin = b_slab + 9 * ne;
  B.bmax = B.bmin + 3 * (size_t)nc;
  B.ph = bi_slab;
  B.ref = bi_slab + ne;
  const int p2g_rows_n = kP2gLight;
  PB.ne = ne;
  // light records, then the appearing cells' masses and flags
  const size_t p2g_bytes = sizeof(double) * (p2g_rows_n * (size_t)ne + PGDG_MAX_PHASE * (size_t)nc) +
                           sizeof(int32_t) * ((size_t)nc + 64);
  PGDG_HIP_TRY(hipMalloc((void**)&PB.light, p2g_bytes));
  PB.mass = PB.light + p2g_rows_n * (size_t)ne;
  PB.flag = (int32_t*)(PB.mass + PGDG_MAX_PHASE * (size_t)nc);
  size_t mb_bytes = 0;
  if (cs.precision == PGDG_PREC_MIXED) {
    // boxes and their unions along the cell lines 2 x 32 B per cell; r1
    // (16 B, plus one spare record: the shifting loop reads one record
    // ahead), r2 (16 B), r3 (8 B), ph, ecell (4 B) per entry
    mb_bytes = 64 * (size_t)nc + (size_t)ne * (16 + 16 + 8 + 4 + 4) + 16 + 64;
    PGDG_HIP_TRY(hipMalloc(&mb_slab, mb_bytes));
    PGDG_HIP_TRY(hipMemset(mb_slab, 0, mb_bytes));
    char* q = (char*)mb_slab;
    MB.box = (float4*)q; q += 32 * (size_t)nc;
    MB.ubox = (float4*)q; q += 32 * (size_t)nc;
    MB.r1 = (float4*)q; q += 16 * ((size_t)ne + 1);
    MB.r2 = (float4*)q; q += 16 * (size_t)ne;
    MB.r3 = (float2*)q; q += 8 * (size_t)ne;
    MB.ph = (int32_t*)q; q += 4 * (size_t)ne;
    MB.ecell = (int32_t*)q;
    MB.ne_cap = ne;
  }
  PGDG_HIP_TRY(hipMalloc((void**)&sc, sizeof(DevScalars)));
  PGDG_HIP_TRY(hipMemset(sc, 0, sizeof(DevScalars)));
  PGDG_TRY(T.init(cap, nwall, nc, st));
  PGDG_HIP_TRY(hipStreamSynchronize(st));
  bytes = sizeof(double) * (2 * (size_t)PA.rows64 * cap + nd + 8 * cap + 7 * nwall +
                            9 * (size_t)ne + 6 * (size_t)nc) +
          p2g_bytes + T.bytes_total + mb_bytes + ff_bytes + mg_bytes;
  return PGDG_OK;
}

void Ctx::release() {
  if (st) (void)hipStreamSynchronize(st);
  drop_graphs();
  PA.release();
  T.release();
  void* ptrs[] = {w_slab, c_slab, ci_slab, used, s_slab, S.in_ele, sc, b_slab, bi_slab, PB.light, mb_slab,
                  ff_slab, mg_slab};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (own_stream && st) (void)hipStreamDestroy(st);
}

int Ctx::upload(const pgdg_particles* p, const pgdg_walls* w, const pgdg_cells* c,
                const pgdg_scalars* s) {
  if (!p || !w || !c || !s) return PGDG_ERR_ARG;
  if (p->n > PA.cap || p->n < 0) return PGDG_ERR_CAPACITY;
  if (w->n != W.n) return PGDG_ERR_ARG;
  np = p->n;
  PtlDev P = PA.view();
  const int64_t hc = p->cap;
  auto up = [&](double* dst, const double* src, int rows) -> int {
    for (int r = 0; r < rows; r++)
      PGDG_HIP_TRY(hipMemcpyAsync(dst + (size_t)r * PA.cap, src + (size_t)r * hc,
                                  sizeof(double) * np, hipMemcpyHostToDevice, st));
    return PGDG_OK;
  };
  if (np > 0) {
    PGDG_TRY(up(P.x, p->x, 3));
    PGDG_TRY(up(P.v, p->v, 3));
    PGDG_TRY(up(P.tv, p->tv, 3));
    PGDG_TRY(up(P.rhov, p->rhov, 3));
    PGDG_TRY(up(P.vol, p->vol, 1));
    PGDG_TRY(up(P.mass, p->mass, 1));
    PGDG_TRY(up(P.h, p->h, 1));
    PGDG_TRY(up(P.P, p->P, 1));
    PGDG_TRY(up(P.rho, p->rho, 1));
    PGDG_TRY(up(P.E, p->E, 1));
    PGDG_TRY(up(P.c, p->c, 1));
    PGDG_TRY(up(P.coef, p->coef, U * M));
    PGDG_HIP_TRY(hipMemcpyAsync(P.phase, p->phase, sizeof(int32_t) * np, hipMemcpyHostToDevice, st));
    PGDG_HIP_TRY(hipMemcpyAsync(P.mark_in, p->mark_in, sizeof(int32_t) * np, hipMemcpyHostToDevice, st));
  }
  if (W.n > 0) {
    PGDG_HIP_TRY(hipMemcpyAsync(W.x, w->x, sizeof(double) * 3 * W.n, hipMemcpyHostToDevice, st));
    PGDG_HIP_TRY(hipMemcpyAsync(W.norm, w->norm, sizeof(double) * 3 * W.n, hipMemcpyHostToDevice, st));
    PGDG_HIP_TRY(hipMemcpyAsync(W.vol, w->vol, sizeof(double) * W.n, hipMemcpyHostToDevice, st));
  }
  size_t fum = (size_t)F * U * M * nc, fu = (size_t)F * U * nc, fn = (size_t)F * nc;
  auto upc = [&](double* dst, const double* src, size_t cnt) -> int {
    PGDG_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyHostToDevice, st));
    return PGDG_OK;
  };
  PGDG_TRY(upc(C.coef, c->coef, fum));
  PGDG_TRY(upc(C.coef0, c->coef0, fum));
  PGDG_TRY(upc(C.dcoef, c->dcoef, fum));
  PGDG_TRY(upc(C.lim, c->lim, fu));
  PGDG_TRY(upc(C.alph, c->alph, fn));
  PGDG_TRY(upc(C.rho, c->rho, fn));
  PGDG_TRY(upc(C.E, c->E, fn));
  PGDG_TRY(upc(C.c, c->c, fn));
  PGDG_TRY(upc(C.ac, c->ac, fn));
  PGDG_TRY(upc(C.rhov, c->rhov, 3 * fn));
  PGDG_TRY(upc(C.v, c->v, 3 * (size_t)nc));
  PGDG_TRY(upc(C.a, c->a, 3 * (size_t)nc));
  PGDG_TRY(upc(C.P, c->P, nc));
  PGDG_TRY(upc(C.G, c->G_v, 9 * (size_t)nc));
  PGDG_HIP_TRY(hipMemcpyAsync(C.ph_state, c->ph_state, sizeof(int32_t) * fn, hipMemcpyHostToDevice, st));
  PGDG_HIP_TRY(hipMemcpyAsync(C.ph_state_b, c->ph_state_b, sizeof(int32_t) * fn, hipMemcpyHostToDevice, st));
  DevScalars h;
  memset(&h, 0, sizeof(h));
  h.dt = s->dt;
  h.max_v = s->max_v;
  h.integral_time = s->integral_time;
  h.run_time = s->run_time;
  h.iterate_num = s->iterate_num;
  PGDG_HIP_TRY(hipMemcpyAsync(sc, &h, sizeof(h), hipMemcpyHostToDevice, st));
  iterate_num = s->iterate_num;
  // no neighbour lists exist at upload time (reference: empty lists)
  hipLaunchKernelGGL(k_used_clear, dim3((nc + 255) / 256), dim3(256), 0, st, used, nc);
  PGDG_TRY(T.reset(np));
  PGDG_HIP_TRY(hipStreamSynchronize(st));
  loaded = true;
  lists_built = false;
  bins_valid = false;
  return PGDG_OK;
}

int Ctx::download(pgdg_particles* p, pgdg_cells* c, pgdg_scalars* s) {
  if (!loaded) return PGDG_ERR_STATE;
  PGDG_HIP_TRY(hipStreamSynchronize(st));
  if (p) {
    if (np > p->cap) return PGDG_ERR_CAPACITY;
    p->n = np;
    PtlDev P = PA.view();
    const int64_t hc = p->cap;
    auto dn = [&](double* dst, const double* src, int rows) -> int {
      for (int r = 0; r < rows; r++)
        PGDG_HIP_TRY(hipMemcpy(dst + (size_t)r * hc, src + (size_t)r * PA.cap, sizeof(double) * np,
                               hipMemcpyDeviceToHost));
      return PGDG_OK;
    };
    if (np > 0) {
      PGDG_TRY(dn(p->x, P.x, 3));
      PGDG_TRY(dn(p->v, P.v, 3));
      PGDG_TRY(dn(p->tv, P.tv, 3));
      PGDG_TRY(dn(p->rhov, P.rhov, 3));
      PGDG_TRY(dn(p->vol, P.vol, 1));
      PGDG_TRY(dn(p->mass, P.mass, 1));
      PGDG_TRY(dn(p->h, P.h, 1));
      PGDG_TRY(dn(p->P, P.P, 1));
      PGDG_TRY(dn(p->rho, P.rho, 1));
      PGDG_TRY(dn(p->E, P.E, 1));
      PGDG_TRY(dn(p->c, P.c, 1));
      PGDG_TRY(dn(p->coef, P.coef, U * M));
      PGDG_HIP_TRY(hipMemcpy(p->phase, P.phase, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
      PGDG_HIP_TRY(hipMemcpy(p->mark_in, P.mark_in, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
    }
  }
  if (c) {
    size_t fum = (size_t)F * U * M * nc, fu = (size_t)F * U * nc, fn = (size_t)F * nc;
    auto dc = [&](double* dst, const double* src, size_t cnt) -> int {
      PGDG_HIP_TRY(hipMemcpy(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToHost));
      return PGDG_OK;
    };
    PGDG_TRY(dc(c->coef, C.coef, fum));
    PGDG_TRY(dc(c->coef0, C.coef0, fum));
    PGDG_TRY(dc(c->dcoef, C.dcoef, fum));
    PGDG_TRY(dc(c->lim, C.lim, fu));
    PGDG_TRY(dc(c->alph, C.alph, fn));
    PGDG_TRY(dc(c->rho, C.rho, fn));
    PGDG_TRY(dc(c->E, C.E, fn));
    PGDG_TRY(dc(c->c, C.c, fn));
    PGDG_TRY(dc(c->ac, C.ac, fn));
    PGDG_TRY(dc(c->rhov, C.rhov, 3 * fn));
    PGDG_TRY(dc(c->v, C.v, 3 * (size_t)nc));
    PGDG_TRY(dc(c->a, C.a, 3 * (size_t)nc));
    PGDG_TRY(dc(c->P, C.P, nc));
    PGDG_TRY(dc(c->G_v, C.G, 9 * (size_t)nc));
    PGDG_HIP_TRY(hipMemcpy(c->ph_state, C.ph_state, sizeof(int32_t) * fn, hipMemcpyDeviceToHost));
    PGDG_HIP_TRY(hipMemcpy(c->ph_state_b, C.ph_state_b, sizeof(int32_t) * fn, hipMemcpyDeviceToHost));
  }
  if (s) {
    DevScalars h;
    PGDG_HIP_TRY(hipMemcpy(&h, sc, sizeof(h), hipMemcpyDeviceToHost));
    s->dt = h.dt;
    s->max_v = h.max_v;
    s->integral_time = h.integral_time;
    s->run_time = h.run_time;
    s->iterate_num = h.iterate_num;
    s->n_split = n_split;
    s->n_inlet = n_inlet;
    s->n_outlet = n_outlet;
    s->n_merge = n_merge;
  }
  return PGDG_OK;
}

template <int D_>
int Ctx::set_timestep() {
  return launch_timestep<D_>(g, PA.view(), C, used, np, iterate_num > 1, sc, cs.output_dt, st);
}

// Stage functions numbered as the oracle's (oracle/step_ref.cpp `call`).
enum {
  FN_SET_TIMESTEP = 0, FN_USED = 1, FN_BIN = 2, FN_SPLIT = 3, FN_MERGE = 4, FN_MEMBERSHIP = 5,
  FN_NEIGHBORS = 6, FN_P2G = 7, FN_LIMITER = 8, FN_RHS = 9, FN_UPDATE_GRID = 10,
  FN_UPDATE_PARTICLE = 11, FN_RUN = 12,
  FN_TOPO_FINISH = 13  // internal: compaction after the split/merge passes
};
static_assert(FN_SET_TIMESTEP == PGDG_FN_SET_TIMESTEP && FN_USED == PGDG_FN_USED &&
                  FN_BIN == PGDG_FN_BIN && FN_SPLIT == PGDG_FN_SPLIT && FN_MERGE == PGDG_FN_MERGE &&
                  FN_MEMBERSHIP == PGDG_FN_MEMBERSHIP && FN_NEIGHBORS == PGDG_FN_NEIGHBORS &&
                  FN_P2G == PGDG_FN_P2G && FN_LIMITER == PGDG_FN_LIMITER && FN_RHS == PGDG_FN_RHS &&
                  FN_UPDATE_GRID == PGDG_FN_UPDATE_GRID &&
                  FN_UPDATE_PARTICLE == PGDG_FN_UPDATE_PARTICLE && FN_RUN == PGDG_FN_RUN,
              "stage numbers must match include/pgdg.h");

// Particle-update argument inside a time step: the stage code, plus
// kCopyLater for the stages before the last (0 of either scheme, 2 of
// SSP-RK3), whose particle coefficient copies and conserved state the last
// stage overwrites before anything reads them.  Stage-level calls always
// write them.
constexpr int kCopyLater = 0x100;
// ... and kBinsKept for the stages after the first: the bins (and the
// mixed form's entry-to-cell map built from them) are those of the first.
constexpr int kBinsKept = 0x400;
// Split/merge argument inside a time step: leave the removed particles for
// the compaction after the last merge pass.
constexpr int kDefer = 0x800;
// RHS argument inside a time step: kFuseRk | stage code (see FN_RHS).
constexpr int kFuseRk = 0x200;
static int g2p_code(int md) { return ((md == 0 || md == 2) ? (md | kCopyLater) : md) | (md != 0 ? kBinsKept : 0); }

static int prof_slot(int fn) {
  switch (fn) {
    case 0: return PS_TIMESTEP;
    case 1: return PS_USED;
    case 2: return PS_BIN;
    case 3: return PS_SPLIT;
    case 4: return PS_MERGE;
    case 5: return PS_MEMBERSHIP;
    case 6: return PS_STAGE_SAVE;
    case 7: return PS_P2G;
    case 8: return PS_LIMITER;
    case 9: return PS_RHS;
    case 10: return PS_UPDATE_GRID;
    case 13: return PS_COMPACT;
    default: return -1;
  }
}

template <int D_>
int Ctx::stage(int fn, int arg) {
  int slot = prof.on ? prof_slot(fn) : -1;
  if (slot < 0) return stage_impl<D_>(fn, arg);
  Profiler::Rec r;
  PGDG_TRY(prof.begin(slot, st, r));
  int rc = stage_impl<D_>(fn, arg);
  PGDG_TRY(prof.end(r, st));
  return rc;
}

template <int D_>
int Ctx::stage_impl(int fn, int arg) {
  PtlDev P = PA.view();
  switch (fn) {
    case FN_SET_TIMESTEP:
      return set_timestep<D_>();
    case FN_USED:
      hipLaunchKernelGGL(k_used_clear, dim3((nc + 255) / 256), dim3(256), 0, st, used, nc);
      hipLaunchKernelGGL(k_used_mark, dim3((nc + 255) / 256), dim3(256), 0, st, g, C.ph_state, used, D_);
      PGDG_HIP_TRY(hipGetLastError());
      return PGDG_OK;
    case FN_BIN:
      PGDG_TRY(T.build_bins<D_>(g, P, W, np));
      bins_valid = true;
      return PGDG_OK;
    case FN_SPLIT:
      if (!bins_valid) return PGDG_ERR_STATE;
      PGDG_TRY(T.split<D_>(g, PA, W));
      return (arg & kDefer) ? PGDG_OK : stage_impl<D_>(FN_TOPO_FINISH, 0);
    case FN_MERGE:
      if (!bins_valid) return PGDG_ERR_STATE;
      PGDG_TRY(T.merge<D_>(g, PA, W));
      return (arg & kDefer) ? PGDG_OK : stage_impl<D_>(FN_TOPO_FINISH, 0);
    case FN_TOPO_FINISH: {
      PGDG_TRY(T.finish(PA, &np));
      n_split = T.hc.n_split;
      n_inlet = T.hc.n_inlet;
      n_outlet = T.hc.n_outlet;
      n_merge = T.hc.n_merge;
      return PGDG_OK;
    }
    case FN_MEMBERSHIP:
      return launch_in_ele<D_>(g, P, S, np, st);
    case FN_NEIGHBORS:
      if (!capturing) PGDG_TRY(sync_counts());
      lists_built = true;
      return launch_stage_save<D_>(g, P, S, capturing ? PA.cap : np, arg, sc, st);
    case FN_P2G:
      if (!bins_valid) return PGDG_ERR_STATE;
      return launch_p2g<D_>(g, C, P, T.bin_off, T.bin_ent, np + W.n, used, PB, st);
    case FN_LIMITER:
      if (tuning_now)
        PGDG_TRY(pick_block([&](int b) {
          return launch_cells<D_>(CELL_LIMITER, g, C, P, T.bin_off, T.bin_ent, used, sc, 0, st, b);
        }, &tune.limiter));
      return launch_cells<D_>(CELL_LIMITER, g, C, P, T.bin_off, T.bin_ent, used, sc, 0, st,
                              tune.limiter);
    case FN_RHS: {
      // mixed form inside a time step (kFuseRk | stage code, and not while
      // the candidate sizes are timed: the fused update is not idempotent):
      // the RHS kernel applies the stage's coefficient update itself
      const bool fuse = cs.precision == PGDG_PREC_MIXED && (arg & kFuseRk) && !tuning_now;
      rhs_fused = fuse;
      const int ra = fuse ? (arg & 3) : -1;
      if (tuning_now)
        PGDG_TRY(pick_block([&](int b) {
          return launch_cells<D_>(CELL_RHS, g, C, P, T.bin_off, T.bin_ent, used, sc, -1, st, b, &ib);
        }, &tune.rhs));
      return launch_cells<D_>(CELL_RHS, g, C, P, T.bin_off, T.bin_ent, used, sc, ra, st, tune.rhs, &ib);
    }
    case FN_UPDATE_GRID: {
      const int ua = (arg & 3) | (rhs_fused ? kCellSkipCoef : 0);
      rhs_fused = false;
      return launch_cells<D_>(CELL_UPDATE, g, C, P, T.bin_off, T.bin_ent, used, sc, ua, st);
    }
    case FN_UPDATE_PARTICLE: {
      if (!bins_valid) return PGDG_ERR_STATE;
      if (!capturing) PGDG_TRY(sync_counts());
      const int64_t n_p = capturing ? PA.cap : np, n_e = capturing ? B.ne_cap : T.n_ent;
      Profiler::Rec r;
      if (prof.on) PGDG_TRY(prof.begin(PS_G2P, st, r));
      PGDG_TRY(launch_g2p<D_>(g, P, C, S, n_p, sc, arg & 3, !(arg & kCopyLater), st));
      // parts: 1 the neighbour records and cell boxes, 2 the shifting
      // kernel, 3 both (timed separately when profiling)
      auto pst = [&](int parts) -> int {
        if (cs.precision == PGDG_PREC_MIXED)
          return launch_pst_mixed<D_>(g, P, W, S, B, MB, n_e, T.bin_off, T.bin_ent, sc, st, tune.pst,
                                      pst_staging, parts, (arg & kBinsKept) != 0);
        return launch_pst_binned<D_>(g, P, W, S, B, n_e, T.bin_off, T.bin_ent, sc, st, tune.pst, parts);
      };
      if (prof.on) {
        PGDG_TRY(prof.end(r, st));
        PGDG_TRY(prof.begin(PS_PST_PREP, st, r));
        PGDG_TRY(pst(1));
        PGDG_TRY(prof.end(r, st));
        PGDG_TRY(prof.begin(PS_PST, st, r));
        PGDG_TRY(pst(2));
      } else {
        PGDG_TRY(pst(3));
      }
      if (cs.precision == PGDG_PREC_MIXED) {
        if (tuning_now) {  // the gathered records are in place: time the shifting kernel alone
          PGDG_TRY(pick_block([&](int b) {
            return launch_pst_mixed_kernel<D_>(g, P, W, S, B, MB, n_e, T.bin_off, sc, st, b, pst_staging);
          }, &tune.pst));
          PGDG_TRY(launch_pst_mixed_kernel<D_>(g, P, W, S, B, MB, n_e, T.bin_off, sc, st, tune.pst,
                                               pst_staging));
        }
      } else {
        if (tuning_now) {
          PGDG_TRY(pick_block([&](int b) {
            return launch_pst_kernel<D_>(g, P, W, S, B, n_e, T.bin_off, sc, st, b);
          }, &tune.pst));
          PGDG_TRY(launch_pst_kernel<D_>(g, P, W, S, B, n_e, T.bin_off, sc, st, tune.pst));
        }
      }
      if (prof.on) PGDG_TRY(prof.end(r, st));
      return PGDG_OK;
    }
    case FN_RUN:
      // MPM_run_simulation(n, ...) receives the caller's iteration number
      // (the Solver increments mpm.iterate_num first, S3:solver.cpp:51-53);
      // the next Set_timestep reads it
      iterate_num = arg;
      return run_simulation<D_>(arg);
    default:
      return PGDG_ERR_ARG;
  }
}

template <int D_>
int Ctx::run_simulation(int64_t n) {
  n_merge = 0;
  PGDG_TRY(T.reset_step_counts());
  if (n == 1) PGDG_TRY(stage<D_>(FN_USED, 0));
  PGDG_TRY(stage<D_>(FN_BIN, 0));
  PGDG_TRY(stage<D_>(FN_SPLIT, kDefer));
  for (int it = 0; it < 4; it++) PGDG_TRY(stage<D_>(FN_MERGE, kDefer));
  PGDG_TRY(stage<D_>(FN_TOPO_FINISH, 0));
  PGDG_TRY(stage<D_>(FN_MEMBERSHIP, 0));
  // stage codes of the cell and particle updates: Heun (reference) 0, 1;
  // SSP-RK3 0, 2, 3
  const int codes3[3] = {0, 2, 3}, codes2[2] = {0, 1};
  const int ns = cs.rk_stages == 3 ? 3 : 2;
  const int* codes = ns == 3 ? codes3 : codes2;
  if (autotune && !tuned && !prof.on) {
    // first step: eager stages, timing the candidate workgroup sizes in place
    for (int q = 0; q < ns; q++) {
      const int md = codes[q];
      tuning_now = md == 0;
      int rc = stage_impl<D_>(FN_NEIGHBORS, md);
      if (rc == PGDG_OK && md == 0) rc = stage_impl<D_>(FN_P2G, 0);
      if (rc == PGDG_OK) rc = stage_impl<D_>(FN_LIMITER, 0);
      if (rc == PGDG_OK) rc = stage_impl<D_>(FN_RHS, kFuseRk | md);
      if (rc == PGDG_OK) rc = stage_impl<D_>(FN_UPDATE_GRID, md);
      if (rc == PGDG_OK) rc = stage_impl<D_>(FN_UPDATE_PARTICLE, g2p_code(md));
      tuning_now = false;
      PGDG_TRY(rc);
    }
    tuned = true;
  } else if (use_graphs && !prof.on) {
    PGDG_TRY(sync_counts());
    for (int q = 0; q < ns; q++) PGDG_TRY(rk_stage<D_>(codes[q]));
  } else {
    for (int q = 0; q < ns; q++) {
      const int md = codes[q];
      PGDG_TRY(stage<D_>(FN_NEIGHBORS, md));
      if (md == 0) PGDG_TRY(stage<D_>(FN_P2G, 0));
      PGDG_TRY(stage<D_>(FN_LIMITER, 0));
      PGDG_TRY(stage<D_>(FN_RHS, kFuseRk | md));
      PGDG_TRY(stage<D_>(FN_UPDATE_GRID, md));
      PGDG_TRY(stage<D_>(FN_UPDATE_PARTICLE, g2p_code(md)));
    }
  }
  // lists the next step's split/merge passes read (built from the last
  // stage's neighbour-list inputs)
  Profiler::Rec r;
  if (prof.on) PGDG_TRY(prof.begin(PS_LISTS, st, r));
  PGDG_TRY(T.end_of_step_lists<D_>(g, PA.view(), S.xb, S.hb, W, np));
  if (prof.on) PGDG_TRY(prof.end(r, st));
  if (ib.mode != IFACE_INLINE) PGDG_TRY(fit_iface_records());
  return PGDG_OK;
}

// The interface-point records of a batched stage live in a buffer sized
// for the largest count seen: a face whose records did not fit solved its
// points in the face kernel (same solutions with the reference's initial
// p*), and the buffer grows before the next step.
int Ctx::fit_iface_records() {
  int32_t peak = 0;
  PGDG_HIP_TRY(hipMemcpyAsync(&peak, C.mg_peak, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PGDG_HIP_TRY(hipStreamSynchronize(st));
  if (peak <= C.mg_cap) return PGDG_OK;
  return alloc_iface_records((int64_t)peak + peak / 4);
}

int Ctx::alloc_iface_records(int64_t cap_rec) {
  if (mg_slab) {
    PGDG_HIP_TRY(hipStreamSynchronize(st));
    drop_graphs();  // the stage graphs hold the buffer's addresses
    (void)hipFree(mg_slab);
    mg_slab = nullptr;
  }
  C.mg_cap = cap_rec;
  const size_t bytes_ = sizeof(double) * 9 * (size_t)cap_rec +
                        sizeof(int32_t) * (2 * (size_t)cap_rec + (size_t)D * nc + 64);
  PGDG_HIP_TRY(hipMalloc(&mg_slab, bytes_));
  PGDG_HIP_TRY(hipMemsetAsync(mg_slab, 0, bytes_, st));
  C.mg_in = (double*)mg_slab;
  C.mg_out = C.mg_in + 6 * (size_t)cap_rec;
  C.mg_pair = (int32_t*)(C.mg_out + 3 * (size_t)cap_rec);
  C.mg_it = C.mg_pair + cap_rec;
  C.mg_base = C.mg_it + cap_rec;
  C.mg_count = C.mg_base + (size_t)D * nc;
  C.mg_peak = C.mg_count + 1;
  mg_bytes = bytes_;
  return PGDG_OK;
}

// One RK stage as a HIP graph: captured on first use for the current pair
// of alternating buffers (grids cover the capacities; the kernels read the
// counts from the device scalars), then replayed.
template <int D_>
int Ctx::rk_stage(int md) {
  const void* k1 = PA.view().x;
  const void* k2 = T.bin_ent;  // fixed; kept in the key for clarity
  for (int q = 0; q < ngraphs; q++)
    if (graphs[q].md == md && graphs[q].k1 == k1 && graphs[q].k2 == k2) {
      PGDG_HIP_TRY(hipGraphLaunch(graphs[q].exec, st));
      return PGDG_OK;
    }
  if (ngraphs == kMaxGraphs) drop_graphs();
  PGDG_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  capturing = true;
  int rc = stage_impl<D_>(FN_NEIGHBORS, md);
  if (rc == PGDG_OK && md == 0) rc = stage_impl<D_>(FN_P2G, 0);
  if (rc == PGDG_OK) rc = stage_impl<D_>(FN_LIMITER, 0);
  if (rc == PGDG_OK) rc = stage_impl<D_>(FN_RHS, kFuseRk | md);
  if (rc == PGDG_OK) rc = stage_impl<D_>(FN_UPDATE_GRID, md);
  if (rc == PGDG_OK) rc = stage_impl<D_>(FN_UPDATE_PARTICLE, g2p_code(md));
  capturing = false;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc != PGDG_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return PGDG_ERR_HIP;
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return PGDG_ERR_HIP;
  graphs[ngraphs++] = StageGraph{md, k1, k2, exec};
  PGDG_HIP_TRY(hipGraphLaunch(exec, st));
  return PGDG_OK;
}

int Ctx::step(int nsteps) {
  for (int s = 0; s < nsteps; s++) {
    if (D == 2) {
      PGDG_TRY(stage<2>(FN_SET_TIMESTEP, 0));
      iterate_num++;
      PGDG_TRY(run_simulation<2>(iterate_num));
    } else {
      PGDG_TRY(stage<3>(FN_SET_TIMESTEP, 0));
      iterate_num++;
      PGDG_TRY(run_simulation<3>(iterate_num));
    }
  }
  return PGDG_OK;
}

}  // namespace pgdg

struct pgdg_ctx {
  pgdg::Ctx c;
};

using pgdg::Ctx;

extern "C" {

int pgdg_create(const pgdg_case* c, int64_t particle_cap, int64_t n_wall, pgdg_ctx** out) {
  if (!c || !out || particle_cap <= 0 || n_wall < 0) return PGDG_ERR_ARG;
  if (c->dim != 2 && c->dim != 3) return PGDG_ERR_ARG;
  if (c->nphase < 2 || c->nphase > PGDG_MAX_PHASE) return PGDG_ERR_ARG;
  if (c->order != 1 && c->order != 2) return PGDG_ERR_ARG;
  if (c->rk_stages != 2 && c->rk_stages != 3) return PGDG_ERR_ARG;
  if (c->cfl < 0 || c->ppc_min < 0 || c->ppc_max < 0) return PGDG_ERR_ARG;
  if (c->ppc_max > 0 && c->ppc_min > c->ppc_max) return PGDG_ERR_ARG;
  if (c->semantics != PGDG_SEM_T2 && c->semantics != PGDG_SEM_S3 && c->semantics != PGDG_SEM_UX)
    return PGDG_ERR_ARG;
  if (c->precision != PGDG_PREC_FP64 && c->precision != PGDG_PREC_MIXED) return PGDG_ERR_ARG;
  if (c->semantics == PGDG_SEM_UX) {
    // the underwater-explosion routines: 2-D, two phases, FP64 form; x sides
    // transmissive, y sides transmissive or reflecting (the reference's
    // tags -1 and 2)
    if (c->dim != 2 || c->nphase != 2 || c->precision != PGDG_PREC_FP64) return PGDG_ERR_ARG;
    for (int f = 0; f < 4; f++) {
      const bool yside = f == PGDG_SIDE_YL || f == PGDG_SIDE_YR;
      if (c->bc[f] != PGDG_BC_OUTLET && !(yside && c->bc[f] == PGDG_BC_WALL)) return PGDG_ERR_ARG;
    }
  }
  // every domain side needs a boundary treatment (interior is no tag for a
  // side of the domain); sides 4 and 5 (z) only in 3-D
  for (int f = 0; f < 2 * c->dim; f++)
    if (c->bc[f] != PGDG_BC_WALL && c->bc[f] != PGDG_BC_INLET && c->bc[f] != PGDG_BC_OUTLET &&
        c->bc[f] != PGDG_BC_SYMMETRY)
      return PGDG_ERR_ARG;
  if (particle_cap > (int64_t)0x7ffffff0) return PGDG_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PGDG_ERR_NODEVICE;
  pgdg_ctx* x = new (std::nothrow) pgdg_ctx();
  if (!x) return PGDG_ERR_CAPACITY;
  int r = x->c.init(c, particle_cap, n_wall);
  if (r != PGDG_OK) {
    x->c.release();
    delete x;
    return r;
  }
  *out = x;
  return PGDG_OK;
}

int pgdg_destroy(pgdg_ctx* x) {
  if (!x) return PGDG_ERR_ARG;
  x->c.release();
  delete x;
  return PGDG_OK;
}

int pgdg_upload_state(pgdg_ctx* x, const pgdg_particles* p, const pgdg_walls* w,
                      const pgdg_cells* c, const pgdg_scalars* s) {
  if (!x) return PGDG_ERR_ARG;
  return x->c.upload(p, w, c, s);
}

int pgdg_download_state(pgdg_ctx* x, pgdg_particles* p, pgdg_cells* c, pgdg_scalars* s) {
  if (!x) return PGDG_ERR_ARG;
  return x->c.download(p, c, s);
}

int pgdg_step(pgdg_ctx* x, int nsteps) {
  if (!x || nsteps < 0) return PGDG_ERR_ARG;
  if (!x->c.loaded) return PGDG_ERR_STATE;
  return x->c.step(nsteps);
}

int pgdg_stage(pgdg_ctx* x, int fn, int arg) {
  if (!x) return PGDG_ERR_ARG;
  if (!x->c.loaded) return PGDG_ERR_STATE;
  if ((fn == pgdg::FN_UPDATE_GRID || fn == pgdg::FN_UPDATE_PARTICLE) && (arg < 0 || arg > 3))
    return PGDG_ERR_ARG;  // stage codes 0-3
  if (fn == pgdg::FN_RHS && arg != 0) return PGDG_ERR_ARG;
  if ((fn == pgdg::FN_SPLIT || fn == pgdg::FN_MERGE) && arg != 0) return PGDG_ERR_ARG;
  if (fn < 0 || fn > pgdg::FN_RUN) return PGDG_ERR_ARG;
  return x->c.D == 2 ? x->c.stage<2>(fn, arg) : x->c.stage<3>(fn, arg);
}

int pgdg_sync(pgdg_ctx* x) {
  if (!x) return PGDG_ERR_ARG;
  PGDG_HIP_TRY(hipStreamSynchronize(x->c.st));
  return PGDG_OK;
}

int pgdg_get_stream(pgdg_ctx* x, void** stream) {
  if (!x || !stream) return PGDG_ERR_ARG;
  *stream = (void*)x->c.st;
  return PGDG_OK;
}

int pgdg_particle_count(pgdg_ctx* x, int64_t* n) {
  if (!x || !n) return PGDG_ERR_ARG;
  *n = x->c.np;
  return PGDG_OK;
}

int pgdg_graphs_enable(pgdg_ctx* x, int on) {
  if (!x) return PGDG_ERR_ARG;
  PGDG_HIP_TRY(hipStreamSynchronize(x->c.st));
  x->c.use_graphs = on != 0;
  if (!on) x->c.drop_graphs();
  return PGDG_OK;
}

int pgdg_autotune_enable(pgdg_ctx* x, int on) {
  if (!x) return PGDG_ERR_ARG;
  x->c.autotune = on != 0;
  x->c.tuned = false;
  x->c.drop_graphs();
  return PGDG_OK;
}

}  // extern "C"

namespace {

// Neighbour-list sizes counted into the device buffer d (np entries, every
// one written) and copied to the host.  Returns 1 in *done when the context
// answered without counting (no particles, no lists yet).
int neighbor_counts_prepare(Ctx& c, int32_t* n_neighbor, int* done) {
  *done = 1;
  if (!c.loaded) return PGDG_ERR_STATE;
  if (c.np == 0) return PGDG_OK;
  if (!c.lists_built) {
    memset(n_neighbor, 0, sizeof(int32_t) * (size_t)c.np);
    return PGDG_OK;
  }
  if (!c.bins_valid) return PGDG_ERR_STATE;
  *done = 0;
  return PGDG_OK;
}

int neighbor_counts_into(Ctx& c, int32_t* d, int32_t* n_neighbor) {
  PGDG_TRY(c.sync_counts());
  int rc = c.D == 2 ? pgdg::launch_neighbor_counts<2>(c.g, c.PA.view(), c.W, c.S, c.B, c.T.n_ent, c.T.bin_off,
                                                       c.T.bin_ent, c.sc, c.st, d)
                    : pgdg::launch_neighbor_counts<3>(c.g, c.PA.view(), c.W, c.S, c.B, c.T.n_ent, c.T.bin_off,
                                                       c.T.bin_ent, c.sc, c.st, d);
  if (rc == PGDG_OK && hipMemcpyAsync(n_neighbor, d, sizeof(int32_t) * (size_t)c.np, hipMemcpyDeviceToHost,
                                      c.st) != hipSuccess)
    rc = PGDG_ERR_HIP;
  if (rc == PGDG_OK && hipStreamSynchronize(c.st) != hipSuccess) rc = PGDG_ERR_HIP;
  return rc;
}

// Scratch of pgdg_neighbor_counts_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends.  Host code uses the address
// hipGetSymbolAddress returns for the current device, never the array's
// name.  Calls that overlap on one GPU share it and race.
constexpr int64_t kNoAllocMaxParticles = (int64_t)1 << 21;  // 8 MiB of int32
__device__ __align__(256) int32_t scratch[kNoAllocMaxParticles];
constexpr int kMaxDevices = 64;
int32_t* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_ptr(int32_t** out) {
  int dev = 0;
  PGDG_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return PGDG_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    PGDG_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return PGDG_ERR_HIP;
    g_scratch[dev] = static_cast<int32_t*>(p);
  }
  *out = g_scratch[dev];
  return PGDG_OK;
}

__global__ void k_zero_i32(int32_t* a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0;
}

}  // namespace

extern "C" {

int pgdg_neighbor_counts(pgdg_ctx* x, int32_t* n_neighbor) {
  if (!x || !n_neighbor) return PGDG_ERR_ARG;
  Ctx& c = x->c;
  int done = 0;
  PGDG_TRY(neighbor_counts_prepare(c, n_neighbor, &done));
  if (done) return PGDG_OK;
  int32_t* d = nullptr;
  PGDG_HIP_TRY(hipMalloc((void**)&d, sizeof(int32_t) * (size_t)c.np));
  const int rc = neighbor_counts_into(c, d, n_neighbor);
  (void)hipFree(d);
  return rc;
}

int pgdg_neighbor_counts_noalloc(pgdg_ctx* x, int32_t* n_neighbor) {
  if (!x || !n_neighbor) return PGDG_ERR_ARG;
  Ctx& c = x->c;
  if (c.np > kNoAllocMaxParticles) return PGDG_ERR_CAPACITY;
  int done = 0;
  PGDG_TRY(neighbor_counts_prepare(c, n_neighbor, &done));
  if (done) return PGDG_OK;
  int32_t* d = nullptr;
  PGDG_TRY(scratch_ptr(&d));
  // the previous call's counts are still there: start from zero
  hipLaunchKernelGGL(k_zero_i32, dim3(256), dim3(256), 0, c.st, d, (int64_t)c.np);
  PGDG_HIP_TRY(hipGetLastError());
  return neighbor_counts_into(c, d, n_neighbor);
}

int pgdg_cell_particle_counts(pgdg_ctx* x, int32_t* counts) {
  if (!x || !counts) return PGDG_ERR_ARG;
  Ctx& c = x->c;
  if (!c.loaded) return PGDG_ERR_STATE;
  if (!c.bins_valid) return PGDG_ERR_STATE;
  PGDG_TRY(c.T.fluid_counts());
  PGDG_HIP_TRY(hipMemcpyAsync(counts, c.T.fcnt, sizeof(int32_t) * (size_t)c.nc, hipMemcpyDeviceToHost, c.st));
  PGDG_HIP_TRY(hipStreamSynchronize(c.st));
  return PGDG_OK;
}

int pgdg_used_cells(pgdg_ctx* x, int32_t* used) {
  if (!x || !used) return PGDG_ERR_ARG;
  Ctx& c = x->c;
  if (!c.loaded) return PGDG_ERR_STATE;
  PGDG_HIP_TRY(hipMemcpyAsync(used, c.used, sizeof(int32_t) * (size_t)c.nc, hipMemcpyDeviceToHost, c.st));
  PGDG_HIP_TRY(hipStreamSynchronize(c.st));
  return PGDG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Checkpoint / restart.  The reference creates a restart directory and never
// writes it (S3:Makefile:95, 115); here a checkpoint holds what a resumed
// run needs to continue bit for bit: the particle, wall and cell arrays, the
// time-loop scalars, the used-cell flags (set in the first step only) and
// the neighbour lists materialised at the end of the last step, which the
// next step's split and merge passes read.
// ---------------------------------------------------------------------------
namespace {

struct CkptHead {
  char magic[8];
  int32_t dim, nphase, order, rk_stages, semantics, precision, n[3], U, M, pad;
  int64_t np, nwall, nc, pool_used;
};
const char kCkptMagic[8] = {'P', 'G', 'D', 'G', 'C', 'K', 'P', '1'};

// Host images of the particle and cell arrays with the pgdg.h layouts.
struct HostImage {
  std::vector<double> pd, cd, wd;
  std::vector<int32_t> pi, ci, used, lcnt, pool;
  std::vector<int64_t> loff;
  pgdg_particles P{};
  pgdg_cells C{};
  pgdg_walls W{};
  pgdg_scalars S{};

  void shape(int64_t n, int64_t nw, int64_t nc, int F, int U, int M) {
    const int64_t pc = n > 0 ? n : 1;
    pd.assign((size_t)(19 + U * M) * pc, 0.0);
    pi.assign((size_t)2 * pc, 0);
    double* q = pd.data();
    auto take = [&](int64_t rows, int64_t len) {
      double* r = q;
      q += (size_t)rows * len;
      return r;
    };
    P.n = n;
    P.cap = pc;
    P.x = take(3, pc), P.v = take(3, pc), P.tv = take(3, pc), P.rhov = take(3, pc);
    P.vol = take(1, pc), P.mass = take(1, pc), P.h = take(1, pc), P.P = take(1, pc);
    P.rho = take(1, pc), P.E = take(1, pc), P.c = take(1, pc);
    P.coef = take(U * M, pc);
    P.phase = pi.data();
    P.mark_in = pi.data() + pc;
    cd.assign((size_t)(3 * F * U * M + F * U + 8 * F + 16) * nc, 0.0);
    ci.assign((size_t)2 * F * nc, 0);
    q = cd.data();
    C.coef = take(F * U * M, nc), C.coef0 = take(F * U * M, nc), C.dcoef = take(F * U * M, nc);
    C.lim = take(F * U, nc);
    C.alph = take(F, nc), C.rho = take(F, nc), C.E = take(F, nc), C.c = take(F, nc), C.ac = take(F, nc);
    C.rhov = take(3 * F, nc);
    C.v = take(3, nc), C.a = take(3, nc), C.P = take(1, nc), C.G_v = take(9, nc);
    C.ph_state = ci.data();
    C.ph_state_b = ci.data() + (size_t)F * nc;
    wd.assign((size_t)7 * (nw > 0 ? nw : 1), 0.0);
    W.n = nw;
    W.x = wd.data();
    W.norm = wd.data() + 3 * nw;
    W.vol = wd.data() + 6 * nw;
    used.assign((size_t)nc, 0);
    lcnt.assign((size_t)pc, 0);
    loff.assign((size_t)pc, 0);
  }
};

bool put_all(FILE* f, const void* p, size_t bytes) { return bytes == 0 || fwrite(p, 1, bytes, f) == bytes; }
bool get_all(FILE* f, void* p, size_t bytes) { return bytes == 0 || fread(p, 1, bytes, f) == bytes; }
template <class T>
size_t nbytes(const std::vector<T>& v) {
  return sizeof(T) * v.size();
}

}  // namespace

extern "C" {

int pgdg_checkpoint_write(pgdg_ctx* x, const char* path) {
  if (!x || !path) return PGDG_ERR_ARG;
  Ctx& c = x->c;
  if (!c.loaded) return PGDG_ERR_STATE;
  PGDG_HIP_TRY(hipStreamSynchronize(c.st));
  HostImage h;
  const int64_t n = c.np, nw = c.W.n, pool_used = c.T.pool_used;
  h.shape(n, nw, c.nc, c.F, c.U, c.M);
  PGDG_TRY(c.download(&h.P, &h.C, &h.S));
  if (nw > 0) {
    PGDG_HIP_TRY(hipMemcpy(h.W.x, c.W.x, sizeof(double) * 3 * nw, hipMemcpyDeviceToHost));
    PGDG_HIP_TRY(hipMemcpy(h.W.norm, c.W.norm, sizeof(double) * 3 * nw, hipMemcpyDeviceToHost));
    PGDG_HIP_TRY(hipMemcpy(h.W.vol, c.W.vol, sizeof(double) * nw, hipMemcpyDeviceToHost));
  }
  PGDG_HIP_TRY(hipMemcpy(h.used.data(), c.used, sizeof(int32_t) * c.nc, hipMemcpyDeviceToHost));
  if (n > 0) {
    PGDG_HIP_TRY(hipMemcpy(h.lcnt.data(), c.T.lcnt, sizeof(int) * n, hipMemcpyDeviceToHost));
    PGDG_HIP_TRY(hipMemcpy(h.loff.data(), c.T.loff, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
  }
  h.pool.assign((size_t)pool_used, 0);
  if (pool_used > 0)
    PGDG_HIP_TRY(hipMemcpy(h.pool.data(), c.T.pool, sizeof(int) * pool_used, hipMemcpyDeviceToHost));
  CkptHead hd{};
  memcpy(hd.magic, kCkptMagic, 8);
  hd.dim = c.cs.dim, hd.nphase = c.cs.nphase, hd.order = c.cs.order, hd.rk_stages = c.cs.rk_stages;
  hd.semantics = c.cs.semantics, hd.precision = c.cs.precision;
  for (int a = 0; a < 3; a++) hd.n[a] = c.cs.n[a];
  hd.U = c.U, hd.M = c.M;
  hd.np = n, hd.nwall = nw, hd.nc = c.nc, hd.pool_used = pool_used;
  FILE* f = fopen(path, "wb");
  if (!f) return PGDG_ERR_ARG;
  bool ok = put_all(f, &hd, sizeof(hd)) && put_all(f, h.pd.data(), nbytes(h.pd)) &&
            put_all(f, h.pi.data(), nbytes(h.pi)) && put_all(f, h.cd.data(), nbytes(h.cd)) &&
            put_all(f, h.ci.data(), nbytes(h.ci)) && put_all(f, h.wd.data(), nbytes(h.wd)) &&
            put_all(f, &h.S, sizeof(h.S)) && put_all(f, h.used.data(), nbytes(h.used)) &&
            put_all(f, h.lcnt.data(), nbytes(h.lcnt)) && put_all(f, h.loff.data(), nbytes(h.loff)) &&
            put_all(f, h.pool.data(), nbytes(h.pool));
  ok = (fclose(f) == 0) && ok;
  return ok ? PGDG_OK : PGDG_ERR_ARG;
}

int pgdg_checkpoint_read(pgdg_ctx* x, const char* path) {
  if (!x || !path) return PGDG_ERR_ARG;
  Ctx& c = x->c;
  FILE* f = fopen(path, "rb");
  if (!f) return PGDG_ERR_ARG;
  CkptHead hd{};
  HostImage h;
  int rc = PGDG_OK;
  if (!get_all(f, &hd, sizeof(hd)) || memcmp(hd.magic, kCkptMagic, 8) != 0) rc = PGDG_ERR_ARG;
  // the context must have been created for the same case
  if (rc == PGDG_OK &&
      (hd.dim != c.cs.dim || hd.nphase != c.cs.nphase || hd.order != c.cs.order ||
       hd.rk_stages != c.cs.rk_stages || hd.semantics != c.cs.semantics || hd.precision != c.cs.precision ||
       hd.n[0] != c.cs.n[0] || hd.n[1] != c.cs.n[1] || hd.n[2] != c.cs.n[2] || hd.U != c.U || hd.M != c.M ||
       hd.nc != c.nc || hd.nwall != c.W.n))
    rc = PGDG_ERR_ARG;
  if (rc == PGDG_OK && (hd.np < 0 || hd.np > c.PA.cap)) rc = PGDG_ERR_CAPACITY;
  if (rc == PGDG_OK && (hd.pool_used < 0 || hd.pool_used > c.T.pool_cap)) rc = PGDG_ERR_CAPACITY;
  if (rc == PGDG_OK) {
    h.shape(hd.np, hd.nwall, hd.nc, c.F, c.U, c.M);
    h.pool.assign((size_t)hd.pool_used, 0);
    const bool ok = get_all(f, h.pd.data(), nbytes(h.pd)) && get_all(f, h.pi.data(), nbytes(h.pi)) &&
                    get_all(f, h.cd.data(), nbytes(h.cd)) && get_all(f, h.ci.data(), nbytes(h.ci)) &&
                    get_all(f, h.wd.data(), nbytes(h.wd)) && get_all(f, &h.S, sizeof(h.S)) &&
                    get_all(f, h.used.data(), nbytes(h.used)) && get_all(f, h.lcnt.data(), nbytes(h.lcnt)) &&
                    get_all(f, h.loff.data(), nbytes(h.loff)) && get_all(f, h.pool.data(), nbytes(h.pool));
    if (!ok) rc = PGDG_ERR_ARG;
  }
  fclose(f);
  if (rc != PGDG_OK) return rc;
  // values the kernels index with: phases, list extents within the pool
  for (int64_t i = 0; i < hd.np; i++) {
    const int32_t ph = h.P.phase[i];
    if (ph < 1 || ph > c.F) return PGDG_ERR_ARG;
    const int n = h.lcnt[i];
    const int64_t o = h.loff[i];
    if (n < 0 || (n > 0 && (o < 0 || o + n > hd.pool_used))) return PGDG_ERR_ARG;
  }
  // a fluid particle, a removed one (-1) or a wall (-(w + 2))
  for (int64_t q = 0; q < hd.pool_used; q++) {
    const int r = h.pool[q];
    if (!((r >= -1 && r < hd.np) || (r <= -2 && -(int64_t)r - 2 < hd.nwall))) return PGDG_ERR_ARG;
  }
  c.drop_graphs();
  PGDG_TRY(c.upload(&h.P, &h.W, &h.C, &h.S));
  PGDG_HIP_TRY(hipMemcpy(c.used, h.used.data(), sizeof(int32_t) * c.nc, hipMemcpyHostToDevice));
  if (hd.np > 0) {
    PGDG_HIP_TRY(hipMemcpy(c.T.lcnt, h.lcnt.data(), sizeof(int) * hd.np, hipMemcpyHostToDevice));
    PGDG_HIP_TRY(hipMemcpy(c.T.loff, h.loff.data(), sizeof(int64_t) * hd.np, hipMemcpyHostToDevice));
  }
  if (hd.pool_used > 0)
    PGDG_HIP_TRY(hipMemcpy(c.T.pool, h.pool.data(), sizeof(int) * hd.pool_used, hipMemcpyHostToDevice));
  PGDG_TRY(c.T.set_pool_used(hd.pool_used));
  PGDG_HIP_TRY(hipStreamSynchronize(c.st));
  return PGDG_OK;
}

int pgdg_pst_staging_enable(pgdg_ctx* x, int on) {
  if (!x) return PGDG_ERR_ARG;
  x->c.pst_staging = on != 0;
  x->c.drop_graphs();
  return PGDG_OK;
}

int pgdg_rhs_mfma_enable(pgdg_ctx* x, int on) {
  if (!x) return PGDG_ERR_ARG;
  x->c.g.mfma_rhs = on != 0;
  x->c.drop_graphs();
  return PGDG_OK;
}

int pgdg_iface_mgfm_mode(pgdg_ctx* x, int mode, pgdg_mlp* m) {
  using namespace pgdg;
  if (!x) return PGDG_ERR_ARG;
  Ctx& c = x->c;
  if (mode != PGDG_IFACE_INLINE && mode != PGDG_IFACE_BATCH && mode != PGDG_IFACE_LEARNED) return PGDG_ERR_ARG;
  if (mode != PGDG_IFACE_INLINE && c.cs.precision != PGDG_PREC_MIXED) return PGDG_ERR_ARG;
  IfaceBatch ib;
  ib.mode = mode;
  if (mode == PGDG_IFACE_LEARNED) {
    if (!m) return PGDG_ERR_ARG;
    PGDG_TRY(mlp_view(m, &ib.P, &ib.nm));
  }
  pgdg_eos_pair eos[kMaxPairs];
  for (int a = 0; a < PGDG_MAX_PHASE; a++)
    for (int b = 0; b < PGDG_MAX_PHASE; b++)
      eos[a * PGDG_MAX_PHASE + b] = pgdg_eos_pair{c.g.gamma[a], c.g.pinf[a], c.g.gamma[b], c.g.pinf[b]};
  ib.tab = make_pair_tab(ib.nm, eos, kMaxPairs);
  c.ib = ib;
  c.drop_graphs();  // the stage graphs hold the face launches' arguments
  return PGDG_OK;
}

int pgdg_iface_pairs(pgdg_ctx* x, pgdg_eos_pair* eos) {
  if (!x || !eos) return PGDG_ERR_ARG;
  const pgdg::Geo& g = x->c.g;
  for (int a = 0; a < PGDG_MAX_PHASE; a++)
    for (int b = 0; b < PGDG_MAX_PHASE; b++)
      eos[a * PGDG_MAX_PHASE + b] = pgdg_eos_pair{g.gamma[a], g.pinf[a], g.gamma[b], g.pinf[b]};
  return PGDG_OK;
}

int pgdg_iface_records(pgdg_ctx* x, double* in, double* out, int32_t* pair, int32_t* iters, int64_t cap,
                       int64_t* n) {
  if (!x || !n || cap < 0) return PGDG_ERR_ARG;
  Ctx& c = x->c;
  if (!c.C.mg_count) return PGDG_ERR_STATE;
  int32_t cnt = 0;
  PGDG_HIP_TRY(hipMemcpyAsync(&cnt, c.C.mg_count, sizeof(int32_t), hipMemcpyDeviceToHost, c.st));
  PGDG_HIP_TRY(hipStreamSynchronize(c.st));
  const int64_t ld = c.C.mg_cap;
  const int64_t have = std::min<int64_t>(cnt, ld);
  *n = have;
  const int64_t k = std::min(have, cap);
  if (k == 0) return PGDG_OK;
  // compact SoA copies with leading dimension k
  for (int q = 0; q < 6 && in; q++)
    PGDG_HIP_TRY(hipMemcpyAsync(in + q * k, c.C.mg_in + q * ld, sizeof(double) * k, hipMemcpyDefault, c.st));
  for (int q = 0; q < 3 && out; q++)
    PGDG_HIP_TRY(hipMemcpyAsync(out + q * k, c.C.mg_out + q * ld, sizeof(double) * k, hipMemcpyDefault, c.st));
  if (pair) PGDG_HIP_TRY(hipMemcpyAsync(pair, c.C.mg_pair, sizeof(int32_t) * k, hipMemcpyDefault, c.st));
  if (iters) PGDG_HIP_TRY(hipMemcpyAsync(iters, c.C.mg_it, sizeof(int32_t) * k, hipMemcpyDefault, c.st));
  PGDG_HIP_TRY(hipStreamSynchronize(c.st));
  return PGDG_OK;
}

int pgdg_tuning_read(pgdg_ctx* x, int* blocks) {
  if (!x || !blocks) return PGDG_ERR_ARG;
  blocks[0] = x->c.tune.pst;
  blocks[1] = x->c.tune.rhs;
  blocks[2] = x->c.tune.limiter;
  return x->c.tuned ? PGDG_OK : PGDG_ERR_STATE;
}

int pgdg_tuning_set(pgdg_ctx* x, const int* blocks) {
  if (!x || !blocks) return PGDG_ERR_ARG;
  for (int q = 0; q < 3; q++)
    if (blocks[q] != 64 && blocks[q] != 128 && blocks[q] != 256) return PGDG_ERR_ARG;
  x->c.tune.pst = blocks[0];
  x->c.tune.rhs = blocks[1];
  x->c.tune.limiter = blocks[2];
  x->c.autotune = false;
  x->c.tuned = true;
  x->c.drop_graphs();
  return PGDG_OK;
}

int pgdg_profile_enable(pgdg_ctx* x, int on) {
  if (!x) return PGDG_ERR_ARG;
  x->c.prof.on = on != 0;
  return PGDG_OK;
}

int pgdg_profile_read(pgdg_ctx* x, double* ms, int64_t* launches, int nslot) {
  if (!x || !ms || !launches || nslot < 0) return PGDG_ERR_ARG;
  return x->c.prof.read(ms, launches, nslot);
}

int pgdg_memory_bytes(pgdg_ctx* x, int64_t* bytes) {
  if (!x || !bytes) return PGDG_ERR_ARG;
  *bytes = (int64_t)x->c.bytes;
  return PGDG_OK;
}

// Reference-named stage entry points (each replaces the MPM member function
// of the same name; see include/pgdg.h).
int pgdg_set_timestep(pgdg_ctx* x) { return pgdg_stage(x, pgdg::FN_SET_TIMESTEP, 0); }
int pgdg_used_state_update(pgdg_ctx* x) { return pgdg_stage(x, pgdg::FN_USED, 0); }
int pgdg_map_particles_to_cells(pgdg_ctx* x) { return pgdg_stage(x, pgdg::FN_BIN, 0); }
int pgdg_split_particle_fast(pgdg_ctx* x) { return pgdg_stage(x, pgdg::FN_SPLIT, 0); }
int pgdg_merge_particle_fast(pgdg_ctx* x) { return pgdg_stage(x, pgdg::FN_MERGE, 0); }
int pgdg_refresh_cell_membership(pgdg_ctx* x) { return pgdg_stage(x, pgdg::FN_MEMBERSHIP, 0); }
int pgdg_refresh_neighbor_info(pgdg_ctx* x, int stage) { return pgdg_stage(x, pgdg::FN_NEIGHBORS, stage); }
int pgdg_mls_p2g(pgdg_ctx* x) { return pgdg_stage(x, pgdg::FN_P2G, 0); }
int pgdg_calculate_limiter(pgdg_ctx* x) { return pgdg_stage(x, pgdg::FN_LIMITER, 0); }
int pgdg_grid_acceleration(pgdg_ctx* x) { return pgdg_stage(x, pgdg::FN_RHS, 0); }
int pgdg_update_grid(pgdg_ctx* x, int md) { return pgdg_stage(x, pgdg::FN_UPDATE_GRID, md); }
int pgdg_update_particle(pgdg_ctx* x, int md) { return pgdg_stage(x, pgdg::FN_UPDATE_PARTICLE, md); }
int pgdg_run_simulation(pgdg_ctx* x, int64_t n) { return pgdg_stage(x, pgdg::FN_RUN, (int)n); }

}  // extern "C"
