// step_topo.hpp - device particle arrays and the topology object (cell bins,
// neighbour-list pool, split/merge scratch).
#pragma once
#include "dev_common.hpp"
#include "dev_scan.hpp"
#include "hip_util.hpp"

namespace pgdg {

// Persistent per-particle arrays with a second set for stable compaction.
struct ParticleArrays {
  int U = 0, M = 0;
  int64_t cap = 0;
  // [0] current, [1] compaction target
  double* f64[2] = {nullptr, nullptr};  // one slab: rows of cap doubles
  int32_t* i32[2] = {nullptr, nullptr}; // phase, mark_in
  int rows64 = 0;                       // number of double rows
  // row layout of the double slab
  enum { R_X = 0, R_V = 3, R_TV = 6, R_RHOV = 9, R_VOL = 12, R_MASS, R_H, R_P, R_RHO, R_E, R_C, R_COEF };

  int init(int U_, int M_, int64_t cap_) {
    U = U_;
    M = M_;
    cap = cap_;
    rows64 = R_COEF + U * M;
    for (int k = 0; k < 2; k++) {
      PGDG_HIP_TRY(hipMalloc((void**)&f64[k], sizeof(double) * rows64 * cap));
      PGDG_HIP_TRY(hipMalloc((void**)&i32[k], sizeof(int32_t) * 2 * cap));
      PGDG_HIP_TRY(hipMemset(f64[k], 0, sizeof(double) * rows64 * cap));
      PGDG_HIP_TRY(hipMemset(i32[k], 0, sizeof(int32_t) * 2 * cap));
    }
    return PGDG_OK;
  }
  void release() {
    for (int k = 0; k < 2; k++) {
      if (f64[k]) (void)hipFree(f64[k]);
      if (i32[k]) (void)hipFree(i32[k]);
      f64[k] = nullptr;
      i32[k] = nullptr;
    }
  }
  PtlDev view() const {
    PtlDev P;
    P.cap = cap;
    double* b = f64[0];
    P.x = b + (size_t)R_X * cap;
    P.v = b + (size_t)R_V * cap;
    P.tv = b + (size_t)R_TV * cap;
    P.rhov = b + (size_t)R_RHOV * cap;
    P.vol = b + (size_t)R_VOL * cap;
    P.mass = b + (size_t)R_MASS * cap;
    P.h = b + (size_t)R_H * cap;
    P.P = b + (size_t)R_P * cap;
    P.rho = b + (size_t)R_RHO * cap;
    P.E = b + (size_t)R_E * cap;
    P.c = b + (size_t)R_C * cap;
    P.coef = b + (size_t)R_COEF * cap;
    P.phase = i32[0];
    P.mark_in = i32[0] + cap;
    return P;
  }
  int gather(const int* newidx, int64_t n, hipStream_t st);
};

// Counters of the topology passes, kept on the device so that binning,
// split and merge run without host round trips.  Particles removed by a
// pass (outlet, merge partners) stay in the arrays, flagged dead, until
// Topo::finish compacts them once after the last pass.
struct TopoCounters {
  long long n_total;    // particles in the arrays, live and dead
  long long n_dead;     // removed, not yet compacted
  long long nent;       // bin entries
  long long pool_used;  // neighbour-list pool entries in use
  long long pool_cap;
  long long n_split, n_inlet, n_outlet, n_merge;  // event counts of the step
  long long tot_split, tot_inlet, tot_outlet;     // this split pass
  long long n_pairs;    // merge pairs of this pass
  long long k_app;      // bin entries appended by this pass
  long long active;     // this pass changes particles or bins
  long long win_len;    // nc when active, else 0
  long long lists_total, list_base, list_ok;
  long long n_contrib;  // per-cell bound contributions of this pass
  long long err;        // PGDG_* error code of the passes since the last finish
};

struct Topo {
  int64_t cap = 0, nwall = 0;
  int nc = 0;
  hipStream_t stream = nullptr;
  TopoCounters* tc = nullptr;  // device
  TopoCounters hc{};           // host copy (read by finish)
  // cell bins (CSR)
  int* bin_off = nullptr;
  int* bin_ent = nullptr;
  int *off_tmp = nullptr, *ent_tmp = nullptr;
  int64_t n_ent = 0;
  uint32_t* keys = nullptr;  // cell of every bin entry being sorted
  int* order = nullptr;      // bucket-sort output
  int *cursor = nullptr, *cnt = nullptr;
  int* fcnt = nullptr;  // fluid entries per cell (per-cell split/merge bounds)
  uint8_t* win = nullptr;
  // per-particle scratch
  uint8_t *rm = nullptr, *dead = nullptr, *src = nullptr, *aff = nullptr, *need_ = nullptr,
          *type = nullptr, *reject = nullptr;
  int *is_a = nullptr, *is_b = nullptr, *is_c = nullptr;  // 0/1 indicators, then their scans
  int* cand = nullptr;
  int* app_ref = nullptr;
  uint32_t* app_cell = nullptr;
  // per-cell bound contributions (cell, candidate, cost), bucketed by cell
  uint32_t* ckey = nullptr;
  int *ccand = nullptr, *ccost = nullptr, *cbase = nullptr, *coff = nullptr;
  int64_t ccap = 0;
  // neighbour-list pool
  int *lcnt = nullptr, *lcnt2 = nullptr, *lc32 = nullptr;
  int64_t *loff = nullptr, *loff2 = nullptr;
  long long* lscan = nullptr;
  int* pool = nullptr;
  int64_t pool_cap = 0, pool_used = 0;
  int* newidx = nullptr;
  int* part32 = nullptr;
  long long* part64 = nullptr;
  size_t bytes_total = 0;

  int init(int64_t cap_, int64_t nwall_, int nc_, hipStream_t st);
  void release();
  // counters after an upload: n live particles, no lists
  int reset(int64_t n);
  int set_pool_used(int64_t used);
  template <int D>
  int build_bins(const Geo& g, const PtlDev& P, const WallDev& W, int64_t n);
  int edit_bins();
  template <int D>
  int build_lists(const Geo& g, const double* xq, const double* hq, const double* xall,
                  const WallDev& W, int status);
  template <int D>
  int mark_in(const Geo& g, const PtlDev& P, const long long* gate = nullptr);
  template <int D>
  int split(const Geo& g, ParticleArrays& PA, const WallDev& W);
  template <int D>
  int merge(const Geo& g, ParticleArrays& PA, const WallDev& W);
  // Waits for the passes, reports their error, compacts the dead particles
  // and returns the live count and the step's event counts.
  int finish(ParticleArrays& PA, int64_t* n_out);
  int reset_step_counts();
  // fluid particles of every cell's bin into fcnt (device)
  int fluid_counts();
  template <int D>
  int end_of_step_lists(const Geo& g, const PtlDev& P, const double* xb, const double* hb,
                        const WallDev& W, int64_t n);
};

}  // namespace pgdg
