// step_topo.hpp - device particle arrays and the topology object (cell bins,
// neighbour-list pool, split/merge scratch).
#pragma once
#include "dev_common.hpp"
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

struct Topo {
  int64_t cap = 0, nwall = 0;
  int nc = 0;
  hipStream_t stream = nullptr;
  int key_bits = 1;
  // cell bins (CSR)
  int* bin_off = nullptr;
  int* bin_ent = nullptr;
  int* bin_ent2 = nullptr;
  int64_t n_ent = 0;
  uint32_t *keys = nullptr, *keys2 = nullptr;
  int* vals = nullptr;
  int *cnt = nullptr, *cnt2 = nullptr, *off2 = nullptr;
  int* fcnt = nullptr;  // fluid entries per cell (per-cell split/merge bounds)
  uint8_t* win = nullptr;
  // per-particle scratch
  uint8_t *rm = nullptr, *removed = nullptr, *src = nullptr, *aff = nullptr, *need_ = nullptr,
          *type = nullptr;
  int *ia = nullptr, *ib = nullptr, *ic = nullptr, *id = nullptr;
  int *sa = nullptr, *sb = nullptr, *sc_ = nullptr, *sd = nullptr;
  int *app_ref = nullptr, *app_ref2 = nullptr;
  uint32_t *app_cell = nullptr, *app_cell2 = nullptr;
  // neighbour-list pool
  int *lcnt = nullptr, *lcnt2 = nullptr;
  int64_t *loff = nullptr, *loff2 = nullptr, *lscan = nullptr, *lwide = nullptr;
  int* pool = nullptr;
  int64_t pool_cap = 0, pool_used = 0;
  int* newidx = nullptr;
  int64_t* host_small = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  size_t bytes_total = 0;

  int init(int64_t cap_, int64_t nwall_, int nc_, hipStream_t st);
  void release();
  int scan_i32(const int* in, int* out, int64_t n);
  int scan_total(const int* in, int* out, int64_t n, int64_t* total);
  template <int D>
  int build_bins(const Geo& g, const PtlDev& P, const WallDev& W, int64_t n);
  int edit_bins(int64_t k_app);
  template <int D>
  int build_lists(const Geo& g, const double* xq, const double* hq, const double* xall,
                  const WallDev& W, int64_t n, int status, bool reset_pool);
  int compact(ParticleArrays& PA, int64_t n, int64_t* n_out);
  template <int D>
  int mark_in(const Geo& g, const PtlDev& P, int64_t n);
  template <int D>
  int split(const Geo& g, ParticleArrays& PA, const WallDev& W, int64_t* n_io, int64_t counts[3]);
  template <int D>
  int merge(const Geo& g, ParticleArrays& PA, const WallDev& W, int64_t* n_io, int64_t* merged);
  template <int D>
  int end_of_step_lists(const Geo& g, const PtlDev& P, const double* xb, const double* hb,
                        const WallDev& W, int64_t n);
};

// small helpers used by step_topo.hip
__global__ void k_widen(const int* in, int64_t n, int64_t* out);
__global__ void k_keep(const uint8_t* removed, int64_t n, int* keep);
__global__ void k_split_src(const uint8_t* type, int64_t n, uint8_t* src);
__global__ void k_mark_refs(const int* refs, int64_t k, uint8_t* flag);
template <int D>
__global__ void k_flag_lists(Geo g, PtlDev P, int64_t n, uint8_t* need);

}  // namespace pgdg
