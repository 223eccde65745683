// step_kernels.hpp - launcher declarations shared by the step translation
// units (cells, particles, topology) and the context.
#pragma once
#include "dev_common.hpp"
#include "mgfm_pairs.hpp"

namespace pgdg {

enum { CELL_P2G = 0, CELL_LIMITER = 1, CELL_RHS = 2, CELL_UPDATE = 3 };
// CELL_RHS arg: -1, or (mixed form) a stage code whose coefficient update
// the RHS kernel applies itself; CELL_UPDATE arg: the stage code, plus
// kCellSkipCoef after such an RHS launch.
constexpr int kCellSkipCoef = 0x100;

// Workgroup sizes of the tuned kernels (64, 128 or 256).
struct Tuning {
  int pst = 256;
  int rhs = 128;
  int limiter = 128;
};

// Mixed form: interior-face point fluxes into C.fflux (before k_rhs<float>);
// ib (may be nullptr: IFACE_INLINE) selects how the interface points' MGFM
// is solved.
template <int D>
int launch_faces(const Geo& g, const CellDev& C, const int32_t* used, hipStream_t st,
                 const IfaceBatch* ib = nullptr);

template <int D>
int launch_cells(int which, const Geo& g, const CellDev& C, const PtlDev& P, const int* bin_off,
                 const int* bin_ent, const int32_t* used, const DevScalars* sc, int arg,
                 hipStream_t st, int block = 256, const IfaceBatch* ib = nullptr);

// Particle data MLS_P2G's kernel weights need (position, volume, phase;
// phase -1 for walls), gathered into bin-entry order.  The state and slopes
// (needed only in cells where a phase appears) are read from the particle
// arrays.
struct P2gBuf {
  double* light;   // [ne][kP2gLight]
  int64_t ne;
  double* mass;    // [PGDG_MAX_PHASE][nc]: phase masses of the flagged cells
  int32_t* flag;   // [nc]: 1 where a phase appears
};
constexpr int kP2gLight = 6;  // x[3] (or x[2], 0), vol, phase, pad

template <int D>
int launch_p2g(const Geo& g, const CellDev& C, const PtlDev& P, const int* bin_off, const int* bin_ent,
               int64_t n_ent, const int32_t* used, const P2gBuf& pb, hipStream_t st);

// Per-particle stage work.
struct StageBufs {
  double *xb, *hb;    // positions / smoothing lengths when the stage's
                      // neighbour lists are built ([3][cap], [cap])
  double *x0, *vol0;  // stage-0 start position / volume
  int32_t* in_ele;    // cell of each particle (membership)
};

template <int D>
int launch_stage_save(const Geo& g, const PtlDev& P, const StageBufs& S, int64_t n, int md,
                      const DevScalars* sc, hipStream_t st);
template <int D>
int launch_in_ele(const Geo& g, const PtlDev& P, const StageBufs& S, int64_t n, hipStream_t st);
template <int D>
int launch_g2p(const Geo& g, const PtlDev& P, const CellDev& C, const StageBufs& S, int64_t n,
               const DevScalars* sc, int md, bool copy, hipStream_t st);
// Neighbour data in cell-bin order (entry e of the bins): positions after
// and before the stage update, volume, density, sound speed, phase (10 for
// walls) and the particle reference; plus the bounding box of the stage-start
// positions of every cell.
struct BinData {
  double *x, *xb;        // [3][ne]
  double *vol, *rho, *c; // [ne]
  int32_t *ph, *ref;     // [ne]
  double *bmin, *bmax;   // [3][nc]
  int64_t ne_cap;
};

template <int D>
int launch_pst_binned(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                      const BinData& B, int64_t n_ent, const int* bin_off, const int* bin_ent,
                      const DevScalars* sc, hipStream_t st, int block, int parts = 3);
template <int D>
int launch_pst_kernel(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                      const BinData& B, int64_t n_ent, const int* bin_off, const DevScalars* sc,
                      hipStream_t st, int block);
// Neighbour-list sizes per particle (the status-2 list of the last stage).
template <int D>
int launch_neighbor_counts(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                           const BinData& B, int64_t n_ent, const int* bin_off, const int* bin_ent,
                           const DevScalars* sc, hipStream_t st, int32_t* cnt);

// Mixed-precision particle shifting: FP32 records in bin-entry order with
// positions relative to the lower corner of the entry's bin cell.
//   r1 = (stage-start position, phase bits), r2 = (current position, vol),
//   r3 = (rho, c); ecell = bin cell of the entry; box[2c], box[2c+1] = lower
//   and upper corner of the box of cell c's stage-start positions in the
//   cell's own frame (x, y, z; w unused).
struct MixedBins {
  float4 *r1, *r2;  // r1: stage-start position (cell frame) and its squared norm; r2: position, volume
  float2* r3;       // density, sound speed
  int32_t* ph;      // phase (10 for walls, -1 removed)
  int32_t* ecell;
  float4* box;   // [nc][2]: lower and upper corner of the entries' stage-start positions (cell frame)
  float4* ubox;  // [nc][2]: x, y union of the boxes of the cells within the search radius along the
                 // cell's line (k +- radius); z: the least lower and the largest upper z of those
                 // boxes, each in its own cell's frame
  int64_t ne_cap;
};

template <int D>
int launch_pst_mixed(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                     const BinData& B, const MixedBins& MB, int64_t n_ent, const int* bin_off,
                     const int* bin_ent, const DevScalars* sc, hipStream_t st, int block, int staging,
                     int parts = 3, bool bins_unchanged = false);
template <int D>
int launch_pst_mixed_kernel(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                            const BinData& B, const MixedBins& MB, int64_t n_ent, const int* bin_off,
                            const DevScalars* sc, hipStream_t st, int block, int staging);

template <int D>
int launch_timestep(const Geo& g, const PtlDev& P, const CellDev& C, const int32_t* used,
                    int64_t n, bool cells, DevScalars* sc, double output_dt, hipStream_t st);

}  // namespace pgdg
