// mgfm_pairs.hpp - MGFM over faces with per-face EOS pairs and an optional
// learned initial guess (kernels in mlp.hip), shared with the time step's
// interface faces.
#pragma once
#include "dev_common.hpp"

namespace pgdg {

// Pair code c selects the left medium l[c] (the reference's `s`) and the
// right medium r[c] (`s1`) of a face; fx[c] holds the standardised EOS
// features 6..9 of the network's input.
constexpr int kMaxPairs = PGDG_MAX_PHASE * PGDG_MAX_PHASE;
struct PairTab {
  Eos l[kMaxPairs], r[kMaxPairs];
  float fx[kMaxPairs][4];
};

// Faces i < n of SoA arrays with leading dimension ld (in[6][ld]: ul, ur,
// pl, pr, rhol, rhor; out[3][ld]: rho*_L, p*, u*_L), pair[i] into tab
// (nullptr: pair 0 for every face).  count (device, may be nullptr): n is
// min(*count, ld), read when the kernel runs, and the grid covers ld.
// P == nullptr: the variant's own initial p*; otherwise the network's
// guess (parameters P, normalisation nm).
int launch_mgfm_pairs(const float* P, const pgdg_mlp_norm& nm, const PairTab& tab, const double* in,
                      const int32_t* pair, double* out, int32_t* iters, int64_t n, int64_t ld,
                      const int32_t* count, int variant, int tol_mode, double rtol, int max_iter,
                      hipStream_t st);

// The network's parameters (device) and normalisation; PGDG_ERR_STATE if it
// has not been trained or set.
int mlp_view(pgdg_mlp* m, const float** params, pgdg_mlp_norm* nm);

// The pair table of np EOS pairs with features standardised by nm.
PairTab make_pair_tab(const pgdg_mlp_norm& nm, const pgdg_eos_pair* eos, int np);

// How the mixed step solves MGFM at its interface face points.
enum { IFACE_INLINE = 0, IFACE_BATCH = 1, IFACE_LEARNED = 2 };
struct IfaceBatch {
  int mode = IFACE_INLINE;  // INLINE: in the face kernel; BATCH: as one batch
                            // from the reference's initial p*; LEARNED: as one
                            // batch from the network's p*
  const float* P = nullptr;
  pgdg_mlp_norm nm{};
  PairTab tab{};  // pair code ss * PGDG_MAX_PHASE + s1
};

}  // namespace pgdg
