/*
 * pgdg_mpm.hpp - C++ host class with the reference's member-function names
 * and argument lists, forwarding to the C ABI of pgdg.h.
 *
 * The reference's hot path is the set of public members of `class MPM`
 * (S3 = 3D_generalized/PGDG_3D_spheric, S3:mpm.h:160-194) driven by
 * `Solver::MPM_Solver` (S3:solver.h:16-17, S3:solver.cpp:30-64).  A caller of
 * the reference can keep its call sequence: each method below replaces the
 * member of the same name and forwards to one pgdg_* entry point.
 *
 * Differences a caller sees:
 *   - `communicator` is a single-rank type: one process drives one GPU, the
 *     reference's two all_reduce calls of Set_timestep are device reductions;
 *   - the case is described by pgdg_case and its initial state is built on
 *     the host by the caller (the reference hard-codes both in
 *     Initialize_case, S3:mpm_ini.cpp:33-673);
 *   - a failing entry point throws pgdg::Error carrying the PGDG_ERR_* code
 *     (the reference has no error reporting; the C ABI never throws);
 *   - the particle and cell state stays on the GPU; download() copies it to
 *     host arrays in the pgdg.h layouts.
 * Header-only; link with libpgdg.so.
 */
#ifndef PGDG_MPM_HPP
#define PGDG_MPM_HPP

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgdg.h"

namespace pgdg {

typedef double Real;

// Single-rank stand-in for boost::mpi::communicator (S3:main.cpp:19-20).
class communicator {
 public:
  int rank() const { return 0; }
  int size() const { return 1; }
};

// Run times of Initialization (S3:system.cpp:12-28).
struct Initialization {
  Real end_time = 0;
  Real output_dt = 0;
};

class Error : public std::runtime_error {
 public:
  Error(const std::string& what, int code) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char* what) {
  if (rc != PGDG_OK) throw Error(std::string(what) + " failed with status " + std::to_string(rc), rc);
}

class MPM {
 public:
  // MPM::MPM (S3:mpm.h:162): device state for the case, with room for
  // particle_cap fluid and n_wall wall particles.
  MPM(const pgdg_case& c, int64_t particle_cap, int64_t n_wall) : case_(c) {
    check(pgdg_create(&c, particle_cap, n_wall, &ctx_), "pgdg_create");
  }
  ~MPM() {
    if (mg_) pgdg_free(mg_);
    if (ctx_) pgdg_destroy(ctx_);
  }
  MPM(const MPM&) = delete;
  MPM& operator=(const MPM&) = delete;

  // MPM::Initialize_case (S3:mpm_ini.cpp:33): the caller's initial state.
  void Initialize_case(const pgdg_particles& p, const pgdg_walls& w, const pgdg_cells& c,
                       const pgdg_scalars& s) {
    pgdg_particles pp = p;
    pgdg_walls ww = w;
    pgdg_cells cc = c;
    pgdg_scalars ss = s;
    check(pgdg_upload_state(ctx_, &pp, &ww, &cc, &ss), "pgdg_upload_state");
    iterate_num = s.iterate_num;
  }

  // MPM::MPM_run_simulation (S3:mpm_ini.cpp:682): step number n (the
  // caller increments iterate_num first, as Solver::MPM_Solver does).  The
  // reference passes the run time and, on the last step of an output
  // interval, flag1 = 1 for Output_point; call Output_point for that.
  void MPM_run_simulation(int n, communicator&, Real, int) {
    check(pgdg_run_simulation(ctx_, n), "pgdg_run_simulation");
  }

  void Set_timestep(communicator&) { check(pgdg_set_timestep(ctx_), "pgdg_set_timestep"); }
  // Reset_cell_list_info clears the cell lists that Map_the_particle_to_tree
  // refills (S3:mpm.cpp:65-96); pgdg_map_particles_to_cells does both.
  void Reset_cell_list_info(communicator&) {}
  void Map_the_particle_to_tree(communicator&) {
    check(pgdg_map_particles_to_cells(ctx_), "pgdg_map_particles_to_cells");
  }
  // Reset_particle_neighbor_info clears what Refresh_neighbor_info rebuilds
  // (S3:mpm.cpp:101-161); the refresh entry points rebuild from scratch.
  void Reset_particle_neighbor_info(communicator&, int) {}
  // flag 4: cell membership (S3:mpm.cpp:137-161); flag 2: particle
  // neighbour lists (S3:level_infor.cpp:153-195).
  void Refresh_neighbor_info(int flag, communicator&) {
    if (flag == 4)
      check(pgdg_refresh_cell_membership(ctx_), "pgdg_refresh_cell_membership");
    else if (flag == 2)
      check(pgdg_refresh_neighbor_info(ctx_, 0), "pgdg_refresh_neighbor_info");
    else
      throw Error("Refresh_neighbor_info: flag must be 2 or 4", PGDG_ERR_ARG);
  }
  void Used_state_update(communicator&) { check(pgdg_used_state_update(ctx_), "pgdg_used_state_update"); }
  void MLS_P2G(communicator&) { check(pgdg_mls_p2g(ctx_), "pgdg_mls_p2g"); }
  void Calculate_limiter(communicator&) { check(pgdg_calculate_limiter(ctx_), "pgdg_calculate_limiter"); }
  void Grid_acceleration(communicator&) { check(pgdg_grid_acceleration(ctx_), "pgdg_grid_acceleration"); }
  void Update_particle(communicator&, int md) { check(pgdg_update_particle(ctx_, md), "pgdg_update_particle"); }
  void Update_grid(communicator&, int md) { check(pgdg_update_grid(ctx_, md), "pgdg_update_grid"); }
  void Split_particle_fast(communicator&) {
    check(pgdg_split_particle_fast(ctx_), "pgdg_split_particle_fast");
  }
  void Merge_particle_fast(communicator&) {
    check(pgdg_merge_particle_fast(ctx_), "pgdg_merge_particle_fast");
  }

  // MPM::EOS_P / EOS_rhoe (S3:mpm.cpp:859-876): stiffened gas of phase s.
  Real EOS_P(Real ee, int s) const { return ee * (case_.gamma[s] - 1) - case_.gamma[s] * case_.pinf[s]; }
  Real EOS_rhoe(Real p1, int s) const { return (p1 + case_.gamma[s] * case_.pinf[s]) / (case_.gamma[s] - 1); }

  // MPM::MGFM (S3:mpm.cpp:882-926): (rho*_L, p*, u*_L) of one face, media
  // s and s1; the variant and tolerance follow the case's semantics.
  std::array<Real, 3> MGFM(Real ul, Real ur, Real pl, Real pr, Real rhol, Real rhor, int s, int s1) {
    if (!mg_) check(pgdg_malloc(&mg_, sizeof(double) * 9), "pgdg_malloc");
    double in[6] = {ul, ur, pl, pr, rhol, rhor};
    double out[3];
    pgdg_eos_pair eos{case_.gamma[s], case_.pinf[s], case_.gamma[s1], case_.pinf[s1]};
    pgdg_mgfm_opts o{case_.semantics == PGDG_SEM_S3 ? PGDG_MGFM_S3 : PGDG_MGFM_T2, PGDG_TOL_REFERENCE, 0.0, 50};
    double* d = static_cast<double*>(mg_);
    check(pgdg_memcpy(d, in, sizeof(in), 0), "pgdg_memcpy");
    check(pgdg_mgfm_batch(d, d + 6, nullptr, 1, &eos, &o, nullptr, nullptr), "pgdg_mgfm_batch");
    check(pgdg_memcpy(out, d + 6, sizeof(out), 1), "pgdg_memcpy");
    return {out[0], out[1], out[2]};
  }

  // MPM::Output_plt_file (S3:mpm_ini.cpp:832-980): particle and cell files
  // of output n into dir (the reference's outdata/).
  void Output_plt_file(int n, int, communicator& world, const std::string& dir = "outdata", Real time = -1) {
    HostState h;
    fetch(h);
    std::vector<int32_t> nb((size_t)(h.P.n > 0 ? h.P.n : 1));
    check(pgdg_neighbor_counts(ctx_, nb.data()), "pgdg_neighbor_counts");
    const std::string tag = "." + std::to_string(n) + "." + std::to_string(world.rank()) + ".plt";
    const Real t = time >= 0 ? time : h.S.run_time;
    check(pgdg_plt_write_particles((dir + "/simulation_sp" + tag).c_str(), &case_, &h.P, nb.data(), t),
          "pgdg_plt_write_particles");
    check(pgdg_plt_write_cells((dir + "/ele_file" + tag).c_str(), &case_, &h.C, t), "pgdg_plt_write_cells");
  }
  // MPM::Output_point (S3:mpm_ini.cpp:774-827): one line of the diagnostic
  // files in dir; first != 0 truncates them.
  void Output_point(communicator&, Real time1, const std::string& dir = ".", int first = 0) {
    HostState h;
    fetch(h);
    std::vector<int32_t> used((size_t)ncells());
    check(pgdg_used_cells(ctx_, used.data()), "pgdg_used_cells");
    pgdg_point_diag d{};
    check(pgdg_point_diagnostics(&case_, &h.P, &h.C, used.data(), time1, &d), "pgdg_point_diagnostics");
    check(pgdg_point_append(dir.c_str(), &case_, &d, first), "pgdg_point_append");
  }

  // The time-loop scalars (glbl_timestep, run_time, ...) and the state.
  pgdg_scalars scalars() {
    pgdg_scalars s{};
    check(pgdg_download_state(ctx_, nullptr, nullptr, &s), "pgdg_download_state");
    return s;
  }
  void download(pgdg_particles* p, pgdg_cells* c, pgdg_scalars* s) {
    check(pgdg_download_state(ctx_, p, c, s), "pgdg_download_state");
  }
  pgdg_ctx* ctx() { return ctx_; }
  int64_t iterate_num = 0;  // the reference's mpm.iterate_num
  const pgdg_case& config() const { return case_; }

 private:
  int64_t ncells() const {
    return (int64_t)case_.n[0] * (case_.dim > 1 ? case_.n[1] : 1) * (case_.dim > 2 ? case_.n[2] : 1);
  }
  // Host arrays in the pgdg.h layouts, sized to the live state.
  struct HostState {
    std::vector<double> pd, cd;
    std::vector<int32_t> pi, ci;
    pgdg_particles P{};
    pgdg_cells C{};
    pgdg_scalars S{};
  };
  void fetch(HostState& h) {
    int64_t n = 0;
    check(pgdg_particle_count(ctx_, &n), "pgdg_particle_count");
    const int U = case_.dim + 2;
    const int M = case_.order == 1 ? case_.dim + 1 : (case_.dim == 2 ? 6 : 10);
    const int F = case_.nphase;
    const int64_t cap = n > 0 ? n : 1, nc = ncells();
    h.pd.assign((size_t)(19 + U * M) * cap, 0.0);
    h.pi.assign((size_t)2 * cap, 0);
    double* q = h.pd.data();
    auto take = [&](int64_t rows, int64_t len) {
      double* r = q;
      q += rows * len;
      return r;
    };
    h.P.n = n;
    h.P.cap = cap;
    h.P.x = take(3, cap), h.P.v = take(3, cap), h.P.tv = take(3, cap), h.P.rhov = take(3, cap);
    h.P.vol = take(1, cap), h.P.mass = take(1, cap), h.P.h = take(1, cap), h.P.P = take(1, cap);
    h.P.rho = take(1, cap), h.P.E = take(1, cap), h.P.c = take(1, cap);
    h.P.coef = take(U * M, cap);
    h.P.phase = h.pi.data();
    h.P.mark_in = h.pi.data() + cap;
    h.cd.assign((size_t)(3 * F * U * M + F * U + 8 * F + 16) * nc, 0.0);
    h.ci.assign((size_t)2 * F * nc, 0);
    q = h.cd.data();
    h.C.coef = take(F * U * M, nc), h.C.coef0 = take(F * U * M, nc), h.C.dcoef = take(F * U * M, nc);
    h.C.lim = take(F * U, nc);
    h.C.alph = take(F, nc), h.C.rho = take(F, nc), h.C.E = take(F, nc), h.C.c = take(F, nc);
    h.C.ac = take(F, nc), h.C.rhov = take(3 * F, nc);
    h.C.v = take(3, nc), h.C.a = take(3, nc), h.C.P = take(1, nc), h.C.G_v = take(9, nc);
    h.C.ph_state = h.ci.data();
    h.C.ph_state_b = h.ci.data() + F * nc;
    download(&h.P, &h.C, &h.S);
  }

  pgdg_case case_;
  pgdg_ctx* ctx_ = nullptr;
  void* mg_ = nullptr;
};

// Solver (S3:solver.h:9-17): the time loop of one output interval.
class Solver {
 public:
  Solver(Initialization&, communicator&) {}
  // Solver::MPM_Solver (S3:solver.cpp:30-64): steps until the interval's
  // integral time reaches output_dt (pgdg_step truncates the last dt onto
  // the output instant and advances the run time), then Output_point and
  // Output_plt_file(n) as the reference's main loop does.
  void MPM_Solver(Initialization& Ini, MPM& mpm, int n, communicator& world, const std::string& dir = "outdata",
                  const std::string& point_dir = ".") {
    for (;;) {
      check(pgdg_step(mpm.ctx(), 1), "pgdg_step");
      pgdg_scalars s = mpm.scalars();
      if (!(s.integral_time < Ini.output_dt - 1.0e-10)) break;
    }
    integral_time = mpm.scalars().integral_time;
    mpm.Output_point(world, run_time + integral_time, point_dir, n == 1);
    mpm.Output_plt_file(n, 1, world, dir, run_time + integral_time);
    run_time += Ini.output_dt;
  }
  Real integral_time = 0, run_time = 0;
};

}  // namespace pgdg

#endif /* PGDG_MPM_HPP */
