// Drives the PG-DG step through the C++ host class pgdg::MPM with the
// reference's call sequence (Solver::MPM_Solver's loop body,
// S3:solver.cpp:33-53: Set_timestep, iterate_num++, MPM_run_simulation) on
// the 2-D triple point at one particle per cell (T2:mpm_ini.cpp:58-243),
// and checks that every particle and cell array equals a second context
// stepped by pgdg_step; then one step stage by stage through the
// reference-named members (MPM_run_simulation's own sequence,
// S3:mpm_ini.cpp:682-766) against pgdg_run_simulation, and one MGFM call
// against pgdg_mgfm_batch.  Prints "mpm host ok" and exits 0 on success.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "pgdg_mpm.hpp"

namespace {

struct State {
  std::vector<double> pd, cd, wd;
  std::vector<int32_t> pi, ci;
  pgdg_particles P{};
  pgdg_walls W{};
  pgdg_cells C{};
  pgdg_scalars S{};
};

const int U = 4, M = 3, F = 2;

void shape(State& s, int64_t cap, int64_t nw, int64_t nc) {
  s.pd.assign((size_t)(19 + U * M) * cap, 0.0);
  s.pi.assign((size_t)2 * cap, 0);
  double* q = s.pd.data();
  auto take = [&](int64_t rows, int64_t len) {
    double* r = q;
    q += rows * len;
    return r;
  };
  s.P.cap = cap;
  s.P.x = take(3, cap), s.P.v = take(3, cap), s.P.tv = take(3, cap), s.P.rhov = take(3, cap);
  s.P.vol = take(1, cap), s.P.mass = take(1, cap), s.P.h = take(1, cap), s.P.P = take(1, cap);
  s.P.rho = take(1, cap), s.P.E = take(1, cap), s.P.c = take(1, cap);
  s.P.coef = take(U * M, cap);
  s.P.phase = s.pi.data();
  s.P.mark_in = s.pi.data() + cap;
  s.cd.assign((size_t)(3 * F * U * M + F * U + 8 * F + 16) * nc, 0.0);
  s.ci.assign((size_t)2 * F * nc, 0);
  q = s.cd.data();
  s.C.coef = take(F * U * M, nc), s.C.coef0 = take(F * U * M, nc), s.C.dcoef = take(F * U * M, nc);
  s.C.lim = take(F * U, nc);
  s.C.alph = take(F, nc), s.C.rho = take(F, nc), s.C.E = take(F, nc), s.C.c = take(F, nc);
  s.C.ac = take(F, nc), s.C.rhov = take(3 * F, nc);
  s.C.v = take(3, nc), s.C.a = take(3, nc), s.C.P = take(1, nc), s.C.G_v = take(9, nc);
  s.C.ph_state = s.ci.data();
  s.C.ph_state_b = s.ci.data() + F * nc;
  s.wd.assign((size_t)7 * (nw > 0 ? nw : 1), 0.0);
  s.W.n = nw;
  s.W.x = s.wd.data();
  s.W.norm = s.wd.data() + 3 * nw;
  s.W.vol = s.wd.data() + 6 * nw;
}

// Triple point on [0,7]x[0,3]: phase 1 (gamma 1.5) left x <= 1 (rho 1,
// p 1) and bottom-right (rho 1, p 0.1), phase 2 (gamma 1.4) top-right
// (rho 0.125, p 0.1); one particle per cell at the centre; walls on the
// boundary nodes of the particle lattice (T2:mpm_ini.cpp:82-92, 149-243,
// 385-409).
pgdg_case make_case(int nx, int ny, State& s) {
  pgdg_case c{};
  c.dim = 2, c.nphase = 2, c.order = 1, c.rk_stages = 2, c.semantics = PGDG_SEM_T2;
  c.n[0] = nx, c.n[1] = ny, c.n[2] = 1;
  for (int f = 0; f < 6; f++) c.bc[f] = PGDG_BC_WALL;
  c.box_r[0] = 7.0, c.box_r[1] = 3.0, c.box_r[2] = 1.0;
  c.p_g = 3.0 / ny;
  c.p_sz = sqrt((7.0 / nx) * (3.0 / ny));
  c.gamma[0] = 1.5, c.gamma[1] = 1.4;
  c.ini_rho[0] = 1.0, c.ini_rho[1] = 0.125;
  c.p2g_cut = 1.5 * c.p_sz;
  c.output_dt = 1e30;
  c.max_v0 = 1e-7;
  const int64_t n = (int64_t)nx * ny;
  // wall particles on the nodes of the (nx + 1) x (ny + 1) lattice boundary
  std::vector<double> wx, wy, wnx, wny;
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j <= ny; j++) {
      if (!(i == 0 || i == nx || j == 0 || j == ny)) continue;
      wx.push_back(7.0 / nx * i);
      wy.push_back(3.0 / ny * j);
      double a = 0, b = 0;
      if (i == 0) a = -1; else if (i == nx) a = 1; else if (j == 0) b = -1; else b = 1;
      wnx.push_back(a);
      wny.push_back(b);
    }
  const int64_t nw = (int64_t)wx.size();
  shape(s, 3 * n + 64, nw, n);
  s.P.n = n;
  const int64_t cap = s.P.cap;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      const int64_t k = (int64_t)i * ny + j;
      const double x = (i + 0.5) * 7.0 / nx, y = (j + 0.5) * 3.0 / ny;
      const bool right = x > 1, top = right && !(y < 1.5);
      const int ph = top ? 2 : 1;
      const double rho = top ? 0.125 : 1.0, p = right ? 0.1 : 1.0, g = c.gamma[ph - 1];
      const double vol = c.p_sz * c.p_sz;
      s.P.x[k] = x;
      s.P.x[cap + k] = y;
      s.P.vol[k] = vol;
      s.P.mass[k] = vol * rho;
      s.P.h[k] = 3.2 / 2.0 * c.p_sz;
      s.P.phase[k] = ph;
      s.P.rho[k] = rho;
      s.P.P[k] = p;
      s.P.E[k] = p / (g - 1);
      s.P.c[k] = sqrt(g * p / rho);
    }
  for (int64_t w = 0; w < nw; w++) {
    s.W.x[w] = wx[w];
    s.W.x[nw + w] = wy[w];
    s.W.norm[w] = wnx[w];
    s.W.norm[nw + w] = wny[w];
    s.W.vol[w] = c.p_sz;
  }
  for (int f = 0; f < F; f++)
    for (int64_t q = 0; q < n; q++) {
      s.C.rho[f * n + q] = c.ini_rho[f];
      for (int u = 0; u < U; u++) s.C.lim[(f * U + u) * n + q] = 1.0;
      s.C.ph_state[f * n + q] = 1;
    }
  s.S.max_v = c.max_v0;
  return c;
}

int compare(const State& a, const State& b, int64_t nc, const char* what) {
  if (a.P.n != b.P.n) {
    fprintf(stderr, "%s: particle counts %lld vs %lld\n", what, (long long)a.P.n, (long long)b.P.n);
    return 1;
  }
  if (memcmp(a.pd.data(), b.pd.data(), sizeof(double) * a.pd.size()) != 0 ||
      memcmp(a.pi.data(), b.pi.data(), sizeof(int32_t) * a.pi.size()) != 0 ||
      memcmp(a.cd.data(), b.cd.data(), sizeof(double) * a.cd.size()) != 0 ||
      memcmp(a.ci.data(), b.ci.data(), sizeof(int32_t) * a.ci.size()) != 0) {
    fprintf(stderr, "%s: states differ\n", what);
    return 1;
  }
  (void)nc;
  return 0;
}

void clear(State& s) {
  std::fill(s.pd.begin(), s.pd.end(), 0.0);
  std::fill(s.pi.begin(), s.pi.end(), 0);
  std::fill(s.cd.begin(), s.cd.end(), 0.0);
  std::fill(s.ci.begin(), s.ci.end(), 0);
}

}  // namespace

int main() {
  int ndev = 0;
  if (pgdg_device_count(&ndev) != PGDG_OK || ndev == 0) {
    printf("no device\n");
    return 2;
  }
  const int nx = 56, ny = 24, steps = 12;
  State init;
  pgdg_case c = make_case(nx, ny, init);
  const int64_t nc = (int64_t)nx * ny;
  try {
    pgdg::communicator world;
    pgdg::MPM mpm(c, init.P.cap, init.W.n);
    mpm.Initialize_case(init.P, init.W, init.C, init.S);
    pgdg_ctx* ref = nullptr;
    pgdg::check(pgdg_create(&c, init.P.cap, init.W.n, &ref), "pgdg_create");
    pgdg::check(pgdg_upload_state(ref, &init.P, &init.W, &init.C, &init.S), "pgdg_upload_state");
    State a, b;
    make_case(nx, ny, a);
    make_case(nx, ny, b);
    for (int k = 0; k < steps; k++) {
      // Solver::MPM_Solver's loop body (output_dt is never reached here)
      mpm.Set_timestep(world);
      mpm.iterate_num++;
      mpm.MPM_run_simulation((int)mpm.iterate_num, world, 0.0, 0);
      pgdg::check(pgdg_step(ref, 1), "pgdg_step");
      clear(a);
      clear(b);
      mpm.download(&a.P, &a.C, &a.S);
      pgdg::check(pgdg_download_state(ref, &b.P, &b.C, &b.S), "pgdg_download_state");
      if (compare(a, b, nc, "MPM_Solver loop vs pgdg_step")) return 1;
    }
    // one step through the reference-named stages (the reference's own
    // sequence of MPM_run_simulation) against pgdg_run_simulation
    mpm.Set_timestep(world);
    pgdg::check(pgdg_set_timestep(ref), "pgdg_set_timestep");
    mpm.iterate_num++;
    mpm.Reset_cell_list_info(world);
    mpm.Map_the_particle_to_tree(world);
    mpm.Split_particle_fast(world);
    for (int it = 0; it < 4; it++) mpm.Merge_particle_fast(world);
    mpm.Reset_particle_neighbor_info(world, 4);
    mpm.Refresh_neighbor_info(4, world);
    for (int md = 0; md < 2; md++) {
      mpm.Reset_particle_neighbor_info(world, 2);
      mpm.Refresh_neighbor_info(2, world);
      if (md == 0) mpm.MLS_P2G(world);
      mpm.Calculate_limiter(world);
      mpm.Grid_acceleration(world);
      mpm.Update_grid(world, md);
      mpm.Update_particle(world, md);
    }
    pgdg::check(pgdg_run_simulation(ref, mpm.iterate_num), "pgdg_run_simulation");
    clear(a);
    clear(b);
    mpm.download(&a.P, &a.C, &a.S);
    pgdg::check(pgdg_download_state(ref, &b.P, &b.C, &b.S), "pgdg_download_state");
    if (compare(a, b, nc, "reference-named stages vs pgdg_run_simulation")) return 1;
    // MGFM of one face against the batched entry point
    std::array<double, 3> r = mpm.MGFM(0.1, -0.2, 1.0, 0.3, 1.0, 0.2, 0, 1);
    double in[6] = {0.1, -0.2, 1.0, 0.3, 1.0, 0.2}, out[3];
    void* d = nullptr;
    pgdg::check(pgdg_malloc(&d, sizeof(double) * 9), "pgdg_malloc");
    pgdg::check(pgdg_memcpy(d, in, sizeof(in), 0), "pgdg_memcpy");
    pgdg_eos_pair eos{c.gamma[0], c.pinf[0], c.gamma[1], c.pinf[1]};
    pgdg_mgfm_opts o{PGDG_MGFM_T2, PGDG_TOL_REFERENCE, 0.0, 50};
    pgdg::check(pgdg_mgfm_batch((double*)d, (double*)d + 6, nullptr, 1, &eos, &o, nullptr, nullptr),
                "pgdg_mgfm_batch");
    pgdg::check(pgdg_memcpy(out, (double*)d + 6, sizeof(out), 1), "pgdg_memcpy");
    pgdg_free(d);
    if (r[0] != out[0] || r[1] != out[1] || r[2] != out[2]) {
      fprintf(stderr, "MGFM differs\n");
      return 1;
    }
    // EOS round trip of the class's host routines
    const double e = mpm.EOS_rhoe(0.7, 1);
    if (fabs(mpm.EOS_P(e, 1) - 0.7) > 1e-15) {
      fprintf(stderr, "EOS round trip\n");
      return 1;
    }
    pgdg_destroy(ref);
  } catch (const pgdg::Error& e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  printf("mpm host ok\n");
  return 0;
}
