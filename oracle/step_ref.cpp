// step_ref.cpp - CPU FP64 restatement of one reference PG-DG step.
//
// TEST INFRASTRUCTURE ONLY: linked into oracle/liboracle.so, loaded by the
// tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker of the HIP library.  Never a compute path of the product.
//
// The code follows the reference routines in their single-thread order
// (the reference runs them under oneTBB; its concurrent_vector appends make
// the order of new particles thread-dependent, so "reference order" here is
// the order one thread would produce):
//   Set_timestep              S3:mpm.cpp:19-57, S3:particle.cpp:554-569
//   Solver::MPM_Solver        S3:solver.cpp:30-64
//   MPM_run_simulation        S3:mpm_ini.cpp:682-766
//   Used_state_update         S3:mpm.cpp:184-263
//   binning                   S3:mpm.cpp:65-96, S3:level_infor.cpp:117-147
//   Split_particle_fast       S3:mpm.cpp:1036-1417 (T2:mpm.cpp:1486-1669)
//   Merge_particle_fast       S3:mpm.cpp:1536-1706 (T2:mpm.cpp:1778-1945)
//   Refresh_neighbor_info     S3:level_infor.cpp:153-215, S3:particle.cpp:527-549
//   MLS_P2G                   S3:mpm.cpp:274-382
//   Calculate_limiter         S3:mpm.cpp:392-738 (T2:mpm.cpp:887-1225)
//   Grid_acceleration         S3:particle_flow.cpp:16-405 (T2:particle_flow.cpp:16-336)
//   Update_grid               S3:mpm.cpp:768-830 (T2:mpm.cpp:1254-1320)
//   Update_particle           S3:mpm.cpp:839-853, S3:particle_flow.cpp:413-622
// Differences from the literal reference, all shared with the HIP library:
//   * Update_particle runs in two passes (every Update_state_onestep, then
//     every Update_PST); the reference's single parallel_for lets PST read
//     neighbours that may or may not have moved yet (a data race).
//   * pow(x, 1./D) is the correctly rounded value (root_ref.hpp); pow(x,2)
//     is x*x and pow(x,4) is (x*x)*(x*x).
//   * mu = 0 in every reference case, so the viscous stress tau is dropped.
//   * Eigen reductions (norm, dot, small matrix-vector products) are
//     evaluated as left-to-right sums.
// Extensions with no reference routine (the same definitions are in the HIP
// library; each is pinned by the tests' closed-form checks):
//   * order 2: the zero-mean P2 basis of UX:glbfunc.h:210-222 (2D: 1, x, y,
//     xy, x^2 - dx^2/12, y^2 - dy^2/12), in 3D 1, x, y, z, xy, xz, yz,
//     x^2 - dx^2/12, y^2 - dy^2/12, z^2 - dz^2/12; 3-point Gauss rules; the
//     diagonal mass of UX:particle_flow.cpp:28-30 extended to 3D.  A
//     particle's coefficient copy holds the gradient at the particle (not at
//     the cell centre) and the second-order coefficients, so the particle's
//     polynomial is U_p + sum_m c_m psi_m(x - x_p) with psi_m = phi_m -
//     phi_m(0).
//   * rk_stages 3: the SSP-RK3 (Shu-Osher) stages, stage codes 0, 2, 3:
//     U1 = U0 + dt L(U0); U2 = 3/4 U0 + 1/4 (U1 + dt L(U1));
//     U = 1/3 U0 + 2/3 (U2 + dt L(U2)); particle positions and volumes the same.
//   * cfl: the time-step factor (0 selects the reference 0.3).
//   * ppc_min / ppc_max (0: off): a particle is a merge candidate only while
//     its cell holds more than ppc_min; a split (merge) is admitted only if
//     no cell it adds to (removes from) passes ppc_max (ppc_min), counting
//     the bins when the pass starts plus the net changes of every earlier
//     candidate of the pass in particle order (split_admit, merge_admit).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <unordered_set>
#include <thread>
#include <vector>

#include "pgdg.h"
#include "root_ref.hpp"

namespace oracle {

void MGFM_S3(double ul, double ur, double pl, double pr, double rhol, double rhor, double gs,
             double ps0, double gs1, double ps1, int max_iter, double guess, double out[3],
             int* iters);
void MGFM_T2(double ul, double ur, double pl, double pr, double rhol, double rhor, double gs,
             double ps0, double gs1, double ps1, int max_iter, double guess, double out[3],
             int* iters);

static const double CELL_RATIO = 3.2;
static const double CUT_OFF = 2.;
static const double PI_REF = 3.1415926;  // S3:glbparam.h:36

// Worker threads of the per-cell and per-particle loops (1: sequential).
// Every parallel loop writes only the item it visits and reads state no
// iteration of the same loop writes, so results do not depend on the count.
static int g_threads = 1;
template <class Fn>
static void par_range(int64_t n, Fn fn) {
  const int T = g_threads;
  if (T <= 1 || n < 64) {
    fn((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + T - 1) / T;
  for (int t = 0; t < T; t++) {
    const int64_t lo = t * chunk, hi = lo + chunk < n ? lo + chunk : n;
    if (lo >= hi) break;
    th.emplace_back(fn, lo, hi);
  }
  for (auto& x : th) x.join();
}

static inline double AMAX1(double a, double b) { return a >= b ? a : b; }
static inline double AMIN1(double a, double b) { return a <= b ? a : b; }
static inline int IMAX(int a, int b) { return a >= b ? a : b; }
static inline int IMIN(int a, int b) { return a <= b ? a : b; }

// Number of modes and of Gauss points per axis.
template <int D, int ORD>
struct Basis {
  static const int M = ORD == 1 ? D + 1 : (D == 2 ? 6 : 10);
  static const int NQ = ORD + 1;
};

// Modes 1..M-1 of the cell polynomial at offset r from the cell centre
// (q12[a] = dx_a^2 / 12).
template <int D, int ORD>
static inline void basis_tail(const double* r, const double* q12, double* phi) {
  for (int a = 0; a < D; a++) phi[a] = r[a];
  if (ORD == 2) {
    if (D == 2) {
      phi[2] = r[0] * r[1];
      phi[3] = r[0] * r[0] - q12[0];
      phi[4] = r[1] * r[1] - q12[1];
    } else {
      phi[3] = r[0] * r[1];
      phi[4] = r[0] * r[2];
      phi[5] = r[1] * r[2];
      phi[6] = r[0] * r[0] - q12[0];
      phi[7] = r[1] * r[1] - q12[1];
      phi[8] = r[2] * r[2] - q12[2];
    }
  }
}

// The same modes without their constants: psi_m(r) = phi_m(r) - phi_m(0).
template <int D, int ORD>
static inline void incr_tail(const double* r, double* psi) {
  const double z[3] = {0, 0, 0};
  basis_tail<D, ORD>(r, z, psi);
}

// Gradient of a polynomial with tail coefficients c[1..M-1] at offset r:
// g[a] = c[1+a] + (second-order terms at r).
template <int D, int ORD>
static inline void grad_at(const double* c, const double* r, double* g) {
  if (ORD == 1) {
    for (int a = 0; a < D; a++) g[a] = c[1 + a];
  } else if (D == 2) {
    g[0] = c[1] + c[3] * r[1] + 2. * c[4] * r[0];
    g[1] = c[2] + c[3] * r[0] + 2. * c[5] * r[1];
  } else {
    g[0] = c[1] + c[4] * r[1] + c[5] * r[2] + 2. * c[7] * r[0];
    g[1] = c[2] + c[4] * r[0] + c[6] * r[2] + 2. * c[8] * r[1];
    g[2] = c[3] + c[5] * r[0] + c[6] * r[1] + 2. * c[9] * r[2];
  }
}

// Volume-integral weight of mode m (1..M-1) given the flux columns col[b]
// (component u) at Gauss offset r: sum_b col[b][u] * d(phi_m)/dx_b.
template <int D, int ORD>
static inline double vol_weight(int m, const double (*col)[D + 2], int u, const double* r) {
  if (m <= D) return col[m - 1][u];
  if (D == 2) {
    if (m == 3) return col[0][u] * r[1] + col[1][u] * r[0];
    if (m == 4) return col[0][u] * (2. * r[0]);
    return col[1][u] * (2. * r[1]);
  }
  switch (m) {
    case 4: return col[0][u] * r[1] + col[1][u] * r[0];
    case 5: return col[0][u] * r[2] + col[2][u] * r[0];
    case 6: return col[1][u] * r[2] + col[2][u] * r[1];
    case 7: return col[0][u] * (2. * r[0]);
    case 8: return col[1][u] * (2. * r[1]);
    default: return col[2][u] * (2. * r[2]);
  }
}

struct Base {
  virtual ~Base() {}
  virtual int load(const pgdg_particles*, const pgdg_walls*, const pgdg_cells*,
                   const pgdg_scalars*) = 0;
  virtual int save(pgdg_particles*, pgdg_cells*, pgdg_scalars*) = 0;
  virtual int step() = 0;
  virtual int call(int fn, int arg) = 0;
  virtual int neighbor_counts(int32_t* out) = 0;
  virtual int used_cells(int32_t* out) = 0;
  virtual int64_t bins_out(int32_t* off, int32_t* ent, int64_t cap) = 0;
  virtual int list_out(int64_t i, int32_t* out, int cap) = 0;
};

// Reference-style references to particles: >= 0 a fluid particle object
// (stable id in the pool), <= -2 a wall particle (-(w+2)), -1 NULL.
static inline bool is_wall(int r) { return r <= -2; }
static inline int wall_of(int r) { return -r - 2; }

template <int D, int ORD>
struct Sim : Base {
  static const int U = D + 2;
  static const int M = Basis<D, ORD>::M;  // P1 (1, x, y[, z]) or P2
  static const int NQ = Basis<D, ORD>::NQ;
  static const int NF = 2 * D;

  struct Ptl {
    double x[3], v[3], tv[3], rhov[3];
    double vol, mass, h, P, rho, E, c;
    double coef[U][M];
    double coef0[U][M];
    int phase, mark_in;
    double x0[3], vol0;
    int in_ele;
    std::vector<int> nb;
    int copy_p, mrg, mrg_mark;
    void clear() {  // Particle_base::Clearup (S3:particle.cpp:61-104)
      nb.clear();
      P = mass = vol = h = 0;
      for (int a = 0; a < 3; a++) x[a] = v[a] = tv[a] = rhov[a] = 0;
      phase = 0;
      rho = E = c = 0;
      std::memset(coef, 0, sizeof(coef));
      std::memset(coef0, 0, sizeof(coef0));
    }
  };
  struct Wall {
    double x[3], norm[3], vol;
  };
  struct Face {
    double cen[3], nrm[3], tan[2][3], S;
    int nbr;  // neighbour cell index or -1
    int b;    // boundary tag
  };
  struct Cell {
    double coef[PGDG_MAX_PHASE][U][M], coef0[PGDG_MAX_PHASE][U][M], dcoef[PGDG_MAX_PHASE][U][M];
    double lim[PGDG_MAX_PHASE][U];
    double alph[PGDG_MAX_PHASE], rho[PGDG_MAX_PHASE], E[PGDG_MAX_PHASE], c[PGDG_MAX_PHASE],
        ac[PGDG_MAX_PHASE];
    double rhov[PGDG_MAX_PHASE][3];
    int ph_state[PGDG_MAX_PHASE], ph_state_b[PGDG_MAX_PHASE];
    double v[3], a[3], P;
    double G[3][3];
    double x[3];
    int idx[3];
    std::vector<int> nb;  // cell membership (fluid particles of the bin)
    int used_st;
    Face f[NF];
    double Vol;
  };

  pgdg_case cs;
  int F;
  int N[3];
  int nc;
  double ps[3];  // p_size (cell size per axis)
  double dcell[3];
  std::vector<Ptl> pool;
  std::vector<int> fluid;  // particle_fluid
  std::vector<Wall> walls;
  std::vector<Cell> cells;
  std::vector<std::vector<int>> bins;  // table_cell_list
  std::vector<int> used;               // Used_Ele_cen
  double dt = 0, max_v = 0, integral_time = 0, run_time = 0;
  int64_t iterate_num = 0;
  int64_t n_split = 0, n_inlet = 0, n_outlet = 0, n_merge = 0;
  int radius = 0;
  double Gp[3], Gw[3];
  double q12[3];  // dx_a^2 / 12
  double cfl;
  double probe_c[4], probe_s[4], sqrt2;

  explicit Sim(const pgdg_case& c) : cs(c) {
    F = cs.nphase;
    for (int a = 0; a < 3; a++) N[a] = a < D ? cs.n[a] : 1;
    nc = N[0] * N[1] * N[2];
    for (int a = 0; a < 3; a++) {
      ps[a] = a < D ? (cs.box_r[a] - cs.box_l[a]) / double(N[a]) : 1.0;
      dcell[a] = ps[a];
    }
    double g_p_rt = cs.p_g / cs.p_sz;
    radius = (int)std::ceil(CELL_RATIO / g_p_rt);
    if (ORD == 1) {
      Gp[0] = -1. / std::sqrt(3.);
      Gp[1] = 1. / std::sqrt(3.);
      Gw[0] = Gw[1] = 0.5;
      Gp[2] = Gw[2] = 0;
    } else {
      Gp[0] = -std::sqrt(3. / 5.);
      Gp[1] = 0.;
      Gp[2] = std::sqrt(3. / 5.);
      Gw[0] = Gw[2] = 5. / 18.;
      Gw[1] = 8. / 18.;
    }
    for (int a = 0; a < 3; a++) q12[a] = ps[a] * ps[a] / 12.;
    cfl = cs.cfl > 0 ? cs.cfl : 0.3;
    for (int jj = 0; jj < 4; jj++) {
      double theta = (jj + 0.5) * PI_REF * 0.5;
      probe_c[jj] = std::cos(theta);
      probe_s[jj] = std::sin(theta);
    }
    sqrt2 = std::sqrt(2.);
    build_cells();
  }

  bool s3() const { return cs.semantics == PGDG_SEM_S3; }
  bool ux() const { return cs.semantics == PGDG_SEM_UX; }
  int cidx(int i, int j, int k) const { return (i * N[1] + j) * N[2] + k; }
  double root(double x) const { return root_dim(D, x); }

  // ---- geometry: S3:mpm_ini.cpp:299-550, T2:mpm_ini.cpp:262-379 --------
  void build_cells() {
    cells.assign(nc, Cell());
    for (int i = 0; i < N[0]; i++)
      for (int j = 0; j < N[1]; j++)
        for (int k = 0; k < N[2]; k++) {
          Cell& C = cells[cidx(i, j, k)];
          std::memset(C.x, 0, sizeof(C.x));
          C.x[0] = ps[0] * (i + 0.5);
          C.x[1] = ps[1] * (j + 0.5);
          if (D == 3) C.x[2] = ps[2] * (k + 0.5);
          C.idx[0] = i;
          C.idx[1] = j;
          C.idx[2] = k;
          double px = ps[0], py = ps[1], pz = ps[2];
          // side order yL, xR, yR, xL, zL, zR
          const int di[6] = {0, 1, 0, -1, 0, 0}, dj[6] = {-1, 0, 1, 0, 0, 0},
                    dk[6] = {0, 0, 0, 0, -1, 1};
          for (int f = 0; f < NF; f++) {
            Face& fc = C.f[f];
            std::memset(&fc, 0, sizeof(fc));
            for (int a = 0; a < 3; a++) fc.cen[a] = C.x[a];
            if (f == 0) { fc.nrm[1] = -1; fc.tan[0][0] = 0.5 * px; fc.cen[1] = C.x[1] - 0.5 * py; fc.S = D == 3 ? px * pz : px; }
            if (f == 1) { fc.nrm[0] = 1; fc.tan[0][1] = 0.5 * py; fc.cen[0] = C.x[0] + 0.5 * px; fc.S = D == 3 ? py * pz : py; }
            if (f == 2) { fc.nrm[1] = 1; fc.tan[0][0] = -0.5 * px; fc.cen[1] = C.x[1] + 0.5 * py; fc.S = D == 3 ? px * pz : px; }
            if (f == 3) { fc.nrm[0] = -1; fc.tan[0][1] = -0.5 * py; fc.cen[0] = C.x[0] - 0.5 * px; fc.S = D == 3 ? py * pz : py; }
            if (f == 4) { fc.nrm[2] = -1; fc.tan[0][0] = 0.5 * px; fc.tan[1][1] = 0.5 * py; fc.cen[2] = C.x[2] - 0.5 * pz; fc.S = py * px; }
            if (f == 5) { fc.nrm[2] = 1; fc.tan[0][0] = 0.5 * px; fc.tan[1][1] = 0.5 * py; fc.cen[2] = C.x[2] + 0.5 * pz; fc.S = py * px; }
            if (D == 3 && f < 4) fc.tan[1][2] = 0.5 * pz;
            int ni = i + di[f], nj = j + dj[f], nk = k + dk[f];
            bool in = ni >= 0 && ni < N[0] && nj >= 0 && nj < N[1] && nk >= 0 && nk < N[2];
            fc.nbr = in ? cidx(ni, nj, nk) : -1;
            fc.b = in ? 0 : cs.bc[f];
          }
          C.Vol = D == 3 ? px * py * pz : px * py;
          C.used_st = 0;
        }
  }

  // ---- state exchange -----------------------------------------------------
  int load(const pgdg_particles* p, const pgdg_walls* w, const pgdg_cells* c,
           const pgdg_scalars* s) override {
    pool.clear();
    fluid.clear();
    for (int64_t i = 0; i < p->n; i++) {
      Ptl q;
      q.clear();
      int64_t cap = p->cap;
      for (int a = 0; a < 3; a++) {
        q.x[a] = p->x[a * cap + i];
        q.v[a] = p->v[a * cap + i];
        q.tv[a] = p->tv[a * cap + i];
        q.rhov[a] = p->rhov[a * cap + i];
      }
      q.vol = p->vol[i]; q.mass = p->mass[i]; q.h = p->h[i]; q.P = p->P[i];
      q.rho = p->rho[i]; q.E = p->E[i]; q.c = p->c[i];
      for (int u = 0; u < U; u++)
        for (int m = 0; m < M; m++) q.coef[u][m] = p->coef[(u * M + m) * cap + i];
      std::memcpy(q.coef0, q.coef, sizeof(q.coef));
      q.phase = p->phase[i];
      q.mark_in = p->mark_in[i];
      q.in_ele = -1;
      q.copy_p = -1; q.mrg = 0; q.mrg_mark = 0;
      std::memset(q.x0, 0, sizeof(q.x0));
      q.vol0 = 0;
      pool.push_back(q);
      fluid.push_back((int)i);
    }
    walls.resize(w->n);
    for (int64_t i = 0; i < w->n; i++) {
      for (int a = 0; a < 3; a++) {
        walls[i].x[a] = w->x[a * w->n + i];
        walls[i].norm[a] = w->norm[a * w->n + i];
      }
      walls[i].vol = w->vol[i];
    }
    for (int ci = 0; ci < nc; ci++) {
      Cell& C = cells[ci];
      for (int f = 0; f < F; f++) {
        for (int u = 0; u < U; u++) {
          for (int m = 0; m < M; m++) {
            int64_t o = ((int64_t)(f * U + u) * M + m) * nc + ci;
            C.coef[f][u][m] = c->coef[o];
            C.coef0[f][u][m] = c->coef0[o];
            C.dcoef[f][u][m] = c->dcoef[o];
          }
          C.lim[f][u] = c->lim[(int64_t)(f * U + u) * nc + ci];
        }
        C.alph[f] = c->alph[f * nc + ci];
        C.rho[f] = c->rho[f * nc + ci];
        C.E[f] = c->E[f * nc + ci];
        C.c[f] = c->c[f * nc + ci];
        C.ac[f] = c->ac[f * nc + ci];
        for (int a = 0; a < 3; a++) C.rhov[f][a] = c->rhov[(int64_t)(f * 3 + a) * nc + ci];
        C.ph_state[f] = c->ph_state[f * nc + ci];
        C.ph_state_b[f] = c->ph_state_b[f * nc + ci];
      }
      for (int a = 0; a < 3; a++) {
        C.v[a] = c->v[a * nc + ci];
        C.a[a] = c->a[a * nc + ci];
      }
      C.P = c->P[ci];
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) C.G[a][b] = c->G_v[(a * 3 + b) * nc + ci];
    }
    dt = s->dt;
    max_v = s->max_v;
    integral_time = s->integral_time;
    run_time = s->run_time;
    iterate_num = s->iterate_num;
    bins.assign(nc, std::vector<int>());
    used.clear();
    return PGDG_OK;
  }

  int save(pgdg_particles* p, pgdg_cells* c, pgdg_scalars* s) override {
    if ((int64_t)fluid.size() > p->cap) return PGDG_ERR_CAPACITY;
    int64_t cap = p->cap;
    p->n = (int64_t)fluid.size();
    for (int64_t i = 0; i < p->n; i++) {
      const Ptl& q = pool[fluid[i]];
      for (int a = 0; a < 3; a++) {
        p->x[a * cap + i] = q.x[a];
        p->v[a * cap + i] = q.v[a];
        p->tv[a * cap + i] = q.tv[a];
        p->rhov[a * cap + i] = q.rhov[a];
      }
      p->vol[i] = q.vol; p->mass[i] = q.mass; p->h[i] = q.h; p->P[i] = q.P;
      p->rho[i] = q.rho; p->E[i] = q.E; p->c[i] = q.c;
      for (int u = 0; u < U; u++)
        for (int m = 0; m < M; m++) p->coef[(u * M + m) * cap + i] = q.coef[u][m];
      p->phase[i] = q.phase;
      p->mark_in[i] = q.mark_in;
    }
    if (c) {
      for (int ci = 0; ci < nc; ci++) {
        const Cell& C = cells[ci];
        for (int f = 0; f < F; f++) {
          for (int u = 0; u < U; u++) {
            for (int m = 0; m < M; m++) {
              int64_t o = ((int64_t)(f * U + u) * M + m) * nc + ci;
              c->coef[o] = C.coef[f][u][m];
              c->coef0[o] = C.coef0[f][u][m];
              c->dcoef[o] = C.dcoef[f][u][m];
            }
            c->lim[(int64_t)(f * U + u) * nc + ci] = C.lim[f][u];
          }
          c->alph[f * nc + ci] = C.alph[f];
          c->rho[f * nc + ci] = C.rho[f];
          c->E[f * nc + ci] = C.E[f];
          c->c[f * nc + ci] = C.c[f];
          c->ac[f * nc + ci] = C.ac[f];
          for (int a = 0; a < 3; a++) c->rhov[(int64_t)(f * 3 + a) * nc + ci] = C.rhov[f][a];
          c->ph_state[f * nc + ci] = C.ph_state[f];
          c->ph_state_b[f * nc + ci] = C.ph_state_b[f];
        }
        for (int a = 0; a < 3; a++) {
          c->v[a * nc + ci] = C.v[a];
          c->a[a * nc + ci] = C.a[a];
        }
        c->P[ci] = C.P;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) c->G_v[(a * 3 + b) * nc + ci] = C.G[a][b];
      }
    }
    if (s) {
      s->dt = dt;
      s->max_v = max_v;
      s->integral_time = integral_time;
      s->run_time = run_time;
      s->iterate_num = iterate_num;
      s->n_split = n_split;
      s->n_inlet = n_inlet;
      s->n_outlet = n_outlet;
      s->n_merge = n_merge;
    }
    return PGDG_OK;
  }

  // ---- small helpers --------------------------------------------------------
  // get_cell_id (S3:glbfunc.h:161-173) on coord - box_l
  void cell_of(const double* x, int pos[3]) const {
    for (int a = 0; a < 3; a++) {
      if (a < D) {
        int p = int(std::floor((x[a] - cs.box_l[a]) / dcell[a]));
        pos[a] = IMAX(0, IMIN(p, N[a] - 1));
      } else {
        pos[a] = 0;
      }
    }
  }
  int cell_index_of(const double* x) const {
    int p[3];
    cell_of(x, p);
    return cidx(p[0], p[1], p[2]);
  }
  const double* ref_x(int r) const { return is_wall(r) ? walls[wall_of(r)].x : pool[r].x; }
  static double norm_d(const double* d) {
    double s = 0;
    for (int a = 0; a < D; a++) s += d[a] * d[a];
    return std::sqrt(s);
  }
  static double sq_d(const double* d) {
    double s = 0;
    for (int a = 0; a < D; a++) s += d[a] * d[a];
    return s;
  }
  double eos_p(double e, int s) const { return e * (cs.gamma[s] - 1) - cs.gamma[s] * cs.pinf[s]; }
  double eos_rhoe(double p, int s) const { return (p + cs.gamma[s] * cs.pinf[s]) / (cs.gamma[s] - 1); }

  // Wendland kernel and derivative (S3:particle.cpp:131-149, 297-315)
  static double kernel_C(double h) {
    if (D == 2) return 1.0 / PI_REF / (h * h) * 7. / 4.;
    return 1.0 / PI_REF / ((h * h) * h) * 21. / 16.;
  }
  static double W(double dist, double h) {
    double C = kernel_C(h);
    double s = dist / h;
    if (s < 2) {
      double q = 1. - 0.5 * s;
      return C * (((q * q) * q) * q) * (1. + 2. * s);
    }
    return 0.0;
  }
  static double dW(double dist, double h) {
    double C = kernel_C(h);
    double s = dist / h;
    if (s < 2) {
      double q = 1. - 0.5 * s;
      return -5. * C * ((q * q) * q) * s / h / dist;
    }
    return 0.0;
  }

  // ---- Set_timestep + the Solver loop body ---------------------------------
  void set_timestep() {
    double lt = 1.e20, mv = -1.e20;
    auto visit = [&](const double* v, double h, const double* c, const double* ac) {
      double Uv = norm_d(v);
      double ds = h / CELL_RATIO * CUT_OFF;
      double c12 = AMAX1(c[0], c[1]);
      double a12 = AMAX1(ac[0], ac[1]);
      double ts = cfl * ds / (c12 + Uv);
      double tmp = cfl * std::sqrt(ds / a12);
      ts = AMIN1(ts, tmp);
      lt = AMIN1(lt, ts);
      mv = AMAX1(mv, Uv);
    };
    if (iterate_num > 1) {
      double ini_h = CELL_RATIO / CUT_OFF * cs.p_sz;
      for (int ci : used) visit(cells[ci].v, ini_h, cells[ci].c, cells[ci].ac);
    } else {
      for (int id : fluid) {
        const Ptl& q = pool[id];
        double c2[2] = {q.c, 0.0}, a2[2] = {0.0, 0.0};
        visit(q.v, q.h, c2, a2);
      }
    }
    dt = lt;
    max_v = mv;
  }

  int step() override {
    if (!(integral_time < cs.output_dt - 1.0e-10)) integral_time = 0.0;
    set_timestep();
    double timestep = dt;
    if (integral_time >= (cs.output_dt - timestep)) {
      timestep = cs.output_dt - integral_time;
      dt = timestep;
    }
    integral_time += timestep;
    run_time += timestep;
    iterate_num++;
    run_simulation(iterate_num);
    return PGDG_OK;
  }

  // ---- MPM_run_simulation (S3:mpm_ini.cpp:682-766) --------------------------
  void run_simulation(int64_t n) {
    n_merge = 0;
    if (n == 1) used_state_update();
    bin_particles();
    split_fast();
    for (int it = 0; it < 4; it++) merge_fast();
    membership();
    const int codes3[3] = {0, 2, 3}, codes2[2] = {0, 1};
    const int ns = cs.rk_stages == 3 ? 3 : 2;
    for (int st = 0; st < ns; st++) {
      const int md = ns == 3 ? codes3[st] : codes2[st];
      refresh_all(2);
      if (md == 0) mls_p2g();
      calculate_limiter();
      grid_acceleration();
      update_grid(md);
      update_particle(md);
    }
  }

  // S3:mpm.cpp:184-263
  void used_state_update() {
    int Ib[3] = {5, 5, D == 3 ? 5 : 0};
    for (auto& C : cells) C.used_st = 0;
    for (int i = 0; i < N[0]; i++)
      for (int j = 0; j < N[1]; j++)
        for (int k = 0; k < N[2]; k++) {
          Cell& C = cells[cidx(i, j, k)];
          if (C.ph_state[0] == 1 || C.ph_state[1] == 1)
            for (int t = IMAX(i - Ib[0], 0); t <= IMIN(i + Ib[0], N[0] - 1); t++)
              for (int s = IMAX(j - Ib[1], 0); s <= IMIN(j + Ib[1], N[1] - 1); s++)
                for (int q = IMAX(k - Ib[2], 0); q <= IMIN(k + Ib[2], N[2] - 1); q++)
                  cells[cidx(t, s, q)].used_st = 1;
        }
    used.clear();
    for (int ci = 0; ci < nc; ci++)
      if (cells[ci].used_st == 1) used.push_back(ci);
  }

  // Reset_cell_list_info + Map_the_particle_to_tree: fluid then walls
  void bin_particles() {
    for (auto& b : bins) b.clear();
    for (int id : fluid) bins[cell_index_of(pool[id].x)].push_back(id);
    for (int w = 0; w < (int)walls.size(); w++) bins[cell_index_of(walls[w].x)].push_back(-(w + 2));
  }

  static void remove_from(std::vector<int>& list, int r) {  // Cell_list::Remove_particle
    std::vector<int> nl;
    for (int e : list)
      if (e != r) nl.push_back(e);
    list.swap(nl);
  }

  // Level_info::Refresh_neighbor_info (S3:level_infor.cpp:153-195)
  void refresh_neighbors(int id, int status) {
    Ptl& q = pool[id];
    q.nb.clear();
    int pos[3];
    cell_of(q.x, pos);
    int rr = status == 3 ? 2 : radius;
    int ic = rr, jc = rr, kc = D == 3 ? rr : 0;
    for (int t = IMAX(pos[0] - ic, 0); t <= IMIN(pos[0] + ic, N[0] - 1); t++)
      for (int s = IMAX(pos[1] - jc, 0); s <= IMIN(pos[1] + jc, N[1] - 1); s++)
        for (int m = IMAX(pos[2] - kc, 0); m <= IMIN(pos[2] + kc, N[2] - 1); m++)
          for (int r : bins[cidx(t, s, m)]) {
            const double* nx = ref_x(r);
            double dr[3];
            for (int a = 0; a < 3; a++) dr[a] = q.x[a] - nx[a];
            double dist = norm_d(dr);
            if (status == 2) {
              if (dist <= q.h * CUT_OFF) q.nb.push_back(r);
            } else {
              if (dist <= cs.p_sz * 2.) q.nb.push_back(r);
            }
          }
  }

  // ---- Split_particle_fast (S3:mpm.cpp:1036-1417) ----------------------------
  // Set_split_ptcl (S3:particle.cpp:575-609)
  void set_split(Ptl& ch, const Ptl& cp) {
    double dr[3];
    for (int a = 0; a < 3; a++) dr[a] = ch.x[a] - cp.x[a];
    ch.phase = cp.phase;
    double div = D == 3 ? 8.0 : 4.0;
    ch.vol = cp.vol / div;
    ch.mass = cp.mass / div;
    ch.h = cp.h / 2.0;
    std::memcpy(ch.coef, cp.coef, sizeof(ch.coef));
    double psi[M - 1];
    incr_tail<D, ORD>(dr, psi);
    auto slope = [&](int u) {
      double s = 0;
      for (int k = 0; k < M - 1; k++) s += cp.coef[u][1 + k] * psi[k];
      return s;
    };
    if (ORD == 2)  // the child's copy carries the gradient at the child
      for (int u = 0; u < U; u++) {
        double g[3];
        grad_at<D, ORD>(cp.coef[u], dr, g);
        for (int a = 0; a < D; a++) ch.coef[u][1 + a] = g[a];
      }
    std::memcpy(ch.coef0, ch.coef, sizeof(ch.coef));
    ch.rho = cp.rho + slope(0);
    for (int a = 0; a < D; a++) ch.rhov[a] = cp.rhov[a] + slope(1 + a);
    ch.E = cp.E + slope(U - 1);
    for (int a = 0; a < 3; a++) ch.tv[a] = ch.v[a] = ch.rhov[a] / ch.rho;
    double sq = 0;
    for (int a = 0; a < D; a++) sq += ch.v[a] * ch.v[a];
    double ei = ch.E - 0.5 * ch.rho * sq;
    int s = ch.phase - 1;
    ch.P = eos_p(ei, s);
    ch.c = std::sqrt(cs.gamma[s] * (cs.pinf[s] + ch.P) / ch.rho);
    ch.mrg_mark = 0; ch.mrg = 0;
  }
  // Copy_split_state (S3:particle.cpp:667-691): does not copy v
  static void copy_split_state(Ptl& dst, const Ptl& src) {
    dst.vol = src.vol; dst.phase = src.phase; dst.mass = src.mass; dst.h = src.h;
    for (int a = 0; a < 3; a++) dst.x[a] = src.x[a];
    std::memcpy(dst.coef, src.coef, sizeof(dst.coef));
    std::memcpy(dst.coef0, src.coef0, sizeof(dst.coef0));
    dst.rho = src.rho;
    for (int a = 0; a < D; a++) dst.rhov[a] = src.rhov[a];
    dst.E = src.E; dst.c = src.c;
    for (int a = 0; a < 3; a++) dst.tv[a] = src.tv[a];
    dst.P = src.P;
    dst.mrg_mark = 0; dst.mrg = 0;
  }
  // Set_ptcl (S3:particle.cpp:696-722)
  void set_ptcl(Ptl& dst, const Ptl& cp) {
    dst.vol = cp.vol; dst.phase = cp.phase; dst.mass = cp.mass; dst.h = cp.h;
    std::memcpy(dst.coef, cp.coef, sizeof(dst.coef));
    std::memcpy(dst.coef0, dst.coef, sizeof(dst.coef));
    dst.rho = cp.rho;
    for (int a = 0; a < D; a++) dst.rhov[a] = cp.rhov[a];
    dst.E = cp.E;
    for (int a = 0; a < 3; a++) dst.tv[a] = dst.v[a] = dst.rhov[a] / dst.rho;
    double sq = 0;
    for (int a = 0; a < D; a++) sq += dst.v[a] * dst.v[a];
    double ei = dst.E - 0.5 * dst.rho * sq;
    int s = dst.phase - 1;
    dst.P = eos_p(ei, s);
    dst.c = std::sqrt(cs.gamma[s] * (cs.pinf[s] + dst.P) / dst.rho);
    dst.mrg_mark = 0; dst.mrg = 0;
  }

  int new_obj() {
    Ptl q;
    q.clear();
    q.mark_in = 0; q.in_ele = -1; q.copy_p = -1; q.mrg = 0; q.mrg_mark = 0;
    std::memset(q.x0, 0, sizeof(q.x0));
    q.vol0 = 0;
    pool.push_back(q);
    return (int)pool.size() - 1;
  }

  void collect_window(int center_cell, std::unordered_set<int>& scanned,
                      const std::function<void(int)>& add) {
    const Cell& C = cells[center_cell];
    int pos[3] = {C.idx[0], C.idx[1], C.idx[2]};
    int kc = D == 3 ? 2 : 0;
    for (int ii = IMAX(pos[0] - 2, 0); ii <= IMIN(pos[0] + 2, N[0] - 1); ++ii)
      for (int jj = IMAX(pos[1] - 2, 0); jj <= IMIN(pos[1] + 2, N[1] - 1); ++jj)
        for (int kk = IMAX(pos[2] - kc, 0); kk <= IMIN(pos[2] + kc, N[2] - 1); ++kk) {
          int cc = cidx(ii, jj, kk);
          if (!scanned.insert(cc).second) continue;
          for (int r : bins[cc]) add(r);
        }
  }

  // Per-cell bounds (no reference counterpart).  A candidate's effect on a
  // cell is the net change of its fluid count; candidates are taken in
  // particle order and each counts the additions (or removals) of every
  // earlier candidate, admitted or not, so that the admission of one never
  // depends on another's.
  //
  // Split: the 2^D children (the parent takes the last one's place) against
  // the parent's cell; admitted if every cell it adds to stays within
  // ppc_max.
  bool split_admit(const double* xp, double dp, const std::vector<int>& fcnt, std::vector<int>& bud) const {
    int cell[9], net[9], k = 0;
    auto add = [&](int c, int d) {
      for (int t = 0; t < k; t++)
        if (cell[t] == c) {
          net[t] += d;
          return;
        }
      cell[k] = c;
      net[k++] = d;
    };
    add(cell_index_of(xp), -1);
    double pref[3] = {xp[0] - dp / 4., xp[1] - dp / 4., D == 3 ? xp[2] - dp / 4. : 0.};
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int c = 0; c < (D == 3 ? 2 : 1); ++c) {
          double x[3] = {pref[0] + dp / 2. * a, pref[1] + dp / 2. * b, 0.};
          if (D == 3) x[2] = pref[2] + dp / 2. * c;
          add(cell_index_of(x), 1);
        }
    bool ok = true;
    for (int t = 0; t < k; t++)
      if (net[t] > 0 && fcnt[cell[t]] + bud[cell[t]] + net[t] > cs.ppc_max) ok = false;
    for (int t = 0; t < k; t++)
      if (net[t] > 0) bud[cell[t]] += net[t];
    return ok;
  }
  // Merge: survivor s and partner p leave their cells, the merged particle
  // enters the cell of the volume-weighted position; admitted if every cell
  // it removes from keeps at least ppc_min.
  bool merge_admit(int sid, int pid, const std::vector<int>& fcnt, std::vector<int>& bud) const {
    const Ptl& a = pool[sid];
    const Ptl& b = pool[pid];
    double vol = a.vol + b.vol;
    double xn[3] = {a.x[0], a.x[1], a.x[2]};
    for (int t = 0; t < D; t++) xn[t] = (a.x[t] * a.vol + b.x[t] * b.vol) / vol;
    int cell[3] = {cell_index_of(a.x), cell_index_of(b.x), cell_index_of(xn)};
    int net[3] = {-1, -1, 1};
    for (int t = 1; t < 3; t++)
      for (int u = 0; u < t; u++)
        if (cell[t] == cell[u] && net[u] != 0) {
          net[u] += net[t];
          net[t] = 0;
          break;
        }
    bool ok = true;
    for (int t = 0; t < 3; t++)
      if (net[t] < 0 && fcnt[cell[t]] - bud[cell[t]] + net[t] < cs.ppc_min) ok = false;
    for (int t = 0; t < 3; t++)
      if (net[t] < 0) bud[cell[t]] -= net[t];
    return ok;
  }

  void split_fast() {
    const int child_N = D == 3 ? 8 : 4;
    double vol_max = (D == 3 ? (cs.p_sz * cs.p_sz) * cs.p_sz : cs.p_sz * cs.p_sz) * 1.6;
    struct SplitRec {
      int parent, old_cell;
      int child[7];
      std::vector<int> old_nb;
    };
    struct OutRec {
      int p, old_cell;
      std::vector<int> old_nb;
    };
    std::vector<int> new_particles;
    std::vector<SplitRec> split_records;
    std::vector<int> inlet_records;
    std::vector<OutRec> outlet_records;
    std::unordered_set<int> outlet_set;
    bool io = s3();
    // 1. split / inlet / outlet, single-thread order
    std::vector<int> fcnt, bud;
    if (cs.ppc_max > 0) {
      fluid_counts(fcnt);
      bud.assign(nc, 0);
    }
    std::vector<int> snapshot = fluid;
    for (int id : snapshot) {
      pool[id].copy_p = -1;
      pool[id].mrg_mark = 0;
      double dp = root(pool[id].vol);
      bool f1 = pool[id].vol > vol_max;
      // per-cell bound: no split that would take a cell above ppc_max
      if (f1 && cs.ppc_max > 0) f1 = split_admit(pool[id].x, dp, fcnt, bud);
      bool f2 = io && pool[id].x[0] - cs.box_l[0] >= dp && pool[id].mark_in == 1;
      bool f3 = io && pool[id].x[0] > cs.box_r[0];
      if (f1) {
        SplitRec rec;
        rec.parent = id;
        rec.old_cell = cell_index_of(pool[id].x);
        double pref[3];
        for (int a = 0; a < 3; a++) pref[a] = pool[id].x[a];
        pref[0] = pool[id].x[0] - dp / 4.;
        pref[1] = pool[id].x[1] - dp / 4.;
        if (D == 3) pref[2] = pool[id].x[2] - dp / 4.;
        int child[8];
        int ci = 0;
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b)
            for (int c = 0; c < (D == 3 ? 2 : 1); ++c) {
              int nid = new_obj();
              Ptl& ch = pool[nid];
              ch.x[0] = pref[0] + dp / 2. * a;
              ch.x[1] = pref[1] + dp / 2. * b;
              ch.x[2] = 0;
              if (D == 3) ch.x[2] = pref[2] + dp / 2. * c;
              set_split(ch, pool[id]);
              child[ci++] = nid;
            }
        rec.old_nb = pool[id].nb;
        copy_split_state(pool[id], pool[child[child_N - 1]]);
        pool[child[child_N - 1]].clear();  // freed; never referenced
        for (int j = 0; j < child_N - 1; ++j) {
          rec.child[j] = child[j];
          new_particles.push_back(child[j]);
        }
        split_records.push_back(rec);
      } else if (f2) {
        int nid = new_obj();
        Ptl& q = pool[nid];
        for (int a = 0; a < 3; a++) q.x[a] = pool[id].x[a];
        q.x[0] = pool[id].x[0] - dp;
        set_ptcl(q, pool[id]);
        inlet_records.push_back(nid);
        new_particles.push_back(nid);
      } else if (f3) {
        OutRec rec;
        rec.p = id;
        rec.old_cell = cell_index_of(pool[id].x);
        rec.old_nb = pool[id].nb;
        outlet_records.push_back(rec);
      }
    }
    n_split = (int64_t)split_records.size();
    n_inlet = (int64_t)inlet_records.size();
    n_outlet = (int64_t)outlet_records.size();
    for (auto& r : outlet_records) outlet_set.insert(r.p);
    // 2. new particles -> particle_fluid
    for (int id : new_particles) fluid.push_back(id);
    // 3. split cell-list edits
    for (auto& rec : split_records) {
      remove_from(bins[rec.old_cell], rec.parent);
      bins[cell_index_of(pool[rec.parent].x)].push_back(rec.parent);
      for (int j = 0; j < child_N - 1; ++j) bins[cell_index_of(pool[rec.child[j]].x)].push_back(rec.child[j]);
    }
    // 4. inlet
    for (int id : inlet_records) bins[cell_index_of(pool[id].x)].push_back(id);
    // 5. outlet
    for (auto& rec : outlet_records) remove_from(bins[rec.old_cell], rec.p);
    // 6. affected particles
    std::unordered_set<int> aff_set, scanned;
    std::vector<int> affected;
    auto add = [&](int r) {
      if (r == -1 || is_wall(r)) return;  // wall lists are never read
      if (outlet_set.count(r)) return;
      if (aff_set.insert(r).second) affected.push_back(r);
    };
    for (auto& rec : split_records) {
      for (int r : rec.old_nb) add(r);
      int sp[8];
      sp[0] = rec.parent;
      for (int j = 0; j < child_N - 1; ++j) sp[j + 1] = rec.child[j];
      for (int j = 0; j < child_N; ++j) {
        add(sp[j]);
        collect_window(cell_index_of(pool[sp[j]].x), scanned, add);
      }
    }
    for (int id : inlet_records) {
      add(id);
      collect_window(cell_index_of(pool[id].x), scanned, add);
    }
    for (auto& rec : outlet_records)
      for (int r : rec.old_nb)
        if (r != rec.p) add(r);
    // 7. refresh (status 3)
    for (int id : affected) refresh_neighbors(id, 3);
    // 8. remove outlet particles, recompute mark_in
    if (io) {
      std::vector<int> nf;
      for (int id : fluid) {
        if (outlet_set.count(id)) continue;
        Ptl& q = pool[id];
        q.mark_in = 0;
        if (q.x[0] - cs.box_l[0] < root(q.vol)) q.mark_in = 1;
        nf.push_back(id);
      }
      fluid.swap(nf);
      for (auto& rec : outlet_records) pool[rec.p].clear();
    }
  }

  // ---- Merge_particle_fast (S3:mpm.cpp:1536-1706) ---------------------------
  void set_merge(int sid) {  // Set_merge_ptcl (S3:particle.cpp:614-661)
    Ptl& cp = pool[sid];
    const Ptl cp1 = pool[cp.copy_p];
    double mass0 = cp.mass, mass1 = cp1.mass, vol0 = cp.vol, vol1 = cp1.vol;
    double x0[3], x1[3];
    for (int a = 0; a < 3; a++) { x0[a] = cp.x[a]; x1[a] = cp1.x[a]; }
    double c0[U][M];
    std::memcpy(c0, cp.coef, sizeof(c0));
    double rho0 = cp.rho, rho1 = cp1.rho, E0 = cp.E, E1 = cp1.E;
    double rv0[3], rv1[3];
    for (int a = 0; a < 3; a++) { rv0[a] = cp.rhov[a]; rv1[a] = cp1.rhov[a]; }
    cp.mass = mass0 + mass1;
    cp.vol = vol0 + vol1;
    cp.h = root(cp.vol) * CELL_RATIO / CUT_OFF;
    for (int a = 0; a < D; a++) cp.x[a] = (x0[a] * vol0 + x1[a] * vol1) / cp.vol;
    for (int u = 0; u < U; u++)
      for (int m = 0; m < M; m++) cp.coef[u][m] = (c0[u][m] * vol0 + cp1.coef[u][m] * vol1) / cp.vol;
    std::memcpy(cp.coef0, cp.coef, sizeof(cp.coef));
    cp.rho = (rho0 * vol0 + rho1 * vol1) / cp.vol;
    for (int a = 0; a < D; a++) cp.rhov[a] = (rv0[a] * vol0 + rv1[a] * vol1) / cp.vol;
    cp.E = (E0 * vol0 + E1 * vol1) / cp.vol;
    for (int a = 0; a < 3; a++) cp.tv[a] = cp.v[a] = cp.rhov[a] / cp.rho;
    double sq = 0;
    for (int a = 0; a < D; a++) sq += cp.v[a] * cp.v[a];
    double ei = cp.E - 0.5 * cp.rho * sq;
    int s = cp.phase - 1;
    cp.P = eos_p(ei, s);
    cp.c = std::sqrt(cs.gamma[s] * (cs.pinf[s] + cp.P) / cp.rho);
    cp.mrg_mark = 0;
    cp.mrg = 0;
  }

  // fluid particles in every cell's bin
  void fluid_counts(std::vector<int>& fcnt) const {
    fcnt.assign(nc, 0);
    for (int c = 0; c < nc; c++)
      for (int r : bins[c])
        if (r >= 0) fcnt[c]++;
  }

  void merge_fast() {
    double vol_min = (D == 3 ? (cs.p_sz * cs.p_sz) * cs.p_sz : cs.p_sz * cs.p_sz) * 2. / 3.;
    struct Rec {
      int s, p, s_old, p_old;
      std::vector<int> old_nb;
    };
    std::vector<Rec> recs;
    // per-cell bound: merge candidates only in cells holding more than
    // ppc_min fluid particles when the pass starts
    std::vector<int> fcnt, bud;
    if (cs.ppc_min > 0) {
      fluid_counts(fcnt);
      bud.assign(nc, 0);
    }
    auto small = [&](const Ptl& q) {
      if (!(q.vol < vol_min)) return false;
      return cs.ppc_min <= 0 || fcnt[cell_index_of(q.x)] > cs.ppc_min;
    };
    for (int id : fluid) {
      Ptl& cp = pool[id];
      cp.mrg = 0;
      cp.mrg_mark = 0;
      int cp1 = -1;
      if (small(cp)) {
        double dss = root(vol_min);
        for (int r : cp.nb) {
          if (r == id || r == -1) continue;
          // stale references to freed objects behave as cleared particles
          // (phase 0), wall particles have phase 10: neither can match
          if (is_wall(r)) continue;
          const Ptl& nb = pool[r];
          double d[3];
          for (int a = 0; a < 3; a++) d[a] = cp.x[a] - nb.x[a];
          double dist = norm_d(d);
          if (nb.vol < vol_min && dist < dss && cp.phase == nb.phase && small(nb)) {
            dss = dist;
            cp1 = r;
            cp.mrg = 1;
          }
        }
      }
      cp.copy_p = cp1;
    }
    std::vector<char> seen(pool.size(), 0);  // partner of an earlier pair
    for (int id : fluid) {
      Ptl& cp = pool[id];
      if (cp.mrg == 1 && cp.copy_p >= 0 && pool[cp.copy_p].copy_p == id && cp.mrg_mark == 0 && !seen[id]) {
        seen[cp.copy_p] = 1;
        // per-cell bound: no merge that would take a cell below ppc_min
        if (cs.ppc_min > 0 && !merge_admit(id, cp.copy_p, fcnt, bud)) continue;
        Rec rec;
        rec.s = id;
        rec.p = cp.copy_p;
        rec.s_old = cell_index_of(cp.x);
        rec.p_old = cell_index_of(pool[rec.p].x);
        rec.old_nb = cp.nb;
        for (int r : pool[rec.p].nb) rec.old_nb.push_back(r);
        set_merge(id);
        pool[rec.p].mrg_mark = 1;
        recs.push_back(rec);
      }
    }
    n_merge += (int64_t)recs.size();
    std::unordered_set<int> aff_set, scanned;
    std::vector<int> affected;
    auto add = [&](int r) {
      if (r == -1 || is_wall(r)) return;
      if (pool[r].mrg_mark == 1) return;
      if (aff_set.insert(r).second) affected.push_back(r);
    };
    for (auto& rec : recs) {
      for (int r : rec.old_nb)
        if (r != rec.s && r != rec.p) add(r);
      remove_from(bins[rec.s_old], rec.s);
      remove_from(bins[rec.p_old], rec.p);
      int nc_ = cell_index_of(pool[rec.s].x);
      bins[nc_].push_back(rec.s);
      add(rec.s);
      collect_window(nc_, scanned, add);
    }
    for (int id : affected) refresh_neighbors(id, 3);
    std::vector<int> nf;
    for (int id : fluid) {
      Ptl& q = pool[id];
      if (q.mrg_mark == 1) {
        q.clear();
        continue;  // mrg_mark stays 1 on the freed object (Clearup keeps it)
      }
      q.mrg = 0;
      q.mrg_mark = 0;
      if (s3()) {
        q.mark_in = 0;
        if (q.x[0] - cs.box_l[0] < root(q.vol)) q.mark_in = 1;
      }
      nf.push_back(id);
    }
    fluid.swap(nf);
  }

  // Reset_particle_neighbor_info(4) + Refresh_neighbor_info(4)
  void membership() {
    for (int id : fluid) pool[id].in_ele = -1;
    for (int ci : used) {
      Cell& C = cells[ci];
      C.nb.clear();
      int b = cell_index_of(C.x);
      for (int r : bins[b]) {
        if (is_wall(r)) continue;  // phase 10
        C.nb.push_back(r);
        if (pool[r].in_ele < 0) pool[r].in_ele = ci;
      }
    }
  }

  // ---- MLS_P2G (S3:mpm.cpp:274-382) -------------------------------------------
  void mls_p2g() {
    double hi = cs.p_g * 1 * 3.2 / CUT_OFF;
    par_range((int64_t)used.size(), [&](int64_t b0_, int64_t b1_) {
      for (int64_t k_ = b0_; k_ < b1_; k_++) {
      const int ci = used[k_];
      Cell& cp = cells[ci];
      double mass_t[PGDG_MAX_PHASE], p_avg[PGDG_MAX_PHASE][U], Cf[PGDG_MAX_PHASE][U][M - 1];
      double q_avg[PGDG_MAX_PHASE][U];  // order 2: the particles' own values
      std::memset(mass_t, 0, sizeof(mass_t));
      std::memset(p_avg, 0, sizeof(p_avg));
      std::memset(Cf, 0, sizeof(Cf));
      std::memset(q_avg, 0, sizeof(q_avg));
      int i = cp.idx[0], j = cp.idx[1], k = cp.idx[2];
      int kc = D == 3 ? 1 : 0;
      for (int t = IMAX(i - 1, 0); t <= IMIN(i + 1, N[0] - 1); t++)
        for (int s = IMAX(j - 1, 0); s <= IMIN(j + 1, N[1] - 1); s++)
          for (int q = IMAX(k - kc, 0); q <= IMIN(k + kc, N[2] - 1); q++) {
            const Cell& nc_ = cells[cidx(t, s, q)];
            for (int id : nc_.nb) {
              const Ptl& nb = pool[id];
              double xij[3];
              for (int a = 0; a < 3; a++) xij[a] = nb.x[a] - cp.x[a];
              double dist = norm_d(xij);
              if (dist < cs.p2g_cut) {  // reference: 1.5 * p_sz
                double ww = W(dist, hi) * nb.vol;
                double pj[U];
                pj[0] = nb.rho;
                for (int a = 0; a < D; a++) pj[1 + a] = nb.rhov[a];
                pj[U - 1] = nb.E;
                double pq[U];
                for (int u = 0; u < U; u++) pq[u] = pj[u];
                // the particle's polynomial at the cell centre
                const double nx[3] = {-xij[0], -xij[1], -xij[2]};
                double psi[M - 1];
                incr_tail<D, ORD>(nx, psi);
                for (int u = 0; u < U; u++) {
                  double sacc = 0;
                  for (int kk = 0; kk < M - 1; kk++) sacc += nb.coef[u][1 + kk] * psi[kk];
                  pj[u] += sacc;
                }
                int ph = nb.phase - 1;
                if (ph >= 0 && ph < F) {
                  for (int u = 0; u < U; u++) p_avg[ph][u] += pj[u] * ww;
                  if (ORD == 2)
                    for (int u = 0; u < U; u++) q_avg[ph][u] += pq[u] * ww;
                  for (int u = 0; u < U; u++)
                    for (int m = 0; m < M - 1; m++) Cf[ph][u][m] += nb.coef[u][1 + m] * ww;
                  mass_t[ph] += ww;
                }
              }
            }
          }
      double mass_total = 0;
      for (int t = 0; t < F; t++) mass_total += mass_t[t];
      for (int ss = 0; ss < F; ss++) {
        cp.ph_state[ss] = mass_t[ss] > 0 ? 1 : 0;
        cp.alph[ss] = mass_t[ss] / (mass_total + 1e-20);
        if (cp.ph_state[ss] == 1 && cp.ph_state_b[ss] == 0) {
          std::memset(cp.coef[ss], 0, sizeof(cp.coef[ss]));
          for (int u = 0; u < U; u++) p_avg[ss][u] /= (mass_t[ss] + 1e-20);
          for (int u = 0; u < U; u++)
            for (int m = 0; m < M - 1; m++) Cf[ss][u][m] /= (mass_t[ss] + 1e-20);
          if (ORD == 2) {  // centre value -> mean of the zero-mean basis
            for (int u = 0; u < U; u++)
              for (int a = 0; a < D; a++) p_avg[ss][u] += Cf[ss][u][(D == 2 ? 3 : 6) + a] * q12[a];
            // positivity: a mean extrapolated from the particles' quadratic
            // copies with non-positive density or pressure below
            // -0.05 Pp0 (the bound of Update_state_onestep's fallback,
            // S3:particle_flow.cpp:460-469) is replaced by the weighted
            // average of the particles' own values, with zero slopes
            double sq = 0;
            for (int a = 0; a < D; a++) sq += p_avg[ss][1 + a] * p_avg[ss][1 + a];
            double e_m = p_avg[ss][U - 1] - 0.5 * sq / p_avg[ss][0];
            double P_m = eos_p(e_m, ss);
            if (!(p_avg[ss][0] > 0) || P_m < -0.05 * cs.pinf[ss]) {
              for (int u = 0; u < U; u++) p_avg[ss][u] = q_avg[ss][u] / (mass_t[ss] + 1e-20);
              for (int u = 0; u < U; u++)
                for (int m = 0; m < M - 1; m++) Cf[ss][u][m] = 0;
            }
          }
          cp.coef[ss][0][0] = cp.rho[ss] = p_avg[ss][0];
          for (int a = 0; a < D; a++) cp.coef[ss][1 + a][0] = cp.rhov[ss][a] = p_avg[ss][1 + a];
          cp.coef[ss][U - 1][0] = cp.E[ss] = p_avg[ss][U - 1];
          for (int u = 0; u < U; u++)
            for (int m = 0; m < M - 1; m++) cp.coef[ss][u][1 + m] = Cf[ss][u][m];
          for (int a = 0; a < 3; a++) cp.v[a] = cp.rhov[ss][a] / cp.rho[ss];
          double sq = 0;
          for (int a = 0; a < D; a++) sq += cp.rhov[ss][a] * cp.rhov[ss][a];
          double e_j = cp.E[ss] - 0.5 * sq / cp.rho[ss];
          cp.P = eos_p(e_j, ss);
          std::memcpy(cp.coef0[ss], cp.coef[ss], sizeof(cp.coef[ss]));
        }
        cp.ph_state_b[ss] = cp.ph_state[ss];
        if (cp.ph_state[ss] == 0) {
          std::memset(cp.coef[ss], 0, sizeof(cp.coef[ss]));
          std::memset(cp.coef0[ss], 0, sizeof(cp.coef0[ss]));
          // T2/S3 reset the absent phase's density and momentum
          // (T2:mpm.cpp:346-352); UX keeps them (UX:mpm.cpp:330-334)
          if (!ux()) {
            cp.rho[ss] = cs.ini_rho[ss];
            for (int a = 0; a < 3; a++) cp.rhov[ss][a] = 0;
          }
        }
      }
    }
    });
  }

  // ---- Calculate_limiter (S3:mpm.cpp:392-738; T2:mpm.cpp:887-1225) ---------
  static double limiter_fn(double ds, int bd_fg) {  // S3:glbfunc.h:283-299
    double val = 1;
    if (ds > 0 && ds < 0.5) val = 2 * ds;
    else if (ds >= 0.5) val = 1;
    if (bd_fg == 1) val = (std::abs(ds) + ds) / (std::abs(ds) + 1);
    return val;
  }

  void calculate_limiter() {
    par_range((int64_t)used.size(), [&](int64_t b0_, int64_t b1_) {
      for (int64_t k_ = b0_; k_ < b1_; k_++) {
      const int ci = used[k_];
      Cell& cp = cells[ci];
      int i = cp.idx[0], j = cp.idx[1], k = cp.idx[2];
      double u_max[PGDG_MAX_PHASE][U], u_min[PGDG_MAX_PHASE][U], pi[PGDG_MAX_PHASE][U];
      for (int s = 0; s < F; s++)
        for (int u = 0; u < U; u++) {
          u_max[s][u] = -1e20;
          u_min[s][u] = 1e20;
        }
      for (int s = 0; s < F; s++) {
        for (int u = 0; u < U; u++) cp.lim[s][u] = 1.;
        pi[s][0] = cp.rho[s];
        for (int a = 0; a < D; a++) pi[s][1 + a] = cp.rhov[s][a];
        pi[s][U - 1] = cp.E[s];
      }
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) cp.G[a][b] = 0.;
      int bd_fg = 0;
      const double* vc = cp.v;
      // G_v columns, S3:mpm.cpp:471-533 (T2:mpm.cpp:922-952); UX takes a
      // zero wall velocity on the x sides too (UX:mpm.cpp:388-399)
      if (i == N[0] - 1) {
        for (int a = 0; a < D; a++) cp.G[a][0] = ux() ? (0.0 - vc[a]) / (0.5 * ps[0]) : (vc[a] - vc[a]) / (0.5 * ps[0]);
        bd_fg = 1;
      } else if (i == 0) {
        for (int a = 0; a < D; a++) cp.G[a][0] = ux() ? (vc[a] - 0.0) / (0.5 * ps[0]) : (vc[a] - vc[a]) / (0.5 * ps[0]);
        bd_fg = 1;
      } else {
        const double* vR = cells[cidx(i + 1, j, k)].v;
        const double* vL = cells[cidx(i - 1, j, k)].v;
        for (int a = 0; a < D; a++) cp.G[a][0] = (vR[a] - vL[a]) / (2 * ps[0]);
      }
      if (j == N[1] - 1) {
        for (int a = 0; a < D; a++) cp.G[a][1] = (0.0 - vc[a]) / (0.5 * ps[1]);
        bd_fg = 1;
      } else if (j == 0) {
        for (int a = 0; a < D; a++) cp.G[a][1] = (vc[a] - 0.0) / (0.5 * ps[1]);
        bd_fg = 1;
      } else {
        const double* vR = cells[cidx(i, j + 1, k)].v;
        const double* vL = cells[cidx(i, j - 1, k)].v;
        for (int a = 0; a < D; a++) cp.G[a][1] = (vR[a] - vL[a]) / (2 * ps[1]);
      }
      if (D == 3) {
        if (k == N[2] - 1) {
          for (int a = 0; a < D; a++) cp.G[a][2] = (0.0 - vc[a]) / (0.5 * ps[2]);
          bd_fg = 1;
        } else if (k == 0) {
          for (int a = 0; a < D; a++) cp.G[a][2] = (vc[a] - 0.0) / (0.5 * ps[2]);
          bd_fg = 1;
        } else {
          const double* vR = cells[cidx(i, j, k + 1)].v;
          const double* vL = cells[cidx(i, j, k - 1)].v;
          for (int a = 0; a < D; a++) cp.G[a][2] = (vR[a] - vL[a]) / (2 * ps[2]);
        }
      }
      // neighbour extrema
      for (int jj = 0; jj < NF; jj++) {
        const Face& fc = cp.f[jj];
        int b = fc.b;
        double pj[U];
        for (int u = 0; u < U; u++) pj[u] = 0;
        if (b == 0) {
          const Cell& nb = cells[fc.nbr];
          for (int s = 0; s < F; s++) {
            for (int u = 0; u < U; u++) pj[u] = 0;
            for (int s1 = 0; s1 < F; s1++) {
              if (nb.alph[s1] == 0) continue;
              double rhoj = nb.rho[s1];
              double rvj[3];
              for (int a = 0; a < D; a++) rvj[a] = nb.rhov[s1][a];
              double Ej = nb.E[s1];
              if (s1 != s) {
                double vj[3];
                for (int a = 0; a < D; a++) vj[a] = rvj[a] / rhoj;
                double sq = sq_d(vj);
                double rhoe_j = Ej - 0.5 * rhoj * sq;
                double P_j = eos_p(rhoe_j, s1);
                rhoj = cp.rho[s];
                for (int a = 0; a < D; a++) rvj[a] = rhoj * vj[a];
                Ej = eos_rhoe(P_j, s) + 0.5 * rhoj * sq;
              }
              double tmp[U];
              tmp[0] = rhoj;
              for (int a = 0; a < D; a++) tmp[1 + a] = rvj[a];
              tmp[U - 1] = Ej;
              for (int u = 0; u < U; u++) pj[u] += tmp[u] * nb.alph[s1];
            }
            for (int u = 0; u < U; u++) {
              double d = pj[u] - pi[s][u];
              u_max[s][u] = AMAX1(u_max[s][u], d);
              u_min[s][u] = AMIN1(u_min[s][u], d);
            }
          }
        } else {
          for (int s = 0; s < F; s++) {
            if (s3()) {
              pj[0] = cp.rho[s];
              for (int a = 0; a < D; a++) pj[1 + a] = cp.rhov[s][a];
              pj[U - 1] = cp.E[s];
              if (b == -1 || b == -2)
                for (int u = 0; u < U; u++) pj[u] = pi[s][u];
              if (j == N[1] - 1 || j == 0) pj[2] = -cp.rhov[s][1];
              if (D == 3 && (k == N[2] - 1 || k == 0)) pj[3] = -cp.rhov[s][2];
            } else if (ux()) {
              // the cell's own means with the momentum normal to a domain
              // edge set to 0, whatever the side's tag (UX:mpm.cpp:478-487)
              pj[0] = cp.rho[s];
              for (int a = 0; a < D; a++) pj[1 + a] = cp.rhov[s][a];
              pj[U - 1] = cp.E[s];
              if (i == N[0] - 1 || i == 0) pj[1] = 0;
              if (j == N[1] - 1 || j == 0) pj[2] = 0;
            } else {
              if (b == 1) {
                pj[0] = cp.rho[s];
                pj[1] = cp.rhov[s][0];
                pj[2] = -cp.rhov[s][1];
                pj[U - 1] = cp.E[s];
              } else if (b == -1 || b == -2) {
                for (int u = 0; u < U; u++) pj[u] = pi[s][u];
              }
            }
            for (int u = 0; u < U; u++) {
              double d = pj[u] - pi[s][u];
              u_max[s][u] = AMAX1(u_max[s][u], d);
              u_min[s][u] = AMIN1(u_min[s][u], d);
            }
          }
        }
      }
      for (int s = 0; s < F; s++)
        for (int u = 0; u < U; u++) {
          u_max[s][u] = AMAX1(u_max[s][u], 0.);
          u_min[s][u] = AMIN1(u_min[s][u], 0.);
        }
      // ratio at neighbour centres / boundary face centres
      for (int jj = 0; jj < NF; jj++) {
        const Face& fc = cp.f[jj];
        double dr[3];
        if (fc.b == 0)
          for (int a = 0; a < 3; a++) dr[a] = cells[fc.nbr].x[a] - cp.x[a];
        else
          for (int a = 0; a < 3; a++) dr[a] = fc.cen[a] - cp.x[a];
        for (int s = 0; s < F; s++) {
          // UX's Limiter is 2 ds / (ds + 1) everywhere (UX:glbfunc.h:291-306, UX:mpm.cpp:522-523)
          int lim_fg = ux() ? 1 : bd_fg;
          if (cp.alph[s] != 1) lim_fg = 1;
          double phi[M - 1];
          basis_tail<D, ORD>(dr, q12, phi);
          for (int u = 0; u < U; u++) {
            double dU = 0;
            for (int kk = 0; kk < M - 1; kk++) dU += cp.coef[s][u][1 + kk] * phi[kk];
            double sd = 1;
            if (dU >= 0) {
              sd = u_max[s][u] / (dU + 1e-20);
              sd = limiter_fn(std::abs(sd), lim_fg);
            } else if (dU < 0) {
              sd = u_min[s][u] / (dU + 1e-20);
              sd = limiter_fn(std::abs(sd), lim_fg);
            }
            cp.lim[s][u] = AMIN1(cp.lim[s][u], sd);
          }
        }
      }
      // positivity probe
      for (int s = 0; s < F; s++) {
        int fg = 0;
        for (int jj = 0; jj < 4 && !fg; jj++) {
          for (int k1 = 0; k1 < (D == 3 ? 2 : 1) && !fg; k1++) {
            double dr[3] = {0, 0, 0};
            dr[0] = sqrt2 * cs.p_g * probe_c[jj];
            dr[1] = sqrt2 * cs.p_g * probe_s[jj];
            if (D == 3) dr[2] = (-0.5 + k1) * cs.p_g;
            double phi[M - 1];
            basis_tail<D, ORD>(dr, q12, phi);
            double Ui[U];
            for (int u = 0; u < U; u++) {
              double acc = 0;
              for (int kk = 0; kk < M - 1; kk++) acc += (cp.lim[s][u] * cp.coef[s][u][1 + kk]) * phi[kk];
              Ui[u] = cp.coef[s][u][0] + acc;
            }
            double rho_p = Ui[0], E_p = Ui[U - 1];
            double sq = 0;
            for (int a = 0; a < D; a++) sq += Ui[1 + a] * Ui[1 + a];
            double e_p = E_p - 0.5 * sq / rho_p;
            double P_p = eos_p(e_p, s);
            if (P_p < -0.05 * cs.pinf[s] || rho_p < 0) fg = 1;
          }
        }
        if (fg == 1)
          for (int u = 0; u < U; u++) cp.lim[s][u] = 0.;
      }
    }
    });
  }

  // ---- Grid_acceleration (S3:particle_flow.cpp:16-405) ---------------------
  struct Trace {
    double rho, rv[3], E, v[3], P, c, vn;
  };
  // value of phase s of cell C at offset dr, slopes scaled by lim
  void trace(const Cell& C, int s, const double* dr, Trace& t) const {
    double phi[M - 1];
    basis_tail<D, ORD>(dr, q12, phi);
    double dU[U];
    for (int u = 0; u < U; u++) {
      double acc = 0;
      for (int kk = 0; kk < M - 1; kk++) acc += C.coef[s][u][1 + kk] * phi[kk];
      dU[u] = C.coef[s][u][0] + C.lim[s][u] * acc;
    }
    t.rho = dU[0];
    t.E = dU[U - 1];
    for (int a = 0; a < 3; a++) t.rv[a] = 0;
    for (int a = 0; a < D; a++) t.rv[a] = dU[1 + a];
  }
  // DF = U v^T + P I (+ v^T P in the energy row); tau = 0 (mu = 0)
  static void flux(double rho, const double* rv, double E, const double* v, double P,
                   double DF[U][D]) {
    for (int b = 0; b < D; b++) {
      DF[0][b] = rho * v[b];
      for (int a = 0; a < D; a++) DF[1 + a][b] = rv[a] * v[b];
      DF[U - 1][b] = E * v[b];
    }
    for (int a = 0; a < D; a++) DF[1 + a][a] += P;
    for (int b = 0; b < D; b++) DF[U - 1][b] += v[b] * P;
  }

  // HLLC flux of an interior face point (S3:particle_flow.cpp:232-272):
  // Davis wave speeds, contact speed S*, star states, four-branch select.
  // Left state i, right state j (conserved rho, rho v, E plus P, c and v),
  // face normal n.  Fh is the full U x D flux tensor; the caller contracts
  // it with n.
  static void hllc(const double* n, double rho_i, const double* rvi, double E_i, double P_i,
                   double ci_, const double* vi, double rho_j, const double* rvj, double E_j,
                   double P_j, double cj, const double* vj, double Fh[U][D]) {
    double Ui[U], Uj[U];
    Ui[0] = rho_i;
    for (int a = 0; a < D; a++) Ui[1 + a] = rvi[a];
    Ui[U - 1] = E_i;
    Uj[0] = rho_j;
    for (int a = 0; a < D; a++) Uj[1 + a] = rvj[a];
    Uj[U - 1] = E_j;
    double DFi[U][D], DFj[U][D];
    flux(rho_i, rvi, E_i, vi, P_i, DFi);
    flux(rho_j, rvj, E_j, vj, P_j, DFj);
    double vni = 0;
    for (int a = 0; a < D; a++) vni += vi[a] * n[a];
    double vnj = 0;
    for (int a = 0; a < D; a++) vnj += vj[a] * n[a];
    double SL = AMIN1(vni - ci_, vnj - cj);
    double SR = AMAX1(vni + ci_, vnj + cj);
    double S_s = ((P_j - P_i) + rho_i * vni * (SL - vni) - rho_j * vnj * (SR - vnj)) /
                 (rho_i * (SL - vni) - rho_j * (SR - vnj));
    double rho_sL = rho_i * (SL - vni) / (SL - S_s);
    double rho_sR = rho_j * (SR - vnj) / (SR - S_s);
    double UsL[U], UsR[U];
    UsL[0] = rho_sL;
    UsR[0] = rho_sR;
    for (int a = 0; a < D; a++) {
      UsL[1 + a] = rho_sL * (vi[a] + (S_s - vni) * n[a]);
      UsR[1 + a] = rho_sR * (vj[a] + (S_s - vnj) * n[a]);
    }
    UsL[U - 1] = rho_sL * (E_i / rho_i + (S_s - vni) * (S_s + (P_i) / (rho_i * (SL - vni))));
    UsR[U - 1] = rho_sR * (E_j / rho_j + (S_s - vnj) * (S_s + (P_j) / (rho_j * (SR - vnj))));
    std::memset(Fh, 0, sizeof(double) * U * D);
    if (SL > 0) {
      std::memcpy(Fh, DFi, sizeof(DFi));
    } else if (SL <= 0 && S_s >= 0) {
      for (int u = 0; u < U; u++)
        for (int bb = 0; bb < D; bb++) Fh[u][bb] = DFi[u][bb] + SL * (UsL[u] - Ui[u]) * n[bb];
    } else if (S_s <= 0 && SR >= 0) {
      for (int u = 0; u < U; u++)
        for (int bb = 0; bb < D; bb++) Fh[u][bb] = DFj[u][bb] + SR * (UsR[u] - Uj[u]) * n[bb];
    } else if (SR <= 0) {
      std::memcpy(Fh, DFj, sizeof(DFj));
    }
  }
  // HLL flux of a boundary face point against its ghost state
  // (S3:particle_flow.cpp:367-381).
  static void hll(const double* n, double rho_i, const double* rvi, double E_i, double P_i,
                  double ci_, const double* vi, double rho_j, const double* rvj, double E_j,
                  double P_j, double cj, const double* vj, double Fh[U][D]) {
    double Ui[U], Uj[U];
    Ui[0] = rho_i;
    for (int a = 0; a < D; a++) Ui[1 + a] = rvi[a];
    Ui[U - 1] = E_i;
    Uj[0] = rho_j;
    for (int a = 0; a < D; a++) Uj[1 + a] = rvj[a];
    Uj[U - 1] = E_j;
    double DFi[U][D], DFj[U][D];
    flux(rho_i, rvi, E_i, vi, P_i, DFi);
    flux(rho_j, rvj, E_j, vj, P_j, DFj);
    double vni = 0;
    for (int a = 0; a < D; a++) vni += vi[a] * n[a];
    double vnj = 0;
    for (int a = 0; a < D; a++) vnj += vj[a] * n[a];
    double SL = AMIN1(vni - ci_, vnj - cj);
    double SR = AMAX1(vni + ci_, vnj + cj);
    std::memset(Fh, 0, sizeof(double) * U * D);
    if (SL > 0) {
      std::memcpy(Fh, DFi, sizeof(DFi));
    } else if (SL <= 0 && SR >= 0) {
      for (int u = 0; u < U; u++)
        for (int bb = 0; bb < D; bb++)
          Fh[u][bb] = (SR * DFi[u][bb] - SL * DFj[u][bb] + SL * SR * (Uj[u] - Ui[u]) * n[bb]) / (SR - SL);
    } else if (SR < 0) {
      std::memcpy(Fh, DFj, sizeof(DFj));
    }
  }

  void grid_acceleration() {
    par_range((int64_t)used.size(), [&](int64_t b0, int64_t b1) {
      for (int64_t k = b0; k < b1; k++) {
        const int ci = used[k];
        for (int s = 0; s < F; s++)
          if (cells[ci].ph_state[s] == 1) cell_rhs(ci, s);
      }
    });
  }

  void cell_rhs(int ci, int ss) {
    Cell& cp = cells[ci];
    double hp[3] = {ps[0] / 2, ps[1] / 2, ps[2] / 2};
    std::memset(cp.dcoef[ss], 0, sizeof(cp.dcoef[ss]));
    double voli = cp.Vol;
    double mij[M];
    mij[0] = 1.;
    for (int a = 0; a < D; a++) mij[1 + a] = hp[a] * hp[a] / 3.;
    if (ORD == 2) {  // UX:particle_flow.cpp:28-29: (hx hy)^2/9 and 4 hx^4/45
      const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
      for (int q = 0; q < (D == 2 ? 1 : 3); q++) {
        const double t = hp[pa[q]] * hp[pb[q]];
        mij[1 + D + q] = t * t / 9.;
      }
      for (int a = 0; a < D; a++) {
        const double t = hp[a] * hp[a];
        mij[M - D + a] = t * t * 4. / 45.;
      }
    }
    for (int m = 0; m < M; m++) mij[m] *= voli;
    const double g = cs.gamma[ss], pinf = cs.pinf[ss];
    // volume integral: Gauss points in g1, g2[, g3] order
    int ng3 = D == 3 ? NQ : 1;
    for (int g1 = 0; g1 < NQ; g1++)
      for (int g2 = 0; g2 < NQ; g2++)
        for (int g3 = 0; g3 < ng3; g3++) {
          double dp[3] = {0, 0, 0};
          dp[0] = Gp[g1] * hp[0];
          dp[1] = Gp[g2] * hp[1];
          double gw = Gw[g1] * Gw[g2];
          if (D == 3) {
            dp[2] = Gp[g3] * hp[2];
            gw = Gw[g1] * Gw[g2] * Gw[g3];
          }
          Trace t;
          trace(cp, ss, dp, t);
          double v[3];
          for (int a = 0; a < D; a++) v[a] = t.rv[a] / t.rho;
          double e_p = t.E - 0.5 * sq_d(t.rv) / t.rho;
          double P_p = eos_p(e_p, ss);
          double DF[U][D];
          flux(t.rho, t.rv, t.E, v, P_p, DF);
          for (int bi = 1; bi <= D; bi++)
            for (int u = 0; u < U; u++) cp.dcoef[ss][u][bi] += DF[u][bi - 1] * gw * voli * 1;
          if (ORD == 2) {
            double col[D][U];
            for (int b2 = 0; b2 < D; b2++)
              for (int u = 0; u < U; u++) col[b2][u] = DF[u][b2];
            for (int bi = D + 1; bi < M; bi++)
              for (int u = 0; u < U; u++)
                cp.dcoef[ss][u][bi] += vol_weight<D, ORD>(bi, col, u, dp) * gw * voli;
          }
        }
    // face integral
    for (int j = 0; j < NF; j++) {
      const Face& fc = cp.f[j];
      const double* n = fc.nrm;
      int b = fc.b;
      double DFij[M][U][D];
      std::memset(DFij, 0, sizeof(DFij));
      int ng2 = D == 3 ? NQ : 1;
      for (int g1 = 0; g1 < NQ; g1++)
        for (int g2 = 0; g2 < ng2; g2++) {
          double dp[3], gw;
          if (D == 3) {
            for (int a = 0; a < 3; a++) dp[a] = Gp[g1] * fc.tan[0][a] + Gp[g2] * fc.tan[1][a];
            gw = Gw[g1] * Gw[g2];
          } else {
            for (int a = 0; a < 3; a++) dp[a] = Gp[g1] * fc.tan[0][a];
            gw = Gw[g1];
          }
          double dri[3];
          for (int a = 0; a < 3; a++) dri[a] = dp[a] + fc.cen[a] - cp.x[a];
          double Nb[M];
          Nb[0] = 1.;
          basis_tail<D, ORD>(dri, q12, Nb + 1);
          Trace ti;
          trace(cp, ss, dri, ti);
          double vi[3] = {0, 0, 0};
          for (int a = 0; a < D; a++) vi[a] = ti.rv[a] / ti.rho;
          double e_i = ti.E - 0.5 * sq_d(ti.rv) / ti.rho;
          double P_i = eos_p(e_i, ss);
          double ci_ = std::sqrt(g * (pinf + P_i) / ti.rho);
          double vni = 0;
          for (int a = 0; a < D; a++) vni += vi[a] * n[a];
          // T2: every boundary face is a wall; S3: tag 1; UX: none (every
          // boundary face takes the HLL ghost flux)
          bool wall_like = s3() ? (b == 1) : (!ux() && b != 0);
          if (b == 0) {
            const Cell& nb = cells[fc.nbr];
            for (int s1 = 0; s1 < F; s1++) {
              if (nb.alph[s1] == 0) continue;
              double drj[3];
              for (int a = 0; a < 3; a++) drj[a] = dp[a] + fc.cen[a] - nb.x[a];
              Trace tj;
              trace(nb, s1, drj, tj);
              double rho_j = tj.rho, E_j = tj.E;
              double rvj[3] = {0, 0, 0}, vj[3] = {0, 0, 0};
              for (int a = 0; a < D; a++) rvj[a] = tj.rv[a];
              for (int a = 0; a < D; a++) vj[a] = rvj[a] / rho_j;
              double rhoe_j = E_j - 0.5 * sq_d(vj) * rho_j;
              double P_j = eos_p(rhoe_j, s1);
              double vnj = 0;
              for (int a = 0; a < D; a++) vnj += vj[a] * n[a];
              if (ss != s1) {
                double gh[3];
                if (s3())
                  MGFM_S3(vni, vnj, P_i, P_j, ti.rho, rho_j, cs.gamma[ss], cs.pinf[ss],
                          cs.gamma[s1], cs.pinf[s1], 50, -1.0, gh, nullptr);
                else
                  MGFM_T2(vni, vnj, P_i, P_j, ti.rho, rho_j, cs.gamma[ss], cs.pinf[ss],
                          cs.gamma[s1], cs.pinf[s1], 50, -1.0, gh, nullptr);
                rho_j = gh[0];
                P_j = gh[1];
                vnj = gh[2];
                for (int a = 0; a < D; a++) vj[a] = vi[a] - vni * n[a] + vnj * n[a];
                for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
                rhoe_j = eos_rhoe(P_j, ss);
                E_j = rhoe_j + 0.5 * rho_j * sq_d(vj);
              }
              double cj = std::sqrt(g * (pinf + P_j) / rho_j);
              double alph_j = nb.alph[s1];
              double Fh[U][D];
              // UX: HLL at interior faces (UX:particle_flow.cpp:280-294)
              if (ux())
                hll(n, ti.rho, ti.rv, ti.E, P_i, ci_, vi, rho_j, rvj, E_j, P_j, cj, vj, Fh);
              else
                hllc(n, ti.rho, ti.rv, ti.E, P_i, ci_, vi, rho_j, rvj, E_j, P_j, cj, vj, Fh);
              for (int bi = 0; bi < M; bi++)
                for (int u = 0; u < U; u++)
                  for (int bb = 0; bb < D; bb++) DFij[bi][u][bb] += Fh[u][bb] * Nb[bi] * gw * alph_j;
            }
          } else if (wall_like) {
            for (int bi = 0; bi < M; bi++)
              for (int a = 0; a < D; a++) DFij[bi][1 + a][a] += P_i * Nb[bi] * gw;
          } else {  // S3 inlet (-1), outlet (-2), symmetry (-3) and every UX
                    // boundary face: HLL with a ghost state
                    // (UX:particle_flow.cpp:314-366)
            double rho_j = ti.rho;
            double vj[3] = {0, 0, 0};
            for (int a = 0; a < D; a++) vj[a] = vi[a];
            double rvj[3] = {0, 0, 0};
            for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
            double P_j = P_i;
            double E_j = ti.E;
            if (b == -1 && s3()) {
              rho_j = cs.inlet_rho;
              for (int a = 0; a < D; a++) vj[a] = cs.inlet_u[a];
              for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
              P_j = cs.inlet_p;
              E_j = eos_rhoe(P_j, ss) + 0.5 * sq_d(vj) * rho_j;
            }
            // S3 reflects v_y in the first and last y rows; UX on wall-tagged
            // sides (its tag 2)
            if (ux() ? b == PGDG_BC_WALL : (cp.idx[1] == N[1] - 1 || cp.idx[1] == 0)) {
              vj[1] = -vi[1];
              for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
            }
            if (D == 3 && (cp.idx[2] == N[2] - 1 || cp.idx[2] == 0)) {
              vj[2] = -vi[2];
              for (int a = 0; a < D; a++) rvj[a] = rho_j * vj[a];
            }
            double cj = std::sqrt(g * (pinf + P_j) / rho_j);
            double Fh[U][D];
            hll(n, ti.rho, ti.rv, ti.E, P_i, ci_, vi, rho_j, rvj, E_j, P_j, cj, vj, Fh);
            for (int bi = 0; bi < M; bi++)
              for (int u = 0; u < U; u++)
                for (int bb = 0; bb < D; bb++) DFij[bi][u][bb] += Fh[u][bb] * Nb[bi] * gw;
          }
        }
      for (int bi = 0; bi < M; bi++)
        for (int u = 0; u < U; u++) {
          double acc = 0;
          for (int bb = 0; bb < D; bb++) acc += -(DFij[bi][u][bb]) * n[bb];
          cp.dcoef[ss][u][bi] += acc * fc.S;
        }
    }
    for (int m = 0; m < M; m++) {
      double inv = 1. / (mij[m] + 1e-20);
      for (int u = 0; u < U; u++) cp.dcoef[ss][u][m] *= inv;
    }
  }

  // ---- Update_grid (S3:mpm.cpp:768-830; T2:mpm.cpp:1254-1320) -------------
  void update_grid(int md) {
    par_range((int64_t)used.size(), [&](int64_t b0_, int64_t b1_) {
      for (int64_t k_ = b0_; k_ < b1_; k_++) {
      const int ci = used[k_];
      Cell& cp = cells[ci];
      double rv_a[3] = {0, 0, 0};
      double rho_a = 0, P_a = 0;
      for (int s = 0; s < F; s++) {
        cp.c[s] = 0;
        cp.ac[s] = 0;
        for (int t = 0; t < U; t++) {
          double l = std::abs(cp.lim[s][t]);
          for (int m = 1; m < M; m++) cp.dcoef[s][t][m] *= l;
          for (int m = 1; m < M; m++) cp.coef[s][t][m] *= l;
        }
        if (cp.ph_state[s] == 1) {
          if (md == 0) {
            std::memcpy(cp.coef0[s], cp.coef[s], sizeof(cp.coef[s]));
            for (int u = 0; u < U; u++)
              for (int m = 0; m < M; m++) cp.coef[s][u][m] += cp.dcoef[s][u][m] * dt;
          } else if (md == 1) {
            for (int u = 0; u < U; u++)
              for (int m = 0; m < M; m++)
                cp.coef[s][u][m] = 0.5 * (cp.dcoef[s][u][m] * dt + cp.coef[s][u][m] + cp.coef0[s][u][m]);
          } else if (md == 2) {  // SSP-RK3 stage 2
            for (int u = 0; u < U; u++)
              for (int m = 0; m < M; m++)
                cp.coef[s][u][m] = 0.75 * cp.coef0[s][u][m] + 0.25 * (cp.coef[s][u][m] + cp.dcoef[s][u][m] * dt);
          } else {  // SSP-RK3 stage 3
            for (int u = 0; u < U; u++)
              for (int m = 0; m < M; m++)
                cp.coef[s][u][m] = (1. / 3.) * cp.coef0[s][u][m] + (2. / 3.) * (cp.coef[s][u][m] + cp.dcoef[s][u][m] * dt);
          }
          cp.rho[s] = cp.coef[s][0][0];
          for (int a = 0; a < D; a++) cp.rhov[s][a] = cp.coef[s][1 + a][0];
          cp.E[s] = cp.coef[s][U - 1][0];
          double sq = 0;
          for (int a = 0; a < D; a++) sq += cp.rhov[s][a] * cp.rhov[s][a];
          double e_j = cp.E[s] - 0.5 * sq / cp.rho[s];
          if (!s3())
            for (int a = 0; a < 3; a++) cp.v[a] = cp.rhov[s][a] / cp.rho[s];
          cp.P = eos_p(e_j, s);
          cp.c[s] = std::sqrt(cs.gamma[s] * (cs.pinf[s] + cp.P) / cp.rho[s]);
          for (int a = 0; a < D; a++) cp.a[a] = (cp.dcoef[s][1 + a][0] - cp.dcoef[s][0][0] * cp.v[a]) / cp.rho[s];
          cp.ac[s] = norm_d(cp.a);
        }
        for (int a = 0; a < 3; a++) rv_a[a] += cp.rhov[s][a] * cp.alph[s];
        rho_a += cp.rho[s] * cp.alph[s];
        P_a += cp.P * cp.alph[s];
      }
      for (int a = 0; a < 3; a++) cp.v[a] = rv_a[a] / rho_a;
      cp.P = P_a;
    }
    });
  }

  // ---- Update_particle: two passes (S3:mpm.cpp:839-853) ----------------------
  void update_particle(int md) {
    par_range((int64_t)fluid.size(), [&](int64_t b0, int64_t b1) {
      for (int64_t k = b0; k < b1; k++) update_state(fluid[k], md);
    });
    par_range((int64_t)fluid.size(), [&](int64_t b0, int64_t b1) {
      for (int64_t k = b0; k < b1; k++) update_pst(fluid[k]);
    });
  }
  void refresh_all(int status) {
    par_range((int64_t)fluid.size(), [&](int64_t b0, int64_t b1) {
      for (int64_t k = b0; k < b1; k++) refresh_neighbors(fluid[k], status);
    });
  }

  // Update_state_onestep (S3:particle_flow.cpp:413-477)
  void update_state(int id, int md) {
    Ptl& q = pool[id];
    const Cell& C = cells[q.in_ele];
    int ss = q.phase - 1;
    double divu = 0;
    for (int a = 0; a < D; a++) divu += C.G[a][a];
    if (md == 0) {
      for (int a = 0; a < 3; a++) q.x0[a] = q.x[a];
      q.vol0 = q.vol;
      for (int a = 0; a < D; a++) q.x[a] += q.tv[a] * dt;
      q.vol = (1. + divu * dt) * q.vol;
    } else if (md == 1) {
      for (int a = 0; a < D; a++) q.x[a] = 0.5 * ((q.x0[a] + q.x[a]) + q.tv[a] * dt);
      q.vol = 0.5 * (q.vol0 + q.vol + divu * dt * q.vol);
    } else if (md == 2) {
      for (int a = 0; a < D; a++) q.x[a] = 0.75 * q.x0[a] + 0.25 * (q.x[a] + q.tv[a] * dt);
      q.vol = 0.75 * q.vol0 + 0.25 * (q.vol + divu * dt * q.vol);
    } else {
      for (int a = 0; a < D; a++) q.x[a] = (1. / 3.) * q.x0[a] + (2. / 3.) * (q.x[a] + q.tv[a] * dt);
      q.vol = (1. / 3.) * q.vol0 + (2. / 3.) * (q.vol + divu * dt * q.vol);
    }
    double dri[3];
    for (int a = 0; a < 3; a++) dri[a] = q.x[a] - C.x[a];
    double phi[M - 1];
    basis_tail<D, ORD>(dri, q12, phi);
    double Ui[U];
    for (int t = 0; t < U; t++) {
      double acc = 0;
      for (int kk = 0; kk < M - 1; kk++) acc += C.coef[ss][t][1 + kk] * phi[kk];
      Ui[t] = C.coef[ss][t][0] + acc;
    }
    std::memcpy(q.coef, C.coef[ss], sizeof(q.coef));
    if (ORD == 2)  // the copy carries the gradient at the particle
      for (int t = 0; t < U; t++) {
        double g[3];
        grad_at<D, ORD>(C.coef[ss][t], dri, g);
        for (int a = 0; a < D; a++) q.coef[t][1 + a] = g[a];
      }
    q.rho = Ui[0];
    for (int a = 0; a < 3; a++) q.rhov[a] = 0;
    for (int a = 0; a < D; a++) q.rhov[a] = Ui[1 + a];
    q.E = Ui[U - 1];
    for (int a = 0; a < 3; a++) q.v[a] = q.rhov[a] / q.rho;
    double ei = q.E - 0.5 * q.rho * sq_d(q.v);
    q.P = eos_p(ei, ss);
    q.c = std::sqrt(cs.gamma[ss] * (cs.pinf[ss] + q.P) / q.rho);
    if (q.P < -0.05 * cs.pinf[ss] || q.rho < 0) {
      for (int u = 0; u < U; u++)
        for (int m = 1; m < M; m++) q.coef[u][m] = 0;
      q.rho = C.rho[ss];
      for (int a = 0; a < 3; a++) q.rhov[a] = C.rhov[ss][a];
      q.E = C.E[ss];
      q.c = C.c[ss];
    }
    q.mass = q.vol * q.rho;
    q.h = root(q.vol) * CELL_RATIO / CUT_OFF;
  }

  // Update_PST + PST + Bound_pen (S3:particle_flow.cpp:485-622)
  void update_pst(int id) {
    Ptl& q = pool[id];
    double pst[3] = {0, 0, 0}, sharp[3] = {0, 0, 0};
    double ca = max_v * 5.;
    double P_b = q.rho * ca * ca;
    double dxk = q.h / CELL_RATIO * CUT_OFF;
    double W_dp = W(dxk, q.h);
    for (int r : q.nb) {
      const double* nx = ref_x(r);
      double dr[3];
      for (int a = 0; a < 3; a++) dr[a] = q.x[a] - nx[a];
      double dist = norm_d(dr);
      if (dist != 0) {
        bool wl = is_wall(r);
        double vol_j = wl ? walls[wall_of(r)].vol : pool[r].vol;
        double aw = dW(dist, q.h);
        double awax[3];
        for (int a = 0; a < D; a++) awax[a] = dr[a] * aw;
        double Wij = W(dist, q.h);
        double rr = Wij / W_dp;
        double r2 = rr * rr;
        double R = 0.2 * (r2 * r2);
        double tmp1 = -1. * P_b * (1. / q.rho) * (1 + R);
        if (!wl) {
          const Ptl& nb = pool[r];
          double tmp2 = 1. / (q.rho + nb.rho);
          for (int a = 0; a < D; a++) pst[a] += awax[a] * tmp1 * vol_j;
          if (nb.phase != q.phase) {
            double k = -0.1 * tmp2 * (q.c * q.c * q.rho + nb.c * nb.c * nb.rho);
            for (int a = 0; a < D; a++) sharp[a] += k * awax[a] * vol_j;
          }
        } else {
          const Wall& w = walls[wall_of(r)];
          for (int a = 0; a < D; a++) pst[a] += w.norm[a] * Wij * tmp1 * w.vol;
        }
      }
    }
    for (int a = 0; a < 3; a++) q.tv[a] = q.v[a];
    double pen[3] = {0, 0, 0};
    {
      const Cell& C = cells[q.in_ele];
      for (int j = 0; j < NF; j++) {
        const Face& fc = C.f[j];
        if (fc.b != 0) {
          double dr[3];
          for (int a = 0; a < 3; a++) dr[a] = q.x[a] - fc.cen[a];
          double ds = 0;
          for (int a = 0; a < D; a++) ds += dr[a] * fc.nrm[a];
          double nrm[3] = {fc.nrm[0], fc.nrm[1], fc.nrm[2]};
          if (ds < 0)
            for (int a = 0; a < 3; a++) nrm[a] = -1 * nrm[a];
          double dis = std::abs(ds);
          if (dis < 0.4 * cs.p_sz) {
            double k = 0.02 / dis * (q.c * q.c);
            for (int a = 0; a < D; a++) pen[a] += k * nrm[a];
          }
        }
      }
    }
    double px = cs.p_sz;
    auto clampv = [&](double* d, double cap) {
      double dis = norm_d(d);
      double lim = AMIN1(cap, dis) / (dis + 1e-20);
      for (int a = 0; a < D; a++) d[a] *= lim;
    };
    double dx[3], dx1[3], dx2[3];
    for (int a = 0; a < D; a++) {
      dx[a] = pst[a] * dt * dt / 2.;
      dx1[a] = sharp[a] * dt * dt / 2.;
      dx2[a] = pen[a] * dt * dt / 2.;
    }
    clampv(dx, px / 10);
    clampv(dx1, px / 10);
    clampv(dx2, px / 2);
    double dv[3] = {0, 0, 0};
    for (int a = 0; a < D; a++) dv[a] = (dx[a] + dx1[a] + dx2[a]) / dt;
    if (s3() && q.mark_in == 1)
      for (int a = 0; a < 3; a++) dv[a] = 0;
    for (int a = 0; a < D; a++) q.tv[a] += dv[a];
  }

  // particle->neighbor.size() of each fluid particle, in particle order
  // (the list the last Refresh_neighbor_info built)
  int neighbor_counts(int32_t* out) override {
    for (size_t i = 0; i < fluid.size(); i++) out[i] = (int32_t)pool[fluid[i]].nb.size();
    return PGDG_OK;
  }

  // 1 for each cell in Used_Ele_cen, else 0 (cell order)
  int used_cells(int32_t* out) override {
    for (size_t ci = 0; ci < cells.size(); ci++) out[ci] = 0;
    for (int ci : used) out[ci] = 1;
    return PGDG_OK;
  }

  // Pool id -> position in particle_fluid (-1 when not a live fluid
  // particle), for the test views below.
  std::vector<int> fluid_pos() const {
    std::vector<int> pos(pool.size(), -1);
    for (size_t i = 0; i < fluid.size(); i++) pos[fluid[i]] = (int)i;
    return pos;
  }
  int map_ref(const std::vector<int>& pos, int r) const {
    if (r >= 0) return pos[r];  // a freed object maps to -1
    return r;                   // wall -(w+2) or NULL -1
  }
  // Cell lists (table_cell_list) as CSR over cells in index order; entries
  // are fluid positions, walls -(w+2).  Returns the entry count, or -1 when
  // cap is too small.
  int64_t bins_out(int32_t* off, int32_t* ent, int64_t cap) override {
    std::vector<int> pos = fluid_pos();
    int64_t k = 0;
    for (int c = 0; c < nc; c++) {
      off[c] = (int32_t)k;
      for (int r : bins[c]) {
        if (k >= cap) return -1;
        ent[k++] = map_ref(pos, r);
      }
    }
    off[nc] = (int32_t)k;
    return k;
  }
  // Neighbour list of the i-th fluid particle (particle->neighbor), entries
  // mapped as in bins_out.  Returns the list size, or -1 when cap is too
  // small or i is out of range.
  int list_out(int64_t i, int32_t* out, int cap) override {
    if (i < 0 || i >= (int64_t)fluid.size()) return -1;
    const std::vector<int>& nb = pool[fluid[i]].nb;
    if ((int)nb.size() > cap) return -1;
    std::vector<int> pos = fluid_pos();
    for (size_t t = 0; t < nb.size(); t++) out[t] = map_ref(pos, nb[t]);
    return (int)nb.size();
  }

  // ---- single-stage entry for unit tests ------------------------------------
  int call(int fn, int arg) override {
    switch (fn) {
      case 0: set_timestep(); return PGDG_OK;
      case 1: used_state_update(); return PGDG_OK;
      case 2: bin_particles(); return PGDG_OK;
      case 3: split_fast(); return PGDG_OK;
      case 4: merge_fast(); return PGDG_OK;
      case 5: membership(); return PGDG_OK;
      case 6: refresh_all(arg); return PGDG_OK;
      case 7: mls_p2g(); return PGDG_OK;
      case 8: calculate_limiter(); return PGDG_OK;
      case 9: grid_acceleration(); return PGDG_OK;
      case 10: update_grid(arg); return PGDG_OK;
      case 11: update_particle(arg); return PGDG_OK;
      case 12: run_simulation(arg); return PGDG_OK;
      default: return PGDG_ERR_ARG;
    }
  }
};

}  // namespace oracle

extern "C" {

void* oracle_create(const pgdg_case* c) {
  if (!c) return nullptr;
  if (c->nphase < 2 || c->nphase > PGDG_MAX_PHASE) return nullptr;
  if (c->order != 1 && c->order != 2) return nullptr;
  if (c->rk_stages != 2 && c->rk_stages != 3) return nullptr;
  if (c->dim == 2) return c->order == 1 ? (void*)new oracle::Sim<2, 1>(*c) : (void*)new oracle::Sim<2, 2>(*c);
  if (c->dim == 3) return c->order == 1 ? (void*)new oracle::Sim<3, 1>(*c) : (void*)new oracle::Sim<3, 2>(*c);
  return nullptr;
}
void oracle_destroy(void* h) { delete static_cast<oracle::Base*>(h); }
int oracle_load(void* h, const pgdg_particles* p, const pgdg_walls* w, const pgdg_cells* c,
                const pgdg_scalars* s) {
  return static_cast<oracle::Base*>(h)->load(p, w, c, s);
}
int oracle_save(void* h, pgdg_particles* p, pgdg_cells* c, pgdg_scalars* s) {
  return static_cast<oracle::Base*>(h)->save(p, c, s);
}
int oracle_step(void* h, int nsteps) {
  for (int i = 0; i < nsteps; i++) {
    int r = static_cast<oracle::Base*>(h)->step();
    if (r) return r;
  }
  return 0;
}
int oracle_call(void* h, int fn, int arg) { return static_cast<oracle::Base*>(h)->call(fn, arg); }
int oracle_neighbor_counts(void* h, int32_t* out) {
  return static_cast<oracle::Base*>(h)->neighbor_counts(out);
}
int oracle_used_cells(void* h, int32_t* out) { return static_cast<oracle::Base*>(h)->used_cells(out); }
// Worker threads of the step's per-cell and per-particle loops.
void oracle_set_threads(int n) { oracle::g_threads = n > 0 ? n : 1; }
int oracle_get_threads(void) { return oracle::g_threads; }
int64_t oracle_bins(void* h, int32_t* off, int32_t* ent, int64_t cap) {
  return static_cast<oracle::Base*>(h)->bins_out(off, ent, cap);
}
int oracle_list(void* h, int64_t i, int32_t* out, int cap) {
  return static_cast<oracle::Base*>(h)->list_out(i, out, cap);
}
// Single routines of the restatement, for the closed-form tests.
// Wendland kernel W and (dW/dr)/r at distance r, width h.
int oracle_kernel(int dim, double dist, double h, double* w, double* dw) {
  if (dim == 2) {
    *w = oracle::Sim<2, 1>::W(dist, h);
    *dw = oracle::Sim<2, 1>::dW(dist, h);
  } else if (dim == 3) {
    *w = oracle::Sim<3, 1>::W(dist, h);
    *dw = oracle::Sim<3, 1>::dW(dist, h);
  } else {
    return PGDG_ERR_ARG;
  }
  return PGDG_OK;
}
double oracle_limiter_fn(double ds, int bd_fg) { return oracle::Sim<3, 1>::limiter_fn(ds, bd_fg); }
// Face flux of two states (kind 0: HLLC of interior faces, 1: HLL of
// boundary faces).  stL/stR: rho, rho v[3], E, P, c (7 doubles); the
// velocity is rho v / rho.  F: U x D row-major (U = dim + 2).
int oracle_face_flux(int dim, int kind, const double* n, const double* stL, const double* stR,
                     double* F) {
  double vi[3] = {0, 0, 0}, vj[3] = {0, 0, 0}, rvi[3] = {0, 0, 0}, rvj[3] = {0, 0, 0};
  for (int a = 0; a < dim; a++) {
    rvi[a] = stL[1 + a];
    rvj[a] = stR[1 + a];
    vi[a] = rvi[a] / stL[0];
    vj[a] = rvj[a] / stR[0];
  }
  if (dim == 2) {
    double Fh[4][2];
    if (kind == 0)
      oracle::Sim<2, 1>::hllc(n, stL[0], rvi, stL[4], stL[5], stL[6], vi, stR[0], rvj, stR[4], stR[5],
                             stR[6], vj, Fh);
    else
      oracle::Sim<2, 1>::hll(n, stL[0], rvi, stL[4], stL[5], stL[6], vi, stR[0], rvj, stR[4], stR[5],
                            stR[6], vj, Fh);
    std::memcpy(F, Fh, sizeof(Fh));
  } else if (dim == 3) {
    double Fh[5][3];
    if (kind == 0)
      oracle::Sim<3, 1>::hllc(n, stL[0], rvi, stL[4], stL[5], stL[6], vi, stR[0], rvj, stR[4], stR[5],
                             stR[6], vj, Fh);
    else
      oracle::Sim<3, 1>::hll(n, stL[0], rvi, stL[4], stL[5], stL[6], vi, stR[0], rvj, stR[4], stR[5],
                            stR[6], vj, Fh);
    std::memcpy(F, Fh, sizeof(Fh));
  } else {
    return PGDG_ERR_ARG;
  }
  return PGDG_OK;
}
void oracle_root_dim(const double* x, double* y, int64_t n, int dim) {
  for (int64_t i = 0; i < n; i++) y[i] = oracle::root_dim(dim, x[i]);
}

}  // extern "C"
