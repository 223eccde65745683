// output.cpp - the reference's output and diagnostics (host code, no GPU):
//   MPM::Output_plt_file  S3:mpm_ini.cpp:832-980, T2:mpm_ini.cpp:662-840
//     particle file: Tecplot binary (#!TDV112), one zone, point data in
//     block order (each variable for every fluid particle in turn);
//     cell file: Tecplot ASCII, one line per cell in i, j(, k) order;
//   MPM::Output_point     S3:mpm_ini.cpp:774-827, T2:mpm_ini.cpp:599-654
//     S3: the phase-2 front positions and mass; T2: particle and grid mass
//     and energy sums; each appended as one text line per call.
// The file contents follow the reference byte for byte: the header words,
// the variable names as one 32-bit word per character, the value types
// (double, 32-bit integer) and the text formatting of C++ streams.
#include <stdint.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "pgdg.h"

namespace {

template <class T>
void put(std::ofstream& out, T v) {
  out.write(reinterpret_cast<const char*>(&v), sizeof(v));
}

// A name as one 32-bit word per character, then a zero word.
void put_name(std::ofstream& out, const std::string& s) {
  for (char ch : s) put<int32_t>(out, (int32_t)(unsigned char)ch);
  put<int32_t>(out, 0);
}

double norm_d(const double* v, int d) {  // the reference's head(DIM).norm()
  double s = 0;
  for (int a = 0; a < d; a++) s += v[a] * v[a];
  return std::sqrt(s);
}

}  // namespace

extern "C" {

int pgdg_plt_write_particles(const char* path, const pgdg_case* c, const pgdg_particles* p,
                             const int32_t* n_neighbor, double run_time) {
  if (!path || !c || !p || !n_neighbor || p->n < 0 || p->n > p->cap || (c->dim != 2 && c->dim != 3))
    return PGDG_ERR_ARG;
  const int D = c->dim;
  std::ofstream out(path, std::ios::binary);
  if (!out) return PGDG_ERR_ARG;
  out.write("#!TDV112", 8);
  put<int32_t>(out, 1);
  put<int32_t>(out, 0);
  put<int32_t>(out, 0);
  const std::vector<std::string> names =
      D == 3 ? std::vector<std::string>{"x", "y", "z", "phase", "vx", "vy", "vol", "rho", "P", "tv", "v",
                                        "n_neighbor"}
             : std::vector<std::string>{"x", "y", "phase", "vx", "vy", "vol", "rho", "P", "tv", "v",
                                        "n_neighbor"};
  // Tecplot value types: 2 double, 3 32-bit integer
  std::vector<int32_t> types;
  for (const auto& s : names) types.push_back(s == "phase" || s == "n_neighbor" ? 3 : 2);
  put<int32_t>(out, (int32_t)names.size());
  for (const auto& s : names) put_name(out, s);
  put<float>(out, 299.0f);  // zone marker
  put_name(out, "grid");
  put<int32_t>(out, -1);  // parent zone
  put<int32_t>(out, 0);   // strand
  put<double>(out, run_time);
  put<int32_t>(out, -1);  // zone colour
  put<int32_t>(out, 0);   // ordered zone
  put<int32_t>(out, 0);   // no variable location list
  put<int32_t>(out, 0);   // no face neighbours
  put<int32_t>(out, 0);   // no miscellaneous connections
  put<int32_t>(out, (int32_t)p->n);  // IMAX
  put<int32_t>(out, 1);              // JMAX
  put<int32_t>(out, 1);              // KMAX
  put<int32_t>(out, 0);              // no auxiliary data
  put<float>(out, 357.0f);           // end of header
  put<float>(out, 299.0f);           // zone data
  for (int32_t t : types) put<int32_t>(out, t);
  put<int32_t>(out, 0);   // no passive variables
  put<int32_t>(out, 0);   // no variable sharing
  put<int32_t>(out, -1);  // no connectivity sharing
  for (size_t j = 0; j < names.size(); j++) {
    put<double>(out, -1.0);  // the reference writes fixed min/max words
    put<double>(out, 1.0);
  }
  const int64_t cap = p->cap, n = p->n;
  for (const auto& s : names) {
    for (int64_t i = 0; i < n; i++) {
      if (s == "x") put<double>(out, p->x[i]);
      else if (s == "y") put<double>(out, p->x[cap + i]);
      else if (s == "z") put<double>(out, p->x[2 * cap + i]);
      else if (s == "phase") put<int32_t>(out, p->phase[i]);
      else if (s == "vx") put<double>(out, p->v[i]);
      else if (s == "vy") put<double>(out, p->v[cap + i]);
      else if (s == "vol") put<double>(out, p->vol[i]);
      else if (s == "rho") put<double>(out, p->rho[i]);
      else if (s == "P") put<double>(out, p->P[i]);
      else if (s == "tv" || s == "v") {
        const double* src = s == "tv" ? p->tv : p->v;
        const double w[3] = {src[i], src[cap + i], src[2 * cap + i]};
        put<double>(out, norm_d(w, D));
      } else {
        put<int32_t>(out, n_neighbor[i]);
      }
    }
  }
  return out.good() ? PGDG_OK : PGDG_ERR_ARG;
}

int pgdg_plt_write_cells(const char* path, const pgdg_case* c, const pgdg_cells* cl, double run_time) {
  if (!path || !c || !cl || (c->dim != 2 && c->dim != 3) || c->nphase < 2) return PGDG_ERR_ARG;
  const int D = c->dim;
  const int n0 = c->n[0], n1 = c->n[1], n2 = D == 3 ? c->n[2] : 1;
  const int64_t nc = (int64_t)n0 * n1 * n2;
  const int U = D + 2, M = c->order == 1 ? D + 1 : (D == 2 ? 6 : 10);
  double ps[3];
  for (int a = 0; a < 3; a++) ps[a] = a < D ? (c->box_r[a] - c->box_l[a]) / (a == 0 ? n0 : a == 1 ? n1 : n2) : 0;
  std::ofstream out(path, std::ios::trunc);
  if (!out) return PGDG_ERR_ARG;
  // one line: the reference's string literal continues across a source line
  // (spliced), leaving eight spaces before "E0"
  const bool ux = c->semantics == PGDG_SEM_UX;
  out << "VARIABLES = \"x\",\"y\",\"z\", \"v1\",\"v2\", \"v\",\"rho0\",\"rho1\", "
         "       \"E0\",\"E1\",\"Lim0\",\"Lim1\",\"st\", \"st1\", "
      << (ux ? "\"ca\", \"P\"\n" : "\"P\"\n");
  out << "ZONE T=" << "\"GRID 1\"," << "SOLUTIONTIME=" << run_time << "\n";
  for (int i = 0; i < n0; i++)
    for (int j = 0; j < n1; j++)
      for (int k = 0; k < n2; k++) {
        const int64_t ci = ((int64_t)i * n1 + j) * n2 + k;
        // the cell-centre particle's coordinates (S3:mpm_ini.cpp:307-310)
        const double x[3] = {ps[0] * (i + 0.5), ps[1] * (j + 0.5), D == 3 ? ps[2] * (k + 0.5) : 0.0};
        const double v[3] = {cl->v[ci], cl->v[nc + ci], cl->v[2 * nc + ci]};
        out << x[0] << ' ' << x[1] << ' ' << x[2] << ' ' << v[0] << ' ' << v[1] << ' ' << norm_d(v, D) << ' '
            << cl->rho[ci] << ' ' << cl->rho[nc + ci] << ' ';
        if (D == 3) {
          // S3: the x-momentum row's x and y slopes, the density limiters
          out << cl->coef[((size_t)1 * M + 1) * nc + ci] << ' ' << cl->coef[((size_t)1 * M + 2) * nc + ci] << ' '
              << cl->lim[0 * nc + ci] << ' ' << cl->lim[(size_t)U * nc + ci] << ' ';
        } else if (ux) {
          // UX: phase 2's y-momentum x and y slopes, the energy limiters
          // (UX:mpm_ini.cpp:781-789)
          out << cl->coef[((size_t)(1 * U + 2) * M + 1) * nc + ci] << ' '
              << cl->coef[((size_t)(1 * U + 2) * M + 2) * nc + ci] << ' ' << cl->lim[(size_t)3 * nc + ci] << ' '
              << cl->lim[((size_t)U + 3) * nc + ci] << ' ';
        } else {
          // T2: each phase's speed, the energy limiters
          double s[2];
          for (int f = 0; f < 2; f++) {
            double w[3];
            for (int a = 0; a < 3; a++) w[a] = cl->rhov[((size_t)f * 3 + a) * nc + ci] / cl->rho[f * nc + ci];
            s[f] = norm_d(w, D);
          }
          out << s[0] << ' ' << s[1] << ' ' << cl->lim[(size_t)3 * nc + ci] << ' '
              << cl->lim[((size_t)U + 3) * nc + ci] << ' ';
        }
        out << cl->alph[ci] << ' ' << cl->alph[nc + ci] << ' ';
        if (ux) out << 0.0 << ' ';  // the cell's ca, never assigned (UX:particle.h:58)
        out << cl->P[ci] << ' ' << std::endl;
      }
  return out.good() ? PGDG_OK : PGDG_ERR_ARG;
}

int pgdg_point_diagnostics(const pgdg_case* c, const pgdg_particles* p, const pgdg_cells* cl,
                           const int32_t* used, double time, pgdg_point_diag* o) {
  if (!c || !p || !cl || !used || !o || p->n < 0 || p->n > p->cap || (c->dim != 2 && c->dim != 3) ||
      c->nphase < 2)
    return PGDG_ERR_ARG;
  const int D = c->dim, F = c->nphase;
  const int64_t cap = p->cap, n = p->n;
  const int n0 = c->n[0], n1 = c->n[1], n2 = D == 3 ? c->n[2] : 1;
  const int64_t nc = (int64_t)n0 * n1 * n2;
  *o = pgdg_point_diag{};
  o->time = time;
  o->n_fluid = n;
  // S3: phase-2 particles near the axis (y, z < 2.5 p_sz) and away from it
  o->dw = 0;
  o->jet = o->ups = c->box_r[0];
  for (int64_t i = 0; i < n; i++) {
    if (p->phase[i] != 2) continue;
    const double x = p->x[i], y = p->x[cap + i], z = p->x[2 * cap + i];
    if (y < 2.5 * c->p_sz && z < 2.5 * c->p_sz) {
      if (x > o->dw) o->dw = x;
      if (x < o->jet) o->jet = x;
    } else if (y > 4 * c->p_sz || z > 4 * c->p_sz) {
      if (x < o->ups) o->ups = x;
    }
    o->mass2 += p->rho[i] * p->vol[i];
  }
  // T2: particle sums, then grid sums over the cells in use (ascending)
  for (int64_t i = 0; i < n; i++) {
    if (p->phase[i] == 1) o->p_mass[0] += p->vol[i] * p->rho[i];
    else if (p->phase[i] == 2) o->p_mass[1] += p->vol[i] * p->rho[i];
    o->p_energy += p->vol[i] * p->E[i];
  }
  double cvol = 1;
  for (int a = 0; a < D; a++) cvol *= (c->box_r[a] - c->box_l[a]) / (a == 0 ? n0 : a == 1 ? n1 : n2);
  int64_t n_used = 0;
  for (int64_t ci = 0; ci < nc; ci++) {
    if (!used[ci]) continue;
    n_used++;
    o->g_mass[0] += cvol * cl->rho[ci] * cl->alph[ci];
    o->g_mass[1] += cvol * cl->rho[nc + ci] * cl->alph[nc + ci];
    for (int f = 0; f < F; f++) o->g_energy += cvol * cl->E[f * nc + ci] * cl->alph[f * nc + ci];
  }
  o->p2g = (double)n / (double)n_used;
  return PGDG_OK;
}

int pgdg_point_append(const char* dir, const pgdg_case* c, const pgdg_point_diag* o, int first) {
  if (!dir || !c || !o) return PGDG_ERR_ARG;
  const std::ios::openmode mode = first ? std::ios::trunc : (std::ios::out | std::ios::ate | std::ios::app);
  const std::string d(dir);
  if (c->semantics == PGDG_SEM_S3) {
    std::ofstream out(d + "/pointa.txt", mode);
    out << o->time << " " << std::setprecision(6) << o->dw << " " << o->jet << " " << o->ups << "\n";
    std::ofstream out1(d + "/mass.txt", mode);
    out1 << o->time << " " << std::setprecision(6) << o->mass2 << " " << (int)o->n_fluid << "\n";
    return out.good() && out1.good() ? PGDG_OK : PGDG_ERR_ARG;
  }
  std::ofstream out(d + "/point_p.txt", mode);
  out << o->time << " " << std::setprecision(6) << o->p_mass[0] << " " << o->p_mass[1] << " " << o->p_energy
      << " " << o->p2g << " " << "\n";
  std::ofstream out1(d + "/point_g.txt", mode);
  out1 << o->time << " " << std::setprecision(6) << o->g_mass[0] << " " << o->g_mass[1] << " " << o->g_energy
       << " " << "\n";
  return out.good() && out1.good() ? PGDG_OK : PGDG_ERR_ARG;
}

}  // extern "C"
