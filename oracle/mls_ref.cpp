// mls_ref.cpp - FP64 oracle of the MLS particle-to-grid moment assembly with
// a per-cell least-squares solve (TEST INFRASTRUCTURE ONLY; loaded by the
// tests and by bench.py's checker, never by the product).
//
// The reference's MLS_P2G (S3:mpm.cpp:274-382) forms, for every cell, a
// Wendland-weighted average of the particles of the 3^D neighbour cells
// within the P2G cut-off (weight W(r, 1.6 p_g) * vol, S3:mpm.cpp:309-310).
// The moment form generalises it: with the cell basis phi (P2, zero-mean
// quadratics in the cell-scaled coordinate xi = (x - x_c)/dx), minimise
//   sum_p w_p (f_p - sum_m c_m phi_m(xi_p))^2
// i.e. solve (sum w phi phi^T) c = sum w phi f.  With one mode it is exactly
// the reference's average sum(w f)/sum(w).
//
// Degenerate moment matrices (a pivot below 1e-12 of the largest diagonal)
// fall back to the P1 system, then to P0.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

static const double PI_REF = 3.1415926;

// P2 zero-mean basis in xi (3-D): 1, x, y, z, x^2-1/12, y^2-1/12, z^2-1/12,
// xy, xz, yz.  order 0 -> first mode, 1 -> first 4, 2 -> all 10.
inline void basis_p2(const double* xi, double* phi) {
  phi[0] = 1.0;
  phi[1] = xi[0];
  phi[2] = xi[1];
  phi[3] = xi[2];
  phi[4] = xi[0] * xi[0] - 1.0 / 12.0;
  phi[5] = xi[1] * xi[1] - 1.0 / 12.0;
  phi[6] = xi[2] * xi[2] - 1.0 / 12.0;
  phi[7] = xi[0] * xi[1];
  phi[8] = xi[0] * xi[2];
  phi[9] = xi[1] * xi[2];
}

// Cholesky of the leading m x m block of A (row-major 10x10 scratch);
// returns false on a pivot <= tol.
static bool chol(double* A, int m, double tol) {
  for (int j = 0; j < m; j++) {
    double d = A[j * 10 + j];
    for (int k = 0; k < j; k++) d -= A[j * 10 + k] * A[j * 10 + k];
    if (!(d > tol)) return false;
    double l = std::sqrt(d);
    A[j * 10 + j] = l;
    for (int i = j + 1; i < m; i++) {
      double s = A[i * 10 + j];
      for (int k = 0; k < j; k++) s -= A[i * 10 + k] * A[j * 10 + k];
      A[i * 10 + j] = s / l;
    }
  }
  return true;
}

static void chol_solve(const double* L, int m, double* b) {
  for (int i = 0; i < m; i++) {
    double s = b[i];
    for (int k = 0; k < i; k++) s -= L[i * 10 + k] * b[k];
    b[i] = s / L[i * 10 + i];
  }
  for (int i = m - 1; i >= 0; i--) {
    double s = b[i];
    for (int k = i + 1; k < m; k++) s -= L[k * 10 + i] * b[k];
    b[i] = s / L[i * 10 + i];
  }
}

// Solve the normal equations of one cell: A (10x10 symmetric, full), B
// (10 x nf).  Writes coef[nf][10]; returns the order used (2, 1 or 0).
int solve_cell(const double* A, const double* B, int nf, int order, double* coef) {
  static const int msize[3] = {1, 4, 10};
  double amax = 0;
  for (int i = 0; i < 10; i++) amax = std::fmax(amax, std::fabs(A[i * 10 + i]));
  for (int o = order; o >= 0; o--) {
    int m = msize[o];
    double L[100];
    std::memcpy(L, A, sizeof(L));
    if (!chol(L, m, 1e-12 * amax)) continue;
    for (int f = 0; f < nf; f++) {
      double b[10];
      for (int i = 0; i < m; i++) b[i] = B[i * 6 + f];
      chol_solve(L, m, b);
      for (int i = 0; i < 10; i++) coef[f * 10 + i] = i < m ? b[i] : 0.0;
    }
    return o;
  }
  for (int f = 0; f < nf; f++)
    for (int i = 0; i < 10; i++) coef[f * 10 + i] = 0.0;
  return -1;
}

}  // namespace oracle

extern "C" {

// n[3] cells on [0, n*dx]; particles sorted by cell (cell index
// (i*n1 + j)*n2 + k), offsets off[nc+1]; x[3][np] absolute positions,
// vol[np], val[nf][np].  Output coef[nf][10][nc]; used order per cell in
// ord[nc] (may be NULL).  cut: particle cut-off, h: kernel width.
int oracle_mls_ls(const int* n, double dx, const int64_t* off, const double* x, const double* vol,
                  const double* val, int64_t np, int nf, int order, double cut, double h,
                  double* coef, int32_t* ord) {
  using namespace oracle;
  const int64_t nc = (int64_t)n[0] * n[1] * n[2];
  const double C = 1.0 / PI_REF / ((h * h) * h) * 21. / 16.;
  for (int i = 0; i < n[0]; i++)
    for (int j = 0; j < n[1]; j++)
      for (int k = 0; k < n[2]; k++) {
        const int64_t c = ((int64_t)i * n[1] + j) * n[2] + k;
        const double xc[3] = {dx * (i + 0.5), dx * (j + 0.5), dx * (k + 0.5)};
        double A[100], B[60];
        std::memset(A, 0, sizeof(A));
        std::memset(B, 0, sizeof(B));
        for (int t = (i > 0 ? i - 1 : 0); t <= (i + 1 < n[0] ? i + 1 : n[0] - 1); t++)
          for (int s = (j > 0 ? j - 1 : 0); s <= (j + 1 < n[1] ? j + 1 : n[1] - 1); s++)
            for (int q = (k > 0 ? k - 1 : 0); q <= (k + 1 < n[2] ? k + 1 : n[2] - 1); q++) {
              const int64_t cc = ((int64_t)t * n[1] + s) * n[2] + q;
              for (int64_t p = off[cc]; p < off[cc + 1]; p++) {
                double d[3], r2 = 0;
                for (int a = 0; a < 3; a++) {
                  d[a] = x[a * np + p] - xc[a];
                  r2 += d[a] * d[a];
                }
                double r = std::sqrt(r2);
                if (!(r < cut)) continue;
                double sk = r / h;
                double w = 0;
                if (sk < 2) {
                  double qq = 1. - 0.5 * sk;
                  w = C * (((qq * qq) * qq) * qq) * (1. + 2. * sk);
                }
                w *= vol[p];
                double xi[3] = {d[0] / dx, d[1] / dx, d[2] / dx};
                double phi[10];
                basis_p2(xi, phi);
                for (int a = 0; a < 10; a++)
                  for (int b = 0; b < 10; b++) A[a * 10 + b] += w * phi[a] * phi[b];
                for (int a = 0; a < 10; a++)
                  for (int f = 0; f < nf; f++) B[a * 6 + f] += w * phi[a] * val[f * np + p];
              }
            }
        double cf[60];
        int o = solve_cell(A, B, nf, order, cf);
        if (ord) ord[c] = o;
        for (int f = 0; f < nf; f++)
          for (int m = 0; m < 10; m++) coef[((int64_t)f * 10 + m) * nc + c] = cf[f * 10 + m];
      }
  return 0;
}

}  // extern "C"
