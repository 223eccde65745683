// mls_ls.hip - MLS particle-to-grid moment assembly with a per-cell P2
// least-squares solve (the moment form of MPM::MLS_P2G, S3:mpm.cpp:274-382).
//
// One lane per cell, 64 cells per workgroup.  The lane walks the particles
// of its 3x3x3 neighbour cells as one flattened candidate stream over 9
// contiguous rows (so lanes of a wave run nearly equal trip counts; the
// per-cell particle counts vary), with the row bounds staged in LDS and the
// next candidate loaded while the current one is summed, and accumulates the
// 55 unique entries of sum w phi phi^T and the 10 x nf entries of
// sum w phi f in registers: FP64 with absolute FP64 positions, or FP32 with
// cell-local FP32 coordinates (the mixed form; the global weight factor
// 21/(16 pi h^3) is dropped there since the solution does not depend on it;
// the FP32 sums run over one plane of 9 neighbour cells at a time and are
// added into FP64 totals, and particles within rounding of the cut-off
// radius are decided by the FP64 form's test on the FP64 positions).
// The 10x10 Cholesky factorisation and the nf triangular solves then run in
// FP64 in the same lane, fully unrolled.  A failing pivot truncates the
// system to P1 (first four modes) or P0, which is what a separate
// factorisation of the leading block gives.  Particles of a cell are
// contiguous (cell-sorted SoA); output is [nf][10][nc], coalesced over cells.
//
// Particle data are not staged in LDS: with one lane per cell each particle
// is read by the 27 cells around it, so a tile of cells needs a halo of
// particles several times its own (a 4x4x4 tile reads 6x6x6 cells' worth),
// more LDS per wave than the occupancy allows; the reads hit L2 instead.
#include <type_traits>

#include "dev_common.hpp"
#include "hip_util.hpp"

namespace pgdg {

constexpr int kNA = 55;        // unique entries of the symmetric moment matrix
constexpr int kNR = kNA + 60;  // + right-hand sides (10 modes x 6 fields)

__host__ __device__ constexpr int tri(int a, int b) { return a * 10 - (a * (a - 1)) / 2 + (b - a); }

template <typename T>
__device__ inline void basis10(T x, T y, T z, T* phi) {
  const T c = T(1.0 / 12.0);
  phi[0] = T(1);
  phi[1] = x;
  phi[2] = y;
  phi[3] = z;
  phi[4] = x * x - c;
  phi[5] = y * y - c;
  phi[6] = z * z - c;
  phi[7] = x * y;
  phi[8] = x * z;
  phi[9] = y * z;
}

// acc[tri(a, b)] (a <= b) += (w phi_b) phi_a and acc[kNA + 6a + f] +=
// (w phi_a) v_f.  FP64: separate multiply and add, in the oracle's operand
// order, so the FP64 form reproduces it bit for bit; FP32: fused.
template <typename T>
__device__ inline void accumulate(T* acc, T w, const T* phi, const T* v, int nf) {
  constexpr bool F32 = std::is_same<T, float>::value;
  T wphi[10];
#pragma unroll
  for (int a = 0; a < 10; a++) wphi[a] = w * phi[a];
#pragma unroll
  for (int a = 0; a < 10; a++)
#pragma unroll
    for (int b = a; b < 10; b++) {
      T& r = acc[tri(a, b)];
      if constexpr (F32) r = fmaf(wphi[b], phi[a], r);
      else r = r + wphi[b] * phi[a];
    }
#pragma unroll
  for (int f = 0; f < 6; f++)
    if (f < nf)
#pragma unroll
      for (int a = 0; a < 10; a++) {
        T& r = acc[kNA + a * 6 + f];
        if constexpr (F32) r = fmaf(wphi[a], v[f], r);
        else r = r + wphi[a] * v[f];
      }
}

// Workgroup size of the fit (one lane per cell).
constexpr int kMlsLanes = 64;

// The 3x3x3 neighbour cells of a cell as 9 rows (di, dj) of up to 3 cells
// along the fastest axis; the cells of a row are consecutive, so their
// particles are one contiguous range [b0, b3) with the middle cell's
// particles at [b1, b2).  Each lane keeps its 9 rows in LDS (b0 and the
// three end offsets relative to it), so the walk moves from row to row
// without a dependent global load.
struct RowTable {
  long long b0[9][kMlsLanes];
  int e1[9][kMlsLanes], e2[9][kMlsLanes], e3[9][kMlsLanes];
};

// One candidate particle of the walk, loaded one step ahead of its use.
template <bool MIXED>
struct Cand;
template <>
struct Cand<true> {
  float x, y, z, vol, v[6];  // cell-scaled coordinates relative to the cell centre
  long long k;
};
template <>
struct Cand<false> {
  double x, y, z, vol, v[6];  // absolute positions
};

// Candidate walk and moment sums of cell c (acc[kNR] in the accumulator
// type of the form).  The walk visits the rows in order (di, dj) and each
// row's particles in order, which is the order of the neighbour cells
// (di, dj, dk) and of their particles; the next candidate's data is loaded
// while the current one is summed.  Mixed form (FOLD): acc holds the FP32
// sums over the particles of one plane (fixed di) of neighbour cells, which
// are added into the FP64 totals tot[kNR] when the walk leaves the plane, so
// that no FP32 sum runs over more than one plane's particles.
template <bool MIXED, bool FOLD, typename T>
__device__ inline void mls_assemble(const pgdg_mls_problem& p, int64_t c, RowTable& rt, T* acc,
                                    double* tot) {
  const int n0 = p.n[0], n1 = p.n[1], n2 = p.n[2];
  const int nf = p.nfields;
  const int lane = threadIdx.x;
  const int ck = (int)(c % n2), cj = (int)((c / n2) % n1), ci = (int)(c / ((int64_t)n1 * n2));
#pragma unroll
  for (int q = 0; q < kNR; q++) acc[q] = T(0);
  if constexpr (FOLD)
#pragma unroll
    for (int q = 0; q < kNR; q++) tot[q] = 0.0;
  const double inv_dx = 1.0 / p.dx;
  const double Ck = 1.0 / kPiRef / ((p.h * p.h) * p.h) * 21. / 16.;
  const double xc0 = p.dx * (ci + 0.5), xc1 = p.dx * (cj + 0.5), xc2 = p.dx * (ck + 0.5);
  const float cut_xi2 = (float)((p.cut * inv_dx) * (p.cut * inv_dx));
  const float dx_h = (float)(p.dx / p.h);

  // row table
#pragma unroll
  for (int r = 0; r < 9; r++) {
    const int t = ci + r / 3 - 1, s = cj + r % 3 - 1;
    long long b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    if (t >= 0 && t < n0 && s >= 0 && s < n1) {
      const int64_t base = ((int64_t)t * n1 + s) * n2 + ck;
      b1 = p.off[base];
      b2 = p.off[base + 1];
      b0 = ck > 0 ? p.off[base - 1] : b1;
      b3 = ck + 1 < n2 ? p.off[base + 2] : b2;
    }
    rt.b0[r][lane] = b0;
    rt.e1[r][lane] = (int)(b1 - b0);
    rt.e2[r][lane] = (int)(b2 - b0);
    rt.e3[r][lane] = (int)(b3 - b0);
  }

  int r = -1;
  long long k = 0, kend = 0, b1 = 0, b2 = 0;
  auto next_row = [&]() {
    while (k >= kend) {
      if (++r >= 9) return;
      const long long b0 = rt.b0[r][lane];
      k = b0;
      b1 = b0 + rt.e1[r][lane];
      b2 = b0 + rt.e2[r][lane];
      kend = b0 + rt.e3[r][lane];
    }
  };
  auto load = [&](Cand<MIXED>& q) {
    if constexpr (MIXED) {
      const int dk = (k >= b1) + (k >= b2) - 1;
      q.x = float(r / 3 - 1) - 0.5f + p.u[k];
      q.y = float(r % 3 - 1) - 0.5f + p.u[p.np + k];
      q.z = float(dk) - 0.5f + p.u[2 * p.np + k];
      q.vol = p.volf[k];
#pragma unroll
      for (int f = 0; f < 6; f++) q.v[f] = f < nf ? p.valf[f * p.np + k] : 0.f;
      q.k = k;
    } else {
      q.x = p.x[k];
      q.y = p.x[p.np + k];
      q.z = p.x[2 * p.np + k];
      q.vol = p.vol[k];
#pragma unroll
      for (int f = 0; f < 6; f++) q.v[f] = f < nf ? p.val[f * p.np + k] : 0.0;
    }
  };

  Cand<MIXED> nx;
  next_row();
  if (r < 9) load(nx);
  while (r < 9) {
    const Cand<MIXED> cu = nx;
    const int rc = r;
    ++k;
    next_row();
    if (r < 9) load(nx);
    if constexpr (MIXED) {
      const float r2 = cu.x * cu.x + cu.y * cu.y + cu.z * cu.z;
      bool in = r2 < cut_xi2;
      // within rounding of the cut-off sphere: the FP64 form's test on the
      // FP64 positions (when given), so that both forms fit the same
      // particles (the kernel is not small at the cut-off, and a particle
      // taken by one form and not the other changes the fit by its weight)
      if (p.x && fabsf(r2 - cut_xi2) <= 1e-5f * cut_xi2) {
        const double d0 = p.x[cu.k] - xc0, d1 = p.x[p.np + cu.k] - xc1, d2 = p.x[2 * p.np + cu.k] - xc2;
        in = sqrt(d0 * d0 + d1 * d1 + d2 * d2) < p.cut;
      }
      if (in) {
        const float sk = sqrtf(r2) * dx_h;
        const float qq = 1.f - 0.5f * sk;
        const float w = sk < 2.f ? (((qq * qq) * qq) * qq) * (1.f + 2.f * sk) * cu.vol : 0.f;
        float phi[10];
        basis10(cu.x, cu.y, cu.z, phi);
        accumulate(acc, w, phi, cu.v, nf);
      }
      if constexpr (FOLD)
        if (rc / 3 != (r < 9 ? r / 3 : 3)) {  // plane done: its FP32 sums into the totals
#pragma unroll
          for (int q = 0; q < kNR; q++) {
            tot[q] += (double)acc[q];
            acc[q] = T(0);
          }
        }
    } else {
      const double d0 = cu.x - xc0, d1 = cu.y - xc1, d2 = cu.z - xc2;
      const double rr = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      if (rr < p.cut) {
        const double sk = rr / p.h;
        double w = 0;
        if (sk < 2) {
          const double qq = 1. - 0.5 * sk;
          w = Ck * (((qq * qq) * qq) * qq) * (1. + 2. * sk);
        }
        w *= cu.vol;
        double phi[10];
        basis10(d0 / p.dx, d1 / p.dx, d2 / p.dx, phi);
        accumulate(acc, w, phi, cu.v, nf);
      }
    }
  }
}

// FP64 factorisation and solves of cell c from its sums (sum(q) returns
// entry q as FP64); writes coef and the order used.
template <typename Sum>
__device__ inline void mls_solve(const pgdg_mls_problem& p, int64_t c, int64_t nc, Sum sum,
                                 double* __restrict__ coef, int32_t* __restrict__ ord) {
  const int nf = p.nfields;
  // FP64 Cholesky of the packed matrix in place: L(i, j), i >= j, at tri(j, i)
  double A[kNA];
#pragma unroll
  for (int q = 0; q < kNA; q++) A[q] = sum(q);
  double amax = 0;
#pragma unroll
  for (int a = 0; a < 10; a++) amax = fmax(amax, fabs(A[tri(a, a)]));
  const double tol = 1e-12 * amax;
  int fail = 10;
#pragma unroll
  for (int j = 0; j < 10; j++) {
    double d = A[tri(j, j)];
#pragma unroll
    for (int q = 0; q < j; q++) d -= A[tri(q, j)] * A[tri(q, j)];
    if (!(d > tol) && fail == 10) fail = j;
    const double l = d > tol ? sqrt(d) : 1.0;
    A[tri(j, j)] = l;
#pragma unroll
    for (int i = j + 1; i < 10; i++) {
      double s = A[tri(j, i)];
#pragma unroll
      for (int q = 0; q < j; q++) s -= A[tri(q, i)] * A[tri(q, j)];
      A[tri(j, i)] = s / l;
    }
  }
  const int used = fail >= 10 ? 2 : fail >= 4 ? 1 : fail >= 1 ? 0 : -1;
  const int o = used < p.order ? used : p.order;
  const int m = o == 2 ? 10 : o == 1 ? 4 : o == 0 ? 1 : 0;
  if (ord) ord[c] = o;
#pragma unroll
  for (int f = 0; f < 6; f++) {
    if (f >= nf) break;
    double b[10];
#pragma unroll
    for (int i = 0; i < 10; i++) {
      double s = sum(kNA + i * 6 + f);
#pragma unroll
      for (int q = 0; q < i; q++) s -= A[tri(q, i)] * b[q];
      b[i] = i < m ? s / A[tri(i, i)] : 0.0;
    }
#pragma unroll
    for (int i = 9; i >= 0; i--) {
      double s = b[i];
#pragma unroll
      for (int q = i + 1; q < 10; q++)
        if (q < m) s -= A[tri(i, q)] * b[q];
      b[i] = i < m ? s / A[tri(i, i)] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 10; i++) coef[((int64_t)f * 10 + i) * nc + c] = b[i];
  }
}

template <bool MIXED>
__global__ void __launch_bounds__(kMlsLanes) k_mls_ls(pgdg_mls_problem p, double* __restrict__ coef,
                                                      int32_t* __restrict__ ord) {
  using T = typename std::conditional<MIXED, float, double>::type;
  __shared__ RowTable rt;
  const int64_t nc = (int64_t)p.n[0] * p.n[1] * p.n[2];
  const int64_t c = (int64_t)blockIdx.x * kMlsLanes + threadIdx.x;
  if (c >= nc) return;
  T acc[kNR];
  if constexpr (MIXED) {
    double tot[kNR];
    mls_assemble<MIXED, true>(p, c, rt, acc, tot);
    mls_solve(p, c, nc, [&](int q) { return tot[q]; }, coef, ord);
  } else {
    mls_assemble<MIXED, false>(p, c, rt, acc, nullptr);
    mls_solve(p, c, nc, [&](int q) { return (double)acc[q]; }, coef, ord);
  }
}

}  // namespace pgdg

namespace {
int mls_check(const pgdg_mls_problem* p, int precision, double* coef) {
  if (!p || !coef) return PGDG_ERR_ARG;
  if (p->nfields < 1 || p->nfields > 6 || p->order < 0 || p->order > 2) return PGDG_ERR_ARG;
  if (p->n[0] < 1 || p->n[1] < 1 || p->n[2] < 1 || !(p->dx > 0) || !(p->h > 0) || !(p->cut > 0))
    return PGDG_ERR_ARG;
  if (!p->off || p->np < 0) return PGDG_ERR_ARG;
  const int64_t nc = (int64_t)p->n[0] * p->n[1] * p->n[2];
  if (nc > (int64_t)INT32_MAX * 64) return PGDG_ERR_ARG;
  if (precision == PGDG_PREC_FP64) {
    if (!p->x || !p->vol || !p->val) return PGDG_ERR_ARG;
  } else if (precision == PGDG_PREC_MIXED) {
    if (!p->u || !p->volf || !p->valf) return PGDG_ERR_ARG;
  } else {
    return PGDG_ERR_ARG;
  }
  return PGDG_OK;
}
}  // namespace

extern "C" int pgdg_mls_ls_fit(const pgdg_mls_problem* p, int precision, double* coef,
                               int32_t* order_used, void* stream) {
  using namespace pgdg;
  PGDG_TRY(mls_check(p, precision, coef));
  const int64_t nc = (int64_t)p->n[0] * p->n[1] * p->n[2];
  const unsigned grid = (unsigned)((nc + kMlsLanes - 1) / kMlsLanes);
  if (precision == PGDG_PREC_FP64)
    hipLaunchKernelGGL(k_mls_ls<false>, dim3(grid), dim3(kMlsLanes), 0, as_stream(stream), *p, coef,
                       order_used);
  else
    hipLaunchKernelGGL(k_mls_ls<true>, dim3(grid), dim3(kMlsLanes), 0, as_stream(stream), *p, coef,
                       order_used);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}
