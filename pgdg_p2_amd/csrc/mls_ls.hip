// mls_ls.hip - MLS particle-to-grid moment assembly with a per-cell P2
// least-squares solve (the moment form of MPM::MLS_P2G, S3:mpm.cpp:274-382).
//
// One lane per cell.  The lane walks the particles of the 3x3x3 neighbour
// cells as one flattened candidate stream (so lanes of a wave run nearly
// equal trip counts; the per-cell particle counts vary) and accumulates the
// 55 unique entries of sum w phi phi^T and the 10 x nf entries of
// sum w phi f in registers: FP64 with absolute FP64 positions, or FP32 with
// cell-local FP32 coordinates (the mixed form; the global weight factor
// 21/(16 pi h^3) is dropped there since the solution does not depend on it).
// The 10x10 Cholesky factorisation and the nf triangular solves then run in
// FP64 in the same lane, fully unrolled.  A failing pivot truncates the
// system to P1 (first four modes) or P0, which is what a separate
// factorisation of the leading block gives.  Particles of a cell are
// contiguous (cell-sorted SoA); output is [nf][10][nc], coalesced over cells.
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

// Candidate walk and moment sums of cell c (acc[kNR] in the accumulator
// type of the form).
template <bool MIXED, typename T>
__device__ inline void mls_assemble(const pgdg_mls_problem& p, int64_t c, T* acc) {
  const int n0 = p.n[0], n1 = p.n[1], n2 = p.n[2];
  const int nf = p.nfields;
  const int ck = (int)(c % n2), cj = (int)((c / n2) % n1), ci = (int)(c / ((int64_t)n1 * n2));
#pragma unroll
  for (int q = 0; q < kNR; q++) acc[q] = T(0);
  const double inv_dx = 1.0 / p.dx;
  const double Ck = 1.0 / kPiRef / ((p.h * p.h) * p.h) * 21. / 16.;
  const double xc0 = p.dx * (ci + 0.5), xc1 = p.dx * (cj + 0.5), xc2 = p.dx * (ck + 0.5);
  const float cut_xi2 = (float)((p.cut * inv_dx) * (p.cut * inv_dx));
  const float dx_h = (float)(p.dx / p.h);

  // flattened walk over the candidates of the 27 neighbour cells
  int nb = -1, di = 0, dj = 0, dk = 0;
  int64_t k = 0, kend = 0;
  for (;;) {
    while (k >= kend && nb < 27) {
      if (++nb >= 27) break;
      di = nb / 9 - 1;
      dj = (nb / 3) % 3 - 1;
      dk = nb % 3 - 1;
      const int t = ci + di, s = cj + dj, q = ck + dk;
      if (t < 0 || t >= n0 || s < 0 || s >= n1 || q < 0 || q >= n2) continue;
      const int64_t cc = ((int64_t)t * n1 + s) * n2 + q;
      k = p.off[cc];
      kend = p.off[cc + 1];
    }
    if (nb >= 27) break;
    if constexpr (MIXED) {
      const float x = float(di) - 0.5f + p.u[k];
      const float y = float(dj) - 0.5f + p.u[p.np + k];
      const float z = float(dk) - 0.5f + p.u[2 * p.np + k];
      const float r2 = x * x + y * y + z * z;
      if (r2 < cut_xi2) {
        const float sk = sqrtf(r2) * dx_h;
        const float qq = 1.f - 0.5f * sk;
        const float w = sk < 2.f ? (((qq * qq) * qq) * qq) * (1.f + 2.f * sk) * p.volf[k] : 0.f;
        float phi[10], v[6];
        basis10(x, y, z, phi);
#pragma unroll
        for (int f = 0; f < 6; f++) v[f] = f < nf ? p.valf[f * p.np + k] : 0.f;
        accumulate(acc, w, phi, v, nf);
      }
    } else {
      const double d0 = p.x[k] - xc0, d1 = p.x[p.np + k] - xc1, d2 = p.x[2 * p.np + k] - xc2;
      const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      if (r < p.cut) {
        const double sk = r / p.h;
        double w = 0;
        if (sk < 2) {
          const double qq = 1. - 0.5 * sk;
          w = Ck * (((qq * qq) * qq) * qq) * (1. + 2. * sk);
        }
        w *= p.vol[k];
        double phi[10], v[6];
        basis10(d0 / p.dx, d1 / p.dx, d2 / p.dx, phi);
#pragma unroll
        for (int f = 0; f < 6; f++) v[f] = f < nf ? p.val[f * p.np + k] : 0.0;
        accumulate(acc, w, phi, v, nf);
      }
    }
    k++;
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
__global__ void __launch_bounds__(256) k_mls_ls(pgdg_mls_problem p, double* __restrict__ coef,
                                                int32_t* __restrict__ ord) {
  using T = typename std::conditional<MIXED, float, double>::type;
  const int64_t nc = (int64_t)p.n[0] * p.n[1] * p.n[2];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  T acc[kNR];
  mls_assemble<MIXED>(p, c, acc);
  mls_solve(p, c, nc, [&](int q) { return (double)acc[q]; }, coef, ord);
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
  const unsigned grid = (unsigned)((nc + 255) / 256);
  if (precision == PGDG_PREC_FP64)
    hipLaunchKernelGGL(k_mls_ls<false>, dim3(grid), dim3(256), 0, as_stream(stream), *p, coef,
                       order_used);
  else
    hipLaunchKernelGGL(k_mls_ls<true>, dim3(grid), dim3(256), 0, as_stream(stream), *p, coef,
                       order_used);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}
