// step_mixed.hip - mixed-precision particle shifting (Particle_base::
// Update_PST / PST / Bound_pen, S3:particle_flow.cpp:485-622).
//
// The list and the pair terms follow the FP64 kernel (k_pst_binned), but
// the pair geometry and kernel values are FP32: positions are stored
// relative to the lower corner of the entry's bin cell, so a pair offset is
// (cell offset) * cell size + (difference of two in-cell offsets), both
// nearly exact in FP32.  The per-pair terms are summed in FP64.  The kernel
// derivative uses dW/(r) = -5 C q^3 / h^2 (the reference's
// -5 C q^3 s / h / r with s = r / h), and the per-target reciprocals are
// formed once.  The boundary penalty and the clamps stay FP64.
#include "dev_common.hpp"
#include "hip_util.hpp"
#include "step_kernels.hpp"

namespace pgdg {

// Bin cell of every entry (binary search over the bin offsets) and its
// FP32 records.
template <int D>
__global__ void k_bin_gather_mixed(Geo g, PtlDev P, WallDev W, StageBufs S, BinData B, MixedBins MB,
                                   const int* __restrict__ bin_off, const int* __restrict__ bin_ent,
                                   const DevScalars* __restrict__ sc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= sc->nent_dev) return;
  int lo = 0, hi = g.nc - 1;  // last cell with bin_off[c] <= e
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bin_off[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  const int c = lo;
  int ci, cj, ck;
  cell_ijk(g, c, ci, cj, ck);
  const double corner[3] = {g.box_l[0] + ci * g.ps[0], g.box_l[1] + cj * g.ps[1],
                            D == 3 ? g.box_l[2] + ck * g.ps[2] : 0.0};
  const int ref = bin_ent[e];
  B.ref[e] = ref;
  MB.ecell[e] = c;
  float ub[3] = {0, 0, 0}, u[3] = {0, 0, 0};
  float vol, rho, cs;
  int ph;
  const int64_t cap = P.cap;
  if (ref_is_wall(ref)) {
    const int w = ref_wall(ref);
#pragma unroll
    for (int a = 0; a < D; a++) ub[a] = u[a] = (float)(W.x[a * W.n + w] - corner[a]);
    vol = (float)W.vol[w];
    rho = cs = 0.f;
    ph = 10;
  } else if (ref >= 0) {
#pragma unroll
    for (int a = 0; a < D; a++) {
      ub[a] = (float)(S.xb[a * cap + ref] - corner[a]);
      u[a] = (float)(P.x[a * cap + ref] - corner[a]);
    }
    vol = (float)P.vol[ref];
    rho = (float)P.rho[ref];
    cs = (float)P.c[ref];
    ph = P.phase[ref];
  } else {  // removed entry: never within reach
    ub[0] = u[0] = INFINITY;
    vol = rho = cs = 0.f;
    ph = -1;
  }
  MB.r1[e] = make_float4(ub[0], ub[1], ub[2], __int_as_float(ph));
  MB.r2[e] = make_float4(u[0], u[1], u[2], vol);
  MB.r3[e] = make_float2(rho, cs);
}

template <int D>
__global__ void k_cell_bbox_mixed(Geo g, MixedBins MB, const int* __restrict__ bin_off) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
  float4 hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
  for (int e = bin_off[c]; e < bin_off[c + 1]; e++) {
    const float4 r = MB.r1[e];
    lo.x = fminf(lo.x, r.x);
    lo.y = fminf(lo.y, r.y);
    hi.x = fmaxf(hi.x, r.x);
    hi.y = fmaxf(hi.y, r.y);
    if (D == 3) {
      lo.z = fminf(lo.z, r.z);
      hi.z = fmaxf(hi.z, r.z);
    }
  }
  MB.box[2 * c] = lo;
  MB.box[2 * c + 1] = hi;
}

template <int D, int BS>
__global__ void __launch_bounds__(BS) k_pst_mixed(Geo g, PtlDev P, WallDev W, StageBufs S, BinData B,
                                                   MixedBins MB, const int* __restrict__ bin_off,
                                                   const DevScalars* __restrict__ sc) {
  const int64_t e = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (e >= sc->nent_dev) return;
  const int i = B.ref[e];
  if (i < 0) return;  // wall or removed entry
  const int64_t cap = P.cap;
  const double dt = sc->dt;
  const double max_v = sc->max_v;
  const int ct = MB.ecell[e];
  int cti, ctj, ctk;
  cell_ijk(g, ct, cti, ctj, ctk);
  const float4 t1 = MB.r1[e], t2 = MB.r2[e];
  const float2 t3 = MB.r3[e];
  const float ubt[3] = {t1.x, t1.y, t1.z}, ut[3] = {t2.x, t2.y, t2.z};
  const int phi = __float_as_int(t1.w);
  double xbi[3], xi[3];
#pragma unroll
  for (int a = 0; a < D; a++) {
    xbi[a] = S.xb[a * cap + i];
    xi[a] = P.x[a * cap + i];
  }
  const double hi = P.h[i], hbi = S.hb[i];
  const double rhoi = P.rho[i], ci = P.c[i];
  const double ca = max_v * 5.;
  const double P_b = rhoi * ca * ca;
  const double dxk = hi / kCellRatio * kCutOff;
  const double Cw = kern_C<D>(hi);
  const float Cf = (float)Cw;
  const float inv_h = (float)(1.0 / hi);
  const float dWc = (float)(-5.0 * Cw / (hi * hi));
  const float inv_Wdp = (float)(1.0 / kern_W<D>(dxk, hi));
  const float rlim = (float)(hbi * kCutOff);
  const float r2 = rlim * rlim;
  const float r2_box = r2 * 1.0001f;
  const float tmp1_base = (float)(-1. * P_b * (1. / rhoi));
  const float rhoi_f = t3.x, ci_f = t3.y;
  const float cri = ci_f * ci_f * rhoi_f;
  double pst[3] = {0, 0, 0}, sharp[3] = {0, 0, 0};
  const int pos = cell_of<D>(g, xbi);
  int pi, pj, pk;
  cell_ijk(g, pos, pi, pj, pk);
  const int r = g.radius;
  const int rk = D == 3 ? r : 0;
  const float psf[3] = {(float)g.ps[0], (float)g.ps[1], (float)g.ps[2]};
  for (int t = max(pi - r, 0); t <= min(pi + r, g.N[0] - 1); t++)
    for (int s = max(pj - r, 0); s <= min(pj + r, g.N[1] - 1); s++)
      for (int m = max(pk - rk, 0); m <= min(pk + rk, g.N[2] - 1); m++) {
        const int cc = (t * g.N[1] + s) * g.N[2] + m;
        // the target's positions in the candidate cell's frame:
        // (target offset) - (cell offset) * cell size
        const float off[3] = {(float)(t - cti) * psf[0], (float)(s - ctj) * psf[1],
                              D == 3 ? (float)(m - ctk) * psf[2] : 0.f};
        float vb[3], vu[3];
#pragma unroll
        for (int a = 0; a < D; a++) {
          vb[a] = ubt[a] - off[a];
          vu[a] = ut[a] - off[a];
        }
        const float4 blo = MB.box[2 * cc], bhi = MB.box[2 * cc + 1];
        const float lo[3] = {blo.x, blo.y, blo.z}, hb_[3] = {bhi.x, bhi.y, bhi.z};
        float bd = 0;
#pragma unroll
        for (int a = 0; a < D; a++) {
          const float d = fmaxf(fmaxf(lo[a] - vb[a], vb[a] - hb_[a]), 0.f);
          bd = fmaf(d, d, bd);
        }
        if (!(bd <= r2_box)) continue;  // empty cell or wholly outside
        const int e1 = bin_off[cc + 1];
        for (int f = bin_off[cc]; f < e1; f++) {
          const float4 c1 = MB.r1[f];
          const float cb[3] = {c1.x, c1.y, c1.z};
          float sqb = 0;
#pragma unroll
          for (int a = 0; a < D; a++) {
            const float d = vb[a] - cb[a];
            sqb = fmaf(d, d, sqb);
          }
          if (!(sqb <= r2)) continue;
          const float4 c2 = MB.r2[f];
          const float cu[3] = {c2.x, c2.y, c2.z};
          float dr[3] = {0, 0, 0};
          float sq = 0;
#pragma unroll
          for (int a = 0; a < D; a++) {
            dr[a] = vu[a] - cu[a];
            sq = fmaf(dr[a], dr[a], sq);
          }
          if (sq == 0.f) continue;
          // Wendland C2: W = C q^4 (1 + 2 s), dW/r = -5 C q^3 / h^2 with
          // q = max(1 - s/2, 0), s = r / h
          const float dist = __builtin_amdgcn_sqrtf(sq);
          const float sk = dist * inv_h;
          const float q = fmaxf(fmaf(-0.5f, sk, 1.f), 0.f);
          const float q3 = (q * q) * q;
          const float W_ = Cf * (q3 * q) * fmaf(2.f, sk, 1.f);
          const float aw = dWc * q3;
          const float rr = W_ * inv_Wdp;
          const float q2 = rr * rr;
          const float tmp1 = tmp1_base * fmaf(0.2f, q2 * q2, 1.f);
          const int phj = __float_as_int(c1.w);
          const float vol_j = c2.w;
          if (phj != 10) {
            const float fac = aw * tmp1 * vol_j;
#pragma unroll
            for (int a = 0; a < D; a++) pst[a] += (double)(dr[a] * fac);
            if (phj != phi) {
              const float2 c3 = MB.r3[f];
              const float kk = -0.1f * fmaf(c3.y * c3.y, c3.x, cri) *
                               __builtin_amdgcn_rcpf(rhoi_f + c3.x);
              const float fs = kk * aw * vol_j;
#pragma unroll
              for (int a = 0; a < D; a++) sharp[a] += (double)(dr[a] * fs);
            }
          } else {
            const int w = ref_wall(B.ref[f]);
            const float fw = W_ * tmp1 * vol_j;
#pragma unroll
            for (int a = 0; a < D; a++) pst[a] += (double)((float)W.norm[a * W.n + w] * fw);
          }
        }
      }
  // boundary penalty, clamps and the transport velocity in FP64, as in the
  // FP64 kernel
  double pen[3] = {0, 0, 0};
  {
    const int c = S.in_ele[i];
    int ci_, cj_, ck_;
    cell_ijk(g, c, ci_, cj_, ck_);
    const double xc[3] = {g.ps[0] * (ci_ + 0.5), g.ps[1] * (cj_ + 0.5),
                          D == 3 ? g.ps[2] * (ck_ + 0.5) : 0.0};
    for (int f = 0; f < 2 * D; f++) {
      int ax, ni = ci_, nj = cj_, nk = ck_;
      double sgn, cen[3] = {xc[0], xc[1], xc[2]};
      switch (f) {
        case 0: ax = 1; sgn = -1; cen[1] = xc[1] - 0.5 * g.ps[1]; nj--; break;
        case 1: ax = 0; sgn = 1; cen[0] = xc[0] + 0.5 * g.ps[0]; ni++; break;
        case 2: ax = 1; sgn = 1; cen[1] = xc[1] + 0.5 * g.ps[1]; nj++; break;
        case 3: ax = 0; sgn = -1; cen[0] = xc[0] - 0.5 * g.ps[0]; ni--; break;
        case 4: ax = 2; sgn = -1; cen[2] = xc[2] - 0.5 * g.ps[2]; nk--; break;
        default: ax = 2; sgn = 1; cen[2] = xc[2] + 0.5 * g.ps[2]; nk++; break;
      }
      bool in = ni >= 0 && ni < g.N[0] && nj >= 0 && nj < g.N[1] && nk >= 0 && nk < g.N[2];
      if (in || g.bc[f] == 0) continue;
      double nrm[3] = {0, 0, 0};
      nrm[ax] = sgn;
      double ds = 0;
#pragma unroll
      for (int a = 0; a < D; a++) ds += (xi[a] - cen[a]) * nrm[a];
      if (ds < 0)
#pragma unroll
        for (int a = 0; a < 3; a++) nrm[a] = -1 * nrm[a];
      double dis = fabs(ds);
      if (dis < 0.4 * g.p_sz) {
        double kk = 0.02 / dis * (ci * ci);
#pragma unroll
        for (int a = 0; a < D; a++) pen[a] += kk * nrm[a];
      }
    }
  }
  const double px = g.p_sz;
  double dx[3], dx1[3], dx2[3];
#pragma unroll
  for (int a = 0; a < D; a++) {
    dx[a] = pst[a] * dt * dt / 2.;
    dx1[a] = sharp[a] * dt * dt / 2.;
    dx2[a] = pen[a] * dt * dt / 2.;
  }
  auto clampv = [&](double* d, double cap_) {
    double s2 = 0;
#pragma unroll
    for (int a = 0; a < D; a++) s2 += d[a] * d[a];
    double dis = sqrt(s2);
    double lim = amin1(cap_, dis) / (dis + 1e-20);
#pragma unroll
    for (int a = 0; a < D; a++) d[a] *= lim;
  };
  clampv(dx, px / 10);
  clampv(dx1, px / 10);
  clampv(dx2, px / 2);
  const bool frozen = g.sem == PGDG_SEM_S3 && P.mark_in[i] == 1;
#pragma unroll
  for (int a = 0; a < D; a++) {
    double dv = frozen ? 0.0 : (dx[a] + dx1[a] + dx2[a]) / dt;
    P.tv[a * cap + i] = P.v[a * cap + i] + dv;
  }
}

template <int D>
int launch_pst_mixed_kernel(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                            const BinData& B, const MixedBins& MB, int64_t n_ent, const int* bin_off,
                            const DevScalars* sc, hipStream_t st, int block) {
  switch (block) {
    case 64:
      hipLaunchKernelGGL((k_pst_mixed<D, 64>), dim3(grid_xcd(n_ent, 64)), dim3(64), 0, st, g, P, W, S, B,
                         MB, bin_off, sc);
      break;
    case 128:
      hipLaunchKernelGGL((k_pst_mixed<D, 128>), dim3(grid_xcd(n_ent, 128)), dim3(128), 0, st, g, P, W, S,
                         B, MB, bin_off, sc);
      break;
    case 256:
      hipLaunchKernelGGL((k_pst_mixed<D, 256>), dim3(grid_xcd(n_ent, 256)), dim3(256), 0, st, g, P, W, S,
                         B, MB, bin_off, sc);
      break;
    default:
      return PGDG_ERR_ARG;
  }
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int launch_pst_mixed(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                     const BinData& B, const MixedBins& MB, int64_t n_ent, const int* bin_off,
                     const int* bin_ent, const DevScalars* sc, hipStream_t st, int block) {
  if (n_ent <= 0) return PGDG_OK;
  if (n_ent > MB.ne_cap) return PGDG_ERR_CAPACITY;
  hipLaunchKernelGGL(k_bin_gather_mixed<D>, dim3((unsigned)((n_ent + 255) / 256)), dim3(256), 0, st, g,
                     P, W, S, B, MB, bin_off, bin_ent, sc);
  hipLaunchKernelGGL(k_cell_bbox_mixed<D>, dim3((g.nc + 255) / 256), dim3(256), 0, st, g, MB, bin_off);
  PGDG_TRY(launch_pst_mixed_kernel<D>(g, P, W, S, B, MB, n_ent, bin_off, sc, st, block));
  return PGDG_OK;
}

#define PGDG_INST_MIXED(D)                                                                         \
  template int launch_pst_mixed<D>(const Geo&, const PtlDev&, const WallDev&, const StageBufs&,     \
                                   const BinData&, const MixedBins&, int64_t, const int*,          \
                                   const int*, const DevScalars*, hipStream_t, int);               \
  template int launch_pst_mixed_kernel<D>(const Geo&, const PtlDev&, const WallDev&,                \
                                          const StageBufs&, const BinData&, const MixedBins&,       \
                                          int64_t, const int*, const DevScalars*, hipStream_t, int);
PGDG_INST_MIXED(2)
PGDG_INST_MIXED(3)
#undef PGDG_INST_MIXED

}  // namespace pgdg
