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
#include <climits>

#include "dev_common.hpp"
#include "hip_util.hpp"
#include "step_kernels.hpp"

namespace pgdg {

// Bin cell of every entry (binary search over the bin offsets; search ==
// false: the bins have not changed since the last search, reuse it) and
// its FP32 records.
template <int D>
__global__ void k_bin_gather_mixed(Geo g, PtlDev P, WallDev W, StageBufs S, BinData B, MixedBins MB,
                                   const int* __restrict__ bin_off, const int* __restrict__ bin_ent,
                                   const DevScalars* __restrict__ sc, bool search) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= sc->nent_dev) return;
  int c;
  if (search) {
    int lo = 0, hi = g.nc - 1;  // last cell with bin_off[c] <= e
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (bin_off[mid] <= e) lo = mid;
      else hi = mid - 1;
    }
    c = lo;
  } else {
    c = MB.ecell[e];
  }
  int ci, cj, ck;
  cell_ijk(g, c, ci, cj, ck);
  const double corner[3] = {g.box_l[0] + ci * g.ps[0], g.box_l[1] + cj * g.ps[1],
                            D == 3 ? g.box_l[2] + ck * g.ps[2] : 0.0};
  const int ref = bin_ent[e];
  B.ref[e] = ref;
  if (search) MB.ecell[e] = c;
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
  MB.r1[e] = make_float4(ub[0], ub[1], ub[2], fmaf(ub[0], ub[0], fmaf(ub[1], ub[1], ub[2] * ub[2])));
  MB.r2[e] = make_float4(u[0], u[1], u[2], vol);
  MB.r3[e] = make_float2(rho, cs);
  MB.ph[e] = ph;
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

// Union of the boxes of the cells k - radius .. k + radius of each cell's
// line (3-D): a target whose position is more than the search radius away
// from it (x, y from the union; z from the cells' least lower and largest
// upper offsets within their own cells) reaches none of those cells.
__global__ void k_cell_ubox_mixed(Geo g, MixedBins MB) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int N2 = g.N[2], k = c % N2, line = c - k;
  float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
  float4 hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
  for (int m = max(k - g.radius, 0); m <= min(k + g.radius, N2 - 1); m++) {
    const float4 a = MB.box[2 * (line + m)], b = MB.box[2 * (line + m) + 1];
    lo.x = fminf(lo.x, a.x);
    lo.y = fminf(lo.y, a.y);
    lo.z = fminf(lo.z, a.z);
    hi.x = fmaxf(hi.x, b.x);
    hi.y = fmaxf(hi.y, b.y);
    hi.z = fmaxf(hi.z, b.z);
  }
  MB.ubox[2 * c] = lo;
  MB.ubox[2 * c + 1] = hi;
}

// 3-D: the frame of cell k's records in the candidate walk: z relative to
// the lower corner of a fixed cell of k's line, so that every walk, staged
// or not, at any workgroup size, forms the same FP32 offsets.  Lines of up
// to kZLine cells (ZL): the middle cell of the line, one frame for the
// whole line; longer lines: the middle cell of k's block of kZBlock cells
// (a window that crosses a block edge switches frames there).  Offsets of
// up to 64 cell sizes keep the rounding of a record's z near 4e-6 cell
// sizes.
constexpr int kZLine = 128, kZBlock = 64;
template <bool ZL>
__device__ __forceinline__ int zframe(int k, int n2) {
  return ZL ? n2 >> 1 : (k & ~(kZBlock - 1)) + kZBlock / 2;
}

// Per-target constants of the FP32 pair terms.
struct PstConst {
  float r2, r2_box, Cf, inv_h, dWc, inv_Wdp, tmp1_base, rhoi_f, cri;
  float cw, ca;  // Cf * inv_Wdp and dWc * tmp1_base (applied to a line's sums)
  float cwall;   // Cf / dWc: a wall term relative to ca
  int phi;
};

// Does the candidate cell's box (its entries' stage-start positions, in its
// own frame) come within the search radius of vb (the target's stage-start
// position in that frame)?
template <int D>
__device__ __forceinline__ bool box_reach(const float4 blo, const float4 bhi, const float* vb, float r2_box) {
  const float lo[3] = {blo.x, blo.y, blo.z}, hb_[3] = {bhi.x, bhi.y, bhi.z};
  float bd = 0;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const float d = fmaxf(fmaxf(lo[a] - vb[a], vb[a] - hb_[a]), 0.f);
    bd = fmaf(d, d, bd);
  }
  return bd <= r2_box;
}

// Pair terms of one target with the candidate entries x in [x0, x1) of one
// cell: rec1(x) / rec2(x) return the entry's records, glob(x) its bin-entry
// index (for the density and sound speed of interface pairs and for wall
// normals).  vb, vu: the target's stage-start and current positions in the
// candidate cell's frame.  The terms are summed in FP32 within the cell
// (pf, sf: at most the cell's entries) and the caller adds each cell's sums
// to the FP64 totals; the particle-shifting sums leave out the per-target
// factor ca (dW/r prefactor times the pressure term), which the caller
// applies to each line's sum.
template <int D, bool SW = false, typename R1, typename R2, typename RP, typename G>
__device__ __forceinline__ void pst_pairs(int x0, int x1, const float* vb_, const float* vu_, const PstConst& k,
                                          R1 rec1, R2 rec2, RP recp, G glob, const MixedBins& MB,
                                          const BinData& B, const WallDev& W, float* pst, float* sharp,
                                          int xs = -1, float vbz2 = 0.f, float vuz2 = 0.f) {
  // list test |vb - cb|^2 <= r2 as |cb|^2 - 2 vb.cb <= r2 - |vb|^2 (the
  // record carries |cb|^2); from entry xs on the target's z positions are
  // vbz2, vuz2 (the records from there on sit in the next block's frame)
  float vu[3], m2vb[3], thr = k.r2;
#pragma unroll
  for (int a = 0; a < D; a++) {
    vu[a] = vu_[a];
    m2vb[a] = -2.f * vb_[a];
    thr = fmaf(-vb_[a], vb_[a], thr);
  }
  if (x0 >= x1) return;
  auto pair = [&](int x, const float4 c1) {
    if (SW && x == xs) {  // (x, y of the frame are unchanged; -0.5 m2vb is vb exactly)
      const float vbx = -0.5f * m2vb[0], vby = -0.5f * m2vb[1];
      thr = fmaf(-vbz2, vbz2, fmaf(-vby, vby, fmaf(-vbx, vbx, k.r2)));
      m2vb[2] = -2.f * vbz2;
      vu[2] = vuz2;
    }
    const float cb[3] = {c1.x, c1.y, c1.z};
    float t = c1.w;
#pragma unroll
    for (int a = 0; a < D; a++) t = fmaf(m2vb[a], cb[a], t);
    if (!(t <= thr)) return;
    const float4 c2 = rec2(x);
    const float cu[3] = {c2.x, c2.y, c2.z};
    float dr[3] = {0, 0, 0};
    float sq = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
      dr[a] = vu[a] - cu[a];
      sq = fmaf(dr[a], dr[a], sq);
    }
    // (a pair at zero distance, the target itself included, adds dr = 0)
    // Wendland C2: W = C q^4 (1 + 2 s), dW/r = -5 C q^3 / h^2 with
    // q = max(1 - s/2, 0), s = r / h; the particle-shifting factor
    // (dW/r) * tmp1 with tmp1 = base (1 + 0.2 (W / W(dx))^4) is formed as
    // (dWc base) q^3 (1 + 0.2 rr^4), rr = (C / W(dx)) q^4 (1 + 2 s)
    const float dist = __builtin_amdgcn_sqrtf(sq);
    const float sk = dist * k.inv_h;
    const float q = fmaxf(fmaf(-0.5f, sk, 1.f), 0.f);
    const float q3 = (q * q) * q;
    const float t2 = fmaf(2.f, sk, 1.f);
    const float rr = k.cw * (q3 * q) * t2;
    const float q2 = rr * rr;
    const float poly = fmaf(0.2f, q2 * q2, 1.f);
    const int phj = recp(x);
    const float vol_j = c2.w;
    if (phj != 10) {
      const float fac = (q3 * poly) * vol_j;
#pragma unroll
      for (int a = 0; a < D; a++) pst[a] = fmaf(dr[a], fac, pst[a]);
      if (phj != k.phi) {
        const float2 c3 = MB.r3[glob(x)];
        const float kk = -0.1f * fmaf(c3.y * c3.y, c3.x, k.cri) * __builtin_amdgcn_rcpf(k.rhoi_f + c3.x);
        const float fs = kk * (k.dWc * q3) * vol_j;
#pragma unroll
        for (int a = 0; a < D; a++) sharp[a] = fmaf(dr[a], fs, sharp[a]);
      }
    } else {
      const int w = ref_wall(B.ref[glob(x)]);
      const float fw = (k.cwall * (q3 * q) * t2) * poly * vol_j;
#pragma unroll
      for (int a = 0; a < D; a++) pst[a] = fmaf((float)W.norm[a * W.n + w], fw, pst[a]);
    }
  };
  // two candidates per iteration, each record loaded one candidate ahead
  // (the reads stay within x1 + 1: both record tables carry a spare record
  // at the end)
  float4 ca = rec1(x0);
  int x = x0;
  for (; x + 1 < x1; x += 2) {
    const float4 cb_ = rec1(x + 1);
    pair(x, ca);
    ca = rec1(x + 2);
    pair(x + 1, cb_);
  }
  if (x < x1) pair(x, ca);
}

// One lane per bin entry (target); workgroups take runs of consecutive
// entries, i.e. of consecutive cells.  3-D: the workgroup's targets lie on
// one or two (i, j) cell lines; for each candidate line offset the
// candidate entries of the cells within reach of those lines (a contiguous
// entry range per line) are staged in LDS, so that every lane reads its
// candidates from LDS rather than through the vector memory path.  The
// staged records are moved into one frame per line (z relative to the
// middle cell of the staged range), so each lane walks a single window of
// entries per candidate line: from the first to the last of its candidate
// cells whose box is within reach (the cells between, out of reach, add no
// pair).  The lanes of a wavefront step through their windows together,
// where a walk cell by cell would step through the longest lane's cell at
// every position.  The candidates and their order are the per-lane walk's
// (the path taken in 2-D, for workgroups spanning more lines, candidate
// ranges wider than the LDS tables, or with staging switched off); the
// pair offsets differ by the FP32 rounding of the frames.
template <int D, int BS, bool ZL>
__global__ void __launch_bounds__(BS, 4) k_pst_mixed(Geo g, PtlDev P, WallDev W, StageBufs S, BinData B,
                                                   MixedBins MB, const int* __restrict__ bin_off,
                                                   const DevScalars* __restrict__ sc, int staging) {
  const int64_t E0 = (int64_t)xcd_block() * BS;
  const int64_t nent = sc->nent_dev;
  if (E0 >= nent) return;  // the whole workgroup is past the entries
  const int64_t e = E0 + threadIdx.x;
  const int i = e < nent ? B.ref[e] : -1;
  const bool active = i >= 0;  // a particle (not a wall, removed or past-the-end entry)
  const int64_t cap = P.cap;
  const double dt = sc->dt;
  const double max_v = sc->max_v;
  int ct = 0, cti = 0, ctj = 0, ctk = 0, pi = 0, pj = 0, pk = 0;
  float ubt[3] = {0, 0, 0}, ut[3] = {0, 0, 0};
  PstConst k{};
  if (active) {
    ct = MB.ecell[e];
    cell_ijk(g, ct, cti, ctj, ctk);
    const float4 t1 = MB.r1[e], t2 = MB.r2[e];
    const float2 t3 = MB.r3[e];
    ubt[0] = t1.x, ubt[1] = t1.y, ubt[2] = t1.z;
    ut[0] = t2.x, ut[1] = t2.y, ut[2] = t2.z;
    k.phi = MB.ph[e];
    double xbi[3];
#pragma unroll
    for (int a = 0; a < D; a++) xbi[a] = S.xb[a * cap + i];
    const double hi = P.h[i], hbi = S.hb[i];
    const double rhoi = P.rho[i];
    const double ca = max_v * 5.;
    const double P_b = rhoi * ca * ca;
    const double dxk = hi / kCellRatio * kCutOff;
    const double Cw = kern_C<D>(hi);
    k.Cf = (float)Cw;
    k.inv_h = (float)(1.0 / hi);
    k.dWc = (float)(-5.0 * Cw / (hi * hi));
    k.inv_Wdp = (float)(1.0 / kern_W<D>(dxk, hi));
    const float rlim = (float)(hbi * kCutOff);
    k.r2 = rlim * rlim;
    k.r2_box = k.r2 * 1.0001f;
    k.tmp1_base = (float)(-1. * P_b * (1. / rhoi));
    k.cw = (float)(Cw / kern_W<D>(dxk, hi));
    k.ca = (float)(-5.0 * Cw / (hi * hi) * (-1. * P_b * (1. / rhoi)));
    k.cwall = (float)(Cw / (-5.0 * Cw / (hi * hi)));
    k.rhoi_f = t3.x;
    k.cri = t3.y * t3.y * t3.x;
    const int pos = cell_of<D>(g, xbi);
    cell_ijk(g, pos, pi, pj, pk);
  }
  double pst[3] = {0, 0, 0}, sharp[3] = {0, 0, 0};
  float pf[3] = {0, 0, 0}, sf[3] = {0, 0, 0};  // the current cell's FP32 sums
  const double ca_d = (double)k.ca;
  auto flush = [&]() {
#pragma unroll
    for (int a = 0; a < D; a++) {
      pst[a] += (double)pf[a] * ca_d;
      sharp[a] += (double)sf[a];
      pf[a] = sf[a] = 0.f;
    }
  };
  const int r = g.radius;
  const float psf[3] = {(float)g.ps[0], (float)g.ps[1], (float)g.ps[2]};
  const float ipsz = (float)(1.0 / g.ps[2]);
  // the target's positions in the frame of a candidate cell (t, s, m):
  // (target offset) - (cell offset) * cell size
  auto frame = [&](int t, int s, int m, float* vb, float* vu) {
    const float off[3] = {(float)(t - cti) * psf[0], (float)(s - ctj) * psf[1],
                          D == 3 ? (float)(m - ctk) * psf[2] : 0.f};
#pragma unroll
    for (int a = 0; a < D; a++) {
      vb[a] = ubt[a] - off[a];
      vu[a] = ut[a] - off[a];
    }
  };
  auto grec1 = [&](int x) { return MB.r1[x]; };
  auto grec2 = [&](int x) { return MB.r2[x]; };
  auto grecp = [&](int x) { return MB.ph[x]; };
  auto gidx = [](int x) { return x; };
  bool staged = false;
  if constexpr (D == 3) {
    constexpr int CAP = BS == 256 ? 640 : BS == 128 ? 384 : 192;  // staged entries
    constexpr int KMAX = BS == 256 ? 96 : BS == 128 ? 64 : 40;    // staged cells per line
    __shared__ float4 s_r1[CAP + 1], s_r2[CAP];  // s_r1[CAP]: the spare record
    __shared__ int s_ph[CAP];
    __shared__ float4 s_box[2][KMAX][2];
    __shared__ int s_off[2][KMAX + 1];
    __shared__ int s_agg[2][6];  // per line: min, max of (pi - i), (pj - j), pk
    const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
    const int64_t e_last = (E0 + BS < nent ? E0 + BS : nent) - 1;
    const int line_lo = MB.ecell[E0] / N2, line_hi = MB.ecell[e_last] / N2;
    staged = staging && line_hi - line_lo <= 1;
    if (staged) {
      if (threadIdx.x < 12) (&s_agg[0][0])[threadIdx.x] = (threadIdx.x & 1) ? INT_MIN : INT_MAX;
      __syncthreads();
      const int sg = active ? ct / N2 - line_lo : 0;
      const int li = (line_lo + sg) / N1, lj = (line_lo + sg) % N1;
      if (active) {
        atomicMin(&s_agg[sg][0], pi - li);
        atomicMax(&s_agg[sg][1], pi - li);
        atomicMin(&s_agg[sg][2], pj - lj);
        atomicMax(&s_agg[sg][3], pj - lj);
        atomicMin(&s_agg[sg][4], pk);
        atomicMax(&s_agg[sg][5], pk);
      }
      __syncthreads();
      // workgroup-uniform plan (scalar registers)
      int di0 = INT_MAX, di1 = INT_MIN, dj0 = INT_MAX, dj1 = INT_MIN;
      int kA[2] = {0, 0}, kB[2] = {-1, -1};
      bool has[2];
#pragma unroll
      for (int q = 0; q < 2; q++) {
        int ag[6];
#pragma unroll
        for (int u = 0; u < 6; u++) ag[u] = __builtin_amdgcn_readfirstlane(s_agg[q][u]);
        has[q] = ag[4] <= ag[5];
        if (has[q]) {
          di0 = min(di0, ag[0]);
          di1 = max(di1, ag[1]);
          dj0 = min(dj0, ag[2]);
          dj1 = max(dj1, ag[3]);
          kA[q] = max(ag[4] - r, 0);
          kB[q] = min(ag[5] + r, N2 - 1);
          if (kB[q] - kA[q] + 1 > KMAX) staged = false;
        }
      }
      const int kA_l = sg ? kA[1] : kA[0];
      if (staged && (has[0] || has[1])) {
        for (int dI = di0 - r; dI <= di1 + r; dI++)
          for (int dJ = dj0 - r; dJ <= dj1 + r; dJ++) {
            int a0[2] = {0, 0}, n0[2] = {0, 0}, cb[2] = {0, 0};
            bool v[2];
#pragma unroll
            for (int q = 0; q < 2; q++) {
              const int I = (line_lo + q) / N1 + dI, J = (line_lo + q) % N1 + dJ;
              v[q] = has[q] && I >= 0 && I < N0 && J >= 0 && J < N1;
              if (v[q]) {
                cb[q] = (I * N1 + J) * N2;
                a0[q] = bin_off[cb[q] + kA[q]];
                n0[q] = bin_off[cb[q] + kB[q] + 1] - a0[q];
              }
            }
            const int ntot = n0[0] + n0[1];
            if (ntot == 0) continue;  // workgroup-uniform
            // the candidate cells' entry offsets and boxes
            for (int q = threadIdx.x; q < 2 * (KMAX + 1); q += BS) {
              const int s_ = q / (KMAX + 1), lc = q % (KMAX + 1);
              if (v[s_] && lc <= kB[s_] - kA[s_] + 1) {
                const int c = cb[s_] + kA[s_] + lc;
                s_off[s_][lc] = bin_off[c];
                if (lc <= kB[s_] - kA[s_]) {
                  s_box[s_][lc][0] = MB.ubox[2 * c];
                  s_box[s_][lc][1] = MB.ubox[2 * c + 1];
                }
              }
            }
            __syncthreads();  // s_off, s_box
            const int I = li + dI, J = lj + dJ;
            bool mine = active && (sg ? v[1] : v[0]) && abs(I - pi) <= r && abs(J - pj) <= r;
            const int base = sg == 0 ? 0 : n0[0];
            const int sh = (sg ? a0[1] : a0[0]) - base;  // entry index - staged index
            // the lane's window: the entries of the candidate cells whose z
            // range comes within rho of the target, rho the search radius
            // left after the lateral distance to the union of the candidate
            // cells' boxes; it holds every cell whose own box is within
            // reach (the others' entries fail the list test).  A window
            // that crosses the edge of a block of cells is walked in two
            // segments, each with the target in its block's frame.
            int wx0 = 0, wx1 = 0, wxb = 0;  // [wx0, wxb) first block, [wxb, wx1) second
            float vb[3], vu[3], vbz2 = 0.f, vuz2 = 0.f;
            if (mine) {
              const int kF = zframe<ZL>(pk, N2);
              frame(I, J, kF, vb, vu);
              const int lp = pk - kA_l;
              const float4 ulo = s_box[sg][lp][0], uhi = s_box[sg][lp][1];
              const float gx = fmaxf(fmaxf(ulo.x - vb[0], vb[0] - uhi.x), 0.f);
              const float gy = fmaxf(fmaxf(ulo.y - vb[1], vb[1] - uhi.y), 0.f);
              const float rho2 = fmaf(-gx, gx, fmaf(-gy, gy, k.r2_box));
              int m0 = max(pk - r, 0), m1 = min(pk + r, N2 - 1);
              if (rho2 >= 0.f) {
                const float rho = __builtin_amdgcn_sqrtf(rho2) * 1.0001f;
                // cell m: (m - kF) psz + ulo.z <= vz + rho, (m - kF) psz + uhi.z >= vz - rho
                m1 = min(m1, kF + (int)floorf((vb[2] + rho - ulo.z) * ipsz));
                m0 = max(m0, kF + (int)ceilf((vb[2] - rho - uhi.z) * ipsz));
              } else {
                m1 = m0 - 1;
              }
              mine = m0 <= m1;
              if (mine) {
                wx0 = s_off[sg][m0 - kA_l] - sh;
                wx1 = s_off[sg][m1 + 1 - kA_l] - sh;
                const int mb = ZL ? 0 : m1 & ~(kZBlock - 1);  // the second block's first cell
                wxb = mb > m0 ? s_off[sg][mb - kA_l] - sh : wx0;
                float vb2[3], vu2[3];
                frame(I, J, zframe<ZL>(m1, N2), vb2, vu2);
                vbz2 = vb2[2];
                vuz2 = vu2[2];
                frame(I, J, zframe<ZL>(m0, N2), vb, vu);
              }
            }
            for (int q0 = 0; q0 < ntot; q0 += CAP) {
              const int qn = min(ntot - q0, CAP);
              for (int q = threadIdx.x; q < qn; q += BS) {
                const int x = q0 + q;
                const bool l1 = x >= n0[0];
                const int f = l1 ? a0[1] + (x - n0[0]) : a0[0] + x;
                const int m = MB.ecell[f] - (l1 ? cb[1] : cb[0]);
                const float oz = (float)(m - zframe<ZL>(m, N2)) * psf[2];
                float4 c1 = MB.r1[f], c2 = MB.r2[f];
                c1.z += oz;
                c2.z += oz;
                c1.w = fmaf(c1.x, c1.x, fmaf(c1.y, c1.y, c1.z * c1.z));
                s_r1[q] = c1;
                s_r2[q] = c2;
                s_ph[q] = MB.ph[f];
              }
              __syncthreads();
              if (mine) {
                auto sr1 = [&](int x) { return s_r1[x - q0]; };
                auto sr2 = [&](int x) { return s_r2[x - q0]; };
                auto sph = [&](int x) { return s_ph[x - q0]; };
                auto sgi = [&](int x) { return x + sh; };
                // (a chunk that starts past wxb starts in the second frame:
                // switch at its first entry)
                const int x0 = max(wx0, q0);
                const int xs = wxb > wx0 ? max(wxb, x0) : -1;
                pst_pairs<D, !ZL>(x0, min(wx1, q0 + qn), vb, vu, k, sr1, sr2, sph, sgi, MB, B, W, pf, sf, xs,
                                  vbz2, vuz2);
              }
              __syncthreads();
            }
            flush();  // the line's FP32 sums (a line may span chunks)
          }
      }
    }
  }
  if (!active) return;
  if (!staged) {
    const int rk = D == 3 ? r : 0;
    for (int t = max(pi - r, 0); t <= min(pi + r, g.N[0] - 1); t++)
      for (int s = max(pj - r, 0); s <= min(pj + r, g.N[1] - 1); s++) {
        for (int m = max(pk - rk, 0); m <= min(pk + rk, g.N[2] - 1); m++) {
          const int cc = (t * g.N[1] + s) * g.N[2] + m;
          float vb[3], vu[3];
          frame(t, s, m, vb, vu);
          if (!box_reach<D>(MB.box[2 * cc], MB.box[2 * cc + 1], vb, k.r2_box)) continue;  // empty or out of reach
          if constexpr (D == 3) {
            // the records moved into the block frame, as the staged walk has them
            const float oz = (float)(m - zframe<ZL>(m, g.N[2])) * psf[2];
            frame(t, s, zframe<ZL>(m, g.N[2]), vb, vu);
            auto br1 = [&](int x) {
              float4 c1 = MB.r1[x];
              c1.z += oz;
              c1.w = fmaf(c1.x, c1.x, fmaf(c1.y, c1.y, c1.z * c1.z));
              return c1;
            };
            auto br2 = [&](int x) {
              float4 c2 = MB.r2[x];
              c2.z += oz;
              return c2;
            };
            pst_pairs<D>(bin_off[cc], bin_off[cc + 1], vb, vu, k, br1, br2, grecp, gidx, MB, B, W, pf, sf);
          } else {
            pst_pairs<D>(bin_off[cc], bin_off[cc + 1], vb, vu, k, grec1, grec2, grecp, gidx, MB, B, W, pf, sf);
          }
        }
        flush();  // the line's FP32 sums
      }
  }
  // boundary penalty, clamps and the transport velocity in FP64, as in the
  // FP64 kernel
  double xi[3];
#pragma unroll
  for (int a = 0; a < D; a++) xi[a] = P.x[a * cap + i];
  const double ci = P.c[i];
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
                            const DevScalars* sc, hipStream_t st, int block, int staging) {
  // 3-D lines of up to kZLine cells: one record frame per line
  const bool zl = D == 3 && g.N[2] <= kZLine;
#define PGDG_PST_MIXED(BSZ)                                                                               \
  do {                                                                                                 \
    if (zl)                                                                                            \
      hipLaunchKernelGGL((k_pst_mixed<D, BSZ, true>), dim3(grid_xcd(n_ent, BSZ)), dim3(BSZ), 0, st, g, P, W, \
                         S, B, MB, bin_off, sc, staging);                                              \
    else                                                                                               \
      hipLaunchKernelGGL((k_pst_mixed<D, BSZ, false>), dim3(grid_xcd(n_ent, BSZ)), dim3(BSZ), 0, st, g, P,  \
                         W, S, B, MB, bin_off, sc, staging);                                           \
  } while (0)
  switch (block) {
    case 64:
      PGDG_PST_MIXED(64);
      break;
    case 128:
      PGDG_PST_MIXED(128);
      break;
    case 256:
      PGDG_PST_MIXED(256);
      break;
    default:
      return PGDG_ERR_ARG;
  }
#undef PGDG_PST_MIXED
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template <int D>
int launch_pst_mixed(const Geo& g, const PtlDev& P, const WallDev& W, const StageBufs& S,
                     const BinData& B, const MixedBins& MB, int64_t n_ent, const int* bin_off,
                     const int* bin_ent, const DevScalars* sc, hipStream_t st, int block, int staging,
                     int parts, bool bins_unchanged) {
  if (n_ent <= 0) return PGDG_OK;
  if (n_ent > MB.ne_cap) return PGDG_ERR_CAPACITY;
  if (parts & 1) {
    hipLaunchKernelGGL(k_bin_gather_mixed<D>, dim3((unsigned)((n_ent + 255) / 256)), dim3(256), 0, st, g,
                       P, W, S, B, MB, bin_off, bin_ent, sc, !bins_unchanged);
    hipLaunchKernelGGL(k_cell_bbox_mixed<D>, dim3((g.nc + 255) / 256), dim3(256), 0, st, g, MB, bin_off);
    if (D == 3) hipLaunchKernelGGL(k_cell_ubox_mixed, dim3((g.nc + 255) / 256), dim3(256), 0, st, g, MB);
    PGDG_HIP_TRY(hipGetLastError());
  }
  if (parts & 2) PGDG_TRY(launch_pst_mixed_kernel<D>(g, P, W, S, B, MB, n_ent, bin_off, sc, st, block, staging));
  return PGDG_OK;
}

#define PGDG_INST_MIXED(D)                                                                         \
  template int launch_pst_mixed<D>(const Geo&, const PtlDev&, const WallDev&, const StageBufs&,     \
                                   const BinData&, const MixedBins&, int64_t, const int*,          \
                                   const int*, const DevScalars*, hipStream_t, int, int, int, bool);\
  template int launch_pst_mixed_kernel<D>(const Geo&, const PtlDev&, const WallDev&,                \
                                          const StageBufs&, const BinData&, const MixedBins&,       \
                                          int64_t, const int*, const DevScalars*, hipStream_t, int, \
                                          int);
PGDG_INST_MIXED(2)
PGDG_INST_MIXED(3)
#undef PGDG_INST_MIXED

}  // namespace pgdg
