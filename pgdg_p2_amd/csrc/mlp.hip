// mlp.hip - learned initial guess for the MGFM Newton solve.
//
// A 12 -> 32 -> 32 -> 32 -> 2 ReLU MLP.  Every matrix product is a set of
// 16x16 output tiles computed with v_mfma_f32_16x16x16_bf16 (BF16 operands
// staged in LDS, FP32 accumulation); a workgroup of four wavefronts handles
// 64 samples.  Training keeps FP32 master parameters; each workgroup writes
// its partial gradient and a second kernel sums the partials in a fixed
// order and applies Adam, so training is deterministic.  The reference has
// no counterpart: its Newton solve starts from max(pL, pR) (S3:mpm.cpp:891)
// or (pL + pR) / 2 (T2:mpm.cpp:1383).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "dev_common.hpp"
#include "hip_util.hpp"
#include "mgfm_pairs.hpp"
#include "riemann.hpp"

namespace pgdg {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kIn = PGDG_MLP_NIN, kInPad = PGDG_MLP_NPAD, kH = PGDG_MLP_WIDTH;
static_assert(kH % 32 == 0, "the hidden width is a multiple of 32 (K = 32 MFMA steps)");
constexpr int kNT = kH / 16;  // 16-unit tiles of a hidden layer
constexpr int kKS = kH / 32;  // K = 32 steps over a hidden layer's inputs
constexpr int kOut = PGDG_MLP_NOUT, kOutPad = 16;
constexpr int oW0 = 0;
constexpr int oB0 = oW0 + kInPad * kH;
constexpr int oW1 = oB0 + kH;
constexpr int oB1 = oW1 + kH * kH;
constexpr int oW2 = oB1 + kH;
constexpr int oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH;
constexpr int oB3 = oW3 + kH * kOutPad;
constexpr int kNParam = oB3 + kOutPad;
static_assert(kNParam == PGDG_MLP_NPARAM, "parameter count");
constexpr int kTile = 64;     // samples per workgroup
constexpr int kThreads = 256;

// padded parameters stay zero
__host__ __device__ inline bool param_live(int p) {
  if (p < oB0) return (p - oW0) / kH < kIn;
  if (p >= oW3 && p < oB3) return (p - oW3) % kOutPad < kOut;
  if (p >= oB3) return p - oB3 < kOut;
  return true;
}

// Natural logarithm and exponential from the hardware log2 and exp2 (1 ulp
// in FP32): the features enter the network as BF16 and the output sets only
// a starting point, so these are ample, and they cost two instructions.
__device__ inline float ln_fast(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
__device__ inline float exp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// FP32 -> BF16, round to nearest even (the conversion instruction of gfx950)
__device__ inline unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// Activation of the hidden layers (one maximum per unit: the network runs
// once per face ahead of a few FP64 Newton iterations, so its activations
// must cost far less than one iteration).
__device__ inline float relu(float x) { return fmaxf(x, 0.f); }

// One 16x16 tile of C = A * B over K (a multiple of 16) from BF16 LDS
// operands with element strides: A(m, k) = A[m * ars + k * acs],
// B(k, n) = B[k * brs + n * bcs].  Lane l holds C(m0 + 4 (l / 16) + i,
// n0 + l % 16) in component i.
__device__ inline f32x4_t mma_tile(const unsigned short* A, int ars, int acs,
                                   const unsigned short* B, int brs, int bcs, int m0, int n0,
                                   int K) {
  const int l = threadIdx.x & 63;
  const int r = l & 15, kq = (l >> 4) * 4;
  f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t a, b;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      a[i] = (short)A[(m0 + r) * ars + (k0 + kq + i) * acs];
      b[i] = (short)B[(k0 + kq + i) * brs + (n0 + r) * bcs];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

struct FwdLds {
  unsigned short X[kTile * kInPad];
  unsigned short W0[kInPad * kH], W1[kH * kH], W2[kH * kH], W3[kH * kOutPad];
  unsigned short H0[kTile * kH], H1[kTile * kH], H2[kTile * kH];
  float b0[kH], b1[kH], b2[kH], b3[kOutPad];
};
// Training adds the output gradient (FP32 and BF16) and the BF16 gradient
// of the current hidden layer; the activation's derivative is taken from
// the stored BF16 activations (1 where h > 0, else 0).  Total below 64 KiB.
struct TrainLds {
  FwdLds f;
  float DYf[kTile * kOutPad];
  unsigned short DY[kTile * kOutPad];
  unsigned short DZ[kTile * kH];
  float lossv[kTile];
};
static_assert(sizeof(TrainLds) <= 65536, "training LDS");

// Load the parameters (FP32 -> BF16 weights, FP32 biases) into LDS.
__device__ inline void load_params(const float* __restrict__ P, FwdLds& s) {
  for (int q = threadIdx.x; q < kInPad * kH; q += kThreads) s.W0[q] = f2bf(P[oW0 + q]);
  for (int q = threadIdx.x; q < kH * kH; q += kThreads) {
    s.W1[q] = f2bf(P[oW1 + q]);
    s.W2[q] = f2bf(P[oW2 + q]);
  }
  for (int q = threadIdx.x; q < kH * kOutPad; q += kThreads) s.W3[q] = f2bf(P[oW3 + q]);
  for (int q = threadIdx.x; q < kH; q += kThreads) {
    s.b0[q] = P[oB0 + q];
    s.b1[q] = P[oB1 + q];
    s.b2[q] = P[oB2 + q];
  }
  for (int q = threadIdx.x; q < kOutPad; q += kThreads) s.b3[q] = P[oB3 + q];
}

// Hidden layer: Hout = relu(Hin * W + b), 64 samples x kH units, K = kin.
__device__ inline void hidden(const unsigned short* Hin, int kin, const unsigned short* W,
                              const float* b, unsigned short* Hout) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int t = w; t < 4 * kNT; t += 4) {
    const int m0 = (t / kNT) * 16, n0 = (t % kNT) * 16;
    f32x4_t acc = mma_tile(Hin, kin, 1, W, kH, 1, m0, n0, kin);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = m0 + (l >> 4) * 4 + i, col = n0 + (l & 15);
      Hout[row * kH + col] = f2bf(relu(acc[i] + b[col]));
    }
  }
}

// Forward through the block's 64 samples (X already in LDS); output layer
// values are passed to `emit(row, col, y)` for col < kOut.
template <typename Emit>
__device__ inline void forward(FwdLds& s, Emit emit) {
  hidden(s.X, kInPad, s.W0, s.b0, s.H0);
  __syncthreads();
  hidden(s.H0, kH, s.W1, s.b1, s.H1);
  __syncthreads();
  hidden(s.H1, kH, s.W2, s.b2, s.H2);
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int m0 = w * 16;
  f32x4_t acc = mma_tile(s.H2, kH, 1, s.W3, kOutPad, 1, m0, 0, kH);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int row = m0 + (l >> 4) * 4 + i, col = l & 15;
    emit(row, col, acc[i] + s.b3[col]);
  }
}

// (pL - pR) / (pL + pR) with the difference and the sum formed in FP64:
// the two sides' small pressure differences survive the features' BF16
// rounding (ln pL and ln pR alone lose them); the quotient needs only the
// hardware reciprocal
__device__ inline float p_contrast(double pl, double pr) {
  const double d = pl - pr, t = pl + pr;
  return t > 0 ? (float)d * __builtin_amdgcn_rcpf((float)t) : 0.f;
}

// standardised face features (FP32, padded to 16)
__device__ inline void face_features(const double* __restrict__ in, int64_t n, int64_t i, Eos ml,
                                     Eos mr, const pgdg_mlp_norm& nm, float* f) {
  const double ul = in[i], ur = in[n + i], pl = in[2 * n + i], pr = in[3 * n + i];
  const double rl = in[4 * n + i], rr = in[5 * n + i];
  // the features enter the network as BF16: FP32 logarithms are ample
  const float raw[kIn] = {ln_fast(fmaxf((float)pl, 1e-30f)), ln_fast(fmaxf((float)pr, 1e-30f)),
                          ln_fast(fmaxf((float)rl, 1e-30f)), ln_fast(fmaxf((float)rr, 1e-30f)),
                          (float)ul, (float)ur, p_contrast(pl, pr), (float)(ul - ur), (float)ml.gamma,
                          (float)log1p(fmax(ml.pinf, 0.0)), (float)mr.gamma,
                          (float)log1p(fmax(mr.pinf, 0.0))};
#pragma unroll
  for (int k = 0; k < kIn; k++) f[k] = (raw[k] - nm.feat_mu[k]) / nm.feat_sd[k];
#pragma unroll
  for (int k = kIn; k < kInPad; k++) f[k] = 0.f;
}

// ---------------------------------------------------------------------------
// inference: p*0 for every face
// ---------------------------------------------------------------------------
// Inference.  Each layer is computed transposed, H^T = W^T X^T: the MFMA
// result tile of a lane (column = face, rows = 4 consecutive units) is then
// the B-operand layout the next layer needs (k = those units, n = face), so
// activations stay in registers from the features to the output.  Hidden
// layers use the K = 32 MFMA of gfx950: a lane's 8 k values are the 4 units
// it holds of two result tiles, and the weights (A operand, W^T) are loaded
// with the same k order, so no lane exchange is needed.  Weights and biases
// are loaded once per wavefront; a wavefront runs 16-face tiles with no LDS
// and no barrier, prefetching the next tile's inputs.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short s16x8_t __attribute__((ext_vector_type(8)));

__device__ inline bf16x4_t pack4(float a, float b, float c, float d) {
  bf16x4_t r;
  r[0] = (short)f2bf(a);
  r[1] = (short)f2bf(b);
  r[2] = (short)f2bf(c);
  r[3] = (short)f2bf(d);
  return r;
}
__device__ inline bf16x8_t cat8(bf16x4_t lo, bf16x4_t hi) {
  s16x8_t r;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    r[q] = lo[q];
    r[4 + q] = hi[q];
  }
  return __builtin_bit_cast(bf16x8_t, r);
}

// A(m = unit, k) = W[k][unit] of a [K][ldw] weight for m-tile mt, 16-wide
// k-step ks (K = 16 MFMA layout: lane group g holds k = 4g .. 4g+3)
__device__ inline bf16x4_t load_wt(const float* __restrict__ W, int ldw, int mt, int ks) {
  const int l = threadIdx.x & 63;
  const int unit = mt * 16 + (l & 15), k = ks * 16 + (l >> 4) * 4;
  return pack4(W[(k + 0) * ldw + unit], W[(k + 1) * ldw + unit], W[(k + 2) * ldw + unit],
               W[(k + 3) * ldw + unit]);
}
// K = 32 step covering input tiles 2j and 2j+1: k values 0..3 of the lane
// are units 4g.. of tile 2j, 4..7 are units 4g.. of tile 2j+1
__device__ inline bf16x8_t load_wt32(const float* __restrict__ W, int ldw, int mt, int j) {
  return cat8(load_wt(W, ldw, mt, 2 * j), load_wt(W, ldw, mt, 2 * j + 1));
}

// The network's operands in the registers of a wavefront (MFMA A-operand
// layouts of W^T and the biases of the units a lane's result entries hold),
// loaded once and used for every 16-face tile the wavefront runs.
struct MlpRegs {
  bf16x4_t A0[kNT];
  bf16x8_t A1[kNT][kKS], A2[kNT][kKS], A3[kKS];
  float bb0[kNT][4], bb1[kNT][4], bb2[kNT][4];
  float b3;
  __device__ void load(const float* __restrict__ P) {
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int mt = 0; mt < kNT; mt++) {
      A0[mt] = load_wt(P + oW0, kH, mt, 0);
#pragma unroll
      for (int j = 0; j < kKS; j++) {
        A1[mt][j] = load_wt32(P + oW1, kH, mt, j);
        A2[mt][j] = load_wt32(P + oW2, kH, mt, j);
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int unit = mt * 16 + g * 4 + q;
        bb0[mt][q] = P[oB0 + unit];
        bb1[mt][q] = P[oB1 + unit];
        bb2[mt][q] = P[oB2 + unit];
      }
    }
#pragma unroll
    for (int j = 0; j < kKS; j++) A3[j] = load_wt32(P + oW3, kOutPad, 0, j);
    b3 = P[oB3];
  }
  // Output unit 0 (standardised ln(p* / max(pL, pR))) of face l % 16 of a
  // tile, valid in the lanes of group 0; x: features 4 g .. 4 g + 3 of that
  // face (lane group g).
  __device__ float run(const float* x) const {
    const bf16x4_t bx = pack4(x[0], x[1], x[2], x[3]);
    bf16x4_t h[kNT];
    {
      f32x4_t acc[kNT];
#pragma unroll
      for (int mt = 0; mt < kNT; mt++) {
        acc[mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A0[mt], bx, acc[mt], 0, 0, 0);
      }
#pragma unroll
      for (int mt = 0; mt < kNT; mt++)
        h[mt] = pack4(relu(acc[mt][0] + bb0[mt][0]), relu(acc[mt][1] + bb0[mt][1]),
                      relu(acc[mt][2] + bb0[mt][2]), relu(acc[mt][3] + bb0[mt][3]));
    }
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
      bf16x8_t bk[kKS];
#pragma unroll
      for (int j = 0; j < kKS; j++) bk[j] = cat8(h[2 * j], h[2 * j + 1]);
      f32x4_t acc[kNT];
#pragma unroll
      for (int mt = 0; mt < kNT; mt++) acc[mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < kKS; j++)
#pragma unroll
        for (int mt = 0; mt < kNT; mt++)
          acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(layer ? A2[mt][j] : A1[mt][j], bk[j], acc[mt],
                                                            0, 0, 0);
#pragma unroll
      for (int mt = 0; mt < kNT; mt++) {
        const float* bb = layer ? bb2[mt] : bb1[mt];
        h[mt] = pack4(relu(acc[mt][0] + bb[0]), relu(acc[mt][1] + bb[1]),
                      relu(acc[mt][2] + bb[2]), relu(acc[mt][3] + bb[3]));
      }
    }
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < kKS; j++)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A3[j], cat8(h[2 * j], h[2 * j + 1]), acc, 0, 0, 0);
    return acc[0] + b3;
  }
};

// Inputs of face i for the lanes of a tile: group 0 (pL, pR, rhoL, rhoR),
// group 1 (uL, uR, pL, pR); 1.0 past the end.
__device__ inline void tile_inputs(const double* __restrict__ in, int64_t n, int64_t i, bool live, double* v) {
  const int g = (threadIdx.x & 63) >> 4;
  v[0] = v[1] = v[2] = v[3] = 1.0;
  if (!live) return;
  if (g == 0) {
    v[0] = in[2 * n + i];
    v[1] = in[3 * n + i];
    v[2] = in[4 * n + i];
    v[3] = in[5 * n + i];
  } else if (g == 1) {
    v[0] = in[i];
    v[1] = in[n + i];
    v[2] = in[2 * n + i];
    v[3] = in[3 * n + i];
  }
}

// Standardised features 4 g .. 4 g + 3 of a face from its lane's inputs
// (fx: the standardised EOS features 8..11 of the face's pair of media).
__device__ inline void tile_features(const double* v, bool live, const pgdg_mlp_norm& nm, float4 fx, float* x) {
  const int g = (threadIdx.x & 63) >> 4;
  x[0] = x[1] = x[2] = x[3] = 0.f;
  if (!live) return;
  if (g == 0) {
#pragma unroll
    for (int q = 0; q < 4; q++)
      x[q] = (ln_fast(fmaxf((float)v[q], 1e-30f)) - nm.feat_mu[q]) * __builtin_amdgcn_rcpf(nm.feat_sd[q]);
  } else if (g == 1) {
    x[0] = ((float)v[0] - nm.feat_mu[4]) * __builtin_amdgcn_rcpf(nm.feat_sd[4]);
    x[1] = ((float)v[1] - nm.feat_mu[5]) * __builtin_amdgcn_rcpf(nm.feat_sd[5]);
    x[2] = (p_contrast(v[2], v[3]) - nm.feat_mu[6]) * __builtin_amdgcn_rcpf(nm.feat_sd[6]);
    x[3] = ((float)(v[0] - v[1]) - nm.feat_mu[7]) * __builtin_amdgcn_rcpf(nm.feat_sd[7]);
  } else if (g == 2) {
    x[0] = fx.x;
    x[1] = fx.y;
    x[2] = fx.z;
    x[3] = fx.w;
  }
}

// Initial p* from the network output y (lanes of group 0): max(pL, pR)
// exp(y sd + mu), or max(pL, pR) where the predicted correction is within
// the network's residual (nm.keep) or the product is not a positive finite
// value.
__device__ inline double guess_of(float y, const double* v, const pgdg_mlp_norm& nm) {
  const double pm = v[0] >= v[1] ? v[0] : v[1];
  const float lr = y * nm.out_sd[0] + nm.out_mu[0];
  if (fabsf(lr) < nm.keep) return pm;
  double gs = pm * (double)exp_fast(lr);
  if (!(gs > 0) || !(gs < 1e300)) gs = pm;
  return gs;
}

__global__ void __launch_bounds__(kThreads) k_mlp_guess(const float* __restrict__ P, pgdg_mlp_norm nm,
                                                        float4 fx, const double* __restrict__ in,
                                                        int64_t n, double* __restrict__ pstar0) {
  const int l = threadIdx.x & 63, g = l >> 4, f = l & 15;
  MlpRegs R;
  R.load(P);
  const int64_t ntile = (n + 15) / 16;
  const int64_t wstride = (int64_t)gridDim.x * (kThreads / 64);
  int64_t tile = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  // the next tile's inputs are loaded while the current one is computed
  double cur[4], nxt[4];
  tile_inputs(in, n, tile * 16 + f, tile < ntile && tile * 16 + f < n, cur);
  for (; tile < ntile; tile += wstride) {
    const int64_t tn = tile + wstride;
    tile_inputs(in, n, tn * 16 + f, tn < ntile && tn * 16 + f < n, nxt);
    const int64_t i = tile * 16 + f;
    const bool live = i < n;
    float x[4];
    tile_features(cur, live, nm, fx, x);
    const float y = R.run(x);
    if (g == 0 && live) pstar0[i] = guess_of(y, cur, nm);
#pragma unroll
    for (int q = 0; q < 4; q++) cur[q] = nxt[q];
  }
}

// MGFM batch with the learned initial guess in one kernel: a wavefront owns
// kLearnedChunk consecutive faces; it first runs the network over the
// chunk's 16-face tiles (operands in registers) and keeps the guesses in
// LDS, then runs mgfm_batch_kernel's work-sharing Newton loop over the chunk
// (a converged lane takes the next face), each face starting from its guess.
// No guess array passes through memory.
constexpr int kLearnedChunk = 1024;

__global__ void __launch_bounds__(kThreads) k_mgfm_learned(
    const float* __restrict__ P, pgdg_mlp_norm nm, float4 fx, const double* __restrict__ in,
    double* __restrict__ out, int32_t* __restrict__ iters, int64_t n, Eos ml, Eos mr, int variant,
    int tol_mode, double rtol, int max_iter) {
  __shared__ double s_g[kThreads / 64][kLearnedChunk];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, f = lane & 15;
  const int64_t wave = (int64_t)blockIdx.x * (kThreads / 64) + w;
  const int64_t c0 = wave * kLearnedChunk;
  const int64_t c1 = c0 + kLearnedChunk < n ? c0 + kLearnedChunk : n;
  if (c0 < n) {
    MlpRegs R;
    R.load(P);
    const int nt = (int)((c1 - c0 + 15) / 16);
    double cur[4], nxt[4];
    tile_inputs(in, n, c0 + f, c0 + f < c1, cur);
    for (int t = 0; t < nt; t++) {
      const int64_t in_next = c0 + (int64_t)(t + 1) * 16 + f;
      tile_inputs(in, n, in_next, t + 1 < nt && in_next < c1, nxt);
      const int64_t i = c0 + (int64_t)t * 16 + f;
      const bool live = i < c1;
      float x[4];
      tile_features(cur, live, nm, fx, x);
      const float y = R.run(x);
      if (g == 0 && live) s_g[w][t * 16 + f] = guess_of(y, cur, nm);
#pragma unroll
      for (int q = 0; q < 4; q++) cur[q] = nxt[q];
    }
  }
  __syncthreads();  // the guesses are read by other lanes of the wavefront
  if (c0 >= n) return;
  int64_t next = c0;  // wave-uniform
  int64_t i = -1;
  bool done = false;
  MgfmState S;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (;;) {
    const bool need = !done && i < 0;
    const unsigned long long needm = __ballot(need);
    if (needm) {
      if (need) {
        const int64_t j = next + __popcll(needm & below);
        if (j < c1) {
          i = j;
          mgfm_begin(S, in[j], in[n + j], in[2 * n + j], in[3 * n + j], in[4 * n + j], in[5 * n + j],
                     variant, s_g[w][j - c0]);
        } else {
          done = true;
        }
      }
      next += __popcll(needm);
    }
    if (__ballot(!done) == 0ull) break;
    if (i >= 0 && !mgfm_iterate(S, ml, mr, variant, tol_mode, rtol, max_iter)) {
      const MgfmOut o = mgfm_end(S, ml, variant, tol_mode);
      out[i] = o.rho_l;
      out[n + i] = o.pstar;
      out[2 * n + i] = o.u_l;
      if (iters) iters[i] = o.iters;
      i = -1;
    }
  }
}

// MGFM over faces with per-face EOS pairs (pair[i] into the table), the
// learned initial guess optional (P == nullptr: the variant's own guess).
// Faces i < n of SoA arrays with leading dimension ld (in[6][ld],
// out[3][ld]); n is min(*count, ld) when count is given (read on the
// device, so a captured launch follows the count of each replay).  Same
// wavefront chunking, guess kept in LDS and lane-refilling Newton loop as
// k_mgfm_learned; every face runs exactly the iterations of mgfm_solve from
// its starting point.
template <bool kNet>
__global__ void __launch_bounds__(kThreads) k_mgfm_pairs(
    const float* __restrict__ P, pgdg_mlp_norm nm, PairTab tab, const double* __restrict__ in,
    const int32_t* __restrict__ pair, double* __restrict__ out, int32_t* __restrict__ iters, int64_t n,
    int64_t ld, const int32_t* __restrict__ count, int variant, int tol_mode, double rtol, int max_iter) {
  __shared__ double s_g[kNet ? kThreads / 64 : 1][kNet ? kLearnedChunk : 1];
  __shared__ PairTab s_tab;
  if (threadIdx.x == 0) s_tab = tab;
  if (count) {
    const int64_t cn = *count;
    n = cn < ld ? cn : ld;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, f = lane & 15;
  const int64_t wave = (int64_t)blockIdx.x * (kThreads / 64) + w;
  const int64_t c0 = wave * kLearnedChunk;
  const int64_t c1 = c0 + kLearnedChunk < n ? c0 + kLearnedChunk : n;
  __syncthreads();
  if constexpr (kNet) {
    if (c0 < n) {
      MlpRegs R;
      R.load(P);
      const int nt = (int)((c1 - c0 + 15) / 16);
      for (int t = 0; t < nt; t++) {
        const int64_t i = c0 + (int64_t)t * 16 + f;
        const bool live = i < c1;
        double v[4];
        tile_inputs(in, ld, i, live, v);
        const int c = live && pair ? pair[i] : 0;
        const float4 fx = make_float4(s_tab.fx[c][0], s_tab.fx[c][1], s_tab.fx[c][2], s_tab.fx[c][3]);
        float x[4];
        tile_features(v, live, nm, fx, x);
        const float y = R.run(x);
        if (g == 0 && live) s_g[w][t * 16 + f] = guess_of(y, v, nm);
      }
    }
    __syncthreads();  // the guesses are read by other lanes of the wavefront
  }
  if (c0 >= n) return;
  int64_t next = c0;  // wave-uniform
  int64_t i = -1;
  bool done = false;
  MgfmState S;
  Eos ml{1.4, 0.0}, mr{1.4, 0.0};
  const unsigned long long below = (1ull << lane) - 1ull;
  for (;;) {
    const bool need = !done && i < 0;
    const unsigned long long needm = __ballot(need);
    if (needm) {
      if (need) {
        const int64_t j = next + __popcll(needm & below);
        if (j < c1) {
          i = j;
          const int c = pair ? pair[j] : 0;
          ml = s_tab.l[c];
          mr = s_tab.r[c];
          mgfm_begin(S, in[j], in[ld + j], in[2 * ld + j], in[3 * ld + j], in[4 * ld + j], in[5 * ld + j],
                     variant, kNet ? s_g[w][j - c0] : -1.0);
        } else {
          done = true;
        }
      }
      next += __popcll(needm);
    }
    if (__ballot(!done) == 0ull) break;
    if (i >= 0 && !mgfm_iterate(S, ml, mr, variant, tol_mode, rtol, max_iter)) {
      const MgfmOut o = mgfm_end(S, ml, variant, tol_mode);
      out[i] = o.rho_l;
      out[ld + i] = o.pstar;
      out[2 * ld + i] = o.u_l;
      if (iters) iters[i] = o.iters;
      i = -1;
    }
  }
}

// Launch of k_mgfm_pairs over capacity ld (count on the device) or n faces.
int launch_mgfm_pairs(const float* P, const pgdg_mlp_norm& nm, const PairTab& tab, const double* in,
                      const int32_t* pair, double* out, int32_t* iters, int64_t n, int64_t ld,
                      const int32_t* count, int variant, int tol_mode, double rtol, int max_iter,
                      hipStream_t st) {
  const int64_t span = count ? ld : n;
  if (span <= 0) return PGDG_OK;
  const int64_t waves = (span + kLearnedChunk - 1) / kLearnedChunk;
  const unsigned grid = (unsigned)((waves + kThreads / 64 - 1) / (kThreads / 64));
  if (P)
    hipLaunchKernelGGL(k_mgfm_pairs<true>, dim3(grid), dim3(kThreads), 0, st, P, nm, tab, in, pair, out,
                       iters, n, ld, count, variant, tol_mode, rtol, max_iter);
  else
    hipLaunchKernelGGL(k_mgfm_pairs<false>, dim3(grid), dim3(kThreads), 0, st, P, nm, tab, in, pair, out,
                       iters, n, ld, count, variant, tol_mode, rtol, max_iter);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

// forward on standardised inputs (y[n][2] standardised)
__global__ void __launch_bounds__(kThreads) k_mlp_forward(const float* __restrict__ P,
                                                          const float* __restrict__ x, int64_t n,
                                                          float* __restrict__ y) {
  __shared__ FwdLds s;
  load_params(P, s);
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int q = threadIdx.x; q < kTile * kInPad; q += kThreads) {
    const int64_t i = base + q / kInPad;
    s.X[q] = f2bf(i < n ? x[(base * kInPad) + q] : 0.f);
  }
  __syncthreads();
  forward(s, [&](int row, int col, float v) {
    const int64_t i = base + row;
    if (col < kOut && i < n) y[i * kOut + col] = v;
  });
}

// ---------------------------------------------------------------------------
// training: forward + backward of 64 samples, partial gradient per block
// ---------------------------------------------------------------------------
// Writes the tile of dW (rows m0.., cols n0..) of a [rows][ncols] weight.
__device__ inline void put_tile(float* g, int ncols, int m0, int n0, f32x4_t acc) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 4; i++) g[(m0 + (l >> 4) * 4 + i) * ncols + n0 + (l & 15)] = acc[i];
}

// Hidden-layer backward: from DZ (BF16 dL/dZ of the layer) compute
// dW = Hin^T DZ and db = column sums of DZ; if `below`, also dH = DZ W^T and
// DZ <- dH * (Hin > 0) for the layer underneath (Hin is its output).
__device__ inline void backward_layer(TrainLds& s, const unsigned short* Hin, int kin,
                                      const unsigned short* W, float* gW, float* gb, bool below) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  // dW[k][n] = sum_s Hin[s][k] DZ[s][n]
  const int ntile_k = kin / 16;
  for (int t = w; t < ntile_k * kNT; t += 4) {
    const int m0 = (t / kNT) * 16, n0 = (t % kNT) * 16;
    put_tile(gW, kH, m0, n0, mma_tile(Hin, 1, kin, s.DZ, kH, 1, m0, n0, kTile));
  }
  if (threadIdx.x < kH) {
    float acc = 0.f;
    for (int r = 0; r < kTile; r++) acc += bf2f(s.DZ[r * kH + threadIdx.x]);
    gb[threadIdx.x] = acc;
  }
  if (!below) return;
  // dH[s][k] = sum_n DZ[s][n] W[k][n]
  f32x4_t dh[kNT];
#pragma unroll
  for (int q = 0; q < kNT; q++) {
    const int t = w + 4 * q;
    const int m0 = (t / kNT) * 16, n0 = (t % kNT) * 16;
    dh[q] = mma_tile(s.DZ, kH, 1, W, 1, kH, m0, n0, kH);
  }
  __syncthreads();  // every read of DZ is done
#pragma unroll
  for (int q = 0; q < kNT; q++) {
    const int t = w + 4 * q;
    const int m0 = (t / kNT) * 16, n0 = (t % kNT) * 16;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = m0 + (l >> 4) * 4 + i, col = n0 + (l & 15);
      const float h = bf2f(Hin[row * kH + col]);
      s.DZ[row * kH + col] = f2bf(h > 0.f ? dh[q][i] : 0.f);
    }
  }
  __syncthreads();
}

// x[n][16] standardised features, t[n][2] standardised targets; samples
// [b0, b0 + gridDim.x * 64) (all < n); inv_b = 1 / batch size.  Partial
// gradient of block b at part[b][kNParam], partial loss at lpart[b].
__global__ void __launch_bounds__(kThreads) k_mlp_fwd_bwd(const float* __restrict__ P,
                                                          const float* __restrict__ x,
                                                          const float* __restrict__ tg, int64_t b0,
                                                          float inv_b, float* __restrict__ part,
                                                          float* __restrict__ lpart) {
  __shared__ TrainLds s;
  load_params(P, s.f);
  const int64_t base = b0 + (int64_t)blockIdx.x * kTile;
  for (int q = threadIdx.x; q < kTile * kInPad; q += kThreads) s.f.X[q] = f2bf(x[base * kInPad + q]);
  __syncthreads();
  forward(s.f, [&](int row, int col, float y) {
    float dy = 0.f;
    if (col < kOut) {
      const float d = y - tg[(base + row) * kOut + col];
      dy = d * inv_b;
    }
    s.DYf[row * kOutPad + col] = dy;
    s.DY[row * kOutPad + col] = f2bf(dy);
  });
  __syncthreads();
  if (threadIdx.x < kTile) {
    // 0.5 * sum_o d^2 of this sample (d = dy / inv_b)
    float a = 0.f;
    for (int o = 0; o < kOut; o++) {
      const float d = s.DYf[threadIdx.x * kOutPad + o] / inv_b;
      a += 0.5f * d * d;
    }
    s.lossv[threadIdx.x] = a;
  }
  float* g = part + (size_t)blockIdx.x * kNParam;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  // output layer: dW3[k][o] = sum_s H2[s][k] DY[s][o]; db3; dH2 = DY W3^T,
  // times the activation's derivative (H2 > 0)
  for (int t = w; t < kNT; t += 4)
    put_tile(g + oW3, kOutPad, t * 16, 0, mma_tile(s.f.H2, 1, kH, s.DY, kOutPad, 1, t * 16, 0, kTile));
  if (threadIdx.x < kOutPad) {
    float acc = 0.f;
    for (int r = 0; r < kTile; r++) acc += s.DYf[r * kOutPad + threadIdx.x];
    g[oB3 + threadIdx.x] = acc;
  }
  if (threadIdx.x == 0) {
    float a = 0.f;
    for (int r = 0; r < kTile; r++) a += s.lossv[r];
    lpart[blockIdx.x] = a;
  }
  {
    f32x4_t dh[kNT];
#pragma unroll
    for (int q = 0; q < kNT; q++) {
      const int t = w + 4 * q;
      const int m0 = (t / kNT) * 16, n0 = (t % kNT) * 16;
      dh[q] = mma_tile(s.DY, kOutPad, 1, s.f.W3, 1, kOutPad, m0, n0, kOutPad);
    }
#pragma unroll
    for (int q = 0; q < kNT; q++) {
      const int t = w + 4 * q;
      const int m0 = (t / kNT) * 16, n0 = (t % kNT) * 16;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = m0 + (l >> 4) * 4 + i, col = n0 + (l & 15);
        const float h = bf2f(s.f.H2[row * kH + col]);
        s.DZ[row * kH + col] = f2bf(h > 0.f ? dh[q][i] : 0.f);
      }
    }
  }
  __syncthreads();
  backward_layer(s, s.f.H1, kH, s.f.W2, g + oW2, g + oB2, true);
  backward_layer(s, s.f.H0, kH, s.f.W1, g + oW1, g + oB1, true);
  backward_layer(s, s.f.X, kInPad, s.f.W0, g + oW0, g + oB0, false);
}

// Sum the block partials in block order (padded parameters get zero) and
// the block losses, scaled to the mean.
__global__ void k_mlp_reduce(const float* __restrict__ part, const float* __restrict__ lpart, int nb,
                             float lscale, float* __restrict__ grad, float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNParam) {
    float acc = 0.f;
    for (int b = 0; b < nb; b++) acc += part[(size_t)b * kNParam + p];
    grad[p] = param_live(p) ? acc : 0.f;
  }
  if (p == 0 && loss) {
    float a = 0.f;
    for (int b = 0; b < nb; b++) a += lpart[b];
    *loss = a * lscale;
  }
}

__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, float lr, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam || !param_live(p)) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gp = grad[p];
  const float mp = b1 * m[p] + (1.f - b1) * gp;
  const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// training set: standardised features and targets from labelled faces
__global__ void k_mlp_build_set(const double* __restrict__ in, const double* __restrict__ out,
                                int64_t n, const int32_t* __restrict__ pair, PairTab tab, pgdg_mlp_norm nm,
                                float* __restrict__ x, float* __restrict__ t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = pair ? pair[i] : 0;
  float f[kInPad];
  face_features(in, n, i, tab.l[c], tab.r[c], nm, f);
#pragma unroll
  for (int k = 0; k < kInPad; k++) x[i * kInPad + k] = f[k];
  const double pl = in[2 * n + i], pr = in[3 * n + i];
  const double pm = pl >= pr ? pl : pr;
  const double y0 = log(out[n + i]) - log(pm), y1 = out[2 * n + i];
  t[i * kOut] = (float)((y0 - nm.out_mu[0]) / nm.out_sd[0]);
  t[i * kOut + 1] = (float)((y1 - nm.out_mu[1]) / nm.out_sd[1]);
}

// raw feature / target moments: per-block FP64 sums of (v, v^2) for the
// 12 features and 2 targets
__global__ void k_mlp_moments(const double* __restrict__ in, const double* __restrict__ out,
                              int64_t n, const int32_t* __restrict__ pair, PairTab tab,
                              double* __restrict__ part) {
  constexpr int NV = kIn + kOut;
  __shared__ double red[kThreads];
  double acc[2 * NV];
  for (int q = 0; q < 2 * NV; q++) acc[q] = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double ul = in[i], ur = in[n + i], pl = in[2 * n + i], pr = in[3 * n + i];
    const double rl = in[4 * n + i], rr = in[5 * n + i];
    const double pm = pl >= pr ? pl : pr;
    const int c = pair ? pair[i] : 0;
    const Eos ml = tab.l[c], mr = tab.r[c];
    const double v[NV] = {log(fmax(pl, 1e-30)), log(fmax(pr, 1e-30)), log(fmax(rl, 1e-30)),
                          log(fmax(rr, 1e-30)), ul, ur, pl + pr > 0 ? (pl - pr) / (pl + pr) : 0.0, ul - ur,
                          ml.gamma, log1p(fmax(ml.pinf, 0.0)), mr.gamma, log1p(fmax(mr.pinf, 0.0)),
                          log(out[n + i]) - log(pm), out[2 * n + i]};
    for (int q = 0; q < NV; q++) {
      acc[q] += v[q];
      acc[NV + q] += v[q] * v[q];
    }
  }
  for (int q = 0; q < 2 * NV; q++) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = 0;
      for (int t = 0; t < kThreads; t++) a += red[t];
      part[(size_t)blockIdx.x * 2 * NV + q] = a;
    }
    __syncthreads();
  }
}

// per-block FP64 sums of (y0 - t0)^2: the residual of the standardised
// ln(p* / max(pL, pR)) output over a training set
__global__ void k_mlp_resid(const float* __restrict__ y, const float* __restrict__ t, int64_t n,
                            double* __restrict__ part) {
  __shared__ double red[kThreads];
  double acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double d = (double)y[i * kOut] - (double)t[i * kOut];
    acc += d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0;
    for (int q = 0; q < kThreads; q++) a += red[q];
    part[blockIdx.x] = a;
  }
}

}  // namespace pgdg

struct pgdg_mlp {
  float* params = nullptr;  // device FP32 master parameters
  float* m = nullptr;
  float* v = nullptr;
  float* grad = nullptr;
  float* part = nullptr;
  float* lpart = nullptr;
  float* loss = nullptr;
  int part_blocks = 0;
  pgdg_mlp_norm norm{};
  bool ready = false;
};

namespace {

int ensure_partials(pgdg_mlp* m, int nb) {
  if (nb <= m->part_blocks) return PGDG_OK;
  if (m->part) (void)hipFree(m->part);
  if (m->lpart) (void)hipFree(m->lpart);
  m->part = nullptr;
  m->lpart = nullptr;
  PGDG_HIP_TRY(hipMalloc((void**)&m->part, sizeof(float) * (size_t)nb * pgdg::kNParam));
  PGDG_HIP_TRY(hipMalloc((void**)&m->lpart, sizeof(float) * (size_t)nb));
  m->part_blocks = nb;
  return PGDG_OK;
}

int launch_fwd_bwd(pgdg_mlp* m, const float* x, const float* t, int64_t b0, int64_t cnt, float* grad,
                   float* loss, hipStream_t st) {
  using namespace pgdg;
  const int nb = (int)(cnt / kTile);
  PGDG_TRY(ensure_partials(m, nb));
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(kThreads), 0, st, m->params, x, t, b0,
                     1.0f / (float)cnt, m->part, m->lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParam + 255) / 256), dim3(256), 0, st, m->part,
                     m->lpart, nb, 1.0f / (float)cnt, grad, loss);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

void init_params(std::vector<float>& P, uint64_t seed) {
  using namespace pgdg;
  std::mt19937_64 rng(seed);
  P.assign(kNParam, 0.f);
  auto fill = [&](int off, int rows, int cols, int live_rows, int live_cols, int fan_in,
                  int fan_out) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    std::uniform_real_distribution<double> U(-a, a);
    for (int r = 0; r < rows; r++)
      for (int c = 0; c < cols; c++) {
        double v = U(rng);
        P[off + r * cols + c] = (r < live_rows && c < live_cols) ? (float)v : 0.f;
      }
  };
  fill(oW0, kInPad, kH, kIn, kH, kIn, kH);
  fill(oW1, kH, kH, kH, kH, kH, kH);
  fill(oW2, kH, kH, kH, kH, kH, kH);
  fill(oW3, kH, kOutPad, kH, kOut, kH, kOut);
}

// The standardised EOS features 6..9 of a face pair (the same for every
// face of a batch).
float4 eos_features(const pgdg_mlp_norm& nm, const pgdg_eos_pair* eos) {
  const double c[4] = {eos->gamma_l, std::log1p(std::max(eos->pinf_l, 0.0)), eos->gamma_r,
                       std::log1p(std::max(eos->pinf_r, 0.0))};
  float fx4[4];
  for (int k = 0; k < 4; k++) fx4[k] = ((float)c[k] - nm.feat_mu[8 + k]) / nm.feat_sd[8 + k];
  return make_float4(fx4[0], fx4[1], fx4[2], fx4[3]);
}

}  // namespace

namespace pgdg {

PairTab make_pair_tab(const pgdg_mlp_norm& nm, const pgdg_eos_pair* eos, int np) {
  PairTab t{};
  for (int c = 0; c < np; c++) {
    t.l[c] = pgdg::Eos{eos[c].gamma_l, eos[c].pinf_l};
    t.r[c] = pgdg::Eos{eos[c].gamma_r, eos[c].pinf_r};
    const float4 fx = eos_features(nm, &eos[c]);
    t.fx[c][0] = fx.x;
    t.fx[c][1] = fx.y;
    t.fx[c][2] = fx.z;
    t.fx[c][3] = fx.w;
  }
  return t;
}

int mlp_view(pgdg_mlp* m, const float** params, pgdg_mlp_norm* nm) {
  if (!m) return PGDG_ERR_ARG;
  if (!m->ready) return PGDG_ERR_STATE;
  *params = m->params;
  *nm = m->norm;
  return PGDG_OK;
}

}  // namespace pgdg

extern "C" {

int pgdg_mlp_create(pgdg_mlp** out) {
  using namespace pgdg;
  if (!out) return PGDG_ERR_ARG;
  pgdg_mlp* m = new (std::nothrow) pgdg_mlp();
  if (!m) return PGDG_ERR_CAPACITY;
  auto fail = [&](int r) {
    pgdg_mlp_destroy(m);
    return r;
  };
  if (hipMalloc((void**)&m->params, sizeof(float) * kNParam) != hipSuccess ||
      hipMalloc((void**)&m->m, sizeof(float) * kNParam) != hipSuccess ||
      hipMalloc((void**)&m->v, sizeof(float) * kNParam) != hipSuccess ||
      hipMalloc((void**)&m->grad, sizeof(float) * kNParam) != hipSuccess ||
      hipMalloc((void**)&m->loss, sizeof(float) * 4) != hipSuccess)
    return fail(PGDG_ERR_HIP);
  std::vector<float> P;
  init_params(P, 0x5EED0004ull);
  if (hipMemcpy(m->params, P.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice) != hipSuccess)
    return fail(PGDG_ERR_HIP);
  for (int k = 0; k < kIn; k++) {
    m->norm.feat_mu[k] = 0.f;
    m->norm.feat_sd[k] = 1.f;
  }
  for (int k = 0; k < kOut; k++) {
    m->norm.out_mu[k] = 0.f;
    m->norm.out_sd[k] = 1.f;
  }
  *out = m;
  return PGDG_OK;
}

int pgdg_mlp_destroy(pgdg_mlp* m) {
  if (!m) return PGDG_ERR_ARG;
  void* ptrs[] = {m->params, m->m, m->v, m->grad, m->part, m->lpart, m->loss};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete m;
  return PGDG_OK;
}

int pgdg_mlp_get_params(pgdg_mlp* m, float* params, pgdg_mlp_norm* norm) {
  if (!m) return PGDG_ERR_ARG;
  if (params)
    PGDG_HIP_TRY(hipMemcpy(params, m->params, sizeof(float) * pgdg::kNParam, hipMemcpyDeviceToHost));
  if (norm) *norm = m->norm;
  return PGDG_OK;
}

int pgdg_mlp_set_params(pgdg_mlp* m, const float* params, const pgdg_mlp_norm* norm) {
  using namespace pgdg;
  if (!m) return PGDG_ERR_ARG;
  if (params) {
    std::vector<float> P(params, params + kNParam);
    for (int p = 0; p < kNParam; p++)
      if (!param_live(p)) P[p] = 0.f;
    PGDG_HIP_TRY(hipMemcpy(m->params, P.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice));
  }
  if (norm) {
    for (int k = 0; k < kIn; k++)
      if (!(norm->feat_sd[k] > 0)) return PGDG_ERR_ARG;
    for (int k = 0; k < kOut; k++)
      if (!(norm->out_sd[k] > 0)) return PGDG_ERR_ARG;
    if (!(norm->keep >= 0)) return PGDG_ERR_ARG;
    m->norm = *norm;
  }
  m->ready = true;
  return PGDG_OK;
}

int pgdg_mlp_forward(pgdg_mlp* m, const float* x, float* y, int64_t n, void* stream) {
  using namespace pgdg;
  if (!m || n < 0 || (n > 0 && (!x || !y))) return PGDG_ERR_ARG;
  if (n == 0) return PGDG_OK;
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + kTile - 1) / kTile)), dim3(kThreads), 0,
                     as_stream(stream), m->params, x, n, y);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int pgdg_mlp_loss_grad(pgdg_mlp* m, const float* x, const float* t, int64_t n, float* grad,
                       float* loss, void* stream) {
  using namespace pgdg;
  if (!m || !x || !t || !grad || n <= 0 || n % kTile) return PGDG_ERR_ARG;
  if (n / kTile > (1 << 20)) return PGDG_ERR_ARG;
  return launch_fwd_bwd(m, x, t, 0, n, grad, loss ? loss : m->loss, as_stream(stream));
}

int pgdg_mlp_train(pgdg_mlp* m, const double* in, int64_t n, const pgdg_eos_pair* eos,
                   const pgdg_mlp_train_opts* o, float* loss_hist, int hist_cap, int* steps,
                   void* stream) {
  return pgdg_mlp_train_pairs(m, in, nullptr, n, eos, 1, o, loss_hist, hist_cap, steps, stream);
}

int pgdg_mlp_train_pairs(pgdg_mlp* m, const double* in, const int32_t* pair, int64_t n,
                         const pgdg_eos_pair* eos, int npairs, const pgdg_mlp_train_opts* o,
                         float* loss_hist, int hist_cap, int* steps, void* stream) {
  using namespace pgdg;
  if (!m || !in || !eos || !o || n < kTile) return PGDG_ERR_ARG;
  if (npairs < 1 || npairs > kMaxPairs || (npairs > 1 && !pair)) return PGDG_ERR_ARG;
  if (o->epochs < 1 || o->batch < kTile || o->batch % kTile || !(o->lr > 0)) return PGDG_ERR_ARG;
  if (o->variant != MGFM_S3 && o->variant != MGFM_T2) return PGDG_ERR_ARG;
  if (hist_cap < 0 || (hist_cap > 0 && !loss_hist)) return PGDG_ERR_ARG;
  hipStream_t st = as_stream(stream);
  PairTab tab = make_pair_tab(m->norm, eos, npairs);  // (the features are standardised below)
  const int64_t nt = n / o->batch * o->batch;  // whole batches
  if (nt == 0) return PGDG_ERR_ARG;
  // 1. label with the exact solve
  double* lab = nullptr;
  float *x = nullptr, *t = nullptr, *hist = nullptr;
  double* mom = nullptr;
  int rc = PGDG_OK;
  const int mblocks = 1024;
  constexpr int NV = kIn + kOut;
  auto cleanup = [&]() {
    void* ptrs[] = {lab, x, t, hist, mom};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  };
#define MLP_TRY(expr)                   \
  do {                                  \
    hipError_t e_ = (expr);             \
    if (e_ != hipSuccess) {             \
      cleanup();                        \
      return PGDG_ERR_HIP;              \
    }                                   \
  } while (0)
  MLP_TRY(hipMalloc((void**)&lab, sizeof(double) * 3 * (size_t)n));
  MLP_TRY(hipMalloc((void**)&x, sizeof(float) * kInPad * (size_t)n));
  MLP_TRY(hipMalloc((void**)&t, sizeof(float) * kOut * (size_t)n));
  MLP_TRY(hipMalloc((void**)&mom, sizeof(double) * 2 * NV * mblocks));
  rc = launch_mgfm_pairs(nullptr, m->norm, tab, in, pair, lab, nullptr, n, n, nullptr, o->variant, TOL_REL,
                         1e-12, 100, st);
  if (rc != PGDG_OK) {
    cleanup();
    return rc;
  }
  // 2. normalisation from the labelled set (block sums added in block order)
  hipLaunchKernelGGL(k_mlp_moments, dim3(mblocks), dim3(kThreads), 0, st, in, lab, n, pair, tab, mom);
  std::vector<double> hm((size_t)2 * NV * mblocks);
  MLP_TRY(hipMemcpyAsync(hm.data(), mom, sizeof(double) * hm.size(), hipMemcpyDeviceToHost, st));
  MLP_TRY(hipStreamSynchronize(st));
  double s1[NV] = {0}, s2[NV] = {0};
  for (int b = 0; b < mblocks; b++)
    for (int q = 0; q < NV; q++) {
      s1[q] += hm[(size_t)b * 2 * NV + q];
      s2[q] += hm[(size_t)b * 2 * NV + NV + q];
    }
  pgdg_mlp_norm nm{};
  for (int q = 0; q < NV; q++) {
    const double mu = s1[q] / (double)n;
    const double var = std::max(0.0, s2[q] / (double)n - mu * mu);
    double sd = std::sqrt(var);
    if (!(sd > 1e-6 * std::max(1.0, std::fabs(mu)))) sd = 1.0;
    if (q < kIn) {
      nm.feat_mu[q] = (float)mu;
      nm.feat_sd[q] = (float)sd;
    } else {
      nm.out_mu[q - kIn] = (float)mu;
      nm.out_sd[q - kIn] = (float)sd;
    }
  }
  m->norm = nm;
  tab = make_pair_tab(nm, eos, npairs);
  hipLaunchKernelGGL(k_mlp_build_set, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, lab,
                     n, pair, tab, nm, x, t);
  // 3. parameters
  std::vector<float> P;
  init_params(P, o->seed);
  MLP_TRY(hipMemcpyAsync(m->params, P.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice, st));
  MLP_TRY(hipMemsetAsync(m->m, 0, sizeof(float) * kNParam, st));
  MLP_TRY(hipMemsetAsync(m->v, 0, sizeof(float) * kNParam, st));
  // 4. Adam over minibatches, cosine-decayed step size
  const int64_t per_epoch = nt / o->batch;
  const int64_t total = per_epoch * o->epochs;
  const int nrec = hist_cap > 0 ? (int)std::min<int64_t>(total, hist_cap) : 0;
  if (nrec > 0) MLP_TRY(hipMalloc((void**)&hist, sizeof(float) * nrec));
  int64_t step = 0;
  for (int ep = 0; ep < o->epochs; ep++)
    for (int64_t bi = 0; bi < per_epoch; bi++, step++) {
      float* lslot = (step < nrec) ? hist + step : m->loss;
      rc = launch_fwd_bwd(m, x, t, bi * o->batch, o->batch, m->grad, lslot, st);
      if (rc != PGDG_OK) {
        cleanup();
        return rc;
      }
      const double frac = (double)step / (double)std::max<int64_t>(1, total - 1);
      const double lr = o->lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * frac)));
      const double c1 = 1.0 - std::pow(0.9, (double)(step + 1));
      const double c2 = 1.0 - std::pow(0.999, (double)(step + 1));
      hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, st, m->params, m->m,
                         m->v, m->grad, (float)lr, (float)c1, (float)c2);
    }
  MLP_TRY(hipGetLastError());
  if (nrec > 0)
    MLP_TRY(hipMemcpyAsync(loss_hist, hist, sizeof(float) * nrec, hipMemcpyDeviceToHost, st));
  // 5. the trained network's residual on its training faces sets where its
  //    p* guess is used (beyond twice the residual RMS of ln p* / max(pL, pR))
  {
    float* y = nullptr;
    MLP_TRY(hipMalloc((void**)&y, sizeof(float) * kOut * (size_t)nt));
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((nt + kTile - 1) / kTile)), dim3(kThreads), 0, st,
                       m->params, x, nt, y);
    hipLaunchKernelGGL(k_mlp_resid, dim3(mblocks), dim3(kThreads), 0, st, y, t, nt, mom);
    std::vector<double> hr(mblocks);
    const hipError_t e1 = hipMemcpyAsync(hr.data(), mom, sizeof(double) * mblocks, hipMemcpyDeviceToHost, st);
    const hipError_t e2 = hipStreamSynchronize(st);
    (void)hipFree(y);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      cleanup();
      return PGDG_ERR_HIP;
    }
    double ss = 0;
    for (int b = 0; b < mblocks; b++) ss += hr[b];
    m->norm.keep = (float)(2.0 * std::sqrt(ss / (double)nt) * nm.out_sd[0]);
  }
  MLP_TRY(hipStreamSynchronize(st));
  if (steps) *steps = (int)total;
  cleanup();
#undef MLP_TRY
  m->ready = true;
  return PGDG_OK;
}

int pgdg_mlp_guess(pgdg_mlp* m, const double* in, int64_t n, const pgdg_eos_pair* eos,
                   double* pstar0, void* stream) {
  using namespace pgdg;
  if (!m || !eos || n < 0 || (n > 0 && (!in || !pstar0))) return PGDG_ERR_ARG;
  if (!m->ready) return PGDG_ERR_STATE;
  if (n == 0) return PGDG_OK;
  const pgdg_mlp_norm& nm = m->norm;
  const float4 fx = eos_features(nm, eos);
  const int64_t ntile = (n + 15) / 16;
  const unsigned grid = (unsigned)std::min<int64_t>((ntile + 3) / 4, 256 * 8);
  hipLaunchKernelGGL(k_mlp_guess, dim3(grid), dim3(kThreads), 0, as_stream(stream), m->params, nm, fx,
                     in, n, pstar0);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int pgdg_mgfm_batch_learned(pgdg_mlp* m, const double* in, double* out, int32_t* iters, int64_t n,
                            const pgdg_eos_pair* eos, const pgdg_mgfm_opts* opts, void* stream) {
  using namespace pgdg;
  if (!m || !eos || !opts || n < 0) return PGDG_ERR_ARG;
  if (!m->ready) return PGDG_ERR_STATE;
  if (n == 0) return PGDG_OK;
  if (!in || !out) return PGDG_ERR_ARG;
  if (opts->variant != MGFM_S3 && opts->variant != MGFM_T2) return PGDG_ERR_ARG;
  if (opts->tol_mode != TOL_REF && opts->tol_mode != TOL_REL) return PGDG_ERR_ARG;
  if (opts->max_iter < 1) return PGDG_ERR_ARG;
  const pgdg_mlp_norm& nm = m->norm;
  const float4 fx = eos_features(nm, eos);
  const int64_t waves = (n + kLearnedChunk - 1) / kLearnedChunk;
  const unsigned grid = (unsigned)((waves + kThreads / 64 - 1) / (kThreads / 64));
  hipLaunchKernelGGL(k_mgfm_learned, dim3(grid), dim3(kThreads), 0, as_stream(stream), m->params, nm, fx,
                     in, out, iters, n, Eos{eos->gamma_l, eos->pinf_l}, Eos{eos->gamma_r, eos->pinf_r},
                     opts->variant, opts->tol_mode, opts->rtol, opts->max_iter);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int pgdg_mgfm_batch_pairs(pgdg_mlp* m, const double* in, const int32_t* pair, double* out,
                          int32_t* iters, int64_t n, const pgdg_eos_pair* eos, int npairs,
                          const pgdg_mgfm_opts* opts, void* stream) {
  using namespace pgdg;
  if (!eos || !opts || n < 0 || npairs < 1 || npairs > kMaxPairs) return PGDG_ERR_ARG;
  if (m && !m->ready) return PGDG_ERR_STATE;
  if (n == 0) return PGDG_OK;
  if (!in || !out || (npairs > 1 && !pair)) return PGDG_ERR_ARG;
  if (opts->variant != MGFM_S3 && opts->variant != MGFM_T2) return PGDG_ERR_ARG;
  if (opts->tol_mode != TOL_REF && opts->tol_mode != TOL_REL) return PGDG_ERR_ARG;
  if (opts->max_iter < 1) return PGDG_ERR_ARG;
  const pgdg_mlp_norm nm = m ? m->norm : pgdg_mlp_norm{};
  return launch_mgfm_pairs(m ? m->params : nullptr, nm, make_pair_tab(nm, eos, npairs), in, pair, out, iters, n, n,
                           nullptr, opts->variant, opts->tol_mode, opts->rtol, opts->max_iter,
                           as_stream(stream));
}

}  // extern "C"
