// mgfm.hip - batched MGFM interface Riemann solve (replaces MPM::MGFM,
// S3:mpm.cpp:882-926 / T2:mpm.cpp:1373-1416, once per face pair).
//
// Layout: SoA FP64 columns so every load and store of a wavefront is one
// contiguous 512-byte segment; one face per lane, grid-stride loop.
#include "hip_util.hpp"
#include "dev_common.hpp"
#include "riemann.hpp"

namespace pgdg {

// Each wavefront owns a chunk of kMgfmChunk consecutive faces and keeps all
// 64 lanes busy: a lane whose face has converged writes it out and takes
// the next face of the chunk (ballot + prefix count, no atomics), so the
// wavefront's time is the chunk's total Newton work rather than 64 times
// its slowest face.  Every face runs exactly the iterations of mgfm_solve.
constexpr int kMgfmChunk = 2048;

__global__ void __launch_bounds__(256) mgfm_batch_kernel(
    const double* __restrict__ in, double* __restrict__ out, int32_t* __restrict__ iters,
    int64_t n, Eos ml, Eos mr, int variant, int tol_mode, double rtol, int max_iter,
    const double* __restrict__ pstar0) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t c0 = wave * kMgfmChunk;
  if (c0 >= n) return;
  const int64_t c1 = c0 + kMgfmChunk < n ? c0 + kMgfmChunk : n;
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
          mgfm_begin(S, in[j], in[n + j], in[2 * n + j], in[3 * n + j], in[4 * n + j],
                     in[5 * n + j], variant, pstar0 ? pstar0[j] : -1.0);
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

}  // namespace pgdg

extern "C" int pgdg_mgfm_batch(const double* in, double* out, int32_t* iters, int64_t n,
                               const pgdg_eos_pair* eos, const pgdg_mgfm_opts* opts,
                               const double* pstar0, void* stream) {
  using namespace pgdg;
  if (n < 0 || !eos || !opts) return PGDG_ERR_ARG;
  if (n == 0) return PGDG_OK;
  if (!in || !out) return PGDG_ERR_ARG;
  if (opts->variant != MGFM_S3 && opts->variant != MGFM_T2) return PGDG_ERR_ARG;
  if (opts->tol_mode != TOL_REF && opts->tol_mode != TOL_REL) return PGDG_ERR_ARG;
  if (opts->max_iter < 1) return PGDG_ERR_ARG;
  Eos ml{eos->gamma_l, eos->pinf_l};
  Eos mr{eos->gamma_r, eos->pinf_r};
  const int block = 256;
  const int64_t waves = (n + kMgfmChunk - 1) / kMgfmChunk;
  const unsigned grid = (unsigned)((waves + block / 64 - 1) / (block / 64));
  hipLaunchKernelGGL(mgfm_batch_kernel, dim3(grid), dim3(block), 0, as_stream(stream), in, out,
                     iters, n, ml, mr, opts->variant, opts->tol_mode, opts->rtol, opts->max_iter,
                     pstar0);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

namespace pgdg {
__global__ void k_root_dim(const double* __restrict__ x, double* __restrict__ y, int64_t n, int dim) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = dim == 2 ? root_d<2>(x[i]) : root_d<3>(x[i]);
}
}  // namespace pgdg

extern "C" int pgdg_root_dim(const double* x, double* y, int64_t n, int dim, void* stream) {
  if (n < 0 || (dim != 2 && dim != 3)) return PGDG_ERR_ARG;
  if (n == 0) return PGDG_OK;
  if (!x || !y) return PGDG_ERR_ARG;
  hipLaunchKernelGGL(pgdg::k_root_dim, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     pgdg::as_stream(stream), x, y, n, dim);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

extern "C" const char* pgdg_version(void) { return "pgdg-mi355x 0.1 gfx950"; }

extern "C" int pgdg_malloc(void** ptr, int64_t bytes) {
  if (!ptr || bytes < 0) return PGDG_ERR_ARG;
  PGDG_HIP_TRY(hipMalloc(ptr, bytes > 0 ? (size_t)bytes : 16));
  return PGDG_OK;
}

extern "C" int pgdg_free(void* ptr) {
  if (ptr) PGDG_HIP_TRY(hipFree(ptr));
  return PGDG_OK;
}

extern "C" int pgdg_memcpy(void* dst, const void* src, int64_t bytes, int kind) {
  if (bytes < 0 || kind < 0 || kind > 2) return PGDG_ERR_ARG;
  if (bytes == 0) return PGDG_OK;
  if (!dst || !src) return PGDG_ERR_ARG;
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  PGDG_HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, k));
  return PGDG_OK;
}

extern "C" int pgdg_device_sync(void) {
  PGDG_HIP_TRY(hipDeviceSynchronize());
  return PGDG_OK;
}

// Plain streaming copy (16-byte vector loads and stores, grid-stride): a
// kernel of known HBM traffic, for calibrating hardware byte counters.
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

extern "C" int pgdg_copy_kernel(void* dst, const void* src, int64_t bytes, void* stream) {
  if (!dst || !src || bytes < 0 || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
    return PGDG_ERR_ARG;
  const int64_t n16 = bytes / 16;
  if (n16 == 0) return PGDG_OK;
  hipLaunchKernelGGL(k_copy16, dim3(pgdg::grid_for(n16, 256, 8192)), dim3(256), 0,
                     pgdg::as_stream(stream), (const uint4*)src, (uint4*)dst, n16);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

extern "C" int pgdg_event_create(void** ev) {
  if (!ev) return PGDG_ERR_ARG;
  hipEvent_t e = nullptr;
  PGDG_HIP_TRY(hipEventCreate(&e));
  *ev = e;
  return PGDG_OK;
}

extern "C" int pgdg_event_record(void* ev, void* stream) {
  if (!ev) return PGDG_ERR_ARG;
  PGDG_HIP_TRY(hipEventRecord((hipEvent_t)ev, pgdg::as_stream(stream)));
  return PGDG_OK;
}

extern "C" int pgdg_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (!start || !stop || !ms) return PGDG_ERR_ARG;
  PGDG_HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
  PGDG_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return PGDG_OK;
}

extern "C" int pgdg_event_destroy(void* ev) {
  if (ev) PGDG_HIP_TRY(hipEventDestroy((hipEvent_t)ev));
  return PGDG_OK;
}

extern "C" int pgdg_device_count(int* count) {
  if (!count) return PGDG_ERR_ARG;
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return PGDG_OK;
}
