// dev_scan.hip - partitioned exclusive scan and stable bucket sort.
//
// Scan: the n elements are split into kScanBlocks contiguous chunks, one per
// workgroup.  Pass 1 reduces each chunk, pass 2 (one workgroup) scans the
// chunk sums, pass 3 rescans each chunk in tiles of 4 elements per lane,
// seeded with its chunk's offset.  Within a tile, lanes scan their 4
// elements serially, wavefronts scan the lane sums with shuffles, and the
// 4 wavefront sums are combined through LDS.
//
// Bucket sort (counting sort by key, stable in element index): a histogram
// with atomic adds, the scan of the counts into bucket offsets, an atomic
// scatter of element indices into their buckets, then each bucket's indices
// sorted in place (buckets hold a few to a few tens of entries: cell bins,
// split children of a cell, appended entries of a cell).  The atomics only
// place an element within its bucket; the final order is the sorted one, so
// the result does not depend on the order in which the atomics land.
#include "dev_scan.hpp"
#include "hip_util.hpp"

namespace pgdg {

namespace {

constexpr int kFillBlocks = 1024;

template <typename T>
__device__ inline T wave_incl_scan(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Exclusive scan of one value per lane over a workgroup of NW wavefronts;
// lds holds NW values.  Returns the lane's exclusive prefix and the total.
template <typename T, int NW>
__device__ inline T block_excl_scan(T v, T* lds, T& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T inc = wave_incl_scan(v);
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < NW; k++) {
    T s = lds[k];
    off += k < w ? s : T(0);
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + inc - v;
}

__device__ inline void chunk_of(long long n, long long& lo, long long& hi) {
  const long long chunk = (n + kScanBlocks - 1) / kScanBlocks;
  lo = (long long)blockIdx.x * chunk;
  if (lo > n) lo = n;
  hi = lo + chunk;
  if (hi > n) hi = n;
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const T* __restrict__ in, Len len, T* part) {
  __shared__ T lds[kScanThreads / 64];
  long long lo, hi;
  chunk_of(len.get(), lo, hi);
  T s = 0;
  for (long long i = lo + threadIdx.x; i < hi; i += kScanThreads) s += in[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
#pragma unroll
    for (int k = 0; k < kScanThreads / 64; k++) t += lds[k];
    part[blockIdx.x] = t;
  }
}

template <typename T>
__global__ void __launch_bounds__(kScanBlocks) k_scan_top(T* part, long long* total, T* out, Len len, int write_end) {
  __shared__ T lds[kScanBlocks / 64];
  if (!len.open()) return;
  T v = part[threadIdx.x];
  T tot;
  T ex = block_excl_scan<T, kScanBlocks / 64>(v, lds, tot);
  part[threadIdx.x] = ex;
  if (threadIdx.x == 0) {
    if (total) *total = (long long)tot;
    if (write_end) out[len.get()] = tot;
  }
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads) k_scan_down(const T* in, T* out, Len len, const T* __restrict__ part) {
  __shared__ T lds[kScanThreads / 64];
  long long lo, hi;
  chunk_of(len.get(), lo, hi);
  T carry = part[blockIdx.x];
  for (long long tile = lo; tile < hi; tile += 4 * kScanThreads) {
    const long long i0 = tile + 4 * (long long)threadIdx.x;
    T v[4];
    T s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      v[k] = i0 + k < hi ? in[i0 + k] : T(0);
      s += v[k];
    }
    T tot;
    T ex = block_excl_scan<T, kScanThreads / 64>(s, lds, tot);
    T run = carry + ex;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (i0 + k < hi) out[i0 + k] = run;
      run += v[k];
    }
    carry += tot;
  }
}

__global__ void k_fill_i32_len(int* a, Len len, int v) {
  const long long n = len.get();
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void k_fill_u8_len(uint8_t* a, Len len, uint8_t v) {
  const long long n = len.get();
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void k_bucket_hist(const uint32_t* __restrict__ key, Len len, int* cnt) {
  const long long n = len.get();
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    atomicAdd(&cnt[key[e]], 1);
}

__global__ void k_copy_i32(const int* __restrict__ src, int* dst, Len len) {
  const long long n = len.get();
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_bucket_scatter(const uint32_t* __restrict__ key, Len len, int* cursor, int* order) {
  const long long n = len.get();
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const int pos = atomicAdd(&cursor[key[e]], 1);
    if (pos < n) order[pos] = (int)e;
  }
}

// Insertion sort of each bucket's element indices (ascending).
__global__ void k_bucket_sort(const int* __restrict__ off, Len nb_len, int* order) {
  const long long nb = nb_len.get();
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < nb; c += (long long)gridDim.x * blockDim.x) {
    const int a = off[c], b = off[c + 1];
    for (int i = a + 1; i < b; i++) {
      const int v = order[i];
      int j = i - 1;
      while (j >= a && order[j] > v) {
        order[j + 1] = order[j];
        j--;
      }
      order[j + 1] = v;
    }
  }
}

}  // namespace

template <typename T>
int scan_excl(const T* in, T* out, Len n, long long* total, bool write_end, T* part, hipStream_t st) {
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3(kScanBlocks), dim3(kScanThreads), 0, st, in, n, part);
  hipLaunchKernelGGL(k_scan_top<T>, dim3(1), dim3(kScanBlocks), 0, st, part, total, out, n,
                     write_end ? 1 : 0);
  hipLaunchKernelGGL(k_scan_down<T>, dim3(kScanBlocks), dim3(kScanThreads), 0, st, in, out, n,
                     (const T*)part);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

template int scan_excl<int>(const int*, int*, Len, long long*, bool, int*, hipStream_t);
template int scan_excl<long long>(const long long*, long long*, Len, long long*, bool, long long*, hipStream_t);

int fill_i32(int* a, Len n, int v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_i32_len, dim3(kFillBlocks), dim3(256), 0, st, a, n, v);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int fill_u8(uint8_t* a, Len n, uint8_t v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u8_len, dim3(kFillBlocks), dim3(256), 0, st, a, n, v);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

int bucket_stable(const uint32_t* key, Len K, int nb, int* off, int* cursor, int* order, int* part,
                  hipStream_t st, const long long* gate) {
  K.gate = gate;
  const Len nb1 = len_host(nb + 1, gate);
  PGDG_TRY(fill_i32(cursor, nb1, 0, st));
  hipLaunchKernelGGL(k_bucket_hist, dim3(kFillBlocks), dim3(256), 0, st, key, K, cursor);
  PGDG_TRY(scan_excl<int>(cursor, off, nb1, nullptr, false, part, st));
  hipLaunchKernelGGL(k_copy_i32, dim3(kFillBlocks), dim3(256), 0, st, (const int*)off, cursor, nb1);
  hipLaunchKernelGGL(k_bucket_scatter, dim3(kFillBlocks), dim3(256), 0, st, key, K, cursor, order);
  hipLaunchKernelGGL(k_bucket_sort, dim3(kFillBlocks * 2), dim3(256), 0, st, (const int*)off,
                     len_host(nb, gate), order);
  PGDG_HIP_TRY(hipGetLastError());
  return PGDG_OK;
}

}  // namespace pgdg
