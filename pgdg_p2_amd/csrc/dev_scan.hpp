// dev_scan.hpp - exclusive scans and a stable bucket sort for the topology
// passes (binning, cell-list edits, split/merge compaction).
//
// Every count can live on the device: a scan or bucket sort takes its length
// either as a host value or as a pointer to a device int64, so a pass whose
// sizes depend on earlier device work runs without a host round trip.  The
// launches use a fixed grid (kScanBlocks workgroups walking their share of
// the elements), so the host never needs the length to size them.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace pgdg {

constexpr int kScanBlocks = 512;   // workgroups of a partitioned scan
constexpr int kScanThreads = 256;  // lanes per workgroup (4 wavefronts)

// Length of a scan, sort or fill: *dev when dev != nullptr, else host; zero
// when a gate is given and *gate == 0 (a pass with nothing to do then costs
// only the launches).
struct Len {
  const long long* dev;
  long long host;
  const long long* gate;
  __device__ inline bool open() const { return !(gate && *gate == 0); }
  __device__ inline long long get() const {
    if (!open()) return 0;
    return dev ? *dev : host;
  }
};
inline Len len_host(long long n, const long long* gate = nullptr) { return Len{nullptr, n, gate}; }
inline Len len_dev(const long long* p, const long long* gate = nullptr) { return Len{p, 0, gate}; }

// Exclusive prefix sum out[i] = in[0] + ... + in[i-1] for i < n (out may be
// in).  total (device, may be null) receives the sum of all n elements; with
// write_end, out[n] = total as well.  part: kScanBlocks elements of scratch.
// A closed gate writes nothing (total and out[n] included).
template <typename T>
int scan_excl(const T* in, T* out, Len n, long long* total, bool write_end, T* part, hipStream_t st);

// Stable bucket sort of K elements by key[e] in [0, nb): on return
// off[0..nb] are the bucket offsets (off[nb] = K) and order[off[c] ..
// off[c+1]) lists the elements of bucket c in increasing e.  cursor: nb + 1
// ints of scratch; part as for scan_excl.  key[e] for e >= K is not read.
// With a gate, nothing is written when *gate == 0.
int bucket_stable(const uint32_t* key, Len K, int nb, int* off, int* cursor, int* order, int* part,
                  hipStream_t st, const long long* gate = nullptr);

// Fill helpers with a device or host length.
int fill_i32(int* a, Len n, int v, hipStream_t st);
int fill_u8(uint8_t* a, Len n, uint8_t v, hipStream_t st);

}  // namespace pgdg
