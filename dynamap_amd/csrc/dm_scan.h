// Device-wide exclusive scan of int arrays and a workgroup scan helper,
// shared by the mesh and non-rigid translation units; plus a scoped device
// buffer for per-call scratch.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include "dm_session.h"

namespace dm {

constexpr int kScanItems = 1024;  // items per workgroup of the device-wide scans (256 threads x 4)

// Exclusive scan over the NT threads of a workgroup (NT a multiple of 64);
// lds holds NT / 64 + 1 values.  total receives the sum.  All threads call it.
template <int NT, class T>
__device__ T wg_exclusive_scan(T x, T *lds, T &total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T incl = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) lds[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    T s = 0;
    for (int i = 0; i < NT / 64; ++i) {
      T t = lds[i];
      lds[i] = s;
      s += t;
    }
    lds[NT / 64] = s;
  }
  __syncthreads();
  const T r = incl - x + lds[w];
  total = lds[NT / 64];
  __syncthreads();
  return r;
}

// Device memory released when the object goes out of scope.
struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T> T *as() const { return (T *)p; }
};

int dev_alloc(DevBuf &b, size_t bytes);

// out[i] = in[0] + ... + in[i-1] (int64); *d_total points at the device
// total (valid until wsum is released).  Asynchronous on st.
int scan_ints(const int *in, int n, long long *out, DevBuf &wsum, hipStream_t st, long long **d_total);
// Workspace (int64 values) the scan of n items needs: one sum per workgroup
// of kScanItems items, plus the total.
inline int scan_wsum_count(int n) { return std::max(1, (n + kScanItems - 1) / kScanItems) + 1; }
// The same scan with a caller-provided workspace of scan_wsum_count(n)
// int64 values; *d_total = wsum + scan_wsum_count(n) - 1.
int scan_ints_ws(const int *in, int n, long long *out, long long *wsum, hipStream_t st, long long **d_total);

}  // namespace dm
