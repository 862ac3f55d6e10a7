// Device-wide exclusive scan (three passes: per-workgroup sums, a scan of
// the sums in one workgroup, per-workgroup rescans with the offsets).
#include "dm_scan.h"

namespace dm {

// per-workgroup sums of kScanItems consecutive items
__global__ __launch_bounds__(256) void k_scan_up(const int *__restrict__ in, int n, long long *__restrict__ wsum) {
  __shared__ long long lds[256 / 64 + 1];
  const int base = blockIdx.x * kScanItems + threadIdx.x * 4;
  long long s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < n) s += in[base + k];
  long long tot;
  wg_exclusive_scan<256>(s, lds, tot);
  if (threadIdx.x == 0) wsum[blockIdx.x] = tot;
}

// exclusive scan of the workgroup sums in place; total in wsum[nw]
__global__ __launch_bounds__(1024) void k_scan_top(long long *__restrict__ wsum, int nw) {
  __shared__ long long lds[1024 / 64 + 1];
  long long carry = 0;
  for (int base = 0; base < nw; base += 1024) {
    const int i = base + threadIdx.x;
    const long long x = i < nw ? wsum[i] : 0;
    long long tot;
    const long long e = wg_exclusive_scan<1024>(x, lds, tot);
    if (i < nw) wsum[i] = carry + e;
    carry += tot;
  }
  if (threadIdx.x == 0) wsum[nw] = carry;
}

__global__ __launch_bounds__(256) void k_scan_down(const int *__restrict__ in, int n, const long long *__restrict__ wsum,
                                                  long long *__restrict__ out) {
  __shared__ long long lds[256 / 64 + 1];
  const int base = blockIdx.x * kScanItems + threadIdx.x * 4;
  int x[4];
  long long s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[k] = base + k < n ? in[base + k] : 0;
    s += x[k];
  }
  long long tot;
  long long e = wg_exclusive_scan<256>(s, lds, tot) + wsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < n) {
      out[base + k] = e;
      e += x[k];
    }
}

int dev_alloc(DevBuf &b, size_t bytes) {
  DM_HIP(hipMalloc(&b.p, std::max<size_t>(bytes, 16)));
  return DM_OK;
}

int scan_ints_ws(const int *in, int n, long long *out, long long *wsum, hipStream_t st, long long **d_total) {
  const int nw = scan_wsum_count(n) - 1;
  hipLaunchKernelGGL(k_scan_up, dim3(nw), dim3(256), 0, st, in, n, wsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, wsum, nw);
  hipLaunchKernelGGL(k_scan_down, dim3(nw), dim3(256), 0, st, in, n, wsum, out);
  DM_HIP(hipGetLastError());
  *d_total = wsum + nw;
  return DM_OK;
}

int scan_ints(const int *in, int n, long long *out, DevBuf &wsum, hipStream_t st, long long **d_total) {
  if (int rc = dev_alloc(wsum, (size_t)scan_wsum_count(n) * 8)) return rc;
  return scan_ints_ws(in, n, out, wsum.as<long long>(), st, d_total);
}

}  // namespace dm
