// scan.hpp -- exclusive prefix sum on the device (block scan of 2048 items per
// workgroup, recursive scan of the block totals, offset add).
#pragma once
#include "device_common.hpp"

namespace mdgpu {

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 8;
constexpr int kScanItems = kScanThreads * kScanPerThread;

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kScanThreads) void k_scan_block(const Tin *__restrict__ in, Tout *__restrict__ out,
                                                             long long n, long long *__restrict__ block_sums) {
  __shared__ long long sh[kScanThreads];
  const long long base = (long long)blockIdx.x * kScanItems + (long long)threadIdx.x * kScanPerThread;
  long long v[kScanPerThread];
  long long tsum = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; k++) {
    const long long idx = base + k;
    v[k] = idx < n ? (long long)in[idx] : 0;
    tsum += v[k];
  }
  sh[threadIdx.x] = tsum;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const long long t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  long long excl = sh[threadIdx.x] - tsum;
#pragma unroll
  for (int k = 0; k < kScanPerThread; k++) {
    const long long idx = base + k;
    if (idx < n) out[idx] = (Tout)excl;
    excl += v[k];
  }
  if (threadIdx.x == kScanThreads - 1) block_sums[blockIdx.x] = sh[kScanThreads - 1];
}

template <typename Tout>
__global__ void k_scan_add(Tout *__restrict__ out, long long n, const long long *__restrict__ offs) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) out[idx] += (Tout)offs[idx / kScanItems];
}

template <typename Tin, typename Tout>
__global__ void k_scan_total(const Tin *__restrict__ in, const Tout *__restrict__ out, long long n,
                             long long *__restrict__ total) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *total = n > 0 ? (long long)out[n - 1] + (long long)in[n - 1] : 0;
}

// Number of long long scratch elements exclusive_scan needs for n items.
inline long long scan_scratch_size(long long n) {
  long long s = 0;
  while (n > 1) {
    n = (n + kScanItems - 1) / kScanItems;
    s += 2 * n + 2;
  }
  return s + 8;
}

// out[k] = sum of in[0..k), device-wide; *total (device) = sum of all.
template <typename Tin, typename Tout>
void exclusive_scan(const Tin *in, Tout *out, long long n, long long *scratch, long long *total,
                    hipStream_t st) {
  if (n > 0) {
    const long long nb = (n + kScanItems - 1) / kScanItems;
    long long *bs = scratch;             // nb block totals
    long long *offs = scratch + nb + 1;  // their exclusive scan
    long long *deeper = offs + nb + 1;
    hipLaunchKernelGGL((k_scan_block<Tin, Tout>), dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, out, n,
                       bs);
    if (nb > 1) {
      exclusive_scan<long long, long long>(bs, offs, nb, deeper, nullptr, st);
      const long long blocks = (n + 255) / 256;
      hipLaunchKernelGGL((k_scan_add<Tout>), dim3((unsigned)blocks), dim3(256), 0, st, out, n, offs);
    }
  }
  if (total)
    hipLaunchKernelGGL((k_scan_total<Tin, Tout>), dim3(1), dim3(1), 0, st, in, out, n, total);
}

}  // namespace mdgpu
