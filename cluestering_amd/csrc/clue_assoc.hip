// Key -> members association map (CSR) on the device: the cluster -> points map
// of get_clusters (utils/detail/get_clusters.hpp:24-49) and the event -> clusters
// map of getSampleAssociations (core/detail/Clusterer.hpp:219-241), both built by
// make_associator (data_structures/internal/MakeAssociator.hpp:15-45):
//   bins     = max key + 1;
//   members  = the indices whose key is >= 0, grouped by key;
//   offsets  = bins + 1 boundaries.
// Members of one key are in ascending index order (a stable sort), which is the
// host map's fill order; the reference's device fill orders them by atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "clue_assoc.h"
#include "clue_device.h"

namespace clue_dev {
namespace {

constexpr int kBlock = 256;

__global__ void k_assoc_init(int32_t* __restrict__ stats) {
  if (threadIdx.x == 0) {
    stats[0] = INT32_MIN;  // max key
    stats[1] = 0;          // keys >= 0
  }
}

__global__ __launch_bounds__(kBlock) void k_assoc_stats(const int32_t* __restrict__ keys, int n,
                                                        int32_t* __restrict__ stats) {
  __shared__ int32_t smax[kBlock / kWave], scnt[kBlock / kWave];
  int mx = INT32_MIN, cnt = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const int k = keys[i];
    mx = max(mx, k);
    cnt += k >= 0 ? 1 : 0;
  }
  mx = wave_max_i(mx);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    smax[w] = mx;
    scnt[w] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = INT32_MIN, c = 0;
    for (int k = 0; k < kBlock / kWave; ++k) {
      m = max(m, smax[k]);
      c += scnt[k];
    }
    atomicMax(&stats[0], m);
    atomicAdd(&stats[1], c);
  }
}

// Sort key: the key itself, negative keys after every bin.
__global__ __launch_bounds__(kBlock) void k_assoc_keys(const int32_t* __restrict__ keys, int n,
                                                       int bins, uint32_t* __restrict__ k0,
                                                       int32_t* __restrict__ v0) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int k = keys[i];
  k0[i] = static_cast<uint32_t>(k >= 0 ? k : bins);
  v0[i] = i;
}

// offsets[b] = first sorted position whose key is >= b (lower bound), b in [0, bins].
__global__ __launch_bounds__(kBlock) void k_assoc_offsets(const uint32_t* __restrict__ sk, int n,
                                                          int bins, int32_t* __restrict__ offsets) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b > bins) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (static_cast<int>(sk[mid]) < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  offsets[b] = lo;
}

}  // namespace

void assoc_scratch_elems(int64_t n, int64_t* hist_elems, int64_t* scan_elems) {
  const int64_t hist = radix_hist_elems(n);
  *hist_elems = hist;
  *scan_elems = std::max(scan_temp_elems(n), scan_temp_elems(hist)) + 16;
}

hipError_t assoc_stats(const int32_t* keys, int n, int32_t* stats_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_assoc_init, dim3(1), dim3(kWave), 0, s, stats_dev);
  const int blocks = static_cast<int>(std::min<int64_t>((n + kBlock - 1) / kBlock, 2048));
  if (n > 0)
    hipLaunchKernelGGL(k_assoc_stats, dim3(blocks), dim3(kBlock), 0, s, keys, n, stats_dev);
  return hipGetLastError();
}

hipError_t assoc_build(const int32_t* keys, int n, int bins, const AssocScratch& w,
                       int32_t* offsets, int32_t* members, int n_members, hipStream_t s) {
  if (n > 0) {
    hipLaunchKernelGGL(k_assoc_keys, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, keys, n,
                       bins, w.k0, w.v0);
    int bits = 0;
    while ((int64_t(1) << bits) <= bins) ++bits;  // keys lie in [0, bins]
    bool in_first = true;
    radix_sort_pairs(w.k0, w.v0, w.k1, w.v1, n, bits, w.hist, w.scan, &in_first, s);
    const uint32_t* sk = in_first ? w.k0 : w.k1;
    const int32_t* sv = in_first ? w.v0 : w.v1;
    hipLaunchKernelGGL(k_assoc_offsets, dim3((bins + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       s, sk, n, bins, offsets);
    if (n_members > 0) {
      hipError_t e = hipMemcpyAsync(members, sv, sizeof(int32_t) * n_members,
                                    hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
    }
  } else {
    hipLaunchKernelGGL(k_assoc_offsets, dim3(1), dim3(kBlock), 0, s, w.k0, 0, 0, offsets);
  }
  return hipGetLastError();
}

}  // namespace clue_dev
