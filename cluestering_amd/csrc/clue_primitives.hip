// Device-wide primitives written for CDNA4 wave64: exclusive scan and a stable
// LSD radix sort of (key, value) pairs. No rocPRIM/rocThrust.
#include <hip/hip_runtime.h>

#include "clue_device.h"
#include "clue_internal.h"

namespace clue_dev {

// ---------------------------------------------------------------------------
// Block reductions / scans (256 threads = 4 waves).
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  return block_excl_scan<kBlock>(v, total);
}

// ---------------------------------------------------------------------------
// Exclusive scan of int32 (reduce -> scan partials -> downsweep).
// ---------------------------------------------------------------------------
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;
static_assert(kScanTile == kScanTileElems, "scan workspace sizing");

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                        int32_t* __restrict__ partial) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  __shared__ int wsum[kWavesPerBlock];
  s = wave_sum_i(s);
  if (lane_id() == 0) wsum[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < kWavesPerBlock; ++k) t += wsum[k];
    partial[blockIdx.x] = t;
  }
}

// One block scans the partials in place (exclusive), in chunks with a carry.
__global__ __launch_bounds__(kBlock) void k_scan_partials(int32_t* __restrict__ partial, int64_t nb,
                                                          int32_t* __restrict__ total_out) {
  int carry = 0;
  for (int64_t base = 0; base < nb; base += kBlock) {
    const int64_t i = base + threadIdx.x;
    const int v = i < nb ? partial[i] : 0;
    int tot;
    const int ex = block_excl_scan(v, &tot);
    if (i < nb) partial[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const int32_t* in, int32_t* out, int64_t n,
                                                      const int32_t* __restrict__ partial) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  int tot;
  int run = block_excl_scan(s, &tot) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

// One block for n <= kScanTile: the downsweep with no partials, and the total
// (small scans, such as one event's tiles or a batch's per-event seed counts,
// then take one launch instead of three).
__global__ __launch_bounds__(kBlock) void k_scan_single(const int32_t* in, int32_t* out, int64_t n,
                                                        int32_t* __restrict__ total_out) {
  const int64_t base = static_cast<int64_t>(threadIdx.x) * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  int tot;
  int run = block_excl_scan(s, &tot);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (threadIdx.x == 0 && total_out) *total_out = tot;
}

void scan_exclusive(const int32_t* in, int32_t* out, int64_t n, int32_t* temp, int32_t* total,
                    hipStream_t s) {
  if (n <= 0) return;
  if (n <= kScanTile) {
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, s, in, out, n, total);
    return;
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_reduce, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s, in, n, temp);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kBlock), 0, s, temp, nb, total);
  hipLaunchKernelGGL(k_scan_down, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s, in, out, n, temp);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (uint32 key, int32 value), 8-bit digits.
//
// Each block owns a contiguous range of kRsTile keys; each of its four waves owns a
// contiguous quarter, walked in rounds of 64. Within a round, lanes with the same
// digit are found with one ballot per digit bit and ranked by lane order; per-wave
// digit counters live in LDS. Ranges, rounds and lanes are all visited in index
// order, so the pass is stable.
// ---------------------------------------------------------------------------
constexpr int kRsBits = 8;
constexpr int kRsBins = 1 << kRsBits;
constexpr int kRsRounds = 16;
constexpr int kRsWaveKeys = kWave * kRsRounds;
constexpr int kRsTile = kRsWaveKeys * kWavesPerBlock;  // 4096
static_assert(kRsTile == kRsTileKeys && kRsBins == kRsDigitBins, "radix workspace sizing");

__global__ __launch_bounds__(kBlock) void k_rs_hist(const uint32_t* __restrict__ keys, int64_t n,
                                                    int shift, int64_t nblocks,
                                                    int32_t* __restrict__ hist) {
  __shared__ int cnt[kRsBins];
  for (int k = threadIdx.x; k < kRsBins; k += kBlock) cnt[k] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kRsTile;
  for (int k = threadIdx.x; k < kRsTile; k += kBlock) {
    const int64_t i = base + k;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kRsBins - 1)], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRsBins; d += kBlock) hist[d * nblocks + blockIdx.x] = cnt[d];
}

__global__ __launch_bounds__(kBlock) void k_rs_scatter(const uint32_t* __restrict__ kin,
                                                       const int32_t* __restrict__ vin,
                                                       uint32_t* __restrict__ kout,
                                                       int32_t* __restrict__ vout, int64_t n,
                                                       int shift, int digit_bits, int64_t nblocks,
                                                       const int32_t* __restrict__ scanned) {
  __shared__ int cnt[kWavesPerBlock][kRsBins];
  const int w = threadIdx.x / kWave;
  const int l = lane_id();
  for (int k = threadIdx.x; k < kWavesPerBlock * kRsBins; k += kBlock) (&cnt[0][0])[k] = 0;
  __syncthreads();
  const int64_t wbase = static_cast<int64_t>(blockIdx.x) * kRsTile + w * kRsWaveKeys;
  const uint64_t lt = lanemask_lt();
  // Phase A: per-wave digit counts.
  for (int r = 0; r < kRsRounds; ++r) {
    const int64_t i = wbase + r * kWave + l;
    const bool valid = i < n;
    const uint32_t dgt = valid ? (kin[i] >> shift) & (kRsBins - 1) : 0u;
    const uint64_t peers = match_peers(dgt, digit_bits, valid);
    if (valid && (peers & lt) == 0) cnt[w][dgt] += __popcll(peers);
  }
  __syncthreads();
  // Per digit: global offset of this block, then the waves in order.
  for (int d = threadIdx.x; d < kRsBins; d += kBlock) {
    int run = scanned[d * nblocks + blockIdx.x];
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; ++k) {
      const int c = cnt[k][d];
      cnt[k][d] = run;
      run += c;
    }
  }
  __syncthreads();
  // Phase B: place keys.
  for (int r = 0; r < kRsRounds; ++r) {
    const int64_t i = wbase + r * kWave + l;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t dgt = (key >> shift) & (kRsBins - 1);
    const uint64_t peers = match_peers(dgt, digit_bits, valid);
    int pos = 0;
    if (valid) pos = cnt[w][dgt] + __popcll(peers & lt);
    __builtin_amdgcn_wave_barrier();
    if (valid && (peers & lt) == 0) cnt[w][dgt] += __popcll(peers);
    if (valid) {
      kout[pos] = key;
      vout[pos] = vin[i];
    }
  }
}

// Sorts (keys, vals) by the low `key_bits` bits of the keys. The result ends in
// (k0, v0) or (k1, v1); *in_first receives true when it is (k0, v0).
void radix_sort_pairs(uint32_t* k0, int32_t* v0, uint32_t* k1, int32_t* v1, int64_t n, int key_bits,
                      int32_t* hist, int32_t* scan_temp, bool* in_first, hipStream_t s) {
  *in_first = true;
  if (n <= 0) return;
  const int64_t nblocks = (n + kRsTile - 1) / kRsTile;
  uint32_t *ka = k0, *kb = k1;
  int32_t *va = v0, *vb = v1;
  for (int shift = 0; shift < key_bits; shift += kRsBits) {
    const int bits = key_bits - shift < kRsBits ? key_bits - shift : kRsBits;
    hipLaunchKernelGGL(k_rs_hist, dim3(static_cast<unsigned>(nblocks)), dim3(kBlock), 0, s, ka, n,
                       shift, nblocks, hist);
    scan_exclusive(hist, hist, kRsBins * nblocks, scan_temp, nullptr, s);
    hipLaunchKernelGGL(k_rs_scatter, dim3(static_cast<unsigned>(nblocks)), dim3(kBlock), 0, s, ka,
                       va, kb, vb, n, shift, bits, nblocks, hist);
    uint32_t* tk = ka;
    ka = kb;
    kb = tk;
    int32_t* tv = va;
    va = vb;
    vb = tv;
    *in_first = !*in_first;
  }
}

}  // namespace clue_dev
