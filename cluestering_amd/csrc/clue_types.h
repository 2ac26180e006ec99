// Plain structs shared by the host orchestration and the device stages.
#pragma once

#include <stdint.h>

namespace clue_dev {

constexpr int kMaxDim = 10;
// Candidates per block of the nearest-higher block layout (one bounding box and
// one largest key per block).
constexpr int kNhBlock = 16;

// Candidates per bounding box of the tile-sorted columns (the density stage
// skips a box that no lane can reach).
constexpr int kSortedBox = 16;

// Largest tile count of an event that the fused per-event tile build takes (its
// per-wave tile counters live in LDS); batches with a larger event take the
// radix-sort path.
constexpr int kEvtMaxTiles = 1024;
// Largest event the fused per-event stages take (one bit per point in LDS).
constexpr int kEvtMaxPoints = 1 << 17;

// Workspace sizes of the scan and radix-sort primitives (clue_primitives.hip),
// here so that host-only code sizes workspaces exactly as the device path does.
constexpr int kScanTileElems = 4096;  // elements per block of the exclusive scan
constexpr int kRsTileKeys = 4096;     // keys per block of a radix pass
constexpr int kRsDigitBins = 256;     // bins of one radix digit
inline int64_t scan_temp_elems(int64_t n) { return (n + kScanTileElems - 1) / kScanTileElems + 1; }
inline int64_t radix_hist_elems(int64_t n) {
  return kRsDigitBins * ((n + kRsTileKeys - 1) / kRsTileKeys);
}

// Constants of the float exp (Cody-Waite reduction x = k ln2 + r with a 15-bit
// ln2_hi, then the degree-6 minimax polynomial of e^r on [-ln2/2, ln2/2] whose
// constant and linear terms are 1), shared by the device code and the host code
// that folds power-of-two argument scales into them.
constexpr float kExpLog2e = 0x1.715476p+0f;
constexpr float kExpLn2Hi = 0x1.62e400p-1f;
constexpr float kExpLn2Lo = 0x1.7f7d1cp-20f;
constexpr float kExpPoly[5] = {0x1.6a244cp-10f, 0x1.1239d4p-7f, 0x1.5558f2p-5f, 0x1.555492p-3f,
                               0x1.fffffcp-2f};  // c6, c5, c4, c3, c2
constexpr float kExpCoreLo = -0x1.9fe368p+6f;  // the core needs no range handling on [lo, hi]
constexpr float kExpCoreHi = 0x1.62e42ep+6f;

// exp(y * 2^-s) evaluated without forming y * 2^-s: every step of the reduction
// and of the Horner polynomial is the float exp's own step scaled by a power of
// two (k = rint(y * (log2e 2^-s)); r' = 2^s r from ln2 2^s; the Horner partial
// sums carry 2^(-s i) through coefficients c_i 2^(-s i)), so with |s| <= 8 no
// step leaves the normal range and the bits equal exp(rn(y * 2^-s)) for every y.
struct ExpScaled {
  float log2e;   // log2e * 2^-s
  float ln2_hi;  // ln2_hi * 2^s
  float ln2_lo;  // ln2_lo * 2^s
  float c[6];    // c6 .. c1, c_i * 2^(-s i)
  float lo, hi;  // the core range, times 2^s
};

// One event of a batch. All offsets are host-computed from the event sizes.
struct EventDesc {
  int32_t p_off;   // first point of the event in the packed arrays
  int32_t n;       // points in the event
  int32_t nper;    // tiles per dimension (SetupTiles.hpp:30-34)
  int32_t ntiles;  // nper^D
  int32_t t_base;  // first global tile key of the event
  int32_t index;   // the caller's event index (empty events are not described)
  int32_t pad[2];
};

// A contiguous range of points of one event, the unit of the per-point grids.
struct Chunk {
  int32_t event;
  int32_t begin;  // global point index
  int32_t end;
  int32_t pad;
};

// A work item of the neighbour stages: up to 64 consecutive tile-sorted points of
// one tile (one lane each).
struct Item {
  int32_t tile;   // global tile key
  int32_t begin;  // sorted position of the first point
  int32_t end;
  int32_t event;
};

// Per-run scalar parameters as the device stages consume them; T is the point
// type (float, or double for Clusterer<Ndim, double>).
template <class T>
struct RunParamsT {
  int32_t D;
  int32_t kernel;
  int32_t metric;
  int32_t has_sigma;
  int32_t has_u;
  int32_t has_tags;
  int32_t n_events;
  int32_t n_total;
  T dc;
  T rhoc;
  T dm;
  T seed_dc;
  T dc_sq_max;           // largest t with sqrt_rn(t) <= dc (sqrt metrics)
  T dm_sq_max;           // same for dm
  T seed_sq_max;         // same for seed_dc
  T kp0, kp1;            // kernel parameters (flat f | exp lambda, A | gauss sigma, A)
  T gauss_den;           // (2 * sigma) * sigma, as the reference evaluates it
  T gauss_inv_den;       // 1 / gauss_den when gauss_den is a power of two
  int32_t gauss_pow2;    // gauss_den is a power of two
  int32_t gauss_fold;    // float runs: gauss_den = 2^s with |s| <= 8, folded into es
  int32_t exp_in_range;  // every kernel exp argument within the radius is in the core range
  ExpScaled es;          // float runs: exp(y * 2^-s) with 2^s = gauss_den
  int32_t wrapped[kMaxDim];
  T mparam[kMaxDim];
  int32_t xcd_chunk_log2;  // neighbour stages: 2^k consecutive blocks per XCD turn; < 0: none
  int32_t block_waves;     // neighbour stages: waves per workgroup (1, 2, 4, 8 or 16)
  int32_t nh_blocks;       // nearest-higher over the Morton-ordered block layout
  int32_t persistent;      // neighbour stages: persistent waves per CU pulling items (0: none)
};
using RunParams = RunParamsT<float>;

}  // namespace clue_dev
