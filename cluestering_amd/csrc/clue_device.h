// Device-side building blocks shared by the CLUE stages: wave-level primitives,
// the exp restatement, tile geometry and distance metrics. gfx950 only.
//
// FP contract: all device code is compiled with -ffp-contract=off, and every
// expression below is evaluated in the order the reference writes it, so that
// results are bit-identical to the serial CPU semantics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clue_types.h"

namespace clue_dev {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Wave primitives (wave64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive scan within a wave.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int l = lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (l >= off) v += t;
  }
  return v;
}

// Exclusive scan across the Block threads of a workgroup; the total in *total.
template <int Block>
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  constexpr int kWaves = Block / kWave;
  __shared__ int wsum[kWaves];
  const int w = threadIdx.x / kWave;
  const int incl = wave_incl_scan(v);
  if (lane_id() == kWave - 1) wsum[w] = incl;
  __syncthreads();
  int before = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kWaves; ++k) {
    const int s = wsum[k];
    before += (k < w) ? s : 0;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return before + incl - v;
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ long long wave_min_i(long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const long long o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ long long wave_max_i(long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Wave-wide reduction of a 32-bit value through DPP lane moves instead of LDS
// permutes (whose results would share a wait counter with the scalar loads):
// within each row of 16 lanes a pair swap, a quad swap, the half-row mirror and
// the row mirror give every lane its row's result; the broadcasts of lane 15 to
// rows 1 and 3 and of lane 31 to rows 2 and 3 then fold rows 0-2 into row 3, and
// lane 63's value is returned in a scalar register. Lanes a move does not write
// take the operation's identity id, so each move can fuse into its operation.
// op must be commutative, associative and idempotent (min, max), and every lane
// of the wave active.
template <class F>
__device__ __forceinline__ int wave_reduce_dpp(int v, int id, F op) {
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xF, 0xF, false));  // row_mirror
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min_uniform(int v) {
  return wave_reduce_dpp(v, INT32_MAX, [](int a, int b) { return a < b ? a : b; });
}
__device__ __forceinline__ int wave_max_uniform(int v) {
  return wave_reduce_dpp(v, INT32_MIN, [](int a, int b) { return a > b ? a : b; });
}

// Peers: the set of active lanes whose key equals this lane's key, built from one
// ballot per key bit. Lanes with valid == false are excluded.
__device__ __forceinline__ uint64_t match_peers(uint32_t key, int bits, bool valid) {
  uint64_t peers = ballot(valid);
  for (int b = 0; b < bits; ++b) {
    const bool bit = (key >> b) & 1u;
    const uint64_t m = ballot(bit);
    peers &= bit ? m : ~m;
  }
  return valid ? peers : 0ull;
}

// ---------------------------------------------------------------------------
// exp: Cody-Waite reduction x = k ln2 + r, the degree-6 minimax polynomial of
// e^r on [-ln2/2, ln2/2] (constant and linear terms 1) in Horner form with fused
// multiply-adds, scaling by 2^k. Same method, constants and operation order as
// the CPU statement of the algorithm, so both give the same bits.
// ---------------------------------------------------------------------------
// The core for x in [LO, HI] = [-103.97, 88.72]: no range handling.
__device__ __forceinline__ float expf_cw_core(float x) {
  const float k = __builtin_rintf(x * kExpLog2e);
  float r = __builtin_fmaf(-k, kExpLn2Hi, x);
  r = __builtin_fmaf(-k, kExpLn2Lo, r);
  float p = kExpPoly[0];
  p = __builtin_fmaf(p, r, kExpPoly[1]);
  p = __builtin_fmaf(p, r, kExpPoly[2]);
  p = __builtin_fmaf(p, r, kExpPoly[3]);
  p = __builtin_fmaf(p, r, kExpPoly[4]);
  p = __builtin_fmaf(p, r, 1.0f);
  p = __builtin_fmaf(p, r, 1.0f);
  return __builtin_ldexpf(p, static_cast<int>(k));
}

// Branch-free for every x: the core runs on x clamped to [LO, HI] and the
// out-of-range results are selected afterwards.
__device__ __forceinline__ float expf_cw(float x) {
  const float xin = x;
  float e = expf_cw_core(__builtin_fminf(__builtin_fmaxf(x, kExpCoreLo), kExpCoreHi));
  e = (xin > kExpCoreHi) ? __builtin_inff() : e;
  e = (xin < kExpCoreLo) ? 0.0f : e;
  return (xin != xin) ? xin : e;
}

// expf_cw_core(y * 2^-s) and expf_cw(y * 2^-s) with the scale folded into the
// constants (ExpScaled): the same bits without the multiplication. A NaN y gives
// y itself, the NaN that y * 2^-s would carry. c0 is e.c[0], which a caller in a
// loop may hold in a vector register (the first step's other two operands are
// scalar).
__device__ __forceinline__ float expf_cw_core_scaled(float y, const ExpScaled& e, float c0) {
  const float k = __builtin_rintf(y * e.log2e);
  float r = __builtin_fmaf(-k, e.ln2_hi, y);
  r = __builtin_fmaf(-k, e.ln2_lo, r);
  float p = c0;
  p = __builtin_fmaf(p, r, e.c[1]);
  p = __builtin_fmaf(p, r, e.c[2]);
  p = __builtin_fmaf(p, r, e.c[3]);
  p = __builtin_fmaf(p, r, e.c[4]);
  p = __builtin_fmaf(p, r, e.c[5]);
  p = __builtin_fmaf(p, r, 1.0f);
  return __builtin_ldexpf(p, static_cast<int>(k));
}
__device__ __forceinline__ float expf_cw_scaled(float y, const ExpScaled& es, float c0) {
  float e = expf_cw_core_scaled(__builtin_fminf(__builtin_fmaxf(y, es.lo), es.hi), es, c0);
  e = (y > es.hi) ? __builtin_inff() : e;
  e = (y < es.lo) ? 0.0f : e;
  return (y != y) ? y : e;
}

// Double precision: the same method with a 29-bit ln2_hi (k*ln2_hi exact for
// |k| <= 1075) and the degree-13 Taylor polynomial; the same constants and
// operation order as the CPU statement. The core is for x in [-745.13, 709.78].
__device__ __forceinline__ double expd_cw_core(double x) {
  const double log2e = 0x1.71547652b82fep+0;
  const double ln2_hi = 0x1.62e42ff000000p-1;
  const double ln2_lo = -0x1.718432a1b0e26p-35;
  const double k = __builtin_rint(x * log2e);
  double r = __builtin_fma(-k, ln2_hi, x);
  r = __builtin_fma(-k, ln2_lo, r);
  double p = 0x1.6124613a86d09p-33;
  p = __builtin_fma(p, r, 0x1.1eed8eff8d898p-29);
  p = __builtin_fma(p, r, 0x1.ae64567f544e4p-26);
  p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);
  p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);
  p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);
  p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);
  p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);
  p = __builtin_fma(p, r, 0x1.1111111111111p-7);
  p = __builtin_fma(p, r, 0x1.5555555555555p-5);
  p = __builtin_fma(p, r, 0x1.5555555555555p-3);
  p = __builtin_fma(p, r, 0x1.0000000000000p-1);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  return __builtin_ldexp(p, static_cast<int>(k));
}

__device__ __forceinline__ double expd_cw(double x) {
  const double lo = -0x1.74910d52d3052p+9, hi = 0x1.62e42fefa39efp+9;
  const double xin = x;
  double e = expd_cw_core(__builtin_fmin(__builtin_fmax(x, lo), hi));
  e = (xin > hi) ? __builtin_inf() : e;
  e = (xin < lo) ? 0.0 : e;
  return (xin != xin) ? xin : e;
}

// Correctly rounded square root. For s >= 2^-96 the hardware estimate r (within
// 1 ulp) is settled by Tuckerman's rounding test on its neighbours r - ulp and
// r + ulp, whose fma residuals s - r' * r carry the sign of the exact error
// there (no denormal intermediates). Zero, tiny, negative, infinite and NaN
// inputs take the IEEE library sqrt.
// The test alone, for finite s >= 2^-96 or s == 0. The answer is r - ulp when
// s - (r - ulp) r <= 0, r + ulp when s - (r + ulp) r > 0, else r; the second
// residual is below the first, so the answer is r - ulp plus one for each
// residual above zero. A residual is above zero exactly when (r' r - s), the fma
// with the sign moved (an exact zero is +0 in round-to-nearest), has its sign bit
// set: two shifts and one three-way integer add replace two compares and two
// selects. r - ulp saturates at +0, so s == 0 gives 0 (both residuals are +0).
__device__ __forceinline__ float sqrt_rn_normal(float s) {
  const float r = __builtin_amdgcn_sqrtf(s);
  const uint32_t rb = __float_as_uint(r);
  const uint32_t mb = __builtin_elementwise_sub_sat(rb, 1u);
  const float rm = __uint_as_float(mb);
  const float rp = __uint_as_float(rb + 1u);
  const uint32_t above_m = __float_as_uint(__builtin_fmaf(rm, r, -s)) >> 31;
  const uint32_t above_p = __float_as_uint(__builtin_fmaf(rp, r, -s)) >> 31;
  return __uint_as_float(mb + above_m + above_p);
}
__device__ __forceinline__ float sqrt_rn(float s) {
  float q = sqrt_rn_normal(s);
  if (__builtin_expect(!(s >= 0x1p-96f && s <= 0x1.fffffep+127f), 0)) q = __builtin_sqrtf(s);
  return q;
}

// Double: the compiler's IEEE sequence (the hardware estimate is not correctly
// rounded in double).
__device__ __forceinline__ double sqrt_rn(double s) { return __builtin_sqrt(s); }

// Ordering key of a double density, for the nearest-higher test in double: the
// high bits order rho (-0 folded onto +0); NaN is the minimum as a candidate and
// the maximum as a query. The tag comparison is done separately (it does not fit).
__device__ __forceinline__ uint64_t rho_key_f64(double r, bool query) {
  if (r != r) return query ? ~0ull : 0ull;
  if (r == 0.0) r = 0.0;
  const uint64_t b = static_cast<uint64_t>(__double_as_longlong(r));
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}

// Ordering key of (rho, tag) such that, for a query point i and a candidate j,
// key(j, candidate) > key(i, query) exactly when j is "higher" than i in the
// reference's sense (ClusteringKernels.hpp:180-182):
//   rho_j > rho_i || (rho_j == rho_i && rho_j > 0 && tag_j > tag_i).
// The high word orders rho (-0 folded onto +0, as the float comparison does);
// the low word holds the tag only for rho > 0. NaN never compares higher: as a
// candidate its key is the minimum, as a query the maximum.
__device__ __forceinline__ uint64_t rho_tag_key(float r, uint32_t tag, bool query) {
  if (r != r) return query ? ~0ull : 0ull;
  if (r == 0.0f) r = 0.0f;
  const uint32_t b = __float_as_uint(r);
  const uint32_t o = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return (static_cast<uint64_t>(o) << 32) | (r > 0.0f ? tag : 0u);
}

// rho of a float ordering key (rho_tag_key's high word inverted; -0 comes back as
// +0, which compares equal).
__device__ __forceinline__ float key_rho(uint64_t key) {
  const uint32_t o = static_cast<uint32_t>(key >> 32);
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// static_cast<int>(float or double) with the x86 semantics the reference's
// serial backend runs with (cvttss2si / cvttsd2si): NaN and out-of-range values
// give INT_MIN.
__device__ __forceinline__ int cvt_trunc(float q) {
  if (!(q >= -2147483648.0f) || !(q < 2147483648.0f)) return INT32_MIN;
  return static_cast<int>(q);
}
__device__ __forceinline__ int cvt_trunc(double q) {
  if (!(q >= -2147483648.0) || !(q < 2147483648.0)) return INT32_MIN;
  return static_cast<int>(q);
}

// Tile geometry of one event (TilesView.hpp:38-104).
template <class T>
struct GeoT {
  int D;
  int nper;
  const T* mn;
  const T* mx;
  const T* side;
  const int32_t* wrapped;
};
using Geo = GeoT<float>;

template <class T>
__device__ __forceinline__ T normalize(const GeoT<T>& g, T c, int d) {
  const T range = g.mx[d] - g.mn[d];
  T r = c - static_cast<T>(cvt_trunc(c / range)) * range;
  if (r >= g.mx[d])
    r -= range;
  else if (r < g.mn[d])
    r += range;
  return r;
}

template <class T>
__device__ __forceinline__ int get_bin(const GeoT<T>& g, T c, int d) {
  int b;
  if (g.wrapped[d])
    b = cvt_trunc((normalize(g, c, d) - g.mn[d]) / g.side[d]);
  else
    b = cvt_trunc((c - g.mn[d]) / g.side[d]);
  b = min(b, g.nper - 1);
  b = max(b, 0);
  return b;
}

// Ordered-integer encodings, so that integer min/max equal floating min/max.
__device__ __forceinline__ int32_t float_to_ordered(float f) {
  const int32_t b = __float_as_int(f);
  return b >= 0 ? b : (b ^ 0x7fffffff);
}
__device__ __forceinline__ float ordered_to_float(int32_t o) {
  return __int_as_float(o >= 0 ? o : (o ^ 0x7fffffff));
}
__device__ __forceinline__ long long float_to_ordered(double f) {
  const long long b = __double_as_longlong(f);
  return b >= 0 ? b : (b ^ 0x7fffffffffffffffll);
}
__device__ __forceinline__ double ordered_to_float(long long o) {
  return __longlong_as_double(o >= 0 ? o : (o ^ 0x7fffffffffffffffll));
}

}  // namespace clue_dev
