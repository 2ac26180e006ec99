// The density stage's gaussian with 2*sigma*sigma = 2^s evaluates
// exp((-d*d) * 2^-s) as exp(y * 2^-s) with the power of two folded into the exp's
// constants (clue_rt::exp_scaled, clue_dev::expf_cw_scaled). This program checks,
// on the CPU, that the folded evaluation gives the oracle's bits,
// oracle_expf(rn(y * 2^-s)), for every y <= 0 of the core range with s = 1 (the
// sigma = 1 gaussian) and on a dense stride of all floats for every |s| <= 8.
// The GPU parity tests then pin the device's own bits to the oracle.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "clue_host.h"
#include "clue_oracle.h"

namespace {

float bits_to_float(uint32_t b) {
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}
uint32_t float_to_bits(float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return b;
}

// The folded evaluation, step for step as clue_device.h writes it.
float core_scaled(float y, const clue_dev::ExpScaled& e) {
  const float k = std::rint(y * e.log2e);
  float r = std::fma(-k, e.ln2_hi, y);
  r = std::fma(-k, e.ln2_lo, r);
  float p = e.c[0];
  for (int i = 1; i < 6; ++i) p = std::fma(p, r, e.c[i]);
  p = std::fma(p, r, 1.0f);
  return std::ldexp(p, static_cast<int>(k));
}
float cw_scaled(float y, const clue_dev::ExpScaled& e) {
  float v = core_scaled(std::fmin(std::fmax(y, e.lo), e.hi), e);
  v = (y > e.hi) ? INFINITY : v;
  v = (y < e.lo) ? 0.0f : v;
  return (y != y) ? y : v;
}

// Returns the number of mismatches over y = bits b0, b0 + step, ... < b1.
long long check(int s, uint64_t b0, uint64_t b1, uint64_t step, bool core_too) {
  const clue_dev::ExpScaled e = clue_rt::exp_scaled(s);
  const float scale = std::ldexp(1.0f, -s);
  long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>((b1 - b0 + step - 1) / step); ++i) {
    const float y = bits_to_float(static_cast<uint32_t>(b0 + static_cast<uint64_t>(i) * step));
    const float x = y * scale;
    const float want = oracle_expf(x);
    const float got = cw_scaled(y, e);
    const bool same = (want != want) ? (got != got) : (float_to_bits(want) == float_to_bits(got));
    if (!same) {
      if (bad < 5)
        std::printf("s=%d y=%a: folded %a, oracle %a\n", s, static_cast<double>(y),
                    static_cast<double>(got), static_cast<double>(want));
      ++bad;
    }
    if (core_too && x >= clue_dev::kExpCoreLo && x <= clue_dev::kExpCoreHi &&
        float_to_bits(core_scaled(y, e)) != float_to_bits(want))
      ++bad;
  }
  return bad;
}

}  // namespace

int main(int argc, char** argv) {
  const bool exhaustive = argc > 1 && std::strcmp(argv[1], "--exhaustive") == 0;
  long long bad = 0;
  // s = 1: every y in [-2 * 103.97, -0] (sign bit set, up to the core's bound).
  const uint64_t neg0 = 0x80000000ull;
  const uint64_t neg_end = float_to_bits(2.0f * clue_dev::kExpCoreLo) + 1ull;
  bad += check(1, neg0, neg_end, exhaustive ? 1 : 7, true);
  // Every |s| <= 8 on a stride of all 2^32 bit patterns (both signs, NaN, inf).
  for (int s = -8; s <= 8; ++s) bad += check(s, 0, 1ull << 32, exhaustive ? 97 : 4099, true);
  // Values around the scaled core bounds.
  for (int s = -8; s <= 8; ++s) {
    const clue_dev::ExpScaled e = clue_rt::exp_scaled(s);
    bad += check(s, float_to_bits(e.hi) - 4096, float_to_bits(e.hi) + 4096, 1, true);
    bad += check(s, float_to_bits(e.lo) - 4096, float_to_bits(e.lo) + 4096, 1, true);
  }
  std::printf("folded exp against the oracle: %lld failed\n", bad);
  return bad ? 1 : 0;
}
