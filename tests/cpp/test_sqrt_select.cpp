// The float square root of the neighbour stages (clue_dev::sqrt_rn_normal) takes
// the hardware estimate r, within one ulp of sqrt(s), and picks among r - ulp, r
// and r + ulp by the signs of two fma residuals, added to r - ulp as integers:
//   q = bits(r - ulp) + [s - (r - ulp) r > 0] + [s - (r + ulp) r > 0],
// with r - ulp saturated at +0 (so that s = 0 gives 0). This program checks, on
// the CPU, that the rule gives the correctly rounded square root for every
// estimate a one-ulp hardware square root can return (the three floats nearest
// the correctly rounded value): for every float s in [1, 4), which holds both
// exponent parities of every mantissa, for s = 0, and on a stride over all finite
// floats from 2^-96 up. The GPU parity tests then check the device itself.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

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

// The selection, step for step as clue_device.h writes it.
float select(float s, float r) {
  const uint32_t rb = float_to_bits(r);
  const uint32_t mb = rb == 0u ? 0u : rb - 1u;  // saturating subtraction
  const float rm = bits_to_float(mb);
  const float rp = bits_to_float(rb + 1u);
  const uint32_t a = float_to_bits(std::fma(rm, r, -s)) >> 31;
  const uint32_t b = float_to_bits(std::fma(rp, r, -s)) >> 31;
  return bits_to_float(mb + a + b);
}

// Correctly rounded: the double square root of a float rounds to the float one
// correctly (53 >= 2 * 24 + 2 bits, so it never lands on a float midpoint).
float sqrt_rn_ref(float s) { return static_cast<float>(std::sqrt(static_cast<double>(s))); }

long check(float s, long& bad_shown) {
  const float q = sqrt_rn_ref(s);
  const uint32_t qb = float_to_bits(q);
  long bad = 0;
  for (int d = -1; d <= 1; ++d) {
    if (qb == 0u && d < 0) continue;
    const float r = bits_to_float(qb + static_cast<uint32_t>(d));
    // Only estimates within one ulp of the exact root are possible.
    const double err = std::fabs(static_cast<double>(r) - std::sqrt(static_cast<double>(s)));
    const double ulp = std::ldexp(1.0, std::ilogb(r == 0.0f ? q : r) - 23);
    if (q != 0.0f && err >= ulp) continue;
    const float got = select(s, r);
    if (float_to_bits(got) != qb) {
      if (bad_shown++ < 10)
        std::printf("s=%a r=%a: got %a, want %a\n", static_cast<double>(s), static_cast<double>(r),
                    static_cast<double>(got), static_cast<double>(q));
      ++bad;
    }
  }
  return bad;
}

}  // namespace

int main() {
  long bad = 0, shown = 0, n = 0;
  bad += check(0.0f, shown);
  ++n;
  for (uint32_t b = float_to_bits(1.0f); b < float_to_bits(4.0f); ++b, ++n)
    bad += check(bits_to_float(b), shown);
  for (uint32_t b = float_to_bits(0x1p-96f); b < 0x7f800000u; b += 997u, ++n)
    bad += check(bits_to_float(b), shown);
  // Both ends of the range exactly.
  bad += check(0x1p-96f, shown);
  bad += check(bits_to_float(0x7f7fffffu), shown);
  n += 2;
  std::printf("%ld inputs, %ld failed\n", n, bad);
  return bad == 0 ? 0 : 1;
}
