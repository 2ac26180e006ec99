// Host-side checks of the facade's metric call operators and point constructors:
// the known values of the reference's tests/test_distance_metrics.cpp:8-63 and the
// host SoA layouts of data_structures/PointsHost.hpp:59-137. Needs no GPU (no
// queue is created); the pytest wrapper tests/test_cpp_api.py runs it on the CPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cluestering_amd/CLUEstering.hpp"

static int g_fail = 0, g_pass = 0;
#define CHECK(cond)                                              \
  do {                                                           \
    if (cond) {                                                  \
      ++g_pass;                                                  \
    } else {                                                     \
      ++g_fail;                                                  \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
    }                                                            \
  } while (0)

int main() {
  // tests/test_distance_metrics.cpp:8-14
  {
    auto m = clue::metrics::Euclidean<2>();
    std::array<float, 3> a{1.f, 2.f, 0.f}, b{4.f, 6.f, 0.f};
    CHECK(m(a, b) == 5.f);
  }
  // :16-32 (array and variadic constructors): 1*3*3 + 2*4*4 = 41
  {
    auto w = std::array<float, 2>{1.f, 2.f};
    auto m1 = clue::metrics::WeightedEuclidean<2>(w);
    auto m2 = clue::metrics::WeightedEuclidean<2>(1.f, 2.f);
    std::array<float, 3> a{1.f, 2.f, 0.f}, b{4.f, 6.f, 0.f};
    CHECK(m1(a, b) == std::sqrt(41.f));
    CHECK(m2(a, b) == std::sqrt(41.f));
  }
  // :36-41
  {
    auto m = clue::metrics::Manhattan<2>();
    std::array<float, 3> a{-1.f, 3.f, 0.f}, b{0.f, 6.f, 0.f};
    CHECK(m(a, b) == 4.f);
  }
  // :43-48
  {
    auto m = clue::metrics::Chebyshev<2>();
    std::array<float, 3> a{-1.f, 4.f, 0.f}, b{0.f, 6.f, 0.f};
    CHECK(m(a, b) == 2.f);
  }
  // :50-63: max(1*1, 2*2) = 4
  {
    auto w = std::array<float, 2>{1.f, 2.f};
    auto m1 = clue::metrics::WeightedChebyshev<2>(w);
    auto m2 = clue::metrics::WeightedChebyshev<2>(1.f, 2.f);
    std::array<float, 3> a{-1.f, 4.f, 0.f}, b{0.f, 6.f, 0.f};
    CHECK(m1(a, b) == 4.f);
    CHECK(m2(a, b) == 4.f);
  }
  // Periodic (the reference's case at :34 is empty): periods (0, 10) on (1,1)-(4,9):
  // dimension 0 is not periodic (diff 3), dimension 1 wraps (min(8, 10-8) = 2).
  {
    auto m = clue::metrics::PeriodicEuclidean<2>(std::array<float, 2>{0.f, 10.f});
    std::array<float, 3> a{1.f, 1.f, 0.f}, b{4.f, 9.f, 0.f};
    CHECK(m(a, b) == std::sqrt(13.f));
    auto md = clue::metrics::PeriodicEuclidean<2, double>(0.0, 10.0);
    std::array<double, 3> c{1., 1., 0.}, d{4., 9., 0.};
    CHECK(md(c, d) == std::sqrt(13.));
  }
  // Mahalanobis on host points (tests/test_mahalanobis.cpp:10-43's points): (2-4)^2/(4+4).
  {
    clue::Queue q(0);  // host points never touch the device
    std::vector<float> x{2.f, 4.f}, w{1.f, 1.f}, s{2.f, 2.f};
    std::vector<int> out(2);
    clue::PointsHost<1> p(q, 2, x.data(), w.data(), out.data());
    p.set_sigma(0, std::span(s));
    CHECK(clue::metrics::Mahalanobis<1>{}(p, 0, 1) == std::sqrt(0.5f));
  }
  // Host constructor forms (PointsHost.hpp:59-137) over the same SoA.
  {
    clue::Queue q(0);
    const int n = 3;
    std::vector<float> soa{1.f, 2.f, 3.f, 10.f, 20.f, 30.f, .5f, .6f, .7f};
    std::vector<int> out(n, -7);
    auto same = [&](const clue::PointsHost<2>& p) {
      bool ok = p.size() == n;
      for (int i = 0; i < n; ++i)
        ok = ok && p.coords(0)[i] == soa[i] && p.coords(1)[i] == soa[n + i] &&
             p.weights()[i] == soa[2 * n + i];
      auto q = p[1];
      return ok && q[0] == 2.f && q[1] == 20.f && q[2] == .6f;
    };
    CHECK(same(clue::PointsHost<2>(q, n, std::span(soa), std::span(out))));
    CHECK(same(clue::PointsHost<2>(q, n, soa.data(), out.data())));
    std::vector<float> xy(soa.begin(), soa.begin() + 2 * n), wt(soa.begin() + 2 * n, soa.end());
    CHECK(same(clue::PointsHost<2>(q, n, std::span(xy), std::span(wt), std::span(out))));
    CHECK(same(clue::PointsHost<2>(q, n, xy.data(), wt.data(), out.data())));
    std::vector<float> x(soa.begin(), soa.begin() + n), y(soa.begin() + n, soa.begin() + 2 * n);
    CHECK(same(clue::PointsHost<2>(q, n, x, y, wt, out)));
    CHECK(same(clue::PointsHost<2>(q, n, x.data(), y.data(), wt.data(), out.data())));
    std::vector<std::byte> bytes(clue::PointsHost<2>::soa_bytes(n));
    std::memcpy(bytes.data(), soa.data(), sizeof(float) * soa.size());
    clue::PointsHost<2> pb(q, n, std::span(bytes));
    CHECK(same(pb));
    pb.clusterIndexes()[2] = 4;
    int tail;
    std::memcpy(&tail, bytes.data() + sizeof(float) * 3 * n + 2 * sizeof(int), sizeof(int));
    CHECK(tail == 4);  // cluster indexes follow the weights in the one-buffer layout
    bool threw = false;
    try {
      clue::PointsHost<2> bad(q, n, std::span(bytes).first(bytes.size() - 1));
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
      clue::PointsHost<2> bad(q, 0);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    CHECK(threw);
  }
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
