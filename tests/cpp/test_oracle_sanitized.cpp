// The oracle under the address and undefined-behaviour sanitizers (built with
// g++ from oracle/clue_oracle.cpp, no GPU): every entry point on small random
// inputs over the kernels, metrics, optional inputs, precisions and batched
// geometries, with the invariants of a CLUE result checked on each.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "clue_oracle.h"

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

template <class T>
static bool consistent(int n, const std::vector<int32_t>& nh, const std::vector<int32_t>& seed,
                       const std::vector<int32_t>& cl, int64_t n_seeds, const std::vector<T>& rho) {
  int64_t seeds = 0;
  bool ok = true;
  for (int i = 0; i < n; ++i) {
    seeds += seed[i];
    ok = ok && nh[i] >= -1 && nh[i] < n && nh[i] != i && cl[i] >= -1 && cl[i] < n_seeds;
    ok = ok && std::isfinite(static_cast<double>(rho[i]));
    if (seed[i]) ok = ok && cl[i] >= 0;
  }
  return ok && seeds == n_seeds;
}

int main() {
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> U(0.f, 10.f), W(0.5f, 1.5f), S(0.5f, 2.f);
  for (int D = 1; D <= 4; ++D)
    for (int kernel = 0; kernel < 3; ++kernel)
      for (int metric = 0; metric < 7; ++metric) {
        const int n = 300 + 37 * D;
        std::vector<float> c(static_cast<size_t>(D) * n), w(n), sig(static_cast<size_t>(D) * n), u(n);
        std::vector<uint32_t> tags(n);
        for (auto& v : c) v = U(rng);
        for (auto& v : w) v = W(rng);
        for (auto& v : sig) v = S(rng);
        for (int i = 0; i < n; ++i) {
          u[i] = W(rng);
          tags[i] = static_cast<uint32_t>((i * 7919) % n);
        }
        oracle_params p{};
        p.ndim = D;
        p.kernel = kernel;
        p.metric = metric;
        p.points_per_tile = 128;
        p.dc = 1.5;
        p.rhoc = 2.0;
        p.dm = 2.0;
        p.seed_dc = 1.5;
        p.kparam[0] = kernel == 0 ? 0.5 : 1.0;
        p.kparam[1] = 1.0;
        for (int d = 0; d < D; ++d) p.mparam[d] = metric == 2 ? 10.0 : 1.0 + 0.25 * d;
        if (metric == 2) p.wrapped[D - 1] = 1;
        const bool mah = metric == 6;
        std::vector<float> rho(n), delta(n);
        std::vector<int32_t> nh(n), seed(n), cl(n), toff(1 << 16), tidx(n);
        int64_t ns = 0;
        oracle_out_f32 out{rho.data(), nh.data(), delta.data(), seed.data(), cl.data(),
                           toff.data(), tidx.data(), &ns};
        const int rc = oracle_run_f32(&p, n, c.data(), w.data(), mah ? sig.data() : nullptr,
                                      u.data(), (metric & 1) ? tags.data() : nullptr, &out);
        CHECK(rc == 0);
        CHECK(consistent(n, nh, seed, cl, ns, rho));
        std::vector<double> r64(n), a64(n);
        std::vector<int32_t> cnt(n);
        CHECK(oracle_brute_density_f32(&p, n, c.data(), w.data(), mah ? sig.data() : nullptr,
                                       r64.data(), a64.data(), cnt.data()) == 0);
        std::vector<int32_t> bnh(n);
        std::vector<float> bd(n);
        CHECK(oracle_brute_nh_f32(&p, n, c.data(), mah ? sig.data() : nullptr, u.data(), nullptr,
                                  rho.data(), bnh.data(), bd.data()) == 0);
        int64_t vd = 0, ir = 0, vn = 0, np = 0;
        CHECK(oracle_pair_counts_f32(&p, n, c.data(), w.data(), mah ? sig.data() : nullptr, nullptr,
                                     nullptr, rho.data(), &vd, &ir, &vn, &np) == 0);
        CHECK(ir >= n && vd >= ir && vn >= np);
        // Double precision on the same points.
        std::vector<double> cd(c.begin(), c.end()), wd(w.begin(), w.end()),
            sd(sig.begin(), sig.end()), ud(u.begin(), u.end()), rd(n), dd(n);
        oracle_out_f64 od{rd.data(), nh.data(), dd.data(), seed.data(), cl.data(), nullptr,
                          nullptr, &ns};
        CHECK(oracle_run_f64(&p, n, cd.data(), wd.data(), mah ? sd.data() : nullptr, ud.data(),
                             nullptr, &od) == 0);
        CHECK(consistent(n, nh, seed, cl, ns, rd));
        // Batched, both geometries, with an empty event in the middle.
        const uint32_t sizes[4] = {static_cast<uint32_t>(n / 3), 0, static_cast<uint32_t>(n / 3),
                                   static_cast<uint32_t>(n - 2 * (n / 3))};
        for (int geo = 0; geo < 2; ++geo) {
          p.geometry = geo;
          std::vector<int32_t> eoc(n);
          oracle_out_f32 ob{rho.data(), nh.data(), delta.data(), seed.data(), cl.data(), nullptr,
                            nullptr, &ns};
          CHECK(oracle_run_batched_f32(&p, n, c.data(), w.data(), mah ? sig.data() : nullptr,
                                       nullptr, nullptr, sizes, 4, &ob, eoc.data()) == 0);
          CHECK(consistent(n, nh, seed, cl, ns, rho));
          bool events_ok = true;
          for (int64_t k = 0; k < ns; ++k) events_ok = events_ok && eoc[k] >= 0 && eoc[k] < 4 && eoc[k] != 1;
          CHECK(events_ok);
        }
      }
  // Invalid arguments are refused, not dereferenced.
  oracle_params bad{};
  bad.ndim = 11;
  float x = 0.f;
  oracle_out_f32 o{};
  CHECK(oracle_run_f32(&bad, 1, &x, &x, nullptr, nullptr, nullptr, &o) != 0);
  CHECK(std::fabs(oracle_expf(-1.f) - std::exp(-1.f)) < 1e-6f);
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
