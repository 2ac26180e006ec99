// C++ drop-in checks: the reference's doctest cases (tests/test_clustering.cpp,
// tests/test_clusterer.cpp, tests/test_mahalanobis.cpp, tests/test_batched_clustering.cpp)
// written against include/cluestering_amd/CLUEstering.hpp. Needs a GPU; the pytest
// wrapper tests/test_cpp_api.py runs the built binary.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numbers>
#include <string>
#include <vector>

#include "cluestering_amd/CLUEstering.hpp"

static int g_fail = 0, g_pass = 0;
#define CHECK(cond)                                                        \
  do {                                                                     \
    if (cond) {                                                            \
      ++g_pass;                                                            \
    } else {                                                               \
      ++g_fail;                                                            \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);           \
    }                                                                      \
  } while (0)
#define CHECK_THROWS(expr)                                                 \
  do {                                                                     \
    bool thrown_ = false;                                                  \
    try {                                                                  \
      (void)(expr);                                                        \
    } catch (const std::exception&) {                                      \
      thrown_ = true;                                                      \
    }                                                                      \
    CHECK(thrown_);                                                        \
  } while (0)

static std::string data_dir;

static void test_constructor_throws() {
  // tests/test_clusterer.cpp:46-55
  auto q = clue::get_queue(0);
  CHECK_THROWS(clue::Clusterer<2>(-1.f, 10.f));
  CHECK_THROWS(clue::Clusterer<2>(1.f, -10.f));
  CHECK_THROWS(clue::Clusterer<2>(q, -1.f, 10.f));
  CHECK_THROWS(clue::Clusterer<2>(q, 1.f, -10.f));
  CHECK_THROWS(clue::FlatKernel<float>(-1.f));
  CHECK_THROWS(clue::ExponentialKernel<float>(-1.f, 1.f));
  CHECK_THROWS(clue::GaussianKernel<float>(1.f, 0.f));
}

static void test_tag_tie_break() {
  // tests/test_clustering.cpp:517-567
  auto q = clue::get_queue(0u);
  clue::Clusterer<2> algo(q, 1.f, 0.f, 1.f);
  std::vector<float> input{0.f, 0.f, 0.f, 0.f, 1.f, 1.f};
  std::vector<int> output(2);
  auto read_back = [&](clue::PointsDevice<2>& d) {
    std::vector<int32_t> nh(2), seed(2);
    clue::detail::check(clue_memcpy_d2h(q.ctx(), nh.data(), d.view().nearest_higher().data(), 8));
    clue::detail::check(clue_memcpy_d2h(q.ctx(), seed.data(), d.view().is_seed().data(), 8));
    return std::make_pair(nh, seed);
  };
  {
    clue::PointsHost<2> h(q, 2, input.data(), output.data());
    clue::PointsDevice<2> d(q, 2);
    algo.make_clusters(q, h, d);
    auto [nh, seed] = read_back(d);
    CHECK(nh[0] == 1 && nh[1] == -1 && seed[0] == 0 && seed[1] == 1);
  }
  {
    clue::PointsHost<2> h(q, 2, input.data(), output.data());
    std::vector<std::uint32_t> tags{100, 1};
    h.set_tags(tags);
    clue::PointsDevice<2> d(q, 2);
    algo.make_clusters(q, h, d);
    auto [nh, seed] = read_back(d);
    CHECK(nh[1] == 0 && nh[0] == -1 && seed[1] == 0 && seed[0] == 1);
  }
}

static void test_mahalanobis() {
  // tests/test_mahalanobis.cpp:10-43
  auto q = clue::get_queue(0u);
  clue::Clusterer<1> algo(q, 1.f, 0.3f);
  {
    clue::PointsHost<1> p(q, 2);
    p.coords(0)[0] = 2.f;
    p.coords(0)[1] = 4.f;
    p.weights()[0] = p.weights()[1] = 1.f;
    algo.make_clusters(q, p);
    CHECK(p.n_clusters() == 2);
  }
  {
    clue::PointsHost<1> p(q, 2);
    p.coords(0)[0] = 2.f;
    p.coords(0)[1] = 4.f;
    p.weights()[0] = p.weights()[1] = 1.f;
    std::vector<float> sigma = {2.f, 2.f};
    p.set_sigma(0, std::span(sigma));
    algo.make_clusters(q, p, clue::metrics::Mahalanobis<1>{});
    CHECK(p.n_clusters() == 1);
  }
}

static void test_periodic() {
  // tests/test_clustering.cpp:237-270
  auto q = clue::get_queue(0u);
  auto points = clue::read_csv<2, float>(q, data_dir + "/opposite_angles.csv");
  clue::Clusterer<2> algo(q, .2f, 5.f, .2f);
  algo.setWrappedCoordinates(0, 1);
  algo.make_clusters(q, points,
                     clue::metrics::PeriodicEuclidean<2>(
                         std::array<float, 2>{0.f, 2.f * std::numbers::pi_v<float>}));
  CHECK(points.n_clusters() == 1);
}

static void test_batched() {
  // tests/test_batched_clustering.cpp:8-65 on 10 copies of data_1024.
  auto q = clue::get_queue(0u);
  auto one = clue::read_csv<2, float>(q, data_dir + "/data_1024.csv");
  clue::Clusterer<2> algo(q, 1.3f, 10.f, 1.3f);
  algo.make_clusters(q, one);
  const auto truth = one.n_clusters();
  clue::PointsHost<2> h(q, 10240);
  for (int k = 0; k < 10; ++k)
    for (int i = 0; i < 1024; ++i) {
      h.coords(0)[k * 1024 + i] = one.coords(0)[i];
      h.coords(1)[k * 1024 + i] = one.coords(1)[i];
      h.weights()[k * 1024 + i] = one.weights()[i];
    }
  clue::PointsDevice<2> d(q, 10240);
  std::vector<uint32_t> sizes(10, 1024);
  algo.make_clusters(q, h, d, sizes);
  CHECK(h.n_clusters() == truth * 10);
  auto assoc = algo.getSampleAssociations(q, h);
  CHECK(assoc.size() == 10);
  for (int e = 0; e < 10; ++e) CHECK(assoc.count(e) == truth);
  // Every event's labels equal the single-event labels shifted by e * truth.
  bool same = true;
  for (int k = 0; k < 10; ++k)
    for (int i = 0; i < 1024; ++i) {
      const int a = one.clusterIndexes()[i];
      const int b = h.clusterIndexes()[k * 1024 + i];
      same = same && (a < 0 ? b == -1 : b == a + k * static_cast<int>(truth));
    }
  CHECK(same);
  auto seeds = algo.getSeeds();
  CHECK(seeds.size() == truth * 10);
  // The device event -> clusters map equals the host one.
  auto dassoc = algo.getSampleAssociations(q, d);
  CHECK(dassoc.size() == 10);
  CHECK(dassoc.n_elements() == truth * 10);
  auto dh = dassoc.to_host();
  CHECK(dh.offsets() == assoc.offsets());
  CHECK(dh.indexes() == assoc.indexes());
}

static void test_device_points_and_clusters() {
  auto q = clue::get_queue(0u);
  auto h = clue::read_csv<2, float>(q, data_dir + "/data_4096.csv");
  clue::PointsDevice<2> d(q, h.size());
  clue::copyToDevice(q, d, h);
  clue::Clusterer<2> algo(q, 1.5f, 10.f, 1.5f);
  algo.make_clusters(q, d);
  clue::copyToHost(q, h, d);
  auto clusters = algo.getClusters(h);
  CHECK(clusters.size() == h.n_clusters());
  std::size_t total = 0;
  for (std::size_t c = 0; c < clusters.size(); ++c) total += clusters.count(int32_t(c));
  std::size_t clustered = 0;
  for (int v : h.clusterIndexes()) clustered += v >= 0;
  CHECK(total == clustered);
  CHECK(h.n_clusters() == 20);
  // getClusters on device points: the same CSR, built on the device
  // (members ascending within a cluster, as the host fill).
  auto dclusters = algo.getClusters(q, d);
  CHECK(dclusters.size() == clusters.size());
  CHECK(dclusters.n_elements() == clustered);
  auto back = dclusters.to_host();
  CHECK(back.offsets() == clusters.offsets());
  CHECK(back.indexes() == clusters.indexes());
  CHECK(dclusters.count(3) == clusters.count(3));
  auto free_fn = clue::get_clusters(q, d);
  CHECK(free_fn.to_host().indexes() == clusters.indexes());
  // read_output on a result file written from these points.
  const std::string out = "/tmp/cluestering_amd_read_output_test.csv";
  {
    std::FILE* fp = std::fopen(out.c_str(), "w");
    std::fprintf(fp, "x0,x1,weight,cluster_ids\n");
    for (int32_t i = 0; i < h.size(); ++i)
      std::fprintf(fp, "%.9g,%.9g,%.9g,%d\n", h.coords(0)[i], h.coords(1)[i], h.weights()[i],
                   h.clusterIndexes()[i]);
    std::fclose(fp);
  }
  auto back2 = clue::read_output<2, float>(q, out);
  CHECK(back2.clustered());
  CHECK(back2.size() == h.size());
  bool same_ids = true;
  for (int32_t i = 0; i < h.size(); ++i)
    same_ids = same_ids && back2.clusterIndexes()[i] == h.clusterIndexes()[i] &&
               back2.coords(0)[i] == h.coords(0)[i];
  CHECK(same_ids);
  CHECK(back2.n_clusters() == h.n_clusters());
  std::remove(out.c_str());
}

static void test_device_buffer_forms() {
  // PointsDevice over caller device memory (data_structures/PointsDevice.hpp:51-122):
  // every form gives the labels and densities of the owning form, and the one-buffer
  // form writes its results where the reference's layout puts them.
  auto q = clue::get_queue(0u);
  auto h = clue::read_csv<2, float>(q, data_dir + "/data_4096.csv");
  const int32_t n = h.size();
  const std::size_t nn = std::size_t(n);
  clue::Clusterer<2> algo(q, 1.5f, 10.f, 1.5f);
  clue::PointsDevice<2> own(q, n);
  clue::copyToDevice(q, own, h);
  algo.make_clusters(q, own);
  std::vector<int32_t> want(nn), got(nn);
  std::vector<float> want_rho(nn), got_rho(nn);
  clue::detail::check(clue_memcpy_d2h(q.ctx(), want.data(), own.view().m_cluster_index, 4 * nn));
  clue::detail::check(clue_memcpy_d2h(q.ctx(), want_rho.data(), own.view().m_rho, 4 * nn));
  CHECK(own.n_clusters() == 20);

  void* mem = nullptr;
  const std::size_t bytes = clue::PointsDevice<2>::soa_bytes(n);
  clue::detail::check(clue_malloc(q.ctx(), bytes, &mem));
  std::byte* buf = static_cast<std::byte*>(mem);
  float* fbuf = static_cast<float*>(mem);
  auto run_and_compare = [&](clue::PointsDevice<2>& d) {
    clue::detail::check(clue_memset_d8(q.ctx(), d.view().m_cluster_index, 0xff, 4 * nn));
    clue::copyToDevice(q, d, h);
    algo.make_clusters(q, d);
    clue::detail::check(clue_memcpy_d2h(q.ctx(), got.data(), d.view().m_cluster_index, 4 * nn));
    clue::detail::check(clue_memcpy_d2h(q.ctx(), got_rho.data(), d.view().m_rho, 4 * nn));
    return got == want && std::memcmp(got_rho.data(), want_rho.data(), 4 * nn) == 0;
  };
  {
    clue::PointsDevice<2> d(q, n, std::span<std::byte>(buf, bytes));
    CHECK(run_and_compare(d));
    // [x][y][w][cluster_index][is_seed][rho][nearest_higher]
    std::vector<int32_t> at(nn);
    std::vector<float> rho_at(nn);
    clue::detail::check(clue_memcpy_d2h(q.ctx(), at.data(), buf + 3 * 4 * nn, 4 * nn));
    clue::detail::check(clue_memcpy_d2h(q.ctx(), rho_at.data(), buf + 5 * 4 * nn, 4 * nn));
    CHECK(at == want);
    CHECK(std::memcmp(rho_at.data(), want_rho.data(), 4 * nn) == 0);
    CHECK_THROWS(clue::PointsDevice<2>(q, n, std::span<std::byte>(buf, bytes - 1)));
  }
  int32_t* out = reinterpret_cast<int32_t*>(fbuf + 3 * nn);
  {
    clue::PointsDevice<2> d(q, n, std::span<float>(fbuf, 3 * nn), std::span<int>(out, nn));
    CHECK(run_and_compare(d));
    CHECK(d.view().m_cluster_index == out && d.view().m_weight == fbuf + 2 * nn);
  }
  {
    clue::PointsDevice<2> d(q, n, fbuf, out);
    CHECK(run_and_compare(d));
  }
  {
    clue::PointsDevice<2> d(q, n, std::span<float>(fbuf, 2 * nn), std::span<float>(fbuf + 2 * nn, nn),
                            std::span<int>(out, nn));
    CHECK(run_and_compare(d));
  }
  {
    clue::PointsDevice<2> d(q, n, fbuf, fbuf + 2 * nn, out);
    CHECK(run_and_compare(d));
  }
  {
    clue::PointsDevice<2> d(q, n, fbuf, fbuf + nn, fbuf + 2 * nn, out);
    CHECK(run_and_compare(d));
    CHECK(d.view().m_coords[1] == fbuf + nn);
  }
  CHECK_THROWS(clue::PointsDevice<2>(q, 0));
  clue::detail::check(clue_free(q.ctx(), mem));
}

static void test_precisions_and_silhouette() {
  // tests/test_clustering.cpp:14-57, 94-130: single and double precision on the
  // benchmark and sissa fixtures, judged by the silhouette score.
  auto q = clue::get_queue(0u);
  for (const char* name : {"data_1024.csv", "data_4096.csv"}) {
    clue::PointsHost<2> hf = clue::read_csv<2, float>(q, data_dir + "/" + name);
    clue::PointsDevice<2> df(q, hf.size());
    clue::Clusterer<2> af(q, 1.5f, 10.f, 1.5f);
    af.make_clusters(q, hf, df);
    CHECK(clue::silhouette(hf) >= 0.9f);

    clue::PointsHost<2, double> hd = clue::read_csv<2, double>(q, data_dir + "/" + name);
    clue::PointsDevice<2, double> dd(q, hd.size());
    clue::Clusterer<2, double> ad(q, 1.5, 10., 1.5);
    ad.make_clusters(q, hd, dd);
    CHECK(clue::silhouette(hd) >= 0.9);
    CHECK(hd.n_clusters() == hf.n_clusters());
    CHECK(ad.getSeeds().size() == hd.n_clusters());
  }
  {
    clue::PointsHost<2, double> h = clue::read_csv<2, double>(q, data_dir + "/sissa_1000.csv");
    clue::PointsDevice<2, double> d(q, h.size());
    clue::Clusterer<2, double> algo(q, 20., 10., 20.);
    algo.make_clusters(q, h, d);
    CHECK(clue::silhouette(h) >= 0.5);
  }
  {
    // Double batched: 3 copies of data_1024, each event clustered alone.
    auto one = clue::read_csv<2, double>(q, data_dir + "/data_1024.csv");
    clue::Clusterer<2, double> algo(q, 1.3, 10., 1.3);
    algo.make_clusters(q, one);
    const auto truth = one.n_clusters();
    clue::PointsHost<2, double> h(q, 3 * 1024);
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 1024; ++i) {
        h.coords(0)[k * 1024 + i] = one.coords(0)[i];
        h.coords(1)[k * 1024 + i] = one.coords(1)[i];
        h.weights()[k * 1024 + i] = one.weights()[i];
      }
    clue::PointsDevice<2, double> d(q, 3 * 1024);
    std::vector<uint32_t> sizes(3, 1024);
    algo.make_clusters(q, h, d, sizes, clue::EuclideanMetric<2, double>{},
                       clue::GaussianKernel<double>(1.0, 1.0));
    clue::PointsHost<2, double> g(q, 1024);
    for (int i = 0; i < 1024; ++i) {
      g.coords(0)[i] = one.coords(0)[i];
      g.coords(1)[i] = one.coords(1)[i];
      g.weights()[i] = one.weights()[i];
    }
    algo.make_clusters(q, g, clue::EuclideanMetric<2, double>{},
                       clue::GaussianKernel<double>(1.0, 1.0));
    const int per = static_cast<int>(g.n_clusters());
    bool same = true;
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 1024; ++i) {
        const int a = g.clusterIndexes()[i];
        const int b = h.clusterIndexes()[k * 1024 + i];
        same = same && (a < 0 ? b == -1 : b == a + k * per);
      }
    CHECK(same);
    CHECK(truth > 0);
  }
}

int main(int argc, char** argv) {
  data_dir = argc > 1 ? argv[1] : "tests/golden/reference_data";
  try {
    test_constructor_throws();
    test_tag_tie_break();
    test_mahalanobis();
    test_periodic();
    test_batched();
    test_device_points_and_clusters();
    test_device_buffer_forms();
    test_precisions_and_silhouette();
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    ++g_fail;
  }
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
