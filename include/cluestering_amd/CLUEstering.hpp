// Header-only C++ drop-in for the reference's clue::Clusterer / PointsHost /
// PointsDevice API (include/CLUEstering/core/Clusterer.hpp:37-312,
// data_structures/PointsHost.hpp:29-257, data_structures/PointsDevice.hpp:32),
// running on libclue_amd.so through the C ABI in clue_amd.h.
//
// Link with -lclue_amd. Points are float or double (PointsHost/PointsDevice<Ndim, T>,
// Clusterer<Ndim, T>); every device stage computes in T.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <fstream>
#include <memory>
#include <mutex>
#include <optional>
#include <ranges>
#include <span>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../clue_amd.h"

namespace clue {

namespace detail {

inline void check(clue_status s) {
  if (s == CLUE_OK) return;
  const std::string msg = clue_last_error();
  switch (s) {
    case CLUE_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case CLUE_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
    default: throw std::runtime_error(msg);
  }
}

template <class T>
struct DeviceDeleter {
  clue_ctx* ctx;
  void operator()(T* p) const {
    if (p) clue_free(ctx, p);
  }
};

// Decimal text to T, rounded once (utils/detail/read_csv.hpp:20-30 parses per type).
template <class T>
T parse_floating(const std::string& s) {
  if constexpr (std::is_same_v<T, float>)
    return std::stof(s);
  else
    return std::stod(s);
}

// The C structs and entry points of each precision (clue_amd.h).
template <class P>
using param_t = std::remove_cvref_t<decltype(std::declval<P&>().dc)>;

template <class T>
struct CApi;
template <>
struct CApi<float> {
  using Params = clue_params;
  using Points = clue_points;
  using Results = clue_results;
  static clue_status run(clue_ctx* c, const Params* p, const Points* pts, Results* r) {
    return clue_make_clusters(c, p, pts, r);
  }
  static clue_status run_batched(clue_ctx* c, const Params* p, const Points* pts,
                                 const uint32_t* sizes, int32_t n_events, Results* r,
                                 int32_t* eoc, int64_t* n_clusters) {
    return clue_make_clusters_batched(c, p, pts, sizes, n_events, r, eoc, n_clusters);
  }
};
template <>
struct CApi<double> {
  using Params = clue_params_f64;
  using Points = clue_points_f64;
  using Results = clue_results_f64;
  static clue_status run(clue_ctx* c, const Params* p, const Points* pts, Results* r) {
    return clue_make_clusters_f64(c, p, pts, r);
  }
  static clue_status run_batched(clue_ctx* c, const Params* p, const Points* pts,
                                 const uint32_t* sizes, int32_t n_events, Results* r,
                                 int32_t* eoc, int64_t* n_clusters) {
    return clue_make_clusters_batched_f64(c, p, pts, sizes, n_events, r, eoc, n_clusters);
  }
};

}  // namespace detail

namespace concepts {
template <class M>
concept distance_metric = M::is_metric && requires(const M& m, clue_params& p) { m.fill(p); };
template <class K>
concept convolutional_kernel = K::is_kernel && requires(const K& k, clue_params& p) { k.fill(p); };
}  // namespace concepts

// The reference's Queue (utils/get_queue.hpp:19-48): a device and a stream. The
// library context behind it is created on first use (copies share it), so host-only
// work such as PointsHost never touches the device.
class Queue {
 public:
  explicit Queue(int device = 0, void* stream = nullptr)
      : m_state{std::make_shared<State>()}, m_device{device} {
    m_state->stream = stream;
  }
  clue_ctx* ctx() const {
    std::call_once(m_state->once, [&] {
      clue_ctx* c = nullptr;
      m_state->status = clue_context_create(m_device, m_state->stream, &c);
      if (m_state->status == CLUE_OK) m_state->ctx = c;
      else m_state->error = clue_last_error();
    });
    if (!m_state->ctx)
      throw std::runtime_error("could not create a context on device " +
                               std::to_string(m_device) + ": " + m_state->error);
    return m_state->ctx;
  }
  int device() const { return m_device; }
  void wait() const { detail::check(clue_synchronize(ctx())); }
  void* stream() const {
    void* s = nullptr;
    detail::check(clue_context_stream(ctx(), &s));
    return s;
  }

 private:
  struct State {
    std::once_flag once;
    clue_ctx* ctx = nullptr;
    void* stream = nullptr;
    clue_status status = CLUE_OK;
    std::string error;
    ~State() {
      if (ctx) clue_context_destroy(ctx);
    }
  };
  std::shared_ptr<State> m_state;
  int m_device = 0;
};

inline Queue get_queue(int device_id = 0) { return Queue(device_id); }
inline void wait(const Queue& q) { q.wait(); }

// ---- Convolutional kernels (core/ConvolutionalKernel.hpp:19-111) -------------------------
template <class T = float>
class FlatKernel {
 public:
  using value_type = T;
  static constexpr bool is_kernel = true;
  FlatKernel(T flat) : m_flat{flat} {
    if (!(flat > T{0})) throw std::invalid_argument("Flat kernel parameter must be positive.");
  }
  template <class P>
  void fill(P& p) const {
    p.kernel = CLUE_KERNEL_FLAT;
    p.kparam[0] = static_cast<detail::param_t<P>>(m_flat);
  }

 private:
  T m_flat;
};

template <class T = float>
class ExponentialKernel {
 public:
  using value_type = T;
  static constexpr bool is_kernel = true;
  ExponentialKernel(T avg, T amplitude) : m_avg{avg}, m_amp{amplitude} {
    if (!(avg > T{0}) || !(amplitude > T{0}))
      throw std::invalid_argument("Exponential kernel parameters must be positive.");
  }
  template <class P>
  void fill(P& p) const {
    p.kernel = CLUE_KERNEL_EXP;
    p.kparam[0] = static_cast<detail::param_t<P>>(m_avg);
    p.kparam[1] = static_cast<detail::param_t<P>>(m_amp);
  }

 private:
  T m_avg, m_amp;
};

template <class T = float>
class GaussianKernel {
 public:
  using value_type = T;
  static constexpr bool is_kernel = true;
  GaussianKernel(T stddev, T amplitude) : m_std{stddev}, m_amp{amplitude} {
    if (!(stddev > T{0}) || !(amplitude > T{0}))
      throw std::invalid_argument("Gaussian kernel parameters must be positive.");
  }
  template <class P>
  void fill(P& p) const {
    p.kernel = CLUE_KERNEL_GAUSS;
    p.kparam[0] = static_cast<detail::param_t<P>>(m_std);
    p.kparam[1] = static_cast<detail::param_t<P>>(m_amp);
  }

 private:
  T m_std, m_amp;
};

// ---- Distance metrics (core/DistanceMetrics.hpp:64-371) ----------------------------------
// Each metric carries its parameters to the device (fill) and can be evaluated on
// the host on two points, coordinates first and the weight last (Point<Ndim, T>,
// DistanceMetrics.hpp:48-49). Sums run as the reference's right fold
// d0 + (d1 + (... + 0)) (internal/meta/accumulate.hpp), maxima as the running
// std::max from 0 (internal/meta/maximum.hpp); the device stages evaluate the
// same expressions in the same order.
template <std::size_t Ndim, class T = float>
using Point = std::array<T, Ndim + 1>;

namespace detail {
template <std::size_t Ndim, class T, class F>
T right_fold_sum(F&& term) {
  T acc = T(0);
  for (std::size_t k = Ndim; k-- > 0;) acc = term(k) + acc;
  return acc;
}
template <std::size_t Ndim, class T, class F>
T running_max(F&& term) {
  T m = T(0);
  for (std::size_t k = 0; k < Ndim; ++k) m = std::max(m, term(k));
  return m;
}
}  // namespace detail

template <std::size_t Ndim, class T = float>
struct EuclideanMetric {
  using value_type = T;
  static constexpr bool is_metric = true;
  template <class P>
  void fill(P& p) const { p.metric = CLUE_METRIC_EUCLIDEAN; }
  // core/DistanceMetrics.hpp:108-113
  T operator()(const Point<Ndim, T>& a, const Point<Ndim, T>& b) const {
    return std::sqrt(detail::right_fold_sum<Ndim, T>(
        [&](std::size_t k) { return (a[k] - b[k]) * (a[k] - b[k]); }));
  }
};
template <std::size_t Ndim, class T = float>
struct WeightedEuclideanMetric {
  using value_type = T;
  static constexpr bool is_metric = true;
  std::array<T, Ndim> weights;
  WeightedEuclideanMetric(const std::array<T, Ndim>& w) : weights{w} {}
  template <class... V>
    requires(sizeof...(V) == Ndim)
  WeightedEuclideanMetric(V... w) : weights{static_cast<T>(w)...} {}
  template <class P>
  void fill(P& p) const {
    p.metric = CLUE_METRIC_WEIGHTED_EUCLIDEAN;
    for (std::size_t d = 0; d < Ndim; ++d) p.mparam[d] = static_cast<detail::param_t<P>>(weights[d]);
  }
  // core/DistanceMetrics.hpp:156-162: (w * diff) * diff per dimension.
  T operator()(const Point<Ndim, T>& a, const Point<Ndim, T>& b) const {
    return std::sqrt(detail::right_fold_sum<Ndim, T>(
        [&](std::size_t k) { return weights[k] * (a[k] - b[k]) * (a[k] - b[k]); }));
  }
};
template <std::size_t Ndim, class T = float>
struct PeriodicEuclideanMetric {
  using value_type = T;
  static constexpr bool is_metric = true;
  std::array<T, Ndim> periods;
  PeriodicEuclideanMetric(const std::array<T, Ndim>& per) : periods{per} {}
  template <class... V>
    requires(sizeof...(V) == Ndim)
  PeriodicEuclideanMetric(V... per) : periods{static_cast<T>(per)...} {}
  template <class P>
  void fill(P& p) const {
    p.metric = CLUE_METRIC_PERIODIC_EUCLIDEAN;
    for (std::size_t d = 0; d < Ndim; ++d) p.mparam[d] = static_cast<detail::param_t<P>>(periods[d]);
  }
  // core/DistanceMetrics.hpp:212-220: min(|diff|, period - |diff|) per dimension.
  T operator()(const Point<Ndim, T>& a, const Point<Ndim, T>& b) const {
    return std::sqrt(detail::right_fold_sum<Ndim, T>([&](std::size_t k) {
      const T diff = std::fabs(a[k] - b[k]);
      const T pd = std::min(diff, periods[k] - diff);
      return pd * pd;
    }));
  }
};
template <std::size_t Ndim, class T = float>
struct ManhattanMetric {
  using value_type = T;
  static constexpr bool is_metric = true;
  template <class P>
  void fill(P& p) const { p.metric = CLUE_METRIC_MANHATTAN; }
  // core/DistanceMetrics.hpp:241-245
  T operator()(const Point<Ndim, T>& a, const Point<Ndim, T>& b) const {
    return detail::right_fold_sum<Ndim, T>([&](std::size_t k) { return std::fabs(a[k] - b[k]); });
  }
};
template <std::size_t Ndim, class T = float>
struct ChebyshevMetric {
  using value_type = T;
  static constexpr bool is_metric = true;
  template <class P>
  void fill(P& p) const { p.metric = CLUE_METRIC_CHEBYSHEV; }
  // core/DistanceMetrics.hpp:266-270
  T operator()(const Point<Ndim, T>& a, const Point<Ndim, T>& b) const {
    return detail::running_max<Ndim, T>([&](std::size_t k) { return std::fabs(a[k] - b[k]); });
  }
};
template <std::size_t Ndim, class T = float>
struct WeightedChebyshevMetric {
  using value_type = T;
  static constexpr bool is_metric = true;
  std::array<T, Ndim> weights;
  WeightedChebyshevMetric(const std::array<T, Ndim>& w) : weights{w} {}
  template <class... V>
    requires(sizeof...(V) == Ndim)
  WeightedChebyshevMetric(V... w) : weights{static_cast<T>(w)...} {}
  template <class P>
  void fill(P& p) const {
    p.metric = CLUE_METRIC_WEIGHTED_CHEBYSHEV;
    for (std::size_t d = 0; d < Ndim; ++d) p.mparam[d] = static_cast<detail::param_t<P>>(weights[d]);
  }
  // core/DistanceMetrics.hpp:313-318
  T operator()(const Point<Ndim, T>& a, const Point<Ndim, T>& b) const {
    return detail::running_max<Ndim, T>(
        [&](std::size_t k) { return weights[k] * std::fabs(a[k] - b[k]); });
  }
};
template <std::size_t Ndim, class T>
class PointsHost;
template <std::size_t Ndim, class T = float>
struct MahalanobisMetric {
  using value_type = T;
  static constexpr bool is_metric = true;
  template <class P>
  void fill(P& p) const { p.metric = CLUE_METRIC_MAHALANOBIS; }
  // core/DistanceMetrics.hpp:75-87: diff^2 / (sigma_i^2 + sigma_j^2) per dimension,
  // over host points whose sigma is set in every dimension.
  T operator()(const PointsHost<Ndim, T>& points, std::size_t i, std::size_t j) const;
};

namespace metrics {
template <std::size_t Ndim, class T = float>
using Euclidean = EuclideanMetric<Ndim, T>;
template <std::size_t Ndim, class T = float>
using WeightedEuclidean = WeightedEuclideanMetric<Ndim, T>;
template <std::size_t Ndim, class T = float>
using PeriodicEuclidean = PeriodicEuclideanMetric<Ndim, T>;
template <std::size_t Ndim, class T = float>
using Manhattan = ManhattanMetric<Ndim, T>;
template <std::size_t Ndim, class T = float>
using Chebyshev = ChebyshevMetric<Ndim, T>;
template <std::size_t Ndim, class T = float>
using WeightedChebyshev = WeightedChebyshevMetric<Ndim, T>;
template <std::size_t Ndim, class T = float>
using Mahalanobis = MahalanobisMetric<Ndim, T>;
}  // namespace metrics

// ---- Host CSR of cluster -> points (AssociationMapHost, data_structures/AssociationMap.hpp) --
class AssociationMapHost {
 public:
  AssociationMapHost() = default;
  AssociationMapHost(std::vector<int32_t> offsets, std::vector<int32_t> indexes)
      : m_off(std::move(offsets)), m_idx(std::move(indexes)) {}
  // Builds key -> values from an association array (utils/detail/get_clusters.hpp:24-49):
  // keys 0..max, values ascending, negative keys dropped.
  static AssociationMapHost from_associations(std::span<const int32_t> assoc) {
    int32_t nkeys = 0;
    for (int32_t k : assoc) nkeys = std::max(nkeys, k + 1);
    std::vector<int32_t> off(static_cast<std::size_t>(nkeys) + 1, 0);
    for (int32_t k : assoc)
      if (k >= 0) off[k + 1]++;
    for (int32_t k = 0; k < nkeys; ++k) off[k + 1] += off[k];
    std::vector<int32_t> fill(off.begin(), off.end() - 1), idx(off.back());
    for (std::size_t i = 0; i < assoc.size(); ++i)
      if (assoc[i] >= 0) idx[fill[assoc[i]]++] = static_cast<int32_t>(i);
    return AssociationMapHost(std::move(off), std::move(idx));
  }
  std::size_t size() const { return m_off.empty() ? 0 : m_off.size() - 1; }
  std::span<const int32_t> operator[](int32_t key) const {
    return {m_idx.data() + m_off[key], static_cast<std::size_t>(m_off[key + 1] - m_off[key])};
  }
  std::size_t count(int32_t key) const { return static_cast<std::size_t>(m_off[key + 1] - m_off[key]); }
  const int32_t* lower_bound(int32_t key) const { return m_idx.data() + m_off[key]; }
  const int32_t* upper_bound(int32_t key) const { return m_idx.data() + m_off[key + 1]; }
  const std::vector<int32_t>& offsets() const { return m_off; }
  const std::vector<int32_t>& indexes() const { return m_idx; }

 private:
  std::vector<int32_t> m_off, m_idx;
};

// ---- Device CSR of key -> members (AssociationMap<Device>, data_structures/AssociationMap.hpp),
// built on the device by clue_make_association (internal/MakeAssociator.hpp:15-45). -----------
class AssociationMapDevice {
 public:
  AssociationMapDevice() = default;
  // keys: device array of n association keys (cluster or event ids; negative = none).
  AssociationMapDevice(clue_ctx* ctx, const int32_t* keys, int32_t n) : m_ctx{ctx} {
    void* mem = nullptr;
    detail::check(clue_malloc(ctx, 4 * std::size_t(std::max(n, 1)), &mem));
    m_idx.reset(static_cast<int32_t*>(mem));
    m_idx.get_deleter().ctx = ctx;
    int32_t cap = n + 2;
    alloc_offsets(cap);
    clue_status st = clue_make_association(ctx, keys, n, m_off.get(), cap, m_idx.get(), &m_bins,
                                           &m_elems);
    if (st == CLUE_ERR_CAPACITY && m_bins + 1 > cap) {  // keys beyond n: size and retry
      cap = m_bins + 1;
      alloc_offsets(cap);
      st = clue_make_association(ctx, keys, n, m_off.get(), cap, m_idx.get(), &m_bins, &m_elems);
    }
    detail::check(st);
  }
  // Number of keys (max key + 1) and of associated elements.
  std::size_t size() const { return static_cast<std::size_t>(m_bins); }
  std::size_t n_elements() const { return static_cast<std::size_t>(m_elems); }
  // Device spans: bins + 1 offsets, n_elements member indices.
  std::span<const int32_t> offsets() const { return {m_off.get(), std::size_t(m_bins) + 1}; }
  std::span<const int32_t> indexes() const { return {m_idx.get(), std::size_t(m_elems)}; }
  std::size_t count(int32_t key) const {
    if (key < 0 || key >= m_bins) throw std::out_of_range("key out of range");
    int32_t b[2];
    detail::check(clue_memcpy_d2h(m_ctx, b, m_off.get() + key, sizeof(b)));
    return static_cast<std::size_t>(b[1] - b[0]);
  }
  AssociationMapHost to_host() const {
    std::vector<int32_t> off(static_cast<std::size_t>(m_bins) + 1);
    std::vector<int32_t> idx(static_cast<std::size_t>(m_elems));
    detail::check(clue_memcpy_d2h(m_ctx, off.data(), m_off.get(), 4 * off.size()));
    if (!idx.empty())
      detail::check(clue_memcpy_d2h(m_ctx, idx.data(), m_idx.get(), 4 * idx.size()));
    return AssociationMapHost(std::move(off), std::move(idx));
  }

 private:
  void alloc_offsets(int32_t count) {
    void* off = nullptr;
    detail::check(clue_malloc(m_ctx, 4 * std::size_t(count), &off));
    m_off.reset(static_cast<int32_t*>(off));
    m_off.get_deleter().ctx = m_ctx;
  }
  clue_ctx* m_ctx = nullptr;
  std::unique_ptr<int32_t, detail::DeviceDeleter<int32_t>> m_off{nullptr, {nullptr}};
  std::unique_ptr<int32_t, detail::DeviceDeleter<int32_t>> m_idx{nullptr, {nullptr}};
  int32_t m_bins = 0, m_elems = 0;
};

// ---- Points ------------------------------------------------------------------------------
namespace detail {
template <class P>
concept pointer = std::is_pointer_v<P>;

inline void check_buffer(std::size_t have, std::size_t need) {
  if (have < need) throw std::invalid_argument("buffer too small for the number of points");
}
inline int32_t check_points(int32_t n, const char* who) {
  // detail/HostViewPartition.hpp:16-19, detail/DeviceViewPartition.hpp:13-19
  if (n <= 0)
    throw std::invalid_argument(std::string("Number of points passed to ") + who +
                                " constructor must be positive.");
  return n;
}
}  // namespace detail

// Host SoA (detail/HostViewPartition.hpp:15-21): Ndim coordinate columns, weights, and
// the cluster indexes, in one owned buffer or in caller buffers. The eight
// constructor forms are those of data_structures/PointsHost.hpp:59-137.
template <std::size_t Ndim, class T = float>
class PointsHost {
  static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>,
                "points are float or double");

 public:
  using value_type = T;
  using element_type = T;

  // Bytes of the one-buffer layout [x_0 .. x_{Ndim-1}][weight][cluster_index]
  // (detail/HostViewPartition.hpp:13-20).
  static std::size_t soa_bytes(int32_t n) {
    return ((Ndim + 1) * sizeof(T) + sizeof(int)) * static_cast<std::size_t>(n);
  }

  // Owned storage.
  PointsHost(Queue&, int32_t n)
      : m_n{detail::check_points(n, "PointsHost")}, m_own(soa_bytes(n)) {
    partition(m_own.data());
  }
  // One caller buffer holding the whole SoA, soa_bytes(n) long.
  PointsHost(Queue&, int32_t n, std::span<std::byte> buffer)
      : m_n{detail::check_points(n, "PointsHost")} {
    detail::check_buffer(buffer.size(), soa_bytes(n));
    partition(buffer.data());
  }
  // (Ndim+1) x n SoA input plus an n-int output (the Python binding's layout, Run.hpp:39).
  PointsHost(Queue& q, int32_t n, std::span<T> input, std::span<int> output)
      : PointsHost(q, n, input.data(), output.data()) {
    detail::check_buffer(input.size(), (Ndim + 1) * std::size_t(n));
    detail::check_buffer(output.size(), std::size_t(n));
  }
  // Ndim x n coordinates, n weights, n-int output.
  PointsHost(Queue& q, int32_t n, std::span<T> coordinates, std::span<T> weights,
             std::span<int> output)
      : PointsHost(q, n, coordinates.data(), weights.data(), output.data()) {
    detail::check_buffer(coordinates.size(), Ndim * std::size_t(n));
    detail::check_buffer(weights.size(), std::size_t(n));
    detail::check_buffer(output.size(), std::size_t(n));
  }
  // One range per coordinate, then the weights and the output.
  template <std::ranges::contiguous_range... R>
    requires(sizeof...(R) == Ndim + 2 && Ndim > 1)
  PointsHost(Queue&, int32_t n, R&&... buffers) : m_n{detail::check_points(n, "PointsHost")} {
    (detail::check_buffer(std::ranges::size(buffers), std::size_t(n)), ...);
    columns(std::ranges::data(buffers)...);
  }
  PointsHost(Queue&, int32_t n, T* input, int* output) : m_n{detail::check_points(n, "PointsHost")} {
    for (std::size_t d = 0; d < Ndim; ++d) m_coords[d] = input + d * n;
    m_weight = input + Ndim * static_cast<std::size_t>(n);
    m_cluster = output;
  }
  PointsHost(Queue&, int32_t n, T* coordinates, T* weights, int* output)
      : m_n{detail::check_points(n, "PointsHost")} {
    for (std::size_t d = 0; d < Ndim; ++d) m_coords[d] = coordinates + d * n;
    m_weight = weights;
    m_cluster = output;
  }
  // One pointer per coordinate, then the weights and the output.
  template <detail::pointer... P>
    requires(sizeof...(P) == Ndim + 2 && Ndim > 1)
  PointsHost(Queue&, int32_t n, P... buffers) : m_n{detail::check_points(n, "PointsHost")} {
    columns(buffers...);
  }
  PointsHost(const PointsHost&) = delete;
  PointsHost& operator=(const PointsHost&) = delete;
  PointsHost(PointsHost&&) = default;
  PointsHost& operator=(PointsHost&&) = default;

  int32_t size() const { return m_n; }
  std::span<T> coords(std::size_t dim) {
    if (dim >= Ndim) throw std::out_of_range("Dimension out of range in call to coords.");
    return {m_coords[dim], static_cast<std::size_t>(m_n)};
  }
  std::span<const T> coords(std::size_t dim) const {
    if (dim >= Ndim) throw std::out_of_range("Dimension out of range in call to coords.");
    return {m_coords[dim], static_cast<std::size_t>(m_n)};
  }
  std::span<T> weights() { return {m_weight, static_cast<std::size_t>(m_n)}; }
  std::span<const T> weights() const { return {m_weight, static_cast<std::size_t>(m_n)}; }
  std::span<const int> clusterIndexes() const { return {m_cluster, static_cast<std::size_t>(m_n)}; }
  std::span<int> clusterIndexes() { return {m_cluster, static_cast<std::size_t>(m_n)}; }
  bool clustered() const { return m_clustered; }
  void mark_clustered() { m_clustered = true; }
  // Point i as coordinates then weight (the argument of the metrics' call operator).
  Point<Ndim, T> operator[](std::size_t i) const {
    Point<Ndim, T> p{};
    for (std::size_t d = 0; d < Ndim; ++d) p[d] = m_coords[d][i];
    p[Ndim] = m_weight[i];
    return p;
  }

  std::size_t n_clusters() const {
    int32_t m = -1;
    for (int32_t i = 0; i < m_n; ++i) m = std::max(m, m_cluster[i]);
    return static_cast<std::size_t>(m + 1);
  }
  void set_density_uncertainty(std::span<T> u) { m_u = check_span(u).data(); }
  void set_sigma(std::size_t dim, std::span<T> s) {
    if (dim >= Ndim) throw std::out_of_range("Dimension out of range in call to set_sigma.");
    m_sigma[dim] = check_span(s).data();
  }
  void set_tags(std::span<std::uint32_t> tags) {
    if (tags.size() < static_cast<std::size_t>(m_n)) throw std::invalid_argument("tags too short");
    m_tags = tags.data();
  }
  void set_tags(std::vector<std::uint32_t>& tags) { set_tags(std::span<std::uint32_t>(tags)); }

  const T* coord_ptr(std::size_t d) const { return m_coords[d]; }
  const T* weight_ptr() const { return m_weight; }
  int* cluster_ptr() const { return m_cluster; }
  const T* sigma_ptr(std::size_t d) const { return m_sigma[d]; }
  const T* uncertainty_ptr() const { return m_u; }
  const std::uint32_t* tags_ptr() const { return m_tags; }

 private:
  void partition(std::byte* b) {
    const std::size_t nn = static_cast<std::size_t>(m_n);
    T* f = reinterpret_cast<T*>(b);
    for (std::size_t d = 0; d < Ndim; ++d) m_coords[d] = f + d * nn;
    m_weight = f + Ndim * nn;
    m_cluster = reinterpret_cast<int*>(f + (Ndim + 1) * nn);
  }
  template <class... P>
  void columns(P... p) {
    auto t = std::make_tuple(p...);
    [&]<std::size_t... D>(std::index_sequence<D...>) {
      ((m_coords[D] = std::get<D>(t)), ...);
    }(std::make_index_sequence<Ndim>{});
    m_weight = std::get<Ndim>(t);
    m_cluster = std::get<Ndim + 1>(t);
  }
  std::span<T> check_span(std::span<T> s) const {
    if (s.size() < static_cast<std::size_t>(m_n)) throw std::invalid_argument("array too short");
    return s;
  }
  int32_t m_n;
  std::vector<std::byte> m_own;
  std::array<T*, Ndim> m_coords{};
  T* m_weight = nullptr;
  int* m_cluster = nullptr;
  std::array<T*, Ndim> m_sigma{};
  T* m_u = nullptr;
  std::uint32_t* m_tags = nullptr;
  bool m_clustered = false;
};

template <std::size_t Ndim, class T>
T MahalanobisMetric<Ndim, T>::operator()(const PointsHost<Ndim, T>& points, std::size_t i,
                                         std::size_t j) const {
  return std::sqrt(detail::right_fold_sum<Ndim, T>([&](std::size_t k) {
    const T* s = points.sigma_ptr(k);
    if (!s) throw std::invalid_argument("sigma must be set in every dimension");
    const T diff = points.coord_ptr(k)[i] - points.coord_ptr(k)[j];
    return diff * diff / (s[i] * s[i] + s[j] * s[j]);
  }));
}

// Device SoA (detail/DeviceViewPartition.hpp:14-20): coordinates, weights, and the
// rho / nearest_higher / is_seed / cluster_index results; optional sigma, density
// uncertainty and tags. The seven constructor forms are those of
// data_structures/PointsDevice.hpp:51-122: owned storage, one caller device
// buffer with the whole layout, or caller device buffers for the inputs and the
// cluster indexes, in which case is_seed, rho and nearest_higher live in storage
// these points own (as the reference's partitionSoAView with an alloc buffer).
// delta, which the reference does not keep, is always owned here.
template <std::size_t Ndim, class T = float>
class PointsDevice {
  static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>,
                "points are float or double");

 public:
  using value_type = T;
  using element_type = T;

  struct View {
    std::array<T*, Ndim> m_coords;
    T* m_weight;
    int32_t* m_cluster_index;
    int32_t* m_is_seed;
    T* m_rho;
    int32_t* m_nearest_higher;
    T* m_delta;
    int32_t m_n;
    std::span<int32_t> cluster_index() const { return {m_cluster_index, std::size_t(m_n)}; }
    std::span<int32_t> is_seed() const { return {m_is_seed, std::size_t(m_n)}; }
    std::span<T> rho() const { return {m_rho, std::size_t(m_n)}; }
    std::span<int32_t> nearest_higher() const { return {m_nearest_higher, std::size_t(m_n)}; }
    std::span<T> delta() const { return {m_delta, std::size_t(m_n)}; }
    int32_t size() const { return m_n; }
  };

  // Bytes of the one-buffer layout [x_0 .. x_{Ndim-1}][weight][cluster_index][is_seed]
  // [rho][nearest_higher] (detail/DeviceViewPartition.hpp:13-40).
  static std::size_t soa_bytes(int32_t n) {
    return ((Ndim + 2) * sizeof(T) + 3 * sizeof(int32_t)) * static_cast<std::size_t>(n);
  }

  PointsDevice(Queue& q, int32_t n) : m_q{q}, m_n{detail::check_points(n, "PointsDevice")} {
    m_block = alloc<std::byte>(soa_bytes(n));
    partition(m_block.get());
    own_delta();
  }
  PointsDevice(Queue& q, int32_t n, std::span<std::byte> buffer)
      : m_q{q}, m_n{detail::check_points(n, "PointsDevice")} {
    detail::check_buffer(buffer.size(), soa_bytes(n));
    partition(buffer.data());
    own_delta();
  }
  PointsDevice(Queue& q, int32_t n, std::span<T> input, std::span<int> output)
      : PointsDevice(q, n, input.data(), output.data()) {
    detail::check_buffer(input.size(), (Ndim + 1) * std::size_t(n));
    detail::check_buffer(output.size(), std::size_t(n));
  }
  PointsDevice(Queue& q, int32_t n, std::span<T> coordinates, std::span<T> weights,
               std::span<int> output)
      : PointsDevice(q, n, coordinates.data(), weights.data(), output.data()) {
    detail::check_buffer(coordinates.size(), Ndim * std::size_t(n));
    detail::check_buffer(weights.size(), std::size_t(n));
    detail::check_buffer(output.size(), std::size_t(n));
  }
  PointsDevice(Queue& q, int32_t n, T* input, int* output)
      : m_q{q}, m_n{detail::check_points(n, "PointsDevice")} {
    for (std::size_t d = 0; d < Ndim; ++d) m_view.m_coords[d] = input + d * std::size_t(n);
    m_view.m_weight = input + Ndim * std::size_t(n);
    m_view.m_cluster_index = output;
    own_results();
  }
  PointsDevice(Queue& q, int32_t n, T* coordinates, T* weights, int* output)
      : m_q{q}, m_n{detail::check_points(n, "PointsDevice")} {
    for (std::size_t d = 0; d < Ndim; ++d) m_view.m_coords[d] = coordinates + d * std::size_t(n);
    m_view.m_weight = weights;
    m_view.m_cluster_index = output;
    own_results();
  }
  // One device pointer per coordinate, then the weights and the cluster indexes.
  template <detail::pointer... P>
    requires(sizeof...(P) == Ndim + 2 && Ndim > 1)
  PointsDevice(Queue& q, int32_t n, P... buffers)
      : m_q{q}, m_n{detail::check_points(n, "PointsDevice")} {
    auto t = std::make_tuple(buffers...);
    [&]<std::size_t... D>(std::index_sequence<D...>) {
      ((m_view.m_coords[D] = std::get<D>(t)), ...);
    }(std::make_index_sequence<Ndim>{});
    m_view.m_weight = std::get<Ndim>(t);
    m_view.m_cluster_index = std::get<Ndim + 1>(t);
    own_results();
  }
  PointsDevice(const PointsDevice&) = delete;
  PointsDevice& operator=(const PointsDevice&) = delete;
  PointsDevice(PointsDevice&&) = default;
  PointsDevice& operator=(PointsDevice&&) = default;

  int32_t size() const { return m_n; }
  const View& view() const { return m_view; }
  std::span<T> coords(std::size_t d) const {
    if (d >= Ndim) throw std::out_of_range("Dimension out of range in call to coords.");
    return {m_view.m_coords[d], std::size_t(m_n)};
  }
  std::span<T> weights() const { return {m_view.m_weight, std::size_t(m_n)}; }
  std::span<int32_t> clusterIndexes() const { return m_view.cluster_index(); }

  void set_sigma(std::size_t dim, std::span<T> device_sigma) {
    if (dim >= Ndim) throw std::out_of_range("Dimension out of range in call to set_sigma.");
    m_sigma[dim] = device_sigma.data();
  }
  void set_density_uncertainty(std::span<T> device_u) { m_u = device_u.data(); }
  void set_tags(std::span<std::uint32_t> device_tags) { m_tags = device_tags.data(); }

  // Copies host sigma / uncertainty / tags into device buffers owned by these points.
  void upload_optionals(const PointsHost<Ndim, T>& h) {
    const std::size_t nn = std::size_t(m_n);
    if (h.sigma_ptr(0)) {
      m_sig_own = alloc<T>(Ndim * nn);
      for (std::size_t d = 0; d < Ndim; ++d) {
        if (!h.sigma_ptr(d)) throw std::invalid_argument("sigma must be set in every dimension");
        detail::check(clue_memcpy_h2d(m_q.ctx(), m_sig_own.get() + d * nn, h.sigma_ptr(d), sizeof(T) * nn));
        m_sigma[d] = m_sig_own.get() + d * nn;
      }
    }
    if (h.uncertainty_ptr()) {
      m_u_own = alloc<T>(nn);
      detail::check(clue_memcpy_h2d(m_q.ctx(), m_u_own.get(), h.uncertainty_ptr(), sizeof(T) * nn));
      m_u = m_u_own.get();
    }
    if (h.tags_ptr()) {
      m_tags_own = alloc<std::uint32_t>(nn);
      detail::check(
          clue_memcpy_h2d(m_q.ctx(), m_tags_own.get(), h.tags_ptr(), sizeof(std::uint32_t) * nn));
      m_tags = m_tags_own.get();
    }
  }

  typename detail::CApi<T>::Points c_points(int32_t offset = 0, int32_t n = -1) const {
    typename detail::CApi<T>::Points p{};
    p.n = n < 0 ? m_n : n;
    for (std::size_t d = 0; d < Ndim; ++d) {
      p.coords[d] = m_view.m_coords[d] + offset;
      p.sigma[d] = m_sigma[d] ? m_sigma[d] + offset : nullptr;
    }
    p.weight = m_view.m_weight + offset;
    p.density_uncertainty = m_u ? m_u + offset : nullptr;
    p.tags = m_tags ? m_tags + offset : nullptr;
    return p;
  }
  typename detail::CApi<T>::Results c_results() const {
    typename detail::CApi<T>::Results r{};
    r.cluster_index = m_view.m_cluster_index;
    r.is_seed = m_view.m_is_seed;
    r.rho = m_view.m_rho;
    r.nearest_higher = m_view.m_nearest_higher;
    r.delta = m_view.m_delta;
    return r;
  }
  void mark_clustered() { m_clustered = true; }
  bool clustered() const { return m_clustered; }

  std::size_t n_clusters() const {
    std::vector<int32_t> h(static_cast<std::size_t>(m_n));
    detail::check(clue_memcpy_d2h(m_q.ctx(), h.data(), m_view.m_cluster_index, 4 * h.size()));
    int32_t m = -1;
    for (int32_t v : h) m = std::max(m, v);
    return static_cast<std::size_t>(m + 1);
  }

 private:
  template <class U>
  std::unique_ptr<U, detail::DeviceDeleter<U>> alloc(std::size_t count) {
    void* p = nullptr;
    detail::check(clue_malloc(m_q.ctx(), count * sizeof(U), &p));
    return std::unique_ptr<U, detail::DeviceDeleter<U>>(static_cast<U*>(p),
                                                        detail::DeviceDeleter<U>{m_q.ctx()});
  }
  // [x_0 .. x_{Ndim-1}][weight][cluster_index][is_seed][rho][nearest_higher] from b.
  void partition(std::byte* b) {
    const std::size_t nn = static_cast<std::size_t>(m_n);
    T* f = reinterpret_cast<T*>(b);
    for (std::size_t d = 0; d < Ndim; ++d) m_view.m_coords[d] = f + d * nn;
    m_view.m_weight = f + Ndim * nn;
    std::byte* r = b + (Ndim + 1) * sizeof(T) * nn;
    m_view.m_cluster_index = reinterpret_cast<int32_t*>(r);
    m_view.m_is_seed = reinterpret_cast<int32_t*>(r + sizeof(int32_t) * nn);
    m_view.m_rho = reinterpret_cast<T*>(r + 2 * sizeof(int32_t) * nn);
    m_view.m_nearest_higher = reinterpret_cast<int32_t*>(r + 2 * sizeof(int32_t) * nn + sizeof(T) * nn);
    m_view.m_n = m_n;
  }
  void own_delta() {
    m_delta_own = alloc<T>(std::size_t(m_n));
    m_view.m_delta = m_delta_own.get();
  }
  // is_seed, rho, nearest_higher (and delta) when the caller gives only inputs and
  // cluster indexes.
  void own_results() {
    const std::size_t nn = static_cast<std::size_t>(m_n);
    m_block = alloc<std::byte>((2 * sizeof(T) + 2 * sizeof(int32_t)) * nn);
    T* f = reinterpret_cast<T*>(m_block.get());
    m_view.m_rho = f;
    m_view.m_delta = f + nn;
    m_view.m_is_seed = reinterpret_cast<int32_t*>(f + 2 * nn);
    m_view.m_nearest_higher = m_view.m_is_seed + nn;
    m_view.m_n = m_n;
  }
  Queue m_q;
  int32_t m_n;
  std::unique_ptr<std::byte, detail::DeviceDeleter<std::byte>> m_block{nullptr, {nullptr}};
  std::unique_ptr<T, detail::DeviceDeleter<T>> m_delta_own{nullptr, {nullptr}};
  std::unique_ptr<T, detail::DeviceDeleter<T>> m_sig_own, m_u_own;
  std::unique_ptr<std::uint32_t, detail::DeviceDeleter<std::uint32_t>> m_tags_own;
  View m_view{};
  std::array<T*, Ndim> m_sigma{};
  T* m_u = nullptr;
  std::uint32_t* m_tags = nullptr;
  bool m_clustered = false;
};

// copyToDevice / copyToHost (data_structures/detail/PointsConversion.hpp:21-111).
template <std::size_t Ndim, class T>
void copyToDevice(Queue& q, PointsDevice<Ndim, T>& d, const PointsHost<Ndim, T>& h) {
  if (d.size() != h.size()) throw std::invalid_argument("host and device sizes differ");
  const std::size_t nn = std::size_t(h.size());
  for (std::size_t k = 0; k < Ndim; ++k)
    detail::check(clue_memcpy_h2d(q.ctx(), d.view().m_coords[k], h.coord_ptr(k), sizeof(T) * nn));
  detail::check(clue_memcpy_h2d(q.ctx(), d.view().m_weight, h.weight_ptr(), sizeof(T) * nn));
  d.upload_optionals(h);
}
template <std::size_t Ndim, class T>
void copyToHost(Queue& q, PointsHost<Ndim, T>& h, const PointsDevice<Ndim, T>& d) {
  if (d.size() != h.size()) throw std::invalid_argument("host and device sizes differ");
  detail::check(clue_memcpy_d2h(q.ctx(), h.cluster_ptr(), d.view().m_cluster_index,
                                sizeof(int32_t) * std::size_t(h.size())));
  h.mark_clustered();
}

inline AssociationMapHost get_clusters(std::span<const int32_t> cluster_index) {
  return AssociationMapHost::from_associations(cluster_index);
}
template <std::size_t Ndim, class T>
AssociationMapHost get_clusters(const PointsHost<Ndim, T>& h) {
  return get_clusters(h.clusterIndexes());
}
// get_clusters(queue, d_points) (utils/detail/get_clusters.hpp:44-48): built on the device.
template <std::size_t Ndim, class T>
AssociationMapDevice get_clusters(Queue& q, const PointsDevice<Ndim, T>& d) {
  return AssociationMapDevice(q.ctx(), d.clusterIndexes().data(), d.size());
}

// ---- silhouette (utils/scores.hpp, detail/scores.hpp:25-99) -------------------------------
// Score of one clustered point: (b - a) / max(a, b), with a the mean Euclidean
// distance to the other members of its cluster and b the smallest mean distance
// to the members of another cluster.
template <std::size_t Ndim, class T>
T silhouette(const AssociationMapHost& clusters, const PointsHost<Ndim, T>& h, int point) {
  auto dist = [&](int i, int j) {
    T s = T(0);
    for (std::size_t d = 0; d < Ndim; ++d) {
      const T dx = h.coords(d)[i] - h.coords(d)[j];
      s += dx * dx;
    }
    return std::sqrt(s);
  };
  const int c = h.clusterIndexes()[point];
  T a = T(0);
  for (int j : clusters[c])
    if (j != point) a += dist(point, j);
  a /= static_cast<T>(clusters.count(c) - 1);
  T b = std::numeric_limits<T>::max();
  for (int k = 0; k < static_cast<int>(clusters.size()); ++k) {
    if (k == c) continue;
    T m = T(0);
    for (int j : clusters[k]) m += dist(point, j);
    b = std::min(b, m / static_cast<T>(clusters.count(k)));
  }
  return (b - a) / std::max(a, b);
}
template <std::size_t Ndim, class T>
T silhouette(const PointsHost<Ndim, T>& h, int point) {
  return silhouette(get_clusters(h), h, point);
}
// Mean over the clustered points of clusters with at least two members.
template <std::size_t Ndim, class T>
T silhouette(const PointsHost<Ndim, T>& h) {
  const AssociationMapHost clusters = get_clusters(h);
  T sum = T(0);
  std::size_t n = 0;
  for (int i = 0; i < h.size(); ++i) {
    const int c = h.clusterIndexes()[i];
    if (c < 0 || clusters.count(c) < 2) continue;
    sum += silhouette(clusters, h, i);
    ++n;
  }
  return sum / static_cast<T>(n);
}

// ---- Clusterer (core/Clusterer.hpp:37-312) -----------------------------------------------
template <std::size_t Ndim, class DataType = float>
class Clusterer {
  static_assert(std::is_same_v<DataType, float> || std::is_same_v<DataType, double>,
                "the device path computes in value_type or double");
  static_assert(Ndim >= 1 && Ndim <= 10, "This library only works up to 10 dimensions");

 public:
  using value_type = DataType;

  Clusterer(value_type dc, value_type rhoc, std::optional<value_type> dm = std::nullopt,
            std::optional<value_type> seed_dc = std::nullopt)
      : m_q{std::make_shared<Queue>(0)} {
    setParameters(dc, rhoc, dm, seed_dc);
  }
  Clusterer(Queue& q, value_type dc, value_type rhoc, std::optional<value_type> dm = std::nullopt,
            std::optional<value_type> seed_dc = std::nullopt)
      : m_q{std::make_shared<Queue>(q)} {
    setParameters(dc, rhoc, dm, seed_dc);
  }

  // core/detail/Clusterer.hpp:75-89
  void setParameters(value_type dc, value_type rhoc, std::optional<value_type> dm = std::nullopt,
                     std::optional<value_type> seed_dc = std::nullopt) {
    m_dc = dc;
    m_rhoc = rhoc;
    m_dm = dm.value_or(dc);
    m_seed_dc = seed_dc.value_or(dc);
    if (!(m_dc > 0) || !(m_rhoc >= 0) || !(m_dm > 0) || !(m_seed_dc > 0))
      throw std::invalid_argument("Invalid clustering parameters. The parameters must be positive.");
  }

  template <std::ranges::contiguous_range R>
    requires std::integral<std::ranges::range_value_t<R>>
  void setWrappedCoordinates(const R& wrapped) {
    std::size_t d = 0;
    for (auto w : wrapped) {
      if (d == Ndim) break;
      m_wrapped[d++] = static_cast<uint8_t>(w);
    }
  }
  template <std::integral... A>
  void setWrappedCoordinates(A... wrapped) {
    m_wrapped = {static_cast<uint8_t>(wrapped)...};
  }

  template <concepts::distance_metric Metric = EuclideanMetric<Ndim, value_type>,
            concepts::convolutional_kernel Kernel = FlatKernel<value_type>>
  void make_clusters(Queue& q, PointsHost<Ndim, value_type>& h, const Metric& metric = Metric{},
                     const Kernel& kernel = Kernel{value_type(.5)}) {
    PointsDevice<Ndim, value_type> d(q, h.size());
    make_clusters(q, h, d, metric, kernel);
  }
  template <concepts::distance_metric Metric = EuclideanMetric<Ndim, value_type>,
            concepts::convolutional_kernel Kernel = FlatKernel<value_type>>
  void make_clusters(PointsHost<Ndim, value_type>& h, const Metric& metric = Metric{},
                     const Kernel& kernel = Kernel{value_type(.5)}) {
    Queue q(0);
    make_clusters(q, h, metric, kernel);
  }
  template <concepts::distance_metric Metric = EuclideanMetric<Ndim, value_type>,
            concepts::convolutional_kernel Kernel = FlatKernel<value_type>>
  void make_clusters(Queue& q, PointsHost<Ndim, value_type>& h, PointsDevice<Ndim, value_type>& d,
                     const Metric& metric = Metric{}, const Kernel& kernel = Kernel{value_type(.5)}) {
    copyToDevice(q, d, h);
    make_clusters(q, d, metric, kernel);
    copyToHost(q, h, d);
  }
  // Device-resident per event (core/detail/Clusterer.hpp:143-151).
  template <concepts::distance_metric Metric = EuclideanMetric<Ndim, value_type>,
            concepts::convolutional_kernel Kernel = FlatKernel<value_type>>
  void make_clusters(Queue& q, PointsDevice<Ndim, value_type>& d, const Metric& metric = Metric{},
                     const Kernel& kernel = Kernel{value_type(.5)}) {
    auto p = params(metric, kernel);
    auto pts = d.c_points();
    auto res = d.c_results();
    detail::check(detail::CApi<value_type>::run(q.ctx(), &p, &pts, &res));
    m_last = q;
    m_event_of_cluster.clear();
    m_eoc_dev.reset();
    d.mark_clustered();
  }
  template <concepts::distance_metric Metric = EuclideanMetric<Ndim, value_type>,
            concepts::convolutional_kernel Kernel = FlatKernel<value_type>>
  void make_clusters(Queue& q, PointsHost<Ndim, value_type>& h, PointsDevice<Ndim, value_type>& d,
                     std::span<const uint32_t> sizes, const Metric& metric = Metric{},
                     const Kernel& kernel = Kernel{value_type(.5)}) {
    copyToDevice(q, d, h);
    make_clusters(q, d, sizes, metric, kernel);
    copyToHost(q, h, d);
  }
  // Batched (core/detail/Clusterer.hpp:171-183, 286-350); every event is tiled alone.
  template <concepts::distance_metric Metric = EuclideanMetric<Ndim, value_type>,
            concepts::convolutional_kernel Kernel = FlatKernel<value_type>>
  void make_clusters(Queue& q, PointsDevice<Ndim, value_type>& d, std::span<const uint32_t> sizes,
                     const Metric& metric = Metric{}, const Kernel& kernel = Kernel{value_type(.5)}) {
    auto p = params(metric, kernel);
    auto pts = d.c_points();
    auto res = d.c_results();
    void* eoc = nullptr;
    detail::check(clue_malloc(q.ctx(), 4 * std::size_t(d.size()), &eoc));
    m_eoc_dev.reset(static_cast<int32_t*>(eoc));
    m_eoc_dev.get_deleter().ctx = q.ctx();
    int64_t nclusters = 0;
    clue_status st = detail::CApi<value_type>::run_batched(
        q.ctx(), &p, &pts, sizes.data(), static_cast<int32_t>(sizes.size()), &res,
        m_eoc_dev.get(), &nclusters);
    if (st == CLUE_OK) {
      m_event_of_cluster.assign(std::size_t(nclusters), 0);
      if (nclusters > 0)
        st = clue_memcpy_d2h(q.ctx(), m_event_of_cluster.data(), eoc, 4 * std::size_t(nclusters));
    }
    detail::check(st);
    m_last = q;
    d.mark_clustered();
  }

  // getSeeds (core/detail/Clusterer.hpp:198-203): host copy of the seed indices.
  std::vector<int32_t> getSeeds() const {
    if (!m_last) throw std::runtime_error("Seeds are not available. Please run make_clusters first.");
    int64_t n = 0;
    detail::check(clue_get_n_clusters(m_last->ctx(), &n));
    std::vector<int32_t> s(std::size_t(std::max<int64_t>(n, 0)));
    detail::check(clue_get_seeds(m_last->ctx(), s.data(), n, &n));
    return s;
  }
  // getClusters (core/detail/Clusterer.hpp:205-217): cluster -> points, on the host
  // for host points and on the device for device points.
  AssociationMapHost getClusters(const PointsHost<Ndim, value_type>& h) const { return get_clusters(h); }
  AssociationMapDevice getClusters(Queue& q, const PointsDevice<Ndim, value_type>& d) const {
    return get_clusters(q, d);
  }
  // getSampleAssociations (core/detail/Clusterer.hpp:219-241): event -> clusters of the
  // last batched run, on the host or on the device.
  AssociationMapHost getSampleAssociations(Queue&, const PointsHost<Ndim, value_type>&) const {
    return AssociationMapHost::from_associations(m_event_of_cluster);
  }
  AssociationMapDevice getSampleAssociations(Queue& q, const PointsDevice<Ndim, value_type>&) const {
    if (!m_eoc_dev) throw std::runtime_error("No batched run to associate. Run make_clusters first.");
    return AssociationMapDevice(q.ctx(), m_eoc_dev.get(),
                                static_cast<int32_t>(m_event_of_cluster.size()));
  }

 private:
  template <class Metric, class Kernel>
  typename detail::CApi<value_type>::Params params(const Metric& metric,
                                                   const Kernel& kernel) const {
    typename detail::CApi<value_type>::Params p{};
    p.ndim = static_cast<int32_t>(Ndim);
    p.dc = m_dc;
    p.rhoc = m_rhoc;
    p.dm = m_dm;
    p.seed_dc = m_seed_dc;
    for (std::size_t d = 0; d < Ndim; ++d) p.wrapped[d] = m_wrapped[d];
    metric.fill(p);
    kernel.fill(p);
    return p;
  }

  std::shared_ptr<Queue> m_q;
  std::optional<Queue> m_last;
  value_type m_dc{}, m_rhoc{}, m_dm{}, m_seed_dc{};
  std::array<uint8_t, Ndim> m_wrapped{};
  std::vector<int32_t> m_event_of_cluster;
  std::unique_ptr<int32_t, detail::DeviceDeleter<int32_t>> m_eoc_dev{nullptr, {nullptr}};
};

// read_csv (utils/detail/read_csv.hpp:22-104): header line, then Ndim coordinates and
// the weight per row.
template <std::size_t Ndim, class T = float>
PointsHost<Ndim, T> read_csv(Queue& q, const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("Could not open file " + path);
  std::string line;
  std::getline(f, line);
  std::vector<std::array<T, Ndim + 1>> rows;
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    std::stringstream ss(line);
    std::array<T, Ndim + 1> r{};
    std::string cell;
    for (std::size_t k = 0; k < Ndim + 1; ++k) {
      if (!std::getline(ss, cell, ',')) throw std::runtime_error("Malformed line in " + path);
      r[k] = detail::parse_floating<T>(cell);
    }
    rows.push_back(r);
  }
  PointsHost<Ndim, T> h(q, static_cast<int32_t>(rows.size()));
  for (std::size_t i = 0; i < rows.size(); ++i) {
    for (std::size_t d = 0; d < Ndim; ++d) h.coords(d)[i] = rows[i][d];
    h.weights()[i] = rows[i][Ndim];
  }
  return h;
}

// read_output (utils/detail/read_csv.hpp:68-104): a clustering result file, one row
// per point with Ndim coordinates, the weight and the cluster id after a header
// line; the points come back marked clustered.
template <std::size_t Ndim, class T = float>
PointsHost<Ndim, T> read_output(Queue& q, const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("Could not open file: " + path);
  std::string line;
  std::getline(f, line);
  std::vector<std::array<T, Ndim + 1>> rows;
  std::vector<int32_t> ids;
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    std::stringstream ss(line);
    std::array<T, Ndim + 1> r{};
    std::string cell;
    for (std::size_t k = 0; k < Ndim + 1; ++k) {
      if (!std::getline(ss, cell, ',')) throw std::runtime_error("Malformed line in " + path);
      r[k] = detail::parse_floating<T>(cell);
    }
    if (!std::getline(ss, cell, ',')) throw std::runtime_error("Malformed line in " + path);
    rows.push_back(r);
    ids.push_back(static_cast<int32_t>(std::stoi(cell)));
  }
  PointsHost<Ndim, T> h(q, static_cast<int32_t>(rows.size()));
  for (std::size_t i = 0; i < rows.size(); ++i) {
    for (std::size_t d = 0; d < Ndim; ++d) h.coords(d)[i] = rows[i][d];
    h.weights()[i] = rows[i][Ndim];
    h.clusterIndexes()[i] = ids[i];
  }
  h.mark_clustered();
  return h;
}

}  // namespace clue
