// CLUE oracle: clean-room CPU restatement of the reference's serial semantics.
// TEST INFRASTRUCTURE ONLY (see clue_oracle.h): never linked into the product.
//
// Build: see oracle/Makefile (-O2 -ffp-contract=off, no fast-math).
#include "clue_oracle.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// exp: Cody-Waite argument reduction x = k ln2 + r, |r| <= ln2/2, then
// p(r) = 1 + r + c2 r^2 + ... + c6 r^6, the minimax fit of e^r in relative
// error on [-ln2/2, ln2/2] (derived by oracle/exp_minimax.py, coefficients
// rounded to float, written as hex literals so that every compiler sees the
// same bits), in Horner form with fused multiply-adds, then scaling by 2^k.
// Within 0.91 ulp of exp over every float of the core range.
// ---------------------------------------------------------------------------
float expf_cw(float x) {
  if (x != x) return x;
  if (x > 0x1.62e42ep+6f) return INFINITY;  // > ln(FLT_MAX)
  if (x < -0x1.9fe368p+6f) return 0.0f;     // e^x < 2^-150 rounds to +0
  const float log2e = 0x1.715476p+0f;
  const float ln2_hi = 0x1.62e400p-1f;  // 15 significant bits: k*ln2_hi is exact
  const float ln2_lo = 0x1.7f7d1cp-20f;
  const float k = std::rint(x * log2e);
  float r = std::fma(-k, ln2_hi, x);
  r = std::fma(-k, ln2_lo, r);
  float p = 0x1.6a244cp-10f;  // minimax c6..c2 (oracle/exp_minimax.py)
  p = std::fma(p, r, 0x1.1239d4p-7f);
  p = std::fma(p, r, 0x1.5558f2p-5f);
  p = std::fma(p, r, 0x1.555492p-3f);
  p = std::fma(p, r, 0x1.fffffcp-2f);
  p = std::fma(p, r, 1.0f);
  p = std::fma(p, r, 1.0f);
  return std::ldexp(p, static_cast<int>(k));
}

// Double precision: the same method with a 29-bit ln2_hi (k*ln2_hi is exact for
// |k| <= 1075) and the degree-13 Taylor polynomial (remainder below 2^-57 on
// |r| <= ln2/2). Coefficients are 1/n! rounded to double.
double expd_cw(double x) {
  if (x != x) return x;
  if (x > 0x1.62e42fefa39efp+9) return INFINITY;  // > ln(DBL_MAX)
  if (x < -0x1.74910d52d3052p+9) return 0.0;      // e^x < 2^-1075 rounds to +0
  const double log2e = 0x1.71547652b82fep+0;
  const double ln2_hi = 0x1.62e42ff000000p-1;
  const double ln2_lo = -0x1.718432a1b0e26p-35;
  const double k = std::rint(x * log2e);
  double r = std::fma(-k, ln2_hi, x);
  r = std::fma(-k, ln2_lo, r);
  double p = 0x1.6124613a86d09p-33;       // 1/13!
  p = std::fma(p, r, 0x1.1eed8eff8d898p-29);  // 1/12!
  p = std::fma(p, r, 0x1.ae64567f544e4p-26);  // 1/11!
  p = std::fma(p, r, 0x1.27e4fb7789f5cp-22);  // 1/10!
  p = std::fma(p, r, 0x1.71de3a556c734p-19);  // 1/9!
  p = std::fma(p, r, 0x1.a01a01a01a01ap-16);  // 1/8!
  p = std::fma(p, r, 0x1.a01a01a01a01ap-13);  // 1/7!
  p = std::fma(p, r, 0x1.6c16c16c16c17p-10);  // 1/6!
  p = std::fma(p, r, 0x1.1111111111111p-7);   // 1/5!
  p = std::fma(p, r, 0x1.5555555555555p-5);   // 1/4!
  p = std::fma(p, r, 0x1.5555555555555p-3);   // 1/3!
  p = std::fma(p, r, 0x1.0000000000000p-1);   // 1/2!
  p = std::fma(p, r, 1.0);
  p = std::fma(p, r, 1.0);
  return std::ldexp(p, static_cast<int>(k));
}

template <class T>
T exp_t(T x);
template <>
float exp_t<float>(float x) {
  return expf_cw(x);
}
template <>
double exp_t<double>(double x) {
  return expd_cw(x);
}

// static_cast<int>(float) as the x86 reference executes it (cvttss2si): NaN and
// out-of-range values give INT_MIN. TilesView.hpp:42-45 and :98 rely on it.
template <class T>
inline int cvt_trunc(T q) {
  if (!(q >= static_cast<T>(-2147483648.0)) || !(q < static_cast<T>(2147483648.0)))
    return INT_MIN;
  return static_cast<int>(q);
}

inline int64_t ipow(int64_t b, int e) {
  int64_t r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

template <class T>
struct Geo {
  int D = 0;
  int nper = 1;
  int ntiles = 1;
  T mn[10] = {};
  T mx[10] = {};
  T side[10] = {};
  int wrapped[10] = {};
  int64_t stride[10] = {};  // nper^(D-1-d)
};

// SetupTiles.hpp:30-34 and ComputeTiles.hpp:17-36.
template <class T>
void make_geometry(Geo<T>& g, const oracle_params* p, int32_t n_for_tiles, int32_t n,
                   const T* coords, int32_t coord_stride) {
  const int D = p->ndim;
  g.D = D;
  const int ppt = p->points_per_tile;
  const int64_t ntiles0 = (static_cast<int64_t>(n_for_tiles) + ppt - 1) / ppt;
  int nper = 1;
  while (ipow(nper, D) < ntiles0) ++nper;
  g.nper = nper;
  g.ntiles = static_cast<int>(ipow(nper, D));
  for (int d = 0; d < D; ++d) {
    T mx = std::numeric_limits<T>::lowest();
    T mn = std::numeric_limits<T>::max();
    const T* c = coords + static_cast<size_t>(d) * coord_stride;
    for (int32_t i = 0; i < n; ++i) {
      mx = std::max(mx, c[i]);
      mn = std::min(mn, c[i]);
    }
    g.mn[d] = mn;
    g.mx[d] = mx;
    g.side[d] = (mx - mn) / static_cast<T>(nper);
    g.wrapped[d] = p->wrapped[d] ? 1 : 0;
    g.stride[d] = ipow(nper, D - 1 - d);
  }
}

// TilesView.hpp:96-104.
template <class T>
inline T normalize(const Geo<T>& g, T c, int d) {
  const T range = g.mx[d] - g.mn[d];
  T r = c - static_cast<T>(cvt_trunc<T>(c / range)) * range;
  if (r >= g.mx[d])
    r -= range;
  else if (r < g.mn[d])
    r += range;
  return r;
}

// TilesView.hpp:38-52.
template <class T>
inline int get_bin(const Geo<T>& g, T c, int d) {
  int b;
  if (g.wrapped[d])
    b = cvt_trunc<T>((normalize(g, c, d) - g.mn[d]) / g.side[d]);
  else
    b = cvt_trunc<T>((c - g.mn[d]) / g.side[d]);
  b = std::min(b, g.nper - 1);
  b = std::max(b, 0);
  return b;
}

// Input view of one event: coordinate d of local point i is c[d*cs + i].
template <class T>
struct View {
  int D;
  int32_t n;        // points in this event
  int32_t g0;       // global index of local point 0 (ids and default tags)
  const T* c;       // coords base (already offset to the event)
  int32_t cs;       // coordinate stride (total points)
  const T* w;
  const T* sigma;   // NULL or same layout as c
  const T* u;
  const uint32_t* tags;
  T coord(int d, int32_t i) const { return c[static_cast<size_t>(d) * cs + i]; }
  T sig(int d, int32_t i) const { return sigma ? sigma[static_cast<size_t>(d) * cs + i] : T(0); }
  uint64_t tag(int32_t i) const {
    return tags ? static_cast<uint64_t>(tags[i]) : static_cast<uint64_t>(g0 + i);
  }
};

// Tile CSR over one event: ascending index within a tile (AssociationMap.hpp:359-383).
struct CSR {
  std::vector<int32_t> off;
  std::vector<int32_t> idx;
};

template <class T>
void build_tiles(const Geo<T>& g, const View<T>& v, CSR& csr) {
  std::vector<int32_t> tile(v.n);
  for (int32_t i = 0; i < v.n; ++i) {
    // TilesView.hpp:54-64: sum over dims of nper^(D-1-d) * bin_d.
    int64_t t = 0;
    for (int d = 0; d < g.D; ++d) t += g.stride[d] * get_bin(g, v.coord(d, i), d);
    tile[i] = static_cast<int32_t>(t);
  }
  csr.off.assign(g.ntiles + 1, 0);
  for (int32_t i = 0; i < v.n; ++i) csr.off[tile[i] + 1]++;
  for (int t = 0; t < g.ntiles; ++t) csr.off[t + 1] += csr.off[t];
  std::vector<int32_t> fill(csr.off.begin(), csr.off.end() - 1);
  csr.idx.assign(v.n, 0);
  for (int32_t i = 0; i < v.n; ++i) csr.idx[fill[tile[i]]++] = i;
}

struct Metric {
  int kind;
  int D;
};

// DistanceMetrics.hpp: every metric takes (lhs = point i, rhs = point j) and folds
// over dimensions with meta::accumulate, a right fold t0 + (t1 + (... + (t_{D-1} + 0))).
template <class T>
inline T distance(const oracle_params* p, const View<T>& v, int32_t i, int32_t j) {
  const int D = v.D;
  switch (p->metric) {
    case ORACLE_MEUCLID: {  // :95-114
      T s = T(0);
      for (int d = D - 1; d >= 0; --d) {
        const T dx = v.coord(d, i) - v.coord(d, j);
        s = dx * dx + s;
      }
      return std::sqrt(s);
    }
    case ORACLE_MWEUCLID: {  // :122-163, w * (l - r) * (l - r)
      T s = T(0);
      for (int d = D - 1; d >= 0; --d) {
        const T dx = v.coord(d, i) - v.coord(d, j);
        const T wd = static_cast<T>(p->mparam[d]);
        s = wd * dx * dx + s;
      }
      return std::sqrt(s);
    }
    case ORACLE_MPERIODIC: {  // :175-221, min(|d|, P - |d|) squared
      T s = T(0);
      for (int d = D - 1; d >= 0; --d) {
        const T a = std::fabs(v.coord(d, i) - v.coord(d, j));
        const T per = static_cast<T>(p->mparam[d]);
        const T pd = std::min(a, per - a);
        s = pd * pd + s;
      }
      return std::sqrt(s);
    }
    case ORACLE_MMANHATTAN: {  // :229-246
      T s = T(0);
      for (int d = D - 1; d >= 0; --d) s = std::fabs(v.coord(d, i) - v.coord(d, j)) + s;
      return s;
    }
    case ORACLE_MCHEBYSHEV: {  // :254-271, meta::maximum folds std::max from 0
      T m = T(0);
      for (int d = 0; d < D; ++d) m = std::max(m, std::fabs(v.coord(d, i) - v.coord(d, j)));
      return m;
    }
    case ORACLE_MWCHEBYSHEV: {  // :279-318
      T m = T(0);
      for (int d = 0; d < D; ++d) {
        const T wd = static_cast<T>(p->mparam[d]);
        m = std::max(m, wd * std::fabs(v.coord(d, i) - v.coord(d, j)));
      }
      return m;
    }
    case ORACLE_MMAHALANOBIS: {  // :64-87, (d*d) / (si*si + sj*sj)
      T s = T(0);
      for (int d = D - 1; d >= 0; --d) {
        const T dx = v.coord(d, i) - v.coord(d, j);
        const T si = v.sig(d, i);
        const T sj = v.sig(d, j);
        s = dx * dx / (si * si + sj * sj) + s;
      }
      return std::sqrt(s);
    }
  }
  return T(0);
}

// ConvolutionalKernel.hpp (detail) :21-57. point ids compare global indices.
template <class T>
inline T kernel_value(const oracle_params* p, T d, int64_t pid, int64_t jid) {
  if (pid == jid) return T(1);
  switch (p->kernel) {
    case ORACLE_KFLAT:
      return static_cast<T>(p->kparam[0]);
    case ORACLE_KEXP: {
      const T lambda = static_cast<T>(p->kparam[0]);
      const T amp = static_cast<T>(p->kparam[1]);
      return amp * exp_t<T>(-lambda * d);
    }
    case ORACLE_KGAUSS: {
      const T sg = static_cast<T>(p->kparam[0]);
      const T amp = static_cast<T>(p->kparam[1]);
      return amp * exp_t<T>(-d * d / (static_cast<T>(2) * sg * sg));
    }
  }
  return T(0);
}

// Search box (ClusteringKernels.hpp:111-121 and TilesView.hpp:77-87).
template <class T>
inline void search_box(const Geo<T>& g, const View<T>& v, int32_t i, T R, int* lo, int* hi) {
  const T sqrt2 = std::sqrt(static_cast<T>(2));
  for (int d = 0; d < g.D; ++d) {
    const T sg = v.sig(d, i);
    const T r = std::max(R, R * sg * sqrt2);
    const T x = v.coord(d, i);
    int a = get_bin(g, x - r, d);
    int b = get_bin(g, x + r, d);
    if (g.wrapped[d] && a > b) b += g.nper;
    lo[d] = a;
    hi[d] = b;
  }
}

// Visit the tiles of a box in the reference's lexicographic order (dim 0
// outermost, ClusteringKernels.hpp:71-89), calling f(tile) for each.
template <class T, class F>
inline void for_each_tile(const Geo<T>& g, const int* lo, const int* hi, F&& f) {
  int b[10];
  for (int d = 0; d < g.D; ++d) b[d] = lo[d];
  while (true) {
    int64_t t = 0;
    for (int d = 0; d < g.D; ++d) t += g.stride[d] * (g.wrapped[d] ? b[d] % g.nper : b[d]);
    f(static_cast<int32_t>(t));
    int d = g.D - 1;
    while (d >= 0) {
      if (b[d] < hi[d]) {
        ++b[d];
        break;
      }
      b[d] = lo[d];
      --d;
    }
    if (d < 0) break;
  }
}

template <class T>
void density(const oracle_params* p, const Geo<T>& g, const View<T>& v, const CSR& csr,
             T* rho) {
  const T dc = static_cast<T>(p->dc);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int32_t i = 0; i < v.n; ++i) {
    int lo[10], hi[10];
    search_box(g, v, i, dc, lo, hi);
    T r = T(0);
    for_each_tile(g, lo, hi, [&](int32_t t) {
      for (int32_t k = csr.off[t]; k < csr.off[t + 1]; ++k) {
        const int32_t j = csr.idx[k];
        const T dist = distance(p, v, i, j);
        const T kv = kernel_value<T>(p, dist, int64_t(v.g0) + i, int64_t(v.g0) + j);
        r += static_cast<T>(static_cast<int>(dist <= dc)) * kv * v.w[j];
      }
    });
    rho[i] = r;
  }
}

template <class T>
void nearest_higher(const oracle_params* p, const Geo<T>& g, const View<T>& v, const CSR& csr,
                    const T* rho, int32_t* nh, T* delta) {
  const T dm = static_cast<T>(p->dm);
  const T sdc = static_cast<T>(p->seed_dc);
  const T rhoc = static_cast<T>(p->rhoc);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int32_t i = 0; i < v.n; ++i) {
    T delta_i = std::numeric_limits<T>::max();
    int32_t nh_i = -1;
    const T rho_i = rho[i];
    const T ui = v.u ? v.u[i] : T(1);
    const T eff_min = rhoc * ui;
    int lo[10], hi[10];
    search_box(g, v, i, dm, lo, hi);
    const T eff = (rho_i >= eff_min) ? sdc : dm;
    const uint64_t tag_i = v.tag(i);
    for_each_tile(g, lo, hi, [&](int32_t t) {
      for (int32_t k = csr.off[t]; k < csr.off[t + 1]; ++k) {
        const int32_t j = csr.idx[k];
        const uint64_t tag_j = v.tag(j);
        const T rho_j = rho[j];
        bool higher = rho_j > rho_i;
        higher = higher || ((rho_j == rho_i) && (rho_j > T(0)) && (tag_j > tag_i));
        if (higher) {
          const T dist = distance(p, v, i, j);
          if (dist <= eff &&
              ((dist < delta_i) ||
               ((dist == delta_i) && (nh_i >= 0) &&
                ((rho_j > rho[nh_i]) || ((rho_j == rho[nh_i]) && (tag_j > v.tag(nh_i))))))) {
            delta_i = dist;
            nh_i = j;
          }
        }
      }
    });
    nh[i] = nh_i;
    delta[i] = delta_i;
  }
}

// Seeds (ClusteringKernels.hpp:300-316) numbered in ascending index; followers
// chase nh to a seed or to a point with nh == -1 (:329-343, :423-446).
template <class T>
int64_t seeds_and_labels(const oracle_params* p, const View<T>& v, const T* rho, const int32_t* nh,
                         int32_t* is_seed, int32_t* cluster, int64_t first_id) {
  const T rhoc = static_cast<T>(p->rhoc);
  int64_t next = first_id;
  for (int32_t i = 0; i < v.n; ++i) {
    const T ui = v.u ? v.u[i] : T(1);
    const bool seed = (nh[i] == -1) && (rho[i] >= rhoc * ui);
    is_seed[i] = seed ? 1 : 0;
    cluster[i] = seed ? static_cast<int32_t>(next++) : -1;
  }
  for (int32_t i = 0; i < v.n; ++i) {
    if (is_seed[i] || nh[i] == -1) continue;
    int32_t cur = i;
    while (!is_seed[cur] && nh[cur] != -1) cur = nh[cur];
    cluster[i] = cluster[cur];
  }
  return next - first_id;
}

template <class T>
bool params_ok(const oracle_params* p) {
  if (!p || p->ndim < 1 || p->ndim > 10) return false;
  if (!(p->dc > 0) || !(p->rhoc >= 0) || !(p->dm > 0) || !(p->seed_dc > 0)) return false;
  if (p->points_per_tile < 1) return false;
  return true;
}

template <class T>
struct Outs {
  T* rho;
  int32_t* nh;
  T* delta;
  int32_t* is_seed;
  int32_t* cluster;
  int32_t* tile_offsets;
  int32_t* tile_index;
  int64_t* n_seeds;
};

// One event, local arrays, with outputs written at local positions.
template <class T>
int64_t run_event(const oracle_params* p, const Geo<T>& g, const View<T>& v, T* rho, int32_t* nh,
                  T* delta, int32_t* is_seed, int32_t* cluster, int64_t first_id, CSR* keep) {
  CSR csr;
  build_tiles(g, v, csr);
  density(p, g, v, csr, rho);
  nearest_higher(p, g, v, csr, rho, nh, delta);
  const int64_t s = seeds_and_labels(p, v, rho, nh, is_seed, cluster, first_id);
  if (keep) *keep = std::move(csr);
  return s;
}

template <class T>
int run_single(const oracle_params* p, int32_t n, const T* coords, const T* weight,
               const T* sigma, const T* u, const uint32_t* tags, const Outs<T>& o) {
  if (!params_ok<T>(p) || n < 1 || !coords || !weight) return -1;
  Geo<T> g;
  make_geometry(g, p, n, n, coords, n);
  View<T> v{p->ndim, n, 0, coords, n, weight, sigma, u, tags};
  std::vector<T> rho(n), delta(n);
  std::vector<int32_t> nh(n), seed(n), cl(n);
  CSR csr;
  const int64_t s = run_event(p, g, v, rho.data(), nh.data(), delta.data(), seed.data(),
                              cl.data(), 0, &csr);
  if (o.rho) std::copy(rho.begin(), rho.end(), o.rho);
  if (o.nh) std::copy(nh.begin(), nh.end(), o.nh);
  if (o.delta) std::copy(delta.begin(), delta.end(), o.delta);
  if (o.is_seed) std::copy(seed.begin(), seed.end(), o.is_seed);
  if (o.cluster) std::copy(cl.begin(), cl.end(), o.cluster);
  if (o.tile_offsets) std::copy(csr.off.begin(), csr.off.end(), o.tile_offsets);
  if (o.tile_index) std::copy(csr.idx.begin(), csr.idx.end(), o.tile_index);
  if (o.n_seeds) *o.n_seeds = s;
  return 0;
}

}  // namespace

extern "C" {

float oracle_expf(float x) { return expf_cw(x); }

double oracle_exp_f64(double x) { return expd_cw(x); }

double oracle_expl_f64(double x) { return static_cast<double>(std::exp(static_cast<long double>(x))); }

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int oracle_geometry_f32(const oracle_params* p, int32_t n, const float* coords,
                        oracle_geometry* out) {
  if (!params_ok<float>(p) || n < 1 || !coords || !out) return -1;
  Geo<float> g;
  make_geometry(g, p, n, n, coords, n);
  out->nper = g.nper;
  out->ntiles = g.ntiles;
  for (int d = 0; d < 10; ++d) {
    out->mn[d] = d < g.D ? g.mn[d] : 0;
    out->mx[d] = d < g.D ? g.mx[d] : 0;
    out->side[d] = d < g.D ? g.side[d] : 0;
  }
  return 0;
}

int oracle_run_f32(const oracle_params* p, int32_t n, const float* coords, const float* weight,
                   const float* sigma, const float* u, const uint32_t* tags, oracle_out_f32* out) {
  if (!out) return -1;
  Outs<float> o{out->rho,     out->nh,           out->delta,      out->is_seed,
                out->cluster, out->tile_offsets, out->tile_index, out->n_seeds};
  return run_single<float>(p, n, coords, weight, sigma, u, tags, o);
}

int oracle_run_f64(const oracle_params* p, int32_t n, const double* coords, const double* weight,
                   const double* sigma, const double* u, const uint32_t* tags,
                   oracle_out_f64* out) {
  if (!out) return -1;
  Outs<double> o{out->rho,     out->nh,           out->delta,      out->is_seed,
                 out->cluster, out->tile_offsets, out->tile_index, out->n_seeds};
  return run_single<double>(p, n, coords, weight, sigma, u, tags, o);
}

int oracle_run_batched_f32(const oracle_params* p, int32_t n_total, const float* coords,
                           const float* weight, const float* sigma, const float* u,
                           const uint32_t* tags, const uint32_t* sizes, int32_t n_events,
                           oracle_out_f32* out, int32_t* event_of_cluster) {
  if (!params_ok<float>(p) || !out || n_events < 1 || !sizes || !coords || !weight) return -1;
  int64_t sum = 0;
  for (int32_t e = 0; e < n_events; ++e) sum += sizes[e];
  if (sum != n_total || n_total < 1) return -2;
  Geo<float> global;
  if (p->geometry == 1) make_geometry(global, p, n_total, n_total, coords, n_total);
  int64_t first_id = 0;
  int64_t offset = 0;
  for (int32_t e = 0; e < n_events; ++e) {
    const int32_t ne = static_cast<int32_t>(sizes[e]);
    if (ne == 0) continue;
    const float* c = coords + offset;
    View<float> v{p->ndim,
                  ne,
                  static_cast<int32_t>(offset),
                  c,
                  n_total,
                  weight + offset,
                  sigma ? sigma + offset : nullptr,
                  u ? u + offset : nullptr,
                  tags ? tags + offset : nullptr};
    Geo<float> g;
    if (p->geometry == 1) {
      g = global;
    } else {
      make_geometry(g, p, ne, ne, c, n_total);
    }
    std::vector<float> rho(ne), delta(ne);
    std::vector<int32_t> nh(ne), seed(ne), cl(ne);
    const int64_t s = run_event(p, g, v, rho.data(), nh.data(), delta.data(), seed.data(),
                                cl.data(), first_id, nullptr);
    for (int32_t i = 0; i < ne; ++i) {
      const int64_t gi = offset + i;
      if (out->rho) out->rho[gi] = rho[i];
      if (out->delta) out->delta[gi] = delta[i];
      if (out->nh) out->nh[gi] = nh[i] < 0 ? -1 : static_cast<int32_t>(nh[i] + offset);
      if (out->is_seed) out->is_seed[gi] = seed[i];
      if (out->cluster) out->cluster[gi] = cl[i];
    }
    if (event_of_cluster)
      for (int64_t k = 0; k < s; ++k) event_of_cluster[first_id + k] = e;
    first_id += s;
    offset += ne;
  }
  if (out->n_seeds) *out->n_seeds = first_id;
  return 0;
}

int oracle_brute_density_f32(const oracle_params* p, int32_t n, const float* coords,
                             const float* weight, const float* sigma, double* rho64,
                             double* abs_sum, int32_t* n_terms) {
  if (!params_ok<float>(p) || n < 1 || !rho64) return -1;
  View<float> v{p->ndim, n, 0, coords, n, weight, sigma, nullptr, nullptr};
  const float dc = static_cast<float>(p->dc);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int32_t i = 0; i < n; ++i) {
    double r = 0, a = 0;
    int32_t cnt = 0;
    for (int32_t j = 0; j < n; ++j) {
      const float dist = distance(p, v, i, j);
      if (!(dist <= dc)) continue;
      double kv;
      if (i == j) {
        kv = 1.0;
      } else if (p->kernel == ORACLE_KFLAT) {
        kv = static_cast<double>(static_cast<float>(p->kparam[0]));
      } else if (p->kernel == ORACLE_KEXP) {
        kv = static_cast<float>(p->kparam[1]) *
             std::exp(-static_cast<double>(static_cast<float>(p->kparam[0])) * dist);
      } else {
        const double sg = static_cast<float>(p->kparam[0]);
        kv = static_cast<float>(p->kparam[1]) *
             std::exp(-static_cast<double>(dist) * dist / (2.0 * sg * sg));
      }
      const double term = kv * weight[j];
      r += term;
      a += std::fabs(term);
      ++cnt;
    }
    rho64[i] = r;
    if (abs_sum) abs_sum[i] = a;
    if (n_terms) n_terms[i] = cnt;
  }
  return 0;
}

int oracle_brute_nh_f32(const oracle_params* p, int32_t n, const float* coords,
                        const float* sigma, const float* u, const uint32_t* tags,
                        const float* rho, int32_t* nh, float* delta) {
  if (!params_ok<float>(p) || n < 1 || !rho || !nh) return -1;
  View<float> v{p->ndim, n, 0, coords, n, nullptr, sigma, u, tags};
  const float dm = static_cast<float>(p->dm);
  const float sdc = static_cast<float>(p->seed_dc);
  const float rhoc = static_cast<float>(p->rhoc);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int32_t i = 0; i < n; ++i) {
    const float ui = u ? u[i] : 1.f;
    const float eff = (rho[i] >= rhoc * ui) ? sdc : dm;
    // Lexicographic minimum of (distance, -rho_j, -tag_j) over higher candidates.
    int32_t best = -1;
    float bd = FLT_MAX;
    for (int32_t j = 0; j < n; ++j) {
      const bool higher = rho[j] > rho[i] ||
                          (rho[j] == rho[i] && rho[j] > 0 && v.tag(j) > v.tag(i));
      if (!higher) continue;
      const float dist = distance(p, v, i, j);
      if (!(dist <= eff)) continue;
      bool better = best < 0 || dist < bd;
      if (!better && dist == bd)
        better = rho[j] > rho[best] || (rho[j] == rho[best] && v.tag(j) > v.tag(best));
      if (better) {
        best = j;
        bd = dist;
      }
    }
    nh[i] = best;
    if (delta) delta[i] = bd;
  }
  return 0;
}

int oracle_pair_counts_f32(const oracle_params* p, int32_t n, const float* coords,
                           const float* weight, const float* sigma, const float* u,
                           const uint32_t* tags, const float* rho, int64_t* visited_density,
                           int64_t* in_radius, int64_t* visited_nh, int64_t* nh_pairs) {
  if (!params_ok<float>(p) || n < 1) return -1;
  Geo<float> g;
  make_geometry(g, p, n, n, coords, n);
  View<float> v{p->ndim, n, 0, coords, n, weight, sigma, u, tags};
  CSR csr;
  build_tiles(g, v, csr);
  const float dc = static_cast<float>(p->dc);
  const float dm = static_cast<float>(p->dm);
  const float sdc = static_cast<float>(p->seed_dc);
  const float rhoc = static_cast<float>(p->rhoc);
  int64_t vd = 0, ir = 0, vn = 0, np = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : vd, ir, vn, np)
#endif
  for (int32_t i = 0; i < n; ++i) {
    int lo[10], hi[10];
    search_box(g, v, i, dc, lo, hi);
    for_each_tile(g, lo, hi, [&](int32_t t) {
      for (int32_t k = csr.off[t]; k < csr.off[t + 1]; ++k) {
        ++vd;
        if (distance(p, v, i, csr.idx[k]) <= dc) ++ir;
      }
    });
    if (!rho) continue;
    search_box(g, v, i, dm, lo, hi);
    const float ui = u ? u[i] : 1.f;
    const float eff = (rho[i] >= rhoc * ui) ? sdc : dm;
    for_each_tile(g, lo, hi, [&](int32_t t) {
      for (int32_t k = csr.off[t]; k < csr.off[t + 1]; ++k) {
        const int32_t j = csr.idx[k];
        ++vn;
        const bool higher = rho[j] > rho[i] ||
                            (rho[j] == rho[i] && rho[j] > 0 && v.tag(j) > v.tag(i));
        if (higher && distance(p, v, i, j) <= eff) ++np;
      }
    });
  }
  if (visited_density) *visited_density = vd;
  if (in_radius) *in_radius = ir;
  if (visited_nh) *visited_nh = vn;
  if (nh_pairs) *nh_pairs = np;
  return 0;
}

}  // extern "C"
