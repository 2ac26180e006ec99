/*
 * CLUE oracle: a clean-room CPU restatement of the serial CLUE semantics.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline). The product path (cluestering_amd/) never
 * links, loads or calls it.
 *
 * Every function follows the reference's serial backend; file:line citations
 * are into the CLUEstering sources:
 *   geometry      core/detail/SetupTiles.hpp:23-57, core/detail/ComputeTiles.hpp:17-37
 *   bins/box      data_structures/internal/TilesView.hpp:38-104
 *   tile CSR      data_structures/detail/AssociationMap.hpp:359-383 (ascending index)
 *   metrics       core/DistanceMetrics.hpp:64-318, internal/meta/accumulate.hpp (right fold)
 *   kernels       core/detail/ConvolutionalKernel.hpp:21-57
 *   density       core/detail/ClusteringKernels.hpp:29-140
 *   nearest-higher core/detail/ClusteringKernels.hpp:142-291
 *   seeds/labels  core/detail/ClusteringKernels.hpp:293-343, 419-446
 *   batched       core/detail/BatchedClusteringKernels.hpp:28-336,
 *                 core/detail/Clusterer.hpp:286-350
 *
 * FP contract: built with -ffp-contract=off and without fast-math, every
 * expression evaluated in the order the reference writes it. The
 * reference's own builds use -funsafe-math-optimizations -march=native, whose
 * bits are host dependent; this oracle fixes the strict IEEE evaluation.
 *
 * exp: the reference calls std::exp (glibc) on the host; its last-ulp bits are
 * library dependent. The oracle uses clue_expf (Cody-Waite reduction + degree-6
 * minimax polynomial, fmaf Horner, see oracle_expf), pinned against libm's double
 * exp to <= 1 ulp in the tests. The same published method is what the device
 * kernels implement.
 */
#ifndef CLUE_ORACLE_H
#define CLUE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum oracle_kernel { ORACLE_KFLAT = 0, ORACLE_KEXP = 1, ORACLE_KGAUSS = 2 };
enum oracle_metric {
  ORACLE_MEUCLID = 0,
  ORACLE_MWEUCLID = 1,
  ORACLE_MPERIODIC = 2,
  ORACLE_MMANHATTAN = 3,
  ORACLE_MCHEBYSHEV = 4,
  ORACLE_MWCHEBYSHEV = 5,
  ORACLE_MMAHALANOBIS = 6
};

typedef struct {
  int32_t ndim;            /* 1..10 */
  int32_t kernel;          /* oracle_kernel */
  int32_t metric;          /* oracle_metric */
  int32_t points_per_tile; /* reference: 128 (core/Clusterer.hpp:57) */
  int32_t wrapped[10];     /* periodic coordinates (TilesView.hpp:96-104) */
  int32_t geometry;        /* batched only: 0 = per-event tiles, 1 = reference global tiles */
  int32_t reserved;
  double dc, rhoc, dm, seed_dc;
  double kparam[2];  /* flat: {f}; exp: {lambda, A}; gauss: {sigma, A} */
  double mparam[10]; /* weights (weighted metrics) or periods (periodic) */
} oracle_params;

/* Outputs of one run. Any pointer may be NULL except where noted. */
typedef struct {
  float* rho;         /* [n] */
  int32_t* nh;        /* [n] nearest higher (index into the input, global in batched) */
  float* delta;       /* [n] distance to nh (FLT_MAX when none) */
  int32_t* is_seed;   /* [n] */
  int32_t* cluster;   /* [n] */
  int32_t* tile_offsets; /* [T+1] per-event tile CSR (single event only) */
  int32_t* tile_index;   /* [n] */
  int64_t* n_seeds;      /* [1] total seeds */
} oracle_out_f32;

typedef struct {
  double* rho;
  int32_t* nh;
  double* delta;
  int32_t* is_seed;
  int32_t* cluster;
  int32_t* tile_offsets;
  int32_t* tile_index;
  int64_t* n_seeds;
} oracle_out_f64;

/* Geometry of one event (A.2): n per dim, tile count, extremes and side. */
typedef struct {
  int32_t nper;
  int32_t ntiles;
  double mn[10], mx[10], side[10];
} oracle_geometry;

/* Returns 0 on success, negative on invalid arguments.
 * coords: ndim x n, SoA (coords[d*n + i]); weight: [n];
 * sigma: NULL or ndim x n; u (density uncertainty): NULL or [n]; tags: NULL or [n]. */
int oracle_run_f32(const oracle_params* p, int32_t n, const float* coords, const float* weight,
                   const float* sigma, const float* u, const uint32_t* tags, oracle_out_f32* out);
int oracle_run_f64(const oracle_params* p, int32_t n, const double* coords, const double* weight,
                   const double* sigma, const double* u, const uint32_t* tags, oracle_out_f64* out);

/* Batched run over events packed back to back (sizes[e] points each).
 * p->geometry selects per-event tiles (0) or the reference's global tiles (1).
 * Cluster ids are global and event-major; nh holds global indices.
 * event_of_cluster (NULL or [n_seeds]) receives the event of each cluster. */
int oracle_run_batched_f32(const oracle_params* p, int32_t n_total, const float* coords,
                           const float* weight, const float* sigma, const float* u,
                           const uint32_t* tags, const uint32_t* sizes, int32_t n_events,
                           oracle_out_f32* out, int32_t* event_of_cluster);

/* Geometry only (for pinning the tile grid). */
int oracle_geometry_f32(const oracle_params* p, int32_t n, const float* coords,
                        oracle_geometry* g);

/* Brute force, O(n^2), no tiles: density accumulated in double (float distances,
 * double kernel), the number of in-radius pairs and the sum of |terms|. */
int oracle_brute_density_f32(const oracle_params* p, int32_t n, const float* coords,
                             const float* weight, const float* sigma, double* rho64,
                             double* abs_sum, int32_t* n_terms);
/* Brute-force nearest-higher from given rho (all candidates, no tiles). */
int oracle_brute_nh_f32(const oracle_params* p, int32_t n, const float* coords,
                        const float* sigma, const float* u, const uint32_t* tags,
                        const float* rho, int32_t* nh, float* delta);

/* Count of candidate pairs within dc (density work) and of "higher" pairs within
 * the effective distance (nearest-higher work), both through the tile walk. */
int oracle_pair_counts_f32(const oracle_params* p, int32_t n, const float* coords,
                           const float* weight, const float* sigma, const float* u,
                           const uint32_t* tags, const float* rho, int64_t* visited_density,
                           int64_t* in_radius, int64_t* visited_nh, int64_t* nh_pairs);

float oracle_expf(float x);
/* The double-precision exp restatement, and exp in long double rounded to double
 * (the accuracy reference of the tests). */
double oracle_exp_f64(double x);
double oracle_expl_f64(double x);
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
