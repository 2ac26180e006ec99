/*
 * clue_amd.h — C ABI of the MI355X-native CLUE library (libclue_amd.so).
 *
 * Plain pointers and sizes only. Every entry point returns a clue_status; on
 * failure clue_last_error() gives a message (thread-local). No exception
 * crosses this boundary.
 *
 * Each entry point names the reference interface it replaces (file:line into
 * the CLUEstering sources).
 *
 * Memory: unless stated otherwise, point and result arrays are DEVICE pointers
 * owned by the caller. A context owns its stream, pooled workspaces and the
 * captured graphs; like the reference's Clusterer (core/Clusterer.hpp:37-51)
 * it is not thread-safe, and separate contexts may run concurrently.
 */
#ifndef CLUE_AMD_H
#define CLUE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t clue_status;
enum {
  CLUE_OK = 0,
  CLUE_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
  CLUE_ERR_OUT_OF_RANGE = 2,     /* std::out_of_range (bad dimension) */
  CLUE_ERR_RUNTIME = 3,          /* std::runtime_error (e.g. seeds before a run) */
  CLUE_ERR_HIP = 4,              /* a HIP runtime call failed */
  CLUE_ERR_CAPACITY = 5,         /* request exceeds a configured capacity/budget */
  CLUE_ERR_NO_DEVICE = 6
};

/* Convolutional kernels (core/ConvolutionalKernel.hpp:19-111). */
enum { CLUE_KERNEL_FLAT = 0, CLUE_KERNEL_EXP = 1, CLUE_KERNEL_GAUSS = 2 };

/* Distance metrics (core/DistanceMetrics.hpp:64-318). */
enum {
  CLUE_METRIC_EUCLIDEAN = 0,
  CLUE_METRIC_WEIGHTED_EUCLIDEAN = 1,
  CLUE_METRIC_PERIODIC_EUCLIDEAN = 2,
  CLUE_METRIC_MANHATTAN = 3,
  CLUE_METRIC_CHEBYSHEV = 4,
  CLUE_METRIC_WEIGHTED_CHEBYSHEV = 5,
  CLUE_METRIC_MAHALANOBIS = 6
};

/* Clustering parameters: Clusterer(dc, rhoc, dm, seed_dc)
 * (core/detail/Clusterer.hpp:33-49), setWrappedCoordinates (core/Clusterer.hpp:291-301),
 * the kernel and the metric template arguments. */
typedef struct {
  int32_t ndim;            /* 1..10 */
  int32_t kernel;          /* CLUE_KERNEL_* */
  int32_t metric;          /* CLUE_METRIC_* */
  int32_t points_per_tile; /* 0 selects the reference's 128 (core/Clusterer.hpp:57) */
  int32_t wrapped[10];     /* 1 = periodic coordinate */
  float dc;                /* density radius, > 0 */
  float rhoc;              /* min density, >= 0 */
  float dm;                /* outlier distance, > 0 (reference default: dc) */
  float seed_dc;           /* seeding distance, > 0 (reference default: dc) */
  float kparam[2];         /* flat {f}; exp {lambda, A}; gauss {sigma, A} */
  float mparam[10];        /* per-dimension weights or periods */
} clue_params;

/* Input points, SoA columns like PointsView (data_structures/internal/PointsCommon.hpp:97-110). */
typedef struct {
  int32_t n;
  const float* coords[10];          /* ndim device columns of n floats */
  const float* weight;              /* n */
  const float* sigma[10];           /* optional per-dimension sigma (NULL = unset) */
  const float* density_uncertainty; /* optional (NULL = 1) */
  const uint32_t* tags;             /* optional (NULL = point index) */
} clue_points;

/* Results, like PointsView's rho/nearest_higher/is_seed/cluster_index
 * (PointsCommon.hpp:103-106). cluster_index is required, the rest may be NULL.
 * delta (distance to the nearest higher, FLT_MAX when none) is not stored by the
 * reference; it is offered for parity checks. */
typedef struct {
  int32_t* cluster_index;
  int32_t* is_seed;
  float* rho;
  int32_t* nearest_higher;
  float* delta;
} clue_results;

typedef struct clue_ctx clue_ctx;

/* Library and device queries. listDevices (CLUEstering/BindingModules/Run.hpp:53-65). */
const char* clue_last_error(void);
const char* clue_version(void);
clue_status clue_device_count(int32_t* count);
clue_status clue_device_name(int32_t device, char* buf, int32_t buflen);

/* Context: replaces the Clusterer's device state (m_tiles, m_seeds, core/Clusterer.hpp:49-51)
 * and the alpaka Queue (utils/get_queue.hpp:19-48). stream == NULL creates an owned
 * stream; otherwise the given hipStream_t is used (not owned). */
clue_status clue_context_create(int32_t device, void* stream, clue_ctx** out);
clue_status clue_context_destroy(clue_ctx* ctx);
clue_status clue_context_stream(clue_ctx* ctx, void** stream);
clue_status clue_synchronize(clue_ctx* ctx);

/* Parameter validation as Clusterer's constructor/setParameters
 * (core/detail/Clusterer.hpp:42-48, 80-88) and the kernel constructors
 * (core/detail/ConvolutionalKernel.hpp:14-18, 28-33, 45-50). */
clue_status clue_validate_params(const clue_params* params);

/* Per-event, device-resident: make_clusters(queue, d_points, metric, kernel)
 * (core/detail/Clusterer.hpp:143-151). Runs on the context stream and waits for
 * completion, as the reference does. */
clue_status clue_make_clusters(clue_ctx* ctx, const clue_params* params,
                               const clue_points* points, clue_results* results);

/* Same, without the final wait (the caller synchronises the stream). */
clue_status clue_make_clusters_async(clue_ctx* ctx, const clue_params* params,
                                     const clue_points* points, clue_results* results);

/* Allocation-free variant of clue_make_clusters. NO COUNTERPART IN THE REFERENCE:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * results and wait as clue_make_clusters, but the pipeline workspace is a static
 * device array of 160 MiB per GPU (held until the process ends; its address is
 * taken when the first context on the device is created) instead of the
 * context's pooled buffers: the call makes no hipMalloc, hipFree, memset or
 * host-to-device copy. One event of up to about a million points in 2-D or 3-D
 * fits (fewer at higher ndim); larger events return CLUE_ERR_CAPACITY. Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (other
 * threads, other contexts, other streams) race and must be serialised by the
 * caller. clue_get_seeds after it reads the array, so it is valid until the
 * next such call on the device. */
clue_status clue_make_clusters_noalloc(clue_ctx* ctx, const clue_params* params,
                                       const clue_points* points, clue_results* results);

/* Batched, device-resident: make_clusters(queue, d_points, batch_item_sizes, ...)
 * (core/detail/Clusterer.hpp:171-183, 286-350). Events are packed back to back in
 * `points`; sizes (HOST array) gives the size of each. Each event is tiled as if
 * it were alone. Cluster ids are global and event-major, nearest_higher holds
 * indices into the packed arrays. event_of_cluster (device, NULL allowed, capacity
 * >= number of clusters) receives the event of each cluster
 * (getSampleAssociations, core/detail/Clusterer.hpp:221-241). n_clusters (host,
 * NULL allowed) receives the total number of clusters; passing it waits. */
clue_status clue_make_clusters_batched(clue_ctx* ctx, const clue_params* params,
                                       const clue_points* points, const uint32_t* sizes,
                                       int32_t n_events, clue_results* results,
                                       int32_t* event_of_cluster, int64_t* n_clusters);

/* Host in, host out: make_clusters(queue, h_points, d_points[, sizes])
 * (core/detail/Clusterer.hpp:127-138, 157-168) and the Python binding's mainRun
 * (CLUEstering/BindingModules/Run.hpp:25-49, 69-142). host_data is the (ndim+1) x n
 * C-contiguous SoA (coordinates then weights); host_cluster_index receives the
 * labels; sizes (host, NULL = one event) selects the batched form. Optional host
 * outputs may be NULL. */
clue_status clue_make_clusters_host(clue_ctx* ctx, const clue_params* params, int32_t n,
                                    const float* host_data, int32_t* host_cluster_index,
                                    const uint32_t* sizes, int32_t n_events,
                                    int32_t* host_is_seed, float* host_rho,
                                    int32_t* host_nearest_higher);

/* Double precision: Clusterer<Ndim, double> with PointsHost / PointsDevice<Ndim, double>
 * (core/Clusterer.hpp:37-51, value_type = double). Same fields and meaning as the
 * single-precision structs above, with double parameters, coordinates, weights and
 * outputs; every stage computes in double. Graph plans and the server are
 * single precision only. */
typedef struct {
  int32_t ndim;
  int32_t kernel;
  int32_t metric;
  int32_t points_per_tile;
  int32_t wrapped[10];
  double dc;
  double rhoc;
  double dm;
  double seed_dc;
  double kparam[2];
  double mparam[10];
} clue_params_f64;

typedef struct {
  int32_t n;
  const double* coords[10];
  const double* weight;
  const double* sigma[10];
  const double* density_uncertainty;
  const uint32_t* tags;
} clue_points_f64;

typedef struct {
  int32_t* cluster_index;
  int32_t* is_seed;
  double* rho;
  int32_t* nearest_higher;
  double* delta;
} clue_results_f64;

clue_status clue_validate_params_f64(const clue_params_f64* params);
/* clue_make_clusters for double points (core/detail/Clusterer.hpp:143-151). Waits. */
clue_status clue_make_clusters_f64(clue_ctx* ctx, const clue_params_f64* params,
                                   const clue_points_f64* points, clue_results_f64* results);
/* clue_make_clusters_batched for double points (core/detail/Clusterer.hpp:171-183). */
clue_status clue_make_clusters_batched_f64(clue_ctx* ctx, const clue_params_f64* params,
                                           const clue_points_f64* points, const uint32_t* sizes,
                                           int32_t n_events, clue_results_f64* results,
                                           int32_t* event_of_cluster, int64_t* n_clusters);
/* clue_make_clusters_host for a (ndim+1) x n double SoA (core/detail/Clusterer.hpp:127-138). */
clue_status clue_make_clusters_host_f64(clue_ctx* ctx, const clue_params_f64* params, int32_t n,
                                        const double* host_data, int32_t* host_cluster_index,
                                        const uint32_t* sizes, int32_t n_events,
                                        int32_t* host_is_seed, double* host_rho,
                                        int32_t* host_nearest_higher);

/* Seeds of the last run, ascending point index (getSeeds, core/detail/Clusterer.hpp:198-203).
 * `seeds` is a HOST array of capacity `capacity`; *count receives the number of seeds.
 * CLUE_ERR_RUNTIME before the first run. */
clue_status clue_get_seeds(clue_ctx* ctx, int32_t* seeds, int64_t capacity, int64_t* count);

/* Number of clusters of the last run (n_clusters = max label + 1,
 * utils/detail/get_cluster_properties.hpp:10-15). Waits. */
clue_status clue_get_n_clusters(clue_ctx* ctx, int64_t* n_clusters);

/* Key -> members association map built on the device (CSR), the device-side
 * make_associator (data_structures/internal/MakeAssociator.hpp:15-45):
 *   bins      = max key + 1 (0 when no key is >= 0);
 *   members   = indices i with keys[i] >= 0, grouped by key, ascending within a key;
 *   offsets   = bins + 1 boundaries (offsets[0] = 0, offsets[bins] = *n_members).
 * keys, offsets and members are DEVICE arrays: offsets of capacity
 * offsets_capacity (>= bins + 1), members of capacity n. *n_bins and *n_members
 * (host) receive the sizes, also when CLUE_ERR_CAPACITY reports that offsets is
 * too small (then nothing is written: call again with bins + 1 entries). Waits.
 * With keys = cluster_index it is getClusters / get_clusters for device points
 * (core/detail/Clusterer.hpp:212-217, utils/detail/get_clusters.hpp:29-48); with
 * keys = event_of_cluster it is getSampleAssociations (:231-241). */
clue_status clue_make_association(clue_ctx* ctx, const int32_t* keys, int32_t n, int32_t* offsets,
                                  int32_t offsets_capacity, int32_t* members, int32_t* n_bins,
                                  int32_t* n_members);

/* Per-stage device time of the last run in microseconds (hipEvents on the context
 * stream); enable with clue_set_profiling. names[i] is a static string. */
clue_status clue_set_profiling(clue_ctx* ctx, int32_t enabled);
clue_status clue_stage_times(clue_ctx* ctx, const char** names, float* us, int32_t capacity,
                             int32_t* count);

/* Launch tuning (no reference counterpart): choices that change the speed of the
 * neighbour stages and never their results. By default each run takes the entry
 * of the built-in table (measured on MI355X per ndim and mean event size by
 * cluestering_amd/autotune.py) that matches its batch. */
typedef struct {
  int32_t group_queries;  /* 1: order each tile's queries by box class and position */
  int32_t xcd_chunk_log2; /* deal 2^k consecutive work-item blocks per XCD turn; -1: none */
  int32_t block_waves;    /* waves per workgroup of the neighbour stages: 1..16, power of 2 (0: 4) */
  int32_t nh_plain_scan;  /* 1: nearest-higher scans whole tiles in sorted order; 0 (default): over
                             Morton-ordered candidate blocks, skipping blocks that cannot hold the
                             answer, whenever tags are the default ones and the metric allows it */
  int32_t persistent;     /* 0: one wave per work item, dispatched by the hardware; k > 0: k
                             persistent waves per CU pull (event, tile) work items from per-XCD
                             device queues (density and the block nearest-higher scan) */
  int32_t reserved[3];
} clue_tuning;
/* Fixes the tuning of the context's later runs and plans; NULL returns to the table. */
clue_status clue_set_tuning(clue_ctx* ctx, const clue_tuning* tuning);
/* The tuning a run of this batch would use (sizes: host, NULL = one event of n). */
clue_status clue_get_tuning(clue_ctx* ctx, const clue_params* params, int32_t n,
                            const uint32_t* sizes, int32_t n_events, clue_tuning* out);

/* Graph capture of the batched pipeline: records every stage of one batch
 * (fixed pointers and sizes) in a hipGraph and replays it with one launch. */
typedef struct clue_plan clue_plan;
clue_status clue_plan_create(clue_ctx* ctx, const clue_params* params, const clue_points* points,
                             const uint32_t* sizes, int32_t n_events, clue_results* results,
                             int32_t* event_of_cluster, clue_plan** out);
clue_status clue_plan_launch(clue_plan* plan); /* async on the context stream */
clue_status clue_plan_destroy(clue_plan* plan);

/* Continuous-batching server (no reference counterpart: the serving form of the
 * batched make_clusters). Clients submit events from any thread; one worker thread
 * folds every waiting event of the same parameters into the next launch on a free
 * lane (stream with pooled device buffers), so uploads, kernels and downloads of
 * consecutive batches overlap. Labels are numbered per event from 0, exactly as a
 * single-event make_clusters numbers them. */
typedef struct clue_server clue_server;
typedef struct {
  int32_t device;
  int32_t n_lanes;          /* concurrent launches (streams); 0 selects 2 */
  int64_t hbm_budget_bytes; /* device memory the lanes may use; 0 selects 64 GiB */
  int64_t max_batch_points; /* points per launch; 0 derives it from the budget (max 16 Mi) */
  int64_t max_batch_events; /* events per launch; 0 = no limit */
  int64_t staging_bytes;    /* pinned host staging ring; 0 selects 1 GiB */
} clue_server_config;
typedef struct {
  int64_t submit_ns; /* clue_now_ns() clock */
  int64_t launch_ns;
  int64_t done_ns;
  int32_t batch_events;
  int32_t reserved;
  int64_t batch_points;
} clue_server_timing;
typedef struct {
  int64_t events_submitted;
  int64_t events_done;
  int64_t points_done;
  int64_t batches;
  int64_t lane_points; /* per-launch point cap in force */
} clue_server_stats;

clue_status clue_server_create(const clue_server_config* cfg, clue_server** out);
clue_status clue_server_destroy(clue_server* server); /* finishes queued work first */
/* host_data: (ndim+1) x n SoA, copied before return (blocks while the staging ring
 * is full). CLUE_ERR_CAPACITY when the event exceeds the per-launch cap. */
clue_status clue_server_submit(clue_server* server, const clue_params* params, int32_t n,
                               const float* host_data, uint64_t* id);
/* Waits for an event (timeout_us < 0: no limit), copies its n labels into
 * host_labels (NULL allowed) and releases it. */
clue_status clue_server_wait(clue_server* server, uint64_t id, int32_t* host_labels,
                             int64_t timeout_us, clue_server_timing* timing);
clue_status clue_server_get_stats(clue_server* server, clue_server_stats* out);
int64_t clue_now_ns(void); /* steady clock of the timings, in nanoseconds */

/* Device memory helpers for callers without their own allocator. */
clue_status clue_malloc(clue_ctx* ctx, size_t bytes, void** ptr);
clue_status clue_free(clue_ctx* ctx, void* ptr);
clue_status clue_memcpy_h2d(clue_ctx* ctx, void* dst, const void* src, size_t bytes);
clue_status clue_memcpy_d2h(clue_ctx* ctx, void* dst, const void* src, size_t bytes);
clue_status clue_memset_d8(clue_ctx* ctx, void* dst, int value, size_t bytes);

/* Device math used by the stages, exposed so that tests can pin its bits against
 * the host: out_sqrt = sqrtf(a), out_exp = the library's expf(a), out_div = a / b.
 * All arrays are device arrays of n floats. No reference counterpart. */
clue_status clue_debug_math(clue_ctx* ctx, int32_t n, const float* a, const float* b,
                            float* out_sqrt, float* out_exp, float* out_div);

/* A plain device copy of `bytes` (multiple of 16), the calibration point of the
 * HBM traffic counters (FETCH_SIZE / WRITE_SIZE). No reference counterpart. */
clue_status clue_debug_copy(clue_ctx* ctx, void* dst, const void* src, size_t bytes);
/* The double-precision stage math (sqrt, exp, division) on device arrays. */
clue_status clue_debug_math_f64(clue_ctx* ctx, int32_t n, const double* a, const double* b,
                                double* out_sqrt, double* out_exp, double* out_div);
/* Test hook: fills every workspace buffer of the context with 0xFF bytes (NaN
 * floats, -1 integers), so that a later run cannot pass on values left by an
 * earlier one. Waits. */
clue_status clue_debug_poison(clue_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* CLUE_AMD_H */
