// Host runtime shared by the C ABI and the server: error reporting, device
// workspaces, batch descriptors and the launch of one batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "clue_amd.h"
#include "clue_internal.h"

namespace clue_rt {

using namespace clue_dev;

// Records the message for clue_last_error() and returns the code.
clue_status fail(clue_status code, const std::string& msg);
const char* last_error();

#define CLUE_HIP(call)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::clue_rt::fail(CLUE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kChunk = 2048;  // points per chunk of the per-point grids
constexpr int kDefaultPointsPerTile = 128;

// Device buffer that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t reserve(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    size_t want = bytes + bytes / 4 + 256;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Workspace slots of one batch run.
enum Slot {
  S_EVENTS, S_CHUNKS, S_EXT, S_GEO, S_K0, S_V0, S_K1, S_V1, S_HIST, S_SCAN, S_TOFF, S_ICOUNT,
  S_NITEMS, S_ITEMS, S_SRHO, S_SKEY, S_SNH, S_SDELTA, S_FLAG, S_RANK, S_NSEEDS, S_SEEDS, S_SX,
  S_SW, S_SSIG, S_SU, S_STAG, S_SFW, S_PERM, S_RHO, S_NH, S_ISSEED, S_DELTA, S_HIN, S_HCL,
  S_QPERM, S_QPERM_NH, S_QFINE, S_QKEY, S_HX, S_HKEY, S_HIDX, S_HBOX, S_HBKEY, S_NUM
};

struct Workspace {
  DevBuf buf[S_NUM];
};

// Host-side description of one batch (all derived from the event sizes).
struct BatchPlan {
  std::vector<EventDesc> events;
  std::vector<Chunk> chunks;
  int64_t n_total = 0;
  int64_t n_tiles_total = 0;
  int64_t max_items = 0;
  int key_bits = 0;
  clue_tuning tuning{};
};

// The C structs of each precision and the point type they carry.
template <class P>
struct ValueOf;
template <>
struct ValueOf<clue_params> {
  using type = float;
};
template <>
struct ValueOf<clue_params_f64> {
  using type = double;
};
template <class P>
using value_t = typename ValueOf<P>::type;

// Tuning of a batch: `fixed` when given, else the built-in table's entry for the
// batch's ndim and mean event size.
clue_tuning table_tuning(int ndim, int64_t n_total, int64_t n_events);
template <class P>
clue_status make_batch_plan(const P* p, const uint32_t* sizes, int32_t n_events, BatchPlan* bp,
                            const clue_tuning* fixed = nullptr);
// Largest t with sqrt(t) <= r (correctly rounded sqrt, in the precision of r).
template <class T>
T sq_max(T r);
template <class P>
clue_status validate(const P* p);
template <class P, class Pts>
RunParamsT<value_t<P>> make_run_params(const P* p, const Pts* pts, const BatchPlan& bp);
template <class P, class Pts>
clue_status check_points(const P* p, const Pts* pts, int64_t n_expected);
template <class P, class Pts>
PointArraysT<value_t<P>> point_arrays(const P* p, const Pts* pts);
// Reserves the workspace of one batch, uploads its descriptors on stream s and
// fills the BatchBuffers and result pointers (NULL results use workspace buffers).
template <class P, class Pts, class Res>
clue_status prepare_batch(Workspace& ws, const BatchPlan& bp, const P* p, const Pts* pts,
                          Res* res, hipStream_t s, BatchBuffersT<value_t<P>>* bb,
                          ResultArraysT<value_t<P>>* ra);

struct ScopedDevice {
  int prev = 0;
  explicit ScopedDevice(int d) {
    (void)hipGetDevice(&prev);
    if (prev != d) (void)hipSetDevice(d);
  }
  ~ScopedDevice() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace clue_rt
