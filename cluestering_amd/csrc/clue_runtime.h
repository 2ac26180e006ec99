// Host runtime shared by the C ABI and the server: error reporting, device
// workspaces, batch descriptors and the launch of one batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "clue_amd.h"
#include "clue_host.h"
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

// Device buffer that only grows. With a stream and a memory pool set, growing
// and releasing are stream-ordered (hipMallocFromPoolAsync / hipFreeAsync on that
// stream): they do not wait for the whole device the way hipFree does, and the
// pool keeps freed blocks for later growth instead of handing them back to the
// driver. Such a buffer is used on that stream only, and its owner calls
// release() before the stream and the pool go away.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipStream_t stream = nullptr;
  hipMemPool_t pool = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p && !pool) (void)hipFree(p);
  }
  hipError_t reserve(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) {
      hipError_t e = pool ? hipFreeAsync(p, stream) : hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    const size_t want = grown_capacity(0, bytes);
    hipError_t e = pool ? hipMallocFromPoolAsync(&p, want, pool, stream) : hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  // Frees the buffer (capacity back to 0).
  void release() {
    if (p) (void)(pool ? hipFreeAsync(p, stream) : hipFree(p));
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Workspace {
  DevBuf buf[S_NUM];
  // Stream-ordered growth of every buffer on stream s from pool mp (see DevBuf).
  void set_stream(hipStream_t s, hipMemPool_t mp) {
    for (DevBuf& b : buf) {
      b.stream = s;
      b.pool = mp;
    }
  }
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

// plan_batch with the error reported through fail(); `fixed` (may be NULL)
// replaces the built-in table's tuning.
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
// The workspace shape of a batch run (what workspace_requests sizes).
template <class P, class Pts, class Res>
WorkspaceShape workspace_shape(const BatchPlan& bp, const P* p, const Pts* pts, const Res* res);
// Slot k of a workspace as a U array.
template <class U>
U* at(void* const (&slot)[S_NUM], int k) {
  return static_cast<U*>(slot[k]);
}
// Fills the BatchBuffers and result pointers of one batch from workspace slots
// that hold at least workspace_requests' bytes each (NULL results use workspace
// slots) and, with `upload`, copies its descriptors from the host on stream s
// (without, the caller writes slot[S_EVENTS] and slot[S_CHUNKS] on the device).
template <class P, class Pts, class Res>
clue_status bind_batch(void* const (&slot)[S_NUM], const BatchPlan& bp, const P* p, const Pts* pts,
                       Res* res, hipStream_t s, BatchBuffersT<value_t<P>>* bb,
                       ResultArraysT<value_t<P>>* ra, bool upload);
// Reserves the workspace of one batch, uploads its descriptors on stream s and
// fills the BatchBuffers and result pointers (NULL results use workspace buffers).
template <class P, class Pts, class Res>
clue_status prepare_batch(Workspace& ws, const BatchPlan& bp, const P* p, const Pts* pts,
                          Res* res, hipStream_t s, BatchBuffersT<value_t<P>>* bb,
                          ResultArraysT<value_t<P>>* ra, bool upload = true);

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
