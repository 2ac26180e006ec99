// Device math probe: lets the tests compare the stages' sqrt, exp and division
// bits with the host's.
#include <hip/hip_runtime.h>

#include <string>

#include "clue_amd.h"
#include "clue_device.h"

namespace {

__global__ void k_debug_math(int n, const float* __restrict__ a, const float* __restrict__ b,
                             float* __restrict__ out_sqrt, float* __restrict__ out_exp,
                             float* __restrict__ out_div) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = a[i];
  if (out_sqrt) out_sqrt[i] = clue_dev::sqrt_rn(x);
  if (out_exp) out_exp[i] = clue_dev::expf_cw(x);
  if (out_div) out_div[i] = x / b[i];
}

__global__ void k_debug_math_f64(int n, const double* __restrict__ a, const double* __restrict__ b,
                                 double* __restrict__ out_sqrt, double* __restrict__ out_exp,
                                 double* __restrict__ out_div) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = a[i];
  if (out_sqrt) out_sqrt[i] = clue_dev::sqrt_rn(x);
  if (out_exp) out_exp[i] = clue_dev::expd_cw(x);
  if (out_div) out_div[i] = x / b[i];
}

// A plain streaming copy, the reference point for the HBM traffic counters.
__global__ void k_debug_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

extern "C" clue_status clue_debug_copy_impl(void* stream, void* dst, const void* src,
                                            size_t bytes) {
  const int64_t n = static_cast<int64_t>(bytes / sizeof(uint4));
  if (n <= 0) return CLUE_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_debug_copy, dim3(2048), dim3(256), 0, s, static_cast<const uint4*>(src),
                     static_cast<uint4*>(dst), n);
  if (hipGetLastError() != hipSuccess) return CLUE_ERR_HIP;
  return hipStreamSynchronize(s) == hipSuccess ? CLUE_OK : CLUE_ERR_HIP;
}

extern "C" clue_status clue_debug_math_impl(void* stream, int32_t n, const float* a,
                                            const float* b, float* out_sqrt, float* out_exp,
                                            float* out_div) {
  if (n <= 0) return CLUE_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_debug_math, dim3((n + 255) / 256), dim3(256), 0, s, n, a, b, out_sqrt,
                     out_exp, out_div);
  if (hipGetLastError() != hipSuccess) return CLUE_ERR_HIP;
  return hipStreamSynchronize(s) == hipSuccess ? CLUE_OK : CLUE_ERR_HIP;
}

extern "C" clue_status clue_debug_math_f64_impl(void* stream, int32_t n, const double* a,
                                                const double* b, double* out_sqrt,
                                                double* out_exp, double* out_div) {
  if (n <= 0) return CLUE_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_debug_math_f64, dim3((n + 255) / 256), dim3(256), 0, s, n, a, b, out_sqrt,
                     out_exp, out_div);
  if (hipGetLastError() != hipSuccess) return CLUE_ERR_HIP;
  return hipStreamSynchronize(s) == hipSuccess ? CLUE_OK : CLUE_ERR_HIP;
}
