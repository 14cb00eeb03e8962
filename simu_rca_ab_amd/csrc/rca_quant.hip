// FP16-I/Q storage: complex64 traces -> half2 samples + one FP32 scale per
// trace (the precision mode's storage format; no reference counterpart, the
// reference keeps I/Q in FP32, bfVirtualSourceGeneric_RCA.m:476-484).
//
// scale = max over the trace of max(|re|, |im|) (1 for an all-zero trace);
// each sample is stored as (re / scale, im / scale) rounded to the nearest
// half.  One workgroup per trace: a max reduction over the trace, then the
// scaled conversion (the trace, 16 KB at 2048 samples, is read again from
// cache).  HBM-bound: 8 B read and 4 B written per sample.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "../../include/rca_das.h"

namespace rca_internal {
void set_error(const char* msg);
}

namespace {

constexpr int kQThreads = 256;

__global__ void __launch_bounds__(kQThreads) quantize_fp16(const float2* __restrict__ iq, int n_sample,
                                                          long long n_trace, __half2* __restrict__ out,
                                                          float* __restrict__ scale) {
  __shared__ float s_max[kQThreads / 64];
  for (long long tr = blockIdx.x; tr < n_trace; tr += gridDim.x) {
    const float2* src = iq + tr * n_sample;
    float m = 0.0f;
    for (int i = threadIdx.x; i < n_sample; i += kQThreads) {
      const float2 v = src[i];
      m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    m = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
    const float sc = m > 0.0f ? m : 1.0f;
    if (threadIdx.x == 0) scale[tr] = sc;
    __half2* dst = out + tr * n_sample;
    for (int i = threadIdx.x; i < n_sample; i += kQThreads) {
      const float2 v = src[i];
      dst[i] = __halves2half2(__float2half_rn(v.x / sc), __float2half_rn(v.y / sc));
    }
    __syncthreads();   // s_max is reused by the next trace
  }
}

}  // namespace

extern "C" {

int rca_quantize_fp16(const void* iq, long long n_trace, int n_sample, void* out, float* scale,
                      hipStream_t stream) {
  if (n_trace < 0 || n_sample < 1 || (n_trace > 0 && (!iq || !out || !scale))) {
    rca_internal::set_error("rca_quantize_fp16: bad arguments");
    return RCA_ERR_INVALID;
  }
  if (n_trace == 0) return RCA_OK;
  const unsigned grid = (unsigned)(n_trace < (1LL << 20) ? n_trace : (1LL << 20));
  hipLaunchKernelGGL(quantize_fp16, dim3(grid), dim3(kQThreads), 0, stream,
                     reinterpret_cast<const float2*>(iq), n_sample, n_trace,
                     reinterpret_cast<__half2*>(out), scale);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    rca_internal::set_error(hipGetErrorString(e));
    return RCA_ERR_HIP;
  }
  return RCA_OK;
}

}  // extern "C"
