// VALU issue ceilings of the GPU, measured per instruction class: the compute
// roofline the DAS kernels are reported against (they are bound by vector-ALU
// issue, not by HBM or matrix cores).
//
// Every lane runs 8 independent chains, 64 instructions of one class per loop
// iteration, so issue is never held by a dependency; 8 waves per SIMD.  The
// lane-operation rate of a launch is the ceiling a stream of that class can
// reach.  Classes (mode):
//   0  v_fma_f32, multiplier in a VGPR      4  v_fma_f64, inline constants
//   1  v_fma_f32, inline constants          5  v_max_f32 across chains
//   2  v_add_f32, inline constant           6  v_add_u32 across chains
//   3  v_mul_f32, literal constant          7  v_fma_f32, three VGPR sources
//                                           8  v_add_f32, two VGPR sources
//                                           9  v_mul_f32, two VGPR sources
//                                          10  v_fma_f32, SGPR + two VGPRs
//                                          11  v_fma_f32, two VGPRs + literal
// (no packed forms: the library is built without packed-FP32 instructions).
// Modes 7-9 add loop-invariant per-lane operands b[k], c[k] held in VGPRs.
#include <hip/hip_runtime.h>

#include "../../include/rca_das.h"

namespace rca_internal {
void set_error(const char* msg);
}

namespace {

constexpr int kProbeThreads = 256;
constexpr int kChains = 8;
constexpr int kUnroll = 8;

template <int MODE>
__global__ void __launch_bounds__(kProbeThreads) valu_probe(float* __restrict__ out, int iters, float m,
                                                           float c) {
  float s = 0.0f;
  if constexpr (MODE == 4) {
    double a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = (double)(threadIdx.x + k) * 1e-3;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
      for (int r = 0; r < kUnroll; ++r) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) a[k] = fma(a[k], 0.5, 1.0);
      }
    }
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += (float)a[k];
  } else if constexpr (MODE == 6) {
    unsigned a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = threadIdx.x * 7u + k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
      for (int r = 0; r < kUnroll; ++r) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) a[k] += a[(k + 1) % kChains];
      }
    }
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += (float)(a[k] & 1023u);
  } else if constexpr (MODE >= 7) {
    float a[kChains], b[kChains], cc[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
      a[k] = (float)(threadIdx.x + k) * 1e-3f;
      b[k] = m - (float)((threadIdx.x + k) & 7) * 1e-6f;
      cc[k] = c + (float)((threadIdx.x * 3 + k) & 7) * 1e-7f;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
      for (int r = 0; r < kUnroll; ++r) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
          if constexpr (MODE == 7) a[k] = fmaf(a[k], b[k], cc[k]);
          else if constexpr (MODE == 8) a[k] = a[k] + cc[k];
          else if constexpr (MODE == 9) a[k] = a[k] * b[k];
          else if constexpr (MODE == 10) a[k] = fmaf(a[k], m, cc[k]);
          else a[k] = fmaf(a[k], b[k], 0.123f);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k] + b[k];
  } else {
    float a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = (float)(threadIdx.x + k) * 1e-3f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
      for (int r = 0; r < kUnroll; ++r) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
          if constexpr (MODE == 0) a[k] = fmaf(a[k], m, c);
          else if constexpr (MODE == 1) a[k] = fmaf(a[k], 0.5f, 1.0f);
          else if constexpr (MODE == 2) a[k] = a[k] + 1.0f;
          else if constexpr (MODE == 3) a[k] = a[k] * 0.999f;
          else a[k] = fmaxf(a[k], a[(k + 1) % kChains]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
  }
  out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = s;
}

}  // namespace

extern "C" {

int rca_probe_valu(int mode, float* out, int n_blocks, int iters, hipStream_t stream,
                   long long* lane_ops) {
  if (!out || n_blocks < 1 || iters < 1 || mode < 0 || mode > 11) {
    rca_internal::set_error("rca_probe_valu: bad arguments");
    return RCA_ERR_INVALID;
  }
  const dim3 g(n_blocks), b(kProbeThreads);
  switch (mode) {
    case 0: hipLaunchKernelGGL(valu_probe<0>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 1: hipLaunchKernelGGL(valu_probe<1>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 2: hipLaunchKernelGGL(valu_probe<2>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 3: hipLaunchKernelGGL(valu_probe<3>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 4: hipLaunchKernelGGL(valu_probe<4>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 5: hipLaunchKernelGGL(valu_probe<5>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 6: hipLaunchKernelGGL(valu_probe<6>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 7: hipLaunchKernelGGL(valu_probe<7>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 8: hipLaunchKernelGGL(valu_probe<8>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 9: hipLaunchKernelGGL(valu_probe<9>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    case 10: hipLaunchKernelGGL(valu_probe<10>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
    default: hipLaunchKernelGGL(valu_probe<11>, g, b, 0, stream, out, iters, 0.999f, 1e-3f); break;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    rca_internal::set_error(hipGetErrorString(e));
    return RCA_ERR_HIP;
  }
  if (lane_ops) *lane_ops = (long long)n_blocks * kProbeThreads * iters * kUnroll * kChains;
  return RCA_OK;
}

}  // extern "C"
