// SVD clutter filter of a beamformed ensemble on MI355X (gfx950).
//
// Reference: Python/clutterfilter/SVDClutterFilter.py:10-23 and
// MATLAB/lib/clutterFilter/SVDClutterFilter.m:1-45.  The ensemble is the
// Casorati matrix Y [nPix x nT] (pixels x frames, column-major: the
// beamformer's [nFrame][nPix] output as is), optionally weighted by
// 1/sqrt(mask) per pixel (.m:21).  The reference takes the SVD of the Gram
// matrix Y^H Y (.py:16, .m:23) - Hermitian positive semi-definite, so its
// singular vectors are its eigenvectors and its singular values its
// eigenvalues - keeps a band of the vectors (.py:17, .m:26-44) and returns
// Y V_k V_k^H (.py:18, .m:44).
//
// Here everything runs in FP64 on the GPU:
//   gram_f64        G = Y^H Y from the complex64 data: split-K FP64 MFMA tiles
//                   (hand-written), then a deterministic sum of the splits
//   weigh_to_c128   Y = iq / sqrt(mask) as complex128 for the projection
//   ZHEEVD          G = V diag(w) V^H, w ascending (rocSOLVER)
//   projector       P = sum over kept k of v_k v_k^H (hand-written)
//   ZGEMM           Y P, in chunks of pixels (rocBLAS)
//   to_c64          complex64 output in the input layout (hand-written)
// Projecting with P = V_k V_k^H equals (Y V_k) V_k^H and does not depend on
// the phase of each eigenvector.  When the kept or the removed set has at
// most kLowRank vectors (a clutter filter removes a few tissue vectors), the
// last four steps become one pass over the complex64 data instead:
//   gather_cols     U = the nr columns of the smaller set (hand-written)
//   project_lowrank Y U U^H, or Y - Y U U^H for the removed set (V is
//                   unitary), one lane per pixel in FP64 (hand-written)
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <climits>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "../../include/rca_das.h"
#include "../../include/rca_svd.h"

namespace rca_internal {
void set_error(const char* msg);
void clear_error();
}  // namespace rca_internal

namespace {

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  rca_internal::set_error(buf);
  return code;
}

constexpr long long kChunkPix = 1LL << 20;   // pixels per projection GEMM
constexpr int kGT = 32;                      // Gram tile: frames per side
constexpr int kGK = 64;                      // Gram: pixels per staged slice
constexpr int kGRow = 2 * kGT + 1;           // LDS row (one pixel): 64 frames + 1; the odd
                                             // stride keeps the staging stores (64 pixels of one
                                             // frame per wave) free of bank conflicts
constexpr int kMaxDevices = 64;
constexpr int kLowRank = 16;                  // largest set projected without ZGEMM

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Workspace: Y (complex128 nPix x nT), G / eigenvectors, P (nT x nT),
// eigenvalues, the eigensolver's off-diagonal scratch and status, the keep
// flags, and one projected chunk (complex128 chunk x nT).
struct WsLayout {
  size_t y, g, p, w, e, info, keep, chunk, part, total;
  long long chunk_pix;
  int nb, npairs, splits;
  long long per_split;
};

WsLayout ws_layout(long long n_pix, int n_t) {
  WsLayout L;
  size_t off = 0;
  L.y = off;
  off = align256(off + 16 * (size_t)n_pix * n_t);
  L.g = off;
  off = align256(off + 16 * (size_t)n_t * n_t);
  L.p = off;
  off = align256(off + 16 * (size_t)n_t * n_t);
  L.w = off;
  off = align256(off + 8 * (size_t)n_t);
  L.e = off;
  off = align256(off + 8 * (size_t)n_t);
  L.info = off;
  off = align256(off + sizeof(rocblas_int));
  L.keep = off;
  off = align256(off + (size_t)n_t);
  L.chunk_pix = n_pix < kChunkPix ? n_pix : kChunkPix;
  L.chunk = off;
  off = align256(off + 16 * (size_t)L.chunk_pix * n_t);
  // Gram tiles: kGT x kGT frames per workgroup over the upper triangle of
  // tile pairs, split over pixel ranges so that about 2048 workgroups run
  L.nb = (n_t + kGT - 1) / kGT;
  L.npairs = L.nb * (L.nb + 1) / 2;
  long long sp = (2048 + L.npairs - 1) / L.npairs;
  const long long max_sp = (n_pix + kGK - 1) / kGK;
  if (sp > max_sp) sp = max_sp;
  L.per_split = ((n_pix + sp - 1) / sp + kGK - 1) / kGK * kGK;
  L.splits = (int)((n_pix + L.per_split - 1) / L.per_split);
  L.part = off;
  off = align256(off + 16 * (size_t)L.splits * L.npairs * kGT * kGT);
  L.total = off;
  return L;
}

// y[t][p] = iq[t][p] / sqrt(mask[p]) in FP64 (SVDClutterFilter.m:21-22).
// Grid: x over pixels, y over frames.
__global__ void weigh_to_c128(const float2* __restrict__ iq, const float* __restrict__ mask,
                              long long n_pix, double2* __restrict__ y) {
  const long long t = blockIdx.y;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n_pix;
       p += (long long)gridDim.x * blockDim.x) {
    const float2 v = iq[t * n_pix + p];
    const double s = mask ? 1.0 / sqrt((double)mask[p]) : 1.0;
    y[t * n_pix + p] = make_double2((double)v.x * s, (double)v.y * s);
  }
}

typedef double d4 __attribute__((ext_vector_type(4)));

// Upper-triangular tile pair of index pr (row-major): bi <= bj < nb.
__host__ __device__ inline void tile_pair(int pr, int nb, int& bi, int& bj) {
  bi = 0;
  while (pr >= nb - bi) {
    pr -= nb - bi;
    ++bi;
  }
  bj = bi + pr;
}

// One workgroup: the 32 x 32 complex block G[I0.., J0..] of Y^H Y over the
// pixels [k_begin, k_end) of its split, in FP64 on the matrix cores.
// Real MFMAs on the complex product conj(a) b = (a_r b_r + a_i b_i) +
// i (a_r b_i - a_i b_r), three per step instead of four (Gauss's product:
// three sums of real products, combined once at the end).
// The mask weight 1/mask_k (= (1/sqrt(mask_k))^2) scales the b operand.
// Slices of kGK pixels x (32 + 32) frames are staged in LDS as [pixel][frame],
// so the 16 lanes that read one pixel's frames read 128 contiguous bytes.
// v_mfma_f64_16x16x4f64 lane layout (checked on MI355X with a probe of known
// products): A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in
// lane j + 16*(i%4), register i/4.
__global__ void __launch_bounds__(256) gram_f64(const float2* __restrict__ iq,
                                                const float* __restrict__ mask, long long n_pix,
                                                int n_t, int nb, long long per_split,
                                                double2* __restrict__ partial) {
  __shared__ float2 tile[kGK][kGRow];
  __shared__ double wk[kGK];
  int bi, bj;
  tile_pair(blockIdx.x, nb, bi, bj);
  const int I0 = bi * kGT, J0 = bj * kGT;
  const long long k_begin = (long long)blockIdx.y * per_split;
  const long long k_end = min(n_pix, k_begin + per_split);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w & 1, wj = w >> 1;   // the wave's 16 x 16 sub-block
  // conj(a) b with three real products (Gauss): K1 = (a_r - a_i) b_r,
  // K2 = a_r (b_i - b_r), K3 = -a_i (b_r + b_i); re = K1 - K3, im = K1 + K2
  d4 k1 = {0.0, 0.0, 0.0, 0.0}, k2 = {0.0, 0.0, 0.0, 0.0}, k3 = {0.0, 0.0, 0.0, 0.0};
  // software pipeline: the next slice's samples (kSlice per thread) and mask
  // weight are loaded into registers while the MFMAs consume the current one
  constexpr int kSlice = 2 * kGT * kGK / 256;
  float2 pre[kSlice];
  double wpre = 0.0;
  auto load_slice = [&](long long k0) {
#pragma unroll
    for (int u = 0; u < kSlice; ++u) {
      const int e = tid + 256 * u;
      const int f = e / kGK, q = e % kGK;   // local frame, pixel (coalesced along pixels)
      const int frame = f < kGT ? I0 + f : J0 + (f - kGT);
      const long long p = k0 + q;
      pre[u] = (frame < n_t && p < k_end) ? iq[(long long)frame * n_pix + p] : make_float2(0.0f, 0.0f);
    }
    if (tid < kGK) {
      const long long p = k0 + tid;
      wpre = p < k_end ? (mask ? 1.0 / (double)mask[p] : 1.0) : 0.0;
    }
  };
  if (k_begin < k_end) load_slice(k_begin);
  for (long long k0 = k_begin; k0 < k_end; k0 += kGK) {
    __syncthreads();   // the previous slice's MFMAs have read the tile
#pragma unroll
    for (int u = 0; u < kSlice; ++u) {
      const int e = tid + 256 * u;
      tile[e % kGK][e / kGK] = pre[u];
    }
    if (tid < kGK) wk[tid] = wpre;
    __syncthreads();
    if (k0 + kGK < k_end) load_slice(k0 + kGK);
#pragma unroll 4
    for (int ks = 0; ks < kGK; ks += 4) {
      const int kk = ks + (lane >> 4);
      const float2 a = tile[kk][wi * 16 + (lane & 15)];
      const float2 b = tile[kk][kGT + wj * 16 + (lane & 15)];
      const double s = wk[kk];
      const double ar = a.x, ai = a.y, br = (double)b.x * s, bim = (double)b.y * s;
      k1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar - ai, br, k1, 0, 0, 0);
      k2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bim - br, k2, 0, 0, 0);
      k3 = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, br + bim, k3, 0, 0, 0);
    }
  }
  double2* dst = partial + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * (kGT * kGT);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = wi * 16 + (lane >> 4) + 4 * r, col = wj * 16 + (lane & 15);
    dst[row * kGT + col] = make_double2(k1[r] - k3[r], k1[r] + k2[r]);
  }
}

// G (column-major n_t x n_t) = sum of the splits' blocks, both triangles.
__global__ void gram_reduce(const double2* __restrict__ partial, int npairs, int splits, int n_t,
                            int nb, double2* __restrict__ G) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)npairs * kGT * kGT) return;
  const int pr = (int)(e / (kGT * kGT)), el = (int)(e % (kGT * kGT));
  int bi, bj;
  tile_pair(pr, nb, bi, bj);
  const int i = bi * kGT + el / kGT, j = bj * kGT + el % kGT;
  if (i >= n_t || j >= n_t) return;
  double re = 0.0, im = 0.0;
  for (int s = 0; s < splits; ++s) {
    const double2 v = partial[((long long)s * npairs + pr) * (kGT * kGT) + el];
    re += v.x;
    im += v.y;
  }
  G[i + (long long)j * n_t] = make_double2(re, im);
  if (bi != bj) G[j + (long long)i * n_t] = make_double2(re, -im);
}

// P[i + j*n] = sum over kept k of V[i][a] conj(V[j][a]), a = n-1-k: the
// eigensolver returns ascending eigenvalues, keep[] counts from the largest.
__global__ void projector(const double2* __restrict__ V, const unsigned char* __restrict__ keep,
                          int n, double2* __restrict__ P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  double re = 0.0, im = 0.0;
  for (int k = 0; k < n; ++k) {
    if (!keep[k]) continue;
    const int col = n - 1 - k;
    const double2 a = V[i + (long long)col * n], b = V[j + (long long)col * n];
    re = fma(a.x, b.x, fma(a.y, b.y, re));
    im = fma(a.y, b.x, fma(-a.x, b.y, im));
  }
  P[i + (long long)j * n] = make_double2(re, im);
}

// Columns of the low-rank projection: U[s * nr + r] = V[s + col[r] * n].
struct LowRankCols {
  int col[kLowRank];
};

__global__ void gather_cols(const double2* __restrict__ V, const LowRankCols cols, int n, int nr,
                            double2* __restrict__ U) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  for (int r = 0; r < nr; ++r) U[(long long)s * nr + r] = V[s + (long long)cols.col[r] * n];
}

// Low-rank projection, one lane per pixel, FP64.  With c_r = sum_s Y[s][p]
// U[s][r] over the nr (<= NR) columns U of the smaller set:
//   kept set (sgn = +1, base 0):     out[t][p] = sum_r c_r conj(U[t][r])
//   removed set (sgn = -1, base Y):  out[t][p] = Y[t][p] - sum_r c_r conj(U[t][r])
// which are Y V_k V_k^H either way (V V^H = I).  Y = iq / sqrt(mask) is read
// from the complex64 data twice (coalesced over pixels; U is the same for
// every lane), so the kernel moves 3 x 8 bytes per sample: HBM-bound.
template <int NR>
__global__ void __launch_bounds__(256) project_lowrank(const float2* __restrict__ iq,
                                                       const float* __restrict__ mask,
                                                       long long n_pix, int n_t,
                                                       const double2* __restrict__ U, int nr,
                                                       int removed, float2* __restrict__ out) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pix) return;
  const double wt = mask ? 1.0 / sqrt((double)mask[p]) : 1.0;
  double cr[NR], ci[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) cr[r] = ci[r] = 0.0;
  const float2* x = iq + p;
  for (int t = 0; t < n_t; ++t) {
    const float2 v = x[(long long)t * n_pix];
    const double yr = v.x * wt, yi = v.y * wt;
    const double2* u = U + (long long)t * nr;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r < nr) {
        const double2 b = u[r];
        cr[r] = fma(yr, b.x, fma(-yi, b.y, cr[r]));
        ci[r] = fma(yr, b.y, fma(yi, b.x, ci[r]));
      }
    }
  }
  const double sgn = removed ? -1.0 : 1.0;
  float2* o = out + p;
  for (int t = 0; t < n_t; ++t) {
    double zr = 0.0, zi = 0.0;
    if (removed) {
      const float2 v = x[(long long)t * n_pix];
      zr = v.x * wt;
      zi = v.y * wt;
    }
    const double2* u = U + (long long)t * nr;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r < nr) {
        const double2 b = u[r];   // c_r conj(b)
        zr = fma(sgn, fma(cr[r], b.x, ci[r] * b.y), zr);
        zi = fma(sgn, fma(ci[r], b.x, -cr[r] * b.y), zi);
      }
    }
    o[(long long)t * n_pix] = make_float2((float)zr, (float)zi);
  }
}

template <int NR>
hipError_t launch_lowrank(const float2* iq, const float* mask, long long n_pix, int n_t,
                          const double2* U, int nr, int removed, float2* out, hipStream_t st) {
  hipLaunchKernelGGL(project_lowrank<NR>, dim3((unsigned)((n_pix + 255) / 256)), dim3(256), 0, st,
                     iq, mask, n_pix, n_t, U, nr, removed, out);
  return hipGetLastError();
}

__global__ void eig_descending(const double* __restrict__ w, int n, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = w[n - 1 - k];
}

// out[t][p0 + p] = c[p + t*ldc] as complex64, p < np.  Grid: x over pixels,
// y over frames.
__global__ void to_c64(const double2* __restrict__ c, long long ldc, long long p0, long long np,
                       long long n_pix, float2* __restrict__ out) {
  const long long t = blockIdx.y;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < np;
       p += (long long)gridDim.x * blockDim.x) {
    const double2 v = c[p + t * ldc];
    out[t * n_pix + p0 + p] = make_float2((float)v.x, (float)v.y);
  }
}

std::mutex g_mu;
rocblas_handle g_handle[kMaxDevices] = {};

int blas_fail(rocblas_status st, const char* what) {
  return fail(RCA_ERR_HIP, "%s: %s", what, rocblas_status_to_string(st));
}

}  // namespace

extern "C" {

size_t rca_svd_workspace_bytes(long long n_pix, int n_t) {
  if (n_pix < 1 || n_t < 1) return 0;
  return ws_layout(n_pix, n_t).total;
}

int rca_svd_filter(const void* iq, const float* mask, long long n_pix, int n_t,
                   const unsigned char* keep, void* out, double* eigval, void* workspace,
                   size_t workspace_bytes, hipStream_t stream) {
  rca_internal::clear_error();
  if (!iq || !out || !keep || !workspace)
    return fail(RCA_ERR_INVALID, "iq, out, keep and workspace must not be NULL");
  if (n_pix < 1 || n_t < 1) return fail(RCA_ERR_INVALID, "nPix and nT must be >= 1");
  if (n_pix > INT_MAX || n_t > 65535)
    return fail(RCA_ERR_UNSUPPORTED, "nPix <= 2^31-1 and nT <= 65535 are supported");
  const WsLayout L = ws_layout(n_pix, n_t);
  if (workspace_bytes < L.total)
    return fail(RCA_ERR_INVALID, "workspace of %zu bytes, %zu needed", workspace_bytes, L.total);
  int dev = 0;
  hipError_t he = hipGetDevice(&dev);
  if (he != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(RCA_ERR_HIP, "hipGetDevice failed");
  rocblas_handle h;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_handle[dev]) {
      const rocblas_status st = rocblas_create_handle(&g_handle[dev]);
      if (st != rocblas_status_success) return blas_fail(st, "rocblas_create_handle");
    }
    h = g_handle[dev];
  }
  rocblas_status st = rocblas_set_stream(h, stream);
  if (st != rocblas_status_success) return blas_fail(st, "rocblas_set_stream");

  unsigned char* ws = static_cast<unsigned char*>(workspace);
  double2* Y = reinterpret_cast<double2*>(ws + L.y);
  double2* G = reinterpret_cast<double2*>(ws + L.g);
  double2* P = reinterpret_cast<double2*>(ws + L.p);
  double* W = reinterpret_cast<double*>(ws + L.w);
  double* E = reinterpret_cast<double*>(ws + L.e);
  rocblas_int* info = reinterpret_cast<rocblas_int*>(ws + L.info);
  unsigned char* keep_d = ws + L.keep;
  double2* CH = reinterpret_cast<double2*>(ws + L.chunk);
  const int ni = (int)n_pix;

  double2* PART = reinterpret_cast<double2*>(ws + L.part);
  hipLaunchKernelGGL(gram_f64, dim3(L.npairs, L.splits), dim3(256), 0, stream,
                     static_cast<const float2*>(iq), mask, n_pix, n_t, L.nb, L.per_split, PART);
  if ((he = hipGetLastError()) != hipSuccess)
    return fail(RCA_ERR_HIP, "gram_f64: %s", hipGetErrorString(he));
  const long long nel = (long long)L.npairs * kGT * kGT;
  hipLaunchKernelGGL(gram_reduce, dim3((unsigned)((nel + 255) / 256)), dim3(256), 0, stream, PART,
                     L.npairs, L.splits, n_t, L.nb, G);
  if ((he = hipGetLastError()) != hipSuccess)
    return fail(RCA_ERR_HIP, "gram_reduce: %s", hipGetErrorString(he));
  st = rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, n_t,
                        reinterpret_cast<rocblas_double_complex*>(G), n_t, W, E, info);
  if (st != rocblas_status_success) return blas_fail(st, "rocsolver_zheevd");

  // the smaller of the kept and removed sets, as eigenvector columns (the
  // eigensolver's order is ascending: keep[k] is column n_t - 1 - k)
  int n_keep = 0;
  for (int k = 0; k < n_t; ++k) n_keep += keep[k] != 0;
  const int removed = n_keep > n_t - n_keep;
  const int nr = removed ? n_t - n_keep : n_keep;
  if (nr <= kLowRank) {
    LowRankCols cols = {};
    int r = 0;
    for (int k = 0; k < n_t; ++k)
      if ((keep[k] != 0) != (removed != 0)) cols.col[r++] = n_t - 1 - k;
    double2* U = P;   // n_t x nr <= n_t x n_t
    if (nr > 0) {
      hipLaunchKernelGGL(gather_cols, dim3((n_t + 255) / 256), dim3(256), 0, stream, G, cols, n_t,
                         nr, U);
      if ((he = hipGetLastError()) != hipSuccess)
        return fail(RCA_ERR_HIP, "gather_cols: %s", hipGetErrorString(he));
    }
    const float2* x = static_cast<const float2*>(iq);
    float2* o = static_cast<float2*>(out);
    he = nr <= 4   ? launch_lowrank<4>(x, mask, n_pix, n_t, U, nr, removed, o, stream)
         : nr <= 8 ? launch_lowrank<8>(x, mask, n_pix, n_t, U, nr, removed, o, stream)
                   : launch_lowrank<16>(x, mask, n_pix, n_t, U, nr, removed, o, stream);
    if (he != hipSuccess) return fail(RCA_ERR_HIP, "project_lowrank: %s", hipGetErrorString(he));
  } else {
    const unsigned gx = (unsigned)((n_pix + 255) / 256 < 1024 ? (n_pix + 255) / 256 : 1024);
    hipLaunchKernelGGL(weigh_to_c128, dim3(gx, n_t), dim3(256), 0, stream,
                       static_cast<const float2*>(iq), mask, n_pix, Y);
    if ((he = hipGetLastError()) != hipSuccess)
      return fail(RCA_ERR_HIP, "weigh_to_c128: %s", hipGetErrorString(he));
    if ((he = hipMemcpyAsync(keep_d, keep, (size_t)n_t, hipMemcpyHostToDevice, stream)) != hipSuccess)
      return fail(RCA_ERR_HIP, "keep upload: %s", hipGetErrorString(he));
    hipLaunchKernelGGL(projector, dim3((n_t + 63) / 64, n_t), dim3(64), 0, stream, G, keep_d, n_t, P);
    if ((he = hipGetLastError()) != hipSuccess)
      return fail(RCA_ERR_HIP, "projector: %s", hipGetErrorString(he));

    const rocblas_double_complex c1(1.0, 0.0), c0(0.0, 0.0);
    for (long long p0 = 0; p0 < n_pix; p0 += L.chunk_pix) {
      const long long np = n_pix - p0 < L.chunk_pix ? n_pix - p0 : L.chunk_pix;
      st = rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, (int)np, n_t, n_t, &c1,
                         reinterpret_cast<const rocblas_double_complex*>(Y + p0), ni,
                         reinterpret_cast<const rocblas_double_complex*>(P), n_t, &c0,
                         reinterpret_cast<rocblas_double_complex*>(CH), (int)np);
      if (st != rocblas_status_success) return blas_fail(st, "rocblas_zgemm");
      const unsigned cx = (unsigned)((np + 255) / 256 < 1024 ? (np + 255) / 256 : 1024);
      hipLaunchKernelGGL(to_c64, dim3(cx, n_t), dim3(256), 0, stream, CH, np, p0, np, n_pix,
                         static_cast<float2*>(out));
      if ((he = hipGetLastError()) != hipSuccess)
        return fail(RCA_ERR_HIP, "to_c64: %s", hipGetErrorString(he));
    }
  }
  if (eigval) {
    hipLaunchKernelGGL(eig_descending, dim3((n_t + 255) / 256), dim3(256), 0, stream, W, n_t, eigval);
    if ((he = hipGetLastError()) != hipSuccess)
      return fail(RCA_ERR_HIP, "eig_descending: %s", hipGetErrorString(he));
  }
  rocblas_int info_h = 0;
  if ((he = hipMemcpyAsync(&info_h, info, sizeof(info_h), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
      (he = hipStreamSynchronize(stream)) != hipSuccess)
    return fail(RCA_ERR_HIP, "status read-back: %s", hipGetErrorString(he));
  if (info_h != 0) return fail(RCA_ERR_HIP, "rocsolver_zheevd did not converge (info=%d)", (int)info_h);
  return RCA_OK;
}

}  // extern "C"
