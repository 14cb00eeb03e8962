// Voxel tiles, sample windows, passes and LDS layouts of the DAS kernels.
//
// The arithmetic that decides which voxels a workgroup owns, which window of
// a trace it stages in LDS for a (tilt chunk, receive line) and in how many
// passes, and where each update's taps fall.  The kernels (rca_das.hip) and
// the host-side window-cover check of the test suite both include this file,
// so that check runs this code rather than a restatement of it.
//
// Invariant the kernels rely on: for every update whose sample position m is
// in range (LO <= m < Ns-1-HI), the taps m-LO .. m+HI lie inside the rows
// staged for the pass that takes it.  The window is bounded from the tile's
// bounding box (transmit distances over the chunk's tilts, receive distances
// over the line's aperture) with one sample of margin on each side.
#ifndef RCA_TILE_H
#define RCA_TILE_H

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

namespace rca_tile {

constexpr int kThreads = 256;
constexpr int kMaxLds = 64 * 1024;
// Longest trace (samples) the kernels accept: sample offsets are clamped to
// +-2^28 and summed in 32-bit unsigned arithmetic (split_pos).
constexpr int kMaxSamples = 1 << 27;

// Forward tiles: columns per wave (the 4 waves stacked in depth), measured
// best of 4, 8, 16 on 128x128x256 volumes.
#ifndef RCA_TILE_COLS
#define RCA_TILE_COLS 16
#endif
constexpr int kTileCols = RCA_TILE_COLS;

// Transmit terms of one (voxel, tilt), cached in LDS for a chunk of tilts.
struct alignas(16) TxTerm {
  int bi;      // transmit sample offset fs*((|P-VS| - d0)/c + t0) = bi + bf,
  float bf;    //   bi = floor, bf in [0, 1)
  float2 et;   // e^{+j 2 pi frac(Bp)}, Bp = fd*((|P-VS| - d0 - 2z)/c + t0)
};

// ----------------------------------------------------------------------------
// Voxel tiles.  A workgroup has 256 lanes x V voxels.  With a depth-run hint
// (the grid's Nz >= 64, voxel p = iz + Nz*column) it owns a tile of
// neighbouring depths and columns: each wave holds 64/CW depths of CW
// neighbouring columns, DW of the 4 waves stacked in depth (4/DW side by side
// in columns), so a tile is DW*64/CW depths x CW*(4/DW)*V columns.  Without
// the hint a workgroup owns 256V consecutive voxels.
// ----------------------------------------------------------------------------
template <int CW, int DW>
__host__ __device__ __forceinline__ long long voxel_of(int depth_run, long long n_pix, unsigned bx,
                                                       int V, int v, int tid, bool& valid) {
  constexpr int WD = 64 / CW, TD = DW * WD, WC = 4 / DW;
  if (depth_run >= 64) {
    const long long ndc = (depth_run + TD - 1) / TD;
    const long long dc = bx % ndc, cg = bx / ndc;
    const int l = tid & 63, w = tid >> 6;
    const long long col = cg * (CW * WC * V) + v * (CW * WC) + (w / DW) * CW + (l % CW);
    const long long dep = dc * TD + (w % DW) * WD + l / CW;
    const long long p = dep + (long long)depth_run * col;
    valid = dep < depth_run && p < n_pix;
    return p;
  }
  const long long p = (long long)bx * (kThreads * V) + (long long)v * kThreads + tid;
  valid = p < n_pix;
  return p;
}

template <int CW, int DW>
__host__ __device__ inline long long tile_blocks(long long n_pix, int depth_run, int V) {
  constexpr int TD = DW * (64 / CW), TC = CW * (4 / DW);
  if (depth_run >= 64) {
    const long long ncol = (n_pix + depth_run - 1) / depth_run;
    return ((ncol + TC * V - 1) / (TC * V)) * ((depth_run + TD - 1) / TD);
  }
  return (n_pix + kThreads * V - 1) / (kThreads * V);
}

__host__ __device__ inline long long fwd_blocks(long long n_pix, int depth_run, int V) {
  return tile_blocks<kTileCols, 4>(n_pix, depth_run, V);
}

// Bounding box of a workgroup's valid voxels: receive-axis coordinate (lat),
// transmit-axis coordinate (txl), depth; and the largest receive aperture
// half-width z/F.
struct TileBox {
  float lat_lo, lat_hi;
  float txl_lo, txl_hi;
  float zlo, zhi;
  double apmax;
};

// |t| over t in [lo, hi]: smallest and largest.
__host__ __device__ __forceinline__ void abs_range(double lo, double hi, double& amin, double& amax) {
  amin = (lo <= 0.0 && hi >= 0.0) ? 0.0 : fmin(fabs(lo), fabs(hi));
  amax = fmax(fabs(lo), fabs(hi));
}

// Transmit bounds of one tilt over a box (transmit-axis range [l_lo, l_hi],
// depth range [z_lo, z_hi]): the sample offset Bu = fs*((|P-VS| - d0)/c + t0)
// grows with the distance, so its extremes come from the extreme distances.
// kfs = fs/c, ut0 = fs*t0.
__host__ __device__ __forceinline__ void tx_bounds(double kfs, double ut0, double l_lo, double l_hi,
                                                   double z_lo, double z_hi, float vl, float vz,
                                                   double d0, float& bmin, float& bmax) {
  double l0, l1, z0, z1;
  abs_range(l_lo - (double)vl, l_hi - (double)vl, l0, l1);
  abs_range(z_lo - (double)vz, z_hi - (double)vz, z0, z1);
  bmin = (float)fma(kfs, sqrt(l0 * l0 + z0 * z0) - d0, ut0);
  bmax = (float)fma(kfs, sqrt(l1 * l1 + z1 * z1) - d0, ut0);
}

// Receive bounds of line (e_lat, ezd) over the box, in samples, including a
// per-line delay uabc (samples).  Returns false when the line's aperture
// misses every voxel (min |lateral| >= max z/F): such a line contributes
// nothing.
__host__ __device__ __forceinline__ bool rx_bounds(double kfs, const TileBox& b, double e_lat,
                                                   double ezd, double uabc, float& r_lo,
                                                   float& r_hi) {
  double A0, A1, Z0, Z1;
  abs_range((double)b.lat_lo - e_lat, (double)b.lat_hi - e_lat, A0, A1);
  abs_range((double)b.zlo - ezd, (double)b.zhi - ezd, Z0, Z1);
  if (!(A0 < b.apmax)) return false;
  A1 = fmin(A1, b.apmax);
  r_lo = (float)(kfs * sqrt(A0 * A0 + Z0 * Z0) + uabc);
  r_hi = (float)(kfs * sqrt(A1 * A1 + Z1 * Z1) + uabc);
  return true;
}

// Sample window [s0, whi) from the range [lo, hi] of sample positions: taps
// floor(u) - LO .. floor(u) + HI, one sample of margin on each side for
// rounding, s0 a multiple of E (samples per 16-byte chunk), clamped to the
// trace.
__host__ __device__ __forceinline__ void window_from(int ns, int LO, int HI, int E, float lo,
                                                     float hi, int& s0, int& whi) {
  const int w_lo = (int)fminf(fmaxf(0.0f, floorf(lo) - (float)(LO + 1)), (float)ns);
  whi = (int)fmaxf(fminf((float)ns, floorf(hi) + (float)(HI + 2)), 0.0f);
  s0 = (w_lo / E) * E;
}

// Passes over a window longer than the rows: rows hold wcap samples; pass k
// stages [s0 + k*step, min(s0 + k*step + wcap, whi)) and takes the updates
// whose first tap lies in [s0 + k*step, s0 + (k+1)*step) (the last pass also
// those above).  An update whose first tap lies below s0 belongs to no pass:
// the window bounds leave none (the host emulation of the launch layouts
// counts any as a violation).  step is at least one 16-byte chunk so the pass
// loop always ends (rows shorter than NT + E only occur with traces too short
// for any tap).
__host__ __device__ __forceinline__ int pass_step(int wcap, int NT, int E) {
  const int s = (wcap - NT + 1) / E * E;
  return s > E ? s : E;
}
__host__ __device__ __forceinline__ int pass_count(int s0, int whi, int wcap, int step) {
  return whi - s0 > wcap ? 1 + (whi - s0 - wcap + step - 1) / step : 1;
}
// (The kernels test the range of .cu:323 and the pass together, as one
// unsigned compare base - cb < hi - cb with hi = last ? ns - NT :
// min(ns - NT, cb + step); this is the host emulation's form.)
__host__ __device__ __forceinline__ bool in_pass(int base, int cb, bool last, int step) {
  return base >= cb && (last || base < cb + step);
}

// x = i + f with i = floor(x) and f = x - i in [0, 1], computed in FP64.
// Sample positions are summed from such parts in integer + FP32 arithmetic:
// the fractions are exact to 2^-25 samples each, far below what the weights
// and the range test resolve.  Offsets beyond +-2^28 samples (and NaN) become
// exactly +-2^28 (NaN: -2^28) with f = 0: out of every trace (traces hold at
// most 2^27 samples), and a sum of up to three such integer parts plus the
// transmit sentinel stays inside 32 bits, which the kernels' unsigned range
// test relies on.
constexpr double kPosClamp = 268435456.0;   // 2^28
__host__ __device__ __forceinline__ void split_pos(double x, int& i, float& f) {
  const double fl = floor(x);
  const double cl = fmin(fmax(fl, -kPosClamp), kPosClamp);
  i = (int)cl;
  f = fl == cl ? (float)(x - fl) : 0.0f;
}

// m = floor(u), d = u - m for u = (bi + bf) + (ri + rf) [+ (ai + af)].
__host__ __device__ __forceinline__ void sample_pos(int i, float f, int& m, float& d) {
  const float fl = floorf(f);
  m = i + (int)fl;
  d = f - fl;
}

// ----------------------------------------------------------------------------
// LDS layouts.
// ----------------------------------------------------------------------------
// Most tilts per chunk: rows of TC x FB windows fit 32 KB at 256 samples.
#ifndef RCA_TC_MAX
#define RCA_TC_MAX 8
#endif
__host__ __device__ constexpr int tc_max(int fb, int sbytes) {
  const int r = (sbytes == 8 ? 16 : 32) / fb;
  return r > RCA_TC_MAX ? RCA_TC_MAX : (r < 1 ? 1 : r);
}

// Forward: tilts per chunk (tc) and samples per row (wcap).  Per tilt of a
// chunk the workgroup holds V x 256 transmit terms and FB rows.  Depth tiles
// stage short windows (rows of kTileRow samples; longer windows take several
// passes); otherwise rows hold whole traces (rounded up to 16 bytes).  tc is
// what fits the budget, at least 1.
constexpr int kTileRow = kTileCols >= 8 ? 128 : 256;
// LDS budget of a forward workgroup (measured on 128x128x256 volumes):
// single-frame blocks run best at 32 KB (6 tilts per chunk), multi-frame
// blocks at 48 KB.
constexpr size_t kFwdLds = 48 * 1024, kFwdLds1 = 32 * 1024;

struct FwdLayout { int tc, wcap; size_t lds; };
inline FwdLayout fwd_layout(int depth_run, int n_sample, int V, int fb, size_t sbytes, int tc_kernel,
                            size_t budget, int tile_row = kTileRow) {
  const int e = (int)(16 / sbytes);
  const int full = (n_sample + e - 1) / e * e;
  FwdLayout L;
  L.wcap = depth_run && full > tile_row ? tile_row : full;
  const size_t per_tilt = (size_t)V * kThreads * sizeof(TxTerm) + (size_t)fb * L.wcap * sbytes;
  const long long fit = (long long)(budget / per_tilt);
  L.tc = (int)(fit < 1 ? 1 : fit > tc_kernel ? tc_kernel : fit);
  L.lds = (size_t)L.tc * per_tilt;
  return L;
}

// Adjoint: per tilt of a chunk 256 V transmit terms and, for each of the 4
// waves, FB rows of complex FP32 accumulators (interleaved re, im); then a
// claim slot per row sample for each wave.  Depth tiles (64 depths x 4V
// columns, V voxels per lane in neighbouring columns) use rows of kAdjTileRow
// samples, other tiles whole traces; rows are capped so one tilt fits 64 KB
// (longer windows take several passes).
constexpr int kAdjTileRow = 192;
constexpr size_t kAdjLds = 48 * 1024;

struct AdjLayout { int tc, wcap; size_t lds; };
inline AdjLayout adj_layout(int depth_run, int n_sample, int fb, int tc_kernel, size_t budget,
                            int V = 1) {
  AdjLayout L;
  const size_t txc = (size_t)kThreads * V * sizeof(TxTerm);   // transmit terms of one tilt
  const int w0 = depth_run && n_sample > kAdjTileRow ? kAdjTileRow : n_sample;
  const int wmax = (int)((kMaxLds - txc) / (4 * 8 * (size_t)fb + 4 * 4));
  L.wcap = w0 < wmax ? w0 : wmax;
  const size_t per_tilt = txc + 4 * 8 * (size_t)fb * L.wcap;
  const size_t claims = 4 * 4 * (size_t)L.wcap;
  const long long fit = budget > claims ? (long long)((budget - claims) / per_tilt) : 0;
  L.tc = (int)(fit < 1 ? 1 : fit > tc_kernel ? tc_kernel : fit);
  L.lds = L.tc * per_tilt + claims;
  return L;
}

}  // namespace rca_tile

#endif  // RCA_TILE_H
