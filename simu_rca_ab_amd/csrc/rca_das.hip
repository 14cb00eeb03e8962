// RCA delay-and-sum I/Q beamformer for MI355X (gfx950).
//
// What is computed is the reference kernel
// MATLAB/lib/beamformer/RCA/bfVirtualSourceGeneric_RCA.cu:103-391: for every
// voxel p, tilt t and receive line c,
//   tau  = (|P - VS_t|_(receive-axis zeroed) - d0_t)/c + hypot(A, z - ez_c)/c + t0
//   u    = tau * fs,  m = floor(u),  delta = u - m
//   out += w(A/z*F) * sum_k q_k(delta) * IQ[t][c][m+k] * exp(+j 2 pi fd (tau - 2 z / c))
// with A = |x - ex_c| (receive on X) or |y - ey_c|, a Hann weight inside the
// f-number aperture A < z/F, and quadratic or Lanczos weights q_k.
//
// How it is computed differs from the reference (one thread per
// voxel/tilt/frame doing global gathers, fast-math FP32 geometry):
//   * A workgroup of 256 lanes owns 256*V consecutive voxels (a depth run:
//     the grid is depth-fastest) and FB frames.  It walks tilt -> line; for
//     each (tilt, line) the FB I/Q traces are staged once in LDS and every
//     lane gathers its 2a interpolation taps from LDS.
//   * Geometry is FP64 (MI355X's vector FP64 peak, 78.6 TFLOP/s, is half its
//     FP32 peak, so it is kept out of the per-update work): transmit delay per
//     (voxel, tilt), receive distance per (voxel, line).  The sample position u and
//     the demodulation phase are range-reduced in FP64 (u - floor(u),
//     p - rint(p) in cycles, with the 4 pi fd z / c term cancelled
//     analytically) before anything is rounded to FP32, so FP32 carries only
//     quantities in [0,1).  Interpolation weights, the rotation and the
//     accumulation are FP32.
//   * Lanczos weights are evaluated as pairs of even/odd polynomials around
//     the half-sample point (46 FMA-class instructions for 8 taps) instead of
//     2a sinpi calls and divides.
//   * The tilt sum (compounding) happens in registers; one output write per
//     voxel and frame.
//   * FP16-I/Q traces (half2 + FP32 per-trace scale) halve LDS and HBM bytes;
//     the math stays FP32 (mixed FMA).
//
// The adjoint (transposed interpolation) scatters into an LDS copy of each
// trace with LDS float atomics and flushes once per block.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>

#include "../../include/rca_das.h"
#include "rca_interp_poly.h"
#include "rca_math.h"
#include "rca_tile.h"

namespace {

using rca_math::sincos2pi;
using rca_math::sqrt_nr;
using rca_tile::AdjLayout;
using rca_tile::FwdLayout;
using rca_tile::TileBox;
using rca_tile::TxTerm;
using rca_tile::kMaxLds;
using rca_tile::kMaxSamples;
using rca_tile::kThreads;
using rca_tile::kTileCols;
using rca_tile::pass_count;
using rca_tile::pass_step;
using rca_tile::sample_pos;
using rca_tile::split_pos;
using rca_tile::tc_max;

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(RCA_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace

// Shared with the other translation units of the library (rca_svd.hip).
namespace rca_internal {
void set_error(const char* msg) { g_err = msg; }
void clear_error() { g_err.clear(); }
}  // namespace rca_internal

namespace {

// Per-launch arguments (kernarg).  Geometry pointers come from rca_das_desc;
// the FP64 scalars are derived on the host from the FP32 parameters.
// Aberration-law instantiations of the forward kernel.
constexpr int kLawLine = 0;     // no law, or one delay per receive line
constexpr int kLawTxLine = 1;   // one delay per (tilt, line)
constexpr int kLawPixel = 2;    // receive and transmit delays per (pixel, line)

struct KArgs {
  const float *gx, *gy, *gz, *ex, *ey, *ez, *vx, *vy, *vz, *d0, *ab;
  const float* abt;    // PIXEL_CHANNEL transmit-side law (ab is the receive side)
  const void* iq;
  const float* scale;
  float2* out;
  float* mask;
  const float2* gout;  // adjoint input
  float2* gin;         // adjoint output
  double kfs, kfd;     // fs/c, fd/c
  double ut0, pt0;     // fs*t0, fd*t0
  double fs, fd;
  float apx, apy, apz, fnum;
  int n_pix, n_ch, n_tx, n_sample, n_frame;
  int receive_x, mask_tx, ab_mode, compound, accumulate, vec_ok, nosum;
  int gout_tx;
  int depth_run, wcap;   // voxel tiling hint; LDS samples per row
  int tc;                // tilts per chunk of the forward kernel
  int adj_lo;            // adjoint: tiles whose columns share the transmit terms go to das_adj_lo
  // deterministic adjoint: 64-bit fixed-point sums instead of float atomics
  unsigned long long* gdet;   // [2 x gradient samples], or null
  const unsigned* gdet_max;   // max over grad_out of max(|re|, |im|), as float bits
};

// ----------------------------------------------------------------------------
// Device math.
// ----------------------------------------------------------------------------

// Hann weight 0.5*(1 + cos(pi*rho)) = cos^2(pi*rho/2) for rho in [0, 1).
__device__ __forceinline__ float hann(float rho) {
  const float t = rho * 1.57079632679490f;
  const float t2 = t * t;
  float c = fmaf(t2, -1.14707456e-11f, 2.08767570e-9f);
  c = fmaf(t2, c, -2.75573192e-7f);
  c = fmaf(t2, c, 2.48015873e-5f);
  c = fmaf(t2, c, -1.38888889e-3f);
  c = fmaf(t2, c, 4.16666667e-2f);
  c = fmaf(t2, c, -0.5f);
  c = fmaf(t2, c, 1.0f);
  return c * c;
}

// Interpolation kinds: taps k = -LO..HI around m = floor(u); W = window divisor.
template <int KIND> struct Interp;
template <> struct Interp<RCA_INTERP_QUADRATIC> { static constexpr int LO = 1, HI = 1, W = 0; };
template <> struct Interp<RCA_INTERP_LANCZOS_REF> { static constexpr int LO = 2, HI = 3, W = 2; };
template <> struct Interp<RCA_INTERP_LANCZOS3> { static constexpr int LO = 2, HI = 3, W = 3; };
template <> struct Interp<RCA_INTERP_LANCZOS4> { static constexpr int LO = 3, HI = 4, W = 4; };

// Weights q[k+LO] = sinc(d-k) * sinc((d-k)/W), d in [0, 1)
// (bfVirtualSourceGeneric_RCA.cu:79-82, 331-336 for W = 2).
//
// With e = d - 1/2 the taps pair up, q_{1-p}(e) = q_p(-e): the even and odd
// parts of q_p are polynomials in u = e^2 (rca_interp_poly.h, fitted
// near-minimax on |e| <= 1/2; degree 3 for a = 4, max |error| 4.3e-7 in FP32,
// degree 4 for the 6-tap kinds, 1.7e-7), stored in monic form lead * M(u).
// M is evaluated by Horner steps whose coefficients are 32-bit literals: an
// add (u + a_top, one register operand), then v_fmaak_f32 (p * u + a_j); the
// leads fold into one multiply for the even part and one for e of the odd
// part, and each tap is one FMA: q_{p,1-p} = fma(+-e*leadO, M_O, leadE*M_E).
// A pair costs 2 adds, 2 (DEG - 1) FMAs, 2 multiplies and 2 FMAs: 40
// instructions for the 8 taps of a = 4 (16 of them with one register
// operand, which issue at twice the rate of two- and three-operand ones on
// gfx950), against 46 for the degree-4 Horner form with the leading
// coefficient moved into a register and 95 for the sine-identity form with
// one reciprocal per update.  Left to itself the compiler keeps the
// loop-invariant constants in SGPRs, and an FMA with an SGPR source issues at
// half rate: hence the literals, in inline asm (the multiply of e by the odd
// lead stays plain C, which gets its literal from the compiler).
template <unsigned K>
__device__ __forceinline__ float fma_lit(float x, float y) {
  float r;
  asm("v_fmaak_f32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(y), "i"(K));
  return r;
}
template <unsigned K>
__device__ __forceinline__ float add_lit(float x) {
  float r;
  asm("v_add_f32 %0, %1, %2" : "=v"(r) : "i"(K), "v"(x));
  return r;
}
template <unsigned K>
__device__ __forceinline__ float mul_lit(float x) {
  float r;
  asm("v_mul_f32 %0, %1, %2" : "=v"(r) : "i"(K), "v"(x));
  return r;
}
template <int KIND, int P, bool ODD>
constexpr unsigned coef_bits(int j) {
  return __builtin_bit_cast(unsigned, ODD ? InterpPoly<KIND>::kOdd[P - 1][j]
                                          : InterpPoly<KIND>::kEven[P - 1][j]);
}
template <int KIND, int P, bool ODD>
constexpr unsigned lead_bits() {
  return __builtin_bit_cast(unsigned, ODD ? InterpPoly<KIND>::kOddLead[P - 1]
                                          : InterpPoly<KIND>::kEvenLead[P - 1]);
}
// p * u + a_J, then the lower coefficients down to a_0.
template <int KIND, int P, bool ODD, int J>
__device__ __forceinline__ float horner_tail(float u, float p) {
  if constexpr (J < 0) return p;
  else return horner_tail<KIND, P, ODD, J - 1>(u, fma_lit<coef_bits<KIND, P, ODD>(J)>(u, p));
}
// M(u) = u^D + a_{D-1} u^{D-1} + ... + a_0.
template <int KIND, int P, bool ODD>
__device__ __forceinline__ float horner_monic(float u) {
  constexpr int D = InterpPoly<KIND>::kDeg;
  return horner_tail<KIND, P, ODD, D - 2>(u, add_lit<coef_bits<KIND, P, ODD>(D - 1)>(u));
}

// lead * M(u) (LEAD) or M(u) as one asm statement.  The compiler's hazard
// check treats every inline asm as a possible 16-bit dst_sel forwarder and
// puts a wait state between an asm result and its next-instruction consumer;
// plain 32-bit VALU results need none, so a chain written as one statement
// issues without those s_nop (14 -> 1 per update in the config-1 forward,
// +0.7 % measured).  Same instructions in the same order as horner_monic
// (+ mul_lit): bitwise the same weights.
template <int KIND, int P, bool ODD, bool LEAD>
__device__ __forceinline__ float horner_block(float u) {
  constexpr int D = InterpPoly<KIND>::kDeg;
  if constexpr (D != 3 && D != 4) {
    const float m = horner_monic<KIND, P, ODD>(u);
    return LEAD ? mul_lit<lead_bits<KIND, P, ODD>()>(m) : m;
  } else {
    float r;
    if constexpr (D == 3 && LEAD) {
      asm("v_add_f32 %0, %2, %1\n\tv_fmaak_f32 %0, %1, %0, %3\n\tv_fmaak_f32 %0, %1, %0, %4\n\t"
          "v_mul_f32 %0, %5, %0"
          : "=&v"(r)
          : "v"(u), "i"(coef_bits<KIND, P, ODD>(2)), "i"(coef_bits<KIND, P, ODD>(1)),
            "i"(coef_bits<KIND, P, ODD>(0)), "i"(lead_bits<KIND, P, ODD>()));
    } else if constexpr (D == 3) {
      asm("v_add_f32 %0, %2, %1\n\tv_fmaak_f32 %0, %1, %0, %3\n\tv_fmaak_f32 %0, %1, %0, %4"
          : "=&v"(r)
          : "v"(u), "i"(coef_bits<KIND, P, ODD>(2)), "i"(coef_bits<KIND, P, ODD>(1)),
            "i"(coef_bits<KIND, P, ODD>(0)));
    } else if constexpr (LEAD) {
      asm("v_add_f32 %0, %2, %1\n\tv_fmaak_f32 %0, %1, %0, %3\n\tv_fmaak_f32 %0, %1, %0, %4\n\t"
          "v_fmaak_f32 %0, %1, %0, %5\n\tv_mul_f32 %0, %6, %0"
          : "=&v"(r)
          : "v"(u), "i"(coef_bits<KIND, P, ODD>(3)), "i"(coef_bits<KIND, P, ODD>(2)),
            "i"(coef_bits<KIND, P, ODD>(1)), "i"(coef_bits<KIND, P, ODD>(0)),
            "i"(lead_bits<KIND, P, ODD>()));
    } else {
      asm("v_add_f32 %0, %2, %1\n\tv_fmaak_f32 %0, %1, %0, %3\n\tv_fmaak_f32 %0, %1, %0, %4\n\t"
          "v_fmaak_f32 %0, %1, %0, %5"
          : "=&v"(r)
          : "v"(u), "i"(coef_bits<KIND, P, ODD>(3)), "i"(coef_bits<KIND, P, ODD>(2)),
            "i"(coef_bits<KIND, P, ODD>(1)), "i"(coef_bits<KIND, P, ODD>(0)));
    }
    return r;
  }
}

// q_p and q_{1-p} from the even and odd parts of pair p.
template <int KIND, int P>
__device__ __forceinline__ void weight_pair(float e, float u, float* q) {
  constexpr int LO = Interp<KIND>::LO;
  const float pe = horner_block<KIND, P, false, true>(u);
  const float mo = horner_block<KIND, P, true, false>(u);
  // plain C: the compiler gives it the literal operand by itself, and an asm
  // result here would cost the next instruction a hazard wait state
  const float eo = e * InterpPoly<KIND>::kOddLead[P - 1];
  q[P + LO] = fmaf(eo, mo, pe);
  q[1 - P + LO] = fmaf(-eo, mo, pe);
  if constexpr (P > 1) weight_pair<KIND, P - 1>(e, u, q);
}

template <int KIND>
__device__ __forceinline__ void interp_weights(float d, float* q) {
  constexpr int LO = Interp<KIND>::LO, HI = Interp<KIND>::HI;
  if constexpr (KIND == RCA_INTERP_QUADRATIC) {
    // bfVirtualSourceGeneric_RCA.cu:355-357
    q[0] = d * (d - 1.0f) * 0.5f;
    q[1] = -(d - 1.0f) * (d + 1.0f);
    q[2] = d * (d + 1.0f) * 0.5f;
  } else {
    static_assert(HI == LO + 1 && InterpPoly<KIND>::kDeg >= 2, "taps pair as p and 1 - p");
    const float e = d - 0.5f;
    weight_pair<KIND, HI>(e, e * e, q);
  }
}

// Transmit insonification cone (bfVirtualSourceGeneric_RCA.cu:85-98, 191-214),
// in FP64.  The reference compares angles, acos(cos_p) <= acos(cos_cone); acos
// is decreasing on [-1, 1] and NaN outside it, so the same test reads
// cos_p in [-1, 1], cos_cone in [-1, 1] and cos_p >= cos_cone, without acos.
__device__ __forceinline__ double cos_ap(float ax, float ay, float bx, float by, float cx,
                                         float cy) {
  const double abx = (double)bx - ax, aby = (double)by - ay;
  const double acx = (double)cx - ax, acy = (double)cy - ay;
  return (abx * acx + aby * acy) / sqrt((abx * abx + aby * aby) * (acx * acx + acy * acy));
}

__device__ __forceinline__ bool within_cone(double c, double cone) {
  return c >= -1.0 && c <= 1.0 && cone >= -1.0 && cone <= 1.0 && c >= cone;
}

__device__ __forceinline__ bool tx_mask(const KArgs& a, float vx, float vy, float vz, float px,
                                        float py, float pz) {
  const float hx = 0.5f * a.apx, hy = 0.5f * a.apy, az = a.apz;
  const double coneX = cos_ap(vx, vz, -hx, -az, hx, -az);
  const double coneY = cos_ap(vy, vz, -hy, -az, hy, -az);
  return within_cone(cos_ap(vx, vz, -hx, -az, px, pz), coneX) &&
         within_cone(cos_ap(vx, vz, hx, -az, px, pz), coneX) &&
         within_cone(cos_ap(vy, vz, -hy, -az, py, pz), coneY) &&
         within_cone(cos_ap(vy, vz, hy, -az, py, pz), coneY);
}

// Sample access: complex64 or half2 (the per-trace scale is applied to the
// rotation weights by the caller).
__device__ __forceinline__ float2 ld_sample(const float2* s, int i) { return s[i]; }
__device__ __forceinline__ float2 ld_sample(const __half2* s, int i) { return __half22float2(s[i]); }

// ----------------------------------------------------------------------------
// Voxel tiles (rca_tile.h).  The forward kernel stacks its 4 waves in depth
// with CW = kTileCols (16 depths x 16 columns): neighbouring lanes read nearby
// (often equal) samples, which spreads its LDS reads over the banks.  The
// adjoint puts one column in each wave (64 depths x 4 columns), so the lanes
// of a wave scatter to distinct samples.
// ----------------------------------------------------------------------------
template <int CW, int DW>
__device__ __forceinline__ long long voxel_of(const KArgs& a, unsigned bx, int V, int v, int tid,
                                              bool& valid) {
  return rca_tile::voxel_of<CW, DW>(a.depth_run, a.n_pix, bx, V, v, tid, valid);
}

// Values every lane of the workgroup holds alike, moved to scalar registers.
__device__ __forceinline__ float uni(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double uni(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Lane j's value of x, in scalar registers (j a compile-time constant).
__device__ __forceinline__ float lane_val(float x, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
}
__device__ __forceinline__ double lane_val(double x, int j) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), j);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Coordinates of a workgroup's voxels and their bounding box.
template <int V, int CW, int DW>
__device__ __forceinline__ TileBox tile_voxels(const KArgs& a, unsigned bx, int tid, bool rx,
                                               double* pl, double* pt, double* pz, double* apF,
                                               double* iapF, bool* valid, float (*s_box)[6]) {
  float bb[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const long long p = voxel_of<CW, DW>(a, bx, V, v, tid, valid[v]);
    const long long pc = valid[v] ? p : 0;
    const float fx = a.gx[pc], fy = a.gy[pc], fz = a.gz[pc];
    pl[v] = rx ? fx : fy;
    pt[v] = rx ? fy : fx;
    pz[v] = fz;
    apF[v] = pz[v] / (double)a.fnum;   // .cu:174
    iapF[v] = (double)a.fnum / pz[v];
    if (valid[v]) {
      bb[0] = fminf(bb[0], fx); bb[1] = fmaxf(bb[1], fx);
      bb[2] = fminf(bb[2], fy); bb[3] = fmaxf(bb[3], fy);
      bb[4] = fminf(bb[4], fz); bb[5] = fmaxf(bb[5], fz);
    }
  }
  // wave shuffles, then LDS
#pragma unroll
  for (int k = 0; k < 6; ++k) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float y = __shfl_xor(bb[k], o, 64);
      bb[k] = (k & 1) ? fmaxf(bb[k], y) : fminf(bb[k], y);
    }
  }
  if ((tid & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) s_box[tid >> 6][k] = bb[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    float r = s_box[0][k];
#pragma unroll
    for (int w = 1; w < 4; ++w) r = (k & 1) ? fmaxf(r, s_box[w][k]) : fminf(r, s_box[w][k]);
    bb[k] = uni(r);
  }
  TileBox b;
  b.lat_lo = rx ? bb[0] : bb[2];
  b.lat_hi = rx ? bb[1] : bb[3];
  b.txl_lo = rx ? bb[2] : bb[0];
  b.txl_hi = rx ? bb[3] : bb[1];
  b.zlo = bb[4];
  b.zhi = bb[5];
  b.apmax = uni(a.fnum > 0.0f ? (double)b.zhi / (double)a.fnum : (double)b.zlo / (double)a.fnum);
  return b;
}

// Transmit sample offset stored for a (voxel, tilt) outside the transmit
// cone when OFF: far beyond any trace, so every sample position built on it
// fails the forward's range test.  Real offsets are clamped to +-2^28
// (split_pos), so the sentinel plus a receive offset and a per-tilt law
// offset is at least 2^29 - 2^27 - NT, above every trace length (<= 2^27),
// and at most 2^30 + 2^29 + 3, with no 32-bit wrap.
constexpr int kTxOff = 1 << 30;

// Transmit terms of tilts [t0, t0 + nt) for the lane's voxels into the LDS
// cache ([j][v][lane]); bit j of act[v] = tilt t0 + j insonifies voxel v
// (.cu:191-214).
template <int V, bool OFF = false>
__device__ __forceinline__ void chunk_tx(const KArgs& a, int t0, int nt, bool rx, const double* pl,
                                         const double* pt, const double* pz, const bool* valid,
                                         TxTerm* txc, int tid, unsigned* act) {
#pragma unroll
  for (int v = 0; v < V; ++v) act[v] = 0u;
  for (int j = 0; j < nt; ++j) {
    const int t = t0 + j;
    const float vxf = a.vx[t], vyf = a.vy[t], vzf = a.vz[t];
    const double d0 = a.d0[t];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      if (!valid[v]) continue;
      // .cu:256-269: |P - VS| with the receive-axis offset dropped
      const double dl = pt[v] - (double)(rx ? vyf : vxf), dz = pz[v] - (double)vzf;
      const double tt = sqrt_nr(fma(dl, dl, dz * dz)) - d0;
      const double Bp = fma(a.kfd, fma(-2.0, pz[v], tt), a.pt0);
      float sn, cs;
      sincos2pi((float)(Bp - rint(Bp)), sn, cs);
      TxTerm x;
      split_pos(fma(a.kfs, tt, a.ut0), x.bi, x.bf);
      x.et = make_float2(cs, sn);
      // (stored before the cone test unless it may change the term: the
      // term's registers then need not live across the test)
      if constexpr (!OFF) txc[(j * V + v) * kThreads + tid] = x;
      const float fx = (float)(rx ? pl[v] : pt[v]), fy = (float)(rx ? pt[v] : pl[v]);
      if (!a.mask_tx || tx_mask(a, vxf, vyf, vzf, fx, fy, (float)pz[v])) act[v] |= 1u << j;
      else if constexpr (OFF) x.bi = kTxOff;
      if constexpr (OFF) txc[(j * V + v) * kThreads + tid] = x;
    }
  }
}

// Range of the transmit sample offset over the box for tilts [t0, t0 + nt)
// (lane L < nt takes tilt t0 + L, then a wave reduction).
__device__ __forceinline__ void chunk_tx_bounds(const KArgs& a, const TileBox& b, int t0, int nt,
                                                int lane, bool rx, float& bu_lo, float& bu_hi) {
  bu_lo = INFINITY;
  bu_hi = -INFINITY;
  if (lane < nt) {
    const int t = t0 + lane;
    rca_tile::tx_bounds(a.kfs, a.ut0, b.txl_lo, b.txl_hi, b.zlo, b.zhi, rx ? a.vy[t] : a.vx[t],
                        a.vz[t], a.d0[t], bu_lo, bu_hi);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    bu_lo = fminf(bu_lo, __shfl_xor(bu_lo, o, 64));
    bu_hi = fmaxf(bu_hi, __shfl_xor(bu_hi, o, 64));
  }
  bu_lo = uni(bu_lo);
  bu_hi = uni(bu_hi);
}

// Sample window [s0, whi) of line c for the chunk's tilts over the box: taps
// floor(u) - LO .. floor(u) + HI, one sample of margin on each side for
// rounding, s0 a multiple of E.  Returns false when the line's aperture
// misses every voxel (min |lateral| >= max z/F): such a line contributes
// nothing.
// PIXEL_CHANNEL law of voxel p on line c in seconds: the receive-side delay
// joins the receive time of flight and the transmit-side one the total delay
// (bfVirtualSourceGenericTxRx.cu:268-275, 292-294); either array may be absent.
__device__ __forceinline__ double pixel_law(const KArgs& a, int c, long long p) {
  const long long i = (long long)c * a.n_pix + p;
  double l = 0.0;
  if (a.ab) l = (double)a.ab[i];
  if (a.abt) l += (double)a.abt[i];
  return l;
}

// Range, in samples, of the PIXEL_CHANNEL law over the workgroup's voxels
// that a tilt of the chunk insonifies (lo > hi when there is none).  The four
// waves meet in buf[slot]; callers alternate the slot between consecutive
// lines, so a wave that reaches the next line early never overwrites what a
// slower wave still reads.
template <int V>
__device__ __forceinline__ void pixel_law_range(const KArgs& a, const double (&law)[V],
                                                const unsigned (&act)[V], const bool (&valid)[V],
                                                float2 (*buf)[4], int slot, int tid, float& lo,
                                                float& hi) {
  float l = INFINITY, h = -INFINITY;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    if (valid[v] && act[v]) {
      const float x = (float)(a.fs * law[v]);
      l = fminf(l, x);
      h = fmaxf(h, x);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    l = fminf(l, __shfl_xor(l, o, 64));
    h = fmaxf(h, __shfl_xor(h, o, 64));
  }
  if ((tid & 63) == 0) buf[slot][tid >> 6] = make_float2(l, h);
  __syncthreads();
  const float2 b0 = buf[slot][0], b1 = buf[slot][1], b2 = buf[slot][2], b3 = buf[slot][3];
  lo = uni(fminf(fminf(b0.x, b1.x), fminf(b2.x, b3.x)));
  hi = uni(fmaxf(fmaxf(b0.y, b1.y), fmaxf(b2.y, b3.y)));
}

// The same from the line's receive bounds (r_lo > r_hi: the aperture misses
// the box).
template <int LO, int HI, int E>
__device__ __forceinline__ bool line_window_r(const KArgs& a, float r_lo, float r_hi, float bu_lo,
                                              float bu_hi, bool ab_tc, int t0, int nt, int c,
                                              int lane, int& s0, int& whi, float px_lo = 0.0f,
                                              float px_hi = 0.0f) {
  if (!(r_lo <= r_hi)) return false;
  if (!(px_lo <= px_hi)) return false;   // PIXEL_CHANNEL law: no active voxel
  r_lo += px_lo;
  r_hi += px_hi;
  if (ab_tc) {   // per-tilt delays of the TX_CHANNEL law: their range over the chunk
    float lo = INFINITY, hi = -INFINITY;
    if (lane < nt) lo = hi = (float)(a.fs * a.ab[(long long)(t0 + lane) * a.n_ch + c]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo = fminf(lo, __shfl_xor(lo, o, 64));
      hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    r_lo += uni(lo);
    r_hi += uni(hi);
  }
  int w0, w1;
  rca_tile::window_from(a.n_sample, LO, HI, E, bu_lo + r_lo, bu_hi + r_hi, w0, w1);
  whi = uni(w1);
  s0 = uni(w0);
  return true;
}

template <int LO, int HI, int E>
__device__ __forceinline__ bool line_window(const KArgs& a, const TileBox& b, double e_lat,
                                            double ezd, double uabc, float bu_lo, float bu_hi,
                                            bool ab_tc, int t0, int nt, int c, int lane, int& s0,
                                            int& whi, float px_lo = 0.0f, float px_hi = 0.0f) {
  float r_lo, r_hi;
  if (!rca_tile::rx_bounds(a.kfs, b, e_lat, ezd, uabc, r_lo, r_hi)) return false;
  return line_window_r<LO, HI, E>(a, r_lo, r_hi, bu_lo, bu_hi, ab_tc, t0, nt, c, lane, s0, whi,
                                  px_lo, px_hi);
}

// Receive sample-offset range of line c over the workgroup's box (r_lo >
// r_hi when the line's aperture misses the box, or for c = -1).  The kernels
// evaluate it for 64 lines at a time, one line per lane, and read a line's
// bounds back with readlane, so the bounds' FP64 square roots run once per
// 64 lines instead of on every lane for every line.
struct LineBlock {
  float r_lo, r_hi;
};

__device__ __forceinline__ LineBlock line_block(const KArgs& a, const TileBox& b, bool rx, int c) {
  LineBlock L{INFINITY, -INFINITY};
  if (c >= 0) {
    const double ab = a.ab_mode == RCA_ABLAW_CHANNEL ? (double)a.ab[c] : 0.0;
    float lo, hi;
    if (rca_tile::rx_bounds(a.kfs, b, rx ? a.ex[c] : a.ey[c], a.ez[c], a.fs * ab, lo, hi)) {
      L.r_lo = lo;
      L.r_hi = hi;
    }
  }
  return L;
}

// floor(f) as an int in one instruction (v_cvt_flr_i32_f32; the compiler
// forms floor + convert, two).  With d = fract(f) for the fraction, a sample
// position costs one VALU fewer per update (+0.85 % on the config-1 forward).
// f is a sum of fractions in [0, 3) here, where fract(f) = f - floor(f)
// exactly.
__device__ __forceinline__ int floor_to_int(float f) {
  int i;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(f));
  return i;
}

// Lane j's value of x (j uniform), in a scalar register.
__device__ __forceinline__ float read_lane(float x, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
}

// Forward DAS of one workgroup: 256 threads x V voxels, frames
// [by*FB, by*FB + FB).  Loop order: chunks of a.tc (<= TC) tilts, then
// receive lines, then the chunk's tilts.  Per (voxel, tilt) the transmit
// sample offset and phasor are computed once per chunk and cached in LDS;
// per (voxel, line) the receive distance, Hann weight and receive phasor are
// computed once and serve every tilt of the chunk.  The phase is the sum of
// the two range-reduced parts, e^{j2pi(p_tx + p_rx)} = e^{j2pi p_tx} e^{j2pi p_rx}.
//
// LDS: the transmit cache [TC][V][256] TxTerm, then per line rows
// [tilt j][frame fi] of wcap samples: the window of trace (f0 + fi, t0 + j, c)
// that the workgroup's voxels can reach, bounded from their bounding box and
// the chunk's tilts (one window for all rows).  A window longer than wcap is
// staged in chunks (passes); pass k takes the voxels whose first tap lies in
// [s0 + k*step, s0 + (k+1)*step) (the last pass: from s0 + k*step on).
//
// CMP: compound over tilts (acc per voxel and frame).  Otherwise TC = 1 and
// accT holds the current tilt's sums over the lines, or one line's terms with
// nosum.  LAW: kLawLine for no law or the per-line law (a scalar per line),
// kLawTxLine for the per-(tilt, line) law, kLawPixel for the per-(pixel, line)
// laws - separate instantiations, so the common kernels carry neither law's
// registers (the per-(tilt, line) path alone held 7-10 VGPRs, which took the
// frame-block kernels past 128).  NS: NoSum with tilt chunks - every update is
// stored as it is computed (the output was cleared before the launch), so a
// chunk of TC tilts shares each line's receive terms.
template <int KIND, int V, int FB, int TC, bool CMP, typename S, int LAW = kLawLine, bool NS = false>
__device__ __forceinline__ void fwd_block(const KArgs& a, unsigned bx, unsigned by,
                                          unsigned char* smem_raw) {
  constexpr int LO = Interp<KIND>::LO, HI = Interp<KIND>::HI, NT = LO + HI + 1;
  constexpr int E = 16 / sizeof(S);   // samples per 16-byte chunk
  constexpr int R = TC * FB;          // LDS rows at most
  constexpr int G = kThreads / R;     // threads per row when staging
  static_assert(R <= 64 && kThreads % R == 0, "rows must divide the workgroup");
  static_assert(CMP || NS || TC == 1, "per-tilt outputs take one tilt per chunk");
  static_assert(!(CMP && NS), "NoSum outputs are not compounded");
  TxTerm* txc = reinterpret_cast<TxTerm*>(smem_raw);
  const S* iq = reinterpret_cast<const S*>(a.iq);
  __shared__ float s_box[4][6];
  __shared__ float s_scl[R];          // FP16 trace scales of the staged rows
  __shared__ int s_uab_i[TC];         // TX_CHANNEL law: per-tilt delay (samples),
  __shared__ float s_uab_f[TC];       //   integer and fractional parts
  __shared__ float2 s_eab[TC];        // and phasor
  __shared__ float2 s_pxr[2][4];      // PIXEL_CHANNEL law: per-wave ranges

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int f0 = by * FB;
  const int nf = min(FB, a.n_frame - f0);
  const int ns = a.n_sample;
  const int wcap = a.wcap;
  const int tcn = (CMP || NS) ? a.tc : 1;   // tilts per chunk (<= TC)
  S* tr = reinterpret_cast<S*>(smem_raw + (size_t)tcn * V * kThreads * sizeof(TxTerm));
  const int step = pass_step(wcap, NT, E);   // pass advance
  const long long tstride = (long long)a.n_ch * ns;
  const bool rx = a.receive_x != 0;
  constexpr bool ab_tc = LAW == kLawTxLine;
  constexpr bool ab_px = LAW == kLawPixel;
  // single-frame kernels mark uninsonified (voxel, tilt) terms with kTxOff
  // instead of testing the tilt's bit per update (the frame-block kernels
  // test the bit: the marking there cost two VGPRs past the 4-wave limit)
  constexpr bool kOff = FB == 1;
  bool ok;   // scratch for voxel_of at the stores

  // receive-axis and transmit-axis coordinates, depth, aperture half-width
  double pl[V], pt[V], pz[V], apF[V], iapF[V];
  bool valid[V];
  const TileBox box = tile_voxels<V, kTileCols, 4>(a, bx, tid, rx, pl, pt, pz, apF, iapF, valid, s_box);
  int line_no = 0;   // lines visited, alternates the s_pxr slot
  long long pv[V];   // NS: voxel indices for the stores
#pragma unroll
  for (int v = 0; v < V; ++v) pv[v] = NS ? voxel_of<kTileCols, 4>(a, bx, V, v, tid, ok) : 0;

  float2 acc[V][FB];
  float macc[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    macc[v] = 0.0f;
#pragma unroll
    for (int fi = 0; fi < FB; ++fi) acc[v][fi] = make_float2(0.0f, 0.0f);
  }

  for (int t0 = 0; t0 < a.n_tx; t0 += tcn) {
    const int nt = min(tcn, a.n_tx - t0);
    unsigned act[V];
    chunk_tx<V, kOff>(a, t0, nt, rx, pl, pt, pz, valid, txc, tid, act);
    float bu_lo, bu_hi;
    chunk_tx_bounds(a, box, t0, nt, lane, rx, bu_lo, bu_hi);

    float2 accT[V][FB];
    float mT[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      mT[v] = 0.0f;
#pragma unroll
      for (int fi = 0; fi < FB; ++fi) accT[v][fi] = make_float2(0.0f, 0.0f);
    }

    LineBlock lb;
    for (int c = 0; c < a.n_ch; ++c) {
      if ((c & 63) == 0) lb = line_block(a, box, rx, c + lane < a.n_ch ? c + lane : -1);
      const int cl = c & 63;
      const double exd = a.ex[c], eyd = a.ey[c], ezd = a.ez[c];
      const double e_lat = rx ? exd : eyd;
      double uabc = 0.0, pabc = 0.0;
      if (a.ab_mode == RCA_ABLAW_CHANNEL) {
        const double ab = a.ab[c];
        uabc = a.fs * ab;
        pabc = a.fd * ab;
      }
      if (!CMP && !NS && a.nosum) {   // per-line outputs: the sums restart at every line
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
          for (int fi = 0; fi < FB; ++fi) accT[v][fi] = make_float2(0.0f, 0.0f);
      }
      double lawv[V];   // PIXEL_CHANNEL law of each voxel on this line (s)
#pragma unroll
      for (int v = 0; v < V; ++v) lawv[v] = 0.0;
      float px_lo = 0.0f, px_hi = 0.0f;
      if constexpr (ab_px) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const long long p = voxel_of<kTileCols, 4>(a, bx, V, v, tid, ok);
          if (valid[v]) lawv[v] = pixel_law(a, c, p);
        }
        pixel_law_range<V>(a, lawv, act, valid, s_pxr, line_no & 1, tid, px_lo, px_hi);
        ++line_no;
      }
      int s0, whi;
      if (line_window_r<LO, HI, E>(a, read_lane(lb.r_lo, cl), read_lane(lb.r_hi, cl), bu_lo, bu_hi,
                                   ab_tc, t0, nt, c, lane, s0, whi, px_lo, px_hi)) {
        const int npass = pass_count(s0, whi, wcap, step);

        for (int k = 0; k < npass; ++k) {
          const int cb = s0 + k * step;
          const bool last = cb + wcap >= whi;
          // stage rows [j][fi] (G threads per row), with their FP16 scales and
          // the TX_CHANNEL law's per-tilt delays and phasors
          __syncthreads();
          {
            const int row = tid / G, li = tid % G;
            const int j = row / FB, fi = row % FB;
            const int len = min(cb + wcap, whi) - cb;
            if (j < nt && fi < nf && len > 0) {
              const S* src = iq + ((long long)(f0 + fi) * a.n_tx + t0 + j) * tstride +
                             (long long)c * ns + cb;
              S* dst = tr + row * wcap;
              if (a.vec_ok) {
                const int n16 = (int)(((long long)len * sizeof(S) + 15) / 16);
                const uint4* s16 = reinterpret_cast<const uint4*>(src);
                uint4* d16 = reinterpret_cast<uint4*>(dst);
                for (int i = li; i < n16; i += G) d16[i] = s16[i];
              } else {
                for (int i = li; i < len; i += G) dst[i] = src[i];
              }
            }
            if constexpr (sizeof(S) == 4) {
              if (li == 0 && j < nt && fi < nf)
                s_scl[row] = a.scale[((long long)(f0 + fi) * a.n_tx + t0 + j) * a.n_ch + c];
            }
            if (ab_tc && k == 0 && tid < nt) {
              const double ab = a.ab[(long long)(t0 + tid) * a.n_ch + c];
              const double pab = a.fd * ab;
              float sn, cs;
              sincos2pi((float)(pab - rint(pab)), sn, cs);
              split_pos(a.fs * ab, s_uab_i[tid], s_uab_f[tid]);
              s_eab[tid] = make_float2(cs, sn);
            }
          }
          __syncthreads();
          // the updates this pass takes: first tap base = m - LO in [cb, hi)
          // (the range test m in [LO, ns - 1 - HI) of .cu:323 included), as
          // t = base - cb < span in unsigned arithmetic
          const int hi = last ? ns - NT : min(ns - NT, cb + step);
          const unsigned span = hi > cb ? (unsigned)(hi - cb) : 0u;

#pragma unroll
          for (int v = 0; v < V; ++v) {
            if (!act[v]) continue;
            const double A = fabs(pl[v] - e_lat);   // .cu:284-291
            if (!(A < apF[v])) continue;            // .cu:294
            const float w = hann((float)(A * iapF[v]));   // .cu:299
            // .cu:301-304: every insonifying tilt, inside the aperture, before
            // the range test
            if (k == 0) {
              if constexpr (CMP) macc[v] = fmaf(w, (float)__popc(act[v]), macc[v]);
              else mT[v] += w;
            }
            const double dz = pz[v] - ezd;
            const double Rr = sqrt_nr(fmax(fma(A, A, dz * dz), 1e-300));   // .cu:308-309
            const double uab = ab_px ? a.fs * lawv[v] : uabc;
            const double pab = ab_px ? a.fd * lawv[v] : pabc;
            int ri;
            float rf;
            split_pos(fma(a.kfs, Rr, uab), ri, rf);                          // .cu:310-311
            const double ph = fma(a.kfd, Rr, pab);                          // .cu:371-373
            float sn, cs;
            sincos2pi((float)(ph - rint(ph)), sn, cs);
            const float er = w * cs, ei = w * sn;
            const unsigned rio = (unsigned)(ri - LO - cb);
            const TxTerm* tv0 = txc + v * kThreads + tid;
            // one tilt of the chunk: tilt j's transmit terms at tv, rows at rows
            const auto tilt = [&](int j, const TxTerm* tv, const S* rows) __attribute__((always_inline)) {
              if (!kOff && !((act[v] >> j) & 1u)) return;   // .cu:191-214
              // .cu:310-319: u = (bi + bf) + (ri + rf) [+ the per-tilt law],
              // m = floor(u), t = m - LO - cb
              const TxTerm x = *tv;
              // the phasor is used only past the range test; this empty use
              // keeps the term one 16-byte LDS read (split in two 8-byte
              // reads at a 16-byte lane stride, it read with 2-way bank
              // conflicts)
              __asm__ volatile("" ::"v"(x.et.x), "v"(x.et.y));
              float f = x.bf + rf;
              unsigned tb = (unsigned)x.bi + rio;
              if (ab_tc) {
                f += s_uab_f[j];
                tb += (unsigned)s_uab_i[j];
              }
              const unsigned t = tb + (unsigned)floor_to_int(f);
              // the bounds put every tap inside the row; the clamp only keeps
              // LDS reads inside it.  Clamped before the range test and held
              // there by an empty asm: sunk past the branch, t was computed a
              // second time (one VALU per update; +0.6 % measured)
              unsigned tcl = min(t, (unsigned)(wcap - NT));
              __asm__ volatile("" : "+v"(tcl));
              if (t >= span) return;
              const float d = __builtin_amdgcn_fractf(f);   // f - floor(f), exact
              const S* trv = rows + tcl;
              // single frame: the taps are read before the weights are
              // evaluated, so their LDS latency overlaps that arithmetic
              // (read after it, the latency was exposed: +10 % measured)
              float2 tap[FB == 1 ? NT : 1];
              if constexpr (FB == 1) {
#pragma unroll
                for (int kk = 0; kk < NT; ++kk) tap[kk] = ld_sample(trv, kk);
              }
              float q[NT];
              interp_weights<KIND>(d, q);
              float pr = er * x.et.x - ei * x.et.y;
              float pi = er * x.et.y + ei * x.et.x;
              if (ab_tc) {
                const float2 eab = s_eab[j];
                const float r2 = pr * eab.x - pi * eab.y;
                pi = pr * eab.y + pi * eab.x;
                pr = r2;
              }
#pragma unroll
              for (int fi = 0; fi < FB; ++fi) {
                if (FB == 1 || fi < nf) {   // nf >= 1
                  const S* trf = trv + fi * wcap;
                  float re = 0.0f, im = 0.0f;
#pragma unroll
                  for (int kk = 0; kk < NT; ++kk) {
                    const float2 smp = FB == 1 ? tap[kk] : ld_sample(trf, kk);
                    re = fmaf(q[kk], smp.x, re);
                    im = fmaf(q[kk], smp.y, im);
                  }
                  float r2 = pr, i2 = pi;
                  if constexpr (sizeof(S) == 4) {
                    const float sc = s_scl[j * FB + fi];
                    r2 *= sc;
                    i2 *= sc;
                  }
                  if constexpr (NS) {   // NoSum.cu:376-380: out[c][f][t][p], one update each
                    float2* o = a.out +
                                (((long long)c * a.n_frame + f0 + fi) * a.n_tx + t0 + j) * a.n_pix +
                                pv[v];
                    float2 r = make_float2(fmaf(re, r2, -im * i2), fmaf(re, i2, im * r2));
                    if (a.accumulate) { const float2 old = *o; r.x += old.x; r.y += old.y; }
                    *o = r;
                  } else {
                    float2& dst = CMP ? acc[v][fi] : accT[v][fi];
                    dst.x = fmaf(re, r2, fmaf(-im, i2, dst.x));
                    dst.y = fmaf(re, i2, fmaf(im, r2, dst.y));
                  }
                }
              }
            };
            // the chunk's tilts, fully unrolled to the chunk's most (TC), so
            // the transmit-term and row offsets are immediates: +1.6 % FP32
            // and +3.2 % FP16 8-frame blocks, measured (single frame: same).
            // NoSum kernels store inside the loop and lost a wave per SIMD
            // that way: they keep the loop.
            if constexpr (NS) {
#pragma unroll 2
              for (int j = 0; j < nt; ++j) tilt(j, tv0 + j * V * kThreads, tr + j * FB * wcap);
            } else {
#pragma unroll
              for (int j = 0; j < TC; ++j) {
                if (j >= nt) break;
                tilt(j, tv0 + j * V * kThreads, tr + j * FB * wcap);
              }
            }
          }
        }
      }
      if (!CMP && !NS && a.nosum) {   // NoSum.cu:228-229,275-278,376-380: out[c][f][t][p]
#pragma unroll
        for (int v = 0; v < V; ++v) {
          if (!valid[v]) continue;
          const long long p = voxel_of<kTileCols, 4>(a, bx, V, v, tid, ok);
#pragma unroll
          for (int fi = 0; fi < FB; ++fi) {
            if (fi < nf) {
              float2* o = a.out + (((long long)c * a.n_frame + f0 + fi) * a.n_tx + t0) * a.n_pix + p;
              float2 r = accT[v][fi];
              if (a.accumulate) { const float2 old = *o; r.x += old.x; r.y += old.y; }
              *o = r;
            }
          }
        }
      }
    }

    if (NS) {   // masks of the chunk's tilts: the line sum for insonified tilts
#pragma unroll
      for (int v = 0; v < V; ++v) {
        if (!valid[v] || !a.mask || by != 0) continue;
        for (int j = 0; j < nt; ++j) {
          float* mk = a.mask + (long long)(t0 + j) * a.n_pix + pv[v];
          const float mv = ((act[v] >> j) & 1u) ? mT[v] : 0.0f;
          *mk = a.accumulate ? *mk + mv : mv;
        }
      }
    } else if (!CMP) {   // per-tilt output and mask of tilt t0
#pragma unroll
      for (int v = 0; v < V; ++v) {
        if (!valid[v]) continue;
        const long long p = voxel_of<kTileCols, 4>(a, bx, V, v, tid, ok);
        if (!a.nosum) {
#pragma unroll
          for (int fi = 0; fi < FB; ++fi) {
            if (fi < nf) {
              float2* o = a.out + ((long long)(f0 + fi) * a.n_tx + t0) * a.n_pix + p;
              float2 r = accT[v][fi];
              if (a.accumulate) { const float2 old = *o; r.x += old.x; r.y += old.y; }
              *o = r;
            }
          }
        }
        if (a.mask && by == 0) {
          float* mk = a.mask + (long long)t0 * a.n_pix + p;
          *mk = a.accumulate ? *mk + mT[v] : mT[v];
        }
      }
    }
  }

  if (CMP) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      if (!valid[v]) continue;
      const long long p = voxel_of<kTileCols, 4>(a, bx, V, v, tid, ok);
#pragma unroll
      for (int fi = 0; fi < FB; ++fi) {
        if (fi < nf) {
          float2* o = a.out + (long long)(f0 + fi) * a.n_pix + p;
          float2 r = acc[v][fi];
          if (a.accumulate) { const float2 old = *o; r.x += old.x; r.y += old.y; }
          *o = r;
        }
      }
      if (a.mask && by == 0) {
        float* mk = a.mask + p;
        *mk = a.accumulate ? *mk + macc[v] : macc[v];
      }
    }
  }
  __syncthreads();   // s_box and the LDS rows may be reused by the caller's next block
}

// FP32 single-frame compounding kernels (the drop-in's, with or without a
// law) are held to 102 VGPRs, 5 waves per SIMD instead of 4: the compiler
// spills a few values to scratch (outside the per-update loop) and the
// kernels run 8-9 % (no law, every interpolation kind) and 5 % (laws)
// faster, measured.  The FP16, frame-block and NoSum kernels ran slower that
// way and keep the default; holding the multi-frame kernels to 4 waves per
// SIMD (<= 128 VGPRs, which they meet by themselves) cost 9-16 % on the
// 16-frame ensembles, measured.
template <int V, int FB, bool CMP, typename S>
constexpr int fwd_waves() { return (V == 1 && FB == 1 && CMP && sizeof(S) == 8) ? 5 : 1; }

template <int KIND, int V, int FB, bool CMP, typename S, int LAW, bool NS>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(fwd_waves<V, FB, CMP, S>())))
das_fwd(const KArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  fwd_block<KIND, V, FB, (CMP || NS) ? tc_max(FB, sizeof(S)) : 1, CMP, S, LAW, NS>(
      a, blockIdx.x, blockIdx.y, smem_raw);
}

// ----------------------------------------------------------------------------
// Batched forward for the serving loop.  The work list is kept in device
// memory, so one captured sequence of launches serves batches of any shape:
//   header  { n items of each type (4) }
//   args    KArgs per request half (pointers, sizes, flags, tilts per chunk
//           and row length of its item type)
//   items   { first arg, number of halves (1|2), voxel block, frame block },
//           grouped by type in type order
// Item types: FP32 or FP16 I/Q, each with whole-trace rows (entries without
// a depth hint: one frame per FP32 item, kBatchFB16 per FP16 item) or with
// depth tiles (kBatchFB32Tiled / kBatchFB16Tiled frames per item, the
// standalone frame-block kernels' layout).  Eight launches per batch: per type, half 0 then half 1.
// Workgroup g of a launch takes item g of its type (a grid of max_items
// workgroups; those past the item count return at once), so the second half
// of an item accumulates onto the first in the next launch and no two
// workgroups of a launch share an output.  One kernel per type and no loop
// around the block keep each at the register count of the plain kernels (a
// persistent kernel over both sample types and halves took 160 VGPRs, 3 waves
// per SIMD, and ran FP32 items at two thirds of the plain kernel's rate).
// ----------------------------------------------------------------------------
struct BatchArgs {
  KArgs k;
  int iq_dtype;
  int pad_[3];
};

// Frames per work item: FP16 entries without depth tiles, and every entry
// with depth tiles (rca_batch_frame_block).
#ifndef RCA_BATCH_FB16
#define RCA_BATCH_FB16 4
#endif
#ifndef RCA_BATCH_FB32T
#define RCA_BATCH_FB32T 8
#endif
#ifndef RCA_BATCH_FB16T
#define RCA_BATCH_FB16T 8
#endif
constexpr int kBatchFB16 = RCA_BATCH_FB16;
constexpr int kBatchFB32Tiled = RCA_BATCH_FB32T;
constexpr int kBatchFB16Tiled = RCA_BATCH_FB16T;
constexpr int kBatchTypes = 4;
__host__ __device__ constexpr int batch_type(bool f32, bool tiled) { return (f32 ? 0 : 2) + (tiled ? 1 : 0); }
__host__ __device__ constexpr bool batch_f32(int type) { return type < 2; }
__host__ __device__ constexpr bool batch_tiled(int type) { return (type & 1) != 0; }
__host__ __device__ constexpr int batch_fb(int type) {
  return batch_tiled(type) ? (batch_f32(type) ? kBatchFB32Tiled : kBatchFB16Tiled)
                           : batch_f32(type) ? 1 : kBatchFB16;
}

// A wave-uniform struct read from memory, word by word through
// readfirstlane: the values land in scalar registers, as kernel arguments do
// (loaded as vector values they took about 30 VGPRs more).
template <typename T>
__device__ __forceinline__ T uniform_load(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  T v;
  const int* src = reinterpret_cast<const int*>(p);
  int* dst = reinterpret_cast<int*>(&v);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return v;
}

// A pointer rebuilt from words is generic; the round trip through the global
// address space lets the compiler emit global instead of flat accesses (flat
// loads also count against the LDS wait counter, so every LDS wait of the
// per-line loop would wait for them too).
template <typename T>
__device__ __forceinline__ T* as_global(T* p) {
  typedef __attribute__((address_space(1))) char gchar;
  return (T*)(gchar*)reinterpret_cast<unsigned long long>(p);
}

__device__ __forceinline__ KArgs global_ptrs(KArgs k) {
  k.gx = as_global(k.gx);   k.gy = as_global(k.gy);   k.gz = as_global(k.gz);
  k.ex = as_global(k.ex);   k.ey = as_global(k.ey);   k.ez = as_global(k.ez);
  k.vx = as_global(k.vx);   k.vy = as_global(k.vy);   k.vz = as_global(k.vz);
  k.d0 = as_global(k.d0);   k.ab = as_global(k.ab);   k.abt = as_global(k.abt);
  k.iq = as_global(k.iq);   k.scale = as_global(k.scale);
  k.out = as_global(k.out); k.mask = as_global(k.mask);
  k.gout = as_global(k.gout); k.gin = as_global(k.gin);
  return k;
}

// The work list's pointers are rebuilt as global pointers (global_ptrs): with
// the generic pointers of a plain word copy the FP16 depth-tiled batch kernel
// ran at two thirds of the standalone frame-block kernel's rate (flat
// accesses, 130 VGPRs); with them it matches it (126 VGPRs, 4 waves per
// SIMD).  Capping it at 4 waves per SIMD with amdgpu_waves_per_eu instead
// measured 17 % slower.
template <int KIND, int TYPE>
__global__ void __launch_bounds__(kThreads) das_fwd_batch(const int* __restrict__ header,
                                                          const BatchArgs* __restrict__ args,
                                                          const int4* __restrict__ items, int h) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using S = typename std::conditional<batch_f32(TYPE), float2, __half2>::type;
  constexpr int FB = batch_fb(TYPE);
  const int n = header[TYPE];
  if ((int)blockIdx.x >= n) return;
  int first = 0;
#pragma unroll
  for (int t = 0; t < TYPE; ++t) first += header[t];
  const int4 it = uniform_load(items + first + blockIdx.x);
  if (h >= it.y) return;
  const KArgs k = global_ptrs(uniform_load(&args[it.x + h].k));
  fwd_block<KIND, 1, FB, tc_max(FB, sizeof(S)), true, S>(k, it.z, it.w, smem_raw);
}

// Workgroup barrier that orders LDS only: waits for the wave's own LDS
// accesses, not for its global stores and atomics (the adjoint's flush
// atomics have no reader inside the workgroup, and waiting for them at every
// barrier stalls the workgroup on L2 atomic latency).  The memory clobber
// keeps the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Deterministic adjoint.  Float atomics add the tiles' partial sums in
// whatever order the tiles finish; 64-bit integer atomics give the same sum
// in any order.  A partial sum v is added as llrint(v * 2^S), with S set so
// that every sum over the voxels of one gradient sample stays below 2^62: a
// voxel adds |q w (g.re cos + g.im sin)| <= 2 gmax (weights |q| <= 1, w <= 1),
// so the bound is 2.25 n_pix gmax (gmax = max over grad_out of max(|re|,
// |im|)).  The quantum is 2^-S ~ 2.25 n_pix gmax / 2^62, about 1e-12 gmax at
// config 4.  S = 62 - e with 2^(e-1) <= bound < 2^e; scale 1 when gmax is 0
// or not finite (det_to_float then writes NaN whatever the sums hold).
__host__ __device__ __forceinline__ double det_scale(float gmax, long long n_pix, bool inverse) {
  if (!(gmax > 0.0f) || !(gmax <= 3.4028235e38f) || n_pix < 1) return 1.0;
  const int e = ilogb(2.25 * (double)n_pix * (double)gmax) + 1;
  return ldexp(1.0, inverse ? e - 62 : 62 - e);
}

// Adds one flushed partial sum (re, im) to gradient sample i.
__device__ __forceinline__ void flush_add(const KArgs& a, double dsc, long long i, float vr,
                                          float vi) {
  if (a.gdet) {
    unsigned long long* w = a.gdet + 2 * i;
    if (vr != 0.0f) atomicAdd(w, (unsigned long long)(long long)llrint((double)vr * dsc));
    if (vi != 0.0f) atomicAdd(w + 1, (unsigned long long)(long long)llrint((double)vi * dsc));
  } else {
    float* dst = reinterpret_cast<float*>(a.gin + i);
    if (vr != 0.0f) unsafeAtomicAdd(dst, vr);
    if (vi != 0.0f) unsafeAtomicAdd(dst + 1, vi);
  }
}

// ----------------------------------------------------------------------------
// Adjoint kernel (the transpose of fwd_block): 256 lanes, one voxel each,
// frames [by*FB, by*FB + FB).  Same depth tiles (one column per wave: the 64
// lanes of a wave hold consecutive depths about two samples apart), tilt
// chunks, transmit cache and per-line windows as the forward.  For every
// update the lane adds q_k * conj(w e^{j2pi p}) * g to taps m-LO .. m+HI of its
// tilt's window rows in LDS with LDS float atomics (real and imaginary parts
// in separate planes, so a wave's atomics spread over the banks).  After each
// line (and pass) the rows are added to the I/Q gradient with global float
// atomics: the tiles that share samples of a trace all add to them, so the
// gradient must start at zero (the host clears it unless accumulating).
// ----------------------------------------------------------------------------
// One voxel's adjoint terms for the chunk's tilts on one line: for tilt j
// with an in-range sample position, q_k * conj(w e^{j2pi p}) * g added to taps
// m-LO .. m+HI of the wave's row [j][fi] (rows hold samples [cb, cb + wcap),
// real plane then imaginary plane).  PT: per-tilt output gradients, read from
// memory; otherwise g holds the voxel's compounded gradient per frame.
//
// The rows are private to the wave, so only the wave's own lanes can touch
// the same sample.  Lanes of one wave hold different depths, whose first taps
// normally differ: then the taps are added with plain LDS read-modify-writes,
// one tap after another in program order (a compiler barrier keeps that
// order; LDS executes a wave's instructions in order), so a lane's tap k+d
// that lands on another lane's tap k sees that lane's write.  When two active
// lanes share a first tap (a claim slot per first tap detects it) the tilt's
// taps go through LDS float atomics instead, which are slower but safe.
template <int KIND, int FB, bool PT, int V>
__device__ __forceinline__ void adj_scatter(const KArgs& a, const TxTerm* tv, float* rw,
                                            int* claim, unsigned act, int nt, float er, float ei,
                                            int ri, float rf, int cb, bool last, int step,
                                            int wcap, int nf, const float2* g, int f0, int t0,
                                            long long p, bool ab_tc, const int* s_uab_i,
                                            const float* s_uab_f,
                                            const float2* s_eab) {
  constexpr int LO = Interp<KIND>::LO, HI = Interp<KIND>::HI, NT = LO + HI + 1;
  const int ns = a.n_sample;
  const int lane = threadIdx.x & 63;
  // updates of this pass: first tap base = m - LO in [cb, hi) (the range
  // test of .cu:323 included), as t = base - cb < span (unsigned)
  const int hi = last ? ns - NT : min(ns - NT, cb + step);
  const unsigned span = hi > cb ? (unsigned)(hi - cb) : 0u;
  const unsigned rio = (unsigned)(ri - LO - cb);
#pragma unroll 1
  for (int j = 0; j < nt; ++j, tv += kThreads * V, rw += FB * 2 * wcap) {
    if (!((act >> j) & 1u)) continue;
    // .cu:310-319: u = (bi + bf) + (ri + rf) [+ the per-tilt law], m = floor(u)
    const TxTerm x = *tv;
    float f = x.bf + rf;
    unsigned tb = (unsigned)x.bi + rio;
    if (ab_tc) {
      f += s_uab_f[j];
      tb += (unsigned)s_uab_i[j];
    }
    const unsigned t = tb + (unsigned)floor_to_int(f);   // m - LO - cb
    if (t >= span) continue;
    const float d = __builtin_amdgcn_fractf(f);   // f - floor(f), exact
    float q[NT];
    interp_weights<KIND>(d, q);
    float pr = er * x.et.x - ei * x.et.y;
    float pi = er * x.et.y + ei * x.et.x;
    if (ab_tc) {
      const float2 eab = s_eab[j];
      const float r2 = pr * eab.x - pi * eab.y;
      pi = pr * eab.y + pi * eab.x;
      pr = r2;
    }
    // the bounds put every tap inside the row; the clamp only keeps LDS
    // accesses inside it
    const int off = (int)min(t, (unsigned)(wcap - NT));
    float2* rv = reinterpret_cast<float2*>(rw) + off;
    float br[FB], bi[FB];
#pragma unroll
    for (int fi = 0; fi < FB; ++fi) {
      float2 gg = make_float2(0.0f, 0.0f);
      if (fi < nf) {
        if constexpr (PT) gg = a.gout[((long long)(f0 + fi) * a.n_tx + t0 + j) * a.n_pix + p];
        else gg = g[fi];
      }
      // conj(w e^{j 2 pi p}) * g
      br[fi] = fmaf(gg.x, pr, gg.y * pi);
      bi[fi] = fmaf(gg.y, pr, -gg.x * pi);
    }
    // duplicate first taps among the active lanes?
    claim[off] = lane;
    __asm__ volatile("" ::: "memory");
    const bool dup = __ballot(claim[off] != lane) != 0;
    if (!dup) {
#pragma unroll
      for (int kk = 0; kk < NT; ++kk) {
        float2 v[FB];
#pragma unroll
        for (int fi = 0; fi < FB; ++fi) v[fi] = rv[fi * wcap + kk];
#pragma unroll
        for (int fi = 0; fi < FB; ++fi)
          rv[fi * wcap + kk] = make_float2(fmaf(q[kk], br[fi], v[fi].x), fmaf(q[kk], bi[fi], v[fi].y));
        __asm__ volatile("" ::: "memory");
      }
    } else {
#pragma unroll
      for (int fi = 0; fi < FB; ++fi) {
#pragma unroll
        for (int kk = 0; kk < NT; ++kk) {
          atomicAdd(&rv[fi * wcap + kk].x, q[kk] * br[fi]);
          atomicAdd(&rv[fi * wcap + kk].y, q[kk] * bi[fi]);
        }
      }
    }
    __asm__ volatile("" ::: "memory");
  }
}

// Adjoint tiles (64 depths x 4V columns) whose columns lie along the receive
// axis: at every depth lane, every valid voxel of the tile has column 0's
// transmit-lateral coordinate and depth, so the transmit terms
// (bfVirtualSourceGeneric_RCA.cu:256-269 drops the receive-axis offset) are
// those of one column.  Column 0 of a tile is valid whenever any of its
// columns is.  Uniform over the workgroup; both adjoint kernels evaluate it
// the same way, so each tile is taken by exactly one of them.
template <int V>
__device__ __forceinline__ bool tile_shares_tx(const double* pt, const double* pz, const bool* valid,
                                               int tid, float (*s_ptz)[64]) {
  const int lane = tid & 63;
  if (tid < 64) {
    s_ptz[0][lane] = valid[0] ? (float)pt[0] : NAN;
    s_ptz[1][lane] = (float)pz[0];
  }
  __syncthreads();
  bool same = true;
#pragma unroll
  for (int v = 0; v < V; ++v)
    if (valid[v]) same = same && (float)pt[v] == s_ptz[0][lane] && (float)pz[v] == s_ptz[1][lane];
  return __syncthreads_and(same) != 0;
}

template <int KIND, int FB, int TC, bool PX, int V>
__device__ __forceinline__ void adj_block(const KArgs& a, unsigned bx, unsigned by,
                                          unsigned char* smem_raw) {
  constexpr int LO = Interp<KIND>::LO, HI = Interp<KIND>::HI, NT = LO + HI + 1;
  constexpr int R = TC * FB;          // rows per wave at most
  constexpr int G = kThreads / R;     // threads per row when clearing and flushing
  static_assert(R <= 64 && kThreads % R == 0, "rows must divide the workgroup");
  TxTerm* txc = reinterpret_cast<TxTerm*>(smem_raw);
  __shared__ float s_box[4][6];
  __shared__ int s_uab_i[TC];         // TX_CHANNEL law: per-tilt delay (samples),
  __shared__ float s_uab_f[TC];       //   integer and fractional parts
  __shared__ float2 s_eab[TC];        // and phasor
  __shared__ float2 s_pxr[2][4];      // PIXEL_CHANNEL law: per-wave ranges

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  // fixed-point scale of the deterministic flush (unused otherwise)
  const double dsc = a.gdet ? det_scale(__uint_as_float(*a.gdet_max), a.n_pix, false) : 0.0;
  const int f0 = by * FB;
  const int nf = min(FB, a.n_frame - f0);
  const int ns = a.n_sample;
  const int wcap = a.wcap;
  const int tcn = a.tc;   // tilts per chunk (<= TC)
  // per wave: rows [j][fi][wcap] of complex samples; then per wave a claim slot per sample
  float* rows = reinterpret_cast<float*>(smem_raw + (size_t)tcn * kThreads * V * sizeof(TxTerm));
  const size_t wrows = (size_t)tcn * FB * 2 * wcap;   // floats per wave
  float* myrows = rows + wave * wrows;
  int* claim = reinterpret_cast<int*>(rows + 4 * wrows) + wave * wcap;
  const int step = pass_step(wcap, NT, 1);   // pass advance
  const bool rx = a.receive_x != 0;
  const bool ab_tc = a.ab_mode == RCA_ABLAW_TX_CHANNEL;
  constexpr bool ab_px = PX;   // PIXEL_CHANNEL law: its own instantiation
  const bool gt1 = a.gout_tx == 1;
  int line_no = 0;   // lines visited, alternates the s_pxr slot

  double pl[V], pt[V], pz[V], apF[V], iapF[V];
  bool valid[V];
  const TileBox box = tile_voxels<V, 1, 1>(a, bx, tid, rx, pl, pt, pz, apF, iapF, valid, s_box);
  if constexpr (V == 4 && !PX) {
    __shared__ float s_ptz[2][64];
    if (a.adj_lo && tile_shares_tx<V>(pt, pz, valid, tid, s_ptz)) return;   // das_adj_lo's tile
  }
  long long p[V];
  float2 g[V][FB];   // the voxels' output gradients per frame (compounded output)
#pragma unroll
  for (int v = 0; v < V; ++v) {
    bool ok;
    p[v] = voxel_of<1, 1>(a, bx, V, v, tid, ok);
#pragma unroll
    for (int fi = 0; fi < FB; ++fi)
      g[v][fi] = (valid[v] && gt1 && fi < nf) ? a.gout[(long long)(f0 + fi) * a.n_pix + p[v]]
                                              : make_float2(0.0f, 0.0f);
  }
  const int row = tid / G, li = tid % G;   // clearing / flushing assignment
  const int rj = row / FB, rfi = row % FB;
  const int c_first = (int)(((bx + (unsigned)by * 7919u) * 2654435761u) >> 8) % a.n_ch;

  for (int t0 = 0; t0 < a.n_tx; t0 += tcn) {
    const int nt = min(tcn, a.n_tx - t0);
    unsigned act[V];
    chunk_tx<V>(a, t0, nt, rx, pl, pt, pz, valid, txc, tid, act);
    float bu_lo, bu_hi;
    chunk_tx_bounds(a, box, t0, nt, lane, rx, bu_lo, bu_hi);

    LineBlock lb;
    for (int ci = 0; ci < a.n_ch; ++ci) {
      // workgroups start at different lines, so the tiles resident at one
      // time add to different traces instead of contending for the same
      // samples in L2
      const int c = (ci + c_first) % a.n_ch;
      if ((ci & 63) == 0)
        lb = line_block(a, box, rx, ci + lane < a.n_ch ? (ci + lane + c_first) % a.n_ch : -1);
      const int cl = ci & 63;
      const double exd = a.ex[c], eyd = a.ey[c], ezd = a.ez[c];
      const double e_lat = rx ? exd : eyd;
      double uabc = 0.0, pabc = 0.0;
      if (a.ab_mode == RCA_ABLAW_CHANNEL) {
        const double ab = a.ab[c];
        uabc = a.fs * ab;
        pabc = a.fd * ab;
      }
      double lawv[V];   // PIXEL_CHANNEL law of the voxels on this line (s)
      float px_lo = 0.0f, px_hi = 0.0f;
      if constexpr (ab_px) {
#pragma unroll
        for (int v = 0; v < V; ++v) lawv[v] = valid[v] ? pixel_law(a, c, p[v]) : 0.0;
        pixel_law_range<V>(a, lawv, act, valid, s_pxr, line_no & 1, tid, px_lo, px_hi);
        ++line_no;
      }
      int s0, whi;
      if (!line_window_r<LO, HI, 1>(a, read_lane(lb.r_lo, cl), read_lane(lb.r_hi, cl), bu_lo, bu_hi,
                                    ab_tc, t0, nt, c, lane, s0, whi, px_lo, px_hi))
        continue;
      const int npass = pass_count(s0, whi, wcap, step);

      for (int k = 0; k < npass; ++k) {
        const int cb = s0 + k * step;
        const bool last = cb + wcap >= whi;
        const int len = min(cb + wcap, whi) - cb;
        // clear every wave's rows [j][fi] (the previous flush has read them),
        // and publish the TX_CHANNEL law's per-tilt delays and phasors
        lds_barrier();
        if (rj < nt && rfi < nf) {
          for (int w = 0; w < 4; ++w) {
            float2* r = reinterpret_cast<float2*>(rows + w * wrows + (size_t)row * 2 * wcap);
            for (int i = li; i < len; i += G) r[i] = make_float2(0.0f, 0.0f);
          }
        }
        if (ab_tc && k == 0 && tid < nt) {
          const double ab = a.ab[(long long)(t0 + tid) * a.n_ch + c];
          const double pab = a.fd * ab;
          float sn, cs;
          sincos2pi((float)(pab - rint(pab)), sn, cs);
          split_pos(a.fs * ab, s_uab_i[tid], s_uab_f[tid]);
          s_eab[tid] = make_float2(cs, sn);
        }
        lds_barrier();

        // the lane's voxels one after another (neighbouring columns, same
        // depth): each adds to the same rows, flushed once for all of them
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const double A = fabs(pl[v] - e_lat);   // .cu:284-291
          if (!(act[v] && A < apF[v])) continue;  // .cu:294
          const float w = hann((float)(A * iapF[v]));   // .cu:299
          const double dz = pz[v] - ezd;
          const double Rr = sqrt_nr(fmax(fma(A, A, dz * dz), 1e-300));   // .cu:308-309
          const double uab = ab_px ? a.fs * lawv[v] : uabc;
          const double pab = ab_px ? a.fd * lawv[v] : pabc;
          int ri;
          float rf;
          split_pos(fma(a.kfs, Rr, uab), ri, rf);                          // .cu:310-311
          const double ph = fma(a.kfd, Rr, pab);                          // .cu:371-373
          float sn, cs;
          sincos2pi((float)(ph - rint(ph)), sn, cs);
          const float er = w * cs, ei = w * sn;
          const TxTerm* tv = txc + v * kThreads + tid;
          if (gt1)
            adj_scatter<KIND, FB, false, V>(a, tv, myrows, claim, act[v], nt, er, ei, ri, rf, cb, last,
                                            step, wcap, nf, g[v], f0, t0, p[v], ab_tc, s_uab_i,
                                            s_uab_f, s_eab);
          else
            adj_scatter<KIND, FB, true, V>(a, tv, myrows, claim, act[v], nt, er, ei, ri, rf, cb, last,
                                           step, wcap, nf, g[v], f0, t0, p[v], ab_tc, s_uab_i,
                                           s_uab_f, s_eab);
        }
        lds_barrier();
        // add the four waves' rows to the gradient traces (other tiles add to
        // the same samples)
        if (rj < nt && rfi < nf) {
          const float2* r = reinterpret_cast<const float2*>(rows + (size_t)row * 2 * wcap);
          const size_t wr2 = wrows / 2;   // complex samples per wave
          const long long i0 = (((long long)(f0 + rfi) * a.n_tx + t0 + rj) * a.n_ch + c) * ns + cb;
          for (int i = li; i < len; i += G) {
            const float2 v0 = r[i], v1 = r[wr2 + i], v2 = r[2 * wr2 + i], v3 = r[3 * wr2 + i];
            flush_add(a, dsc, i0 + i, v0.x + v1.x + v2.x + v3.x, v0.y + v1.y + v2.y + v3.y);
          }
        }
      }
    }
  }
  __syncthreads();
}

// Line-outer adjoint of the tiles whose columns share the transmit terms
// (tile_shares_tx; a tile of 64 depths x 16 columns along the receive axis,
// 4 voxels per lane).  The transmit terms of a tilt chunk (up to
// kAdjLoTilts tilts) are one column's: 64 per tilt, cached in LDS.  The
// receive terms of a voxel (aperture, Hann weight, FP64 receive distance,
// sample offset, receive phasor) are computed once per (line, chunk) instead
// of once per (line, tilt), and the rows hold one tilt at a time: per (line,
// tilt) the workgroup clears its rows, scatters (adj_scatter), and flushes
// them to the gradient with global float atomics as adj_block does.
// LDS: [tc][64] TxTerm, then per wave FB rows of wcap complex samples, then
// per wave a claim slot per row sample.
constexpr int kAdjLoTilts = 32;

template <int KIND, int FB>
__device__ __forceinline__ void adj_block_lo(const KArgs& a, unsigned bx, unsigned by,
                                             unsigned char* smem_raw) {
  constexpr int V = 4;
  constexpr int LO = Interp<KIND>::LO, HI = Interp<KIND>::HI, NT = LO + HI + 1;
  constexpr int G = kThreads / FB;    // threads per row when clearing and flushing
  __shared__ float s_box[4][6];
  __shared__ float s_ptz[2][64];
  __shared__ int s_uab_i[1];          // TX_CHANNEL law of the current (tilt, line)
  __shared__ float s_uab_f[1];
  __shared__ float2 s_eab[1];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  // fixed-point scale of the deterministic flush (unused otherwise)
  const double dsc = a.gdet ? det_scale(__uint_as_float(*a.gdet_max), a.n_pix, false) : 0.0;
  const int f0 = by * FB;
  const int nf = min(FB, a.n_frame - f0);
  const int ns = a.n_sample;
  const int wcap = a.wcap;
  const int tcn = a.tc;   // tilts per chunk (<= kAdjLoTilts)
  TxTerm* txc = reinterpret_cast<TxTerm*>(smem_raw);
  float* rows = reinterpret_cast<float*>(smem_raw + (size_t)tcn * 64 * sizeof(TxTerm));
  const size_t wrows = (size_t)FB * 2 * wcap;   // floats per wave
  float* myrows = rows + wave * wrows;
  int* claim = reinterpret_cast<int*>(rows + 4 * wrows) + wave * wcap;
  const int step = pass_step(wcap, NT, 1);
  const bool rx = true;   // the receive_x half (transmit terms follow y and z)
  const bool ab_tc = a.ab_mode == RCA_ABLAW_TX_CHANNEL;
  const bool gt1 = a.gout_tx == 1;

  double pl[V], pt[V], pz[V], apF[V], iapF[V];
  bool valid[V];
  const TileBox box = tile_voxels<V, 1, 1>(a, bx, tid, rx, pl, pt, pz, apF, iapF, valid, s_box);
  if (!tile_shares_tx<V>(pt, pz, valid, tid, s_ptz)) return;   // das_adj's tile
  // column 0's transmit-lateral coordinate and depth at this lane's depth
  const float pt0 = s_ptz[0][lane], pz0 = s_ptz[1][lane];
  const bool valid0 = !(pt0 != pt0);
  long long p[V];
  float2 g[V][FB];   // the voxels' output gradients per frame (compounded output)
#pragma unroll
  for (int v = 0; v < V; ++v) {
    bool ok;
    p[v] = voxel_of<1, 1>(a, bx, V, v, tid, ok);
#pragma unroll
    for (int fi = 0; fi < FB; ++fi)
      g[v][fi] = (valid[v] && gt1 && fi < nf) ? a.gout[(long long)(f0 + fi) * a.n_pix + p[v]]
                                              : make_float2(0.0f, 0.0f);
  }
  const int row = tid / G, li = tid % G;   // clearing / flushing assignment (row = frame)
  const int c_first = (int)(((bx + (unsigned)by * 7919u) * 2654435761u) >> 8) % a.n_ch;

  for (int t0 = 0; t0 < a.n_tx; t0 += tcn) {
    const int nt = min(tcn, a.n_tx - t0);
    lds_barrier();   // the previous chunk's transmit terms are no longer read
    // transmit terms of column 0 (.cu:256-269), wave w taking tilts w, w + 4, ...
    for (int j = wave; j < nt; j += 4) {
      const int t = t0 + j;
      TxTerm x;
      x.bi = 0;
      x.bf = 0.0f;
      x.et = make_float2(1.0f, 0.0f);
      if (valid0) {
        const double dl = (double)pt0 - (double)a.vy[t], dz = (double)pz0 - (double)a.vz[t];
        const double tt = sqrt_nr(fma(dl, dl, dz * dz)) - (double)a.d0[t];
        const double Bp = fma(a.kfd, fma(-2.0, (double)pz0, tt), a.pt0);
        float sn, cs;
        sincos2pi((float)(Bp - rint(Bp)), sn, cs);
        split_pos(fma(a.kfs, tt, a.ut0), x.bi, x.bf);
        x.et = make_float2(cs, sn);
      }
      txc[j * 64 + lane] = x;
    }
    // transmit cone bits of the lane's voxels for the chunk's tilts
    unsigned act[V];
#pragma unroll
    for (int v = 0; v < V; ++v) act[v] = 0u;
    for (int j = 0; j < nt; ++j) {
      const int t = t0 + j;
      const float vxf = a.vx[t], vyf = a.vy[t], vzf = a.vz[t];
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (valid[v] && (!a.mask_tx || tx_mask(a, vxf, vyf, vzf, (float)pl[v], (float)pt[v], (float)pz[v])))
          act[v] |= 1u << j;
    }
    // transmit sample-offset bounds over the tile, lane j holding tilt t0 + j
    float bu_lo = INFINITY, bu_hi = -INFINITY;
    if (lane < nt)
      rca_tile::tx_bounds(a.kfs, a.ut0, box.txl_lo, box.txl_hi, box.zlo, box.zhi, a.vy[t0 + lane],
                          a.vz[t0 + lane], a.d0[t0 + lane], bu_lo, bu_hi);
    lds_barrier();   // transmit terms published

    LineBlock lb;
    for (int ci = 0; ci < a.n_ch; ++ci) {
      const int c = (ci + c_first) % a.n_ch;
      if ((ci & 63) == 0)
        lb = line_block(a, box, true, ci + lane < a.n_ch ? (ci + lane + c_first) % a.n_ch : -1);
      const float r_lo = read_lane(lb.r_lo, ci & 63), r_hi = read_lane(lb.r_hi, ci & 63);
      if (!(r_lo <= r_hi)) continue;
      const double exd = a.ex[c], ezd = a.ez[c];
      double uabc = 0.0, pabc = 0.0;
      if (a.ab_mode == RCA_ABLAW_CHANNEL) {
        const double ab = a.ab[c];
        uabc = a.fs * ab;
        pabc = a.fd * ab;
      }
      // receive terms of the lane's voxels on this line (.cu:284-311, 371-373)
      bool inap[V];
      int ri[V];
      float rf[V], er[V], ei[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const double A = fabs(pl[v] - exd);
        inap[v] = valid[v] && A < apF[v];
        ri[v] = 0;
        rf[v] = er[v] = ei[v] = 0.0f;
        if (inap[v]) {
          const float w = hann((float)(A * iapF[v]));
          const double dz = pz[v] - ezd;
          const double Rr = sqrt_nr(fmax(fma(A, A, dz * dz), 1e-300));
          split_pos(fma(a.kfs, Rr, uabc), ri[v], rf[v]);
          const double ph = fma(a.kfd, Rr, pabc);
          float sn, cs;
          sincos2pi((float)(ph - rint(ph)), sn, cs);
          er[v] = w * cs;
          ei[v] = w * sn;
        }
      }
      for (int j = 0; j < nt; ++j) {
        const int t = t0 + j;
        int s0, whi;
        if (!line_window_r<LO, HI, 1>(a, r_lo, r_hi, __shfl(bu_lo, j, 64), __shfl(bu_hi, j, 64),
                                      ab_tc, t, 1, c, lane, s0, whi))
          continue;
        const int npass = pass_count(s0, whi, wcap, step);
        for (int k = 0; k < npass; ++k) {
          const int cb = s0 + k * step;
          const bool last = cb + wcap >= whi;
          const int len = min(cb + wcap, whi) - cb;
          lds_barrier();   // the previous flush has read the rows
          if (row < nf) {
            for (int w = 0; w < 4; ++w) {
              float2* r = reinterpret_cast<float2*>(rows + w * wrows + (size_t)row * 2 * wcap);
              for (int i = li; i < len; i += G) r[i] = make_float2(0.0f, 0.0f);
            }
          }
          if (ab_tc && k == 0 && tid == 0) {
            const double ab = a.ab[(long long)t * a.n_ch + c];
            const double pab = a.fd * ab;
            float sn, cs;
            sincos2pi((float)(pab - rint(pab)), sn, cs);
            split_pos(a.fs * ab, s_uab_i[0], s_uab_f[0]);
            s_eab[0] = make_float2(cs, sn);
          }
          lds_barrier();
#pragma unroll
          for (int v = 0; v < V; ++v) {
            if (!(inap[v] && ((act[v] >> j) & 1u))) continue;
            const TxTerm* tv = txc + j * 64 + lane;
            if (gt1)
              adj_scatter<KIND, FB, false, V>(a, tv, myrows, claim, 1u, 1, er[v], ei[v], ri[v], rf[v], cb,
                                              last, step, wcap, nf, g[v], f0, t, p[v], ab_tc, s_uab_i,
                                              s_uab_f, s_eab);
            else
              adj_scatter<KIND, FB, true, V>(a, tv, myrows, claim, 1u, 1, er[v], ei[v], ri[v], rf[v], cb,
                                             last, step, wcap, nf, g[v], f0, t, p[v], ab_tc, s_uab_i,
                                             s_uab_f, s_eab);
          }
          lds_barrier();
          // the four waves' rows added to the gradient traces of tilt t
          if (row < nf) {
            const float2* r = reinterpret_cast<const float2*>(rows + (size_t)row * 2 * wcap);
            const size_t wr2 = wrows / 2;
            const long long i0 = (((long long)(f0 + row) * a.n_tx + t) * a.n_ch + c) * ns + cb;
            for (int i = li; i < len; i += G) {
              const float2 v0 = r[i], v1 = r[wr2 + i], v2 = r[2 * wr2 + i], v3 = r[3 * wr2 + i];
              flush_add(a, dsc, i0 + i, v0.x + v1.x + v2.x + v3.x, v0.y + v1.y + v2.y + v3.y);
            }
          }
        }
      }
    }
  }
  __syncthreads();
}

template <int KIND, int FB>
__global__ void __launch_bounds__(kThreads) das_adj_lo(const KArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  adj_block_lo<KIND, FB>(a, blockIdx.x, blockIdx.y, smem_raw);
}

template <int KIND, int FB, bool PX, int V>
__global__ void __launch_bounds__(kThreads) das_adj(const KArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  adj_block<KIND, FB, tc_max(FB, 8), PX, V>(a, blockIdx.x, blockIdx.y, smem_raw);
}

__global__ void fill_zero(float* __restrict__ p, long long n) {
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long long i = i0; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0.0f;
}

// Deterministic adjoint, before the scatter: *out = max |x[i]| over n floats
// (the re and im parts of grad_out), as float bits (*out starts at 0; the
// bits of non-negative floats order like the floats; a NaN wins).
__global__ void abs_max(const float* __restrict__ x, long long n, unsigned* __restrict__ out) {
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned m = 0u;
  for (long long i = i0; i < n; i += (long long)gridDim.x * blockDim.x)
    m = max(m, __float_as_uint(fabsf(x[i])));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

// Deterministic adjoint, after the scatter: the fixed-point sums back to FP32,
// stored or added to grad_iq (a non-finite grad_out gives NaN throughout).
__global__ void det_to_float(const unsigned long long* __restrict__ w, const unsigned* __restrict__ gmax,
                             float* __restrict__ out, long long n, long long n_pix, int accumulate) {
  const float gm = __uint_as_float(*gmax);
  const double inv = det_scale(gm, n_pix, true);
  const bool finite = gm <= 3.4e38f;
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long long i = i0; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = finite ? (float)((double)(long long)w[i] * inv) : __int_as_float(0x7fc00000);
    out[i] = accumulate ? out[i] + v : v;
  }
}

// ----------------------------------------------------------------------------
// Host side.
// ----------------------------------------------------------------------------
// process-wide launch-shape overrides (rca_set_tuning); atomics, so a sweep
// thread may change them while other threads launch
std::atomic<int> g_tune_v{0}, g_tune_fb{0}, g_tune_lds{0};

// LDS budget of a workgroup: the override when set, else the default
size_t lds_budget(size_t dflt) {
  const int kib = g_tune_lds.load();
  return kib ? (size_t)kib * 1024 : dflt;
}

int check_desc(const rca_das_desc* d) {
  if (!d) return fail(RCA_ERR_INVALID, "desc is NULL");
  if (d->n_pix < 0 || d->n_channel < 1 || d->n_tx < 1 || d->n_sample < 1)
    return fail(RCA_ERR_INVALID, "bad sizes: nPix=%d nChannel=%d nTx=%d nSample=%d", d->n_pix,
                d->n_channel, d->n_tx, d->n_sample);
  if (d->n_sample > kMaxSamples)
    return fail(RCA_ERR_UNSUPPORTED, "nSample=%d exceeds %d samples per trace", d->n_sample,
                kMaxSamples);
  if (d->n_pix > 0 && (!d->grid_x || !d->grid_y || !d->grid_z))
    return fail(RCA_ERR_INVALID, "grid pointers are NULL");
  if (!d->ele_x || !d->ele_y || !d->ele_z) return fail(RCA_ERR_INVALID, "element pointers are NULL");
  if (!d->vs_x || !d->vs_y || !d->vs_z || !d->min_dist)
    return fail(RCA_ERR_INVALID, "virtual-source pointers are NULL");
  if (d->interp < RCA_INTERP_QUADRATIC || d->interp > RCA_INTERP_LANCZOS4)
    return fail(RCA_ERR_INVALID, "unknown interpolation %d", d->interp);
  if (d->ab_law_mode < RCA_ABLAW_NONE || d->ab_law_mode > RCA_ABLAW_PIXEL_CHANNEL)
    return fail(RCA_ERR_INVALID, "unknown aberration-law mode %d", d->ab_law_mode);
  if (d->ab_law_mode == RCA_ABLAW_PIXEL_CHANNEL) {
    if (!d->ab_law_rx && !d->ab_law_tx)
      return fail(RCA_ERR_INVALID, "ab_law_rx and ab_law_tx are both NULL");
  } else if (d->ab_law_mode && !d->ab_law_rx) {
    return fail(RCA_ERR_INVALID, "ab_law_rx is NULL");
  }
  if (!(d->speed_of_sound > 0.0f) || !(d->sampling_frequency > 0.0f))
    return fail(RCA_ERR_INVALID, "speed of sound and sampling frequency must be > 0");
  if (!(d->fnumber != 0.0f)) return fail(RCA_ERR_INVALID, "fnumber must be non-zero");
  return RCA_OK;
}

KArgs make_args(const rca_das_desc* d, int n_frame) {
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.gx = d->grid_x; a.gy = d->grid_y; a.gz = d->grid_z;
  a.ex = d->ele_x; a.ey = d->ele_y; a.ez = d->ele_z;
  a.vx = d->vs_x; a.vy = d->vs_y; a.vz = d->vs_z; a.d0 = d->min_dist;
  a.ab = d->ab_law_rx;
  a.abt = d->ab_law_mode == RCA_ABLAW_PIXEL_CHANNEL ? d->ab_law_tx : nullptr;
  a.ab_mode = d->ab_law_mode;
  const double fs = d->sampling_frequency, fd = d->demodulation_frequency, c = d->speed_of_sound;
  a.kfs = fs / c; a.kfd = fd / c;
  a.ut0 = fs * (double)d->start_time; a.pt0 = fd * (double)d->start_time;
  a.fs = fs; a.fd = fd;
  a.apx = d->aperture_x; a.apy = d->aperture_y; a.apz = d->aperture_z; a.fnum = d->fnumber;
  a.n_pix = d->n_pix; a.n_ch = d->n_channel; a.n_tx = d->n_tx; a.n_sample = d->n_sample;
  a.n_frame = n_frame;
  a.receive_x = d->receive_x != 0; a.mask_tx = d->mask_transmit != 0;
  a.depth_run = d->depth_run >= 64 ? d->depth_run : 0;
  return a;
}

using rca_tile::kFwdLds;
using rca_tile::kFwdLds1;

template <typename S> struct SName;

template <int KIND, int V, int FB, bool CMP, typename S, int LAW = kLawLine, bool NS = false>
int launch_fwd_t(KArgs a, hipStream_t st) {
  const FwdLayout L = rca_tile::fwd_layout(a.depth_run, a.n_sample, V, FB, sizeof(S),
                                           (CMP || NS) ? tc_max(FB, sizeof(S)) : 1,
                                           lds_budget(FB == 1 ? kFwdLds1 : kFwdLds));
  if (L.lds > (size_t)kMaxLds)
    return fail(RCA_ERR_UNSUPPORTED, "FB=%d traces of %d samples do not fit in LDS", FB, a.n_sample);
  a.tc = L.tc;
  a.wcap = L.wcap;
  const unsigned gx = (unsigned)rca_tile::fwd_blocks(a.n_pix, a.depth_run, V);
  const unsigned gy = (unsigned)((a.n_frame + FB - 1) / FB);
  if (gx == 0 || gy == 0) return RCA_OK;
  hipLaunchKernelGGL((das_fwd<KIND, V, FB, CMP, S, LAW, NS>), dim3(gx, gy, 1), dim3(kThreads), L.lds, st,
                     a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? RCA_OK : hip_fail(e, "das_fwd launch");
}

// Instantiated shapes: compounding (V, FB) in (1|2, 1), (1, 2|4|8); per-tilt
// and per-line outputs (1, 1|4); with the PIXEL_CHANNEL law (1, 1|4) and the
// TX_CHANNEL law (1, 1|4), NoSum included.
template <int KIND, typename S>
int launch_fwd_k(const KArgs& a, int V, int FB, hipStream_t st) {
  if (a.ab_mode == RCA_ABLAW_PIXEL_CHANNEL) {
    if (V != 1 || (FB != 1 && FB != 4))
      return fail(RCA_ERR_UNSUPPORTED, "no pixel-law kernel for V=%d FB=%d", V, FB);
    if (a.compound)
      return FB == 1 ? launch_fwd_t<KIND, 1, 1, true, S, kLawPixel>(a, st)
                     : launch_fwd_t<KIND, 1, 4, true, S, kLawPixel>(a, st);
    return FB == 1 ? launch_fwd_t<KIND, 1, 1, false, S, kLawPixel>(a, st)
                   : launch_fwd_t<KIND, 1, 4, false, S, kLawPixel>(a, st);
  }
  if (a.ab_mode == RCA_ABLAW_TX_CHANNEL) {
    if (V != 1 || (FB != 1 && FB != 4))
      return fail(RCA_ERR_UNSUPPORTED, "no per-(tilt, line) law kernel for V=%d FB=%d", V, FB);
    if (a.compound)
      return FB == 1 ? launch_fwd_t<KIND, 1, 1, true, S, kLawTxLine>(a, st)
                     : launch_fwd_t<KIND, 1, 4, true, S, kLawTxLine>(a, st);
    if (a.nosum)
      return FB == 1 ? launch_fwd_t<KIND, 1, 1, false, S, kLawTxLine, true>(a, st)
                     : launch_fwd_t<KIND, 1, 4, false, S, kLawTxLine, true>(a, st);
    return FB == 1 ? launch_fwd_t<KIND, 1, 1, false, S, kLawTxLine>(a, st)
                   : launch_fwd_t<KIND, 1, 4, false, S, kLawTxLine>(a, st);
  }
  if (a.compound) {
#define RCA_FWD_CASE(VV, FF) \
    if (V == VV && FB == FF) return launch_fwd_t<KIND, VV, FF, true, S>(a, st);
    RCA_FWD_CASE(1, 1) RCA_FWD_CASE(2, 1) RCA_FWD_CASE(1, 2) RCA_FWD_CASE(1, 4) RCA_FWD_CASE(1, 8)
#undef RCA_FWD_CASE
  } else if (a.nosum) {   // NoSum with tilt chunks (output cleared by the caller)
    if (V == 1 && FB == 1) return launch_fwd_t<KIND, 1, 1, false, S, kLawLine, true>(a, st);
    if (V == 1 && FB == 4) return launch_fwd_t<KIND, 1, 4, false, S, kLawLine, true>(a, st);
  } else {
#define RCA_FWD_CASE(FF) \
    if (V == 1 && FB == FF) return launch_fwd_t<KIND, 1, FF, false, S>(a, st);
    RCA_FWD_CASE(1) RCA_FWD_CASE(4)
#undef RCA_FWD_CASE
  }
  return fail(RCA_ERR_UNSUPPORTED, "no kernel for V=%d FB=%d", V, FB);
}

template <typename S>
int launch_fwd_s(const KArgs& a, int interp, int V, int FB, hipStream_t st) {
  switch (interp) {
    case RCA_INTERP_QUADRATIC: return launch_fwd_k<RCA_INTERP_QUADRATIC, S>(a, V, FB, st);
    case RCA_INTERP_LANCZOS_REF: return launch_fwd_k<RCA_INTERP_LANCZOS_REF, S>(a, V, FB, st);
    case RCA_INTERP_LANCZOS3: return launch_fwd_k<RCA_INTERP_LANCZOS3, S>(a, V, FB, st);
    case RCA_INTERP_LANCZOS4: return launch_fwd_k<RCA_INTERP_LANCZOS4, S>(a, V, FB, st);
  }
  return fail(RCA_ERR_INVALID, "unknown interpolation %d", interp);
}

int pick_fb(int n_frame, int n_sample, size_t bytes, bool tiled) {
  int fb = 8;   // measured best for FP32 and FP16 compounding (frame-block kernels at <= 128 VGPRs)
  // depth tiles stage short windows, so any FB fits; whole traces must fit
  while (fb > 1 && (fb > n_frame || (!tiled && (size_t)fb * n_sample * bytes +
                                      kThreads * sizeof(TxTerm) > (size_t)kMaxLds))) fb >>= 1;
  return fb;
}

int pick_v(long long, int, int, int, bool) {
  // V = 1: the tilt-chunked kernel holds more per-voxel state, and one voxel
  // per lane measured fastest at every LDS budget on cfg1-sized volumes.
  return 1;
}

template <int KIND, int FB, bool PX = false, int V = 1>
int launch_adj_t(KArgs a, hipStream_t st) {
  const AdjLayout L = rca_tile::adj_layout(a.depth_run, a.n_sample, FB, tc_max(FB, 8),
                                           lds_budget(rca_tile::kAdjLds), V);
  if (L.lds > (size_t)kMaxLds || L.wcap < Interp<KIND>::LO + Interp<KIND>::HI + 1)
    return fail(RCA_ERR_UNSUPPORTED, "FB=%d adjoint rows do not fit in LDS", FB);
  a.tc = L.tc;
  a.wcap = L.wcap;
  const dim3 grid((unsigned)rca_tile::tile_blocks<1, 1>(a.n_pix, a.depth_run, V),
                  (unsigned)((a.n_frame + FB - 1) / FB));
  if constexpr (V == 4 && !PX) {
    // receive-x half with depth tiles: tiles along the receive axis share
    // the transmit terms and go to the line-outer kernel, the rest (checked
    // per tile, so any grid is handled) to das_adj
    if (a.receive_x && a.depth_run >= 64) {
      KArgs b = a;
      b.tc = a.n_tx < kAdjLoTilts ? a.n_tx : kAdjLoTilts;
      b.wcap = a.n_sample > rca_tile::kAdjTileRow ? rca_tile::kAdjTileRow : a.n_sample;
      const auto lo_lds = [&](int tc) {
        return (size_t)tc * 64 * sizeof(TxTerm) + 4 * 8 * (size_t)FB * b.wcap + 4 * 4 * (size_t)b.wcap;
      };
      while (b.tc > 1 && lo_lds(b.tc) > (size_t)kMaxLds) --b.tc;
      if (lo_lds(b.tc) <= (size_t)kMaxLds && b.wcap >= Interp<KIND>::LO + Interp<KIND>::HI + 1) {
        a.adj_lo = b.adj_lo = 1;
        hipLaunchKernelGGL((das_adj_lo<KIND, FB>), grid, dim3(kThreads), lo_lds(b.tc), st, b);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "das_adj_lo launch");
      }
    }
  }
  hipLaunchKernelGGL((das_adj<KIND, FB, PX, V>), grid, dim3(kThreads), L.lds, st, a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? RCA_OK : hip_fail(e, "das_adj launch");
}

template <int KIND>
int launch_adj_k(const KArgs& a, int FB, int V, hipStream_t st) {
  if (a.ab_mode == RCA_ABLAW_PIXEL_CHANNEL) {   // instantiated for FB = 1, V = 1 only
    if (FB != 1 || V != 1)
      return fail(RCA_ERR_UNSUPPORTED, "no pixel-law adjoint kernel for FB=%d V=%d", FB, V);
    return launch_adj_t<KIND, 1, true>(a, st);
  }
#define RCA_ADJ_CASE(FF, VV) \
  if (FB == FF && V == VV) return launch_adj_t<KIND, FF, false, VV>(a, st);
  RCA_ADJ_CASE(1, 1) RCA_ADJ_CASE(1, 2) RCA_ADJ_CASE(1, 4)
  RCA_ADJ_CASE(2, 1) RCA_ADJ_CASE(2, 2) RCA_ADJ_CASE(2, 4)
  RCA_ADJ_CASE(4, 1) RCA_ADJ_CASE(4, 2) RCA_ADJ_CASE(4, 4)
  RCA_ADJ_CASE(8, 1) RCA_ADJ_CASE(8, 2)
#undef RCA_ADJ_CASE
  return fail(RCA_ERR_UNSUPPORTED, "no adjoint kernel for FB=%d V=%d", FB, V);
}

// Adjoint launch shared by rca_adjoint (float atomics) and rca_adjoint_det
// (ws non-null: fixed-point sums in ws, then converted into grad_iq).
constexpr size_t kDetHeader = 256;

// Interpolation taps NT = LO + HI + 1 of a (validated) kind.
int taps_of(int interp) {
  switch (interp) {
    case RCA_INTERP_QUADRATIC: return Interp<RCA_INTERP_QUADRATIC>::LO + Interp<RCA_INTERP_QUADRATIC>::HI + 1;
    case RCA_INTERP_LANCZOS_REF: return Interp<RCA_INTERP_LANCZOS_REF>::LO + Interp<RCA_INTERP_LANCZOS_REF>::HI + 1;
    case RCA_INTERP_LANCZOS3: return Interp<RCA_INTERP_LANCZOS3>::LO + Interp<RCA_INTERP_LANCZOS3>::HI + 1;
    default: return Interp<RCA_INTERP_LANCZOS4>::LO + Interp<RCA_INTERP_LANCZOS4>::HI + 1;
  }
}

int adjoint_run(const rca_das_desc* d, const void* grad_out, int grad_tx, int n_frame,
                void* grad_iq, int flags, void* ws, hipStream_t stream) {
  g_err.clear();
  int rc = check_desc(d);
  if (rc) return rc;
  if (!grad_out || !grad_iq) return fail(RCA_ERR_INVALID, "grad pointers are NULL");
  if (grad_tx != 1 && grad_tx != d->n_tx)
    return fail(RCA_ERR_INVALID, "grad_tx must be 1 or nTx (got %d)", grad_tx);
  if (n_frame < 0) return fail(RCA_ERR_INVALID, "nFrame < 0");
  if (flags & ~RCA_FLAG_ACCUMULATE) return fail(RCA_ERR_INVALID, "unknown flags 0x%x", flags);
  if (n_frame == 0) return RCA_OK;
  KArgs a = make_args(d, n_frame);
  a.gout = reinterpret_cast<const float2*>(grad_out);
  a.gin = reinterpret_cast<float2*>(grad_iq);
  a.gout_tx = grad_tx;
  a.accumulate = (flags & RCA_FLAG_ACCUMULATE) != 0;
  // Frames per workgroup (FB) and voxels per lane (V).  FB = 4 with V = 4
  // measured fastest on a 96x96x192 volume (8 frames) and on both config-1
  // halves (4 and 8 frames).  A partial frame block costs a full one (the
  // scatter runs all FB frames), so the frames go in launches of
  // power-of-two blocks: the 4-frame blocks that fit, then a 2-frame and a
  // 1-frame launch for the rest, on the same fill and workspace.  Those take
  // V = 2 (3 waves per SIMD), except 2-frame blocks on the receive-x half
  // with depth tiles, where V = 4 sends the tiles to the line-outer kernel.
  // Config 1, ms: 2 frames 74.9 (V = 4) vs 80.5 on receive x, 85 vs 113
  // (V = 4) on receive y; receive y, 3 frames as 2 + 1: 134 vs 161 in one
  // launch of 2-frame blocks; 5 and 6 frames 167 and 193 vs 234 in 4-frame
  // blocks (receive x).  A tuned FB or V (rca_set_tuning) gives one launch
  // of that shape, as does the pixel law (its one instantiated shape).
  struct Part { int f0, nf, fb, v; };
  Part parts[3];
  int n_parts = 0;
  const int tfb = g_tune_fb.load(), tv = g_tune_v.load();
  const bool lo_tiles = d->receive_x != 0 && d->depth_run >= 64;
  if (d->ab_law_mode == RCA_ABLAW_PIXEL_CHANNEL) {
    parts[n_parts++] = {0, n_frame, 1, 1};
  } else if (tfb || tv) {
    int fb = tfb ? tfb : 4;
    while (fb > 1 && fb > n_frame) fb >>= 1;
    int v = tv ? tv : ((fb == 4 || (fb == 2 && lo_tiles)) ? 4 : 2);
    if (fb == 8 && v > 2) v = 2;
    parts[n_parts++] = {0, n_frame, fb, v};
  } else if (n_frame % 4 == 3 && lo_tiles) {
    // three frames left on the line-outer half: one partial 4-frame block
    // (118 ms) beats 2 + 1 (124 ms); on the other half 2 + 1 wins (134 vs 147)
    parts[n_parts++] = {0, n_frame, 4, 4};
  } else {
    const int n4 = n_frame / 4 * 4;
    if (n4) parts[n_parts++] = {0, n4, 4, 4};
    if (n_frame - n4 >= 2) parts[n_parts++] = {n4, 2, 2, lo_tiles ? 4 : 2};
    if ((n_frame - n4) % 2) parts[n_parts++] = {n_frame - 1, 1, 1, 2};
  }
  const long long n = 2LL * n_frame * d->n_tx * d->n_channel * d->n_sample;   // floats of grad_iq
  hipError_t e;
  if (ws) {
    // fixed-point sums start at zero; grad_out's largest component sets their scale
    unsigned char* w = reinterpret_cast<unsigned char*>(ws);
    hipLaunchKernelGGL(fill_zero, dim3(4096), dim3(256), 0, stream, reinterpret_cast<float*>(w),
                       (long long)((kDetHeader + 8 * (size_t)n) / 4));
    a.gdet_max = reinterpret_cast<const unsigned*>(w);
    a.gdet = reinterpret_cast<unsigned long long*>(w + kDetHeader);
    const long long ng = 2LL * n_frame * grad_tx * d->n_pix;
    hipLaunchKernelGGL(abs_max, dim3(2048), dim3(256), 0, stream,
                       reinterpret_cast<const float*>(grad_out), ng, reinterpret_cast<unsigned*>(w));
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "deterministic adjoint set-up");
  } else if (!a.accumulate) {
    // every workgroup adds its windows to the traces: start from zero
    hipLaunchKernelGGL(fill_zero, dim3(2048), dim3(256), 0, stream,
                       reinterpret_cast<float*>(grad_iq), n);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "fill_zero launch");
  }
  // a trace of at most NT samples holds no update (.cu:323 needs m - LO >= 0
  // and m + HI < nSample - 1): A^H grad_out is zero, which the fill above wrote
  if (d->n_pix > 0 && d->n_sample > taps_of(d->interp)) {
    const long long trace_samples = (long long)d->n_tx * d->n_channel * d->n_sample;
    for (int k = 0; k < n_parts; ++k) {
      const Part& q = parts[k];
      // the part's frames: grad_out, grad_iq and the fixed-point sums from
      // frame q.f0 on (the kernels index frames from their base pointers)
      KArgs b = a;
      b.n_frame = q.nf;
      b.gout = a.gout + (long long)q.f0 * grad_tx * d->n_pix;
      b.gin = a.gin + (long long)q.f0 * trace_samples;
      if (a.gdet) b.gdet = a.gdet + 2 * (long long)q.f0 * trace_samples;
      switch (d->interp) {
        case RCA_INTERP_QUADRATIC: rc = launch_adj_k<RCA_INTERP_QUADRATIC>(b, q.fb, q.v, stream); break;
        case RCA_INTERP_LANCZOS_REF: rc = launch_adj_k<RCA_INTERP_LANCZOS_REF>(b, q.fb, q.v, stream); break;
        case RCA_INTERP_LANCZOS3: rc = launch_adj_k<RCA_INTERP_LANCZOS3>(b, q.fb, q.v, stream); break;
        case RCA_INTERP_LANCZOS4: rc = launch_adj_k<RCA_INTERP_LANCZOS4>(b, q.fb, q.v, stream); break;
        default: return fail(RCA_ERR_INVALID, "unknown interpolation %d", d->interp);
      }
      if (rc) return rc;
    }
  }
  if (ws) {
    hipLaunchKernelGGL(det_to_float, dim3(4096), dim3(256), 0, stream,
                       reinterpret_cast<const unsigned long long*>(a.gdet), a.gdet_max,
                       reinterpret_cast<float*>(grad_iq), n, (long long)d->n_pix, a.accumulate);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "det_to_float launch");
  }
  return RCA_OK;
}

}  // namespace

extern "C" {

const char* rca_last_error(void) { return g_err.c_str(); }

const char* rca_version(void) { return "rca_das 0.1.0 gfx950"; }

// Tuning override for sweeps: workgroup voxels per lane (V), frames per block
// (FB) and adjoint voxel splits; 0 restores the automatic choice.
int rca_set_tuning(int v, int fb, int split, int lds_kib) {
  if ((v != 0 && v != 1 && v != 2 && v != 4) || (fb != 0 && fb != 1 && fb != 2 && fb != 4 && fb != 8) ||
      split < 0 || (lds_kib != 0 && (lds_kib < 16 || lds_kib > 64)))
    return fail(RCA_ERR_INVALID, "bad tuning V=%d FB=%d split=%d lds=%d KiB", v, fb, split, lds_kib);
  g_tune_v = v;
  g_tune_fb = fb;
  g_tune_lds = lds_kib;
  return RCA_OK;
}

int rca_forward(const rca_das_desc* d, const void* iq, int iq_dtype, const float* trace_scale,
                int n_frame, void* out, float* mask, int flags, hipStream_t stream) {
  g_err.clear();
  int rc = check_desc(d);
  if (rc) return rc;
  if (n_frame < 0) return fail(RCA_ERR_INVALID, "nFrame < 0");
  if (!iq || !out) return fail(RCA_ERR_INVALID, "iq/out pointers are NULL");
  if (iq_dtype != RCA_IQ_C64 && iq_dtype != RCA_IQ_F16X2)
    return fail(RCA_ERR_INVALID, "unknown iq dtype %d", iq_dtype);
  if (iq_dtype == RCA_IQ_F16X2 && !trace_scale)
    return fail(RCA_ERR_INVALID, "trace_scale is required for FP16 I/Q");
  if (flags & ~(RCA_FLAG_COMPOUND | RCA_FLAG_ACCUMULATE | RCA_FLAG_NOSUM))
    return fail(RCA_ERR_INVALID, "unknown flags 0x%x", flags);
  if ((flags & RCA_FLAG_NOSUM) && (flags & RCA_FLAG_COMPOUND))
    return fail(RCA_ERR_INVALID, "RCA_FLAG_NOSUM excludes RCA_FLAG_COMPOUND");
  if (n_frame == 0 || d->n_pix == 0) return RCA_OK;
  const size_t sbytes = iq_dtype == RCA_IQ_C64 ? 8 : 4;
  const bool tiled = d->depth_run >= 64;
  // whole-trace rows (no depth tiles) must fit LDS beside one tilt's transmit terms
  const auto trace_fits = [&](int fb, int v) {
    return tiled || (size_t)fb * d->n_sample * sbytes + (size_t)v * kThreads * sizeof(TxTerm) <=
                        (size_t)kMaxLds;
  };
  if (!trace_fits(1, 1))
    return fail(RCA_ERR_UNSUPPORTED, "nSample=%d exceeds the LDS trace capacity (%d bytes)",
                d->n_sample, kMaxLds);
  KArgs a = make_args(d, n_frame);
  a.iq = iq;
  a.scale = trace_scale;
  a.out = reinterpret_cast<float2*>(out);
  a.mask = mask;
  a.compound = (flags & RCA_FLAG_COMPOUND) != 0;
  a.accumulate = (flags & RCA_FLAG_ACCUMULATE) != 0;
  a.nosum = (flags & RCA_FLAG_NOSUM) != 0;
  a.vec_ok = (((uintptr_t)iq & 15) == 0) && (((size_t)d->n_sample * sbytes) % 16 == 0);
  const int tfb = g_tune_fb.load(), tv = g_tune_v.load();
  int fb = tfb ? tfb : pick_fb(n_frame, d->n_sample, sbytes, a.depth_run != 0);
  const int nfb = (n_frame + fb - 1) / fb;
  int v = tv ? tv : pick_v(d->n_pix, a.depth_run, nfb, fb, a.compound != 0);
  if (v > 2) v = 2;                                  // instantiated shapes
  if (!a.compound || fb > 1) v = 1;
  if (!a.compound && fb > 1) fb = fb >= 4 ? 4 : 1;
  while (fb > 1 && !trace_fits(fb, v)) fb >>= 1;    // a tuned FB that does not fit
  if (!trace_fits(fb, v)) v = 1;
  if (d->ab_law_mode == RCA_ABLAW_PIXEL_CHANNEL || d->ab_law_mode == RCA_ABLAW_TX_CHANNEL) {
    v = 1;                                           // the instantiated law shapes
    fb = fb >= 4 ? 4 : 1;
  }
  if (a.nosum && !a.accumulate && d->ab_law_mode != RCA_ABLAW_PIXEL_CHANNEL) {
    // the tilt-chunked NoSum kernel stores only the updates it computes
    const long long n = 2LL * d->n_channel * n_frame * d->n_tx * d->n_pix;
    hipLaunchKernelGGL(fill_zero, dim3(4096), dim3(256), 0, stream, reinterpret_cast<float*>(out), n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "fill_zero launch");
  }
  if (iq_dtype == RCA_IQ_C64) return launch_fwd_s<float2>(a, d->interp, v, fb, stream);
  return launch_fwd_s<__half2>(a, d->interp, v, fb, stream);
}

int rca_adjoint(const rca_das_desc* d, const void* grad_out, int grad_tx, int n_frame,
                void* grad_iq, int flags, hipStream_t stream) {
  return adjoint_run(d, grad_out, grad_tx, n_frame, grad_iq, flags, nullptr, stream);
}

size_t rca_adjoint_det_workspace_bytes(const rca_das_desc* d, int n_frame) {
  if (!d || n_frame < 0 || d->n_tx < 0 || d->n_channel < 0 || d->n_sample < 0) return 0;
  return kDetHeader + 16 * (size_t)n_frame * d->n_tx * d->n_channel * d->n_sample;
}

int rca_adjoint_det(const rca_das_desc* d, const void* grad_out, int grad_tx, int n_frame,
                    void* grad_iq, int flags, void* workspace, size_t workspace_bytes,
                    hipStream_t stream) {
  g_err.clear();
  if (!workspace) return fail(RCA_ERR_INVALID, "workspace is NULL");
  if (((uintptr_t)workspace & 255) != 0) return fail(RCA_ERR_INVALID, "workspace must be 256-byte aligned");
  if (!d) return fail(RCA_ERR_INVALID, "desc is NULL");
  const size_t need = rca_adjoint_det_workspace_bytes(d, n_frame);
  if (workspace_bytes < need)
    return fail(RCA_ERR_INVALID, "workspace of %zu bytes, %zu needed", workspace_bytes, need);
  return adjoint_run(d, grad_out, grad_tx, n_frame, grad_iq, flags, workspace, stream);
}

int bfVirtualSourceGeneric_RCA(float* dataBf, float* mask, const float* data,
                               const float* gridBfX, const float* gridBfY, const float* gridBfZ,
                               const float* posEleX, const float* posEleY, const float* posEleZ,
                               const float* virtualSourceX, const float* virtualSourceY,
                               const float* virtualSourceZ, const float* minDistVirtualSource,
                               int flagReceiveX, float apertureX, float apertureY, float apertureZ,
                               float samplingFrequency, float demodulationFrequency,
                               float speedOfSound, float startTime, float fnumber, int nPix,
                               int nChannel, int nTx, int nFrame, int nSample,
                               int flagInterpolationLanczos, int flagMaskTransmit,
                               hipStream_t stream) {
  rca_das_desc d;
  memset(&d, 0, sizeof(d));
  d.grid_x = gridBfX; d.grid_y = gridBfY; d.grid_z = gridBfZ;
  d.ele_x = posEleX; d.ele_y = posEleY; d.ele_z = posEleZ;
  d.vs_x = virtualSourceX; d.vs_y = virtualSourceY; d.vs_z = virtualSourceZ;
  d.min_dist = minDistVirtualSource;
  d.n_pix = nPix; d.n_channel = nChannel; d.n_tx = nTx; d.n_sample = nSample;
  d.receive_x = flagReceiveX; d.mask_transmit = flagMaskTransmit;
  d.interp = flagInterpolationLanczos ? RCA_INTERP_LANCZOS_REF : RCA_INTERP_QUADRATIC;
  d.aperture_x = apertureX; d.aperture_y = apertureY; d.aperture_z = apertureZ;
  d.sampling_frequency = samplingFrequency; d.demodulation_frequency = demodulationFrequency;
  d.speed_of_sound = speedOfSound; d.start_time = startTime; d.fnumber = fnumber;
  if (!dataBf || !mask) return fail(RCA_ERR_INVALID, "dataBf/mask pointers are NULL");
  return rca_forward(&d, data, RCA_IQ_C64, nullptr, nFrame, dataBf, mask, 0, stream);
}

// ---------------------------------------------------------------------------
// Batched launches for the serving loop (persistent, graph-capturable).
// ---------------------------------------------------------------------------
// LDS layout of a batch item type, shared by every entry of that type so the
// launch knows its size: with depth tiles, the standalone frame-block
// kernels' layout (rows of kTileRow samples, 48 KB budget); with whole-trace
// rows, the standalone budget (32 KB for FP32 one-frame items - at 48 KB only
// 3 workgroups fit a CU - and 48 KB for FP16 frame blocks) or one tilt of
// whole traces of n_sample_max samples when that is larger.
static FwdLayout batch_layout(int n_sample_max, int type) {
  const int fb = batch_fb(type);
  const size_t sb = batch_f32(type) ? 8 : 4;
  if (batch_tiled(type))
    return rca_tile::fwd_layout(64, n_sample_max, 1, fb, sb, tc_max(fb, (int)sb), kFwdLds);
  const size_t ns = ((size_t)n_sample_max + 3) / 4 * 4;
  const size_t need = (size_t)kThreads * sizeof(TxTerm) + (size_t)fb * ns * sb;
  const size_t dflt = fb == 1 ? kFwdLds1 : kFwdLds;
  return rca_tile::fwd_layout(0, n_sample_max, 1, fb, sb, tc_max(fb, (int)sb), need > dflt ? need : dflt);
}

static size_t batch_args_off() { return 256; }
static size_t batch_items_off(int max_halves) {
  return (batch_args_off() + (size_t)max_halves * sizeof(BatchArgs) + 255) / 256 * 256;
}

size_t rca_batch_workspace_bytes(int max_halves, int max_items) {
  if (max_halves < 1 || max_items < 1) return 0;
  return batch_items_off(max_halves) + (size_t)max_items * sizeof(int4);
}

int rca_batch_encode(const rca_batch_entry* e, int n_entries, int max_halves, int max_items,
                     int n_sample_max, void* host_ws, size_t ws_bytes, int* n_items_out) {
  g_err.clear();
  if (!host_ws || !n_items_out || n_entries < 0)
    return fail(RCA_ERR_INVALID, "null workspace / counter or negative entry count");
  if (ws_bytes < rca_batch_workspace_bytes(max_halves, max_items))
    return fail(RCA_ERR_INVALID, "workspace too small for %d halves / %d items", max_halves, max_items);
  unsigned char* ws = reinterpret_cast<unsigned char*>(host_ws);
  BatchArgs* args = reinterpret_cast<BatchArgs*>(ws + batch_args_off());
  int4* items = reinterpret_cast<int4*>(ws + batch_items_off(max_halves));
  int na = 0, ni = 0;
  int count[kBatchTypes] = {0, 0, 0, 0};
  for (int q = 0; q < n_entries; ++q) {
    if (e[q].iq_dtype != RCA_IQ_C64 && e[q].iq_dtype != RCA_IQ_F16X2)
      return fail(RCA_ERR_INVALID, "entry %d: unknown iq dtype", q);
    if (e[q].n_half < 1 || e[q].n_half > 2)
      return fail(RCA_ERR_INVALID, "entry %d: n_half must be 1 or 2", q);
    if (!e[q].half[0]) return fail(RCA_ERR_INVALID, "entry %d: desc is NULL", q);
  }
  // the items grouped by type, in type order
  for (int pass = 0; pass < kBatchTypes; ++pass)
  for (int q = 0; q < n_entries; ++q) {
    const rca_batch_entry& E = e[q];
    const bool tiled = E.half[0]->depth_run >= 64;
    const int type = batch_type(E.iq_dtype == RCA_IQ_C64, tiled);
    if (type != pass) continue;
    const int fb = batch_fb(type);
    if (!E.out || E.n_frame < 1 || E.frame_begin < 0 || E.frame_end > E.n_frame ||
        E.frame_begin >= E.frame_end || E.frame_begin % fb ||
        (E.frame_end % fb && E.frame_end != E.n_frame))
      return fail(RCA_ERR_INVALID, "entry %d: bad output or frame range [%d, %d) of %d", q,
                  E.frame_begin, E.frame_end, E.n_frame);
    if (na + E.n_half > max_halves) return fail(RCA_ERR_UNSUPPORTED, "batch exceeds %d halves", max_halves);
    const int first = na;
    int n_pix = -1, depth_run = 0;
    for (int h = 0; h < E.n_half; ++h) {
      const rca_das_desc* d = E.half[h];
      int rc = check_desc(d);
      if (rc) return rc;
      if (d->ab_law_mode == RCA_ABLAW_PIXEL_CHANNEL || d->ab_law_mode == RCA_ABLAW_TX_CHANNEL)
        return fail(RCA_ERR_UNSUPPORTED, "entry %d: per-(pixel, line) and per-(tilt, line) laws are not batched", q);
      const int dr = d->depth_run >= 64 ? d->depth_run : 0;
      if (n_pix >= 0 && (d->n_pix != n_pix || dr != depth_run))
        return fail(RCA_ERR_INVALID, "entry %d: halves differ in nPix or depth_run", q);
      n_pix = d->n_pix;
      depth_run = dr;
      if (d->n_sample > n_sample_max)
        return fail(RCA_ERR_UNSUPPORTED, "entry %d: nSample %d > batch maximum %d", q, d->n_sample, n_sample_max);
      if (!E.iq[h] || (E.iq_dtype == RCA_IQ_F16X2 && !E.trace_scale[h]))
        return fail(RCA_ERR_INVALID, "entry %d: missing I/Q or scale of half %d", q, h);
      KArgs a = make_args(d, E.n_frame);
      a.iq = E.iq[h];
      a.scale = E.trace_scale[h];
      a.out = reinterpret_cast<float2*>(E.out);
      a.mask = E.mask;
      a.compound = 1;
      a.accumulate = h > 0;
      a.nosum = 0;
      const size_t sb = E.iq_dtype == RCA_IQ_C64 ? 8 : 4;
      a.vec_ok = (((uintptr_t)E.iq[h] & 15) == 0) && (((size_t)d->n_sample * sb) % 16 == 0);
      if (tiled) {
        const FwdLayout L = batch_layout(n_sample_max, type);
        a.tc = L.tc;
        a.wcap = L.wcap;
      } else {   // whole traces of this entry in rows sized by the type's budget
        const FwdLayout L = rca_tile::fwd_layout(0, a.n_sample, 1, fb, sb, tc_max(fb, (int)sb),
                                                 batch_layout(n_sample_max, type).lds);
        a.tc = L.tc;
        a.wcap = L.wcap;
      }
      args[na].k = a;
      args[na].iq_dtype = E.iq_dtype;
      ++na;
    }
    const long long nvb = rca_tile::fwd_blocks(n_pix, depth_run, 1);
    for (int fbk = E.frame_begin / fb; fbk * fb < E.frame_end; ++fbk) {
      for (long long vb = 0; vb < nvb; ++vb) {
        if (ni >= max_items) return fail(RCA_ERR_UNSUPPORTED, "batch exceeds %d items", max_items);
        items[ni++] = make_int4(first, E.n_half, (int)vb, fbk);
        ++count[type];
      }
    }
  }
  for (int t = 0; t < kBatchTypes; ++t) reinterpret_cast<int*>(ws)[t] = count[t];
  *n_items_out = ni;
  return RCA_OK;
}

int rca_batch_launch(const void* dev_ws, int max_halves, int max_items, int interp,
                     int n_sample_max, hipStream_t stream) {
  g_err.clear();
  if (!dev_ws || max_halves < 1 || max_items < 1 || n_sample_max < 1)
    return fail(RCA_ERR_INVALID, "bad batch launch arguments");
  size_t lds[kBatchTypes];
  for (int t = 0; t < kBatchTypes; ++t) {
    lds[t] = batch_layout(n_sample_max, t).lds;
    if (lds[t] > (size_t)kMaxLds) return fail(RCA_ERR_UNSUPPORTED, "nSample %d exceeds LDS", n_sample_max);
  }
  const unsigned char* ws = reinterpret_cast<const unsigned char*>(dev_ws);
  const int* hdr = reinterpret_cast<const int*>(ws);
  const BatchArgs* args = reinterpret_cast<const BatchArgs*>(ws + batch_args_off());
  const int4* items = reinterpret_cast<const int4*>(ws + batch_items_off(max_halves));
  const dim3 grid(max_items), block(kThreads);
  switch (interp) {
#define RCA_BATCH_CASE(K)                                                                       \
    case K:                                                                                     \
      for (int h = 0; h < 2; ++h)                                                               \
        hipLaunchKernelGGL((das_fwd_batch<K, 0>), grid, block, lds[0], stream, hdr, args, items, h); \
      for (int h = 0; h < 2; ++h)                                                               \
        hipLaunchKernelGGL((das_fwd_batch<K, 1>), grid, block, lds[1], stream, hdr, args, items, h); \
      for (int h = 0; h < 2; ++h)                                                               \
        hipLaunchKernelGGL((das_fwd_batch<K, 2>), grid, block, lds[2], stream, hdr, args, items, h); \
      for (int h = 0; h < 2; ++h)                                                               \
        hipLaunchKernelGGL((das_fwd_batch<K, 3>), grid, block, lds[3], stream, hdr, args, items, h); \
      break;
    RCA_BATCH_CASE(RCA_INTERP_QUADRATIC)
    RCA_BATCH_CASE(RCA_INTERP_LANCZOS_REF)
    RCA_BATCH_CASE(RCA_INTERP_LANCZOS3)
    RCA_BATCH_CASE(RCA_INTERP_LANCZOS4)
#undef RCA_BATCH_CASE
    default: return fail(RCA_ERR_INVALID, "unknown interpolation %d", interp);
  }
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? RCA_OK : hip_fail(err, "das_fwd_batch launch");
}

int rca_batch_frame_block(int iq_dtype, int depth_tiled) {
  if (iq_dtype != RCA_IQ_C64 && iq_dtype != RCA_IQ_F16X2) return 0;
  return batch_fb(batch_type(iq_dtype == RCA_IQ_C64, depth_tiled != 0));
}

}  // extern "C"
