// SIMUS-style RCA simulator on MI355X (gfx950): the upstream generator of the
// beamformer's data.
//
// Reference: MATLAB/lib/simulator/SimUS_CUDARCA.cu (kernels TXfield :209-298
// and RXfield :301-379) and SimUS_CUDARCA.m (wrapper :136-390), with the
// receive-aberration variant SimUS_CUDARCA_ab.m.  Frequency domain: for every
// frequency f of the grid, scatterer s and element e,
//   term = exp(j k r)/r x sinc(k w sin(az)/2) x elevation x exp(j w lens)
// summed over the element's sub-elements (k complex: wavenumber + j
// attenuation; elevation = the same sinc with the width, or the four-Gaussian
// focused term when the elevation focus is <= 10 m).  Transmit weighs the
// elements with Apod exp(j w Delay) and the pulse and impulse-response
// spectra; receive multiplies by the impulse response, the apodisation and
// the input filter, sums over scatterers against the transmit field, and an
// inverse FFT gives the time signal, decimated per sample mode.
//
// Kernels (hand-written):
//   sim_tx       TX per (tilt, scatterer, frequency); lanes = scatterers, each
//                wave a block of F frequencies, so element positions,
//                wavenumbers and laws are wave-uniform (scalar registers)
//   sim_rx       the fused receive: lanes = receive elements, a workgroup a
//                frequency block whose four waves take every fourth
//                scatterer (wave-uniform position and TX value): the
//                sub-element sum accumulated against TX, the four partial
//                sums added in LDS; writes the conjugated, zero-padded
//                spectra (and the aberrated ones)
//   sim_rxfield  the reference RXfield output, materialised (reference ABI)
//   sim_raw      real part / N, trim or pad, sample-mode decimation
// The inverse FFT is hipFFT.  Distances and every phase (k r, w Delay,
// w lens, the focused term's k y^2/2r) are formed in FP64 and reduced to
// [-1/2, 1/2] cycle before the FP32 sine and cosine; the rest is FP32.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>

#include "../../include/rca_das.h"
#include "rca_math.h"
#include "../../include/rca_sim.h"

namespace rca_internal {
void set_error(const char* msg);
void clear_error();
}  // namespace rca_internal

namespace {

using rca_math::sincos2pi_lean;
using rca_math::sqrt_nr;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  rca_internal::set_error(buf);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(RCA_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

constexpr int kThreads = 256;
// Frequencies per lane.  The kernels are instantiated per elevation model
// (flat sinc or focused four-Gaussian term) and baffle, so each carries only
// its own path's registers, and the block size is chosen per model
// (measured): flat 6 (about 100 VGPRs, 5 waves per SIMD; 8 is 1 % slower),
// focused 2 (about 140 VGPRs; the four-Gaussian term holds ~47 VGPRs per
// frequency, so 4 frequencies take ~240).
#ifndef RCA_SIM_F_FLAT
#define RCA_SIM_F_FLAT 6
#endif
#ifndef RCA_SIM_F_FOCUSED
#define RCA_SIM_F_FOCUSED 2
#endif
template <bool FLAT>
constexpr int block_f() { return FLAT ? RCA_SIM_F_FLAT : RCA_SIM_F_FOCUSED; }
constexpr float kPiRef = 3.14159265359f;         // the kernels' float PI (SimUS_CUDARCA.cu:8)
constexpr double kInv2Pi = 0.15915494309189533577;

// ----------------------------------------------------------------------------
// Device math
// ----------------------------------------------------------------------------

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// exp(2 pi j x) for a phase given in cycles (FP64), reduced first.
__device__ __forceinline__ float2 cis_cycles(double x) {
  float s, c;
  sincos2pi_lean((float)(x - rint(x)), s, c);
  return make_float2(c, s);
}

// sin(x)/x of x = a + jb: 1 at 0, 0 where |x| >= pi (SimUS_CUDARCA.cu:43-66).
// sin(a + jb) = sin a cosh b + j cos a sinh b.  |a| < pi here, so sin a and
// cos a come from the half-cycle polynomial; cosh and sinh of small b from
// their series (no cancellation: two terms below 1e-2, where the next is
// under 1e-9 relative; five below 0.3), of larger b from exp.  The division
// by x uses the hardware reciprocal (1 ulp) of |x|^2; below |x|^2 = 1e-12
// the value is 1 (the exact one differs by |x|^2/6 < 2e-13).
__device__ __forceinline__ float2 csinc(float a, float b) {
  const float m2 = fmaf(a, a, b * b);
  if (!(m2 > 1e-12f)) return make_float2(1.0f, 0.0f);
  if (!(m2 < kPiRef * kPiRef)) return make_float2(0.0f, 0.0f);
  float sa, ca;
  sincos2pi_lean(a * 0.159154943091895f, sa, ca);
  float ch, sh;
  const float ab = fabsf(b);
  if (ab < 1e-2f) {
    const float b2 = b * b;
    ch = fmaf(b2, 0.5f, 1.0f);
    sh = b * fmaf(b2, 1.66666667e-1f, 1.0f);
  } else if (ab < 0.3f) {
    const float b2 = b * b;
    ch = fmaf(b2, fmaf(b2, fmaf(b2, fmaf(b2, 2.48015873e-5f, 1.38888889e-3f), 4.16666667e-2f), 0.5f),
              1.0f);
    sh = b * fmaf(b2, fmaf(b2, fmaf(b2, 1.98412698e-4f, 8.33333333e-3f), 1.66666667e-1f), 1.0f);
  } else {
    const float e = expf(b), ie = 1.0f / e;
    ch = 0.5f * (e + ie);
    sh = 0.5f * (e - ie);
  }
  const float nr = sa * ch, ni = ca * sh;
  const float inv = __builtin_amdgcn_rcpf(m2);
  return make_float2((nr * a + ni * b) * inv, (ni * a - nr * b) * inv);
}

// Principal square root, in the form without cancellation on either half plane.
__device__ __forceinline__ float2 csqrt(float2 z) {
  const float m = sqrtf(fmaf(z.x, z.x, z.y * z.y));
  if (z.x >= 0.0f) {
    const float t = sqrtf(0.5f * (m + z.x));
    return make_float2(t, z.y / (2.0f * t));
  }
  const float t = sqrtf(0.5f * (m - z.x));
  return make_float2(fabsf(z.y) / (2.0f * t), copysignf(t, z.y));
}

// Elevation-focused directivity (SimUS_CUDARCA.cu:166-203): sum over four
// Gaussians of A_g sqrt(pi/alpha_g) exp(beta^2/(4 alpha_g) + gamma),
// alpha_g = B_g/h^2 + j k (r - rf)/(2 r rf), beta = -j k y/r, gamma = j k y^2/(2r).
// The coefficient table is the reference's (.cu:10-13).
__device__ __forceinline__ float2 dir_curv(float kr, float ki, double cyc, double r, double y,
                                           float yr, float inv_h2, double rfocus) {
  constexpr float Ar[4] = {0.187f, 0.288f, 0.187f, 0.288f};
  constexpr float Ai[4] = {0.275f, -1.954f, -0.275f, 1.954f};
  constexpr float Br[4] = {4.558f, 8.598f, 4.558f, 8.598f};
  constexpr float Bi[4] = {-25.59f, -7.924f, 25.59f, 7.924f};
  const float q = (float)((r - rfocus) / (r * rfocus) * 0.5);
  const double y2r = y * y / (2.0 * r);
  const float yr2 = yr * yr;
  const float al_r = -ki * q, al_i = kr * q;                 // alpha = j k q
  const float b2r = (ki * ki - kr * kr) * yr2;               // beta^2 = -(k y/r)^2
  const float b2i = -2.0f * kr * ki * yr2;
  const float g_re = -ki * (float)y2r;                       // gamma = j k y^2/2r:
  const double gc = cyc * y2r;                               //   phase in cycles, FP64
  const float g_ph = 6.28318530717959f * (float)(gc - rint(gc));
  float2 out = make_float2(0.0f, 0.0f);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float ar = fmaf(Br[g], inv_h2, al_r), ai = fmaf(Bi[g], inv_h2, al_i);
    const float m2 = fmaf(ar, ar, ai * ai);
    const float i4 = 0.25f / m2;                              // beta^2/(4 alpha) = beta^2 conj(alpha)/(4|alpha|^2)
    const float tr = (b2r * ar + b2i * ai) * i4;
    const float ti = (b2i * ar - b2r * ai) * i4;
    const float mag = expf(tr + g_re);
    float s, c;
    sincosf(ti + g_ph, &s, &c);
    const float ip = kPiRef / m2;                             // pi/alpha = pi conj(alpha)/|alpha|^2
    const float2 sq = csqrt(make_float2(ar * ip, -ai * ip));
    const float2 v = cmul(sq, make_float2(mag * c, mag * s));
    out = cadd(out, cmul(make_float2(Ar[g], Ai[g]), v));
  }
  return out;
}

// Geometry of one (scatterer, sub-element) pair: distance in FP64, the sines
// and cosines of az = asin(dx/r) - Az and el = asin(dy/r) - El
// (SimUS_CUDARCA.cu:262-264) by the angle-difference identities, and the
// focused term's y = r sin(el).
struct Geo {
  double r, y;
  float ir, saz, caz, sel, cel;
};

__device__ __forceinline__ Geo geometry(double sx, double sy, double sz, const float* p, int n_ele) {
  Geo g;
  const double dx = sx - (double)p[0], dy = sy - (double)p[n_ele], dz = sz - (double)p[2 * n_ele];
  g.r = sqrt_nr(fma(dx, dx, fma(dy, dy, dz * dz)));
  const float rf = (float)g.r;
  g.ir = __builtin_amdgcn_rcpf(rf);   // 1 ulp
  const float ux = (float)dx * g.ir, uy = (float)dy * g.ir;
  // hardware square root (1 ulp): the correctly rounded one costs a sequence
  const float cx = __builtin_amdgcn_sqrtf(fmaxf(1.0f - ux * ux, 0.0f));
  const float cy = __builtin_amdgcn_sqrtf(fmaxf(1.0f - uy * uy, 0.0f));
  const float az0 = p[3 * n_ele], el0 = p[4 * n_ele];
  if (az0 == 0.0f) {
    g.saz = ux;
    g.caz = cx;
  } else {
    float sa, ca;
    sincosf(az0, &sa, &ca);
    g.saz = ux * ca - cx * sa;
    g.caz = cx * ca + ux * sa;
  }
  if (el0 == 0.0f) {
    g.sel = uy;
    g.cel = cy;
    g.y = dy;
  } else {
    float se, ce;
    sincosf(el0, &se, &ce);
    g.sel = uy * ce - cy * se;
    g.cel = cy * ce + uy * se;
    g.y = g.r * (double)g.sel;
  }
  return g;
}

// One sub-element's field at one frequency, without the lens and spectra:
// exp(j k r)/r x azimuth directivity x elevation directivity
// (SimUS_CUDARCA.cu:267-274).
__device__ __forceinline__ float2 sub_term(const Geo& g, float kr, float ki, double cyc, float hw,
                                           bool baffle, bool flat, float inv_h2, double rfocus) {
  float2 v = cis_cycles(g.r * cyc);
  // exp(-Im(k) r) / r, with the hardware base-2 exponential
  const float att = __builtin_amdgcn_exp2f(-1.44269504088896f * ki * (float)g.r) * g.ir;
  v = cscale(v, att);
  const float xa = g.saz * hw;
  float2 d = csinc(xa * kr, xa * ki);
  if (baffle) d = cscale(d, g.caz);
  v = cmul(v, d);
  if (flat) {
    const float xe = g.sel * hw;
    float2 de = csinc(xe * kr, xe * ki);
    if (baffle) de = cscale(de, g.cel);
    v = cmul(v, de);
  } else {
    v = cmul(v, dir_curv(kr, ki, cyc, g.r, g.y, g.sel, inv_h2, rfocus));
  }
  return v;
}

struct Tr {            // one transducer's scalars
  const float* ele;    // [nSub][5][nEle]
  int n_ele, n_sub;
  float hw, inv_h2;    // width/2, 1/height^2
  double rfocus;
  bool baffle, flat;
  double lens, c;      // lens delay (s), speed of sound
};

Tr make_tr(const float* ele, int n_ele, int n_sub, float width, float height, float rfocus,
           float baffle, float lens, float c) {
  Tr t;
  t.ele = ele;
  t.n_ele = n_ele;
  t.n_sub = n_sub;
  t.hw = 0.5f * width;
  t.inv_h2 = 1.0f / (height * height);
  t.rfocus = rfocus;
  t.baffle = baffle == 1.0f;    // .cu:125 (baffle==1)
  t.flat = rfocus > 10.0f;      // .cu:269
  t.lens = lens;
  t.c = c;
  return t;
}

// Per-lane frequency constants of a wave's block.
template <int KF>
struct FBlock {
  float kr[KF], ki[KF];
  double cyc[KF];   // Re(k)/(2 pi): cycles per metre
};

template <int KF>
__device__ __forceinline__ FBlock<KF> load_block(const float2* k, int f0, int n_f) {
  FBlock<KF> b;
#pragma unroll
  for (int i = 0; i < KF; ++i) {
    const float2 kk = k[min(f0 + i, n_f - 1)];
    b.kr[i] = kk.x;
    b.ki[i] = kk.y;
    b.cyc[i] = (double)kk.x * kInv2Pi;
  }
  return b;
}

// Sum over one element's sub-elements, for the block's frequencies (FLAT and
// BAFFLE are the transducer's T.flat and T.baffle, fixed per instantiation).
template <bool FLAT, bool BAFFLE, int KF>
__device__ __forceinline__ void element_sum(const Tr& T, const FBlock<KF>& B, int e, double sx,
                                            double sy, double sz, float2 (&h)[KF]) {
#pragma unroll
  for (int i = 0; i < KF; ++i) h[i] = make_float2(0.0f, 0.0f);
  for (int u = 0; u < T.n_sub; ++u) {
    const float* p = T.ele + e + (long long)u * 5 * T.n_ele;
    const Geo g = geometry(sx, sy, sz, p, T.n_ele);
#pragma unroll
    for (int i = 0; i < KF; ++i)
      h[i] = cadd(h[i], sub_term(g, B.kr[i], B.ki[i], B.cyc[i], T.hw, BAFFLE, FLAT, T.inv_h2,
                                 T.rfocus));
  }
}

// ----------------------------------------------------------------------------
// Transmit field (SimUS_CUDARCA.cu:209-298)
// ----------------------------------------------------------------------------
struct TxArgs {
  Tr T;
  float2* out;
  const float* r_scat;
  const float2 *k, *pulse, *ir;
  const float *delay, *apod, *bsc;
  int n_scat, n_f, n_tx;
};

template <bool FLAT, bool BAFFLE>
__global__ void __launch_bounds__(kThreads) sim_tx(const TxArgs a) {
  constexpr int kF = block_f<FLAT>();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int f0 = (blockIdx.y * 4 + wave) * kF;
  if (f0 >= a.n_f) return;
  const int t = blockIdx.z;
  const int s = blockIdx.x * 64 + lane;
  const int sc = min(s, a.n_scat - 1);
  const float* rs = a.r_scat + (long long)t * 3 * a.n_scat;
  const double sx = rs[sc], sy = rs[sc + a.n_scat], sz = rs[sc + 2 * a.n_scat];
  const FBlock<kF> B = load_block<kF>(a.k, f0, a.n_f);
  const int ne = a.T.n_ele;
  float2 acc[kF];
#pragma unroll
  for (int i = 0; i < kF; ++i) acc[i] = make_float2(0.0f, 0.0f);
  for (int e = 0; e < ne; ++e) {
    float2 h[kF];
    element_sum<FLAT, BAFFLE>(a.T, B, e, sx, sy, sz, h);
    // transmit wave front Apod exp(j w Delay) (.cu:281-285), w = Re(k) c
    const float ap = a.apod[e + t * ne];
    const double dl = (double)a.delay[e + t * ne] * a.T.c;
#pragma unroll
    for (int i = 0; i < kF; ++i) acc[i] = cadd(acc[i], cmul(h[i], cscale(cis_cycles(B.cyc[i] * dl), ap)));
  }
  if (s >= a.n_scat) return;
  const double lc = a.T.lens * a.T.c;
#pragma unroll
  for (int i = 0; i < kF; ++i) {
    const int f = f0 + i;
    if (f >= a.n_f) break;
    // lens phase and conj(PULSE IR), common to every term (.cu:275-277)
    const float2 p = a.pulse[f], r = a.ir[f];
    const float2 spec = cmul(make_float2(p.x, -p.y), make_float2(r.x, -r.y));
    float2 v = cmul(acc[i], cmul(spec, cis_cycles(B.cyc[i] * lc)));
    if (a.bsc) {   // wrapper: TX .* bsc, NaN -> 0 (SimUS_CUDARCA.m:331-335)
      v = cscale(v, a.bsc[s + (long long)t * a.n_scat]);
      if (isnan(v.x) || isnan(v.y)) v = make_float2(0.0f, 0.0f);
    }
    a.out[f + (long long)s * a.n_f + (long long)t * a.n_f * a.n_scat] = v;
  }
}

// ----------------------------------------------------------------------------
// Receive field, reference layout (SimUS_CUDARCA.cu:301-379)
// ----------------------------------------------------------------------------
struct RxfArgs {
  Tr T;
  float2* out;
  const float* r_scat;
  const float2 *k, *ir;
  int n_scat, n_f;
};

template <bool FLAT, bool BAFFLE>
__global__ void __launch_bounds__(kThreads) sim_rxfield(const RxfArgs a) {
  constexpr int kF = block_f<FLAT>();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int f0 = (blockIdx.y * 4 + wave) * kF;
  if (f0 >= a.n_f) return;
  const int s = blockIdx.z;
  const int ne = a.T.n_ele;
  const int e = blockIdx.x * 64 + lane;
  const int ec = min(e, ne - 1);
  const double sx = a.r_scat[s], sy = a.r_scat[s + a.n_scat], sz = a.r_scat[s + 2 * a.n_scat];
  const FBlock<kF> B = load_block<kF>(a.k, f0, a.n_f);
  float2 h[kF];
  element_sum<FLAT, BAFFLE>(a.T, B, ec, sx, sy, sz, h);
  if (e >= ne) return;
  const double lc = a.T.lens * a.T.c;
#pragma unroll
  for (int i = 0; i < kF; ++i) {
    const int f = f0 + i;
    if (f >= a.n_f) break;
    a.out[f + (long long)e * a.n_f + (long long)s * a.n_f * ne] =
        cmul(h[i], cmul(a.ir[f], cis_cycles(B.cyc[i] * lc)));
  }
}

// ----------------------------------------------------------------------------
// Fused receive and sum over scatterers (SimUS_CUDARCA.m:338-363)
// ----------------------------------------------------------------------------
struct RxArgs {
  Tr T;
  const float2* tx;
  const float* r_scat;
  const float2 *k, *ir, *ipf;
  const float *agl, *apod, *ab_delay, *ab_apod;
  float2 *spec, *spec_ab;
  int n_scat, n_f, n_tx, nfft;
};

template <bool FLAT, bool BAFFLE>
__global__ void __launch_bounds__(kThreads) sim_rx(const RxArgs a) {
  constexpr int kF = block_f<FLAT>();
  // the four waves of a workgroup share one frequency block and take every
  // fourth scatterer; their sums meet in LDS (fixed order)
  __shared__ float2 s_part[3][kF][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int f0 = blockIdx.y * kF;
  const int t = blockIdx.z;
  const int ne = a.T.n_ele;
  const int e = blockIdx.x * 64 + lane;
  const int ec = min(e, ne - 1);
  const FBlock<kF> B = load_block<kF>(a.k, f0, a.n_f);
  const float* rs = a.r_scat + (long long)t * 3 * a.n_scat;
  const float2* txt = a.tx + (long long)t * a.n_scat * a.n_f;
  float2 acc[kF];
#pragma unroll
  for (int i = 0; i < kF; ++i) acc[i] = make_float2(0.0f, 0.0f);
  for (int s = wave; s < a.n_scat; s += 4) {
    float2 tv[kF];
    bool any = false;
#pragma unroll
    for (int i = 0; i < kF; ++i) {
      tv[i] = txt[(long long)s * a.n_f + min(f0 + i, a.n_f - 1)];
      any |= tv[i].x != 0.0f || tv[i].y != 0.0f;
    }
    if (!any) continue;   // wave-uniform: a scatterer with no transmit field adds nothing
    const double sx = rs[s], sy = rs[s + a.n_scat], sz = rs[s + 2 * a.n_scat];
    float2 h[kF];
    element_sum<FLAT, BAFFLE>(a.T, B, ec, sx, sy, sz, h);
#pragma unroll
    for (int i = 0; i < kF; ++i) acc[i] = cadd(acc[i], cmul(tv[i], h[i]));
  }
  if (wave > 0) {
#pragma unroll
    for (int i = 0; i < kF; ++i) s_part[wave - 1][i][lane] = acc[i];
  }
  __syncthreads();
  if (wave > 0 || e >= ne) return;
#pragma unroll
  for (int i = 0; i < kF; ++i)
    acc[i] = cadd(cadd(acc[i], s_part[0][i][lane]), cadd(s_part[1][i][lane], s_part[2][i][lane]));
  const double lc = a.T.lens * a.T.c;
  const float ap = a.apod[(long long)t * ne + e];
  const long long row = ((long long)t * ne + e) * a.nfft;
  const double abd = a.ab_delay ? (double)a.ab_delay[e] * kInv2Pi : 0.0;
  const float aba = a.ab_apod ? a.ab_apod[e] : 1.0f;
#pragma unroll
  for (int i = 0; i < kF; ++i) {
    const int f = f0 + i;
    if (f >= a.n_f) break;
    // RX x IR x lens (.cu:367-368) x Apod (.m:358), summed, x IPFilt (.m:361)
    const float2 pre = cmul(cmul(a.ir[f], cis_cycles(B.cyc[i] * lc)), a.ipf[f]);
    const float2 v = cscale(cmul(acc[i], pre), ap);
    a.spec[row + f] = make_float2(v.x, -v.y);                      // conj (.m:364)
    if (a.spec_ab) {   // x AbLaw.Apod exp(j 2 pi f AbLaw.Delay) (SimUS_CUDARCA_ab.m:369-374)
      const float2 w = cmul(v, cscale(cis_cycles((double)a.agl[f] * abd), aba));
      a.spec_ab[row + f] = make_float2(w.x, -w.y);
    }
  }
}

// Bins n_f .. nfft-1 of every spectrum row: zero (the transform's padding).
__global__ void sim_zero_tail(float2* spec, long long rows, int n_f, int nfft) {
  const int tail = nfft - n_f;
  const long long n = rows * tail;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long r = i / tail;
    spec[r * nfft + n_f + (i - r * tail)] = make_float2(0.0f, 0.0f);
  }
}

// Real part / nfft, trimmed or padded to n_sample rows, samples 1 and 2 of
// every `sub` (SimUS_CUDARCA.m:365, 378-390).
__global__ void sim_raw(const float2* spec, long long n_batch, int nfft, int n_out, int sub,
                        float inv_n, float* raw) {
  const long long n = n_batch * n_out;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long b = i / n_out;
    const int j = (int)(i - b * n_out);
    const int s0 = j * sub, s1 = j * sub + 1;
    const float2* x = spec + b * nfft;
    raw[b * 2 * n_out + 2 * j] = s0 < nfft ? x[s0].x * inv_n : 0.0f;
    raw[b * 2 * n_out + 2 * j + 1] = s1 < nfft ? x[s1].x * inv_n : 0.0f;
  }
}

// hipFFT plans per (device, nfft, batch).
std::mutex g_fft_mu;
std::map<std::tuple<int, int, int>, hipfftHandle> g_plans;

dim3 field_grid(int lanes, int n_f, int z, int kf) {
  const int per_block = 4 * kf;
  return dim3((unsigned)((lanes + 63) / 64), (unsigned)((n_f + per_block - 1) / per_block),
              (unsigned)z);
}

// Launch the instantiation of a field kernel for the transducer's elevation
// model and baffle.
template <bool F, bool B>
void launch_tx(const TxArgs& a, hipStream_t st) {
  hipLaunchKernelGGL((sim_tx<F, B>), field_grid(a.n_scat, a.n_f, a.n_tx, block_f<F>()), dim3(kThreads),
                     0, st, a);
}
template <bool F, bool B>
void launch_rxfield(const RxfArgs& a, hipStream_t st) {
  hipLaunchKernelGGL((sim_rxfield<F, B>), field_grid(a.T.n_ele, a.n_f, a.n_scat, block_f<F>()),
                     dim3(kThreads), 0, st, a);
}
template <bool F, bool B>
void launch_rx(const RxArgs& a, hipStream_t st) {
  constexpr int kf = block_f<F>();
  hipLaunchKernelGGL((sim_rx<F, B>),
                     dim3((unsigned)((a.T.n_ele + 63) / 64), (unsigned)((a.n_f + kf - 1) / kf),
                          (unsigned)a.n_tx),
                     dim3(kThreads), 0, st, a);
}
#define RCA_SIM_DISPATCH(fn, T, ...)                                    \
  do {                                                                  \
    if ((T).flat) {                                                     \
      if ((T).baffle) fn<true, true>(__VA_ARGS__);                      \
      else fn<true, false>(__VA_ARGS__);                                \
    } else {                                                            \
      if ((T).baffle) fn<false, true>(__VA_ARGS__);                     \
      else fn<false, false>(__VA_ARGS__);                               \
    }                                                                   \
  } while (0)

int check_tr(const float* ele, int n_ele, int n_sub, float height, float c) {
  if (!ele) return fail(RCA_ERR_INVALID, "element positions are NULL");
  if (n_ele < 1 || n_sub < 1) return fail(RCA_ERR_INVALID, "nEle=%d nSub=%d", n_ele, n_sub);
  if (!(c > 0.0f)) return fail(RCA_ERR_INVALID, "speed of sound must be > 0");
  if (!(height != 0.0f)) return fail(RCA_ERR_INVALID, "element height must be non-zero");
  return RCA_OK;
}

}  // namespace

extern "C" {

int rca_sim_txfield(void* TX, const float* r_scat, const float* r_ele, const void* kWaveNumber,
                    const void* PULSE, const void* IR, const float* Delay, const float* Apod,
                    float width, float height, float rfocus, float baffle, float lensCorrection,
                    float SpeedOfSound, int NscatPerEvent, int Nele, int Nf, int Ntx, int Nsubele,
                    const float* bsc, hipStream_t stream) {
  rca_internal::clear_error();
  if (!TX || !r_scat || !kWaveNumber || !PULSE || !IR || !Delay || !Apod)
    return fail(RCA_ERR_INVALID, "rca_sim_txfield: NULL pointer");
  if (NscatPerEvent < 0 || Nf < 0 || Ntx < 0) return fail(RCA_ERR_INVALID, "negative size");
  int rc = check_tr(r_ele, Nele, Nsubele, height, SpeedOfSound);
  if (rc) return rc;
  if (NscatPerEvent == 0 || Nf == 0 || Ntx == 0) return RCA_OK;
  if (Ntx > 65535) return fail(RCA_ERR_UNSUPPORTED, "nTx=%d > 65535", Ntx);
  TxArgs a;
  a.T = make_tr(r_ele, Nele, Nsubele, width, height, rfocus, baffle, lensCorrection, SpeedOfSound);
  a.out = reinterpret_cast<float2*>(TX);
  a.r_scat = r_scat;
  a.k = reinterpret_cast<const float2*>(kWaveNumber);
  a.pulse = reinterpret_cast<const float2*>(PULSE);
  a.ir = reinterpret_cast<const float2*>(IR);
  a.delay = Delay;
  a.apod = Apod;
  a.bsc = bsc;
  a.n_scat = NscatPerEvent;
  a.n_f = Nf;
  a.n_tx = Ntx;
  RCA_SIM_DISPATCH(launch_tx, a.T, a, stream);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? RCA_OK : hip_fail(e, "sim_tx launch");
}

int rca_sim_rxfield(void* RX, const float* r_scat, const float* r_ele, const void* kWaveNumber,
                    const void* IR, float width, float height, float rfocus, float baffle,
                    float lensCorrection, float SpeedOfSound, int Nscat, int Nele, int Nf,
                    int Nsubele, hipStream_t stream) {
  rca_internal::clear_error();
  if (!RX || !r_scat || !kWaveNumber || !IR) return fail(RCA_ERR_INVALID, "rca_sim_rxfield: NULL pointer");
  if (Nscat < 0 || Nf < 0) return fail(RCA_ERR_INVALID, "negative size");
  int rc = check_tr(r_ele, Nele, Nsubele, height, SpeedOfSound);
  if (rc) return rc;
  if (Nscat == 0 || Nf == 0) return RCA_OK;
  if (Nscat > 65535) return fail(RCA_ERR_UNSUPPORTED, "nScat=%d > 65535 per call", Nscat);
  RxfArgs a;
  a.T = make_tr(r_ele, Nele, Nsubele, width, height, rfocus, baffle, lensCorrection, SpeedOfSound);
  a.out = reinterpret_cast<float2*>(RX);
  a.r_scat = r_scat;
  a.k = reinterpret_cast<const float2*>(kWaveNumber);
  a.ir = reinterpret_cast<const float2*>(IR);
  a.n_scat = Nscat;
  a.n_f = Nf;
  RCA_SIM_DISPATCH(launch_rxfield, a.T, a, stream);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? RCA_OK : hip_fail(e, "sim_rxfield launch");
}

int rca_sim_receive(const rca_sim_receive_args* p, hipStream_t stream) {
  rca_internal::clear_error();
  if (!p) return fail(RCA_ERR_INVALID, "rca_sim_receive: args are NULL");
  if (!p->tx || !p->r_scat || !p->k || !p->ir || !p->ipfilt || !p->apod || !p->spec)
    return fail(RCA_ERR_INVALID, "rca_sim_receive: NULL pointer");
  if (p->spec_ab && (!p->ab_delay || !p->agl))
    return fail(RCA_ERR_INVALID, "rca_sim_receive: the aberrated output needs ab_delay and agl");
  if (p->n_scat < 0 || p->n_f < 0 || p->n_tx < 0 || p->nfft < p->n_f)
    return fail(RCA_ERR_INVALID, "bad sizes: nScat=%d Nf=%d nTx=%d nfft=%d", p->n_scat, p->n_f,
                p->n_tx, p->nfft);
  int rc = check_tr(p->r_ele, p->n_ele, p->n_sub, p->height, p->speed_of_sound);
  if (rc) return rc;
  if (p->n_tx == 0 || p->nfft == 0) return RCA_OK;
  if (p->n_tx > 65535) return fail(RCA_ERR_UNSUPPORTED, "nTx=%d > 65535", p->n_tx);
  RxArgs a;
  a.T = make_tr(p->r_ele, p->n_ele, p->n_sub, p->width, p->height, p->rfocus, p->baffle, p->lens,
                p->speed_of_sound);
  a.tx = reinterpret_cast<const float2*>(p->tx);
  a.r_scat = p->r_scat;
  a.k = reinterpret_cast<const float2*>(p->k);
  a.ir = reinterpret_cast<const float2*>(p->ir);
  a.ipf = reinterpret_cast<const float2*>(p->ipfilt);
  a.agl = p->agl;
  a.apod = p->apod;
  a.ab_delay = p->spec_ab ? p->ab_delay : nullptr;
  a.ab_apod = p->spec_ab ? p->ab_apod : nullptr;
  a.spec = reinterpret_cast<float2*>(p->spec);
  a.spec_ab = reinterpret_cast<float2*>(p->spec_ab);
  a.n_scat = p->n_scat;
  a.n_f = p->n_f;
  a.n_tx = p->n_tx;
  a.nfft = p->nfft;
  const long long rows = (long long)p->n_tx * p->n_ele;
  if (p->nfft > p->n_f) {
    hipLaunchKernelGGL(sim_zero_tail, dim3(1024), dim3(256), 0, stream, a.spec, rows, p->n_f, p->nfft);
    if (a.spec_ab)
      hipLaunchKernelGGL(sim_zero_tail, dim3(1024), dim3(256), 0, stream, a.spec_ab, rows, p->n_f,
                         p->nfft);
  }
  if (p->n_f > 0) RCA_SIM_DISPATCH(launch_rx, a.T, a, stream);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? RCA_OK : hip_fail(e, "sim_rx launch");
}

int rca_sim_spectra_to_raw(void* spec, int n_batch, int nfft, int n_sample, int sub, float* raw,
                           hipStream_t stream) {
  rca_internal::clear_error();
  if (!spec || !raw) return fail(RCA_ERR_INVALID, "rca_sim_spectra_to_raw: NULL pointer");
  if (n_batch < 0 || nfft < 1 || n_sample < 0 || sub < 2)
    return fail(RCA_ERR_INVALID, "bad sizes: batch=%d nfft=%d nSample=%d sub=%d", n_batch, nfft,
                n_sample, sub);
  if (n_batch == 0) return RCA_OK;
  int dev = 0;
  hipError_t he = hipGetDevice(&dev);
  if (he != hipSuccess) return hip_fail(he, "hipGetDevice");
  {
    std::lock_guard<std::mutex> lk(g_fft_mu);
    const auto key = std::make_tuple(dev, nfft, n_batch);
    auto it = g_plans.find(key);
    hipfftHandle plan;
    if (it == g_plans.end()) {
      if (hipfftPlan1d(&plan, nfft, HIPFFT_C2C, n_batch) != HIPFFT_SUCCESS)
        return fail(RCA_ERR_HIP, "hipfftPlan1d(%d, batch %d) failed", nfft, n_batch);
      g_plans[key] = plan;
    } else {
      plan = it->second;
    }
    if (hipfftSetStream(plan, stream) != HIPFFT_SUCCESS) return fail(RCA_ERR_HIP, "hipfftSetStream failed");
    hipfftComplex* x = reinterpret_cast<hipfftComplex*>(spec);
    if (hipfftExecC2C(plan, x, x, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
      return fail(RCA_ERR_HIP, "hipfftExecC2C failed");
  }
  const int n_out = n_sample / sub;
  if (n_out > 0) {
    hipLaunchKernelGGL(sim_raw, dim3(1024), dim3(256), 0, stream,
                       reinterpret_cast<const float2*>(spec), (long long)n_batch, nfft, n_out, sub,
                       1.0f / (float)nfft, raw);
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? RCA_OK : hip_fail(e, "sim_raw launch");
}

}  // extern "C"
