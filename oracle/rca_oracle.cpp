// ORACLE — test infrastructure only.  Nothing in the product links, loads or
// calls this file; only tests, the smoke check and the CPU-baseline leg of the
// benchmark do, as the checker and the CPU reference path.
//
// CPU restatement of the row-column-addressed (RCA) delay-and-sum I/Q
// beamformer of Jeanrop/Simu_RCA_ab:
//   * kernel body  MATLAB/lib/beamformer/RCA/bfVirtualSourceGeneric_RCA.cu:79-391
//   * host prep    MATLAB/lib/beamformer/RCA/bfVirtualSourceGeneric_RCA.m:491-558
//   * variants     NoSum/bfVirtualSourceGeneric_RCA_Nosum.cu:228-229,275-278,376-380
//                  Hyp1_save/bfVirtualSourceGeneric_RCA_ab_1.cu:306-311 (AbLawRx[ch])
//                  Hyp2/bfVirtualSourceGeneric_RCA_ab_2.cu:306-311   (AbLawRx[tx*nCh+ch])
//                  exemple_paul/bfVirtualSourceGenericTxRx.cu:268-275, 292-294
//                    (AbLawRx/AbLawTx[idxPix + ch*nPix], on the RCA geometry)
//
// The arithmetic follows the reference formula term by term, in the precision
// R of the instantiation: R=double is the FP64 oracle, R=float is the FP32
// "reference CPU path" used as the CPU throughput baseline.  Inputs are the
// FP32 values the reference kernel receives (RCA.m:598-636), promoted to R.
//
// Interpolation kinds:
//   0  quadratic Lagrange on m-1,m,m+1            (RCA.cu:355-365, bounds 1/1)
//   1  reference "Lanczos 5-lobes": taps m-2..m+3,
//      sinc(x)*sinc(x/2), not truncated           (RCA.cu:331-350, bounds 2/3)
//   2  textbook Lanczos a: taps m-(a-1)..m+a, sinc(x)*sinc(x/a),
//      bounds a-1 / a (the reference bounds generalised; a=3 gives the
//      reference's tap set with window sinc(x/3))
//
// Build: oracle/Makefile (g++ -O3 -fopenmp, no fast-math).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <omp.h>

namespace {

template <typename R> inline R pi_r();
template <> inline double pi_r<double>() { return 3.14159265358979323846; }
template <> inline float pi_r<float>() { return 3.14159265358979323846f; }

// sin(pi*x) with exact reduction by the nearest integer.
template <typename R> inline R sinpi_r(R x) {
  R n = std::nearbyint(x);
  R r = x - n;
  R s = std::sin(pi_r<R>() * r);
  long long ni = (long long)n;
  return (ni & 1) ? -s : s;
}

// cos(pi*x) = sin(pi*(x + 1/2)) evaluated with the same reduction.
template <typename R> inline R cospi_r(R x) {
  R n = std::nearbyint(x);
  R r = x - n;
  R c = std::cos(pi_r<R>() * r);
  long long ni = (long long)n;
  return (ni & 1) ? -c : c;
}

// RCA.cu:79-82 — fminf(sinpi(x)/(PI*x), 1): the NaN at x=0 becomes 1.
template <typename R> inline R sinc_ref(R x) {
  R v = sinpi_r(x) / (pi_r<R>() * x);
  return std::fmin(v, R(1));
}

// RCA.cu:85-98 — angle at A between AB and AC; double intermediates, the
// final acos in the instantiation's precision.
template <typename R>
inline R angle_aperture(float ax, float ay, float bx, float by, float cx, float cy) {
  double abx = (double)bx - ax, aby = (double)by - ay;
  double acx = (double)cx - ax, acy = (double)cy - ay;
  double v = (abx * acx + aby * acy) / std::sqrt((abx * abx + aby * aby) * (acx * acx + acy * acy));
  return std::acos((R)v);
}

struct Interp {
  int kind;   // 0 quadratic, 1 reference Lanczos, 2 textbook Lanczos
  int a;      // textbook order
  int lo;     // taps k = -lo .. hi around floor(u)
  int hi;
  double win; // window divisor W in sinc(x)*sinc(x/W)
};

inline Interp make_interp(int kind, int a) {
  Interp I;
  I.kind = kind;
  I.a = a;
  if (kind == 0) { I.lo = 1; I.hi = 1; I.win = 0; }
  else if (kind == 1) { I.lo = 2; I.hi = 3; I.win = 2; }
  else { I.lo = a - 1; I.hi = a; I.win = a; }
  return I;
}

// Weights q[k+lo] for k = -lo..hi at fractional position d in [0,1).
template <typename R> inline void weights(const Interp& I, R d, R* q) {
  if (I.kind == 0) {  // RCA.cu:355-357
    q[0] = d * (d - 1) / 2;
    q[1] = -(d - 1) * (d + 1);
    q[2] = d * (d + 1) / 2;
    return;
  }
  const R W = (R)I.win;
  for (int k = -I.lo; k <= I.hi; ++k) {  // RCA.cu:331-336
    R x = d - (R)k;
    q[k + I.lo] = sinc_ref(x) * sinc_ref(x / W);
  }
}

struct Args {
  const float *gx, *gy, *gz, *ex, *ey, *ez, *vx, *vy, *vz, *d0;
  const float* ablaw;     // may be null
  int ablaw_mode;         // 0 none, 1 per channel, 2 per (tx, channel), 3 per (channel, pixel)
  const float* ablaw_tx;  // mode 3 transmit-side law, may be null
  int receive_x;
  float apx, apy, apz, fs, fd, c, t0, fnum;
  int nPix, nCh, nTx, nFrame, nSample;
  int mask_tx;
  Interp I;
};

// Transmit insonification test, RCA.cu:191-219.
template <typename R> inline bool mask_transmit(const Args& A, int p, int t) {
  if (!A.mask_tx) return true;
  float hx = 0.5f * A.apx, hy = 0.5f * A.apy;
  float vx = A.vx[t], vy = A.vy[t], vz = A.vz[t];
  float px = A.gx[p], py = A.gy[p], pz = A.gz[p];
  R coneX = angle_aperture<R>(vx, vz, -hx, -A.apz, hx, -A.apz);
  R coneY = angle_aperture<R>(vy, vz, -hy, -A.apz, hy, -A.apz);
  return angle_aperture<R>(vx, vz, -hx, -A.apz, px, pz) <= coneX &&
         angle_aperture<R>(vx, vz, hx, -A.apz, px, pz) <= coneX &&
         angle_aperture<R>(vy, vz, -hy, -A.apz, py, pz) <= coneY &&
         angle_aperture<R>(vy, vz, hy, -A.apz, py, pz) <= coneY;
}

// Per-(pixel, tilt) geometry shared by the forward and the adjoint.
template <typename R> struct PixTx {
  R px, py, pz, invC, delayTx, apF, phase0;
  int p;
};

template <typename R> inline PixTx<R> pix_tx(const Args& A, int p, int t) {
  PixTx<R> G;
  G.p = p;
  G.px = A.gx[p]; G.py = A.gy[p]; G.pz = A.gz[p];
  G.invC = R(1) / (R)A.c;                      // RCA.cu:164
  G.apF = G.pz / (R)A.fnum;                    // RCA.cu:174
  R vx = A.vx[t], vy = A.vy[t], vz = A.vz[t], d0 = A.d0[t];
  R dx, dy, dz = G.pz - vz;
  if (A.receive_x) { dx = 0; dy = G.py - vy; }  // RCA.cu:256-262
  else { dx = G.px - vx; dy = 0; }              // RCA.cu:263-269
  G.delayTx = (std::sqrt(dx * dx + dy * dy + dz * dz) - d0) * G.invC;
  // 4*PI*fd*pixZ/c term of RCA.cu:371-373
  G.phase0 = R(4) * pi_r<R>() * (R)A.fd * G.pz * G.invC;
  return G;
}

// One (pixel, tilt, channel) term: returns false when the channel is outside
// the f-number aperture (apod undefined) and sets *inrange when the sample
// window is inside the trace.  Outputs apod, m (floor), q[], cos/sin(phi).
template <typename R>
inline bool channel_term(const Args& A, const PixTx<R>& G, int t, int ch, R* apod, long long* m,
                         R* q, R* cphi, R* sphi, bool* inrange) {
  R eX = A.ex[ch], eY = A.ey[ch], eZ = A.ez[ch];
  R aper = A.receive_x ? std::fabs(G.px - eX) : std::fabs(G.py - eY);  // RCA.cu:284-291
  *inrange = false;
  if (!(aper < G.apF)) return false;                                    // RCA.cu:294
  *apod = R(0.5) * (R(1) + cospi_r(aper / G.apF));                      // RCA.cu:299
  R dzr = G.pz - eZ;
  R delayRx = std::sqrt(aper * aper + dzr * dzr) * G.invC;              // RCA.cu:308-309
  if (A.ablaw_mode == 1) delayRx = (R)A.ablaw[ch] + delayRx;            // Hyp1.cu:306-311
  else if (A.ablaw_mode == 2) delayRx = (R)A.ablaw[t * A.nCh + ch] + delayRx;  // Hyp2.cu:306-311
  R abTx = 0;
  if (A.ablaw_mode == 3) {                                              // TxRx.cu:268-275
    const long long i = G.p + (long long)ch * A.nPix;
    if (A.ablaw) delayRx = (R)A.ablaw[i] + delayRx;                     // TxRx.cu:292-293
    if (A.ablaw_tx) abTx = (R)A.ablaw_tx[i];
  }
  R u = (G.delayTx + delayRx + (R)A.t0 + abTx) * (R)A.fs;             // RCA.cu:310-311, TxRx.cu:294
  R mf = std::floor(u);                                                 // RCA.cu:319
  // RCA.cu:323 — bounds kept in floating point so no int overflow can occur
  if (!(mf >= (R)A.I.lo && mf < (R)(A.nSample - 1 - A.I.hi))) return true;
  *inrange = true;
  *m = (long long)mf;
  weights<R>(A.I, u - mf, q);                                           // RCA.cu:325-366
  R phi = -R(2) * pi_r<R>() * u * ((R)A.fd / (R)A.fs) + G.phase0;       // RCA.cu:371-373
  *cphi = std::cos(phi);
  *sphi = std::sin(phi);
  return true;
}

// Forward: dataBf[2*(p + t*nPix + f*nPix*nTx) + {0,1}] (per tilt, as
// RCA.cu:230) and mask[p + t*nPix] (RCA.cu:229).  nosum != 0 writes one
// output per channel at +ch*2*nFrame*nPix*nTx (NoSum.cu:228-229).
template <typename R, typename D>
void forward(const Args& A, R* dataBf, R* mask, const D* data, int nosum, int nthreads) {
  const long long nPix = A.nPix, nTx = A.nTx, nFrame = A.nFrame;
  const long long total = nPix * nTx * nFrame;
  const long long nPixBf = 2 * nFrame * nPix * nTx;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
  for (long long it = 0; it < total; ++it) {
    const int p = (int)(it % nPix);
    const int t = (int)((it / nPix) % nTx);
    const int f = (int)(it / (nPix * nTx));
    if (!mask_transmit<R>(A, p, t)) continue;  // RCA.cu:224 — outputs untouched
    const long long idxPixTx = p + (long long)t * nPix;
    const long long idxPixBf = 2 * (p + (long long)t * nPix + (long long)f * nPix * nTx);
    if (f == 0) mask[idxPixTx] = 0;
    if (!nosum) { dataBf[idxPixBf] = 0; dataBf[idxPixBf + 1] = 0; }
    PixTx<R> G = pix_tx<R>(A, p, t);
    R accR = 0, accI = 0;
    R q[16];
    for (int ch = 0; ch < A.nCh; ++ch) {
      const long long idxCh = idxPixBf + (long long)ch * nPixBf;
      if (nosum) { dataBf[idxCh] = 0; dataBf[idxCh + 1] = 0; }  // NoSum.cu:275-278
      R apod, cphi, sphi;
      long long m;
      bool inr;
      if (!channel_term<R>(A, G, t, ch, &apod, &m, q, &cphi, &sphi, &inr)) continue;
      if (f == 0) mask[idxPixTx] += apod;                       // RCA.cu:301-304
      if (!inr) continue;
      const long long idxData = 2 * ((long long)ch * A.nSample + (long long)t * A.nSample * A.nCh +
                                     (long long)f * A.nSample * A.nCh * nTx);  // RCA.cu:315-317
      R re = 0, im = 0;
      for (int k = -A.I.lo; k <= A.I.hi; ++k) {
        const long long j = 2 * (m + k) + idxData;
        re += q[k + A.I.lo] * (R)data[j];
        im += q[k + A.I.lo] * (R)data[j + 1];
      }
      R oR = apod * (re * cphi + im * sphi);   // RCA.cu:375-380
      R oI = apod * (-re * sphi + im * cphi);
      if (nosum) { dataBf[idxCh] = oR; dataBf[idxCh + 1] = oI; }
      else { accR += oR; accI += oI; }
    }
    if (!nosum) { dataBf[idxPixBf] += accR; dataBf[idxPixBf + 1] += accI; }
  }
}

// Adjoint of the per-tilt forward map (linear in data): grad_data += A^H g.
// g is [nFrame][gTx][nPix] interleaved with gTx in {1, nTx} (1 = the
// compounded output, broadcast over tilts).  One thread owns one trace, so
// the result is deterministic.
template <typename R>
void adjoint(const Args& A, R* grad, const R* g, int gTx, int nthreads) {
  const long long nTr = (long long)A.nFrame * A.nTx * A.nCh;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (long long tr = 0; tr < nTr; ++tr) {
    const int ch = (int)(tr % A.nCh);
    const int t = (int)((tr / A.nCh) % A.nTx);
    const int f = (int)(tr / ((long long)A.nCh * A.nTx));
    R* gd = grad + 2 * tr * (long long)A.nSample;
    R q[16];
    for (int p = 0; p < A.nPix; ++p) {
      if (!mask_transmit<R>(A, p, t)) continue;
      PixTx<R> G = pix_tx<R>(A, p, t);
      R apod, cphi, sphi;
      long long m;
      bool inr;
      if (!channel_term<R>(A, G, t, ch, &apod, &m, q, &cphi, &sphi, &inr) || !inr) continue;
      const long long gi = 2 * (p + (long long)(gTx == 1 ? 0 : t) * A.nPix +
                                (long long)f * A.nPix * gTx);
      const R gr = g[gi], gim = g[gi + 1];
      // conj(apod * e^{-j phi}) * g
      const R br = apod * (gr * cphi - gim * sphi);
      const R bi = apod * (gr * sphi + gim * cphi);
      for (int k = -A.I.lo; k <= A.I.hi; ++k) {
        gd[2 * (m + k)] += q[k + A.I.lo] * br;
        gd[2 * (m + k) + 1] += q[k + A.I.lo] * bi;
      }
    }
  }
}

Args make_args(const float* gx, const float* gy, const float* gz, const float* ex, const float* ey,
               const float* ez, const float* vx, const float* vy, const float* vz, const float* d0,
               int receive_x, float apx, float apy, float apz, float fs, float fd, float c, float t0,
               float fnum, int nPix, int nCh, int nTx, int nFrame, int nSample, int interp,
               int lanczos_a, int mask_tx, const float* ablaw, int ablaw_mode,
               const float* ablaw_tx) {
  Args A;
  A.gx = gx; A.gy = gy; A.gz = gz; A.ex = ex; A.ey = ey; A.ez = ez;
  A.vx = vx; A.vy = vy; A.vz = vz; A.d0 = d0;
  A.ablaw = ablaw; A.ablaw_tx = ablaw_mode == 3 ? ablaw_tx : nullptr;
  A.ablaw_mode = (ablaw || A.ablaw_tx) ? ablaw_mode : 0;
  A.receive_x = receive_x;
  A.apx = apx; A.apy = apy; A.apz = apz;
  A.fs = fs; A.fd = fd; A.c = c; A.t0 = t0; A.fnum = fnum;
  A.nPix = nPix; A.nCh = nCh; A.nTx = nTx; A.nFrame = nFrame; A.nSample = nSample;
  A.mask_tx = mask_tx;
  A.I = make_interp(interp, lanczos_a);
  return A;
}

int check(int interp, int lanczos_a, int nPix, int nCh, int nTx, int nFrame, int nSample) {
  if (interp < 0 || interp > 2) return -1;
  if (interp == 2 && (lanczos_a < 1 || lanczos_a > 8)) return -1;
  if (nPix < 0 || nCh < 0 || nTx < 0 || nFrame < 0 || nSample < 0) return -1;
  return 0;
}

}  // namespace

extern "C" {

#define ORACLE_ARGS                                                                              \
  const float *gx, const float *gy, const float *gz, const float *ex, const float *ey,           \
      const float *ez, const float *vx, const float *vy, const float *vz, const float *d0,       \
      int receive_x, float apx, float apy, float apz, float fs, float fd, float c, float t0,     \
      float fnum, int nPix, int nCh, int nTx, int nFrame, int nSample, int interp, int lanczos_a, \
      int mask_tx, const float *ablaw, int ablaw_mode, const float *ablaw_tx
#define ORACLE_PASS                                                                            \
  gx, gy, gz, ex, ey, ez, vx, vy, vz, d0, receive_x, apx, apy, apz, fs, fd, c, t0, fnum, nPix, \
      nCh, nTx, nFrame, nSample, interp, lanczos_a, mask_tx, ablaw, ablaw_mode, ablaw_tx

// FP64 oracle of the reference kernel (data given in FP64; FP32 inputs are
// promoted exactly by the caller).
int oracle_rca_forward_f64(double* dataBf, double* mask, const double* data, ORACLE_ARGS, int nosum,
                           int nthreads) {
  if (check(interp, lanczos_a, nPix, nCh, nTx, nFrame, nSample)) return -1;
  Args A = make_args(ORACLE_PASS);
  forward<double, double>(A, dataBf, mask, data, nosum, nthreads);
  return 0;
}

// FP32 restatement ("reference CPU path"): same loop nest and formulas in float.
int oracle_rca_forward_f32(float* dataBf, float* mask, const float* data, ORACLE_ARGS, int nosum,
                           int nthreads) {
  if (check(interp, lanczos_a, nPix, nCh, nTx, nFrame, nSample)) return -1;
  Args A = make_args(ORACLE_PASS);
  forward<float, float>(A, dataBf, mask, data, nosum, nthreads);
  return 0;
}

// FP64 adjoint: grad (zeroed by the caller) += A^H g.
int oracle_rca_adjoint_f64(double* grad, const double* g, int gTx, ORACLE_ARGS, int nthreads) {
  if (check(interp, lanczos_a, nPix, nCh, nTx, nFrame, nSample)) return -1;
  if (gTx != 1 && gTx != nTx) return -1;
  Args A = make_args(ORACLE_PASS);
  adjoint<double>(A, grad, g, gTx, nthreads);
  return 0;
}

// Host prep of RCA.m:491-558, in double, cast to single at the end
// (RCA.m:601-617).  ap_in[i] is NaN when the aperture was not provided.
// theta != 0: tilts (thetaX, thetaY) are converted to virtual sources at
// r = 10 m; otherwise vs_in holds the virtual sources.
int oracle_rca_prepare(const double* posX, const double* posY, const double* posZ, int nCh,
                       const double* ap_in, int use_theta, const double* thetaX,
                       const double* thetaY, const double* vsX_in, const double* vsY_in,
                       const double* vsZ_in, int nTx, float* ex, float* ey, float* ez, float* ap_out,
                       float* vx, float* vy, float* vz, float* d0) {
  if (nCh < 1 || nTx < 0) return -1;
  double mx = 0, my = 0, mz = 0;
  for (int i = 0; i < nCh; ++i) { mx += posX[i]; my += posY[i]; mz += posZ[i]; }
  mx /= nCh; my /= nCh; mz /= nCh;
  double *cx = new double[nCh], *cy = new double[nCh], *cz = new double[nCh];
  for (int i = 0; i < nCh; ++i) { cx[i] = posX[i] - mx; cy[i] = posY[i] - my; cz[i] = posZ[i] - mz; }
  const double* cc[3] = {cx, cy, cz};
  for (int a = 0; a < 3; ++a) {  // RCA.m:505-521
    if (std::isnan(ap_in[a])) {
      double lo = cc[a][0], hi = cc[a][0];
      for (int i = 1; i < nCh; ++i) { lo = std::fmin(lo, cc[a][i]); hi = std::fmax(hi, cc[a][i]); }
      ap_out[a] = (float)(hi - lo);
    } else {
      ap_out[a] = (float)ap_in[a];
    }
  }
  for (int s = 0; s < nTx; ++s) {
    double X, Y, Z;
    if (use_theta) {  // RCA.m:529-541
      const double r = 10;
      double tx = std::tan(thetaX[s]), ty = std::tan(thetaY[s]);
      Z = -r / std::sqrt(1 + tx * tx + ty * ty);
      X = Z * tx;
      Y = Z * ty;
    } else {
      X = vsX_in[s]; Y = vsY_in[s]; Z = vsZ_in[s];
    }
    // RCA.m:551-558 — note the X/Y swap between elements and sources
    double best = INFINITY;
    for (int i = 0; i < nCh; ++i) {
      double a = cy[i] - X, b = cx[i] - Y, c3 = cz[i] - Z;
      best = std::fmin(best, std::sqrt(a * a + b * b + c3 * c3));
    }
    vx[s] = (float)X; vy[s] = (float)Y; vz[s] = (float)Z; d0[s] = (float)best;
  }
  for (int i = 0; i < nCh; ++i) { ex[i] = (float)cx[i]; ey[i] = (float)cy[i]; ez[i] = (float)cz[i]; }
  delete[] cx; delete[] cy; delete[] cz;
  return 0;
}

// Interpolation weights q[0 .. lo+hi] at fractional position d (for pinning
// the weights against closed forms); returns the tap count.
int oracle_interp_weights(int interp, int lanczos_a, double d, double* q) {
  if (check(interp, lanczos_a, 0, 0, 0, 0, 0)) return -1;
  Interp I = make_interp(interp, lanczos_a);
  weights<double>(I, d, q);
  return I.lo + I.hi + 1;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
