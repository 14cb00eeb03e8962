// Resident plans: the wrapper's host prep in FP64 and one upload of the
// geometry (rca_plan_create / rca_plan_destroy of include/rca_das.h).
//
// The prep restates bfVirtualSourceGeneric_RCA.m:491-558: element positions
// centred on their mean (:500-502), apertures defaulting to the span of the
// centred positions (:505-521), tilt angles turned into virtual sources at
// r = 10 m (:527-541), and minDist_t = min_c |(ey_c - vx_t, ex_c - vy_t,
// ez_c - vz_t)| with the reference's X/Y swap (:551-558); every value is
// then rounded to FP32 as the wrapper's single() does (:598-636).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <vector>

#include "../../include/rca_das.h"

namespace rca_internal {
void set_error(const char* msg);
void clear_error();
}  // namespace rca_internal

struct rca_plan {
  rca_das_desc desc;
  void* mem;   // one device allocation holding every array
};

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

int check_params(const rca_plan_params* p) {
  if (!p) return fail(RCA_ERR_INVALID, "plan params are NULL");
  if (p->n_pix < 0 || p->n_channel < 1 || p->n_tx < 1 || p->n_sample < 1)
    return fail(RCA_ERR_INVALID, "bad sizes: nPix=%d nChannel=%d nTx=%d nSample=%d", p->n_pix,
                p->n_channel, p->n_tx, p->n_sample);
  if (!p->ele_x || !p->ele_y || !p->ele_z) return fail(RCA_ERR_INVALID, "element positions are NULL");
  const bool tilts = p->theta_x && p->theta_y;
  const bool sources = p->vs_x && p->vs_y && p->vs_z;
  if (tilts == sources)   // RCA.m:316-367: angles XOR virtual sources
    return fail(RCA_ERR_INVALID, "give either theta_x/theta_y or vs_x/vs_y/vs_z");
  return RCA_OK;
}

double mean(const double* a, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i];
  return s / n;
}

}  // namespace

extern "C" {

int rca_prepare_geometry(const rca_plan_params* p, float* ele, float* vs, float* min_dist,
                         float* aperture) {
  rca_internal::clear_error();
  int rc = check_params(p);
  if (rc) return rc;
  if (!ele || !vs || !min_dist || !aperture) return fail(RCA_ERR_INVALID, "output arrays are NULL");
  const int nc = p->n_channel, nt = p->n_tx;
  const double* pos[3] = {p->ele_x, p->ele_y, p->ele_z};
  std::vector<double> cen(3 * (size_t)nc);
  for (int d = 0; d < 3; ++d) {
    const double m = mean(pos[d], nc);                       // RCA.m:500-502
    double lo = INFINITY, hi = -INFINITY;
    for (int c = 0; c < nc; ++c) {
      const double v = pos[d][c] - m;
      cen[d * nc + c] = v;
      ele[d * nc + c] = (float)v;
      lo = std::fmin(lo, v);
      hi = std::fmax(hi, v);
    }
    const double given[3] = {p->aperture_x, p->aperture_y, p->aperture_z};
    aperture[d] = (float)(p->has_aperture ? given[d] : hi - lo);   // RCA.m:505-521
  }
  std::vector<double> v(3 * (size_t)nt);
  for (int t = 0; t < nt; ++t) {
    if (p->theta_x) {                                        // RCA.m:527-541
      const double tx = std::tan(p->theta_x[t]), ty = std::tan(p->theta_y[t]);
      const double vz = -10.0 / std::sqrt(1.0 + tx * tx + ty * ty);
      v[t] = vz * tx;
      v[nt + t] = vz * ty;
      v[2 * nt + t] = vz;
    } else {
      v[t] = p->vs_x[t];
      v[nt + t] = p->vs_y[t];
      v[2 * nt + t] = p->vs_z[t];
    }
    double best = INFINITY;                                  // RCA.m:551-558 (X/Y swapped)
    for (int c = 0; c < nc; ++c) {
      const double dx = cen[nc + c] - v[t], dy = cen[c] - v[nt + t], dz = cen[2 * nc + c] - v[2 * nt + t];
      best = std::fmin(best, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    min_dist[t] = (float)best;
    for (int d = 0; d < 3; ++d) vs[d * nt + t] = (float)v[d * nt + t];
  }
  return RCA_OK;
}

int rca_plan_create(const rca_plan_params* p, rca_plan** plan) {
  rca_internal::clear_error();
  if (!plan) return fail(RCA_ERR_INVALID, "plan pointer is NULL");
  *plan = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  if (p->n_pix > 0 && (!p->grid_x || !p->grid_y || !p->grid_z))
    return fail(RCA_ERR_INVALID, "grid arrays are NULL");
  if (p->ab_law_mode < RCA_ABLAW_NONE || p->ab_law_mode > RCA_ABLAW_PIXEL_CHANNEL)
    return fail(RCA_ERR_INVALID, "unknown aberration-law mode %d", p->ab_law_mode);
  const long long np = p->n_pix, nc = p->n_channel, nt = p->n_tx;
  const long long n_law = p->ab_law_mode == RCA_ABLAW_CHANNEL ? nc
                          : p->ab_law_mode == RCA_ABLAW_TX_CHANNEL ? nt * nc
                          : p->ab_law_mode == RCA_ABLAW_PIXEL_CHANNEL ? nc * np : 0;
  if (p->ab_law_mode && !p->ab_law_rx && !(p->ab_law_mode == RCA_ABLAW_PIXEL_CHANNEL && p->ab_law_tx))
    return fail(RCA_ERR_INVALID, "aberration law arrays are NULL");
  // host staging: grid [3][nPix], elements [3][nCh], sources [3][nTx], minDist [nTx], laws
  const bool rx_law = p->ab_law_mode && p->ab_law_rx;
  const bool tx_law = p->ab_law_mode == RCA_ABLAW_PIXEL_CHANNEL && p->ab_law_tx;
  const long long n_f = 3 * np + 3 * nc + 4 * nt + (rx_law ? n_law : 0) + (tx_law ? n_law : 0);
  std::vector<float> h((size_t)n_f);
  float* hg = h.data();
  float* he = hg + 3 * np;
  float* hv = he + 3 * nc;
  float* hd = hv + 3 * nt;
  float* hr = hd + nt;
  float* ht = hr + (rx_law ? n_law : 0);
  float ap[3];
  rc = rca_prepare_geometry(p, he, hv, hd, ap);
  if (rc) return rc;
  for (long long i = 0; i < np; ++i) {
    hg[i] = (float)p->grid_x[i];
    hg[np + i] = (float)p->grid_y[i];
    hg[2 * np + i] = (float)p->grid_z[i];
  }
  for (long long i = 0; rx_law && i < n_law; ++i) hr[i] = (float)p->ab_law_rx[i];
  for (long long i = 0; tx_law && i < n_law; ++i) ht[i] = (float)p->ab_law_tx[i];
  rca_plan* pl = new rca_plan();
  hipError_t e = hipMalloc(&pl->mem, (size_t)n_f * sizeof(float));
  if (e != hipSuccess) {
    delete pl;
    return fail(RCA_ERR_HIP, "hipMalloc(%lld bytes): %s", n_f * 4, hipGetErrorString(e));
  }
  e = hipMemcpy(pl->mem, h.data(), (size_t)n_f * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(pl->mem);
    delete pl;
    return fail(RCA_ERR_HIP, "hipMemcpy: %s", hipGetErrorString(e));
  }
  const float* d = static_cast<const float*>(pl->mem);
  rca_das_desc& s = pl->desc;
  s = rca_das_desc{};
  s.grid_x = d;
  s.grid_y = d + np;
  s.grid_z = d + 2 * np;
  s.ele_x = d + 3 * np;
  s.ele_y = s.ele_x + nc;
  s.ele_z = s.ele_x + 2 * nc;
  s.vs_x = s.ele_x + 3 * nc;
  s.vs_y = s.vs_x + nt;
  s.vs_z = s.vs_x + 2 * nt;
  s.min_dist = s.vs_x + 3 * nt;
  s.ab_law_rx = rx_law ? s.min_dist + nt : nullptr;
  s.ab_law_tx = tx_law ? s.min_dist + nt + (rx_law ? n_law : 0) : nullptr;
  s.ab_law_mode = p->ab_law_mode;
  s.n_pix = p->n_pix;
  s.n_channel = p->n_channel;
  s.n_tx = p->n_tx;
  s.n_sample = p->n_sample;
  s.receive_x = p->receive_x != 0;
  s.mask_transmit = p->mask_transmit != 0;
  s.interp = p->interp;
  s.aperture_x = ap[0];
  s.aperture_y = ap[1];
  s.aperture_z = ap[2];
  s.sampling_frequency = (float)p->sampling_frequency;
  s.demodulation_frequency = (float)p->demodulation_frequency;
  s.speed_of_sound = (float)p->speed_of_sound;
  s.start_time = (float)p->start_time;
  s.fnumber = (float)p->fnumber;
  s.depth_run = p->depth_run > 0 && p->n_pix % p->depth_run == 0 ? p->depth_run : 0;
  *plan = pl;
  return RCA_OK;
}

const rca_das_desc* rca_plan_desc(const rca_plan* plan) { return plan ? &plan->desc : nullptr; }

int rca_plan_destroy(rca_plan* plan) {
  rca_internal::clear_error();
  if (!plan) return RCA_OK;
  const hipError_t e = hipFree(plan->mem);
  delete plan;
  return e == hipSuccess ? RCA_OK : fail(RCA_ERR_HIP, "hipFree: %s", hipGetErrorString(e));
}

}  // extern "C"
