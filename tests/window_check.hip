// Window-cover check of the DAS kernels' LDS staging, on the host (test
// infrastructure; built and loaded by tests/window_check.py).
//
// The forward and adjoint kernels stage, per (tilt chunk, receive line), only
// the window of each trace that their workgroup's voxels can reach, and read
// an update's taps m-LO .. m+HI from it; a tap outside the staged rows would
// be clamped to the row edge (the LDS index is kept in bounds) and give a
// wrong value without any error.  This walks the same workgroups, tilt chunks,
// lines and passes as the kernels, with the kernels' own tile, bound, window,
// pass and sample-position code (simu_rca_ab_amd/csrc/rca_tile.h), and counts
// every in-range update whose taps are not inside the rows of the pass that
// takes it, every update taken by no pass or by two, and every skipped line
// that a voxel of the tile would have used.  FP64 square roots are the
// correctly rounded ones here and the hardware rsq + Newton pair in the
// kernels (a few ulp apart); the windows keep one sample of margin on each
// side, so such differences cannot move a tap across a window edge unless a
// bound is wrong by about a sample.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../simu_rca_ab_amd/csrc/rca_tile.h"

namespace {

using namespace rca_tile;

constexpr int kAbChannel = 1, kAbTxChannel = 2, kAbPixel = 3;

}  // namespace

extern "C" {

struct WinCase {
  const float *gx, *gy, *gz;         // n_pix voxel coordinates (grid order)
  const float *ex, *ey, *ez;         // n_ch centred element positions
  const float *vx, *vy, *vz, *d0;    // n_tx virtual sources and minimal distances
  const float* ab;                   // receive law: [n_ch], [n_tx][n_ch] or [n_ch][n_pix]
  const float* abt;                  // per-(line, voxel) transmit-side law or NULL
  int ab_mode;                       // 0 none, 1 per line, 2 per (tilt, line), 3 per (line, voxel)
  int n_pix, n_ch, n_tx, n_sample, depth_run;
  int receive_x;
  float fs, fd, c, t0, fnumber;
  int lo, hi;                        // taps m-lo .. m+hi
  int kernel;                        // 0 forward, 1 adjoint
  int sbytes;                        // forward: bytes per sample (8 complex64, 4 half2)
  int fb, v;                         // forward: frame block, voxels per lane
  int compound;                      // forward: tilt chunks (compounding / NoSum) or one tilt
  int lds_budget;                    // bytes; 0 = the kernel's default
  int shrink;                        // narrows every window by this many samples per side
  const long long* blocks;           // workgroups to check, or NULL for all
  long long n_blocks;
  int n_threads;
};

struct WinStats {
  long long blocks, lines, lines_skipped, updates_in_range, violations, skipped_in_aperture,
      multi_pass_lines, max_passes;
  int tc, wcap, step, e;
  long long first_bx;
  int first_c, first_t, first_base, first_cb, first_end, first_kind;
};

int rca_window_check(const WinCase* wc, WinStats* st) {
  const WinCase& K = *wc;
  memset(st, 0, sizeof(*st));
  st->first_bx = -1;
  const int depth_run = K.depth_run >= 64 ? K.depth_run : 0;
  const bool fwd = K.kernel == 0;
  const int V = K.v;
  int tc, wcap, E;
  if (fwd) {
    const size_t budget = K.lds_budget ? (size_t)K.lds_budget : (K.fb == 1 ? kFwdLds1 : kFwdLds);
    const FwdLayout L = fwd_layout(depth_run, K.n_sample, V, K.fb, (size_t)K.sbytes,
                                   K.compound ? tc_max(K.fb, K.sbytes) : 1, budget);
    tc = L.tc;
    wcap = L.wcap;
    E = 16 / K.sbytes;
  } else {
    const size_t budget = K.lds_budget ? (size_t)K.lds_budget : kAdjLds;
    const AdjLayout L = adj_layout(depth_run, K.n_sample, K.fb, tc_max(K.fb, 8), budget, V);
    tc = L.tc;
    wcap = L.wcap;
    E = 1;
  }
  const int LO = K.lo, HI = K.hi, NT = LO + HI + 1;
  const int step = pass_step(wcap, NT, E);
  st->tc = tc;
  st->wcap = wcap;
  st->step = step;
  st->e = E;
  const long long nblk = fwd ? fwd_blocks(K.n_pix, depth_run, V) : tile_blocks<1, 1>(K.n_pix, depth_run, V);
  const long long nchk = K.blocks ? K.n_blocks : nblk;
  const bool rx = K.receive_x != 0;
  // the kernels' FP64 constants (make_args)
  const double fs = K.fs, fd = K.fd, cc = K.c;
  const double kfs = fs / cc, ut0 = fs * (double)K.t0;
  (void)fd;
  const int ns = K.n_sample;

  std::mutex mu;
  std::atomic<long long> next{0};
  auto work = [&]() {
    WinStats loc;
    memset(&loc, 0, sizeof(loc));
    loc.first_bx = -1;
    const int nvox = kThreads * V;
    std::vector<long long> pidx(nvox);
    std::vector<char> valid(nvox);
    std::vector<double> pl(nvox), pt(nvox), pz(nvox), apF(nvox), lawv(nvox);
    std::vector<int> bi((size_t)tc * nvox);
    std::vector<float> bf((size_t)tc * nvox);
    for (;;) {
      const long long ib = next.fetch_add(1);
      if (ib >= nchk) break;
      const long long bxl = K.blocks ? K.blocks[ib] : ib;
      if (bxl < 0 || bxl >= nblk) continue;
      const unsigned bx = (unsigned)bxl;
      ++loc.blocks;
      // voxels of the workgroup and their bounding box (tile_voxels)
      float bb[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
      bool any = false;
      for (int v = 0; v < V; ++v)
        for (int tid = 0; tid < kThreads; ++tid) {
          const int i = v * kThreads + tid;
          bool ok;
          const long long p = fwd ? voxel_of<kTileCols, 4>(depth_run, K.n_pix, bx, V, v, tid, ok)
                                  : voxel_of<1, 1>(depth_run, K.n_pix, bx, V, v, tid, ok);
          pidx[i] = p;
          valid[i] = ok;
          if (!ok) continue;
          any = true;
          const float fx = K.gx[p], fy = K.gy[p], fz = K.gz[p];
          pl[i] = rx ? fx : fy;
          pt[i] = rx ? fy : fx;
          pz[i] = fz;
          apF[i] = pz[i] / (double)K.fnumber;
          bb[0] = fminf(bb[0], fx); bb[1] = fmaxf(bb[1], fx);
          bb[2] = fminf(bb[2], fy); bb[3] = fmaxf(bb[3], fy);
          bb[4] = fminf(bb[4], fz); bb[5] = fmaxf(bb[5], fz);
        }
      if (!any) continue;
      TileBox box;
      box.lat_lo = rx ? bb[0] : bb[2];
      box.lat_hi = rx ? bb[1] : bb[3];
      box.txl_lo = rx ? bb[2] : bb[0];
      box.txl_hi = rx ? bb[3] : bb[1];
      box.zlo = bb[4];
      box.zhi = bb[5];
      box.apmax = K.fnumber > 0.0f ? (double)box.zhi / (double)K.fnumber : (double)box.zlo / (double)K.fnumber;

      for (int t0 = 0; t0 < K.n_tx; t0 += tc) {
        const int nt = std::min(tc, K.n_tx - t0);
        // transmit terms (chunk_tx) and bounds (chunk_tx_bounds)
        float bu_lo = INFINITY, bu_hi = -INFINITY;
        for (int j = 0; j < nt; ++j) {
          const int t = t0 + j;
          float b0, b1;
          tx_bounds(kfs, ut0, box.txl_lo, box.txl_hi, box.zlo, box.zhi, rx ? K.vy[t] : K.vx[t], K.vz[t],
                    (double)K.d0[t], b0, b1);
          bu_lo = fminf(bu_lo, b0);
          bu_hi = fmaxf(bu_hi, b1);
          for (int i = 0; i < nvox; ++i) {
            if (!valid[i]) continue;
            const double dl = pt[i] - (double)(rx ? K.vy[t] : K.vx[t]), dz = pz[i] - (double)K.vz[t];
            const double tt = sqrt(fma(dl, dl, dz * dz)) - (double)K.d0[t];
            split_pos(fma(kfs, tt, ut0), bi[(size_t)j * nvox + i], bf[(size_t)j * nvox + i]);
          }
        }
        for (int c = 0; c < K.n_ch; ++c) {
          ++loc.lines;
          const double e_lat = rx ? (double)K.ex[c] : (double)K.ey[c], ezd = K.ez[c];
          const double uabc = K.ab_mode == kAbChannel ? fs * (double)K.ab[c] : 0.0;
          float px_lo = 0.0f, px_hi = 0.0f;
          if (K.ab_mode == kAbPixel) {
            px_lo = INFINITY;
            px_hi = -INFINITY;
            for (int i = 0; i < nvox; ++i) {
              lawv[i] = 0.0;
              if (!valid[i]) continue;
              const long long q = (long long)c * K.n_pix + pidx[i];
              if (K.ab) lawv[i] = (double)K.ab[q];
              if (K.abt) lawv[i] += (double)K.abt[q];
              const float x = (float)(fs * lawv[i]);
              px_lo = fminf(px_lo, x);
              px_hi = fmaxf(px_hi, x);
            }
          }
          float tc_lo = 0.0f, tc_hi = 0.0f;
          std::vector<int> ui(nt, 0);
          std::vector<float> uf(nt, 0.0f);
          if (K.ab_mode == kAbTxChannel) {
            tc_lo = INFINITY;
            tc_hi = -INFINITY;
            for (int j = 0; j < nt; ++j) {
              const double ab = K.ab[(long long)(t0 + j) * K.n_ch + c];
              const float x = (float)(fs * ab);
              tc_lo = fminf(tc_lo, x);
              tc_hi = fmaxf(tc_hi, x);
              split_pos(fs * ab, ui[j], uf[j]);
            }
          }
          // line_window
          float r_lo = 0.0f, r_hi = 0.0f;
          bool use = rx_bounds(kfs, box, e_lat, ezd, uabc, r_lo, r_hi);
          if (use && !(px_lo <= px_hi)) use = false;
          if (!use) {
            ++loc.lines_skipped;
            for (int i = 0; i < nvox; ++i)
              if (valid[i] && fabs(pl[i] - e_lat) < apF[i]) {
                ++loc.skipped_in_aperture;
                if (loc.first_bx < 0) {
                  loc.first_bx = bx; loc.first_c = c; loc.first_t = t0; loc.first_kind = 3;
                }
                break;
              }
            continue;
          }
          r_lo += px_lo;   // 0 unless the per-(line, voxel) law
          r_hi += px_hi;
          if (K.ab_mode == kAbTxChannel) {
            r_lo += tc_lo;
            r_hi += tc_hi;
          }
          int s0, whi;
          window_from(ns, LO, HI, E, bu_lo + r_lo, bu_hi + r_hi, s0, whi);
          s0 += K.shrink;
          whi -= K.shrink;
          const int npass = pass_count(s0, whi, wcap, step);
          if (npass > 1) ++loc.multi_pass_lines;
          loc.max_passes = std::max<long long>(loc.max_passes, npass);
          for (int i = 0; i < nvox; ++i) {
            if (!valid[i]) continue;
            const double A = fabs(pl[i] - e_lat);
            if (!(A < apF[i])) continue;
            const double dz = pz[i] - ezd;
            const double Rr = sqrt(fmax(fma(A, A, dz * dz), 1e-300));
            const double uab = K.ab_mode == kAbPixel ? fs * lawv[i] : uabc;
            int ri;
            float rf;
            split_pos(fma(kfs, Rr, uab), ri, rf);
            for (int j = 0; j < nt; ++j) {
              int m;
              float d;
              const size_t q = (size_t)j * nvox + i;
              if (K.ab_mode == kAbTxChannel) sample_pos(bi[q] + ri + ui[j], bf[q] + rf + uf[j], m, d);
              else sample_pos(bi[q] + ri, bf[q] + rf, m, d);
              if (!(m >= LO && m < ns - 1 - HI)) continue;
              ++loc.updates_in_range;
              const int base = m - LO;
              int taken = 0, kk = -1;
              for (int k = 0; k < npass; ++k) {
                const int cb = s0 + k * step;
                const bool last = cb + wcap >= whi;
                if (in_pass(base, cb, last, step)) { ++taken; kk = k; }
              }
              int kind = 0, cb = 0, end = 0;
              if (taken != 1) {
                kind = 2;
              } else {
                cb = s0 + kk * step;
                end = std::min(cb + wcap, whi);
                if (base < cb || base + NT > end) kind = 1;
              }
              if (kind) {
                ++loc.violations;
                if (loc.first_bx < 0) {
                  loc.first_bx = bx; loc.first_c = c; loc.first_t = t0 + j; loc.first_base = base;
                  loc.first_cb = cb; loc.first_end = end; loc.first_kind = kind;
                }
              }
            }
          }
        }
      }
    }
    std::lock_guard<std::mutex> g(mu);
    st->blocks += loc.blocks;
    st->lines += loc.lines;
    st->lines_skipped += loc.lines_skipped;
    st->updates_in_range += loc.updates_in_range;
    st->violations += loc.violations;
    st->skipped_in_aperture += loc.skipped_in_aperture;
    st->multi_pass_lines += loc.multi_pass_lines;
    st->max_passes = std::max(st->max_passes, loc.max_passes);
    if (st->first_bx < 0 && loc.first_bx >= 0) {
      st->first_bx = loc.first_bx; st->first_c = loc.first_c; st->first_t = loc.first_t;
      st->first_base = loc.first_base; st->first_cb = loc.first_cb; st->first_end = loc.first_end;
      st->first_kind = loc.first_kind;
    }
  };
  const int nth = std::max(1, K.n_threads);
  std::vector<std::thread> th;
  for (int i = 0; i < nth; ++i) th.emplace_back(work);
  for (auto& t : th) t.join();
  return 0;
}

}  // extern "C"
