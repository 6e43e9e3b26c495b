// C ABI of the gfx950 3D RK4 FDTD stepper (declared in include/wap_fdtd.h).
//
// The handle owns every device buffer of one grid.  Per-plane coefficient
// tables (PML weights of fdtd_3d.cu:96-123 and :215-234, the KG sponge) and
// the soft-source phasors (fdtd_3d.cu:139-156) are evaluated on the host in
// float64 once per handle / once per step, so no transcendental runs per
// voxel on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/wap_fdtd.h"
#include "fdtd3d_kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define WAP_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? WAP_ERR_OOM : WAP_ERR_HIP,               \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                  \
  } while (0)

}  // namespace

struct wap_handle {
  wap_desc d{};
  size_t n = 0;       // voxels
  size_t plane = 0;   // nx*ny
  size_t cs = 0;      // bytes per complex element
  size_t rs = 0;      // bytes per real element
  hipStream_t stream = nullptr;
  void *psi = nullptr, *pi = nullptr;
  void *acc_psi = nullptr, *acc_pi = nullptr;
  void *ya_psi = nullptr, *ya_pi = nullptr;
  void *yb_psi = nullptr, *yb_pi = nullptr;
  void *cmp_psi = nullptr, *cmp_pi = nullptr;
  void *V = nullptr;
  void *zc = nullptr;
  void *gam = nullptr;
  void *prof = nullptr;
  void *src_tab = nullptr;  // wap::SrcCoef<R>[kMaxBand]
  double *partial = nullptr;
  double *dsum = nullptr;
  void *staging = nullptr;
  size_t staging_bytes = 0;
  int64_t bytes = 0;
  bool v_set = false;
  int tile_y = 4, tile_z = 8;
  int ntx = 1, nty = 1, ntz = 1;
  double hx = 0, hy = 0, m2 = 0;
  int gz_lo = 1, gz_hi = 0;  // the longest run of planes with gamma = 0 (empty if lo > hi)
  // Schrodinger: the longest run of planes whose two z weights are one real number hz (in
  // the stepper's precision), i.e. outside the PML (empty if lo > hi)
  int pz_lo = 1, pz_hi = 0;
  double pz_w = 0;
  int nsrc = 0;
  int src_k0 = 0;
  double src_keep[wap::kMaxBand] = {};
  std::complex<double> src_spatial[wap::kMaxBand];
  int n_partial = 0;
};

namespace {

bool is_kg(const wap_handle *h) { return h->d.equation == WAP_EQ_KLEIN_GORDON; }
bool is_f64(const wap_handle *h) { return h->d.precision == WAP_PREC_F64; }
bool is_kahan(const wap_handle *h) { return h->d.precision == WAP_PREC_F32_KAHAN; }

int dev_alloc(wap_handle *h, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? WAP_ERR_OOM : WAP_ERR_HIP,
                std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
  }
  h->bytes += (int64_t)bytes;
  return WAP_OK;
}

void free_all(wap_handle *h) {
  void **bufs[] = {&h->psi, &h->pi, &h->acc_psi, &h->acc_pi, &h->ya_psi, &h->ya_pi,
                   &h->yb_psi, &h->yb_pi, &h->cmp_psi, &h->cmp_pi, &h->V, &h->zc,
                   &h->gam, &h->prof, &h->src_tab, (void **)&h->partial, (void **)&h->dsum,
                   &h->staging};
  for (void **b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
}

int ensure_staging(wap_handle *h, size_t bytes) {
  if (h->staging_bytes >= bytes) return WAP_OK;
  if (h->staging) {
    WAP_HIP(hipStreamSynchronize(h->stream));
    (void)hipFree(h->staging);
    h->bytes -= (int64_t)h->staging_bytes;
    h->staging = nullptr;
    h->staging_bytes = 0;
  }
  int rc = dev_alloc(h, &h->staging, bytes);
  if (rc) return rc;
  h->staging_bytes = bytes;
  return WAP_OK;
}

int grid_1d(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

// Per-plane coefficients, float64 on the host.
template <typename R>
int upload_src_tab(wap_handle *h, std::complex<double> mi_w) {
  wap::SrcCoef<R> tab[wap::kMaxBand];
  for (int b = 0; b < wap::kMaxBand; ++b) {
    const bool on = b < h->nsrc;
    const std::complex<double> c = on ? h->src_spatial[b] : 0.0;
    const std::complex<double> cq = mi_w * c;
    tab[b].keep = on ? (R)h->src_keep[b] : (R)1;
    tab[b].pad = (R)0;
    tab[b].c = wap::cplx<R>{(R)c.real(), (R)c.imag()};
    tab[b].cq = wap::cplx<R>{(R)cq.real(), (R)cq.imag()};
  }
  WAP_HIP(hipMemcpy(h->src_tab, tab, sizeof(tab), hipMemcpyHostToDevice));
  return WAP_OK;
}

int build_tables(wap_handle *h) {
  const wap_desc &d = h->d;
  const double dx = (d.xmax - d.xmin) / (d.nx - 1.0);
  const double dy = d.ny > 1 ? (d.ymax - d.ymin) / (d.ny - 1.0) : 0.0;
  const double dz = (d.zmax - d.zmin) / (d.nz - 1.0);
  const int nz = d.nz;
  std::vector<double> ramp(nz), z(nz);
  for (int k = 0; k < nz; ++k) {
    z[k] = d.zmin + k * (d.zmax - d.zmin) / (nz - 1.0);
    const double rb = std::fmax((d.zmin + d.pml_thick) - z[k], 0.0) / d.pml_thick;
    const double rt = std::fmax(z[k] - (d.zmax - d.pml_thick), 0.0) / d.pml_thick;
    ramp[k] = std::fmax(rb, rt);
  }
  std::vector<std::complex<double>> a(nz, std::complex<double>(1.0, 0.0));
  if (!is_kg(h)) {
    for (int k = 0; k < nz; ++k) {
      if (ramp[k] > 0.0) {
        const double sig = d.pml_sigma0 * std::pow(ramp[k], d.pml_power);
        if (sig > 0.0) a[k] = 1.0 / (1.0 + sig * std::polar(1.0, d.pml_theta));
      }
    }
  }
  const double scale = is_kg(h) ? d.hbar_c * d.hbar_c : -d.lap_coeff;
  h->hx = scale / (dx * dx);
  h->hy = d.ny > 1 ? scale / (dy * dy) : 0.0;  // ny = 1: a 2D grid, no y term
  h->m2 = d.rest_energy * d.rest_energy;
  std::vector<std::complex<double>> zc(2 * (size_t)nz, std::complex<double>(0, 0));
  for (int k = 1; k + 1 < nz; ++k) {
    zc[2 * k] = scale * a[k] * 0.5 * (a[k] + a[k + 1]) / (dz * dz);
    zc[2 * k + 1] = scale * a[k] * 0.5 * (a[k - 1] + a[k]) / (dz * dz);
  }
  {
    // rounded as the stepper sees them: equal real weights compare equal after rounding
    auto rd = [&](double x) { return is_f64(h) ? x : (double)(float)x; };
    h->pz_lo = 1;
    h->pz_hi = 0;
    for (int k = 1, run = 0; k + 1 < nz; ++k) {
      const bool real_eq = zc[2 * k].imag() == 0.0 && zc[2 * k + 1].imag() == 0.0 &&
                           rd(zc[2 * k].real()) == rd(zc[2 * k + 1].real()) &&
                           (run == 0 || rd(zc[2 * k].real()) == rd(zc[2 * (k - 1)].real()));
      run = real_eq ? run + 1 : 0;
      if (run > h->pz_hi - h->pz_lo + 1) {
        h->pz_lo = k - run + 1;
        h->pz_hi = k;
        h->pz_w = zc[2 * k].real();
      }
    }
  }
  std::vector<double> gam(nz, 0.0);
  if (is_kg(h))
    for (int k = 0; k < nz; ++k)
      if (ramp[k] > 0.0) gam[k] = d.kg_gamma_max * std::pow(ramp[k], d.pml_power);
  h->gz_lo = 1;
  h->gz_hi = 0;
  for (int k = 0, run = 0; k < nz; ++k) {
    run = gam[k] == 0.0 ? run + 1 : 0;
    if (run > h->gz_hi - h->gz_lo + 1) {
      h->gz_lo = k - run + 1;
      h->gz_hi = k;
    }
  }

  // upload in the stepper's precision
  if (is_f64(h)) {
    WAP_HIP(hipMemcpy(h->zc, zc.data(), zc.size() * 16, hipMemcpyHostToDevice));
    WAP_HIP(hipMemcpy(h->gam, gam.data(), gam.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> zf(4 * (size_t)nz), gf(nz);
    for (size_t t = 0; t < zc.size(); ++t) {
      zf[2 * t] = (float)zc[t].real();
      zf[2 * t + 1] = (float)zc[t].imag();
    }
    for (int k = 0; k < nz; ++k) gf[k] = (float)gam[k];
    WAP_HIP(hipMemcpy(h->zc, zf.data(), zf.size() * 4, hipMemcpyHostToDevice));
    WAP_HIP(hipMemcpy(h->gam, gf.data(), gf.size() * 4, hipMemcpyHostToDevice));
  }

  // soft source: window weights and spatial phasors (fdtd_3d.cu:134-156)
  h->nsrc = 0;
  h->src_k0 = d.src_k;
  if (d.src_enabled) {
    const double zK = d.zmin + d.src_k * (d.zmax - d.zmin) / (nz - 1.0);
    for (int b = 0; b < d.src_band; ++b) {
      const int k = d.src_k + b;
      const double s = d.src_band > 1 ? (double)b / (d.src_band - 1.0) : 0.0;
      const double win = 0.5 * (1.0 - std::cos(M_PI * s));
      const double al = d.src_alpha * win;
      const double zk = d.zmin + k * (d.zmax - d.zmin) / (nz - 1.0);
      h->src_keep[b] = 1.0 - al;
      h->src_spatial[b] = al * std::polar(1.0, d.k_beam * (zk - zK));
    }
    h->nsrc = d.src_band;
  }
  const std::complex<double> mi_w(0.0, -d.e_sim);
  if (h->rs == 8) return upload_src_tab<double>(h, mi_w);
  return upload_src_tab<float>(h, mi_w);
}

// The float32 clock of fdtd_3d.cu:365 has no time origin: callers pass t0 = 0 there
// (check_t0), so every consumer of step times (source, demodulation, adjoint seeds)
// agrees on t_n.
int check_t0(const wap_handle *h, double t0) {
  if (h->d.time_mode == WAP_TIME_F32 && t0 != 0.0)
    return fail(WAP_ERR_ARG, "t0 must be 0 with the float32 clock (WAP_TIME_F32)");
  return WAP_OK;
}

double step_time(const wap_handle *h, int64_t step0, int64_t n, double t0) {
  if (h->d.time_mode == WAP_TIME_F32) {
    const float sf = (float)(step0 + n);
    return (double)(sf * (float)h->d.dt);
  }
  return t0 + (double)(step0 + n) * h->d.dt;
}

// Source coefficients of a launch: time-free band weights/phasors plus the
// step's e^{-i w t} (the kernel multiplies them, as batched launches do).
template <typename A>
void fill_source(const wap_handle *h, double t, A &a) {
  using R = decltype(a.hx);
  a.src_k0 = h->src_k0;
  a.src_n = h->nsrc;
  a.src_tab = (const wap::SrcCoef<R> *)h->src_tab;
  const std::complex<double> ph = std::polar(1.0, -h->d.e_sim * t);
  a.src_ph = wap::cplx<R>{(R)ph.real(), (R)ph.imag()};
}

template <typename R>
wap::PassArgs<R> base_args(const wap_handle *h) {
  using C = wap::cplx<R>;
  wap::PassArgs<R> a;
  std::memset((void *)&a, 0, sizeof(a));
  a.V = (const R *)h->V;
  a.zc = (const C *)h->zc;
  a.gam = (const R *)h->gam;
  a.prof = (const C *)h->prof;
  a.nx = h->d.nx;
  a.ny = h->d.ny;
  a.nz = h->d.nz;
  a.tile_y = h->tile_y;
  a.tile_z = h->tile_z;
  a.ntx = h->ntx;
  a.nty = h->nty;
  a.ntz = h->ntz;
  a.hx = (R)h->hx;
  a.hy = (R)h->hy;
  a.m2 = (R)h->m2;
  a.st_psi = (const C *)h->psi;
  a.st_pi = (const C *)h->pi;
  a.acc_psi = (C *)h->acc_psi;
  a.acc_pi = (C *)h->acc_pi;
  a.cmp_psi = (C *)h->cmp_psi;
  a.cmp_pi = (C *)h->cmp_pi;
  a.zero_ends = h->d.end_planes == WAP_ENDS_ZERO ? 1 : 0;
  a.rec_psi = nullptr;
  a.rec_pi = nullptr;
  return a;
}

// Arguments of stage pass s (1..4) of the 4-pass schedule.  ty overrides the
// handle's rows per workgroup (batched launches share one workgroup shape).
template <typename R>
wap::PassArgs<R> stage_args(const wap_handle *h, int s, int ty) {
  using C = wap::cplx<R>;
  wap::PassArgs<R> a = base_args<R>(h);
  a.tile_y = ty;
  a.nty = (h->d.ny + ty - 1) / ty;
  const R dt = (R)h->d.dt;
  C *psi = (C *)h->psi, *pi = (C *)h->pi;
  C *ya = (C *)h->ya_psi, *yaq = (C *)h->ya_pi;
  C *yb = (C *)h->yb_psi, *ybq = (C *)h->yb_pi;
  switch (s) {
    case 1:  // psi_n -> y2, acc
      a.yin_psi = psi; a.yin_pi = pi; a.out_psi = ya; a.out_pi = yaq;
      a.w_acc = dt / (R)6; a.w_y = dt / (R)2;
      break;
    case 2:  // y2 -> y3
      a.yin_psi = ya; a.yin_pi = yaq; a.out_psi = yb; a.out_pi = ybq;
      a.w_acc = dt / (R)3; a.w_y = dt / (R)2;
      break;
    case 3:  // y3 -> y4
      a.yin_psi = yb; a.yin_pi = ybq; a.out_psi = ya; a.out_pi = yaq;
      a.w_acc = dt / (R)3; a.w_y = dt;
      break;
    default:  // y4 -> psi_{n+1}
      a.yin_psi = ya; a.yin_pi = yaq; a.out_psi = psi; a.out_pi = pi;
      a.w_acc = dt / (R)6; a.w_y = (R)0;
      break;
  }
  return a;
}

template <typename R, int EQ, bool K>
int run_steps(wap_handle *h, int64_t nsteps, int64_t step0, double t0, hipEvent_t *ev,
              double *pass_ms) {
  const dim3 block(wap::kWaveX, h->tile_y);
  const dim3 grid((unsigned)(h->ntx * h->nty * h->ntz));
  wap::PassArgs<R> a[4];
  for (int s = 0; s < 4; ++s) a[s] = stage_args<R>(h, s + 1, h->tile_y);
  for (int64_t n = 0; n < nsteps; ++n) {
    const double t = step_time(h, step0, n, t0);
    for (int s = 0; s < 4; ++s) fill_source(h, t, a[s]);
    if (ev) WAP_HIP(hipEventRecord(ev[0], h->stream));
    hipLaunchKernelGGL((wap::rk4_pass<R, EQ, 1, K>), grid, block, 0, h->stream, a[0]);
    if (ev) WAP_HIP(hipEventRecord(ev[1], h->stream));
    hipLaunchKernelGGL((wap::rk4_pass<R, EQ, 2, K>), grid, block, 0, h->stream, a[1]);
    if (ev) WAP_HIP(hipEventRecord(ev[2], h->stream));
    hipLaunchKernelGGL((wap::rk4_pass<R, EQ, 3, K>), grid, block, 0, h->stream, a[2]);
    if (ev) WAP_HIP(hipEventRecord(ev[3], h->stream));
    hipLaunchKernelGGL((wap::rk4_pass<R, EQ, 4, K>), grid, block, 0, h->stream, a[3]);
    if (ev) {
      WAP_HIP(hipEventRecord(ev[4], h->stream));
      WAP_HIP(hipEventSynchronize(ev[4]));
      for (int s = 0; s < 4; ++s) {
        float ms = 0.f;
        WAP_HIP(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
        pass_ms[s] += ms;
      }
    }
  }
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

// The fused passes address each field through buffer descriptors based at the first plane
// of their workgroup's window (its z chunk and the chunk's halo: at most tile_z + 9
// planes) with 32-bit byte offsets, and skip lanes with an offset of 2^31: a window must
// stay below 2 GiB.  The longest chunk whose window fits (capped at the LDS tables' 256
// planes); the fused schedules need one of at least 1 plane, i.e. a plane under
// 2^31 / 10 bytes (26.8 M voxels in complex64), whatever nz.
int window_tz_max(const wap_desc &d) {
  const int64_t cs = d.precision == WAP_PREC_F64 ? 16 : 8;
  const int64_t planes = ((((int64_t)1 << 31) - 1) / ((int64_t)d.nx * d.ny * cs)) - 9;
  return (int)std::min<int64_t>(planes, wap::kKG1MaxTileZ);
}
bool fused_plane_fits(const wap_desc &d) { return window_tz_max(d) >= 1; }

// The fused passes store their results with non-temporal stores when the fields one
// step touches (4 of the state's size: in and out of two components, or input, u3 and
// output) outgrow the 256 MB Infinity Cache; below that the next pass or step rereads
// them from the cache, which non-temporal stores would bypass.  A batch counts its
// instances together.
constexpr size_t kStreamStoreBytes = (size_t)256 << 20;
bool stream_stores(size_t voxels, size_t cs) { return voxels * cs * 4 > kStreamStoreBytes; }

// (hbar c)^2 / dz^2: the uniform z weight of the Klein-Gordon operator (no stretch)
double kg_hz(const wap_handle *h) {
  const double dz = (h->d.zmax - h->d.zmin) / (h->d.nz - 1.0);
  return h->d.hbar_c * h->d.hbar_c / (dz * dz);
}

// x = hi + lo with hi the nearest R and lo the nearest R to the remainder
template <typename R>
void split_hi_lo(double x, R &hi, R &lo) {
  hi = (R)x;
  lo = (R)(x - (double)hi);
}

// the Schrodinger two-pass schedule (every precision): frozen end planes (its levels have
// no end-plane factors for the 2D rule), chunk windows under 2 GiB
bool schr_fused_ok(const wap_desc &d) {
  return d.equation == WAP_EQ_SCHRODINGER && d.end_planes == WAP_ENDS_FROZEN &&
         fused_plane_fits(d);
}
// the one-pass Schrodinger schedule: float32 and float64 (3D grids: float32 with
// compensation too), frozen or zero end planes (its levels' end-plane factors), chunk
// windows under 2 GiB
bool schr_1pass_ok(const wap_desc &d) {
  return d.equation == WAP_EQ_SCHRODINGER && fused_plane_fits(d) &&
         (d.precision != WAP_PREC_F32_KAHAN || d.ny > 1);
}

int sched_of(const wap_handle *h) {
  if (!is_kg(h)) {
    if (h->d.schedule == WAP_SCHED_AUTO)
      return schr_1pass_ok(h->d)   ? WAP_SCHED_SCHR_1PASS
             : schr_fused_ok(h->d) ? WAP_SCHED_SCHR_FUSED
                                   : WAP_SCHED_STAGES;
    return h->d.schedule;
  }
  if (h->d.schedule == WAP_SCHED_AUTO)
    return fused_plane_fits(h->d) ? WAP_SCHED_KG_FUSED : WAP_SCHED_KG_2PASS;
  return h->d.schedule;
}

bool kg_two_pass(const wap_handle *h) {
  const int s = sched_of(h);
  return s == WAP_SCHED_KG_2PASS || s == WAP_SCHED_KG_FUSED;
}

// Fused-pass workgroups: tile_y = 4, 8 or 16 waves (default 16), i.e. rows = waves - 1
// rows of 64 columns (two of them halo rows) plus the side-column wave.  The float64 KG
// pass takes at most 8 (it needs ~180 VGPRs: 2 waves per SIMD).
int kg1_waves(const wap_handle *h) {
  const int ty = h->d.tile_y;
  const int w = (ty == 4 || ty == 8 || ty == 16) ? ty : 16;
  return (is_f64(h) && is_kg(h) && w > 8) ? 8 : w;
}
int kg1_ntx(const wap_handle *h) { return (h->d.nx + wap::kWaveX - 1) / wap::kWaveX; }
int kg1_nty(const wap_handle *h, int waves) {
  const int out_rows = waves - 3;
  return (h->d.ny + out_rows - 1) / out_rows;
}
int kg1_nty(const wap_handle *h) { return kg1_nty(h, kg1_waves(h)); }

// The fused passes hold per_cu workgroups per CU on 256 CUs.  Pick the z chunk
// that minimises (full rounds of workgroups) x (planes per chunk + halo planes):
// few, long chunks amortise the halo planes, but a last round that is mostly empty
// costs a whole round.  A lone grid searches chunks from 2 planes: a small grid needs
// short chunks to occupy the CUs, and a plane step costs its latency whatever the grid
// (64^3: 2-plane chunks run 1.6x as fast as 8-plane ones).  A grid in a batch shares
// the CUs with the other grids, where short chunks only add halo planes: from 8.
int chunk_tile_z(const wap_handle *h, long cols, int per_cu, int halo, bool batch) {
  const int cap = window_tz_max(h->d);  // <= kKG1MaxTileZ
  if (h->d.tile_z > 0) return std::min(std::min(h->d.tile_z, h->d.nz), cap);
  const long slots = 256L * per_cu;
  const int tmax = std::min(h->d.nz, cap);
  const int tmin = batch ? 8 : 2;
  int best = tmax;
  double best_cost = 1e300;
  for (int tz = tmin < tmax ? tmin : tmax; tz <= tmax; ++tz) {
    const long chunks = (h->d.nz + tz - 1) / tz;
    const double cost = (double)((cols * chunks + slots - 1) / slots) * (tz + halo);
    if (cost < best_cost) {  // ties keep the shorter chunk (more workgroups in flight)
      best_cost = cost;
      best = tz;
    }
  }
  return best;
}
int kg1_tile_z(const wap_handle *h, int waves, int per_cu, bool batch = false) {
  return chunk_tile_z(h, (long)kg1_ntx(h) * kg1_nty(h, waves), per_cu, 2, batch);
}
int kg1_tile_z(const wap_handle *h) { return kg1_tile_z(h, kg1_waves(h), 16 / kg1_waves(h)); }

// tiles of the fused passes over one grid with `waves` waves per workgroup
int64_t fused_tiles(const wap_handle *h, int waves, bool batch = false) {
  const int tz = kg1_tile_z(h, waves, 16 / waves, batch);
  return (int64_t)kg1_ntx(h) * kg1_nty(h, waves) * ((h->d.nz + tz - 1) / tz);
}

// the launch-invariant arguments of kg1_pass (buffers and source are set per launch)
template <typename R>
wap::KG1Args<R> kg1_args(const wap_handle *h, int waves, bool batch = false) {
  using C = wap::cplx<R>;
  wap::KG1Args<R> a;
  std::memset((void *)&a, 0, sizeof(a));
  const int tz = kg1_tile_z(h, waves, 16 / waves, batch);
  a.V = (const R *)h->V;
  a.gam = (const R *)h->gam;
  a.prof = (const C *)h->prof;
  a.nx = h->d.nx; a.ny = h->d.ny; a.nz = h->d.nz;
  a.tile_z = tz;
  a.ntx = kg1_ntx(h); a.nty = kg1_nty(h, waves); a.ntz = (h->d.nz + tz - 1) / tz;
  const double hz = kg_hz(h);
  a.hx = (R)h->hx; a.hy = (R)h->hy;
  split_hi_lo(hz, a.hz_hi, a.hz_lo);
  split_hi_lo(2.0 * ((double)a.hx + (double)a.hy + hz) + h->m2, a.d0_hi, a.d0_lo);
  a.h = (R)h->d.dt;
  a.stream_out = stream_stores(h->n, h->cs) ? 1 : 0;
  a.cmp_psi = (C *)h->cmp_psi;  // null unless f32_kahan
  a.cmp_pi = (C *)h->cmp_pi;
  return a;
}

template <typename R>
int run_steps_kg1(wap_handle *h, int64_t nsteps, int64_t step0, double t0, hipEvent_t *ev,
                  double *pass_ms, bool record = false) {
  using C = wap::cplx<R>;
  const int waves = kg1_waves(h);
  const dim3 block(wap::kWaveX, waves);
  const dim3 grid((unsigned)fused_tiles(h, waves));
  wap::KG1Args<R> a = kg1_args<R>(h, waves);
  // A record step's output is the next record's input but then waits in its slot for the
  // backward sweep: non-temporal stores keep the records from evicting the cache-resident
  // fields (37 against 39.5 us per record step at 128 x 128 x 256, where a plain step,
  // whose output the next step rereads from the cache, is 14 % slower with them)
  if (record) a.stream_out = 1;
  const bool kahan = is_kahan(h);  // compensation buffers updated in place
  for (int64_t n = 0; n < nsteps; ++n) {
    fill_source(h, step_time(h, step0, n, t0), a);
    a.in_psi = (const C *)h->psi;
    a.in_pi = (const C *)h->pi;
    a.out_psi = (C *)h->ya_psi;
    a.out_pi = (C *)h->ya_pi;
    if (ev) WAP_HIP(hipEventRecord(ev[0], h->stream));
    if constexpr (sizeof(R) == 4) {
      if (kahan) {
        switch (waves) {
          case 4: hipLaunchKernelGGL((wap::kg1_pass<R, 3, true>), grid, block, 0, h->stream, a); break;
          case 8: hipLaunchKernelGGL((wap::kg1_pass<R, 7, true>), grid, block, 0, h->stream, a); break;
          default: hipLaunchKernelGGL((wap::kg1_pass<R, 15, true>), grid, block, 0, h->stream, a); break;
        }
        goto launched;
      }
    }
    switch (waves) {
      case 4: hipLaunchKernelGGL((wap::kg1_pass<R, 3, false>), grid, block, 0, h->stream, a); break;
      case 8: hipLaunchKernelGGL((wap::kg1_pass<R, 7, false>), grid, block, 0, h->stream, a); break;
      default: hipLaunchKernelGGL((wap::kg1_pass<R, 15, false>), grid, block, 0, h->stream, a); break;
    }
  launched:
    // the new state is in the other buffer
    std::swap(h->psi, h->ya_psi);
    std::swap(h->pi, h->ya_pi);
    if (ev) {
      WAP_HIP(hipEventRecord(ev[1], h->stream));
      WAP_HIP(hipEventSynchronize(ev[1]));
      float ms = 0.f;
      WAP_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
      pass_ms[0] += ms;
    }
  }
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

// the launch-invariant arguments of schr2_pass; pass 0 (A) or 1 (B) sets the level factors
template <typename R>
wap::Schr2Args<R> schr2_args(const wap_handle *h, int waves, int pass, bool batch = false) {
  using C = wap::cplx<R>;
  wap::Schr2Args<R> a;
  std::memset((void *)&a, 0, sizeof(a));
  const int tz = kg1_tile_z(h, waves, 16 / waves, batch);
  a.V = (const R *)h->V;
  a.zc = (const C *)h->zc;
  a.prof = (const C *)h->prof;
  a.nx = h->d.nx; a.ny = h->d.ny; a.nz = h->d.nz;
  a.tile_z = tz;
  a.ntx = kg1_ntx(h); a.nty = kg1_nty(h, waves); a.ntz = (h->d.nz + tz - 1) / tz;
  a.hx = (R)h->hx; a.hy = (R)h->hy;
  const double dt = h->d.dt;
  a.c1 = (R)(pass == 0 ? dt / 4.0 : dt / 2.0);
  a.c2 = (R)(pass == 0 ? dt / 3.0 : dt);
  a.stream_out = stream_stores(h->n, h->cs) ? 1 : 0;
  a.cmp = (pass == 1 && is_kahan(h)) ? (C *)h->cmp_psi : nullptr;
  return a;
}

// Schrodinger step in two fused passes (schr2_pass): A psi -> u3 (ya), B (u3, psi) ->
// psi' (yb), then the state buffer and yb swap; K: f32_kahan (pass B also updates the
// compensation in place)
template <typename R, bool K>
int run_steps_schr2(wap_handle *h, int64_t nsteps, int64_t step0, double t0, hipEvent_t *ev,
                    double *pass_ms) {
  using C = wap::cplx<R>;
  const int waves = kg1_waves(h);
  const dim3 block(wap::kWaveX, waves);
  const dim3 grid((unsigned)fused_tiles(h, waves));
  wap::Schr2Args<R> pa = schr2_args<R>(h, waves, 0), pb = schr2_args<R>(h, waves, 1);
  auto launch = [&](auto pass, const wap::Schr2Args<R> &a) {
    constexpr int PB = decltype(pass)::value;
    constexpr bool KB = K && PB == 1;
    switch (waves) {
      case 4: hipLaunchKernelGGL((wap::schr2_pass<R, 3, PB, KB>), grid, block, 0, h->stream, a); break;
      case 8: hipLaunchKernelGGL((wap::schr2_pass<R, 7, PB, KB>), grid, block, 0, h->stream, a); break;
      default: hipLaunchKernelGGL((wap::schr2_pass<R, 15, PB, KB>), grid, block, 0, h->stream, a); break;
    }
  };
  for (int64_t n = 0; n < nsteps; ++n) {
    const double t = step_time(h, step0, n, t0);
    fill_source(h, t, pa);
    fill_source(h, t, pb);
    pa.in = (const C *)h->psi;
    pa.out = (C *)h->ya_psi;
    pb.in = (const C *)h->ya_psi;
    pb.base = (const C *)h->psi;
    pb.out = (C *)h->yb_psi;
    if (ev) WAP_HIP(hipEventRecord(ev[0], h->stream));
    launch(std::integral_constant<int, 0>{}, pa);
    if (ev) WAP_HIP(hipEventRecord(ev[1], h->stream));
    launch(std::integral_constant<int, 1>{}, pb);
    std::swap(h->psi, h->yb_psi);
    if (ev) {
      WAP_HIP(hipEventRecord(ev[2], h->stream));
      WAP_HIP(hipEventSynchronize(ev[2]));
      for (int s = 0; s < 2; ++s) {
        float ms = 0.f;
        WAP_HIP(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
        pass_ms[s] += ms;
      }
    }
  }
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

// One-pass Schrodinger step (schr1_pass): kSchr1Rows row waves (8 interior rows of 64
// columns) and a side wave per workgroup, kSchr1PerCu workgroups per CU.  A chunk of tz
// planes evaluates 4 tz + 12 levels (the z halo of levels 1..3), which the chunk choice
// counts as 3 extra planes.
int schr1_nty(const wap_handle *h) {
  constexpr int out_rows = wap::kSchr1Rows - 6;
  return (h->d.ny + out_rows - 1) / out_rows;
}
int schr1_tile_z(const wap_handle *h, bool batch = false) {
  return chunk_tile_z(h, (long)kg1_ntx(h) * schr1_nty(h), wap::kSchr1PerCu, 3, batch);
}
int64_t schr1_tiles(const wap_handle *h, bool batch = false) {
  const int tz = schr1_tile_z(h, batch);
  return (int64_t)kg1_ntx(h) * schr1_nty(h) * ((h->d.nz + tz - 1) / tz);
}

template <typename R>
wap::Schr1Args<R> schr1_args(const wap_handle *h, bool batch = false) {
  using C = wap::cplx<R>;
  wap::Schr1Args<R> a;
  std::memset((void *)&a, 0, sizeof(a));
  const int tz = schr1_tile_z(h, batch);
  a.V = (const R *)h->V;
  a.zc = (const C *)h->zc;
  a.prof = (const C *)h->prof;
  a.nx = h->d.nx; a.ny = h->d.ny; a.nz = h->d.nz;
  a.tile_z = tz;
  a.ntx = kg1_ntx(h); a.nty = schr1_nty(h); a.ntz = (h->d.nz + tz - 1) / tz;
  a.hx = (R)h->hx; a.hy = (R)h->hy;
  const double dt = h->d.dt;
  a.f1 = (R)(dt / 4.0);
  a.f2 = (R)(dt / 3.0);
  a.f3 = (R)(dt / 2.0);
  a.f4 = (R)dt;
  // end-plane outputs of levels 1..3: b (frozen planes) or b/2, b/3, b/6 (the 2D rule)
  a.cmp = is_kahan(h) ? (C *)h->cmp_psi : nullptr;  // compensation, updated in place
  const bool zero_ends = h->d.end_planes == WAP_ENDS_ZERO;
  a.g1 = zero_ends ? (R)0.5 : (R)1;
  a.g2 = zero_ends ? (R)(1.0 / 3.0) : (R)1;
  a.g3 = zero_ends ? (R)(1.0 / 6.0) : (R)1;
  a.pz_lo = h->pz_lo;
  a.pz_hi = h->pz_hi;
  a.hz = (R)h->pz_w;
  // the state ping-pongs between psi and yb: input, output and V per step
  a.stream_out = stream_stores(h->n, h->cs) ? 1 : 0;
  return a;
}

// A single 2D grid (ny = 1) runs schr2d_pass: one wave per 56-column x tile and z chunk.
// The waves are independent and up to 8 share a SIMD, which hides a step's latency once
// there are a few per SIMD; beyond that a chunk's 6 ramp steps are the cost.  Measured (us
// per step, f32): 1398 x 1024: 4 planes 15.5, 8: 16.8, 16: 21.3; 4096 x 2048: 4: 46.6,
// 8: 44.7, 16: 48.8, 32: 56.7.  8 planes, halved while the grid makes fewer than 4096
// waves (four per SIMD), doubled while it still makes 65536.
int schr2d_ntx(const wap_handle *h) { return (h->d.nx + wap::kSchr2dCols - 1) / wap::kSchr2dCols; }
int schr2d_tile_z(const wap_handle *h) {
  const int nz = h->d.nz;
  const int cap = window_tz_max(h->d);
  if (h->d.tile_z > 0) return std::min(std::min(h->d.tile_z, nz), cap);
  const long ntx = schr2d_ntx(h);
  auto waves = [&](int tz) { return ntx * ((nz + tz - 1) / tz); };
  int tz = 8;
  while (tz > 2 && waves(tz) < 4096) tz /= 2;
  while (2 * tz <= 64 && waves(2 * tz) >= 65536) tz *= 2;
  return std::min(std::min(tz, nz), cap);
}
int64_t schr2d_tiles(const wap_handle *h) {
  const int tz = schr2d_tile_z(h);
  return (int64_t)schr2d_ntx(h) * ((h->d.nz + tz - 1) / tz);
}
bool schr2d(const wap_handle *h) { return h->d.ny == 1; }
template <typename R>
wap::Schr1Args<R> schr2d_args(const wap_handle *h) {
  wap::Schr1Args<R> a = schr1_args<R>(h);
  const int tz = schr2d_tile_z(h);
  a.tile_z = tz;
  a.ntx = schr2d_ntx(h);
  a.nty = 1;
  a.ntz = (h->d.nz + tz - 1) / tz;
  a.hy = (R)0;
  return a;
}

// Schrodinger step in one fused pass: psi -> psi' (yb), then the two buffers swap.
// record: non-temporal result stores (an adjoint record step, whose output waits in its slot)
template <typename R>
int run_steps_schr1(wap_handle *h, int64_t nsteps, int64_t step0, double t0, hipEvent_t *ev,
                    double *pass_ms, bool record = false) {
  using C = wap::cplx<R>;
  const bool two_d = schr2d(h);
  constexpr int kRows = wap::kSchr1Rows;
  const dim3 block = two_d ? dim3(wap::kWaveX) : dim3(wap::kWaveX, kRows + 1);
  const dim3 grid((unsigned)(two_d ? schr2d_tiles(h) : schr1_tiles(h)));
  wap::Schr1Args<R> a = two_d ? schr2d_args<R>(h) : schr1_args<R>(h);
  if (record) a.stream_out = 1;
  for (int64_t n = 0; n < nsteps; ++n) {
    fill_source(h, step_time(h, step0, n, t0), a);
    a.in = (const C *)h->psi;
    a.out = (C *)h->yb_psi;
    if (ev) WAP_HIP(hipEventRecord(ev[0], h->stream));
    if (two_d) {
      hipLaunchKernelGGL((wap::schr2d_pass<R>), grid, block, 0, h->stream, a);
    } else if constexpr (sizeof(R) == 4) {
      if (a.cmp)
        hipLaunchKernelGGL((wap::schr1_pass<R, kRows, true>), grid, block, 0, h->stream, a);
      else
        hipLaunchKernelGGL((wap::schr1_pass<R, kRows>), grid, block, 0, h->stream, a);
    } else {
      hipLaunchKernelGGL((wap::schr1_pass<R, kRows>), grid, block, 0, h->stream, a);
    }
    std::swap(h->psi, h->yb_psi);
    if (ev) {
      WAP_HIP(hipEventRecord(ev[1], h->stream));
      WAP_HIP(hipEventSynchronize(ev[1]));
      float ms = 0.f;
      WAP_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
      pass_ms[0] += ms;
    }
  }
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

template <typename R>
wap::KG2Args<R> kg2_args(const wap_handle *h, int ty = 0) {
  using C = wap::cplx<R>;
  wap::KG2Args<R> a;
  std::memset((void *)&a, 0, sizeof(a));
  a.st_psi = (const C *)h->psi;
  a.st_pi = (const C *)h->pi;
  a.y3 = (C *)h->ya_psi;
  a.y4 = (C *)h->yb_psi;
  a.z3 = (C *)h->ya_pi;
  a.pp = (C *)h->acc_psi;
  a.qq = (C *)h->acc_pi;
  a.out_psi = (C *)h->psi;
  a.out_pi = (C *)h->pi;
  a.cmp_psi = (C *)h->cmp_psi;
  a.cmp_pi = (C *)h->cmp_pi;
  a.V = (const R *)h->V;
  a.zc = (const C *)h->zc;
  a.gam = (const R *)h->gam;
  a.prof = (const C *)h->prof;
  a.nx = h->d.nx;
  a.ny = h->d.ny;
  a.nz = h->d.nz;
  a.tile_y = h->tile_y;
  a.tile_z = h->tile_z;
  a.ntx = h->ntx;
  a.nty = h->nty;
  a.ntz = h->ntz;
  a.hx = (R)h->hx;
  a.hy = (R)h->hy;
  split_hi_lo(kg_hz(h), a.hz_hi, a.hz_lo);
  split_hi_lo(h->m2, a.m2_hi, a.m2_lo);
  a.h = (R)h->d.dt;
  if (ty > 0) {
    a.tile_y = ty;
    a.nty = (h->d.ny + ty - 1) / ty;
  }
  return a;
}

template <typename R, bool K>
int run_steps_kg2(wap_handle *h, int64_t nsteps, int64_t step0, double t0, hipEvent_t *ev,
                  double *pass_ms) {
  const dim3 block(wap::kWaveX, h->tile_y);
  const dim3 grid((unsigned)(h->ntx * h->nty * h->ntz));
  wap::KG2Args<R> a = kg2_args<R>(h);
  for (int64_t n = 0; n < nsteps; ++n) {
    fill_source(h, step_time(h, step0, n, t0), a);
    if (ev) WAP_HIP(hipEventRecord(ev[0], h->stream));
    hipLaunchKernelGGL((wap::kg2_pass<R, 1, K>), grid, block, 0, h->stream, a);
    if (ev) WAP_HIP(hipEventRecord(ev[1], h->stream));
    hipLaunchKernelGGL((wap::kg2_pass<R, 2, K>), grid, block, 0, h->stream, a);
    if (ev) {
      WAP_HIP(hipEventRecord(ev[2], h->stream));
      WAP_HIP(hipEventSynchronize(ev[2]));
      for (int s = 0; s < 2; ++s) {
        float ms = 0.f;
        WAP_HIP(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
        pass_ms[s] += ms;
      }
    }
  }
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

int dispatch_steps(wap_handle *h, int64_t nsteps, int64_t step0, double t0, hipEvent_t *ev,
                   double *pass_ms) {
  const bool kg = is_kg(h);
  if (sched_of(h) == WAP_SCHED_SCHR_1PASS)
    return is_f64(h) ? run_steps_schr1<double>(h, nsteps, step0, t0, ev, pass_ms)
                     : run_steps_schr1<float>(h, nsteps, step0, t0, ev, pass_ms);
  if (sched_of(h) == WAP_SCHED_SCHR_FUSED) {
    switch (h->d.precision) {
      case WAP_PREC_F64: return run_steps_schr2<double, false>(h, nsteps, step0, t0, ev, pass_ms);
      case WAP_PREC_F32: return run_steps_schr2<float, false>(h, nsteps, step0, t0, ev, pass_ms);
      default: return run_steps_schr2<float, true>(h, nsteps, step0, t0, ev, pass_ms);
    }
  }
  if (sched_of(h) == WAP_SCHED_KG_FUSED) {
    if (h->d.precision == WAP_PREC_F64) return run_steps_kg1<double>(h, nsteps, step0, t0, ev, pass_ms);
    return run_steps_kg1<float>(h, nsteps, step0, t0, ev, pass_ms);
  }
  if (kg_two_pass(h)) {
    switch (h->d.precision) {
      case WAP_PREC_F64: return run_steps_kg2<double, false>(h, nsteps, step0, t0, ev, pass_ms);
      case WAP_PREC_F32: return run_steps_kg2<float, false>(h, nsteps, step0, t0, ev, pass_ms);
      default: return run_steps_kg2<float, true>(h, nsteps, step0, t0, ev, pass_ms);
    }
  }
  switch (h->d.precision) {
    case WAP_PREC_F64:
      return kg ? run_steps<double, 1, false>(h, nsteps, step0, t0, ev, pass_ms)
                : run_steps<double, 0, false>(h, nsteps, step0, t0, ev, pass_ms);
    case WAP_PREC_F32:
      return kg ? run_steps<float, 1, false>(h, nsteps, step0, t0, ev, pass_ms)
                : run_steps<float, 0, false>(h, nsteps, step0, t0, ev, pass_ms);
    default:
      return kg ? run_steps<float, 1, true>(h, nsteps, step0, t0, ev, pass_ms)
                : run_steps<float, 0, true>(h, nsteps, step0, t0, ev, pass_ms);
  }
}

void choose_shape(wap_handle *h) {
  const wap_desc &d = h->d;
  int ty = d.tile_y > 0 ? d.tile_y : 4;
  if (ty > 8) ty = 8;
  if (ty > d.ny) ty = d.ny;   // a 2D grid (ny = 1) runs one-wave workgroups
  h->tile_y = ty;
  h->ntx = (d.nx + wap::kWaveX - 1) / wap::kWaveX;
  h->nty = (d.ny + ty - 1) / ty;
  int tz = d.tile_z;
  if (tz <= 0) {
    // aim for >= 2048 workgroups (8 per CU) with the longest z run that allows
    const long cols = (long)h->ntx * h->nty;
    long t = (long)d.nz * cols / 2048;
    if (t < 2) t = 2;
    if (t > 32) t = 32;
    tz = (int)t;
  }
  if (tz > d.nz) tz = d.nz;
  h->tile_z = tz;
  h->ntz = (d.nz + tz - 1) / tz;
}

int validate(const wap_desc *d) {
  if (!d) return fail(WAP_ERR_ARG, "desc is NULL");
  if (d->nx < 2 || d->ny < 1 || d->nz < 3)
    return fail(WAP_ERR_ARG, "grid must be at least 2 x 1 x 3 (nx >= 2, ny >= 1, nz >= 3)");
  if (d->end_planes != WAP_ENDS_FROZEN && d->end_planes != WAP_ENDS_ZERO)
    return fail(WAP_ERR_ARG, "unknown end_planes");
  if (d->end_planes == WAP_ENDS_ZERO &&
      (d->equation != WAP_EQ_SCHRODINGER || d->precision == WAP_PREC_F32_KAHAN))
    return fail(WAP_ERR_ARG, "zero end planes are for the f32/f64 Schrodinger stepper");
  if ((int64_t)d->nx * d->ny > (int64_t)1 << 30)
    return fail(WAP_ERR_ARG, "nx*ny exceeds the 32-bit in-plane index range");
  if (d->equation != WAP_EQ_SCHRODINGER && d->equation != WAP_EQ_KLEIN_GORDON)
    return fail(WAP_ERR_ARG, "unknown equation");
  if (d->precision < WAP_PREC_F64 || d->precision > WAP_PREC_F32_KAHAN)
    return fail(WAP_ERR_ARG, "unknown precision");
  if (d->time_mode != WAP_TIME_EXACT && d->time_mode != WAP_TIME_F32)
    return fail(WAP_ERR_ARG, "unknown time_mode");
  if (!(d->xmax > d->xmin) || !(d->zmax > d->zmin) || (d->ny > 1 && !(d->ymax > d->ymin)))
    return fail(WAP_ERR_ARG, "extent must be increasing on every axis");
  if (!(d->dt > 0.0)) return fail(WAP_ERR_ARG, "dt must be > 0");
  if (!(d->pml_thick > 0.0)) return fail(WAP_ERR_ARG, "pml_thick must be > 0");
  if (d->src_enabled) {
    if (d->src_band < 1 || d->src_band > WAP_MAX_SRC_BAND)
      return fail(WAP_ERR_ARG, "src_band must be in [1, 8]");
    if (d->src_k < 0 || d->src_k >= d->nz)
      return fail(WAP_ERR_ARG, "source plane k_source must lie inside the grid");
  }
  if (d->tile_y < 0 || d->tile_y > 16 || (d->tile_y & (d->tile_y - 1)))
    return fail(WAP_ERR_ARG, "tile_y must be 0 (auto) or 1, 2, 4, 8, 16");
  if (d->tile_z < 0) return fail(WAP_ERR_ARG, "tile_z must be >= 0");
  if (d->schedule < WAP_SCHED_AUTO || d->schedule > WAP_SCHED_SCHR_1PASS)
    return fail(WAP_ERR_ARG, "unknown schedule");
  if (d->schedule == WAP_SCHED_SCHR_1PASS && !schr_1pass_ok(*d))
    return fail(WAP_ERR_ARG, "the one-pass Schrodinger schedule needs the Schrodinger equation "
                             "in f32 or f64 (f32_kahan on 3D grids only) and planes under "
                             "2^31 / 10 bytes");
  if (d->schedule == WAP_SCHED_SCHR_FUSED && !schr_fused_ok(*d))
    return fail(WAP_ERR_ARG, "the Schrodinger fused schedule needs the Schrodinger equation, "
                             "frozen end planes and planes under 2^31 / 10 bytes");
  if ((d->schedule == WAP_SCHED_KG_2PASS || d->schedule == WAP_SCHED_KG_FUSED) &&
      d->equation != WAP_EQ_KLEIN_GORDON)
    return fail(WAP_ERR_ARG, "the 2-pass and fused schedules are for the Klein-Gordon equation");
  if (d->schedule == WAP_SCHED_KG_FUSED && !fused_plane_fits(*d))
    return fail(WAP_ERR_ARG, "the fused schedule needs planes under 2^31 / 10 bytes");
  return WAP_OK;
}

}  // namespace

extern "C" {

int wap_abi_version(void) { return WAP_ABI_VERSION; }

const char *wap_last_error(void) { return g_err.c_str(); }

int wap_create(const wap_desc *desc, wap_handle **out) {
  if (!out) return fail(WAP_ERR_ARG, "out is NULL");
  *out = nullptr;
  int rc = validate(desc);
  if (rc) return rc;
  wap_handle *h = new (std::nothrow) wap_handle();
  if (!h) return fail(WAP_ERR_OOM, "host allocation failed");
  h->d = *desc;
  h->plane = (size_t)desc->nx * desc->ny;
  h->n = h->plane * desc->nz;
  h->cs = is_f64(h) ? 16 : 8;
  h->rs = is_f64(h) ? 8 : 4;
  choose_shape(h);
  const size_t fb = h->n * h->cs;
  const bool kg = is_kg(h);
  void **fields[] = {&h->psi, &h->acc_psi, &h->ya_psi, &h->yb_psi};
  void **fields_pi[] = {&h->pi, &h->acc_pi, &h->ya_pi, &h->yb_pi};
  for (int s = 0; s < 4 && !rc; ++s) rc = dev_alloc(h, fields[s], fb);
  if (kg) {
    // the 2-pass KG schedule uses 7 complex fields (state + Y3 Y4 Z3 P Q), not 8
    const int npi = kg_two_pass(h) ? 3 : 4;
    for (int s = 0; s < npi && !rc; ++s) rc = dev_alloc(h, fields_pi[s], fb);
  }
  if (!rc && is_kahan(h)) {
    rc = dev_alloc(h, &h->cmp_psi, fb);
    if (!rc && kg) rc = dev_alloc(h, &h->cmp_pi, fb);
  }
  if (!rc) rc = dev_alloc(h, &h->V, h->n * h->rs);
  if (!rc) rc = dev_alloc(h, &h->zc, 2 * (size_t)desc->nz * h->cs);
  if (!rc) rc = dev_alloc(h, &h->gam, (size_t)desc->nz * h->rs);
  if (!rc) rc = dev_alloc(h, &h->src_tab, wap::kMaxBand * sizeof(wap::SrcCoef<double>));
  h->n_partial = grid_1d(h->n) < 1024 ? grid_1d(h->n) : 1024;
  if (!rc) rc = dev_alloc(h, (void **)&h->partial, (size_t)h->n_partial * 8);
  if (!rc) rc = dev_alloc(h, (void **)&h->dsum, 8);
  if (!rc) rc = build_tables(h);
  if (!rc) {
    // zero state (fdtd_3d.cu:322-326); V must be set explicitly
    void *zero_bufs[] = {h->psi, h->pi, h->cmp_psi, h->cmp_pi};
    for (void *b : zero_bufs) {
      if (!b) continue;
      hipError_t e = hipMemset(b, 0, fb);
      if (e != hipSuccess) {
        rc = fail(WAP_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
        break;
      }
    }
  }
  if (rc) {
    free_all(h);
    delete h;
    return rc;
  }
  *out = h;
  return WAP_OK;
}

int wap_destroy(wap_handle *h) {
  if (!h) return WAP_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  else (void)hipDeviceSynchronize();
  free_all(h);
  delete h;
  return WAP_OK;
}

int wap_set_stream(wap_handle *h, void *stream) {
  if (!h) return fail(WAP_ERR_ARG, "handle is NULL");
  h->stream = (hipStream_t)stream;
  return WAP_OK;
}

int wap_set_potential(wap_handle *h, const void *v, int dtype, int on_device) {
  if (!h || !v) return fail(WAP_ERR_ARG, "handle or V is NULL");
  if (dtype != WAP_DT_F32 && dtype != WAP_DT_F64)
    return fail(WAP_ERR_ARG, "V dtype must be WAP_DT_F32 or WAP_DT_F64");
  const size_t in_b = (dtype == WAP_DT_F64 ? 8 : 4);
  const bool same = (in_b == h->rs);
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if (same) {
    WAP_HIP(hipMemcpyAsync(h->V, v, h->n * in_b, kind, h->stream));
  } else {
    const void *src = v;
    if (!on_device) {
      int rc = ensure_staging(h, h->n * in_b);
      if (rc) return rc;
      WAP_HIP(hipMemcpyAsync(h->staging, v, h->n * in_b, hipMemcpyHostToDevice, h->stream));
      src = h->staging;
    }
    if (dtype == WAP_DT_F32)
      hipLaunchKernelGGL((wap::import_real<float, double>), dim3(grid_1d(h->n)), dim3(256), 0,
                         h->stream, (const float *)src, (double *)h->V, h->n);
    else
      hipLaunchKernelGGL((wap::import_real<double, float>), dim3(grid_1d(h->n)), dim3(256), 0,
                         h->stream, (const double *)src, (float *)h->V, h->n);
    WAP_HIP(hipGetLastError());
  }
  WAP_HIP(hipStreamSynchronize(h->stream));
  h->v_set = true;
  return WAP_OK;
}

namespace {
int import_one(wap_handle *h, const void *src, int dtype, int on_device, void *dst_hi,
               void *dst_cmp) {
  const size_t in_b = (dtype == WAP_DT_C128 ? 16 : 8);
  if (!src) {
    WAP_HIP(hipMemsetAsync(dst_hi, 0, h->n * h->cs, h->stream));
    if (dst_cmp) WAP_HIP(hipMemsetAsync(dst_cmp, 0, h->n * h->cs, h->stream));
    return WAP_OK;
  }
  const void *s = src;
  if (!on_device) {
    int rc = ensure_staging(h, h->n * in_b);
    if (rc) return rc;
    WAP_HIP(hipMemcpyAsync(h->staging, src, h->n * in_b, hipMemcpyHostToDevice, h->stream));
    s = h->staging;
  }
  const dim3 g(grid_1d(h->n)), b(256);
  if (is_f64(h)) {
    if (dtype == WAP_DT_C128)
      WAP_HIP(hipMemcpyAsync(dst_hi, s, h->n * 16, hipMemcpyDeviceToDevice, h->stream));
    else
      hipLaunchKernelGGL((wap::import_state<float, double>), g, b, 0, h->stream,
                         (const float2 *)s, (double2 *)dst_hi, (double2 *)nullptr, h->n);
  } else {
    if (dtype == WAP_DT_C128)
      hipLaunchKernelGGL((wap::import_state<double, float>), g, b, 0, h->stream,
                         (const double2 *)s, (float2 *)dst_hi, (float2 *)dst_cmp, h->n);
    else {
      WAP_HIP(hipMemcpyAsync(dst_hi, s, h->n * 8, hipMemcpyDeviceToDevice, h->stream));
      if (dst_cmp) WAP_HIP(hipMemsetAsync(dst_cmp, 0, h->n * 8, h->stream));
    }
  }
  WAP_HIP(hipGetLastError());
  // the staging buffer is reused by the next import, and a device source may be released
  // by the caller: finish this one first, unless the caller orders that itself
  if (on_device != WAP_STREAM_ORDERED) WAP_HIP(hipStreamSynchronize(h->stream));
  return WAP_OK;
}

int export_one(wap_handle *h, void *dst, int dtype, int on_device, const void *src_hi,
               const void *src_cmp) {
  const size_t out_b = (dtype == WAP_DT_C128 ? 16 : 8);
  const bool direct = (out_b == h->cs) && !src_cmp;
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (direct) {
    WAP_HIP(hipMemcpyAsync(dst, src_hi, h->n * out_b, kind, h->stream));
  } else {
    void *d = dst;
    if (!on_device) {
      int rc = ensure_staging(h, h->n * out_b);
      if (rc) return rc;
      d = h->staging;
    }
    const dim3 g(grid_1d(h->n)), b(256);
    if (is_f64(h)) {
      if (dtype == WAP_DT_C64)
        hipLaunchKernelGGL((wap::export_state<double, float>), g, b, 0, h->stream,
                           (const double2 *)src_hi, (const double2 *)nullptr, (float2 *)d, h->n);
      else
        hipLaunchKernelGGL((wap::export_state<double, double>), g, b, 0, h->stream,
                           (const double2 *)src_hi, (const double2 *)nullptr, (double2 *)d, h->n);
    } else {
      if (dtype == WAP_DT_C64)
        hipLaunchKernelGGL((wap::export_state<float, float>), g, b, 0, h->stream,
                           (const float2 *)src_hi, (const float2 *)src_cmp, (float2 *)d, h->n);
      else
        hipLaunchKernelGGL((wap::export_state<float, double>), g, b, 0, h->stream,
                           (const float2 *)src_hi, (const float2 *)src_cmp, (double2 *)d, h->n);
    }
    WAP_HIP(hipGetLastError());
    if (!on_device)
      WAP_HIP(hipMemcpyAsync(dst, h->staging, h->n * out_b, hipMemcpyDeviceToHost, h->stream));
  }
  if (on_device != WAP_STREAM_ORDERED) WAP_HIP(hipStreamSynchronize(h->stream));
  return WAP_OK;
}
}  // namespace

int wap_set_state(wap_handle *h, const void *psi, const void *pi, int dtype, int on_device) {
  if (!h) return fail(WAP_ERR_ARG, "handle is NULL");
  if (dtype != WAP_DT_C64 && dtype != WAP_DT_C128)
    return fail(WAP_ERR_ARG, "state dtype must be WAP_DT_C64 or WAP_DT_C128");
  int rc = import_one(h, psi, dtype, on_device, h->psi, h->cmp_psi);
  if (rc) return rc;
  if (is_kg(h)) rc = import_one(h, psi ? pi : nullptr, dtype, on_device, h->pi, h->cmp_pi);
  return rc;
}

int wap_get_state(wap_handle *h, void *psi, void *pi, int dtype, int on_device) {
  if (!h || !psi) return fail(WAP_ERR_ARG, "handle or psi is NULL");
  if (dtype != WAP_DT_C64 && dtype != WAP_DT_C128)
    return fail(WAP_ERR_ARG, "state dtype must be WAP_DT_C64 or WAP_DT_C128");
  int rc = export_one(h, psi, dtype, on_device, h->psi, h->cmp_psi);
  if (rc) return rc;
  if (pi) {
    if (!is_kg(h)) return fail(WAP_ERR_ARG, "pi requested from a Schrodinger handle");
    rc = export_one(h, pi, dtype, on_device, h->pi, h->cmp_pi);
  }
  return rc;
}

int wap_set_source_profile(wap_handle *h, const void *prof, int dtype, int on_device) {
  if (!h) return fail(WAP_ERR_ARG, "handle is NULL");
  if (!prof) {
    if (h->prof) {
      WAP_HIP(hipStreamSynchronize(h->stream));
      (void)hipFree(h->prof);
      h->bytes -= (int64_t)(h->plane * h->cs);
      h->prof = nullptr;
    }
    return WAP_OK;
  }
  if (dtype != WAP_DT_C64 && dtype != WAP_DT_C128)
    return fail(WAP_ERR_ARG, "profile dtype must be WAP_DT_C64 or WAP_DT_C128");
  if (!h->prof) {
    int rc = dev_alloc(h, &h->prof, h->plane * h->cs);
    if (rc) return rc;
  }
  // convert on the host side (one plane)
  const size_t in_b = (dtype == WAP_DT_C128 ? 16 : 8);
  std::vector<unsigned char> hostbuf(h->plane * in_b);
  if (on_device)
    WAP_HIP(hipMemcpy(hostbuf.data(), prof, hostbuf.size(), hipMemcpyDeviceToHost));
  else
    std::memcpy(hostbuf.data(), prof, hostbuf.size());
  std::vector<unsigned char> conv(h->plane * h->cs);
  for (size_t t = 0; t < 2 * h->plane; ++t) {
    const double v = (in_b == 16) ? ((const double *)hostbuf.data())[t]
                                  : (double)((const float *)hostbuf.data())[t];
    if (h->cs == 16) ((double *)conv.data())[t] = v;
    else ((float *)conv.data())[t] = (float)v;
  }
  WAP_HIP(hipMemcpyAsync(h->prof, conv.data(), conv.size(), hipMemcpyHostToDevice, h->stream));
  WAP_HIP(hipStreamSynchronize(h->stream));
  return WAP_OK;
}

int wap_step(wap_handle *h, int64_t nsteps, int64_t step0, double t0) {
  if (!h) return fail(WAP_ERR_ARG, "handle is NULL");
  if (!h->v_set) return fail(WAP_ERR_STATE, "potential not set (wap_set_potential)");
  if (nsteps < 0) return fail(WAP_ERR_ARG, "nsteps < 0");
  if (int rc = check_t0(h, t0)) return rc;
  return dispatch_steps(h, nsteps, step0, t0, nullptr, nullptr);
}

int wap_step_profiled(wap_handle *h, int64_t nsteps, int64_t step0, double t0, double *pass_ms,
                      double *total_ms) {
  if (!h || !pass_ms || !total_ms) return fail(WAP_ERR_ARG, "NULL argument");
  if (!h->v_set) return fail(WAP_ERR_STATE, "potential not set (wap_set_potential)");
  if (int rc = check_t0(h, t0)) return rc;
  hipEvent_t ev[5], t_start, t_stop;
  for (int s = 0; s < 5; ++s) WAP_HIP(hipEventCreate(&ev[s]));
  WAP_HIP(hipEventCreate(&t_start));
  WAP_HIP(hipEventCreate(&t_stop));
  for (int s = 0; s < 4; ++s) pass_ms[s] = 0.0;
  WAP_HIP(hipEventRecord(t_start, h->stream));
  int rc = dispatch_steps(h, nsteps, step0, t0, ev, pass_ms);
  if (!rc) {
    WAP_HIP(hipEventRecord(t_stop, h->stream));
    WAP_HIP(hipEventSynchronize(t_stop));
    float ms = 0.f;
    WAP_HIP(hipEventElapsedTime(&ms, t_start, t_stop));
    *total_ms = ms;
  }
  for (int s = 0; s < 5; ++s) (void)hipEventDestroy(ev[s]);
  (void)hipEventDestroy(t_start);
  (void)hipEventDestroy(t_stop);
  return rc;
}

int wap_get_plane(wap_handle *h, int k, void *out, int dtype, int on_device) {
  if (!h || !out) return fail(WAP_ERR_ARG, "handle or out is NULL");
  if (k < 0 || k >= h->d.nz) return fail(WAP_ERR_ARG, "plane index out of range");
  if (dtype != WAP_DT_C64 && dtype != WAP_DT_C128)
    return fail(WAP_ERR_ARG, "plane dtype must be WAP_DT_C64 or WAP_DT_C128");
  int rc = ensure_staging(h, h->plane * 16);
  if (rc) return rc;
  WAP_HIP(hipMemsetAsync(h->staging, 0, h->plane * 16, h->stream));
  const int g = (int)((h->plane + 255) / 256);
  const size_t off = (size_t)k * h->plane;
  if (is_f64(h))
    hipLaunchKernelGGL((wap::accumulate_plane<double>), dim3(g), dim3(256), 0, h->stream,
                       (const double2 *)h->psi, (const double2 *)nullptr, off, (int)h->plane, 1.0,
                       0.0, (double2 *)h->staging);
  else
    hipLaunchKernelGGL((wap::accumulate_plane<float>), dim3(g), dim3(256), 0, h->stream,
                       (const float2 *)h->psi, (const float2 *)h->cmp_psi, off, (int)h->plane, 1.0,
                       0.0, (double2 *)h->staging);
  WAP_HIP(hipGetLastError());
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (dtype == WAP_DT_C128) {
    WAP_HIP(hipMemcpyAsync(out, h->staging, h->plane * 16, kind, h->stream));
    WAP_HIP(hipStreamSynchronize(h->stream));
  } else {
    std::vector<double> tmp(2 * h->plane);
    WAP_HIP(hipMemcpyAsync(tmp.data(), h->staging, h->plane * 16, hipMemcpyDeviceToHost,
                           h->stream));
    WAP_HIP(hipStreamSynchronize(h->stream));
    std::vector<float> f(2 * h->plane);
    for (size_t t = 0; t < f.size(); ++t) f[t] = (float)tmp[t];
    if (on_device) WAP_HIP(hipMemcpy(out, f.data(), f.size() * 4, hipMemcpyHostToDevice));
    else std::memcpy(out, f.data(), f.size() * 4);
  }
  return WAP_OK;
}

int wap_accumulate_plane(wap_handle *h, int k, double cos_p, double sin_p, void *acc) {
  if (!h || !acc) return fail(WAP_ERR_ARG, "handle or acc is NULL");
  if (k < 0 || k >= h->d.nz) return fail(WAP_ERR_ARG, "plane index out of range");
  const int g = (int)((h->plane + 255) / 256);
  const size_t off = (size_t)k * h->plane;
  if (is_f64(h))
    hipLaunchKernelGGL((wap::accumulate_plane<double>), dim3(g), dim3(256), 0, h->stream,
                       (const double2 *)h->psi, (const double2 *)nullptr, off, (int)h->plane,
                       cos_p, sin_p, (double2 *)acc);
  else
    hipLaunchKernelGGL((wap::accumulate_plane<float>), dim3(g), dim3(256), 0, h->stream,
                       (const float2 *)h->psi, (const float2 *)h->cmp_psi, off, (int)h->plane,
                       cos_p, sin_p, (double2 *)acc);
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

int wap_norm2(wap_handle *h, double *out) {
  if (!h || !out) return fail(WAP_ERR_ARG, "handle or out is NULL");
  const int g = h->n_partial;
  if (is_f64(h))
    hipLaunchKernelGGL((wap::norm2_partial<double>), dim3(g), dim3(256), 0, h->stream,
                       (const double2 *)h->psi, (const double2 *)nullptr, h->n, h->partial);
  else
    hipLaunchKernelGGL((wap::norm2_partial<float>), dim3(g), dim3(256), 0, h->stream,
                       (const float2 *)h->psi, (const float2 *)h->cmp_psi, h->n, h->partial);
  hipLaunchKernelGGL(wap::sum_partials, dim3(1), dim3(256), 0, h->stream, h->partial, g, h->dsum);
  WAP_HIP(hipGetLastError());
  WAP_HIP(hipMemcpyAsync(out, h->dsum, 8, hipMemcpyDeviceToHost, h->stream));
  WAP_HIP(hipStreamSynchronize(h->stream));
  return WAP_OK;
}

int wap_synchronize(wap_handle *h) {
  if (!h) return fail(WAP_ERR_ARG, "handle is NULL");
  WAP_HIP(hipStreamSynchronize(h->stream));
  return WAP_OK;
}

int wap_device_bytes(wap_handle *h, int64_t *out) {
  if (!h || !out) return fail(WAP_ERR_ARG, "NULL argument");
  *out = h->bytes;
  return WAP_OK;
}

int wap_launch_shape(wap_handle *h, int32_t *tile_y, int32_t *tile_z, int64_t *groups) {
  if (!h) return fail(WAP_ERR_ARG, "handle is NULL");
  if (sched_of(h) == WAP_SCHED_SCHR_1PASS) {
    if (tile_y) *tile_y = schr2d(h) ? 1 : wap::kSchr1Rows + 1;
    if (tile_z) *tile_z = schr2d(h) ? schr2d_tile_z(h) : schr1_tile_z(h);
    if (groups) *groups = schr2d(h) ? schr2d_tiles(h) : schr1_tiles(h);
    return WAP_OK;
  }
  if (sched_of(h) == WAP_SCHED_KG_FUSED || sched_of(h) == WAP_SCHED_SCHR_FUSED) {
    const int tz = kg1_tile_z(h);
    if (tile_y) *tile_y = kg1_waves(h);
    if (tile_z) *tile_z = tz;
    if (groups) *groups = (int64_t)kg1_ntx(h) * kg1_nty(h) * ((h->d.nz + tz - 1) / tz);
    return WAP_OK;
  }
  if (tile_y) *tile_y = h->tile_y;
  if (tile_z) *tile_z = h->tile_z;
  if (groups) *groups = (int64_t)h->ntx * h->nty * h->ntz;
  return WAP_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches
struct wap_batch {
  std::vector<wap_handle *> hs;
  bool kg2 = false;
  int sched = WAP_SCHED_STAGES;  // the handles' schedule (all equal)
  int waves = 16;                // fused schedules: waves per workgroup
  int precision = 0;
  int equation = 0;
  int tile_y = 4;
  int total_tiles = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void *d_args[4] = {nullptr, nullptr, nullptr, nullptr};
  int *d_tile_begin = nullptr;
  // [steps][ninst] complex e^{-i w_i t_n}: one table for eager steps (regrown on demand)
  // and one the captured graph reads, sized at capture and never moved while the graph
  // lives (the graph's kernel arguments hold its address)
  void *d_phase = nullptr;
  int64_t phase_steps = 0;
  void *d_phase_graph = nullptr;
  int64_t phase_graph_steps = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t graph_steps = 0;
  std::vector<void *> baked_psi;  // state pointers the uploaded arguments refer to
};

namespace {

void batch_free(wap_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void *&p : b->d_args)
    if (p) { (void)hipFree(p); p = nullptr; }
  if (b->d_tile_begin) (void)hipFree(b->d_tile_begin);
  if (b->d_phase) (void)hipFree(b->d_phase);
  if (b->d_phase_graph) (void)hipFree(b->d_phase_graph);
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  b->exec = nullptr;
  b->graph = nullptr;
  b->d_tile_begin = nullptr;
  b->d_phase = nullptr;
  b->d_phase_graph = nullptr;
}

// batched launches read the per-step phasor from a device table instead of src_ph
template <typename A>
void timeless_source(const wap_handle *h, A &a) {
  fill_source(h, 0.0, a);
}

// (Re)write the per-instance arguments into the batch's device arrays (the
// arrays themselves never move: a captured graph keeps pointing at them).
bool batch_fused(const wap_batch *b) {
  return b->sched == WAP_SCHED_KG_FUSED || b->sched == WAP_SCHED_SCHR_FUSED ||
         b->sched == WAP_SCHED_SCHR_1PASS;
}

// Fused batches ping-pong every instance's state: step t of a launch sequence uses the
// argument set of parity t & 1 (parity 0 reads the handle's current state buffer).
template <typename R>
int batch_upload_fused(wap_batch *b) {
  using C = wap::cplx<R>;
  const int n = (int)b->hs.size();
  size_t voxels = 0;
  for (const wap_handle *h : b->hs) voxels += h->n;
  const int stream = stream_stores(voxels, sizeof(C)) ? 1 : 0;  // the batch's fields together
  WAP_HIP(hipStreamSynchronize(b->stream));
  for (int par = 0; par < 2; ++par) {
    if (b->sched == WAP_SCHED_KG_FUSED) {
      std::vector<wap::KG1Args<R>> a(n);
      for (int i = 0; i < n; ++i) {
        const wap_handle *h = b->hs[i];
        a[i] = kg1_args<R>(h, b->waves, true);
        timeless_source(h, a[i]);
        a[i].in_psi = (const C *)(par ? h->ya_psi : h->psi);
        a[i].in_pi = (const C *)(par ? h->ya_pi : h->pi);
        a[i].out_psi = (C *)(par ? h->psi : h->ya_psi);
        a[i].out_pi = (C *)(par ? h->pi : h->ya_pi);
        a[i].stream_out = stream;
      }
      WAP_HIP(hipMemcpy(b->d_args[par], a.data(), n * sizeof(a[0]), hipMemcpyHostToDevice));
    } else if (b->sched == WAP_SCHED_SCHR_1PASS) {
      std::vector<wap::Schr1Args<R>> a(n);
      for (int i = 0; i < n; ++i) {
        const wap_handle *h = b->hs[i];
        a[i] = schr1_args<R>(h, true);
        timeless_source(h, a[i]);
        a[i].in = (const C *)(par ? h->yb_psi : h->psi);
        a[i].out = (C *)(par ? h->psi : h->yb_psi);
        a[i].stream_out = stream;
      }
      WAP_HIP(hipMemcpy(b->d_args[par], a.data(), n * sizeof(a[0]), hipMemcpyHostToDevice));
    } else {
      std::vector<wap::Schr2Args<R>> pa(n), pb(n);
      for (int i = 0; i < n; ++i) {
        const wap_handle *h = b->hs[i];
        const C *cur = (const C *)(par ? h->yb_psi : h->psi);
        pa[i] = schr2_args<R>(h, b->waves, 0, true);
        pb[i] = schr2_args<R>(h, b->waves, 1, true);
        timeless_source(h, pa[i]);
        timeless_source(h, pb[i]);
        pa[i].in = cur;
        pa[i].out = (C *)h->ya_psi;
        pb[i].in = (const C *)h->ya_psi;
        pb[i].base = cur;
        pb[i].out = (C *)(par ? h->psi : h->yb_psi);
        pa[i].stream_out = pb[i].stream_out = stream;
      }
      WAP_HIP(hipMemcpy(b->d_args[2 * par], pa.data(), n * sizeof(pa[0]), hipMemcpyHostToDevice));
      WAP_HIP(hipMemcpy(b->d_args[2 * par + 1], pb.data(), n * sizeof(pb[0]), hipMemcpyHostToDevice));
    }
  }
  b->baked_psi.resize(n);
  for (int i = 0; i < n; ++i) b->baked_psi[i] = b->hs[i]->psi;
  return WAP_OK;
}

// after an odd number of fused batched steps the state sits in the other buffer
void batch_flip(wap_batch *b, int64_t nsteps) {
  if (!batch_fused(b) || !(nsteps & 1)) return;
  for (wap_handle *h : b->hs) {
    if (b->sched == WAP_SCHED_KG_FUSED) {
      std::swap(h->psi, h->ya_psi);
      std::swap(h->pi, h->ya_pi);
    } else {
      std::swap(h->psi, h->yb_psi);
    }
  }
}

template <typename R>
int batch_upload(wap_batch *b) {
  const int n = (int)b->hs.size();
  if (batch_fused(b)) return batch_upload_fused<R>(b);
  if (b->kg2) {
    std::vector<wap::KG2Args<R>> a(n);
    for (int i = 0; i < n; ++i) {
      a[i] = kg2_args<R>(b->hs[i], b->tile_y);
      timeless_source(b->hs[i], a[i]);
    }
    WAP_HIP(hipStreamSynchronize(b->stream));
    WAP_HIP(hipMemcpy(b->d_args[0], a.data(), n * sizeof(wap::KG2Args<R>), hipMemcpyHostToDevice));
  } else {
    WAP_HIP(hipStreamSynchronize(b->stream));
    for (int s = 0; s < 4; ++s) {
      std::vector<wap::PassArgs<R>> a(n);
      for (int i = 0; i < n; ++i) {
        a[i] = stage_args<R>(b->hs[i], s + 1, b->tile_y);
        timeless_source(b->hs[i], a[i]);
      }
      WAP_HIP(hipMemcpy(b->d_args[s], a.data(), n * sizeof(wap::PassArgs<R>),
                        hipMemcpyHostToDevice));
    }
  }
  b->baked_psi.resize(n);
  for (int i = 0; i < n; ++i) b->baked_psi[i] = b->hs[i]->psi;
  return WAP_OK;
}

template <typename R>
int batch_build(wap_batch *b) {
  const int n = (int)b->hs.size();
  std::vector<int> begin(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const wap_handle *h = b->hs[i];
    const int nty = (h->d.ny + b->tile_y - 1) / b->tile_y;
    const int64_t t = b->sched == WAP_SCHED_SCHR_1PASS ? schr1_tiles(h, true)
                      : batch_fused(b)                   ? fused_tiles(h, b->waves, true)
                                                         : (int64_t)h->ntx * nty * h->ntz;
    if (begin[i] + t > INT32_MAX) return fail(WAP_ERR_ARG, "too many tiles in one batch");
    begin[i + 1] = begin[i] + (int)t;
  }
  b->total_tiles = begin[n];
  WAP_HIP(hipMalloc((void **)&b->d_tile_begin, (n + 1) * sizeof(int)));
  WAP_HIP(hipMemcpy(b->d_tile_begin, begin.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
  size_t one = b->kg2 ? sizeof(wap::KG2Args<R>) : sizeof(wap::PassArgs<R>);
  int nargs = b->kg2 ? 1 : 4;
  if (b->sched == WAP_SCHED_KG_FUSED) {
    one = sizeof(wap::KG1Args<R>);
    nargs = 2;
  } else if (b->sched == WAP_SCHED_SCHR_FUSED) {
    one = sizeof(wap::Schr2Args<R>);
    nargs = 4;
  } else if (b->sched == WAP_SCHED_SCHR_1PASS) {
    one = sizeof(wap::Schr1Args<R>);
    nargs = 2;
  }
  for (int s = 0; s < nargs; ++s) WAP_HIP(hipMalloc(&b->d_args[s], n * one));
  return batch_upload<R>(b);
}

// single-grid fused steps swap a handle's state buffers: re-point the arguments
int batch_refresh(wap_batch *b) {
  for (size_t i = 0; i < b->hs.size(); ++i)
    if (b->hs[i]->psi != b->baked_psi[i])
      return b->precision == WAP_PREC_F64 ? batch_upload<double>(b) : batch_upload<float>(b);
  return WAP_OK;
}

// e^{-i w_i t_n} for n in [0, nsteps) into the device table *table of *cap steps, grown
// when nsteps exceeds it (synchronous upload).  A table a live graph reads is never
// passed here with more steps than it holds (the graph's table is sized at capture).
template <typename R>
int batch_phases(wap_batch *b, int64_t nsteps, int64_t step0, void **table, int64_t *cap) {
  const int n = (int)b->hs.size();
  if (nsteps > *cap) {
    if (*table) {
      WAP_HIP(hipStreamSynchronize(b->stream));
      (void)hipFree(*table);
      *table = nullptr;
      *cap = 0;
    }
    WAP_HIP(hipMalloc(table, (size_t)nsteps * n * sizeof(wap::cplx<R>)));
    *cap = nsteps;
  }
  std::vector<wap::cplx<R>> ph((size_t)nsteps * n);
  for (int64_t t = 0; t < nsteps; ++t)
    for (int i = 0; i < n; ++i) {
      const wap_handle *h = b->hs[i];
      const std::complex<double> z = std::polar(1.0, -h->d.e_sim * step_time(h, step0, t, 0.0));
      ph[(size_t)t * n + i] = wap::cplx<R>{(R)z.real(), (R)z.imag()};
    }
  WAP_HIP(hipStreamSynchronize(b->stream));
  WAP_HIP(hipMemcpy(*table, ph.data(), ph.size() * sizeof(wap::cplx<R>),
                    hipMemcpyHostToDevice));
  return WAP_OK;
}

template <typename R>
void batch_launch_fused(wap_batch *b, int64_t t, const wap::cplx<R> *ph) {
  const int n = (int)b->hs.size();
  const dim3 block(wap::kWaveX, b->waves);
  const dim3 grid((unsigned)b->total_tiles);
  const int par = (int)(t & 1);
  if (b->sched == WAP_SCHED_KG_FUSED) {
    const wap::KG1Args<R> *a = (const wap::KG1Args<R> *)b->d_args[par];
    if constexpr (sizeof(R) == 4) {
      if (b->precision == WAP_PREC_F32_KAHAN) {
        switch (b->waves) {
          case 4: hipLaunchKernelGGL((wap::kg1_pass_batched<R, 3, true>), grid, block, 0, b->stream, a, b->d_tile_begin, n, ph); break;
          case 8: hipLaunchKernelGGL((wap::kg1_pass_batched<R, 7, true>), grid, block, 0, b->stream, a, b->d_tile_begin, n, ph); break;
          default: hipLaunchKernelGGL((wap::kg1_pass_batched<R, 15, true>), grid, block, 0, b->stream, a, b->d_tile_begin, n, ph); break;
        }
        return;
      }
    }
    switch (b->waves) {
      case 4: hipLaunchKernelGGL((wap::kg1_pass_batched<R, 3, false>), grid, block, 0, b->stream, a, b->d_tile_begin, n, ph); break;
      case 8: hipLaunchKernelGGL((wap::kg1_pass_batched<R, 7, false>), grid, block, 0, b->stream, a, b->d_tile_begin, n, ph); break;
      default: hipLaunchKernelGGL((wap::kg1_pass_batched<R, 15, false>), grid, block, 0, b->stream, a, b->d_tile_begin, n, ph); break;
    }
    return;
  }
  if (b->sched == WAP_SCHED_SCHR_1PASS) {
    constexpr int kRows = wap::kSchr1Rows;
    const wap::Schr1Args<R> *a = (const wap::Schr1Args<R> *)b->d_args[par];
    const dim3 blk(wap::kWaveX, kRows + 1);
    if (b->precision == WAP_PREC_F32_KAHAN) {
      if constexpr (sizeof(R) == 4)
        hipLaunchKernelGGL((wap::schr1_pass_batched<R, kRows, true>), grid, blk, 0, b->stream, a,
                           b->d_tile_begin, n, ph);
    } else {
      hipLaunchKernelGGL((wap::schr1_pass_batched<R, kRows>), grid, blk, 0, b->stream, a,
                         b->d_tile_begin, n, ph);
    }
    return;
  }
  const wap::Schr2Args<R> *pa = (const wap::Schr2Args<R> *)b->d_args[2 * par];
  const wap::Schr2Args<R> *pb = (const wap::Schr2Args<R> *)b->d_args[2 * par + 1];
  auto two = [&](auto rows, auto kahan) {
    constexpr int RW = decltype(rows)::value;
    constexpr bool KB = decltype(kahan)::value;
    hipLaunchKernelGGL((wap::schr2_pass_batched<R, RW, 0, false>), grid, block, 0, b->stream, pa, b->d_tile_begin, n, ph);
    hipLaunchKernelGGL((wap::schr2_pass_batched<R, RW, 1, KB>), grid, block, 0, b->stream, pb, b->d_tile_begin, n, ph);
  };
  auto rows = [&](auto kahan) {
    switch (b->waves) {
      case 4: two(std::integral_constant<int, 3>{}, kahan); break;
      case 8: two(std::integral_constant<int, 7>{}, kahan); break;
      default: two(std::integral_constant<int, 15>{}, kahan); break;
    }
  };
  if constexpr (sizeof(R) == 4) {
    if (b->precision == WAP_PREC_F32_KAHAN) {
      rows(std::true_type{});
      return;
    }
  }
  rows(std::false_type{});
}

template <typename R, int EQ, bool K>
int batch_launch(wap_batch *b, int64_t nsteps, const void *phase) {
  using C = wap::cplx<R>;
  const int n = (int)b->hs.size();
  const dim3 block(wap::kWaveX, b->tile_y);
  const dim3 grid((unsigned)b->total_tiles);
  for (int64_t t = 0; t < nsteps; ++t) {
    const C *ph = (const C *)phase + (size_t)t * n;
    if (batch_fused(b)) {
      batch_launch_fused<R>(b, t, ph);
    } else if (b->kg2) {
      const wap::KG2Args<R> *a = (const wap::KG2Args<R> *)b->d_args[0];
      hipLaunchKernelGGL((wap::kg2_pass_batched<R, 1, K>), grid, block, 0, b->stream, a,
                         b->d_tile_begin, n, ph);
      hipLaunchKernelGGL((wap::kg2_pass_batched<R, 2, K>), grid, block, 0, b->stream, a,
                         b->d_tile_begin, n, ph);
    } else {
      const wap::PassArgs<R> *const *a = (const wap::PassArgs<R> *const *)b->d_args;
      hipLaunchKernelGGL((wap::rk4_pass_batched<R, EQ, 1, K>), grid, block, 0, b->stream, a[0],
                         b->d_tile_begin, n, ph);
      hipLaunchKernelGGL((wap::rk4_pass_batched<R, EQ, 2, K>), grid, block, 0, b->stream, a[1],
                         b->d_tile_begin, n, ph);
      hipLaunchKernelGGL((wap::rk4_pass_batched<R, EQ, 3, K>), grid, block, 0, b->stream, a[2],
                         b->d_tile_begin, n, ph);
      hipLaunchKernelGGL((wap::rk4_pass_batched<R, EQ, 4, K>), grid, block, 0, b->stream, a[3],
                         b->d_tile_begin, n, ph);
    }
  }
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

int batch_dispatch(wap_batch *b, int64_t nsteps, const void *ph) {
  const bool kg = b->equation == WAP_EQ_KLEIN_GORDON;
  switch (b->precision) {
    case WAP_PREC_F64:
      return kg ? batch_launch<double, 1, false>(b, nsteps, ph) : batch_launch<double, 0, false>(b, nsteps, ph);
    case WAP_PREC_F32:
      return kg ? batch_launch<float, 1, false>(b, nsteps, ph) : batch_launch<float, 0, false>(b, nsteps, ph);
    default:
      return kg ? batch_launch<float, 1, true>(b, nsteps, ph) : batch_launch<float, 0, true>(b, nsteps, ph);
  }
}

int batch_phase_dispatch(wap_batch *b, int64_t nsteps, int64_t step0, void **table, int64_t *cap) {
  return b->precision == WAP_PREC_F64 ? batch_phases<double>(b, nsteps, step0, table, cap)
                                      : batch_phases<float>(b, nsteps, step0, table, cap);
}

}  // namespace

extern "C" {

int wap_batch_create(wap_handle *const *handles, int n, int tile_y, wap_batch **out) {
  if (!out || !handles || n < 1) return fail(WAP_ERR_ARG, "need at least one handle");
  *out = nullptr;
  if (tile_y < 0 || tile_y > 16 || (tile_y & (tile_y - 1)))
    return fail(WAP_ERR_ARG, "tile_y must be 0 (auto) or 1, 2, 4, 8, 16");
  const wap_handle *h0 = handles[0];
  if (!h0) return fail(WAP_ERR_ARG, "NULL handle");
  for (int i = 0; i < n; ++i) {
    const wap_handle *h = handles[i];
    if (!h) return fail(WAP_ERR_ARG, "NULL handle");
    if (h->d.equation != h0->d.equation || h->d.precision != h0->d.precision ||
        sched_of(h) != sched_of(h0))
      return fail(WAP_ERR_ARG, "batched handles must share equation, precision and schedule");
    if (!h->v_set) return fail(WAP_ERR_STATE, "potential not set on a batched handle");
  }
  const int s0 = sched_of(h0);
  for (int i = 0; i < n; ++i)
    if (s0 == WAP_SCHED_SCHR_1PASS && is_f64(h0) && handles[i]->d.ny == 1)
      return fail(WAP_ERR_ARG, "batched float64 one-pass Schrodinger grids are 3D (float64 2D "
                               "grids: create the handles with the stage schedule)");
  const bool fused = s0 == WAP_SCHED_KG_FUSED || s0 == WAP_SCHED_SCHR_FUSED ||
                     s0 == WAP_SCHED_SCHR_1PASS;
  if (fused ? (tile_y != 0 && tile_y < 4) : tile_y > 8)
    return fail(WAP_ERR_ARG, fused ? "fused batches take tile_y 0 (auto), 4, 8 or 16"
                                   : "tile_y must be 0 (auto) or 1, 2, 4, 8");
  wap_batch *b = new (std::nothrow) wap_batch();
  if (!b) return fail(WAP_ERR_OOM, "host allocation failed");
  b->hs.assign(handles, handles + n);
  b->sched = s0;
  b->kg2 = kg_two_pass(h0) && !fused;
  b->waves = tile_y > 0 ? tile_y : kg1_waves(h0);
  if (is_f64(h0) && is_kg(h0) && b->waves > 8) b->waves = 8;  // as kg1_waves
  if (s0 == WAP_SCHED_SCHR_1PASS) b->waves = wap::kSchr1Rows + 1;  // one fixed shape
  b->precision = h0->d.precision;
  b->equation = h0->d.equation;
  b->tile_y = tile_y > 0 ? tile_y : h0->tile_y;
  b->stream = h0->stream;
  int rc = b->precision == WAP_PREC_F64 ? batch_build<double>(b) : batch_build<float>(b);
  if (rc) {
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return WAP_OK;
}

int wap_batch_destroy(wap_batch *b) {
  if (!b) return WAP_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  else (void)hipDeviceSynchronize();
  batch_free(b);
  delete b;
  return WAP_OK;
}

int wap_batch_set_stream(wap_batch *b, void *stream) {
  if (!b) return fail(WAP_ERR_ARG, "batch is NULL");
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  b->own_stream = false;
  b->stream = (hipStream_t)stream;
  return WAP_OK;
}

int wap_batch_tiles(wap_batch *b, int64_t *tiles, int32_t *tile_y) {
  if (!b) return fail(WAP_ERR_ARG, "batch is NULL");
  if (tiles) *tiles = b->total_tiles;
  if (tile_y) *tile_y = batch_fused(b) ? b->waves : b->tile_y;
  return WAP_OK;
}

int wap_batch_step(wap_batch *b, int64_t nsteps, int64_t step0) {
  if (!b) return fail(WAP_ERR_ARG, "batch is NULL");
  if (nsteps < 0) return fail(WAP_ERR_ARG, "nsteps < 0");
  if (nsteps == 0) return WAP_OK;
  int rc = batch_refresh(b);
  if (rc) return rc;
  rc = batch_phase_dispatch(b, nsteps, step0, &b->d_phase, &b->phase_steps);
  if (rc) return rc;
  rc = batch_dispatch(b, nsteps, b->d_phase);
  if (!rc) batch_flip(b, nsteps);
  return rc;
}

int wap_batch_capture(wap_batch *b, int64_t nsteps) {
  if (!b) return fail(WAP_ERR_ARG, "batch is NULL");
  if (nsteps < 1) return fail(WAP_ERR_ARG, "nsteps < 1");
  // a replay of the previous graph may still be running: let it end before its graph
  // and phase table go
  if (b->exec) WAP_HIP(hipStreamSynchronize(b->stream));
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  if (!b->stream) {
    // capture needs a stream of its own (default priority)
    WAP_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->own_stream = true;
  }
  // the graph's own phase table, sized here (the previous graph, its only reader, is
  // gone); its contents are written per replay
  int rc = batch_refresh(b);
  if (rc) return rc;
  if (b->phase_graph_steps != nsteps && b->d_phase_graph) {
    WAP_HIP(hipStreamSynchronize(b->stream));
    (void)hipFree(b->d_phase_graph);
    b->d_phase_graph = nullptr;
    b->phase_graph_steps = 0;
  }
  rc = batch_phase_dispatch(b, nsteps, 0, &b->d_phase_graph, &b->phase_graph_steps);
  if (rc) return rc;
  WAP_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  rc = batch_dispatch(b, nsteps, b->d_phase_graph);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(WAP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  b->graph = g;
  WAP_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_steps = nsteps;
  return WAP_OK;
}

int wap_batch_replay(wap_batch *b, int64_t step0, double *ms) {
  if (!b) return fail(WAP_ERR_ARG, "batch is NULL");
  if (!b->exec) return fail(WAP_ERR_STATE, "no captured graph (wap_batch_capture)");
  int rc = batch_refresh(b);
  if (rc) return rc;
  // graph_steps == phase_graph_steps: the table is rewritten in place, never regrown
  rc = batch_phase_dispatch(b, b->graph_steps, step0, &b->d_phase_graph, &b->phase_graph_steps);
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms) {
    WAP_HIP(hipEventCreate(&e0));
    WAP_HIP(hipEventCreate(&e1));
    WAP_HIP(hipEventRecord(e0, b->stream));
  }
  WAP_HIP(hipGraphLaunch(b->exec, b->stream));
  batch_flip(b, b->graph_steps);
  if (ms) {
    WAP_HIP(hipEventRecord(e1, b->stream));
    WAP_HIP(hipEventSynchronize(e1));
    float f = 0.f;
    WAP_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return WAP_OK;
}

int wap_batch_synchronize(wap_batch *b) {
  if (!b) return fail(WAP_ERR_ARG, "batch is NULL");
  WAP_HIP(hipStreamSynchronize(b->stream));
  return WAP_OK;
}

}  // extern "C"

// ------------------------------------------------------------- batch sets
// One hipGraph over several batches (one per equation / precision / schedule group): the
// graph forks to one stream per batch, runs each batch's steps there and joins them, so
// the groups' launches overlap where the CUs allow.  The set owns a phase table per batch
// (sized at capture), separate from the batches' own eager and graph tables.
struct wap_batch_set {
  std::vector<wap_batch *> bs;
  std::vector<hipStream_t> streams;  // [0] the launch stream; one per batch
  std::vector<hipEvent_t> events;    // [0] the fork, [i > 0] batch i's join
  hipEvent_t done = nullptr;         // the end of a replay, which the handles' streams wait for
  std::vector<hipStream_t> handle_streams;  // the distinct streams of the batches' handles
  std::vector<void *> phase;         // per batch: [steps][ninst] phasors
  std::vector<int64_t> phase_steps;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t graph_steps = 0;
};

namespace {

void batch_set_free(wap_batch_set *s) {
  if (!s->streams.empty() && s->streams[0]) (void)hipStreamSynchronize(s->streams[0]);
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->exec = nullptr;
  s->graph = nullptr;
  for (void *&p : s->phase)
    if (p) { (void)hipFree(p); p = nullptr; }
  for (hipStream_t &st : s->streams)
    if (st) { (void)hipStreamDestroy(st); st = nullptr; }
  for (hipEvent_t &ev : s->events)
    if (ev) { (void)hipEventDestroy(ev); ev = nullptr; }
  if (s->done) (void)hipEventDestroy(s->done);
  s->done = nullptr;
}

// The steps of every batch, each on its stream, forked from and joined into streams[0]
// (called inside stream capture).
int batch_set_record(wap_batch_set *s, int64_t nsteps) {
  const size_t n = s->bs.size();
  const hipEvent_t fork = s->events[0];
  int rc = WAP_OK;
  hipError_t e = hipEventRecord(fork, s->streams[0]);
  for (size_t i = 0; i < n && e == hipSuccess && !rc; ++i) {
    wap_batch *b = s->bs[i];
    if (i > 0) e = hipStreamWaitEvent(s->streams[i], fork, 0);
    if (e != hipSuccess) break;
    hipStream_t own = b->stream;
    b->stream = s->streams[i];  // the batch's launches go to the set's stream
    rc = batch_dispatch(b, nsteps, s->phase[i]);
    b->stream = own;
    if (!rc && i > 0) e = hipEventRecord(s->events[i], s->streams[i]);
  }
  for (size_t i = 1; i < n && e == hipSuccess && !rc; ++i)
    e = hipStreamWaitEvent(s->streams[0], s->events[i], 0);
  if (rc) return rc;
  if (e != hipSuccess) return fail(WAP_ERR_HIP, std::string("batch set capture: ") + hipGetErrorString(e));
  return WAP_OK;
}

}  // namespace

extern "C" {

int wap_batch_set_create(wap_batch *const *batches, int n, wap_batch_set **out) {
  if (!out || !batches || n < 1) return fail(WAP_ERR_ARG, "need at least one batch");
  *out = nullptr;
  for (int i = 0; i < n; ++i)
    if (!batches[i]) return fail(WAP_ERR_ARG, "NULL batch");
  wap_batch_set *s = new (std::nothrow) wap_batch_set();
  if (!s) return fail(WAP_ERR_OOM, "host allocation failed");
  s->bs.assign(batches, batches + n);
  s->streams.assign(n, nullptr);
  s->events.assign(n, nullptr);
  s->phase.assign(n, nullptr);
  s->phase_steps.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    // default priority, non-blocking (capture needs streams of its own)
    hipError_t e = hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->events[i], hipEventDisableTiming);
    if (e == hipSuccess && i == 0) e = hipEventCreateWithFlags(&s->done, hipEventDisableTiming);
    if (e != hipSuccess) {
      batch_set_free(s);
      delete s;
      return fail(WAP_ERR_HIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
    }
  }
  for (wap_batch *b : s->bs)
    for (wap_handle *h : b->hs)
      if (std::find(s->handle_streams.begin(), s->handle_streams.end(), h->stream) ==
          s->handle_streams.end())
        s->handle_streams.push_back(h->stream);
  *out = s;
  return WAP_OK;
}

int wap_batch_set_destroy(wap_batch_set *s) {
  if (!s) return WAP_OK;
  batch_set_free(s);
  delete s;
  return WAP_OK;
}

int wap_batch_set_capture(wap_batch_set *s, int64_t nsteps) {
  if (!s) return fail(WAP_ERR_ARG, "batch set is NULL");
  if (nsteps < 1) return fail(WAP_ERR_ARG, "nsteps < 1");
  WAP_HIP(hipStreamSynchronize(s->streams[0]));  // a replay of the previous graph
  if (s->exec) { (void)hipGraphExecDestroy(s->exec); s->exec = nullptr; }
  if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
  for (size_t i = 0; i < s->bs.size(); ++i) {
    wap_batch *b = s->bs[i];
    // eager work of the batch (its own stream) ends before its arguments are rewritten
    WAP_HIP(hipStreamSynchronize(b->stream));
    int rc = batch_refresh(b);
    if (rc) return rc;
    if (s->phase_steps[i] != nsteps && s->phase[i]) {
      (void)hipFree(s->phase[i]);
      s->phase[i] = nullptr;
      s->phase_steps[i] = 0;
    }
    rc = batch_phase_dispatch(b, nsteps, 0, &s->phase[i], &s->phase_steps[i]);
    if (rc) return rc;
  }
  WAP_HIP(hipStreamBeginCapture(s->streams[0], hipStreamCaptureModeThreadLocal));
  int rc = batch_set_record(s, nsteps);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s->streams[0], &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(WAP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  s->graph = g;
  WAP_HIP(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
  s->graph_steps = nsteps;
  return WAP_OK;
}

int wap_batch_set_replay(wap_batch_set *s, int64_t step0, double *ms) {
  if (!s) return fail(WAP_ERR_ARG, "batch set is NULL");
  if (!s->exec) return fail(WAP_ERR_STATE, "no captured graph (wap_batch_set_capture)");
  WAP_HIP(hipStreamSynchronize(s->streams[0]));  // the previous replay reads the tables
  // earlier work on the handles (eager steps, state uploads) ends before the graph starts
  for (hipStream_t hs : s->handle_streams) WAP_HIP(hipStreamSynchronize(hs));
  for (size_t i = 0; i < s->bs.size(); ++i) {
    wap_batch *b = s->bs[i];
    WAP_HIP(hipStreamSynchronize(b->stream));
    int rc = batch_refresh(b);
    if (rc) return rc;
    // phase_steps == graph_steps: rewritten in place, never regrown
    rc = batch_phase_dispatch(b, s->graph_steps, step0, &s->phase[i], &s->phase_steps[i]);
    if (rc) return rc;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms) {
    WAP_HIP(hipEventCreate(&e0));
    WAP_HIP(hipEventCreate(&e1));
    WAP_HIP(hipEventRecord(e0, s->streams[0]));
  }
  WAP_HIP(hipGraphLaunch(s->exec, s->streams[0]));
  // later work on the handles' streams waits for the graph
  WAP_HIP(hipEventRecord(s->done, s->streams[0]));
  for (hipStream_t hs : s->handle_streams) WAP_HIP(hipStreamWaitEvent(hs, s->done, 0));
  for (wap_batch *b : s->bs) batch_flip(b, s->graph_steps);
  if (ms) {
    WAP_HIP(hipEventRecord(e1, s->streams[0]));
    WAP_HIP(hipEventSynchronize(e1));
    float f = 0.f;
    WAP_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return WAP_OK;
}

int wap_batch_set_synchronize(wap_batch_set *s) {
  if (!s) return fail(WAP_ERR_ARG, "batch set is NULL");
  WAP_HIP(hipStreamSynchronize(s->streams[0]));
  return WAP_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ adjoint
struct wap_adjoint {
  wap_handle *h = nullptr;
  int slots = 0;
  int ncomp = 1;
  // Klein-Gordon: one fused reverse pass per step from the step's input state
  bool fused = false;
  std::vector<void *> st;        // [slot][comp]: the input state u_n of the recorded step
  std::vector<int64_t> st_step;  // its step index
  std::vector<double> st_t0;
  int cur = 0;                   // lam[cur] holds dL/du of the latest step
  // Schrodinger: the 4-stage records and the stage-by-stage reverse passes
  std::vector<void *> rec;   // [slot][stage 0..3][comp]
  void *lam2[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [buffer][comp]
  void *acc[2] = {nullptr, nullptr};
  void *kba[2] = {nullptr, nullptr};
  void *kbb[2] = {nullptr, nullptr};
  double *vbar = nullptr;
  void *vacc = nullptr;  // fused path: dL/dV accumulated in the handle's real type,
                         // folded into vbar every kGradFlush reverse steps
  int since_flush = 0;
  bool schr = false;     // fused Schrodinger (one state component; with two-pass records u3,
                         // v3, S in kba, kbb, acc)
  bool owns = true;      // false: the buffers are carved from a caller's workspace
  void *orig[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [comp]: the handle's
                                                                  // state, scratch at create
  void *lam(int c) const { return lam2[cur][c]; }
  // fused Klein-Gordon: lam2[b][0] holds LamPair (lam_psi, lam_pi) per voxel, lam2[b][1] is
  // unused
  bool pairs() const { return fused && !schr; }
};

namespace {

// The device buffers of an adjoint and their sizes (st / rec already sized).
std::vector<std::pair<void **, size_t>> adjoint_buffers(wap_adjoint *a) {
  const wap_handle *h = a->h;
  const size_t fb = h->n * h->cs;
  std::vector<std::pair<void **, size_t>> v;
  if (a->fused) {
    for (void *&b : a->st) v.push_back({&b, fb});
    if (a->pairs()) {
      v.push_back({&a->lam2[0][0], 2 * fb});
      v.push_back({&a->lam2[1][0], 2 * fb});
    } else {
      v.push_back({&a->lam2[0][0], fb});
      v.push_back({&a->lam2[1][0], fb});
    }
    if (a->schr && sched_of(h) == WAP_SCHED_SCHR_FUSED) {  // schr_adj_pass I's u3, v3 and S
      v.push_back({&a->kba[0], fb});
      v.push_back({&a->kbb[0], fb});
      v.push_back({&a->acc[0], fb});
    }
    v.push_back({&a->vacc, h->n * h->rs});
  } else {
    for (void *&b : a->rec) v.push_back({&b, fb});
    for (int c = 0; c < a->ncomp; ++c) {
      v.push_back({&a->lam2[0][c], fb});
      v.push_back({&a->acc[c], fb});
      v.push_back({&a->kba[c], fb});
      v.push_back({&a->kbb[c], fb});
    }
  }
  v.push_back({(void **)&a->vbar, h->n * sizeof(double)});
  return v;
}

constexpr size_t kWsAlign = 256;
size_t ws_round(size_t b) { return (b + kWsAlign - 1) / kWsAlign * kWsAlign; }

// Give the handle back its own state buffers.  A fused record swaps the step's output
// buffer with a slot, so after records the handle may hold record buffers and the slots
// the handle's; the state is copied when it sits in a record buffer.  A no-op when
// nothing was swapped.
int adjoint_restore(wap_adjoint *a) {
  if (!a->fused) return WAP_OK;
  wap_handle *h = a->h;
  // the state buffer and the one the fused step's output swaps with the slots
  void **cur[2] = {&h->psi, &h->pi};
  void **ya[2] = {a->schr ? &h->yb_psi : &h->ya_psi, &h->ya_pi};
  const int nc = a->ncomp;
  for (int c = 0; c < nc; ++c) {
    void *o1 = a->orig[c][0], *o2 = a->orig[c][1];
    auto where = [&](void *p) -> void ** {
      if (*ya[c] == p) return ya[c];
      for (int s = 0; s < a->slots; ++s)
        if (a->st[(size_t)s * nc + c] == p) return &a->st[(size_t)s * nc + c];
      return nullptr;
    };
    if (*cur[c] != o1 && *cur[c] != o2) {
      void **x = where(o1);
      if (!x) return fail(WAP_ERR_STATE, "adjoint: a state buffer of the handle is missing");
      WAP_HIP(hipMemcpyAsync(o1, *cur[c], h->n * h->cs, hipMemcpyDeviceToDevice, h->stream));
      *x = *cur[c];
      *cur[c] = o1;
    }
    void *other = (*cur[c] == o1) ? o2 : o1;
    if (*ya[c] != other) {
      void **x = where(other);
      if (!x || x == ya[c]) return fail(WAP_ERR_STATE, "adjoint: a state buffer of the handle is missing");
      std::swap(*x, *ya[c]);
    }
  }
  return WAP_OK;
}

void adjoint_free(wap_adjoint *a) {
  for (auto &b : adjoint_buffers(a)) {
    if (*b.first && a->owns) (void)hipFree(*b.first);
    *b.first = nullptr;
  }
}

void *rec_ptr(const wap_adjoint *a, int slot, int stage, int comp) {
  return a->rec[((size_t)slot * 4 + stage) * a->ncomp + comp];
}

template <typename R, int EQ>
int adjoint_record_t(wap_adjoint *ad, int64_t step, double t0, int slot) {
  using C = wap::cplx<R>;
  wap_handle *h = ad->h;
  const dim3 block(wap::kWaveX, h->tile_y);
  const dim3 grid((unsigned)(h->ntx * h->nty * h->ntz));
  wap::PassArgs<R> a[4];
  for (int s = 0; s < 4; ++s) a[s] = stage_args<R>(h, s + 1, h->tile_y);
  const int q = EQ == 1 ? 1 : 0;
  // P1: state -> Y2 (records Y1 = injected state)
  a[0].out_psi = (C *)rec_ptr(ad, slot, 1, 0);
  a[0].out_pi = (C *)rec_ptr(ad, slot, 1, q);
  a[0].rec_psi = (C *)rec_ptr(ad, slot, 0, 0);
  a[0].rec_pi = (C *)rec_ptr(ad, slot, 0, q);
  // P2: Y2 -> Y3, P3: Y3 -> Y4, P4: Y4 -> state
  for (int s = 1; s < 4; ++s) {
    a[s].yin_psi = (const C *)rec_ptr(ad, slot, s, 0);
    a[s].yin_pi = (const C *)rec_ptr(ad, slot, s, q);
    if (s < 3) {
      a[s].out_psi = (C *)rec_ptr(ad, slot, s + 1, 0);
      a[s].out_pi = (C *)rec_ptr(ad, slot, s + 1, q);
    }
  }
  if (EQ == 0)
    for (int s = 0; s < 4; ++s) { a[s].yin_pi = nullptr; a[s].out_pi = nullptr; a[s].rec_pi = nullptr; }
  const double t = step_time(h, step, 0, t0);
  for (int s = 0; s < 4; ++s) fill_source(h, t, a[s]);
  hipLaunchKernelGGL((wap::rk4_pass<R, EQ, 1, false>), grid, block, 0, h->stream, a[0]);
  hipLaunchKernelGGL((wap::rk4_pass<R, EQ, 2, false>), grid, block, 0, h->stream, a[1]);
  hipLaunchKernelGGL((wap::rk4_pass<R, EQ, 3, false>), grid, block, 0, h->stream, a[2]);
  hipLaunchKernelGGL((wap::rk4_pass<R, EQ, 4, false>), grid, block, 0, h->stream, a[3]);
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

template <typename R, int EQ>
int adjoint_step_t(wap_adjoint *ad, int slot) {
  using C = wap::cplx<R>;
  wap_handle *h = ad->h;
  const dim3 block(wap::kWaveX, h->tile_y);
  const dim3 grid((unsigned)(h->ntx * h->nty * h->ntz));
  wap::AdjArgs<R> a;
  std::memset((void *)&a, 0, sizeof(a));
  a.lam_psi = (C *)ad->lam(0);
  a.lam_pi = (C *)ad->lam(1);
  a.acc_psi = (C *)ad->acc[0];
  a.acc_pi = (C *)ad->acc[1];
  a.vbar = ad->vbar;
  a.V = (const R *)h->V;
  a.zc = (const C *)h->zc;
  a.gam = (const R *)h->gam;
  a.nx = h->d.nx; a.ny = h->d.ny; a.nz = h->d.nz;
  a.tile_y = h->tile_y; a.tile_z = h->tile_z;
  a.ntx = h->ntx; a.nty = h->nty; a.ntz = h->ntz;
  a.hx = (R)h->hx; a.hy = (R)h->hy; a.m2 = (R)h->m2; a.h = (R)h->d.dt;
  a.src_k0 = h->src_k0;
  a.src_n = h->nsrc;
  a.src_tab = (const wap::SrcCoef<R> *)h->src_tab;
  const int q = EQ == 1 ? 1 : 0;
  void *const *kin[4] = {nullptr, ad->kba, ad->kbb, ad->kba};   // by stage S = 4, 3, 2, 1
  void *const *kout[4] = {ad->kba, ad->kbb, ad->kba, nullptr};
  for (int r = 0; r < 4; ++r) {
    const int S = 4 - r;
    a.y_psi = (const C *)rec_ptr(ad, slot, S - 1, 0);
    a.y_pi = (const C *)rec_ptr(ad, slot, S - 1, q);
    a.kb_in_psi = kin[r] ? (const C *)kin[r][0] : nullptr;
    a.kb_in_pi = kin[r] ? (const C *)kin[r][1] : nullptr;
    a.kb_out_psi = kout[r] ? (C *)kout[r][0] : nullptr;
    a.kb_out_pi = kout[r] ? (C *)kout[r][1] : nullptr;
    switch (S) {
      case 4: hipLaunchKernelGGL((wap::adj_pass<R, EQ, 4>), grid, block, 0, h->stream, a); break;
      case 3: hipLaunchKernelGGL((wap::adj_pass<R, EQ, 3>), grid, block, 0, h->stream, a); break;
      case 2: hipLaunchKernelGGL((wap::adj_pass<R, EQ, 2>), grid, block, 0, h->stream, a); break;
      default: hipLaunchKernelGGL((wap::adj_pass<R, EQ, 1>), grid, block, 0, h->stream, a); break;
    }
  }
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

// Fused Klein-Gordon adjoint workgroups: the handle's tile_y (16 or 8 waves; 16 by
// default), at most 8 for double (its LDS planes are twice as large)
template <typename R>
int kg1_adj_waves(const wap_handle *h) {
  const int w = kg1_waves(h) >= 16 ? 16 : 8;
  return sizeof(R) == 4 ? w : 8;
}

// Record step `step`: keep its input state in `slot`, then advance the handle.  The
// fused forward pass ping-pongs its state buffers, so the input state is already in
// the handle's spare buffer afterwards and changes owner with the slot's (no copy).
int adjoint_record_fused(wap_adjoint *ad, int64_t step, double t0, int slot) {
  wap_handle *h = ad->h;
  void **sp = &ad->st[(size_t)slot * ad->ncomp];
  int rc;
  if (ad->schr) {
    // the fused steps leave psi_n in the buffer their output swapped with (yb)
    rc = sched_of(h) == WAP_SCHED_SCHR_1PASS
             ? run_steps_schr1<float>(h, 1, step, t0, nullptr, nullptr, true)
             : run_steps_schr2<float, false>(h, 1, step, t0, nullptr, nullptr);
    if (rc) return rc;
    std::swap(h->yb_psi, sp[0]);
  } else if (sched_of(h) == WAP_SCHED_KG_FUSED) {
    rc = is_f64(h) ? run_steps_kg1<double>(h, 1, step, t0, nullptr, nullptr, true)
                   : run_steps_kg1<float>(h, 1, step, t0, nullptr, nullptr, true);
    if (rc) return rc;
    std::swap(h->ya_psi, sp[0]);
    std::swap(h->ya_pi, sp[1]);
  } else {
    const size_t fb = h->n * h->cs;
    WAP_HIP(hipMemcpyAsync(sp[0], h->psi, fb, hipMemcpyDeviceToDevice, h->stream));
    WAP_HIP(hipMemcpyAsync(sp[1], h->pi, fb, hipMemcpyDeviceToDevice, h->stream));
    rc = dispatch_steps(h, 1, step, t0, nullptr, nullptr);
    if (rc) return rc;
  }
  ad->st_step[slot] = step;
  ad->st_t0[slot] = t0;
  return WAP_OK;
}

// both passes of one Schrodinger reverse step (NT: non-temporal result stores)
template <typename R, int ROWS, bool NT>
void launch_schr_adj(dim3 grid, dim3 block, hipStream_t s, const wap::SchrAdjArgs<R> &a) {
  hipLaunchKernelGGL((wap::schr_adj_pass<R, ROWS, 1, NT>), grid, block, 0, s, a);
  hipLaunchKernelGGL((wap::schr_adj_pass<R, ROWS, 2, NT>), grid, block, 0, s, a);
}

template <typename R>
int adjoint_step_fused(wap_adjoint *ad, int slot) {
  using C = wap::cplx<R>;
  wap_handle *h = ad->h;
  const int waves = kg1_adj_waves<R>(h);
  const int ntx = kg1_ntx(h), nty = kg1_nty(h, waves);
  const int tz = kg1_tile_z(h, waves, 16 / waves);
  const int ntz = (h->d.nz + tz - 1) / tz;
  wap::KG1AdjArgs<R> a;
  std::memset((void *)&a, 0, sizeof(a));
  a.u_psi = (const C *)ad->st[(size_t)slot * 2];
  a.u_pi = (const C *)ad->st[(size_t)slot * 2 + 1];
  a.lam_in = (const wap::LamPair<R> *)ad->lam2[ad->cur][0];
  a.lam_out = (wap::LamPair<R> *)ad->lam2[ad->cur ^ 1][0];
  a.vbar = (R *)ad->vacc;
  a.V = (const R *)h->V;
  a.gam = (const R *)h->gam;
  a.prof = (const C *)h->prof;
  a.nx = h->d.nx; a.ny = h->d.ny; a.nz = h->d.nz;
  a.tile_z = tz;
  a.ntx = ntx; a.nty = nty; a.ntz = ntz;
  a.hx = (R)h->hx; a.hy = (R)h->hy; a.hz = (R)kg_hz(h);
  split_hi_lo(2.0 * ((double)a.hx + (double)a.hy + (double)a.hz) + h->m2, a.d0_hi, a.d0_lo);
  a.h = (R)h->d.dt;
  a.gz_lo = h->gz_lo;
  a.gz_hi = h->gz_hi;
  a.stream_out = stream_stores(h->n, h->cs) ? 1 : 0;
  fill_source(h, step_time(h, ad->st_step[slot], 0, ad->st_t0[slot]), a);
  const dim3 block(wap::kWaveX, waves);
  const dim3 grid((unsigned)(ntx * nty * ntz));
  constexpr int kRows = sizeof(R) == 4 ? 15 : 7;
  if (sizeof(R) == 4 && waves == 16) {
    if (a.stream_out)
      hipLaunchKernelGGL((wap::kg1_adj_pass<R, kRows, true>), grid, block, 0, h->stream, a);
    else
      hipLaunchKernelGGL((wap::kg1_adj_pass<R, kRows, false>), grid, block, 0, h->stream, a);
  } else {
    if (a.stream_out)
      hipLaunchKernelGGL((wap::kg1_adj_pass<R, 7, true>), grid, block, 0, h->stream, a);
    else
      hipLaunchKernelGGL((wap::kg1_adj_pass<R, 7, false>), grid, block, 0, h->stream, a);
  }
  WAP_HIP(hipGetLastError());
  ad->cur ^= 1;
  return WAP_OK;
}

// The fused adjoints add each step's dL/dV into vacc (the handle's real type); every
// kGradFlush steps vacc is added into the float64 vbar and cleared, so a float32
// accumulator never sums more than that many steps (a run of 1500 reverse steps would
// otherwise carry 1500 float32 roundings of a growing partial sum per voxel).  32 steps
// bound the partial sum's rounding near 32 ulp while a flush (read vacc, read and write
// the float64 vbar, clear vacc: 24 B per voxel) costs under 1 % of the reverse sweep.
constexpr int kGradFlush = 32;

int adjoint_flush(wap_adjoint *a) {
  if (!a->vacc) return WAP_OK;
  const size_t n = a->h->n;
  const dim3 grid(grid_1d(n)), block(256);
  if (is_f64(a->h))
    hipLaunchKernelGGL(wap::flush_grad<double>, grid, block, 0, a->h->stream,
                       (double *)a->vacc, a->vbar, n);
  else
    hipLaunchKernelGGL(wap::flush_grad<float>, grid, block, 0, a->h->stream,
                       (float *)a->vacc, a->vbar, n);
  WAP_HIP(hipGetLastError());
  a->since_flush = 0;
  return WAP_OK;
}

int adjoint_count_step(wap_adjoint *a) {
  return ++a->since_flush >= kGradFlush ? adjoint_flush(a) : WAP_OK;
}

// Reverse of one recorded one-pass Schrodinger step: schr_adj1_pass, the tiles of
// schr1_pass (one workgroup per CU: 135 KB of LDS)
int adjoint_step_schr1(wap_adjoint *ad, int slot) {
  using R = float;
  using C = wap::cplx<R>;
  wap_handle *h = ad->h;
  wap::SchrAdj1Args<R> a;
  std::memset((void *)&a, 0, sizeof(a));
  const int tz = schr1_tile_z(h);
  a.rec = (const C *)ad->st[slot];
  a.lam = (const C *)ad->lam2[ad->cur][0];
  a.lam_out = (C *)ad->lam2[ad->cur ^ 1][0];
  a.vbar = (R *)ad->vacc;
  a.V = (const R *)h->V;
  a.zc = (const C *)h->zc;
  a.prof = (const C *)h->prof;
  a.nx = h->d.nx; a.ny = h->d.ny; a.nz = h->d.nz;
  a.tile_z = tz;
  a.ntx = kg1_ntx(h); a.nty = schr1_nty(h); a.ntz = (h->d.nz + tz - 1) / tz;
  a.hx = (R)h->hx; a.hy = (R)h->hy;
  a.h = (R)h->d.dt;
  a.pz_lo = h->pz_lo;
  a.pz_hi = h->pz_hi;
  a.hz = (R)h->pz_w;
  a.stream_out = stream_stores(h->n, h->cs) ? 1 : 0;
  fill_source(h, step_time(h, ad->st_step[slot], 0, ad->st_t0[slot]), a);
  const dim3 block(wap::kWaveX, wap::kSchr1Rows + 1);
  const dim3 grid((unsigned)schr1_tiles(h));
  if (a.stream_out)
    hipLaunchKernelGGL((wap::schr_adj1_pass<R, wap::kSchr1Rows, true>), grid, block, 0, h->stream, a);
  else
    hipLaunchKernelGGL((wap::schr_adj1_pass<R, wap::kSchr1Rows, false>), grid, block, 0, h->stream, a);
  WAP_HIP(hipGetLastError());
  ad->cur ^= 1;
  return WAP_OK;
}

// Reverse of one recorded Schrodinger 2-pass step: schr_adj_pass I then II.
int adjoint_step_schr(wap_adjoint *ad, int slot) {
  using R = float;
  using C = wap::cplx<R>;
  wap_handle *h = ad->h;
  if (sched_of(h) == WAP_SCHED_SCHR_1PASS) return adjoint_step_schr1(ad, slot);
  const int waves = kg1_waves(h);
  const int tz = kg1_tile_z(h, waves, 16 / waves);
  wap::SchrAdjArgs<R> a;
  std::memset((void *)&a, 0, sizeof(a));
  a.rec = (const C *)ad->st[slot];
  a.lam = (const C *)ad->lam2[ad->cur][0];
  a.u3 = (C *)ad->kba[0];
  a.v3 = (C *)ad->kbb[0];
  a.sacc = (C *)ad->acc[0];
  a.lam_out = (C *)ad->lam2[ad->cur ^ 1][0];
  a.vbar = (R *)ad->vacc;
  a.V = (const R *)h->V;
  a.zc = (const C *)h->zc;
  a.prof = (const C *)h->prof;
  a.nx = h->d.nx; a.ny = h->d.ny; a.nz = h->d.nz;
  a.tile_z = tz;
  a.ntx = kg1_ntx(h); a.nty = kg1_nty(h, waves); a.ntz = (h->d.nz + tz - 1) / tz;
  a.hx = (R)h->hx; a.hy = (R)h->hy;
  a.h = (R)h->d.dt;
  a.stream_out = stream_stores(h->n, h->cs) ? 1 : 0;
  fill_source(h, step_time(h, ad->st_step[slot], 0, ad->st_t0[slot]), a);
  const dim3 block(wap::kWaveX, waves);
  const dim3 grid((unsigned)fused_tiles(h, waves));
  const bool nt = a.stream_out != 0;
  switch (waves) {
    case 4:
      nt ? launch_schr_adj<R, 3, true>(grid, block, h->stream, a)
         : launch_schr_adj<R, 3, false>(grid, block, h->stream, a);
      break;
    case 8:
      nt ? launch_schr_adj<R, 7, true>(grid, block, h->stream, a)
         : launch_schr_adj<R, 7, false>(grid, block, h->stream, a);
      break;
    default:
      nt ? launch_schr_adj<R, 15, true>(grid, block, h->stream, a)
         : launch_schr_adj<R, 15, false>(grid, block, h->stream, a);
      break;
  }
  WAP_HIP(hipGetLastError());
  ad->cur ^= 1;
  return WAP_OK;
}

// validation and the host object of an adjoint (no device memory yet)
int adjoint_new(wap_handle *h, int slots, wap_adjoint **out) {
  if (!h || !out) return fail(WAP_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (slots < 1) return fail(WAP_ERR_ARG, "slots must be >= 1");
  if (is_kahan(h)) return fail(WAP_ERR_ARG, "the adjoint supports f64 and f32 handles");
  if (h->d.end_planes != WAP_ENDS_FROZEN)
    return fail(WAP_ERR_ARG, "the adjoint supports frozen end planes");
  wap_adjoint *a = new (std::nothrow) wap_adjoint();
  if (!a) return fail(WAP_ERR_OOM, "host allocation failed");
  a->h = h;
  a->slots = slots;
  a->ncomp = is_kg(h) ? 2 : 1;
  // the fused Schrodinger adjoint is float32 (float64 handles reverse stage by stage)
  a->schr = !is_kg(h) && h->d.precision == WAP_PREC_F32 &&
            (sched_of(h) == WAP_SCHED_SCHR_FUSED || sched_of(h) == WAP_SCHED_SCHR_1PASS);
  a->fused = (is_kg(h) && fused_plane_fits(h->d)) || a->schr;
  if (a->fused) {
    a->st.assign((size_t)slots * a->ncomp, nullptr);
    a->st_step.assign(slots, 0);
    a->st_t0.assign(slots, 0.0);
    a->orig[0][0] = h->psi;
    a->orig[0][1] = a->schr ? h->yb_psi : h->ya_psi;
    a->orig[1][0] = h->pi;
    a->orig[1][1] = h->ya_pi;
  } else {
    a->rec.assign((size_t)slots * 4 * a->ncomp, nullptr);
  }
  *out = a;
  return WAP_OK;
}

}  // namespace

extern "C" {

int wap_adjoint_create(wap_handle *h, int slots, wap_adjoint **out) {
  wap_adjoint *a = nullptr;
  int rc = adjoint_new(h, slots, &a);
  if (rc) return rc;
  for (auto &b : adjoint_buffers(a)) {
    hipError_t e = hipMalloc(b.first, b.second);
    if (e != hipSuccess) {
      rc = fail(WAP_ERR_OOM, std::string("adjoint: ") + hipGetErrorString(e));
      break;
    }
  }
  if (rc) {
    adjoint_free(a);
    delete a;
    return rc;
  }
  *out = a;
  return wap_adjoint_reset(a);
}

int wap_adjoint_workspace_bytes(wap_handle *h, int slots, int64_t *bytes) {
  if (!bytes) return fail(WAP_ERR_ARG, "NULL argument");
  wap_adjoint *a = nullptr;
  int rc = adjoint_new(h, slots, &a);
  if (rc) return rc;
  size_t total = 0;
  for (auto &b : adjoint_buffers(a)) total += ws_round(b.second);
  delete a;
  *bytes = (int64_t)total;
  return WAP_OK;
}

int wap_adjoint_create_in(wap_handle *h, int slots, void *workspace, int64_t bytes,
                          wap_adjoint **out) {
  if (!workspace) return fail(WAP_ERR_ARG, "NULL workspace");
  if ((uintptr_t)workspace % kWsAlign) return fail(WAP_ERR_ARG, "workspace must be 256-byte aligned");
  wap_adjoint *a = nullptr;
  int rc = adjoint_new(h, slots, &a);
  if (rc) return rc;
  a->owns = false;
  auto bufs = adjoint_buffers(a);
  size_t total = 0;
  for (auto &b : bufs) total += ws_round(b.second);
  if (bytes < 0 || (size_t)bytes < total) {
    delete a;
    return fail(WAP_ERR_ARG, "workspace smaller than wap_adjoint_workspace_bytes");
  }
  char *p = (char *)workspace;
  for (auto &b : bufs) {
    *b.first = p;
    p += ws_round(b.second);
  }
  *out = a;
  return wap_adjoint_reset(a);
}

int wap_adjoint_destroy(wap_adjoint *a) {
  if (!a) return WAP_OK;
  const int rc = adjoint_restore(a);
  (void)hipStreamSynchronize(a->h->stream);
  if (rc) return rc;  // the handle still holds record buffers: keep them allocated
  adjoint_free(a);
  delete a;
  return WAP_OK;
}

int wap_adjoint_reset(wap_adjoint *a) {
  if (!a) return fail(WAP_ERR_ARG, "adjoint is NULL");
  const size_t fb = a->h->n * a->h->cs;
  if (a->pairs())
    WAP_HIP(hipMemsetAsync(a->lam(0), 0, 2 * fb, a->h->stream));
  else
    for (int c = 0; c < a->ncomp; ++c) WAP_HIP(hipMemsetAsync(a->lam(c), 0, fb, a->h->stream));
  WAP_HIP(hipMemsetAsync(a->vbar, 0, a->h->n * sizeof(double), a->h->stream));
  if (a->vacc) WAP_HIP(hipMemsetAsync(a->vacc, 0, a->h->n * a->h->rs, a->h->stream));
  a->since_flush = 0;
  return WAP_OK;
}

int wap_adjoint_record(wap_adjoint *a, int64_t step, double t0, int slot) {
  if (!a) return fail(WAP_ERR_ARG, "adjoint is NULL");
  if (slot < 0 || slot >= a->slots) return fail(WAP_ERR_ARG, "slot out of range");
  if (!a->h->v_set) return fail(WAP_ERR_STATE, "potential not set (wap_set_potential)");
  if (int rc = check_t0(a->h, t0)) return rc;
  if (a->fused) return adjoint_record_fused(a, step, t0, slot);
  const bool kg = is_kg(a->h);
  if (is_f64(a->h))
    return kg ? adjoint_record_t<double, 1>(a, step, t0, slot) : adjoint_record_t<double, 0>(a, step, t0, slot);
  return kg ? adjoint_record_t<float, 1>(a, step, t0, slot) : adjoint_record_t<float, 0>(a, step, t0, slot);
}

int wap_adjoint_step(wap_adjoint *a, int slot) {
  if (!a) return fail(WAP_ERR_ARG, "adjoint is NULL");
  if (slot < 0 || slot >= a->slots) return fail(WAP_ERR_ARG, "slot out of range");
  if (a->schr) {
    const int rc = adjoint_step_schr(a, slot);
    return rc ? rc : adjoint_count_step(a);
  }
  if (a->fused) {
    const int rc = is_f64(a->h) ? adjoint_step_fused<double>(a, slot)
                                : adjoint_step_fused<float>(a, slot);
    return rc ? rc : adjoint_count_step(a);
  }
  const bool kg = is_kg(a->h);
  if (is_f64(a->h)) return kg ? adjoint_step_t<double, 1>(a, slot) : adjoint_step_t<double, 0>(a, slot);
  return kg ? adjoint_step_t<float, 1>(a, slot) : adjoint_step_t<float, 0>(a, slot);
}

int wap_adjoint_seed(wap_adjoint *a, int k, const void *g, double c_re, double c_im) {
  if (!a || !g) return fail(WAP_ERR_ARG, "NULL argument");
  wap_handle *h = a->h;
  if (k < 0 || k >= h->d.nz) return fail(WAP_ERR_ARG, "plane index out of range");
  const int n = (int)h->plane;
  const dim3 grid((n + 255) / 256), block(256);
  const size_t off = (size_t)k * h->plane;
  const int stride = a->pairs() ? 2 : 1;
  if (is_f64(h))
    hipLaunchKernelGGL(wap::seed_plane<double>, grid, block, 0, h->stream, (double2 *)a->lam(0),
                       off, n, stride, (const double2 *)g, c_re, c_im);
  else
    hipLaunchKernelGGL(wap::seed_plane<float>, grid, block, 0, h->stream, (float2 *)a->lam(0), off,
                       n, stride, (const double2 *)g, c_re, c_im);
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}

int wap_adjoint_get_grad(wap_adjoint *a, void *out, int on_device) {
  if (!a || !out) return fail(WAP_ERR_ARG, "NULL argument");
  if (int rc = adjoint_flush(a)) return rc;  // the fused path's unflushed steps
  WAP_HIP(hipMemcpyAsync(out, a->vbar, a->h->n * sizeof(double),
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, a->h->stream));
  WAP_HIP(hipStreamSynchronize(a->h->stream));
  return WAP_OK;
}

}  // extern "C"

extern "C" int wap_stream_bandwidth(int64_t bytes, int reps, double *copy_gbs, double *read_gbs) {
  if (bytes < (1 << 20) || reps < 1 || !copy_gbs || !read_gbs)
    return fail(WAP_ERR_ARG, "need bytes >= 1 MiB, reps >= 1 and output pointers");
  const size_t n = (size_t)bytes / 16;
  wap::f32x4 *a = nullptr, *b = nullptr;
  float *sink = nullptr;
  WAP_HIP(hipMalloc((void **)&a, n * 16));
  hipError_t e = hipMalloc((void **)&b, n * 16);
  if (e == hipSuccess) e = hipMalloc((void **)&sink, 16);
  if (e != hipSuccess) {
    (void)hipFree(a);
    (void)hipFree(b);
    return fail(WAP_ERR_OOM, std::string("wap_stream_bandwidth: ") + hipGetErrorString(e));
  }
  int rc = WAP_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  do {
    if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
      rc = fail(WAP_ERR_HIP, "wap_stream_bandwidth setup failed");
      break;
    }
    const dim3 grid((unsigned)((n + wap::kStreamChunk - 1) / wap::kStreamChunk)), block(256);
    float ms = 0.f;
    hipLaunchKernelGGL(wap::stream_copy, grid, block, 0, nullptr, a, b, n);  // warm-up
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(wap::stream_copy, grid, block, 0, nullptr, a, b, n);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    *copy_gbs = 2.0 * (double)n * 16 * reps / (ms * 1e-3) / 1e9;
    hipLaunchKernelGGL(wap::stream_read, grid, block, 0, nullptr, a, n, sink);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(wap::stream_read, grid, block, 0, nullptr, a, n, sink);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    *read_gbs = (double)n * 16 * reps / (ms * 1e-3) / 1e9;
    if (hipGetLastError() != hipSuccess) rc = fail(WAP_ERR_HIP, "stream kernels failed");
  } while (0);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(sink);
  return rc;
}

namespace {

int atom_args_ok(const double *atoms, int n_atoms, int nx, int ny, int nz, const double *extent,
                 double cutoff, void *out, int dtype) {
  if (!atoms || n_atoms < 0 || !extent || !out) return fail(WAP_ERR_ARG, "NULL argument");
  if (nx < 2 || ny < 2 || nz < 2) return fail(WAP_ERR_ARG, "grid must be at least 2 x 2 x 2");
  if (dtype != WAP_DT_F32 && dtype != WAP_DT_F64)
    return fail(WAP_ERR_ARG, "V dtype must be WAP_DT_F32 or WAP_DT_F64");
  if (!(cutoff > 0.0)) return fail(WAP_ERR_ARG, "cutoff must be > 0");
  return WAP_OK;
}

// Copy the host atom records into the device staging d_atoms, run the atom-sum kernel on
// `st` and wait for it (the caller may reuse d_atoms afterwards).
hipError_t atom_potential_run(double *d_atoms, const double *atoms, int n_atoms, int nx, int ny,
                              int nz, const double *extent, double r_soft, double cutoff,
                              void *out, int dtype, hipStream_t st) {
  const double dx = (extent[1] - extent[0]) / (nx - 1.0);
  const double dy = (extent[3] - extent[2]) / (ny - 1.0);
  const double dz = (extent[5] - extent[4]) / (nz - 1.0);
  hipError_t e = hipSuccess;
  if (n_atoms > 0)
    e = hipMemcpyAsync(d_atoms, atoms, (size_t)n_atoms * 5 * sizeof(double),
                       hipMemcpyHostToDevice, st);
  const size_t total = (size_t)nx * ny * nz;
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  const double ke2 = 14.399645;  // e^2 / (4 pi eps0) in eV A
  if (e == hipSuccess) {
    if (dtype == WAP_DT_F64)
      hipLaunchKernelGGL(wap::atom_potential_kernel<double>, grid, block, 0, st, d_atoms, n_atoms,
                         nx, ny, nz, extent[0], extent[2], extent[4], dx, dy, dz, r_soft * r_soft,
                         cutoff, ke2, (double *)out);
    else
      hipLaunchKernelGGL(wap::atom_potential_kernel<float>, grid, block, 0, st, d_atoms, n_atoms,
                         nx, ny, nz, extent[0], extent[2], extent[4], dx, dy, dz, r_soft * r_soft,
                         cutoff, ke2, (float *)out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}

}  // namespace

extern "C" int wap_atom_potential(const double *atoms, int n_atoms, int nx, int ny, int nz,
                                  const double *extent, double r_soft, double cutoff, void *out,
                                  int dtype, void *stream) {
  if (int rc = atom_args_ok(atoms, n_atoms, nx, ny, nz, extent, cutoff, out, dtype)) return rc;
  double *d_atoms = nullptr;
  const size_t ab = (size_t)(n_atoms > 0 ? n_atoms : 1) * 5 * sizeof(double);
  WAP_HIP(hipMalloc((void **)&d_atoms, ab));
  const hipError_t e = atom_potential_run(d_atoms, atoms, n_atoms, nx, ny, nz, extent, r_soft,
                                          cutoff, out, dtype, (hipStream_t)stream);
  (void)hipFree(d_atoms);
  if (e != hipSuccess) return fail(WAP_ERR_HIP, std::string("wap_atom_potential: ") + hipGetErrorString(e));
  return WAP_OK;
}

// ------------------------------------------- allocation-free atom potential
// The atom records of wap_atom_potential_static are staged in this file-scope device
// array instead of a per-call hipMalloc: each GPU holds one copy of it (160 KiB) until
// the process ends.  Host code never uses the array's name as an address (on the host it
// is only a stand-in for the symbol): the device address comes from hipGetSymbolAddress,
// once per device, and is kept in g_scratch_dev.
constexpr int kStaticAtoms = 4096;  // WAP_STATIC_MAX_ATOMS
static_assert(kStaticAtoms == WAP_STATIC_MAX_ATOMS, "header and library disagree");
static __device__ __align__(256) double scratch[kStaticAtoms * 5];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// The current device's address of `scratch` (looked up on the first call on that device).
int scratch_address(double **out) {
  int dev = -1;
  WAP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(WAP_ERR_ARG, "device index beyond the static scratch table (64 devices)");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    WAP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(WAP_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_dev[dev] = (double *)p;
  }
  *out = g_scratch_dev[dev];
  return WAP_OK;
}
}  // namespace

extern "C" int wap_atom_potential_static(const double *atoms, int n_atoms, int nx, int ny,
                                         int nz, const double *extent, double r_soft,
                                         double cutoff, void *out, int dtype, void *stream) {
  if (int rc = atom_args_ok(atoms, n_atoms, nx, ny, nz, extent, cutoff, out, dtype)) return rc;
  if (n_atoms > kStaticAtoms)
    return fail(WAP_ERR_ARG, "wap_atom_potential_static: " + std::to_string(n_atoms) +
                                 " atoms exceed the static staging of " +
                                 std::to_string(kStaticAtoms) + " (use wap_atom_potential)");
  double *d_atoms = nullptr;
  if (int rc = scratch_address(&d_atoms)) return rc;
  const hipError_t e = atom_potential_run(d_atoms, atoms, n_atoms, nx, ny, nz, extent, r_soft,
                                          cutoff, out, dtype, (hipStream_t)stream);
  if (e != hipSuccess)
    return fail(WAP_ERR_HIP, std::string("wap_atom_potential_static: ") + hipGetErrorString(e));
  return WAP_OK;
}

extern "C" int wap_slice_transmit(void *psi, int psi_dtype, const void *v, int v_dtype, int64_t n,
                                  double energy_ev, double rest_energy_ev, double slice_thickness,
                                  double wavelength, void *stream) {
  if (n < 0) return fail(WAP_ERR_ARG, "n must be >= 0");
  if (n > 0 && (!psi || !v)) return fail(WAP_ERR_ARG, "NULL argument");
  if (psi_dtype != WAP_DT_C64 && psi_dtype != WAP_DT_C128)
    return fail(WAP_ERR_ARG, "psi dtype must be WAP_DT_C64 or WAP_DT_C128");
  if (v_dtype != WAP_DT_F32 && v_dtype != WAP_DT_F64)
    return fail(WAP_ERR_ARG, "V dtype must be WAP_DT_F32 or WAP_DT_F64");
  if (!(energy_ev > 0.0) || !(rest_energy_ev > 0.0) || !(wavelength > 0.0))
    return fail(WAP_ERR_ARG, "energy, rest energy and wavelength must be > 0");
  if (n == 0) return WAP_OK;
  const double kdz = 2.0 * M_PI * slice_thickness / wavelength;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 16384);
  const dim3 grid((unsigned)blocks), block(256);
  hipStream_t st = (hipStream_t)stream;
  const bool c128 = psi_dtype == WAP_DT_C128, f64 = v_dtype == WAP_DT_F64;
  if (c128 && f64)
    hipLaunchKernelGGL((wap::slice_transmit_kernel<double, double>), grid, block, 0, st,
                       (wap::cplx<double> *)psi, (const double *)v, n, energy_ev, rest_energy_ev, kdz);
  else if (c128)
    hipLaunchKernelGGL((wap::slice_transmit_kernel<double, float>), grid, block, 0, st,
                       (wap::cplx<double> *)psi, (const float *)v, n, energy_ev, rest_energy_ev, kdz);
  else if (f64)
    hipLaunchKernelGGL((wap::slice_transmit_kernel<float, double>), grid, block, 0, st,
                       (wap::cplx<float> *)psi, (const double *)v, n, energy_ev, rest_energy_ev, kdz);
  else
    hipLaunchKernelGGL((wap::slice_transmit_kernel<float, float>), grid, block, 0, st,
                       (wap::cplx<float> *)psi, (const float *)v, n, energy_ev, rest_energy_ev, kdz);
  WAP_HIP(hipGetLastError());
  return WAP_OK;
}
