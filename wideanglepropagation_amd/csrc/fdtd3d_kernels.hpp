// gfx950 device code of the fused RK4 FDTD stepper.
//
// One step of the reference (fdtd_3d.cu:363-382) is 2 memsets, a source
// launch, 4 rk4_stage_3d launches that each read-modify-write a global
// accumulator, and rk4_final_3d.  Here a step is 4 streaming passes:
//
//   P1  in = inject(psi_n)          acc  = b + dt/6 k1     out = b + dt/2 k1  (y2)
//   P2  in = y2,  base = inject(psi_n)  acc += dt/3 k2     out = b + dt/2 k2  (y3)
//   P3  in = y3,  base = inject(psi_n)  acc += dt/3 k3     out = b + dt   k3  (y4)
//   P4  in = y4                     psi_{n+1} = acc + dt/6 k4
//
// with b = inject(psi_n) (the soft source of fdtd_3d.cu:126-161 applied on
// load, so psi is never rewritten in place), k = F(in) on interior planes and
// k = 0 on k = 0, nz-1 (the frozen planes of fdtd_3d.cu:184-190).  Stage
// inputs and outputs are distinct buffers (no in-place stencil race).
//
// F32_KAHAN keeps acc as the increment sum only (P1 writes acc = dt/6 k1) and
// P4 adds it to b with a Kahan-compensated update against a persistent
// per-voxel compensation buffer.
//
// Thread mapping: a workgroup is 64 x TY threads over a 64 x TY column tile;
// each thread marches up a chunk of tile_z planes keeping y(k-1), y(k), y(k+1)
// in registers (z reuse), the +-x / +-y neighbours come from L1/L2 (they are
// the loads of adjacent lanes / rows of the same workgroup).  Tiles are
// numbered so each XCD (workgroup id mod 8) gets one contiguous range: the
// lateral halo rows of a tile are then read by workgroups on the same XCD and
// hit its L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wap {

constexpr int kMaxBand = 8;
constexpr int kWaveX = 64;   // lanes along x per row (one wavefront wide)
constexpr int kNumXcd = 8;   // MI355X: 8 XCDs, workgroups dispatched round-robin

template <typename R> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };
template <typename R> using cplx = typename cplx_of<R>::type;

template <typename R>
__device__ __forceinline__ cplx<R> mkc(R x, R y) {
  cplx<R> c;
  c.x = x;
  c.y = y;
  return c;
}
template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return mkc<R>(a.x + b.x, a.y + b.y); }
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return mkc<R>(a.x - b.x, a.y - b.y); }
template <typename R>
__device__ __forceinline__ cplx<R> cscale(cplx<R> a, R s) { return mkc<R>(a.x * s, a.y * s); }
template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return mkc<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a + s*b
template <typename R>
__device__ __forceinline__ cplx<R> caxpy(cplx<R> a, R s, cplx<R> b) {
  return mkc<R>(a.x + s * b.x, a.y + s * b.y);
}

// Per-launch arguments of one pass.  Passed by value (kernarg segment).
template <typename R>
struct PassArgs {
  using C = cplx<R>;
  const C *yin_psi;    // stage input, psi component (P1: the state psi_n)
  const C *yin_pi;     // stage input, pi component (KG)
  const C *st_psi;     // state psi_n (base of P2/P3, Kahan P4)
  const C *st_pi;
  C *acc_psi;          // accumulator
  C *acc_pi;
  C *out_psi;          // stage output y, or the new state (P4)
  C *out_pi;
  C *cmp_psi;          // Kahan compensation (P4 only)
  C *cmp_pi;
  const R *V;          // potential (nz,ny,nx)
  const C *zc;         // per plane k: zc[2k] = z+ weight, zc[2k+1] = z- weight
  const R *gam;        // KG sponge damping per plane
  const C *prof;       // source profile (ny,nx) or nullptr
  int nx, ny, nz;
  int tile_y, tile_z;
  int ntx, nty, ntz;   // tile counts
  R hx, hy;            // lateral weights
  R m2;                // KG: (m c^2)^2
  R w_acc, w_y;        // stage weights
  int src_k0, src_n;   // source planes src_k0 .. src_k0+src_n-1
  R src_keep[kMaxBand];
  C src_coef[kMaxBand];     // a_b e^{i k (z-zK)} e^{-i w t}
  C src_coef_pi[kMaxBand];  // (-i w) * src_coef   (KG)
};

// Workgroup id -> tile id such that the workgroups of XCD x (id % 8 == x)
// cover one contiguous range of tile ids.
__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
  const int q = ntiles / kNumXcd;
  const int rem = ntiles % kNumXcd;
  const int xcd = bid % kNumXcd;
  const int pos = bid / kNumXcd;
  return xcd * q + (xcd < rem ? xcd : rem) + pos;
}

// Source band selection without dynamic indexing of the kernarg arrays.
template <typename R>
__device__ __forceinline__ bool src_pick(const PassArgs<R> &p, int k, R &keep, cplx<R> &coef,
                                         cplx<R> &coef_pi) {
  const int b = k - p.src_k0;
  if (b < 0 || b >= p.src_n) return false;
#pragma unroll
  for (int q = 0; q < kMaxBand; ++q) {
    if (q == b) {
      keep = p.src_keep[q];
      coef = p.src_coef[q];
      coef_pi = p.src_coef_pi[q];
    }
  }
  return true;
}

// Load psi component of the stage input; P1 applies the soft source on load.
template <typename R, int STAGE>
__device__ __forceinline__ cplx<R> load_in(const PassArgs<R> &p, const cplx<R> *arr, size_t idx,
                                           int k, int ij) {
  cplx<R> v = arr[idx];
  if (STAGE == 1) {
    R keep;
    cplx<R> coef, coef_pi;
    if (src_pick(p, k, keep, coef, coef_pi)) {
      cplx<R> inc = coef;
      if (p.prof) inc = cmul<R>(coef, p.prof[ij]);
      v = caxpy<R>(cscale<R>(v, keep), R(1), inc);
    }
  }
  return v;
}

template <typename R, int EQ, int STAGE, bool KAHAN>
__global__ __launch_bounds__(512) void rk4_pass(const PassArgs<R> p) {
  using C = cplx<R>;
  const int ntiles = p.ntx * p.nty * p.ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int tx = tile % p.ntx;
  const int rest = tile / p.ntx;
  const int ty = rest % p.nty;
  const int tz = rest / p.nty;
  const int i = tx * kWaveX + threadIdx.x;
  const int j = ty * p.tile_y + threadIdx.y;
  if (i >= p.nx || j >= p.ny) return;  // no workgroup barriers in this kernel
  const int k0 = tz * p.tile_z;
  const int k1 = min(p.nz, k0 + p.tile_z);

  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const size_t plane = (size_t)nx * ny;
  const int ij = j * nx + i;
  const int ip = (i + 1 == nx) ? 0 : i + 1;
  const int im = (i == 0) ? nx - 1 : i - 1;
  const int jp = (j + 1 == ny) ? 0 : j + 1;
  const int jm = (j == 0) ? ny - 1 : j - 1;
  const int o_xp = j * nx + ip, o_xm = j * nx + im;
  const int o_yp = jp * nx + i, o_ym = jm * nx + i;

  const C zero = mkc<R>(R(0), R(0));
  C ym = (k0 > 0) ? load_in<R, STAGE>(p, p.yin_psi, (size_t)(k0 - 1) * plane + ij, k0 - 1, ij) : zero;
  C yc = load_in<R, STAGE>(p, p.yin_psi, (size_t)k0 * plane + ij, k0, ij);

  for (int k = k0; k < k1; ++k) {
    const size_t base = (size_t)k * plane;
    const size_t idx = base + ij;
    C yp = (k + 1 < nz) ? load_in<R, STAGE>(p, p.yin_psi, idx + plane, k + 1, ij) : zero;
    const bool interior = (k > 0) && (k < nz - 1);
    const R v = p.V[idx];

    // stage input, pi component (KG), and its injected value on P1
    C q = zero;
    R keep = R(1);
    C coef = zero, coef_pi = zero;
    const bool in_band = src_pick(p, k, keep, coef, coef_pi);
    C inc = coef, inc_pi = coef_pi;
    if (in_band && p.prof) {
      const C pr = p.prof[ij];
      inc = cmul<R>(coef, pr);
      inc_pi = cmul<R>(coef_pi, pr);
    }
    if (EQ == 1) {
      q = p.yin_pi[idx];
      if (STAGE == 1 && in_band) q = caxpy<R>(cscale<R>(q, keep), R(1), inc_pi);
    }

    C kpsi = zero, kpi = zero;
    if (interior) {
      const C xp = load_in<R, STAGE>(p, p.yin_psi, base + o_xp, k, o_xp);
      const C xm = load_in<R, STAGE>(p, p.yin_psi, base + o_xm, k, o_xm);
      const C yyp = load_in<R, STAGE>(p, p.yin_psi, base + o_yp, k, o_yp);
      const C yym = load_in<R, STAGE>(p, p.yin_psi, base + o_ym, k, o_ym);
      const C wzp = p.zc[2 * k];
      const C wzm = p.zc[2 * k + 1];
      const C lx = cadd<R>(csub<R>(xp, yc), csub<R>(xm, yc));
      const C ly = cadd<R>(csub<R>(yyp, yc), csub<R>(yym, yc));
      const C lz = csub<R>(cmul<R>(wzp, csub<R>(yp, yc)), cmul<R>(wzm, csub<R>(yc, ym)));
      C lap = cadd<R>(cadd<R>(cscale<R>(lx, p.hx), cscale<R>(ly, p.hy)), lz);
      if (EQ == 0) {
        // H y = -A Lap y + V y (weights carry -A); k = -i H y
        const C h = caxpy<R>(lap, v, yc);
        kpsi = mkc<R>(h.y, -h.x);
      } else {
        // d psi = pi ; d pi = c^2 Lap psi - (m^2 - V^2) psi - 2 i V pi - 2 g pi
        const R g = p.gam[k];
        const R w = p.m2 - v * v;
        kpsi = q;
        kpi = mkc<R>(lap.x - w * yc.x - R(2) * g * q.x + R(2) * v * q.y,
                     lap.y - w * yc.y - R(2) * v * q.x - R(2) * g * q.y);
      }
    }

    // base b = inject(psi_n): P1 has it in yc/q; P2, P3 and Kahan P4 reload it.
    C b = yc, bq = q;
    if (STAGE == 2 || STAGE == 3 || (STAGE == 4 && KAHAN)) {
      b = p.st_psi[idx];
      if (in_band) b = caxpy<R>(cscale<R>(b, keep), R(1), inc);
      if (EQ == 1) {
        bq = p.st_pi[idx];
        if (in_band) bq = caxpy<R>(cscale<R>(bq, keep), R(1), inc_pi);
      }
    }

    if (STAGE == 1) {
      p.acc_psi[idx] = KAHAN ? cscale<R>(kpsi, p.w_acc) : caxpy<R>(b, p.w_acc, kpsi);
      p.out_psi[idx] = caxpy<R>(b, p.w_y, kpsi);
      if (EQ == 1) {
        p.acc_pi[idx] = KAHAN ? cscale<R>(kpi, p.w_acc) : caxpy<R>(bq, p.w_acc, kpi);
        p.out_pi[idx] = caxpy<R>(bq, p.w_y, kpi);
      }
    } else if (STAGE == 2 || STAGE == 3) {
      p.acc_psi[idx] = caxpy<R>(p.acc_psi[idx], p.w_acc, kpsi);
      p.out_psi[idx] = caxpy<R>(b, p.w_y, kpsi);
      if (EQ == 1) {
        p.acc_pi[idx] = caxpy<R>(p.acc_pi[idx], p.w_acc, kpi);
        p.out_pi[idx] = caxpy<R>(bq, p.w_y, kpi);
      }
    } else {  // STAGE 4
      if (!KAHAN) {
        p.out_psi[idx] = caxpy<R>(p.acc_psi[idx], p.w_acc, kpsi);
        if (EQ == 1) p.out_pi[idx] = caxpy<R>(p.acc_pi[idx], p.w_acc, kpi);
      } else {
        // true state = hi - comp; the source scales both by keep.
        C c = p.cmp_psi[idx];
        if (in_band) c = cscale<R>(c, keep);
        const C tot = caxpy<R>(p.acc_psi[idx], p.w_acc, kpsi);
        const C yk = csub<R>(tot, c);
        const C t = cadd<R>(b, yk);
        p.cmp_psi[idx] = csub<R>(csub<R>(t, b), yk);
        p.out_psi[idx] = t;
        if (EQ == 1) {
          C cq = p.cmp_pi[idx];
          if (in_band) cq = cscale<R>(cq, keep);
          const C totq = caxpy<R>(p.acc_pi[idx], p.w_acc, kpi);
          const C ykq = csub<R>(totq, cq);
          const C tq = cadd<R>(bq, ykq);
          p.cmp_pi[idx] = csub<R>(csub<R>(tq, bq), ykq);
          p.out_pi[idx] = tq;
        }
      }
    }
    ym = yc;
    yc = yp;
  }
}

// ---------------------------------------------------------------- utilities

// out = (complex TO) (hi - cmp)   (cmp may be null); n elements
template <typename RI, typename RO>
__global__ void export_state(const cplx<RI> *hi, const cplx<RI> *cmp, cplx<RO> *out, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) {
    cplx<RI> v = hi[t];
    if (cmp) {
      // compensated value, rounded once into the output type
      const cplx<RI> c = cmp[t];
      out[t] = mkc<RO>((RO)((double)v.x - (double)c.x), (RO)((double)v.y - (double)c.y));
    } else {
      out[t] = mkc<RO>((RO)v.x, (RO)v.y);
    }
  }
}

// hi = (RO) in; cmp = hi - in (so that hi - cmp reproduces in to the extra bits)
template <typename RI, typename RO>
__global__ void import_state(const cplx<RI> *in, cplx<RO> *hi, cplx<RO> *cmp, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) {
    const cplx<RI> v = in ? in[t] : mkc<RI>(RI(0), RI(0));
    const cplx<RO> h = mkc<RO>((RO)v.x, (RO)v.y);
    hi[t] = h;
    if (cmp) cmp[t] = mkc<RO>((RO)((double)h.x - (double)v.x), (RO)((double)h.y - (double)v.y));
  }
}

template <typename RI, typename RO>
__global__ void import_real(const RI *in, RO *out, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) out[t] = (RO)in[t];
}

template <typename R>
__global__ void accumulate_plane(const cplx<R> *psi, const cplx<R> *cmp, size_t off, int n, double c,
                                 double s, double2 *acc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  cplx<R> v = psi[off + t];
  double vr = (double)v.x, vi = (double)v.y;
  if (cmp) {
    const cplx<R> e = cmp[off + t];
    vr -= (double)e.x;
    vi -= (double)e.y;
  }
  double2 a = acc[t];
  a.x += vr * c - vi * s;
  a.y += vr * s + vi * c;
  acc[t] = a;
}

// Deterministic two-pass FP64 sum of |psi|^2.
template <typename R>
__global__ __launch_bounds__(256) void norm2_partial(const cplx<R> *psi, const cplx<R> *cmp, size_t n,
                                                     double *partial) {
  __shared__ double red[256];
  double acc = 0.0;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) {
    const cplx<R> v = psi[t];
    double vr = (double)v.x, vi = (double)v.y;
    if (cmp) {
      const cplx<R> e = cmp[t];
      vr -= (double)e.x;
      vi -= (double)e.y;
    }
    acc += vr * vr + vi * vi;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void sum_partials(const double *partial, int n, double *out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) acc += partial[t];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace wap
