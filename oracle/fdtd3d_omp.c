/* FP64 C + OpenMP restatement of the oracle's RK4 FDTD step.
 *
 * TEST INFRASTRUCTURE ONLY: the multi-threaded CPU baseline of bench.py and a
 * second CPU path checked against the NumPy oracle (oracle/fdtd3d.py) in
 * tests/test_oracle_c.py.  Nothing in the package loads it.
 *
 * It follows the reference's wide_angle_propagation/fdtd_3d.cu:
 *   grid, index, spacing     fdtd_3d.cu:31-33, 81-83, 98
 *   PML weights of L_z        fdtd_3d.cu:96-123, 215-234 (per-plane cp, cm from the caller)
 *   soft plane-wave source    fdtd_3d.cu:126-161
 *   RK4 stage                 fdtd_3d.cu:164-257 (k = -iH y; frozen planes k = 0, nz-1)
 *   final update              fdtd_3d.cu:259-273
 *   per-step order            fdtd_3d.cu:363-382
 * with the in-place race of fdtd_3d.cu:374-377 removed (every stage reads an
 * input no thread of the same stage writes), as the NumPy oracle does, and the
 * Klein-Gordon extension of that oracle:
 *   d psi/dt = pi,  d pi/dt = c2 Lap psi - (m2 - V^2) psi - 2 i V pi - 2 gamma(z) pi.
 *
 * The step is written as one streaming sweep per stage: acc = b + dt/6 k1,
 * y2 = b + dt/2 k1; acc += dt/3 k2, y3 = b + dt/2 k2; acc += dt/3 k3,
 * y4 = b + dt k3; psi' = acc + dt/6 k4.  That sums the same terms as the
 * oracle's b + dt/6 (k1 + 2 k2 + 2 k3 + k4) in another order, so the two
 * agree to rounding, not bitwise.
 *
 * Fields are complex128, C order (nz, ny, nx), x fastest.
 */
#include <complex.h>
#include <omp.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cd;

typedef struct {
  int eq;               /* 0 Schrodinger, 1 Klein-Gordon */
  int nx, ny, nz;
  double idx2, idy2;    /* 1/dx^2, 1/dy^2 (the y term is skipped when ny == 1) */
  const cd *cp, *cm;    /* L_z weights per plane (nz) */
  const double *gam;    /* KG sponge rate per plane (nz), or NULL */
  double A, c2, m2, dt;
  int zero_ends;        /* 2D rule: stage inputs are 0 on the end planes */
  const double *V;      /* (nz, ny, nx) */
} Grid;

/* Laplacian of component y at voxel (k, j, i) of an interior plane, summed in the
 * oracle's order: ((x+ - c) + (x- - c)) / dx^2 + ((y+ - c) + (y- - c)) / dy^2
 * + cp (z+ - c) - cm (c - z-) (fdtd_3d.cu:192-238). */
static inline cd lap_at(const Grid *g, const cd *y, int k, int j, int i) {
  const size_t nx = (size_t)g->nx, plane = nx * (size_t)g->ny;
  const size_t o = (size_t)k * plane + (size_t)j * nx + (size_t)i;
  const cd c = y[o];
  const int ip = (i + 1 == g->nx) ? 0 : i + 1, im = (i == 0) ? g->nx - 1 : i - 1;
  const size_t row = (size_t)k * plane + (size_t)j * nx;
  cd l = ((y[row + ip] - c) + (y[row + im] - c)) * g->idx2;
  if (g->ny > 1) {
    const int jp = (j + 1 == g->ny) ? 0 : j + 1, jm = (j == 0) ? g->ny - 1 : j - 1;
    const size_t base = (size_t)k * plane + (size_t)i;
    l += ((y[base + (size_t)jp * nx] - c) + (y[base + (size_t)jm * nx] - c)) * g->idy2;
  }
  l += g->cp[k] * (y[o + plane] - c) - g->cm[k] * (c - y[o - plane]);
  return l;
}

/* One stage: k = F(y) on the interior, then acc (+)= wa k and ynext = b + cn k
 * (stage 1 sets acc = b + wa k; the last stage writes acc + wa k into out). */
static void stage(const Grid *g, int first, int last, double wa, double cn,
                  const cd *ys, const cd *yp,      /* stage input: psi (and pi) */
                  const cd *bs, const cd *bp,      /* injected base */
                  cd *as, cd *ap,                  /* accumulators */
                  cd *ns, cd *np) {                /* next stage input, or the result */
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const size_t plane = (size_t)nx * (size_t)ny;
  const int kg = g->eq == 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      const size_t row = (size_t)k * plane + (size_t)j * (size_t)nx;
      const int frozen = (k == 0 || k == nz - 1);
      for (int i = 0; i < nx; ++i) {
        const size_t o = row + (size_t)i;
        cd ks = 0, kp = 0;
        if (!frozen) {
          if (kg) {
            const double v = g->V[o];
            ks = yp[o];
            kp = g->c2 * lap_at(g, ys, k, j, i) - (g->m2 - v * v) * ys[o] - 2.0 * I * v * yp[o] -
                 2.0 * g->gam[k] * yp[o];
          } else {
            const cd h = -g->A * lap_at(g, ys, k, j, i) + g->V[o] * ys[o];
            ks = -I * h;
          }
        }
        if (last) {
          ns[o] = as[o] + wa * ks;
          if (kg) np[o] = ap[o] + wa * kp;
          continue;
        }
        if (first) {
          as[o] = bs[o] + wa * ks;
          if (kg) ap[o] = bp[o] + wa * kp;
        } else {
          as[o] += wa * ks;
          if (kg) ap[o] += wa * kp;
        }
        const int zero = frozen && g->zero_ends;
        ns[o] = zero ? 0 : bs[o] + cn * ks;
        if (kg) np[o] = zero ? 0 : bp[o] + cn * kp;
      }
    }
}

/* nsteps RK4 steps in place on psi (and pi for Klein-Gordon).
 * Source (fdtd_3d.cu:126-161): before step n, on the nsrc planes src_k[] with weights
 * src_w[], psi <- (1 - a) psi + a inc and pi <- (1 - a) pi + a (-i e_sim) inc, where
 * inc = src_ph[s] * step_ph[n] * prof[j, i] (prof NULL: 1); src_ph[s] = e^{i k (z_s - z_K)}
 * and step_ph[n] = e^{-i e_sim t_n} come from the caller.  Returns 0, or -1 when the
 * scratch memory cannot be allocated. */
int oracle_rk4_run(int eq, int nx, int ny, int nz, double idx2, double idy2, const cd *cp,
                   const cd *cm, const double *gam, double A, double c2, double m2, double dt,
                   int zero_ends, const double *V, int nsrc, const int *src_k,
                   const double *src_w, const cd *src_ph, const cd *prof, double e_sim,
                   const cd *step_ph, int nsteps, cd *psi, cd *pi) {
  Grid g = {eq, nx, ny, nz, idx2, idy2, cp, cm, gam, A, c2, m2, dt, zero_ends, V};
  const size_t n = (size_t)nx * (size_t)ny * (size_t)nz, plane = (size_t)nx * (size_t)ny;
  const int kg = eq == 1;
  const int nc = kg ? 2 : 1;
  /* per component: acc, y (two, ping-pong) */
  cd *buf = (cd *)malloc(sizeof(cd) * n * 3 * (size_t)nc);
  if (!buf) return -1;
  cd *as = buf, *ya = buf + n, *yb = buf + 2 * n;
  cd *ap = kg ? buf + 3 * n : NULL, *pa = kg ? buf + 4 * n : NULL, *pb = kg ? buf + 5 * n : NULL;
  for (int s = 0; s < nsteps; ++s) {
    for (int q = 0; q < nsrc; ++q) {
      const int k = src_k[q];
      const double a = src_w[q];
      const cd base = src_ph[q] * step_ph[s];
#pragma omp parallel for schedule(static)
      for (size_t t = 0; t < plane; ++t) {
        const cd inc = prof ? base * prof[t] : base;
        const size_t o = (size_t)k * plane + t;
        psi[o] = (1.0 - a) * psi[o] + a * inc;
        if (kg) pi[o] = (1.0 - a) * pi[o] + a * (-I * e_sim) * inc;
      }
    }
    /* psi (pi) now hold b for the whole step */
    stage(&g, 1, 0, dt / 6.0, 0.5 * dt, psi, pi, psi, pi, as, ap, ya, pa);
    stage(&g, 0, 0, dt / 3.0, 0.5 * dt, ya, pa, psi, pi, as, ap, yb, pb);
    stage(&g, 0, 0, dt / 3.0, dt, yb, pb, psi, pi, as, ap, ya, pa);
    stage(&g, 0, 1, dt / 6.0, 0.0, ya, pa, psi, pi, as, ap, yb, pb);
    memcpy(psi, yb, sizeof(cd) * n);
    if (kg) memcpy(pi, pb, sizeof(cd) * n);
  }
  free(buf);
  return 0;
}

/* OpenMP threads used by the next call (n <= 0: leave the runtime's choice). */
void oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
int oracle_max_threads(void) { return omp_get_max_threads(); }
