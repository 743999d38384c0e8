/*
 * CPU oracle of MISO's grey short-characteristics radiative transfer --
 * TEST INFRASTRUCTURE ONLY (only tests/, smoke() and bench.py's cpu_baseline
 * may load it).
 *
 * Restates miso/include/miso/rt_quadrature.hpp and rt_host.hpp for one rank,
 * Real = double:
 *   - Carlson A4 quadrature (rt_quadrature.hpp:39-55), single +z ray
 *     otherwise (:57-62);
 *   - the RT lattice ib0 = i_margin - is ... ib1 = ib0 + i_size (rt_host.hpp:127-154);
 *   - the periodic halo self-exchange, z then y then x (rt_host.hpp:229-452);
 *   - incoming-boundary setting on physical faces (rt_host.hpp:50-101);
 *   - single_sweep (rt_host.hpp:456-610), get_max_diff (612-629),
 *     solve (632-651), the segment integrals (654-689).
 * Q_rad has no counterpart in the reference: Q_rad = 4*pi*kappa*(J - S),
 * J = sum_r w_r I_r, on every lattice node.
 *
 * Arrays: kappa, S are (i_total, j_total, k_total) in MISO order (k fastest);
 * rint is (num_rays, i_total, j_total, k_total).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mhd_oracle.h"

/* OG_PLANT (default 0) selects a deliberately wrong variant for the tests
 * that show the pins catch it (tests/test_oracle_pins.py). */
#ifndef OG_PLANT
#define OG_PLANT 0
#endif

static inline double max2n(double a, double b) { return (double)fmaxf((float)a, (float)b); }
static inline double min3n(double a, double b, double c) {
  return (double)fminf(fminf((float)a, (float)b), (float)c);
}

/* rt_quadrature.hpp:24-62 */
int og_rt_quadrature(int num_rays, double *w, double *mx, double *my, double *mz) {
  if (num_rays == 24) {
    const double mu0 = 1.0 / 3.0;
    const double mu1 = sqrt(1.0 - 2.0 * mu0 * mu0);
    const double ww0 = 1.0 / 24.0;
    const double X[24] = {-mu1, +mu1, -mu1, +mu1, -mu1, +mu1, -mu1, +mu1,
                          -mu0, +mu0, -mu0, +mu0, -mu0, +mu0, -mu0, +mu0,
                          -mu0, +mu0, -mu0, +mu0, -mu0, +mu0, -mu0, +mu0};
    const double Y[24] = {-mu0, -mu0, +mu0, +mu0, -mu0, -mu0, +mu0, +mu0,
                          -mu1, -mu1, +mu1, +mu1, -mu1, -mu1, +mu1, +mu1,
                          -mu0, -mu0, +mu0, +mu0, -mu0, -mu0, +mu0, +mu0};
    const double Z[24] = {-mu0, -mu0, -mu0, -mu0, +mu0, +mu0, +mu0, +mu0,
                          -mu0, -mu0, -mu0, -mu0, +mu0, +mu0, +mu0, +mu0,
                          -mu1, -mu1, -mu1, -mu1, +mu1, +mu1, +mu1, +mu1};
    for (int r = 0; r < 24; ++r) {
      w[r] = ww0;
      mx[r] = X[r];
      my[r] = Y[r];
      mz[r] = Z[r];
    }
    return 24;
  }
  w[0] = 1.0;
  mx[0] = 0.0;
  my[0] = 0.0;
  mz[0] = 1.0;
  return 1;
}

typedef struct {
  int It, Jt, Kt;
  int b0[3], b1[3], st[3], size[3];
} rtg;

static rtg rt_make(const og_grid *g) {
  rtg r;
  r.It = g->i_total;
  r.Jt = g->j_total;
  r.Kt = g->k_total;
  r.st[0] = g->is; r.st[1] = g->js; r.st[2] = g->ks;
  r.size[0] = g->i_size; r.size[1] = g->j_size; r.size[2] = g->k_size;
  r.b0[0] = g->i_margin - g->is;
  r.b0[1] = g->j_margin - g->js;
  r.b0[2] = g->k_margin - g->ks;
  /* rt_host.hpp:146-154 sets b1 = b0 + size on every axis, which on an axis
   * of one cell addresses node 1 of a length-1 axis (it aliases the next row
   * and overruns the array on the last ray); the oracle -- like the GPU
   * solver -- keeps the single node 0 there.  Axes with more than one cell
   * are unchanged. */
  for (int a = 0; a < 3; ++a) r.b1[a] = r.b0[a] + (r.size[a] > 1 ? r.size[a] : 0);
  return r;
}

#define NODE(r, n, i, j, k) ((((size_t)(n) * (r).It + (i)) * (r).Jt + (j)) * (r).Kt + (k))
#define CELL(r, i, j, k) ((((size_t)(i)) * (r).Jt + (j)) * (r).Kt + (k))

/* one axis of the periodic halo self-exchange (rt_host.hpp:236-452): rays
 * moving + along the axis receive node b1 into node b0, rays moving - receive
 * b0 into b1, over the full transverse extent */
static void rt_halo_axis(const rtg *r, int nr, const double *mu, int axis, double *I) {
  const int b0 = r->b0[axis], b1 = r->b1[axis];
  const int n1 = axis == 0 ? r->Jt : r->It, n2 = axis == 2 ? r->Jt : r->Kt;
  for (int n = 0; n < nr; ++n) {
    const int pos = mu[n] >= 0.0;
    for (int u = 0; u < n1; ++u)
      for (int v = 0; v < n2; ++v) {
        int i, j, k;
        if (axis == 0) { j = u; k = v; i = 0; }
        else if (axis == 1) { i = u; k = v; j = 0; }
        else { i = u; j = v; k = 0; }
        int src[3] = {i, j, k}, dst[3] = {i, j, k};
        src[axis] = pos ? b1 : b0;
        dst[axis] = pos ? b0 : b1;
        I[NODE(*r, n, dst[0], dst[1], dst[2])] = I[NODE(*r, n, src[0], src[1], src[2])];
      }
  }
}

/* set_incoming_boundary (rt_host.hpp:50-89) for one face.  kind 1: I = S at
 * the face node (thermal), kind 2: I = 0, kind 3: I = value. */
static void rt_set_face(const rtg *r, int nr, const double *mu, int axis, int side, int kind,
                        double value, const double *S, double *I) {
  int lo[3] = {r->b0[0], r->b0[1], r->b0[2]}, hi[3] = {r->b1[0], r->b1[1], r->b1[2]};
  lo[axis] = hi[axis] = side == 0 ? r->b0[axis] : r->b1[axis];
  for (int n = 0; n < nr; ++n) {
    const int incoming = side == 0 ? (mu[n] > 0.0) : (mu[n] < 0.0);
    if (!incoming) continue;
    for (int i = lo[0]; i <= hi[0]; ++i)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int k = lo[2]; k <= hi[2]; ++k)
          I[NODE(*r, n, i, j, k)] = kind == 1 ? S[CELL(*r, i, j, k)] : kind == 2 ? 0.0 : value;
  }
}

/* rt_host.hpp:680-689 */
static inline double bilinear(double cx, double cy, double u00, double u10, double u01,
                              double u11) {
  if (OG_PLANT == 15) { const double t = u10; u10 = u01; u01 = t; }
  return (1.0 - cx) * (1.0 - cy) * u00 + cx * (1.0 - cy) * u10 + (1.0 - cx) * cy * u01 +
         cx * cy * u11;
}

/* rt_host.hpp:667-678 */
static inline double integrate_intensity(double Iu, double Su, double Sd, double dtau) {
  const double eps_tau = 1.e-4;
  const double exp_dtau = exp(-dtau);
  const double u0 = 1.0 - exp_dtau;
  if (dtau < eps_tau) {
    const double u1 = 0.5 * dtau;
    return Iu * exp_dtau + (u0 - u1) * Su + u1 * Sd;
  }
  const double u1 = 1.0 - (u0 / dtau);
  if (OG_PLANT == 14) return Iu * exp_dtau + u1 * Su + (u0 - u1) * Sd;
  return Iu * exp_dtau + (u0 - u1) * Su + u1 * Sd;
}

/* single_sweep (rt_host.hpp:456-610) */
static void rt_sweep(const rtg *r, const og_grid *g, const double *dxyz, int nr, const double *mx,
                     const double *my, const double *mz, const double *kap, const double *S,
                     double *I) {
  for (int n = 0; n < nr; ++n) {
    const double eps_mu = 1.e-30;
    const double amx = max2n(fabs(mx[n]), eps_mu), amy = max2n(fabs(my[n]), eps_mu),
                 amz = max2n(fabs(mz[n]), eps_mu);
    const double dx = g->i_size > 1 ? dxyz[0] : 1.e30;
    const double dy = g->j_size > 1 ? dxyz[1] : 1.e30;
    const double dz = g->k_size > 1 ? dxyz[2] : 1.e30;
    const double dsx = dx / amx, dsy = dy / amy, dsz = dz / amz;
    const double ds = min3n(dsx, dsy, dsz);
    const double ccx = amx * ds / dx, ccy = amy * ds / dy, ccz = amz * ds / dz;
    int ib, ie, it, jb, je, jt, kb, ke, kt;
    if (mx[n] >= 0.0) { ib = r->b0[0] + g->is; ie = r->b1[0]; it = 1; }
    else { ib = r->b1[0] - g->is; ie = r->b0[0]; it = -1; }
    if (my[n] >= 0.0) { jb = r->b0[1] + g->js; je = r->b1[1]; jt = 1; }
    else { jb = r->b1[1] - g->js; je = r->b0[1]; jt = -1; }
    if (mz[n] >= 0.0) { kb = r->b0[2] + g->ks; ke = r->b1[2]; kt = 1; }
    else { kb = r->b1[2] - g->ks; ke = r->b0[2]; kt = -1; }
    const int ist = it * g->is, jst = jt * g->js, kst = kt * g->ks;
    const int jlo = jb < je ? jb : je, jhi = jb < je ? je : jb;
    const int klo = kb < ke ? kb : ke, khi = kb < ke ? ke : kb;
    const int ilo = ib < ie ? ib : ie, ihi = ib < ie ? ie : ib;
#define RINT(i, j, k) I[NODE(*r, n, i, j, k)]
#define SF(i, j, k) S[CELL(*r, i, j, k)]
#define KA(i, j, k) kap[CELL(*r, i, j, k)]
    if (dsx <= dsy && dsx <= dsz) {
      for (int i = ib; i != ie + it; i += it)
        for (int j = jlo; j <= jhi; ++j)
          for (int k = klo; k <= khi; ++k) {
            const double iu = bilinear(ccy, ccz, RINT(i - ist, j, k), RINT(i - ist, j - jst, k),
                                       RINT(i - ist, j, k - kst), RINT(i - ist, j - jst, k - kst));
            const double su = bilinear(ccy, ccz, SF(i - ist, j, k), SF(i - ist, j - jst, k),
                                       SF(i - ist, j, k - kst), SF(i - ist, j - jst, k - kst));
            const double ku = bilinear(ccy, ccz, KA(i - ist, j, k), KA(i - ist, j - jst, k),
                                       KA(i - ist, j, k - kst), KA(i - ist, j - jst, k - kst));
            const double dtau = 0.5 * (ku + KA(i, j, k)) * ds;
            RINT(i, j, k) = integrate_intensity(iu, su, SF(i, j, k), dtau);
          }
    } else if (dsy <= dsx && dsy <= dsz) {
      for (int j = jb; j != je + jt; j += jt)
        for (int i = ilo; i <= ihi; ++i)
          for (int k = klo; k <= khi; ++k) {
            const double iu = bilinear(ccx, ccz, RINT(i, j - jst, k), RINT(i - ist, j - jst, k),
                                       RINT(i, j - jst, k - kst), RINT(i - ist, j - jst, k - kst));
            const double su = bilinear(ccx, ccz, SF(i, j - jst, k), SF(i - ist, j - jst, k),
                                       SF(i, j - jst, k - kst), SF(i - ist, j - jst, k - kst));
            const double ku = bilinear(ccx, ccz, KA(i, j - jst, k), KA(i - ist, j - jst, k),
                                       KA(i, j - jst, k - kst), KA(i - ist, j - jst, k - kst));
            const double dtau = 0.5 * (ku + KA(i, j, k)) * ds;
            RINT(i, j, k) = integrate_intensity(iu, su, SF(i, j, k), dtau);
          }
    } else {
      for (int k = kb; k != ke + kt; k += kt)
        for (int i = ilo; i <= ihi; ++i)
          for (int j = jlo; j <= jhi; ++j) {
            const double iu = bilinear(ccx, ccy, RINT(i, j, k - kst), RINT(i - ist, j, k - kst),
                                       RINT(i, j - jst, k - kst), RINT(i - ist, j - jst, k - kst));
            const double su = bilinear(ccx, ccy, SF(i, j, k - kst), SF(i - ist, j, k - kst),
                                       SF(i, j - jst, k - kst), SF(i - ist, j - jst, k - kst));
            const double ku = bilinear(ccx, ccy, KA(i, j, k - kst), KA(i - ist, j, k - kst),
                                       KA(i, j - jst, k - kst), KA(i - ist, j - jst, k - kst));
            const double dtau = 0.5 * (ku + KA(i, j, k)) * ds;
            RINT(i, j, k) = integrate_intensity(iu, su, SF(i, j, k), dtau);
          }
    }
#undef RINT
#undef SF
#undef KA
  }
}

/* solve (rt_host.hpp:632-651).  periodic[a]: halo self-exchange on axis a;
 * face_kind[a][side] (non-periodic axes): 0 none, 1 thermal (I = S),
 * 2 vacuum (I = 0), 3 constant face_value[a][side].  Returns iterations
 * done; *max_diff receives the last maximum difference. */
int og_rt_solve(const og_grid *g, const double *dxyz, int num_rays, const double *kap,
                const double *S, double *I, const int *periodic, const int *face_kind,
                const double *face_value, double tolerance, int max_iters, double *max_diff) {
  double w[24], mx[24], my[24], mz[24];
  const int nr = og_rt_quadrature(num_rays, w, mx, my, mz);
  const rtg r = rt_make(g);
  const size_t n = (size_t)nr * r.It * r.Jt * r.Kt;
  double *old = (double *)malloc(sizeof(double) * n);
  const double *mus[3] = {mx, my, mz};
  int iter;
  double md = 0.0;
  for (iter = 0; iter < max_iters; ++iter) {
    memcpy(old, I, sizeof(double) * n);
    /* mpi_exchange_halo: z, y, x (rt_host.hpp:229-233) */
    for (int a = 2; a >= 0; --a)
      if (periodic[a] && r.size[a] > 1) rt_halo_axis(&r, nr, mus[a], a, I);
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side)
        if (!periodic[a] && r.size[a] > 1 && face_kind[2 * a + side])
          rt_set_face(&r, nr, mus[a], a, side, face_kind[2 * a + side],
                      face_value ? face_value[2 * a + side] : 0.0, S, I);
    rt_sweep(&r, g, dxyz, nr, mx, my, mz, kap, S, I);
    /* get_max_diff over the lattice (rt_host.hpp:612-629) */
    md = 0.0;
    for (int q = 0; q < nr; ++q)
      for (int i = r.b0[0]; i <= r.b1[0]; ++i)
        for (int j = r.b0[1]; j <= r.b1[1]; ++j)
          for (int k = r.b0[2]; k <= r.b1[2]; ++k) {
            const size_t c = NODE(r, q, i, j, k);
            md = max2n(md, fabs(I[c] - old[c]));
          }
    if (md < tolerance) {
      ++iter;
      break;
    }
  }
  free(old);
  if (max_diff) *max_diff = md;
  return iter;
}

/* Q_rad = 4 pi kappa (J - S) on every lattice node, J = sum_r w_r I_r
 * (summed in ray order); zero elsewhere. */
void og_rt_qrad(const og_grid *g, int num_rays, const double *kap, const double *S,
                const double *I, double *qrad) {
  double w[24], mx[24], my[24], mz[24];
  const int nr = og_rt_quadrature(num_rays, w, mx, my, mz);
  const rtg r = rt_make(g);
  memset(qrad, 0, sizeof(double) * r.It * r.Jt * r.Kt);
  for (int i = r.b0[0]; i <= r.b1[0]; ++i)
    for (int j = r.b0[1]; j <= r.b1[1]; ++j)
      for (int k = r.b0[2]; k <= r.b1[2]; ++k) {
        double J = 0.0;
        for (int q = 0; q < nr; ++q) J += w[q] * I[NODE(r, q, i, j, k)];
        const size_t c = CELL(r, i, j, k);
        qrad[c] = (OG_PLANT == 16 ? 2.0 : 4.0) * M_PI * kap[c] * (J - S[c]);
      }
}
