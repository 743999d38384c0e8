// C ABI of the GPU radiative transfer (include/misohip.h, misohip_rt_*).
//
// Host side of RT::solve (miso/include/miso/rt_host.hpp:632-651): per-ray
// sweep constants are computed here with the reference's expressions
// (rt_host.hpp:461-508, float-narrowed max/min included), then every
// iteration is: save the rays' upwind faces into rint_old, periodic halo
// (z, y, x), incoming boundary on physical faces, the sweeps (which take
// the change of every node they write), the change on the upwind faces,
// and the max change read back to decide convergence.  That is the
// reference's max over the lattice of |rint - rint_old| with rint_old a
// full copy (rt_host.hpp:612-651), without the copy.
//
// Rays whose dominant axis is x run in the y-fastest frame (node (i, j, k)
// at (k * i_total + i) * j_total + j): their constants are those of the
// x-fastest frame with the strides of x and y swapped, and each halo, face
// and reduction launch comes in two, one per frame, with x and y swapped in
// the second.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ctx_internal.hpp"
#include "launch.hpp"

using namespace misohip;

namespace {

struct RtSetup {
  int nrays;
  double w[24], mu[3][24];
  RtRays rays;
  int max_nodes, max_planes;
  int b0[3], b1[3], size[3], st[3];
  long ray_stride;
  unsigned pos_mask[3];
  unsigned tr_mask;  // rays kept y-fastest
};

// rt_quadrature.hpp:39-62
int quadrature(int num_rays, double *w, double *mx, double *my, double *mz) {
  if (num_rays == 24) {
    const double mu0 = 1.0 / 3.0;
    const double mu1 = std::sqrt(1.0 - 2.0 * mu0 * mu0);
    const double ww0 = 1.0 / 24.0;
    for (int r = 0; r < 24; ++r) {
      // rays 0-7: |mu_x| = mu1, 8-15: |mu_y| = mu1, 16-23: |mu_z| = mu1; signs
      // cycle x fastest, then y, then z (the tables of rt_quadrature.hpp:46-54)
      const int g = r / 8, b = r % 8;
      const double sx = (b & 1) ? 1.0 : -1.0, sy = (b & 2) ? 1.0 : -1.0, sz = (b & 4) ? 1.0 : -1.0;
      w[r] = ww0;
      mx[r] = sx * (g == 0 ? mu1 : mu0);
      my[r] = sy * (g == 1 ? mu1 : mu0);
      mz[r] = sz * (g == 2 ? mu1 : mu0);
    }
    return 24;
  }
  w[0] = 1.0;
  mx[0] = 0.0;
  my[0] = 0.0;
  mz[0] = 1.0;
  return 1;
}

inline double nmax2(double a, double b) { return (double)std::fmax((float)a, (float)b); }
inline double nmin3(double a, double b, double c) {
  return (double)std::fmin(std::fmin((float)a, (float)b), (float)c);
}

int setup(misohip_ctx *ctx, const misohip_grid *g, const misohip_rt_config *rt, RtSetup &S) {
  S.nrays = quadrature(rt->num_rays, S.w, S.mu[0], S.mu[1], S.mu[2]);
  const int It = g->i_total, Jt = g->j_total, Kt = g->k_total;
  const int sizes[3] = {g->i_size, g->j_size, g->k_size};
  const int margins[3] = {g->i_margin, g->j_margin, g->k_margin};
  const int sts[3] = {g->is, g->js, g->ks};
  const long strides_x[3] = {1, It, (long)It * Jt};
  const long strides_y[3] = {Jt, 1, (long)It * Jt};
  S.tr_mask = 0u;
  for (int a = 0; a < 3; ++a) {
    S.size[a] = sizes[a];
    S.st[a] = sts[a];
    S.b0[a] = margins[a] - sts[a];
    // on an axis of one cell the lattice is the single node 0 (the
    // reference's b0 + size would address index 1 of a length-1 axis)
    S.b1[a] = S.b0[a] + (sizes[a] > 1 ? sizes[a] : 0);
  }
  S.ray_stride = (long)It * Jt * Kt;
  // uniform spacings (rt_host.hpp:467-469 uses dx[0] etc.), read in stream
  // order into the context's pinned words and waited for once
  double d[3] = {1.e30, 1.e30, 1.e30};
  const void *dptr[3] = {g->dx, g->dy, g->dz};
  double *hd = reinterpret_cast<double *>(reinterpret_cast<char *>(ctx->h_slot) + 64);
  bool any = false;
  for (int a = 0; a < 3; ++a)
    if (sizes[a] > 1) {
      MISOHIP_TRY(hipMemcpyAsync(hd + a, dptr[a], sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->stream));
      any = true;
    }
  if (any) MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int a = 0; a < 3; ++a)
    if (sizes[a] > 1) d[a] = hd[a];
  S.max_nodes = 0;
  S.max_planes = 0;
  for (int a = 0; a < 3; ++a) S.pos_mask[a] = 0u;
  for (int n = 0; n < S.nrays; ++n) {
    const double mx = S.mu[0][n], my = S.mu[1][n], mz = S.mu[2][n];
    const double eps_mu = 1.e-30;
    const double am[3] = {nmax2(std::fabs(mx), eps_mu), nmax2(std::fabs(my), eps_mu),
                          nmax2(std::fabs(mz), eps_mu)};
    const double dsa[3] = {d[0] / am[0], d[1] / am[1], d[2] / am[2]};
    const double ds = nmin3(dsa[0], dsa[1], dsa[2]);
    const double cc[3] = {am[0] * ds / d[0], am[1] * ds / d[1], am[2] * ds / d[2]};
    int dom;
    if (dsa[0] <= dsa[1] && dsa[0] <= dsa[2]) dom = 0;
    else if (dsa[1] <= dsa[0] && dsa[1] <= dsa[2]) dom = 1;
    else dom = 2;
    const double mus[3] = {mx, my, mz};
    int beg[3], end_[3], t[3];
    for (int a = 0; a < 3; ++a) {
      if (mus[a] >= 0.0) {
        beg[a] = S.b0[a] + sts[a];
        end_[a] = S.b1[a];
        t[a] = 1;
        S.pos_mask[a] |= 1u << n;
      } else {
        beg[a] = S.b1[a] - sts[a];
        end_[a] = S.b0[a];
        t[a] = -1;
      }
    }
    const int b = dom == 0 ? 1 : 0, c = dom == 2 ? 1 : 2;
    RtRay &R = S.rays.r[n];
    R.tr = (rt->x_rays_yfast && dom == 0 && sizes[0] > 1 && sizes[1] > 1) ? 1 : 0;
    if (R.tr) S.tr_mask |= 1u << n;
    const long *strides = R.tr ? strides_y : strides_x;
    const int lo_b = std::min(beg[b], end_[b]), lo_c = std::min(beg[c], end_[c]);
    int idx[3];
    idx[dom] = beg[dom];
    idx[b] = lo_b;
    idx[c] = lo_c;
    R.start = (int)(idx[0] * strides[0] + idx[1] * strides[1] + idx[2] * strides[2]);
    R.da = (int)(t[dom] * strides[dom]);
    R.db = (int)strides[b];
    R.dc = (int)strides[c];
    R.ua = (int)(t[dom] * sts[dom] * strides[dom]);
    R.ub = (int)(t[b] * sts[b] * strides[b]);
    R.uc = (int)(t[c] * sts[c] * strides[c]);
    R.nb = std::abs(end_[b] - beg[b]) + 1;
    R.nc = std::abs(end_[c] - beg[c]) + 1;
    R.sb = t[b] * sts[b];
    R.sc = t[c] * sts[c];
    R.nplanes = std::abs(end_[dom] - beg[dom]) + 1;
    const double cx = cc[b], cy = cc[c];
    R.w00 = (1.0 - cx) * (1.0 - cy);
    R.w10 = cx * (1.0 - cy);
    R.w01 = (1.0 - cx) * cy;
    R.w11 = cx * cy;
    R.ds = ds;
    R.wq = S.w[n];
    R.id = n;
    S.max_nodes = std::max(S.max_nodes, R.nb * R.nc);
    S.max_planes = std::max(S.max_planes, R.nplanes);
  }
  return MISOHIP_OK;
}

// (lo, cnt, It, Jt, axis) of the y-fastest frame: x and y swapped
inline void swap_xy(const int in[3], int out[3]) {
  out[0] = in[1];
  out[1] = in[0];
  out[2] = in[2];
}
inline int frame_axis(int a, bool y) { return y && a < 2 ? 1 - a : a; }

int check_rt(const misohip_grid *g, const misohip_rt_config *rt) {
  if (!g || !rt) return MISOHIP_E_ARG;
  if (rt->num_rays != 24 && rt->num_rays != 1) return MISOHIP_E_ARG;
  if (rt->max_iters < 1) return MISOHIP_E_ARG;
  if (rt->sweep_planes < 0 || rt->sweep_planes > RT_MULTI_PLANES) return MISOHIP_E_ARG;
  const double n = (double)g->i_total * g->j_total * g->k_total;
  if (n >= 2147483647.0) return MISOHIP_E_RANGE;  // int indices within one ray
  if (!g->dx || !g->dy || !g->dz) return MISOHIP_E_ARG;
  return MISOHIP_OK;
}

}  // namespace

extern "C" {

int misohip_rt_quadrature(int num_rays, double *w, double *mu_x, double *mu_y, double *mu_z) {
  if (!w || !mu_x || !mu_y || !mu_z) return MISOHIP_E_ARG;
  return quadrature(num_rays, w, mu_x, mu_y, mu_z);
}

int misohip_rt_solve(misohip_ctx *ctx, const misohip_grid *g, const misohip_rt_config *rt,
                     const void *kappa, const void *S, void *rint, void *rint_old, int *iters,
                     double *max_diff) {
  if (!ctx || !kappa || !S || !rint || !rint_old) return MISOHIP_E_ARG;
  if (int rc = check_rt(g, rt)) return rc;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  RtSetup su;
  if (int rc = setup(ctx, g, rt, su)) return rc;
  double *I = static_cast<double *>(rint);
  double *old = static_cast<double *>(rint_old);
  const double *kap = static_cast<const double *>(kappa);
  const double *src = static_cast<const double *>(S);
  const int It = g->i_total, Jt = g->j_total, Kt = g->k_total;
  int lo[3], cnt[3], lo_y[3], cnt_y[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = su.b0[a];
    cnt[a] = su.b1[a] - su.b0[a] + 1;
  }
  swap_xy(lo, lo_y);
  swap_xy(cnt, cnt_y);
  const unsigned all = su.nrays >= 32 ? ~0u : ((1u << su.nrays) - 1u);
  const unsigned x_mask = all & ~su.tr_mask;
  const double *kap_y = nullptr, *src_y = nullptr;
  if (su.tr_mask) {
    const size_t cells = (size_t)su.ray_stride;
    if (ctx->rt_cells < cells) {
      if (ctx->rt_kap_y) MISOHIP_TRY(hipFree(ctx->rt_kap_y));
      if (ctx->rt_src_y) MISOHIP_TRY(hipFree(ctx->rt_src_y));
      ctx->rt_kap_y = ctx->rt_src_y = nullptr;
      ctx->rt_cells = 0;
      MISOHIP_TRY(hipMalloc(&ctx->rt_kap_y, sizeof(double) * cells));
      MISOHIP_TRY(hipMalloc(&ctx->rt_src_y, sizeof(double) * cells));
      ctx->rt_cells = cells;
    }
    Timed t_(ctx, MISOHIP_KIND_RT);
    MISOHIP_TRY(launch_rt_transpose_xy(ctx->stream, kap, ctx->rt_kap_y, It, Jt, Kt));
    MISOHIP_TRY(launch_rt_transpose_xy(ctx->stream, src, ctx->rt_src_y, It, Jt, Kt));
    kap_y = ctx->rt_kap_y;
    src_y = ctx->rt_src_y;
  }
  // Rays with a short c axis advance several planes per launch
  // (rt_sweep_multi_kernel); the others go one plane per launch.
  const int P = rt->sweep_planes;
  RtRays mrays;
  int nm = 0, m_planes = 0, m_nb = 0;
  std::vector<int> plane_rays;
  for (int n = 0; n < su.nrays; ++n) {
    const RtRay &R = su.rays.r[n];
    if (P >= 2 && R.sb != 0 && R.sc != 0 && R.ua == R.da && R.nc + 1 <= RT_MULTI_ROWS) {
      mrays.r[nm++] = R;
      m_planes = std::max(m_planes, R.nplanes);
      m_nb = std::max(m_nb, R.nb);
    } else {
      plane_rays.push_back(n);
    }
  }
  // Sweep phases: plane ranges over which the set of rays that still have
  // planes is constant; each phase launches only those rays, sized for
  // their largest plane (a thin slab's z rays end long before its x/y rays).
  struct Phase {
    int p0, p1, n, max_nodes;
    RtRays rays;
  };
  std::vector<Phase> phases;
  {
    std::vector<int> ends;
    for (int n : plane_rays) ends.push_back(su.rays.r[n].nplanes);
    std::sort(ends.begin(), ends.end());
    int p0 = 0;
    for (int e : ends) {
      if (e <= p0) continue;
      Phase ph;
      ph.p0 = p0;
      ph.p1 = e;
      ph.n = 0;
      ph.max_nodes = 0;
      for (int n : plane_rays)
        if (su.rays.r[n].nplanes > p0) {
          ph.rays.r[ph.n++] = su.rays.r[n];
          ph.max_nodes = std::max(ph.max_nodes, su.rays.r[n].nb * su.rays.r[n].nc);
        }
      phases.push_back(ph);
      p0 = e;
    }
  }
  RtFaces uf;  // the rays' upwind faces
  for (int a = 0; a < 3; ++a) {
    uf.lo[a] = lo[a];
    uf.n[a] = cnt[a];
    uf.pos[a] = su.pos_mask[a];
  }
  uf.tr = su.tr_mask;
  uf.It = It;
  uf.Jt = Jt;
  if (!ctx->rt_parts) MISOHIP_TRY(hipMalloc(&ctx->rt_parts, sizeof(unsigned) * 32 * RT_PARTS));
  // The multi-plane rays and the plane rays touch disjoint intensities and
  // meet only in the change maximum (atomic), so with both present the
  // multi-plane sweeps run on a second stream beside the plane sweeps.
  const bool split = nm > 0 && !phases.empty();
  if (split && !ctx->rt_side) {
    MISOHIP_TRY(hipStreamCreateWithFlags(&ctx->rt_side, hipStreamNonBlocking));
    MISOHIP_TRY(hipEventCreateWithFlags(&ctx->rt_fork, hipEventDisableTiming));
    MISOHIP_TRY(hipEventCreateWithFlags(&ctx->rt_join, hipEventDisableTiming));
  }
  hipStream_t mst = split ? ctx->rt_side : ctx->stream;
  unsigned *parts = ctx->rt_parts;
  MISOHIP_TRY(launch_rt_parts_reduce(ctx->stream, parts, nullptr, true));
  int it = 0;
  double md = 0.0;
  for (it = 0; it < rt->max_iters; ++it) {
    {
      Timed t_(ctx, MISOHIP_KIND_RT);
      MISOHIP_TRY(launch_rt_faces(ctx->stream, I, old, true, uf, su.ray_stride, su.nrays, parts));
      for (int a = 2; a >= 0; --a)
        if (rt->periodic[a] && su.size[a] > 1) {
          if (x_mask)
            MISOHIP_TRY(launch_rt_halo(ctx->stream, I, a, su.b0[a], su.b1[a], It, Jt, Kt,
                                       su.pos_mask[a], x_mask, su.ray_stride, su.nrays));
          if (su.tr_mask)
            MISOHIP_TRY(launch_rt_halo(ctx->stream, I, frame_axis(a, true), su.b0[a], su.b1[a],
                                       Jt, It, Kt, su.pos_mask[a], su.tr_mask, su.ray_stride,
                                       su.nrays));
        }
      for (int a = 0; a < 3; ++a)
        for (int side = 0; side < 2; ++side) {
          const int kind = rt->face_kind[2 * a + side];
          if (rt->periodic[a] || su.size[a] <= 1 || kind == 0) continue;
          unsigned in_mask = 0u;
          for (int n = 0; n < su.nrays; ++n) {
            const double mu = su.mu[a][n];
            if (side == 0 ? mu > 0.0 : mu < 0.0) in_mask |= 1u << n;
          }
          for (int y = 0; y < 2; ++y) {
            const unsigned m = in_mask & (y ? su.tr_mask : x_mask);
            if (!m) continue;
            const int fa = frame_axis(a, y);
            const int t0 = fa == 0 ? 1 : 0, t1 = fa == 2 ? 1 : 2;
            const int *fl = y ? lo_y : lo, *fc = y ? cnt_y : cnt;
            MISOHIP_TRY(launch_rt_face(ctx->stream, I, y ? src_y : src, fa,
                                       side == 0 ? su.b0[a] : su.b1[a], fl[t0], fl[t1], fc[t0],
                                       fc[t1], y ? Jt : It, y ? It : Jt, kind,
                                       rt->face_value[2 * a + side], m, su.ray_stride,
                                       su.nrays));
          }
        }
      if (split) {
        MISOHIP_TRY(hipEventRecord(ctx->rt_fork, ctx->stream));
        MISOHIP_TRY(hipStreamWaitEvent(ctx->rt_side, ctx->rt_fork, 0));
      }
      for (int p0 = 0; p0 < m_planes; p0 += P)
        MISOHIP_TRY(launch_rt_sweep_multi(mst, I, kap, src, kap_y, src_y, mrays, nm, p0, P,
                                          su.ray_stride, m_nb, parts));
      for (const Phase &ph : phases)
        for (int p = ph.p0; p < ph.p1; ++p)
          MISOHIP_TRY(launch_rt_sweep_plane(ctx->stream, I, kap, src, kap_y, src_y, ph.rays,
                                            ph.n, p, su.ray_stride, ph.max_nodes, parts));
      if (split) {
        MISOHIP_TRY(hipEventRecord(ctx->rt_join, ctx->rt_side));
        MISOHIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->rt_join, 0));
      }
      MISOHIP_TRY(launch_rt_faces(ctx->stream, I, old, false, uf, su.ray_stride, su.nrays, parts));
      MISOHIP_TRY(launch_rt_parts_reduce(ctx->stream, parts, ctx->d_slot + 1, false));
    }
    MISOHIP_TRY(hipMemcpyAsync(ctx->h_slot + 1, ctx->d_slot + 1, sizeof(unsigned),
                               hipMemcpyDeviceToHost, ctx->stream));
    MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
    float f;
    std::memcpy(&f, ctx->h_slot + 1, sizeof(float));
    md = (double)f;
    if (md < rt->tolerance) {
      ++it;
      break;
    }
  }
  if (iters) *iters = it;
  if (max_diff) *max_diff = md;
  return MISOHIP_OK;
}

int misohip_rt_qrad(misohip_ctx *ctx, const misohip_grid *g, const misohip_rt_config *rt,
                    const void *kappa, const void *S, const void *rint, void *qrad) {
  if (!ctx || !kappa || !S || !rint || !qrad) return MISOHIP_E_ARG;
  if (int rc = check_rt(g, rt)) return rc;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  RtSetup su;
  if (int rc = setup(ctx, g, rt, su)) return rc;
  int lo[3], cnt[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = su.b0[a];
    cnt[a] = su.b1[a] - su.b0[a] + 1;
  }
  Timed t_(ctx, MISOHIP_KIND_RT);
  MISOHIP_TRY(hipMemsetAsync(qrad, 0, sizeof(double) * su.ray_stride, ctx->stream));
  MISOHIP_TRY(launch_rt_qrad(ctx->stream, static_cast<const double *>(rint),
                             static_cast<const double *>(kappa), static_cast<const double *>(S),
                             static_cast<double *>(qrad), lo, cnt, g->i_total, g->j_total,
                             su.ray_stride, su.nrays, su.tr_mask, su.rays));
  return MISOHIP_OK;
}

int misohip_rt_ray_layout(misohip_ctx *ctx, const misohip_grid *g, const misohip_rt_config *rt,
                          unsigned *yfast_mask) {
  if (!ctx || !yfast_mask) return MISOHIP_E_ARG;
  if (int rc = check_rt(g, rt)) return rc;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  RtSetup su;
  if (int rc = setup(ctx, g, rt, su)) return rc;
  *yfast_mask = su.tr_mask;
  return MISOHIP_OK;
}

int misohip_rt_source(misohip_ctx *ctx, const misohip_grid *g, const misohip_fields *qq,
                      double gm, double kappa0, double s_rad, void *kappa, void *S) {
  if (!ctx || !g || !qq || !kappa || !S || !qq->q[0] || !qq->q[7]) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  const long n = (long)g->i_total * g->j_total * g->k_total;
  Timed t_(ctx, MISOHIP_KIND_RT);
  MISOHIP_TRY(launch_rt_source(ctx->stream, static_cast<const double *>(qq->q[0]),
                               static_cast<const double *>(qq->q[7]), static_cast<double *>(kappa),
                               static_cast<double *>(S), n, gm, kappa0, s_rad));
  return MISOHIP_OK;
}

int misohip_rt_heat(misohip_ctx *ctx, const misohip_grid *g, const misohip_fields *qq,
                    const void *qrad, double dt) {
  if (!ctx || !g || !qq || !qrad || !qq->q[0] || !qq->q[7]) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  Shape sh;
  sh.It = g->i_total; sh.Jt = g->j_total; sh.Kt = g->k_total;
  sh.mi = g->i_margin; sh.mj = g->j_margin; sh.mk = g->k_margin;
  sh.is = g->is; sh.js = g->js; sh.ks = g->ks;
  Timed t_(ctx, MISOHIP_KIND_RT);
  MISOHIP_TRY(launch_rt_heat(ctx->stream, static_cast<const double *>(qq->q[0]),
                             static_cast<double *>(qq->q[7]), static_cast<const double *>(qrad),
                             sh, dt));
  return MISOHIP_OK;
}

}  // extern "C"
