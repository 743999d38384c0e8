// Directional artificial-viscosity sweep of MISO
// (ArtificialViscosity::update, miso/include/miso/mhd_artificial_viscosity_host.hpp:70-348;
// CUDA: mhd_artificial_viscosity_cuda.cuh:41-574 uses nine kernels per direction).
//
// One fused kernel updates all nine fields of a cell for one direction.  The
// characteristic speed cc is frozen for the whole AV step: it is evaluated
// from the post-RK state (characteristic_velocity_eval, :50-68) and read by
// all three sweeps (mhd_integrator_host.hpp:284-299).  The x sweep computes it
// in-kernel at i-1, i, i+1 and stores it for the y and z sweeps, over the
// interior plus a one-cell ring in y and z (the only cells those sweeps read).
#include "launch.hpp"

namespace misohip {

// cc = cs*cs_fac + vv*vv_fac + ca*ca_fac with cs = sqrt(gm*(gm-1)*ei)
// (eos.hpp:35-43, mhd_artificial_viscosity_host.hpp:57-64)
template <typename R>
__device__ __forceinline__ R char_speed(const Fld<const R> &q, int c, R gm, R cs_fac, R ca_fac,
                                        R vv_fac) {
  const R ro = q.q[0][c], vx = q.q[1][c], vy = q.q[2][c], vz = q.q[3][c];
  const R bx = q.q[4][c], by = q.q[5][c], bz = q.q[6][c], ei = q.q[7][c];
  const R cs = dsqrt(gm * (gm - R(1)) * ei);
  const R vv = dsqrt(vx * vx + vy * vy + vz * vz);
  const R ca = dsqrt((bx * bx + by * by + bz * bz) / ro * Const<R>::pii4);
  return cs * cs_fac + vv * vv_fac + ca * ca_fac;
}

template <typename R>
__device__ __forceinline__ void flux_pair(const R v[5], R cdw, R cup, R ep, R fh, R &f_dw,
                                          R &f_up) {
  const R d_dw = dqq_eval(v[0], v[1], v[2], ep);
  const R d_cn = dqq_eval(v[1], v[2], v[3], ep);
  const R d_up = dqq_eval(v[2], v[3], v[4], ep);
  f_dw = flux_core(v[1], v[2], d_dw, d_cn, cdw, fh);
  f_up = flux_core(v[2], v[3], d_cn, d_up, cup, fh);
}

// DIR: sweep direction.  The launch covers interior x-columns; for DIR == 0
// it also covers the y/z ring where only cc is produced.
template <typename R, int DIR>
__global__ __launch_bounds__(256) void av_sweep_kernel(Fld<const R> qq, Fld<R> rslt,
                                                       R *__restrict__ cc_store,
                                                       const R *__restrict__ cc_in,
                                                       const R *__restrict__ dd_arr, Shape sh,
                                                       int j_lo, int k_lo, int j_n, int k_n,
                                                       AvScalars sc, int nbx, int nby, int nb) {
  const int V = xcd_virtual_block(blockIdx.x, nb);
  const int blk_x = V % nbx, rest = V / nbx;
  const int i = blk_x * 64 + (threadIdx.x & 63) + sh.mi;
  const int j = (rest % nby) * 4 + (threadIdx.x >> 6) + j_lo;
  const int k = rest / nby + k_lo;
  if (i >= sh.It - sh.mi || j >= j_lo + j_n || k >= k_lo + k_n) return;
  const int c = sh.idx(i, j, k);
  const R gm = R(sc.gm), cs_fac = R(sc.cs_fac), ca_fac = R(sc.ca_fac), vv_fac = R(sc.vv_fac);
  const int s = sh.stride(DIR);
  R ccl, ccc, ccr;
  if (DIR == 0) {
    ccc = char_speed(qq, c, gm, cs_fac, ca_fac, vv_fac);
    cc_store[c] = ccc;
    const bool ring = (j < sh.mj) || (j >= sh.Jt - sh.mj) || (k < sh.mk) || (k >= sh.Kt - sh.mk);
    if (ring) return;
    ccl = s ? char_speed(qq, c - s, gm, cs_fac, ca_fac, vv_fac) : ccc;
    ccr = s ? char_speed(qq, c + s, gm, cs_fac, ca_fac, vv_fac) : ccc;
  } else {
    ccl = cc_in[c - s];
    ccc = cc_in[c];
    ccr = cc_in[c + s];
  }
  const R cdw = R(0.5) * (ccl + ccc), cup = R(0.5) * (ccc + ccr);
  const R ep = R(sc.ep), fh = R(sc.fh), dt = R(sc.dt);
  const int pos = DIR == 0 ? i : (DIR == 1 ? j : k);
  const R dd = dd_arr[s ? pos : 0];

  R ro5[5], v[5], fdw, fup;
#pragma unroll
  for (int p = 0; p < 5; ++p) ro5[p] = qq.q[0][c + (p - 2) * s];

  // density
  flux_pair(ro5, cdw, cup, ep, fh, fdw, fup);
  const R ro_n = ro5[2] - (fup - fdw) * dd * dt;
  // momenta (flux of ro*v, then v = (ro v)' / ro')
  R vn[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
#pragma unroll
    for (int p = 0; p < 5; ++p) v[p] = ro5[p] * qq.q[1 + m][c + (p - 2) * s];
    flux_pair(v, cdw, cup, ep, fh, fdw, fup);
    vn[m] = (v[2] - (fup - fdw) * dd * dt) / ro_n;
  }
  // magnetic field and psi
  R bn[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int f = m < 3 ? 4 + m : 8;
#pragma unroll
    for (int p = 0; p < 5; ++p) v[p] = qq.q[f][c + (p - 2) * s];
    flux_pair(v, cdw, cup, ep, fh, fdw, fup);
    bn[m] = v[2] - (fup - fdw) * dd * dt;
  }
  // internal energy: flux of ro*ei applied to the total energy
#pragma unroll
  for (int p = 0; p < 5; ++p) v[p] = ro5[p] * qq.q[7][c + (p - 2) * s];
  flux_pair(v, cdw, cup, ep, fh, fdw, fup);
  const R ro = ro5[2], vx = qq.q[1][c], vy = qq.q[2][c], vz = qq.q[3][c];
  const R bx = qq.q[4][c], by = qq.q[5][c], bz = qq.q[6][c], ei = qq.q[7][c];
  const double pii8 = double(Const<R>::pii8);
  const double Et = double(ro) * ei + 0.5 * double(ro) * (double(vx) * vx + double(vy) * vy + double(vz) * vz) +
                    pii8 * (double(bx) * bx + double(by) * by + double(bz) * bz);
  const double dro = ro_n, dvx = vn[0], dvy = vn[1], dvz = vn[2];
  const double dbx = bn[0], dby = bn[1], dbz = bn[2];
  const double ein = (Et - double((fup - fdw) * dd * dt) - 0.5 * dro * (dvx * dvx + dvy * dvy + dvz * dvz) -
                      pii8 * (dbx * dbx + dby * dby + dbz * dbz)) /
                     dro;

  rslt.q[0][c] = ro_n;
  rslt.q[1][c] = vn[0];
  rslt.q[2][c] = vn[1];
  rslt.q[3][c] = vn[2];
  rslt.q[4][c] = bn[0];
  rslt.q[5][c] = bn[1];
  rslt.q[6][c] = bn[2];
  rslt.q[7][c] = R(ein);
  rslt.q[8][c] = bn[3];
}

// Characteristic speed over every cell (misohip_char_vel, kept for the
// per-function parity tests; the step computes it inside the x sweep).
template <typename R>
__global__ __launch_bounds__(256) void char_vel_kernel(Fld<const R> qq, R *__restrict__ cc, int n,
                                                       AvScalars sc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  cc[c] = char_speed(qq, c, R(sc.gm), R(sc.cs_fac), R(sc.ca_fac), R(sc.vv_fac));
}

template <typename R, int DIR>
static hipError_t launch_av_t(hipStream_t st, const Fld<const R> &q, const Fld<R> &r, R *cc_store,
                              const R *cc_in, const R *dd, const Shape &sh, const AvScalars &sc) {
  int j_lo = sh.mj, j_n = sh.Jt - 2 * sh.mj, k_lo = sh.mk, k_n = sh.Kt - 2 * sh.mk;
  if (DIR == 0) {
    // ring of one cell in y and z for the frozen cc
    j_lo -= sh.js; j_n += 2 * sh.js;
    k_lo -= sh.ks; k_n += 2 * sh.ks;
  }
  const int ni = sh.It - 2 * sh.mi;
  const int nbx = (ni + 63) / 64, nby = (j_n + 3) / 4;
  const long nbl = (long)nbx * nby * k_n;
  if (nbl > 2147483647L) return hipErrorInvalidValue;
  const int nb = (int)nbl;
  hipLaunchKernelGGL((av_sweep_kernel<R, DIR>), dim3(nb), dim3(256), 0, st, q, r, cc_store, cc_in,
                     dd, sh, j_lo, k_lo, j_n, k_n, sc, nbx, nby, nb);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_av_sweep(hipStream_t st, int dir, const Fld<const R> &q, const Fld<R> &r,
                           R *cc_store, const R *cc_in, const R *dd, const Shape &sh,
                           const AvScalars &sc) {
  if (dir == 0) return launch_av_t<R, 0>(st, q, r, cc_store, cc_in, dd, sh, sc);
  if (dir == 1) return launch_av_t<R, 1>(st, q, r, cc_store, cc_in, dd, sh, sc);
  return launch_av_t<R, 2>(st, q, r, cc_store, cc_in, dd, sh, sc);
}

template <typename R>
hipError_t launch_char_vel(hipStream_t st, const Fld<const R> &q, R *cc, int n, const AvScalars &sc) {
  hipLaunchKernelGGL((char_vel_kernel<R>), dim3((n + 255) / 256), dim3(256), 0, st, q, cc, n, sc);
  return hipGetLastError();
}

template hipError_t launch_av_sweep<double>(hipStream_t, int, const Fld<const double> &,
                                            const Fld<double> &, double *, const double *,
                                            const double *, const Shape &, const AvScalars &);
template hipError_t launch_av_sweep<float>(hipStream_t, int, const Fld<const float> &,
                                           const Fld<float> &, float *, const float *,
                                           const float *, const Shape &, const AvScalars &);
template hipError_t launch_char_vel<double>(hipStream_t, const Fld<const double> &, double *, int,
                                            const AvScalars &);
template hipError_t launch_char_vel<float>(hipStream_t, const Fld<const float> &, float *, int,
                                           const AvScalars &);

}  // namespace misohip
