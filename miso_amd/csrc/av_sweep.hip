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
__device__ __forceinline__ R char_speed_v(const R f[9], R gm, R cs_fac, R ca_fac, R vv_fac) {
  const R ro = f[0], vx = f[1], vy = f[2], vz = f[3];
  const R bx = f[4], by = f[5], bz = f[6], ei = f[7];
  const R cs = dsqrt(gm * (gm - R(1)) * ei);
  const R vv = dsqrt(vx * vx + vy * vy + vz * vz);
  const R ca = dsqrt((bx * bx + by * by + bz * bz) / ro * Const<R>::pii4);
  return cs * cs_fac + vv * vv_fac + ca * ca_fac;
}

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
__device__ __forceinline__ void av_generic_block(const Fld<const R> &qq, const Fld<R> &rslt,
                                                 R *__restrict__ cc_store,
                                                 const R *__restrict__ cc_in,
                                                 const R *__restrict__ dd_arr, const Shape &sh,
                                                 int j_lo, int k_lo, int j_n, int k_n,
                                                 const AvScalars &sc, const R dt, int nbx,
                                                 int nby, int V) {
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
  const R ep = R(sc.ep), fh = R(sc.fh);
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

template <typename R, int DIR>
__global__ __launch_bounds__(256) void av_sweep_kernel(Fld<const R> qq, Fld<R> rslt,
                                                       R *__restrict__ cc_store,
                                                       const R *__restrict__ cc_in,
                                                       const R *__restrict__ dd_arr, Shape sh,
                                                       int j_lo, int k_lo, int j_n, int k_n,
                                                       AvScalars sc, int nbx, int nby, int nb) {
  av_generic_block<R, DIR>(qq, rslt, cc_store, cc_in, dd_arr, sh, j_lo, k_lo, j_n, k_n, sc,
                           R(sc.dt), nbx, nby, xcd_virtual_block(blockIdx.x, nb));
}

// The sweep along a direction of extent one (the reference runs it, with
// stride 0: every stencil point is the cell itself) evaluated on one cell's
// values in registers, with exactly the generic kernel's arithmetic; dd is
// that direction's inverse spacing (0).  in/out: ro vx vy vz bx by bz ei ph.
template <typename R>
__device__ __forceinline__ void degenerate_point(const R in[9], R cc, R ep, R fh, R dd, R dt,
                                                 R out[9]) {
  const R cdw = R(0.5) * (cc + cc), cup = R(0.5) * (cc + cc);
  R v[5], fdw, fup;
  const R ro = in[0];
#pragma unroll
  for (int p = 0; p < 5; ++p) v[p] = ro;
  flux_pair(v, cdw, cup, ep, fh, fdw, fup);
  const R ro_n = ro - (fup - fdw) * dd * dt;
  R vn[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
#pragma unroll
    for (int p = 0; p < 5; ++p) v[p] = ro * in[1 + m];
    flux_pair(v, cdw, cup, ep, fh, fdw, fup);
    vn[m] = (v[2] - (fup - fdw) * dd * dt) / ro_n;
  }
  R bn[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int f = m < 3 ? 4 + m : 8;
#pragma unroll
    for (int p = 0; p < 5; ++p) v[p] = in[f];
    flux_pair(v, cdw, cup, ep, fh, fdw, fup);
    bn[m] = v[2] - (fup - fdw) * dd * dt;
  }
#pragma unroll
  for (int p = 0; p < 5; ++p) v[p] = ro * in[7];
  flux_pair(v, cdw, cup, ep, fh, fdw, fup);
  const R vx = in[1], vy = in[2], vz = in[3], bx = in[4], by = in[5], bz = in[6], ei = in[7];
  const double pii8 = double(Const<R>::pii8);
  const double Et = double(ro) * ei + 0.5 * double(ro) * (double(vx) * vx + double(vy) * vy + double(vz) * vz) +
                    pii8 * (double(bx) * bx + double(by) * by + double(bz) * bz);
  const double dro = ro_n, dvx = vn[0], dvy = vn[1], dvz = vn[2];
  const double dbx = bn[0], dby = bn[1], dbz = bn[2];
  const double ein = (Et - double((fup - fdw) * dd * dt) - 0.5 * dro * (dvx * dvx + dvy * dvy + dvz * dvz) -
                      pii8 * (dbx * dbx + dby * dby + dbz * dbz)) /
                     dro;
  out[0] = ro_n; out[1] = vn[0]; out[2] = vn[1]; out[3] = vn[2];
  out[4] = bn[0]; out[5] = bn[1]; out[6] = bn[2]; out[7] = R(ein); out[8] = bn[3];
}

// ---------------------------------------------------------------------------
// y and z sweeps that walk along the sweep axis.  Lanes run along x
// (coalesced); each thread walks `chunk` cells along DIR keeping a window of
// the conserved values, so the slope at every point and the flux at every
// interface are computed once and reused by both cells that share them (the
// reference recomputes them three and two times with identical operands, so
// the reuse is exact).  Consecutive blocks are consecutive chunks of the
// same column, so chunk halos are L2 hits.
template <typename R> struct AvCons {
  R v[9];  // ro, ro*vx, ro*vy, ro*vz, bx, by, bz, ph, ro*ei
  R vx, vy, vz, ei;
};

// conserved values of a cell from its nine fields (ro vx vy vz bx by bz ei ph)
template <typename R>
__device__ __forceinline__ void cons_of(const R f[9], AvCons<R> &a) {
  const R ro = f[0];
  a.vx = f[1];
  a.vy = f[2];
  a.vz = f[3];
  a.ei = f[7];
  a.v[0] = ro;
  a.v[1] = ro * a.vx;
  a.v[2] = ro * a.vy;
  a.v[3] = ro * a.vz;
  a.v[4] = f[4];
  a.v[5] = f[5];
  a.v[6] = f[6];
  a.v[7] = f[8];
  a.v[8] = ro * a.ei;
}

template <typename R>
__device__ __forceinline__ void load_cons(const Fld<const R> &q, int c, AvCons<R> &a) {
  R f[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) f[m] = q.q[m][c];
  cons_of(f, a);
}

// conservative update of one cell from its two interface fluxes
// (mhd_artificial_viscosity_host.hpp:131-344)
// DEGEN: apply the sweep of a following direction of extent one (2D runs:
// z) to the result before storing; dz0 is its inverse spacing, cc the cell's
// frozen speed.
// the nine updated fields of one cell from its two interface fluxes
template <typename R>
__device__ __forceinline__ void av_cell_values(const AvCons<R> &a, const R fdw[9], const R fup[9],
                                              R dd, R dt, R o[9]) {
  const R ro_n = a.v[0] - (fup[0] - fdw[0]) * dd * dt;
  const R vxn = (a.v[1] - (fup[1] - fdw[1]) * dd * dt) / ro_n;
  const R vyn = (a.v[2] - (fup[2] - fdw[2]) * dd * dt) / ro_n;
  const R vzn = (a.v[3] - (fup[3] - fdw[3]) * dd * dt) / ro_n;
  const R bxn = a.v[4] - (fup[4] - fdw[4]) * dd * dt;
  const R byn = a.v[5] - (fup[5] - fdw[5]) * dd * dt;
  const R bzn = a.v[6] - (fup[6] - fdw[6]) * dd * dt;
  const R phn = a.v[7] - (fup[7] - fdw[7]) * dd * dt;
  const double pii8 = double(Const<R>::pii8);
  const double ro = a.v[0];
  const double Et = ro * double(a.ei) +
                    0.5 * ro * (double(a.vx) * a.vx + double(a.vy) * a.vy + double(a.vz) * a.vz) +
                    pii8 * (double(a.v[4]) * a.v[4] + double(a.v[5]) * a.v[5] +
                            double(a.v[6]) * a.v[6]);
  const double dro = ro_n, dvx = vxn, dvy = vyn, dvz = vzn, dbx = bxn, dby = byn, dbz = bzn;
  const double ein = (Et - double((fup[8] - fdw[8]) * dd * dt) -
                      0.5 * dro * (dvx * dvx + dvy * dvy + dvz * dvz) -
                      pii8 * (dbx * dbx + dby * dby + dbz * dbz)) /
                     dro;
  o[0] = ro_n; o[1] = vxn; o[2] = vyn; o[3] = vzn; o[4] = bxn;
  o[5] = byn; o[6] = bzn; o[7] = R(ein); o[8] = phn;
}

template <typename R, bool DEGEN = false>
__device__ __forceinline__ void av_cell_update(const AvCons<R> &a, const R fdw[9], const R fup[9],
                                              R dd, R dt, const Fld<R> &out, int c, R cc = R(0),
                                              R ep = R(0), R fh = R(0), R dz0 = R(0),
                                              R *stored = nullptr) {
  const R ro_n = a.v[0] - (fup[0] - fdw[0]) * dd * dt;
  const R vxn = (a.v[1] - (fup[1] - fdw[1]) * dd * dt) / ro_n;
  const R vyn = (a.v[2] - (fup[2] - fdw[2]) * dd * dt) / ro_n;
  const R vzn = (a.v[3] - (fup[3] - fdw[3]) * dd * dt) / ro_n;
  const R bxn = a.v[4] - (fup[4] - fdw[4]) * dd * dt;
  const R byn = a.v[5] - (fup[5] - fdw[5]) * dd * dt;
  const R bzn = a.v[6] - (fup[6] - fdw[6]) * dd * dt;
  const R phn = a.v[7] - (fup[7] - fdw[7]) * dd * dt;
  const double pii8 = double(Const<R>::pii8);
  const double ro = a.v[0];
  const double Et = ro * double(a.ei) +
                    0.5 * ro * (double(a.vx) * a.vx + double(a.vy) * a.vy + double(a.vz) * a.vz) +
                    pii8 * (double(a.v[4]) * a.v[4] + double(a.v[5]) * a.v[5] +
                            double(a.v[6]) * a.v[6]);
  const double dro = ro_n, dvx = vxn, dvy = vyn, dvz = vzn, dbx = bxn, dby = byn, dbz = bzn;
  const double ein = (Et - double((fup[8] - fdw[8]) * dd * dt) -
                      0.5 * dro * (dvx * dvx + dvy * dvy + dvz * dvz) -
                      pii8 * (dbx * dbx + dby * dby + dbz * dbz)) /
                     dro;
  if (DEGEN) {
    const R in[9] = {ro_n, vxn, vyn, vzn, bxn, byn, bzn, R(ein), phn};
    R o[9];
    degenerate_point(in, cc, ep, fh, dz0, dt, o);
#pragma unroll
    for (int f = 0; f < 9; ++f) out.q[f][c] = o[f];
    if (stored) {
#pragma unroll
      for (int f = 0; f < 9; ++f) stored[f] = o[f];
    }
    return;
  }
  if (stored) {
    stored[0] = ro_n; stored[1] = vxn; stored[2] = vyn; stored[3] = vzn; stored[4] = bxn;
    stored[5] = byn; stored[6] = bzn; stored[7] = R(ein); stored[8] = phn;
  }
  out.q[0][c] = ro_n;
  out.q[1][c] = vxn;
  out.q[2][c] = vyn;
  out.q[3][c] = vzn;
  out.q[4][c] = bxn;
  out.q[5][c] = byn;
  out.q[6][c] = bzn;
  out.q[7][c] = R(ein);
  out.q[8][c] = phn;
}

// CFLF: also reduce the CFL candidates of the stored results into cf.slot
// (one atomicMin per wave), the next step's dt.
template <typename R, int DIR, bool DEGEN = false, bool CFLF = false>
__device__ __forceinline__ void av_walk_block(const Fld<const R> &qq, const Fld<R> &rslt,
                                              const R *__restrict__ cc,
                                              const R *__restrict__ dd_arr, const Shape &sh,
                                              int chunk, int nchunks, int nbx,
                                              const AvScalars &sc, const R dt, int V,
                                              R dz0 = R(0), const CflFuse<R> *cf = nullptr,
                                              const R *__restrict__ hw = nullptr) {
  const int nchl = sc.nparts ? sc.nparts : nchunks;  // chunks in this launch
  const int ch = sc.part0 + V % nchl, r = V / nchl;
  const int bx = r % nbx, bo = r / nbx;
  const int wx = walk_waves_x(DIR == 1 ? sh.Kt - 2 * sh.mk : sh.Jt - 2 * sh.mj);
  const int wave = threadIdx.x >> 6;
  const int i = (bx * wx + wave % wx) * 64 + (threadIdx.x & 63) + sh.mi;
  const int o = bo * (4 / wx) + wave / wx;  // the other transverse axis
  unsigned best = __float_as_uint(1.e10f);
  int j, k, m_dir, n_dir;
  bool valid = i < sh.It - sh.mi;
  if (DIR == 1) {
    k = o + sh.mk;
    valid = valid && k < sh.Kt - sh.mk;
    m_dir = sh.mj; n_dir = sh.Jt - 2 * sh.mj;
    j = m_dir + ch * chunk;
  } else {
    j = o + sh.mj;
    valid = valid && j < sh.Jt - sh.mj;
    m_dir = sh.mk; n_dir = sh.Kt - 2 * sh.mk;
    k = m_dir + ch * chunk;
  }
  const int p0 = ch * chunk;  // offset along DIR from the first interior cell
  const int len = valid ? min(chunk, n_dir - p0) : 0;
  if (len > 0) {
  const int s = sh.stride(DIR);
  const R ep = R(sc.ep), fh = R(sc.fh);
  int c = sh.idx(i, j, k);
  // the fields of the cell q steps along DIR from c: a ghost cell of a
  // wrapped axis is read at its periodic source (cc is not: the x-y walk
  // stores it on the one-cell ring the walk reads)
  const bool wd = DIR == 2 && (sc.wrap & 4);
  int pd = DIR == 1 ? j : k;  // c's coordinate along DIR, advanced with c
  const int td = DIR == 1 ? sh.Jt : sh.Kt;
  auto fc = [&](int q) { return wd ? c + (wrap_coord(pd + q, td, m_dir, true) - pd) * s : c + q * s; };

  // window: a0 = p, a1 = p+1 (and p+2 loaded per step); slopes at p; flux p-1/2
  AvCons<R> am, a0, a1, a2;
  load_cons(qq, fc(-2), am);
  load_cons(qq, fc(-1), a0);
  load_cons(qq, c, a1);
  load_cons(qq, fc(1), a2);
  R d0[9], fdw[9];
  R cc_m = cc[c - s], cc_0 = cc[c], cc_1 = cc[c + s];
  {
    const R cdw = R(0.5) * (cc_m + cc_0);
#pragma unroll
    for (int f = 0; f < 9; ++f) {
      const R dm = dqq_eval(am.v[f], a0.v[f], a1.v[f], ep);  // slope at p0-1
      d0[f] = dqq_eval(a0.v[f], a1.v[f], a2.v[f], ep);       // slope at p0
      fdw[f] = flux_core(a0.v[f], a1.v[f], dm, d0[f], cdw, fh);
    }
  }
  // shift so that a0 = p0, a1 = p0+1
  a0 = a1;
  a1 = a2;
  // the fields and speed of row p+2 arrive two steps ahead of their use
  // (row p+3 is in flight meanwhile): two waves per SIMD leave the registers
  // for it, and one step of arithmetic does not cover a load from HBM.  The
  // two rows in flight sit in two register sets that the loop, unrolled by
  // two, uses in turn: a copy from one set to the other at the end of a
  // step would wait for the load issued in that step.
  R nb0[9], nb1[9] = {}, cb0, cb1 = R(0);
  {
    const int c2 = fc(2);
#pragma unroll
    for (int m = 0; m < 9; ++m) nb0[m] = qq.q[m][c2];
  }
  cb0 = cc[c + 2 * s];
  if (len > 1) {  // row p0+3 is inside the array when p0+1 is an output cell
    const int c3 = fc(3);
#pragma unroll
    for (int m = 0; m < 9; ++m) nb1[m] = qq.q[m][c3];
    cb1 = cc[c + 3 * s];
  }
  // the cell widths of the CFL candidates along the two other axes (those
  // along DIR, hw, are read through a kernel argument the compiler knows no
  // store aliases: a scalar load, which the waits for the rows in flight
  // do not count)
  R hx = R(0), hy = R(0), hz = R(0);
  if (CFLF) {
    hx = cf->dx[i];
    if (DIR == 2) hy = cf->dy[j];
    if (DIR == 1) hz = cf->dz[k];
  }
  // one output cell: nb, cb hold row p+2, and take row p+4 afterwards
  auto step = [&](R (&nb)[9], R &cb, int t) {
    AvCons<R> an;
    {
      R fr[9];
#pragma unroll
      for (int m = 0; m < 9; ++m) fr[m] = reg_copy(nb[m]);
      cons_of(fr, an);
    }
    const R cc_2 = reg_copy(cb);
    {
      // row p+4 is inside the array when p+2 is an output cell; otherwise
      // the load takes row p+2 again (unused), so that every step issues the
      // same loads: a load under a branch makes the wait for the rows in
      // flight assume it was not issued, and wait for younger loads
      const int q = t + 2 < len ? 4 : 2;
      const int c4 = fc(q);
#pragma unroll
      for (int m = 0; m < 9; ++m) nb[m] = qq.q[m][c4];
      cb = cc[c + q * s];
    }
    const R cup = R(0.5) * (cc_0 + cc_1);
    R fup[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) {
      const R d1 = dqq_eval(a0.v[f], a1.v[f], an.v[f], ep);  // slope at p+1
      fup[f] = flux_core(a0.v[f], a1.v[f], d0[f], d1, cup, fh);
      d0[f] = d1;
    }
    const R dd = dd_arr[__builtin_amdgcn_readfirstlane(m_dir + p0 + t)];
    if (CFLF) {
      R st[9];
      av_cell_update<R, DEGEN>(a0, fdw, fup, dd, dt, rslt, c, cc_0, ep, fh, dz0, st);
      if (cf->heat_q) {  // the heating kernel's expression on the stored values
        const double e = double(st[7]) + cf->heat_dt * cf->heat_q[c] / double(st[0]);
        st[7] = R(e);
        rslt.q[7][c] = st[7];
      }
      const int pw = __builtin_amdgcn_readfirstlane(m_dir + p0 + t);
      const R hmin = nmin3(hx, DIR == 1 ? hw[pw] : hy, DIR == 2 ? hw[pw] : hz);
      const float cand = cfl_candidate_v(st[0], st[1], st[2], st[3], st[4], st[5], st[6], st[7],
                                         R(cf->gm), R(cf->cfl), hmin);
      best = min(best, __float_as_uint(cand));
    } else {
      av_cell_update<R, DEGEN>(a0, fdw, fup, dd, dt, rslt, c, cc_0, ep, fh, dz0);
    }
#pragma unroll
    for (int f = 0; f < 9; ++f) fdw[f] = fup[f];
    a0 = a1;
    a1 = an;
    cc_0 = cc_1;
    cc_1 = cc_2;
    c += s;
    ++pd;
  };
  // the loop is entered with no load in flight: the waits the compiler
  // places inside it hold for every way in, and with rows p0+2 and p0+3
  // still in flight on entry they would wait at every step for the row
  // loaded one step before
  __builtin_amdgcn_s_waitcnt(0);
  int t = 0;
  for (; t + 1 < len; t += 2) {
    step(nb0, cb0, t);
    step(nb1, cb1, t + 1);
  }
  if (t < len) step(nb0, cb0, t);
  }
  if (CFLF) {
    best = wave_umin(best);
    if ((threadIdx.x & 63) == 0) atomicMin(cf->slot, best);
  }
}

template <typename R, int DIR, bool DEGEN>
__global__ __launch_bounds__(256) void av_walk_cfl_kernel(Fld<const R> qq, Fld<R> rslt,
                                                         const R *__restrict__ cc,
                                                         const R *__restrict__ dd_arr, Shape sh,
                                                         int chunk, int nchunks, int nbx, int nb,
                                                         AvScalars sc,
                                                         const R *__restrict__ dz_arr,
                                                         CflFuse<R> cf,
                                                         const R *__restrict__ hw) {
  av_walk_block<R, DIR, DEGEN, true>(qq, rslt, cc, dd_arr, sh, chunk, nchunks, nbx, sc,
                                     R(sc.dt), xcd_virtual_block(blockIdx.x, nb),
                                     DEGEN ? dz_arr[0] : R(0), &cf, hw);
}

template <typename R, int DIR, bool DEGEN>
__global__ __launch_bounds__(256) void av_walk_kernel(Fld<const R> qq, Fld<R> rslt,
                                                     const R *__restrict__ cc,
                                                     const R *__restrict__ dd_arr, Shape sh,
                                                     int chunk, int nchunks, int nbx, int nbo,
                                                     int nb, AvScalars sc,
                                                     const R *__restrict__ dz_arr) {
  (void)nbo;
  av_walk_block<R, DIR, DEGEN>(qq, rslt, cc, dd_arr, sh, chunk, nchunks, nbx, sc, R(sc.dt),
                               xcd_virtual_block(blockIdx.x, nb), DEGEN ? dz_arr[0] : R(0));
}

// ---------------------------------------------------------------------------
// x sweep with lane shuffles: a wave covers x cells [i0-2, i0+62) and
// outputs [i0, i0+60); every lane evaluates cc, the slope and the right
// interface flux at its own point once and takes its neighbours' from the
// adjacent lanes.  It also stores the frozen cc for the y and z sweeps over
// the interior plus a one-cell ring in y and z.
// Lane shifts: dpp_move, lane_up, lane_dn (device_common.hpp).

template <typename R>
__device__ __forceinline__ void av_xshfl_block(const Fld<const R> &qq, const Fld<R> &rslt,
                                               R *__restrict__ cc_store,
                                               const R *__restrict__ dxi, const Shape &sh,
                                               int j_lo, int k_lo, int j_n, int k_n, int nseg,
                                               const AvScalars &sc, const R dt, int V) {
  const int seg = V % nseg, r = V / nseg;
  const int lane = threadIdx.x & 63;
  const int row = r * 4 + (threadIdx.x >> 6);  // (j, k) row, j fastest
  if (row >= j_n * k_n) return;  // whole wave leaves together
  const int j = j_lo + row % j_n, k = k_lo + row / j_n;
  const int i0 = sh.mi + seg * 60;
  const int i = i0 - 2 + lane;
  const int ic = min(i, sh.It - 1);
  const int c = sh.idx(ic, j, k);
  const R gm = R(sc.gm);
  const R ccv = char_speed(qq, c, gm, R(sc.cs_fac), R(sc.ca_fac), R(sc.vv_fac));
  const bool out_lane = lane >= 2 && lane < 62 && i < sh.It - sh.mi;
  if (out_lane) cc_store[c] = ccv;
  const bool ring = (j < sh.mj) || (j >= sh.Jt - sh.mj) || (k < sh.mk) || (k >= sh.Kt - sh.mk);
  if (ring) return;  // uniform per wave
  AvCons<R> a;
  load_cons(qq, c, a);
  const R ep = R(sc.ep), fh = R(sc.fh);
  const R cup = R(0.5) * (ccv + lane_dn(ccv));
  R fdw[9], fup[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) {
    const R vm = lane_up(a.v[f]), vp = lane_dn(a.v[f]);
    const R d = dqq_eval(vm, a.v[f], vp, ep);
    const R dn = lane_dn(d);
    fup[f] = flux_core(a.v[f], vp, d, dn, cup, fh);
    fdw[f] = lane_up(fup[f]);
  }
  if (out_lane) av_cell_update(a, fdw, fup, dxi[i], dt, rslt, c);
}

template <typename R>
__global__ __launch_bounds__(256) void av_xshfl_kernel(Fld<const R> qq, Fld<R> rslt,
                                                      R *__restrict__ cc_store,
                                                      const R *__restrict__ dxi, Shape sh,
                                                      int j_lo, int k_lo, int j_n, int k_n,
                                                      int nseg, int nb, AvScalars sc) {
  av_xshfl_block<R>(qq, rslt, cc_store, dxi, sh, j_lo, k_lo, j_n, k_n, nseg, sc, R(sc.dt),
                    xcd_virtual_block(blockIdx.x, nb));
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

// ---------------------------------------------------------------------------
// x and y sweeps in one walk (3D).  A wave owns 64 x lanes (60 outputs, two
// halo lanes each side, as the x-shuffle kernel) of one plane and walks a
// chunk of y rows.  Every row entering the y window is x-swept in registers
// first, with the x-shuffle kernel's arithmetic; a ghost row is what the
// ghost fill after the x sweep would hold: the x-swept source row for a
// periodic or mirrored (sign-flipped) boundary, the row itself for a fixed
// one (its ghost values are the fixed fields both before and after the x
// sweep).  The frozen speed cc of a row is evaluated from the row's post-RK
// values (the source row's for a mapped ghost row: the same bits, since cc
// depends on squares only), and stored for the interior and the z ring,
// which is all the z sweep reads.  One read pass and one write pass of the
// separate x sweep, and the ghost fill between the sweeps, drop out.

// the post-RK fields of the row that supplies row jr (the row itself, or
// the source row of a periodic or mirrored ghost row)
template <typename R>
__device__ __forceinline__ void av_xrow_load(const Fld<const R> &qq, const Shape &sh,
                                             const BcDev &bc, int ic, int jr, int k, R f[9]) {
  const bool ghost = jr < sh.mj || jr >= sh.Jt - sh.mj;
  bool mir = false;
  const int js = ghost ? map_coord(jr, sh.Jt, sh.mj, bc.periodic[1], mir) : jr;
  const int c = sh.idx(ic, js, k);
#pragma unroll
  for (int m = 0; m < 9; ++m) f[m] = qq.q[m][c];
}

// Conserved component m (ro, ro vx, ro vy, ro vz, bx, by, bz, ph, ro ei: the
// order of AvCons) of a cell from its nine fields; the products are the ones
// load_cons forms.
template <typename R, int M> __device__ __forceinline__ R cons_m(const R f[9]) {
  return M == 0 ? f[0] : (M <= 3 ? f[0] * f[M] : (M <= 6 ? f[M] : (M == 7 ? f[8] : f[0] * f[7])));
}

// Total energy of a cell (av_cell_values' Et, in double).
template <typename R> __device__ __forceinline__ double cell_et(const R f[9]) {
  const double pii8 = double(Const<R>::pii8);
  const double ro = f[0];
  return ro * double(f[7]) + 0.5 * ro * (double(f[1]) * f[1] + double(f[2]) * f[2] + double(f[3]) * f[3]) +
         pii8 * (double(f[4]) * f[4] + double(f[5]) * f[5] + double(f[6]) * f[6]);
}

// One field of av_cell_values from its flux difference df = (fup - fdw) dd dt:
// the new value into o (ro first: the velocities divide by it), the
// energy's term into efl.
template <typename R, int M>
__device__ __forceinline__ void cell_field(R v, R df, R o[9], R &efl) {
  if (M == 0) o[0] = v - df;
  else if (M <= 3) o[M] = (v - df) / o[0];
  else if (M <= 6) o[M] = v - df;
  else if (M == 7) o[8] = v - df;
  else efl = df;
}

// The internal energy (av_cell_values' recovery) once every other field is new.
template <typename R>
__device__ __forceinline__ void cell_energy(double Et, R efl, R o[9]) {
  const double pii8 = double(Const<R>::pii8);
  const double dro = o[0], dvx = o[1], dvy = o[2], dvz = o[3], dbx = o[4], dby = o[5], dbz = o[6];
  const double ein = (Et - double(efl) - 0.5 * dro * (dvx * dvx + dvy * dvy + dvz * dvz) -
                      pii8 * (dbx * dbx + dby * dby + dbz * dbz)) /
                     dro;
  o[7] = R(ein);
}

// x-sweep of one field M at every lane (the x-shuffle kernel's slope, flux
// and update), accumulated into o / efl
template <typename R, int M>
__device__ __forceinline__ void xsweep_field(const R f[9], R cup, R ep, R fh, R dd, R dt, R o[9],
                                             R &efl) {
  const R v = cons_m<R, M>(f);
  const R vm = lane_up(v), vp = lane_dn(v);
  const R d = dqq_eval(vm, v, vp, ep);
  const R dn = lane_dn(d);
  const R fup = flux_core(v, vp, d, dn, cup, fh);
  const R fdw = lane_up(fup);
  cell_field<R, M>(v, (fup - fdw) * dd * dt, o, efl);
}

// x-swept fields of row jr (primitive) and its frozen speed, from the fields
// av_xrow_load gave for it; one field at a time, so only one field's
// slopes and fluxes are live
template <typename R>
__device__ __forceinline__ void av_xrow_sweep(const R f[9], const Shape &sh, const BcDev &bc,
                                              int jr, R ddx, const AvScalars &sc, R dt, R o[9],
                                              R &ccv) {
  const bool ghost = jr < sh.mj || jr >= sh.Jt - sh.mj;
  const int kind = bc.periodic[1];
  const bool mir = ghost && kind == 0;
  ccv = char_speed_v(f, R(sc.gm), R(sc.cs_fac), R(sc.ca_fac), R(sc.vv_fac));
  if (ghost && kind == 2) {  // uniform per wave: a fixed ghost row is not swept
#pragma unroll
    for (int m = 0; m < 9; ++m) o[m] = f[m];
    return;
  }
  const R ep = R(sc.ep), fh = R(sc.fh);
  const R cup = R(0.5) * (ccv + lane_dn(ccv));
  const double Et = cell_et(f);
  R efl;
  xsweep_field<R, 0>(f, cup, ep, fh, ddx, dt, o, efl);
  xsweep_field<R, 1>(f, cup, ep, fh, ddx, dt, o, efl);
  xsweep_field<R, 2>(f, cup, ep, fh, ddx, dt, o, efl);
  xsweep_field<R, 3>(f, cup, ep, fh, ddx, dt, o, efl);
  xsweep_field<R, 4>(f, cup, ep, fh, ddx, dt, o, efl);
  xsweep_field<R, 5>(f, cup, ep, fh, ddx, dt, o, efl);
  xsweep_field<R, 6>(f, cup, ep, fh, ddx, dt, o, efl);
  xsweep_field<R, 7>(f, cup, ep, fh, ddx, dt, o, efl);
  xsweep_field<R, 8>(f, cup, ep, fh, ddx, dt, o, efl);
  cell_energy<R>(Et, efl, o);
  if (mir) {
#pragma unroll
    for (int m = 0; m < 9; ++m) o[m] = R(bc.sign[1][m]) * o[m];
  }
}

// y step of one field M of the window (p, p+1, p+2 as primitive fields):
// slope at p+1, flux at p+1/2, the update of cell p; d0 and fdw carry the
// slope at p and the flux at p-1/2
template <typename R, int M>
__device__ __forceinline__ void ystep_field(const R w0[9], const R w1[9], const R wn[9], R cup,
                                            R ep, R fh, R dd, R dt, R d0[9], R fdw[9], R o[9],
                                            R &efl) {
  const R v0 = cons_m<R, M>(w0), v1 = cons_m<R, M>(w1), vn = cons_m<R, M>(wn);
  const R d1 = dqq_eval(v0, v1, vn, ep);
  const R fup = flux_core(v0, v1, d0[M], d1, cup, fh);
  d0[M] = d1;
  const R df = (fup - fdw[M]) * dd * dt;
  fdw[M] = fup;
  cell_field<R, M>(v0, df, o, efl);
}

// W: waves per SIMD the FP64 walk is compiled for -- 3 (168 registers; the
// compiler parks about 20 of them in scratch, and a scratch reload waits
// for every earlier global load and store) or 2 (190 registers, nothing
// parked).  FP32 needs about 100 registers and runs 4.
template <typename R, int W>
__global__ __launch_bounds__(256) __attribute__((
    amdgpu_waves_per_eu(sizeof(R) == 8 ? W : 4, sizeof(R) == 8 ? W : 4))) void av_xy_walk_kernel(Fld<const R> qq, Fld<R> rslt,
                                                        R *__restrict__ cc_store,
                                                        const R *__restrict__ dxi,
                                                        const R *__restrict__ dyi, Shape sh,
                                                        BcDev bc, int chunk, int nchunks,
                                                        int nseg, int nwaves, AvScalars sc) {
  const int w = xcd_virtual_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (w >= nwaves) return;  // whole wave; the kernel has no block barrier
  const int seg = w % nseg, rr = w / nseg, ch = rr % nchunks;
  int k = (sc.nparts ? sc.part0 : sh.mk - 1) + rr / nchunks;
  if (k > sh.Kt - sh.mk) k -= sh.Kt - 2 * sh.mk + 2;  // past the top ring plane: the bottom one on
  const int lane = threadIdx.x & 63;
  const int i = sh.mi + seg * 60 - 2 + lane;
  // a lane on a ghost column of a wrapped x axis reads its periodic source
  const int ic = wrap_coord(min(i, sh.It - 1), sh.It, sh.mi, sc.wrap & 1);
  const bool out_lane = lane >= 2 && lane < 62 && i < sh.It - sh.mi;
  const int nj = sh.Jt - 2 * sh.mj;
  const int p0 = ch * chunk, len = min(chunk, nj - p0);
  const R gm = R(sc.gm), csf = R(sc.cs_fac), caf = R(sc.ca_fac), vvf = R(sc.vv_fac);
  if (k < sh.mk || k >= sh.Kt - sh.mk) {  // z ring plane: the frozen speed only
    // (of the source plane when z is wrapped: cc depends on the cell alone)
    const int ks = wrap_coord(k, sh.Kt, sh.mk, sc.wrap & 4);
    for (int t = 0; t < len; ++t) {
      const int c = sh.idx(ic, sh.mj + p0 + t, k);
      const R v = char_speed(qq, sh.idx(ic, sh.mj + p0 + t, ks), gm, csf, caf, vvf);
      if (out_lane) cc_store[c] = v;
    }
    return;
  }
  const R dt = R(sc.dt), ep = R(sc.ep), fh = R(sc.fh);
  const R ddx = out_lane ? dxi[i] : R(0);
  const int jb = sh.mj + p0;  // first output row
  // window: w0 = p, w1 = p+1 (x-swept, primitive); slopes at p; flux p-1/2
  R wm[9], w0[9], w1[9], w2[9];
  R cc_m, cc_0, cc_1;
  R nf[9];  // fields of the next row to sweep, loaded one row ahead
  av_xrow_load(qq, sh, bc, ic, jb - 2, k, nf);
  av_xrow_sweep(nf, sh, bc, jb - 2, ddx, sc, dt, wm, cc_m);
  av_xrow_load(qq, sh, bc, ic, jb - 1, k, nf);
  av_xrow_sweep(nf, sh, bc, jb - 1, ddx, sc, dt, w0, cc_m);
  av_xrow_load(qq, sh, bc, ic, jb, k, nf);
  av_xrow_sweep(nf, sh, bc, jb, ddx, sc, dt, w1, cc_0);
  av_xrow_load(qq, sh, bc, ic, jb + 1, k, nf);
  av_xrow_sweep(nf, sh, bc, jb + 1, ddx, sc, dt, w2, cc_1);
  av_xrow_load(qq, sh, bc, ic, jb + 2, k, nf);
  R d0[9], fdw[9];
  {
    const R cdw = R(0.5) * (cc_m + cc_0);
    R am[9], a0[9], a1[9], a2[9];
    am[0] = cons_m<R, 0>(wm); a0[0] = cons_m<R, 0>(w0); a1[0] = cons_m<R, 0>(w1); a2[0] = cons_m<R, 0>(w2);
    am[1] = cons_m<R, 1>(wm); a0[1] = cons_m<R, 1>(w0); a1[1] = cons_m<R, 1>(w1); a2[1] = cons_m<R, 1>(w2);
    am[2] = cons_m<R, 2>(wm); a0[2] = cons_m<R, 2>(w0); a1[2] = cons_m<R, 2>(w1); a2[2] = cons_m<R, 2>(w2);
    am[3] = cons_m<R, 3>(wm); a0[3] = cons_m<R, 3>(w0); a1[3] = cons_m<R, 3>(w1); a2[3] = cons_m<R, 3>(w2);
    am[4] = cons_m<R, 4>(wm); a0[4] = cons_m<R, 4>(w0); a1[4] = cons_m<R, 4>(w1); a2[4] = cons_m<R, 4>(w2);
    am[5] = cons_m<R, 5>(wm); a0[5] = cons_m<R, 5>(w0); a1[5] = cons_m<R, 5>(w1); a2[5] = cons_m<R, 5>(w2);
    am[6] = cons_m<R, 6>(wm); a0[6] = cons_m<R, 6>(w0); a1[6] = cons_m<R, 6>(w1); a2[6] = cons_m<R, 6>(w2);
    am[7] = cons_m<R, 7>(wm); a0[7] = cons_m<R, 7>(w0); a1[7] = cons_m<R, 7>(w1); a2[7] = cons_m<R, 7>(w2);
    am[8] = cons_m<R, 8>(wm); a0[8] = cons_m<R, 8>(w0); a1[8] = cons_m<R, 8>(w1); a2[8] = cons_m<R, 8>(w2);
#pragma unroll
    for (int f = 0; f < 9; ++f) {
      const R dm = dqq_eval(am[f], a0[f], a1[f], ep);
      d0[f] = dqq_eval(a0[f], a1[f], a2[f], ep);
      fdw[f] = flux_core(a0[f], a1[f], dm, d0[f], cdw, fh);
    }
  }
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    w0[m] = w1[m];
    w1[m] = w2[m];
  }
  R dd_n = dyi[jb];  // the next output row's 1/dy, loaded a row ahead
  // one output row: A, B hold rows p and p+1 (x-swept, primitive), row p+2
  // is swept into Cn; c0, c1, c2 are the rows' frozen speeds
  auto step = [&](R (&A)[9], R (&B)[9], R (&Cn)[9], R &c0, R &c1, R &c2, int t) {
    const int j = jb + t;
    av_xrow_sweep(nf, sh, bc, j + 2, ddx, sc, dt, Cn, c2);
    // row j+3 exists (at most the last ghost row) when j+1 is an output row;
    // 1/dy of row j+1 comes after it, so waiting for it waits for the row
    const R dd = dd_n;
    if (t + 1 < len) {
      av_xrow_load(qq, sh, bc, ic, j + 3, k, nf);
      dd_n = dyi[j + 1];
    }
    const R cup = R(0.5) * (c0 + c1);
    const double Et = cell_et(A);
    R o[9], efl;
    ystep_field<R, 0>(A, B, Cn, cup, ep, fh, dd, dt, d0, fdw, o, efl);
    ystep_field<R, 1>(A, B, Cn, cup, ep, fh, dd, dt, d0, fdw, o, efl);
    ystep_field<R, 2>(A, B, Cn, cup, ep, fh, dd, dt, d0, fdw, o, efl);
    ystep_field<R, 3>(A, B, Cn, cup, ep, fh, dd, dt, d0, fdw, o, efl);
    ystep_field<R, 4>(A, B, Cn, cup, ep, fh, dd, dt, d0, fdw, o, efl);
    ystep_field<R, 5>(A, B, Cn, cup, ep, fh, dd, dt, d0, fdw, o, efl);
    ystep_field<R, 6>(A, B, Cn, cup, ep, fh, dd, dt, d0, fdw, o, efl);
    ystep_field<R, 7>(A, B, Cn, cup, ep, fh, dd, dt, d0, fdw, o, efl);
    ystep_field<R, 8>(A, B, Cn, cup, ep, fh, dd, dt, d0, fdw, o, efl);
    cell_energy<R>(Et, efl, o);
    const int c = sh.idx(ic, j, k);
    if (out_lane) {
#pragma unroll
      for (int f = 0; f < 9; ++f) rslt.q[f][c] = o[f];
      cc_store[c] = c0;
    }
  };
  R wx[9], cc_2 = R(0);
  if constexpr (sizeof(R) == 8 && W == 2) {
    // the two-wave FP64 build has the registers for the window's rows to
    // rotate through three sets (the loop unrolled by three) instead of
    // being copied at every row: 4 % fewer VALU instructions
    int t = 0;
    for (; t + 2 < len; t += 3) {
      step(w0, w1, wx, cc_0, cc_1, cc_2, t);
      step(w1, wx, w0, cc_1, cc_2, cc_0, t + 1);
      step(wx, w0, w1, cc_2, cc_0, cc_1, t + 2);
    }
    if (t < len) step(w0, w1, wx, cc_0, cc_1, cc_2, t);
    if (t + 1 < len) step(w1, wx, w0, cc_1, cc_2, cc_0, t + 1);
  } else {
    for (int t = 0; t < len; ++t) {
      step(w0, w1, wx, cc_0, cc_1, cc_2, t);
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        w0[m] = w1[m];
        w1[m] = wx[m];
      }
      cc_0 = cc_1;
      cc_1 = cc_2;
    }
  }
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

template <typename R, int DIR, bool DEGEN = false>
static hipError_t launch_walk_t(hipStream_t st, const Fld<const R> &q, const Fld<R> &r,
                                const R *cc_in, const R *dd, const Shape &sh,
                                const AvScalars &sc, const R *dz_arr = nullptr) {
  const int n_dir = DIR == 1 ? sh.Jt - 2 * sh.mj : sh.Kt - 2 * sh.mk;
  const int n_oth = DIR == 1 ? sh.Kt - 2 * sh.mk : sh.Jt - 2 * sh.mj;
  const int wx = walk_waves_x(n_oth);
  const int nbx = ((sh.It - 2 * sh.mi + 63) / 64 + wx - 1) / wx;
  const int nbo = (n_oth + 4 / wx - 1) / (4 / wx);
  int chunk, nchunks;  // even walk lengths
  walk_split(n_dir, 256L * nbx * nbo, sc.chunk, chunk, nchunks);
  if (sc.nparts && (sc.part0 < 0 || sc.nparts < 0 || sc.part0 + sc.nparts > nchunks))
    return hipErrorInvalidValue;
  const long nbl = (long)(sc.nparts ? sc.nparts : nchunks) * nbx * nbo;
  if (nbl > 2147483647L) return hipErrorInvalidValue;
  hipLaunchKernelGGL((av_walk_kernel<R, DIR, DEGEN>), dim3((unsigned)nbl), dim3(256), 0, st, q, r,
                     cc_in, dd, sh, chunk, nchunks, nbx, nbo, (int)nbl, sc, dz_arr);
  return hipGetLastError();
}

template <typename R>
static hipError_t launch_xshfl(hipStream_t st, const Fld<const R> &q, const Fld<R> &r,
                               R *cc_store, const R *dxi, const Shape &sh, const AvScalars &sc) {
  const int j_lo = sh.mj - sh.js, j_n = sh.Jt - 2 * sh.mj + 2 * sh.js;
  const int k_lo = sh.mk - sh.ks, k_n = sh.Kt - 2 * sh.mk + 2 * sh.ks;
  const int nseg = (sh.It - 2 * sh.mi + 59) / 60;
  const long rows = (long)j_n * k_n;
  const long nbl = (long)nseg * ((rows + 3) / 4);
  if (nbl > 2147483647L) return hipErrorInvalidValue;
  hipLaunchKernelGGL((av_xshfl_kernel<R>), dim3((unsigned)nbl), dim3(256), 0, st, q, r, cc_store,
                     dxi, sh, j_lo, k_lo, j_n, k_n, nseg, (int)nbl, sc);
  return hipGetLastError();
}

// Active directions use the x-shuffle and walking kernels; a direction of
// extent one (2D, 1D runs) is the pointwise degenerate sweep of the
// reference, done by the generic kernel with stride 0.
template <typename R>
hipError_t launch_av_sweep(hipStream_t st, int dir, const Fld<const R> &q, const Fld<R> &r,
                           R *cc_store, const R *cc_in, const R *dd, const Shape &sh,
                           const AvScalars &sc) {
  if (dir == 0) {
    if (sh.is) return launch_xshfl<R>(st, q, r, cc_store, dd, sh, sc);
    return launch_av_t<R, 0>(st, q, r, cc_store, cc_in, dd, sh, sc);
  }
  if (dir == 1) {
    if (sh.js) return launch_walk_t<R, 1>(st, q, r, cc_in, dd, sh, sc);
    return launch_av_t<R, 1>(st, q, r, cc_store, cc_in, dd, sh, sc);
  }
  if (sh.ks) return launch_walk_t<R, 2>(st, q, r, cc_in, dd, sh, sc);
  return launch_av_t<R, 2>(st, q, r, cc_store, cc_in, dd, sh, sc);
}

// 2D runs (x and y active, z of extent one): the y sweep with the
// degenerate z sweep applied to each result in registers; dzi is the z
// metric (its single inverse spacing is read on the device).
template <typename R>
hipError_t launch_av_sweep_y_degen_z(hipStream_t st, const Fld<const R> &q, const Fld<R> &r,
                                     const R *cc_in, const R *dyi, const R *dzi, const Shape &sh,
                                     const AvScalars &sc) {
  if (!sh.is || !sh.js || sh.ks) return hipErrorInvalidValue;
  return launch_walk_t<R, 1, true>(st, q, r, cc_in, dyi, sh, sc, dzi);
}

template <typename R, int DIR, bool DEGEN>
static hipError_t launch_walk_cfl_t(hipStream_t st, const Fld<const R> &q, const Fld<R> &r,
                                    const R *cc_in, const R *dd, const Shape &sh,
                                    const AvScalars &sc, const R *dz_arr, const CflFuse<R> &cf) {
  const int n_dir = DIR == 1 ? sh.Jt - 2 * sh.mj : sh.Kt - 2 * sh.mk;
  const int n_oth = DIR == 1 ? sh.Kt - 2 * sh.mk : sh.Jt - 2 * sh.mj;
  const int wx = walk_waves_x(n_oth);
  const int nbx = ((sh.It - 2 * sh.mi + 63) / 64 + wx - 1) / wx;
  const int nbo = (n_oth + 4 / wx - 1) / (4 / wx);
  int chunk, nchunks;  // even walk lengths
  walk_split(n_dir, 256L * nbx * nbo, sc.chunk, chunk, nchunks);
  if (sc.nparts && (sc.part0 < 0 || sc.nparts < 0 || sc.part0 + sc.nparts > nchunks))
    return hipErrorInvalidValue;
  const long nbl = (long)(sc.nparts ? sc.nparts : nchunks) * nbx * nbo;
  if (nbl > 2147483647L) return hipErrorInvalidValue;
  hipLaunchKernelGGL((av_walk_cfl_kernel<R, DIR, DEGEN>), dim3((unsigned)nbl), dim3(256), 0, st,
                     q, r, cc_in, dd, sh, chunk, nchunks, nbx, (int)nbl, sc, dz_arr, cf,
                     DIR == 2 ? cf.dz : cf.dy);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_av_last_sweep_cfl(hipStream_t st, const Fld<const R> &q, const Fld<R> &r,
                                    const R *cc_in, const R *dyi, const R *dzi, const Shape &sh,
                                    const AvScalars &sc, const CflFuse<R> &cf) {
  if (sh.is && sh.js && sh.ks) return launch_walk_cfl_t<R, 2, false>(st, q, r, cc_in, dzi, sh, sc,
                                                                      nullptr, cf);
  if (sh.is && sh.js && !sh.ks) return launch_walk_cfl_t<R, 1, true>(st, q, r, cc_in, dyi, sh, sc,
                                                                      dzi, cf);
  return hipErrorInvalidValue;
}

// The x and y sweeps of a 3D run in one walk (y extent of at least 2, so
// every ghost row maps into the interior); cc is written for the z sweep.
template <typename R>
hipError_t launch_av_sweep_xy(hipStream_t st, const Fld<const R> &q, const Fld<R> &r,
                              R *cc_store, const R *dxi, const R *dyi, const Shape &sh,
                              const BcDev &bc, const AvScalars &sc) {
  const int nj = sh.Jt - 2 * sh.mj;
  if (!sh.is || !sh.js || !sh.ks || nj < 2 || sh.mj < 2) return hipErrorInvalidValue;
  const int nseg = (sh.It - 2 * sh.mi + 59) / 60;
  int chunk, nchunks;  // even walk lengths
  walk_split(nj, 64L * nseg * (sh.Kt - 2 * sh.mk + 2), sc.chunk, chunk, nchunks);
  const int npl = sh.Kt - 2 * sh.mk + 2;  // interior planes and the two z ring planes
  if (sc.nparts && (sc.nparts < 0 || sc.nparts > npl || sc.part0 < sh.mk - 1 ||
                    sc.part0 > sh.Kt - sh.mk))
    return hipErrorInvalidValue;
  const long nw = (long)nseg * nchunks * (sc.nparts ? sc.nparts : npl);
  const long nbl = (nw + 3) / 4;
  if (nbl > 2147483647L) return hipErrorInvalidValue;
  if (sc.xy_waves == 2)
    hipLaunchKernelGGL((av_xy_walk_kernel<R, 2>), dim3((unsigned)nbl), dim3(256), 0, st, q, r,
                       cc_store, dxi, dyi, sh, bc, chunk, nchunks, nseg, (int)nw, sc);
  else
    hipLaunchKernelGGL((av_xy_walk_kernel<R, 3>), dim3((unsigned)nbl), dim3(256), 0, st, q, r,
                       cc_store, dxi, dyi, sh, bc, chunk, nchunks, nseg, (int)nw, sc);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_char_vel(hipStream_t st, const Fld<const R> &q, R *cc, int n, const AvScalars &sc) {
  hipLaunchKernelGGL((char_vel_kernel<R>), dim3((n + 255) / 256), dim3(256), 0, st, q, cc, n, sc);
  return hipGetLastError();
}

// ---------------------------------------------------------------- batches
template <typename R> __device__ __forceinline__ Fld<const R> av_const(const Fld<R> &f) {
  Fld<const R> o;
#pragma unroll
  for (int m = 0; m < 9; ++m) o.q[m] = f.q[m];
  return o;
}

// One AV sweep of every instance of a batch: MODE 0 the x-shuffle sweep, 1
// the walking y/z sweep, 2 the generic kernel (degenerate direction).  The
// block decomposition per instance is the single-run launcher's.
template <typename R, int DIR, int MODE>
__global__ __launch_bounds__(256) void batch_av_kernel(const BatchInst<R> *__restrict__ inst,
                                                       const BatchDyn<R> *__restrict__ dyn,
                                                       const int *__restrict__ start, int n,
                                                       int total, int in_set, int out_set,
                                                       AvScalars sc) {
  const int V = xcd_virtual_block(blockIdx.x, total);
  const int m = batch_find(start, n, V);
  const int b = V - start[m];
  const BatchInst<R> &I = inst[m];
  const Shape sh = I.sh;
  const R dt = dyn[m].dt;
  const Fld<const R> q = av_const(I.st[in_set]);
  const Fld<R> out = I.st[out_set];
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nk = sh.Kt - 2 * sh.mk;
  if (MODE == 0) {
    const int j_lo = sh.mj - sh.js, j_n = nj + 2 * sh.js;
    const int k_lo = sh.mk - sh.ks, k_n = nk + 2 * sh.ks;
    av_xshfl_block<R>(q, out, I.cc, I.dxi, sh, j_lo, k_lo, j_n, k_n, (ni + 59) / 60, sc, dt, b);
  } else if (MODE == 1 || MODE == 3) {
    const int n_dir = DIR == 1 ? nj : nk;
    const int chunk = I.chunk;
    const int wx = walk_waves_x(DIR == 1 ? nk : nj);
    av_walk_block<R, DIR, MODE == 3>(q, out, I.cc, DIR == 1 ? I.dyi : I.dzi, sh, chunk,
                                     (n_dir + chunk - 1) / chunk, ((ni + 63) / 64 + wx - 1) / wx,
                                     sc, dt, b, MODE == 3 ? I.dzi[0] : R(0));
  } else {
    const R *dd = DIR == 0 ? I.dxi : (DIR == 1 ? I.dyi : I.dzi);
    av_generic_block<R, DIR>(q, out, I.cc, I.cc, dd, sh, sh.mj, sh.mk, nj, nk, sc, dt,
                             (ni + 63) / 64, (nj + 3) / 4, b);
  }
}

template <typename R, int DIR, int MODE>
static hipError_t launch_batch_av_t(hipStream_t st, const BatchInst<R> *inst,
                                    const BatchDyn<R> *dyn, const int *start, int n, int total,
                                    int in_set, int out_set, const AvScalars &sc) {
  hipLaunchKernelGGL((batch_av_kernel<R, DIR, MODE>), dim3(total), dim3(256), 0, st, inst, dyn,
                     start, n, total, in_set, out_set, sc);
  return hipGetLastError();
}

// mode: 0 x-shuffle (dir 0), 1 walk (dir 1, 2), 2 degenerate (dir 1, 2),
// 3 walk along y with the degenerate z sweep fused (dir 1, 2D runs)
template <typename R>
hipError_t launch_batch_av(hipStream_t st, int dir, int mode, const BatchInst<R> *inst,
                           const BatchDyn<R> *dyn, const int *start, int n, int total, int in_set,
                           int out_set, const AvScalars &sc) {
  if (dir == 0 && mode == 0)
    return launch_batch_av_t<R, 0, 0>(st, inst, dyn, start, n, total, in_set, out_set, sc);
  if (dir == 1 && mode == 1)
    return launch_batch_av_t<R, 1, 1>(st, inst, dyn, start, n, total, in_set, out_set, sc);
  if (dir == 1 && mode == 2)
    return launch_batch_av_t<R, 1, 2>(st, inst, dyn, start, n, total, in_set, out_set, sc);
  if (dir == 1 && mode == 3)
    return launch_batch_av_t<R, 1, 3>(st, inst, dyn, start, n, total, in_set, out_set, sc);
  if (dir == 2 && mode == 1)
    return launch_batch_av_t<R, 2, 1>(st, inst, dyn, start, n, total, in_set, out_set, sc);
  if (dir == 2 && mode == 2)
    return launch_batch_av_t<R, 2, 2>(st, inst, dyn, start, n, total, in_set, out_set, sc);
  return hipErrorInvalidValue;
}

void av_walk_split(const Shape &sh, int dir, int chunk_req, int *chunk, int *nchunks) {
  const int n_dir = dir == 1 ? sh.Jt - 2 * sh.mj : sh.Kt - 2 * sh.mk;
  const int n_oth = dir == 1 ? sh.Kt - 2 * sh.mk : sh.Jt - 2 * sh.mj;
  const int wx = walk_waves_x(n_oth);
  const int nbx = ((sh.It - 2 * sh.mi + 63) / 64 + wx - 1) / wx;
  const int nbo = (n_oth + 4 / wx - 1) / (4 / wx);
  walk_split(n_dir, 256L * nbx * nbo, chunk_req, *chunk, *nchunks);
}

template hipError_t launch_batch_av<double>(hipStream_t, int, int, const BatchInst<double> *,
                                            const BatchDyn<double> *, const int *, int, int, int,
                                            int, const AvScalars &);
template hipError_t launch_batch_av<float>(hipStream_t, int, int, const BatchInst<float> *,
                                           const BatchDyn<float> *, const int *, int, int, int,
                                           int, const AvScalars &);

template hipError_t launch_av_sweep<double>(hipStream_t, int, const Fld<const double> &,
                                            const Fld<double> &, double *, const double *,
                                            const double *, const Shape &, const AvScalars &);
template hipError_t launch_av_sweep<float>(hipStream_t, int, const Fld<const float> &,
                                           const Fld<float> &, float *, const float *,
                                           const float *, const Shape &, const AvScalars &);
template hipError_t launch_av_sweep_y_degen_z<double>(hipStream_t, const Fld<const double> &,
                                                      const Fld<double> &, const double *,
                                                      const double *, const double *,
                                                      const Shape &, const AvScalars &);
template hipError_t launch_av_sweep_y_degen_z<float>(hipStream_t, const Fld<const float> &,
                                                     const Fld<float> &, const float *,
                                                     const float *, const float *, const Shape &,
                                                     const AvScalars &);
template hipError_t launch_av_last_sweep_cfl<double>(hipStream_t, const Fld<const double> &,
                                                     const Fld<double> &, const double *,
                                                     const double *, const double *,
                                                     const Shape &, const AvScalars &,
                                                     const CflFuse<double> &);
template hipError_t launch_av_last_sweep_cfl<float>(hipStream_t, const Fld<const float> &,
                                                    const Fld<float> &, const float *,
                                                    const float *, const float *, const Shape &,
                                                    const AvScalars &, const CflFuse<float> &);
template hipError_t launch_av_sweep_xy<double>(hipStream_t, const Fld<const double> &,
                                               const Fld<double> &, double *, const double *,
                                               const double *, const Shape &, const BcDev &,
                                               const AvScalars &);
template hipError_t launch_av_sweep_xy<float>(hipStream_t, const Fld<const float> &,
                                              const Fld<float> &, float *, const float *,
                                              const float *, const Shape &, const BcDev &,
                                              const AvScalars &);
template hipError_t launch_char_vel<double>(hipStream_t, const Fld<const double> &, double *, int,
                                            const AvScalars &);
template hipError_t launch_char_vel<float>(hipStream_t, const Fld<const float> &, float *, int,
                                           const AvScalars &);

}  // namespace misohip
