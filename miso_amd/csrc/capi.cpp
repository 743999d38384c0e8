// C ABI of the MI355X MHD backend (include/misohip.h).
//
// Host-side orchestration of MISO's Integrator (mhd_integrator_host.hpp:
// 258-350): the scalar set-up (divb_parameters_set, the stage steps dt/4,
// dt/3, dt/2, dt and exp(-dt_s/tau)) is done here in the Real type of the
// run with the same expressions as the reference, so the device receives the
// very same scalars the reference's kernels would.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

#include "../../include/misohip.h"
#include "tail_plan.hpp"
#include "launch.hpp"


using namespace misohip;

#include "ctx_internal.hpp"

namespace misohip_detail {
thread_local char g_last_hip_error[256];
int hip_fail(hipError_t e) {
  std::snprintf(g_last_hip_error, sizeof(g_last_hip_error), "%s", hipGetErrorString(e));
  return MISOHIP_E_HIP;
}
}  // namespace misohip_detail
using misohip_detail::g_last_hip_error;

namespace {

int check_grid(const misohip_grid *g) {
  if (!g) return MISOHIP_E_ARG;
  const int sz[3] = {g->i_size, g->j_size, g->k_size};
  const int tot[3] = {g->i_total, g->j_total, g->k_total};
  const int mar[3] = {g->i_margin, g->j_margin, g->k_margin};
  const int st[3] = {g->is, g->js, g->ks};
  for (int a = 0; a < 3; ++a) {
    if (sz[a] <= 0 || mar[a] < 0) return MISOHIP_E_ARG;
    if (st[a] != (sz[a] > 1 ? 1 : 0)) return MISOHIP_E_ARG;
    if (sz[a] == 1 && mar[a] != 0) return MISOHIP_E_ARG;
    if (sz[a] > 1 && mar[a] < 2) return MISOHIP_E_ARG;  // radius-2 stencils
    if (tot[a] != sz[a] + 2 * mar[a]) return MISOHIP_E_ARG;
  }
  const double n = (double)g->i_total * g->j_total * g->k_total;
  if (n >= 2147483647.0) return MISOHIP_E_RANGE;  // int indexing inside kernels
  if (!g->dx || !g->dy || !g->dz || !g->dxi || !g->dyi || !g->dzi) return MISOHIP_E_ARG;
  return MISOHIP_OK;
}

int check_fields(const misohip_fields *f) {
  if (!f) return MISOHIP_E_ARG;
  for (int m = 0; m < 9; ++m)
    if (!f->q[m]) return MISOHIP_E_ARG;
  return MISOHIP_OK;
}

Shape make_shape(const misohip_grid *g) {
  Shape s;
  s.It = g->i_total; s.Jt = g->j_total; s.Kt = g->k_total;
  s.mi = g->i_margin; s.mj = g->j_margin; s.mk = g->k_margin;
  s.is = g->is; s.js = g->js; s.ks = g->ks;
  return s;
}

template <typename R> Fld<R> fld(const misohip_fields *f) {
  Fld<R> r;
  for (int m = 0; m < 9; ++m) r.q[m] = static_cast<R *>(f->q[m]);
  return r;
}
template <typename R> Fld<const R> cfld(const misohip_fields *f) {
  Fld<const R> r;
  for (int m = 0; m < 9; ++m) r.q[m] = static_cast<const R *>(f->q[m]);
  return r;
}

// a fixed axis needs its nine value fields
bool bc_invalid(const misohip_bc *bc) {
  if (!bc) return true;
  for (int a = 0; a < 3; ++a)
    if (bc->periodic[a] == 2)
      for (int f = 0; f < 9; ++f)
        if (!bc->fixed[f]) return true;
  return false;
}

// an axis of kind 3 (user) needs the context's boundary callback
bool bc_invalid(const misohip_ctx *ctx, const misohip_bc *bc) {
  if (bc_invalid(bc)) return true;
  for (int a = 0; a < 3; ++a)
    if (bc->periodic[a] == 3 && !ctx->bc_fn) return true;
  return false;
}

// a source description the context can carry out: gravity on a real axis,
// or the user callback with its four arrays
bool src_invalid(const misohip_ctx *ctx, const misohip_params *p) {
  if (p->src_kind < 0 || p->src_kind > 2) return true;
  if (p->src_kind == 1 && (p->grav_axis < 0 || p->grav_axis > 2)) return true;
  if (p->src_kind == 2) {
    if (!ctx->src_fn) return true;
    for (int d = 0; d < 4; ++d)
      if (!ctx->src_arr[d]) return true;
  }
  return false;
}

BcDev make_bc(const misohip_bc *bc) {
  BcDev b;
  for (int a = 0; a < 3; ++a) {
    const int k = bc->periodic[a];
    b.periodic[a] = k == 2 ? 2 : k == 3 ? 3 : (k ? 1 : 0);
    for (int f = 0; f < 9; ++f) b.sign[a][f] = bc->sym_sign[a][f] >= 0 ? 1.0f : -1.0f;
  }
  for (int f = 0; f < 9; ++f) b.fixed[f] = bc->fixed[f];
  return b;
}

template <typename R>
int rk_stage_t(misohip_ctx *ctx, const misohip_grid *g, const misohip_params *p,
               const misohip_fields *o, const misohip_fields *a, const misohip_fields *r,
               R dt_s, R ch2, R damp, int wrap = 0) {
  // a user source first writes src.vx, src.vy, src.vz, src.ei of this
  // stage's argm (the reference evaluates the functor on argm inside the
  // stage, mhd_integrator_host.hpp:134,151,168,241)
  if (p->src_kind == 2 && ctx->src_fn(ctx->src_user, a, ctx->src_arr, ctx->stream) != 0)
    return MISOHIP_E_CALLBACK;
  Timed t_(ctx, MISOHIP_KIND_RK_STAGE);
  RkScalars sc;
  for (int d = 0; d < 4; ++d) sc.src_arr[d] = p->src_kind == 2 ? ctx->src_arr[d] : nullptr;
  sc.dt = dt_s; sc.ch2 = ch2; sc.damp = damp; sc.gm = (R)p->gm;
  sc.grav = (R)p->grav; sc.src_kind = p->src_kind; sc.grav_axis = p->grav_axis;
  sc.pencil_bx = ctx->tile.rk_pencil_bx;
  sc.pencil_by = ctx->tile.rk_pencil_by;
  sc.variant = ctx->tile.rk_variant;
  sc.zchunk = ctx->tile.rk_zchunk;
  sc.wrap = wrap;
  MISOHIP_TRY(launch_rk_stage<R>(ctx->stream, cfld<R>(o), cfld<R>(a), fld<R>(r),
                                 static_cast<const R *>(g->dxi), static_cast<const R *>(g->dyi),
                                 static_cast<const R *>(g->dzi), make_shape(g), sc));
  return MISOHIP_OK;
}

template <typename R> AvScalars av_scalars(const misohip_params *p, R dt, int chunk = 64) {
  AvScalars s;
  s.chunk = chunk;
  s.dt = dt; s.ep = (R)p->ep; s.fh = (R)p->fh; s.gm = (R)p->gm;
  s.cs_fac = (R)p->cs_fac; s.ca_fac = (R)p->ca_fac; s.vv_fac = (R)p->vv_fac;
  return s;
}

template <typename R>
int av_sweep_t(misohip_ctx *ctx, const misohip_grid *g, const misohip_params *p, int dir, R dt,
               const misohip_fields *q, const misohip_fields *r, void *cc, int wrap = 0) {
  const R *dd = static_cast<const R *>(dir == 0 ? g->dxi : dir == 1 ? g->dyi : g->dzi);
  Timed t_(ctx, MISOHIP_KIND_AV_X + dir);
  AvScalars sc = av_scalars<R>(p, dt, ctx->tile.sweep_chunk);
  sc.wrap = wrap;
  MISOHIP_TRY(launch_av_sweep<R>(ctx->stream, dir, cfld<R>(q), fld<R>(r), static_cast<R *>(cc),
                                 static_cast<const R *>(cc), dd, make_shape(g), sc));
  return MISOHIP_OK;
}

template <typename R>
int cfl_t(misohip_ctx *ctx, const misohip_grid *g, const misohip_params *p,
          const misohip_fields *q, double *dt_out) {
  {
    Timed t_(ctx, MISOHIP_KIND_CFL);
    MISOHIP_TRY(launch_cfl<R>(ctx->stream, cfld<R>(q), static_cast<const R *>(g->dx),
                            static_cast<const R *>(g->dy), static_cast<const R *>(g->dz),
                            make_shape(g), (double)(R)p->gm, (double)(R)p->cfl_number,
                            ctx->d_slot));
  }
  MISOHIP_TRY(hipMemcpyAsync(ctx->h_slot, ctx->d_slot, sizeof(unsigned), hipMemcpyDeviceToHost,
                             ctx->stream));
  MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
  float f;
  std::memcpy(&f, ctx->h_slot, sizeof(float));
  *dt_out = (double)(R)f;
  return MISOHIP_OK;
}

// axes: the faces to fill (bit 0 x, 1 y, 2 z); none: nothing to do (only
// passes without a user boundary leave faces out)
template <typename R>
int ghost_t(misohip_ctx *ctx, const misohip_fields *q, const Shape &sh, const BcDev &b,
            int axes = 7) {
  if (!axes && !ctx->bc_fn) return MISOHIP_OK;
  Timed t_(ctx, MISOHIP_KIND_GHOST);
  // the user boundary first, then the fill (BC::apply, then the halo
  // exchange: mhd_integrator_host.hpp:250-255)
  if (ctx->bc_fn && ctx->bc_fn(ctx->bc_user, q, ctx->stream) != 0) return MISOHIP_E_CALLBACK;
  if (axes) MISOHIP_TRY(launch_ghost_fill<R>(ctx->stream, fld<R>(q), sh, b, axes));
  return MISOHIP_OK;
}

// Integrator::update (mhd_integrator_host.hpp:258-300,337-350) with the
// workspace ping-pong: every stage writes rslt, which becomes the next argm.
// fuse_cfl: the last AV sweep also reduces the next step's CFL candidates
// into ctx->d_slot (3D and 2D x-y runs).
// The AV sweeps of a 3D step with a periodic or fixed z axis as a pipeline
// over z slabs (tail_plan.hpp): the one-walk x-y sweeps in slabs on a
// second stream of the context, the z walk in slabs on a third, each z slab
// once the x-y slabs it reads from are done, beside the x-y walk's next
// slabs.  With a fixed z axis the swept fields' z ghost planes, which the
// sequential path fills after the x-y walk from the fixed values alone, are
// filled first.
// The x-y slabs write planes no z slab running beside them reads and read
// the post-RK state, which no z slab writes; the z slabs write disjoint
// planes of qq.  So every value is what the two walks in sequence produce:
// the same bits.
// TAIL_NOT_TAKEN: no plan for this shape, or the stream is being captured;
// the sequential walks run instead.
constexpr int TAIL_NOT_TAKEN = 1;
static_assert(sizeof(misohip_ctx::tail_ev) / sizeof(hipEvent_t) == TAIL_SLABS + 3,
              "one event per slab, the fork and two joins");

template <typename R>
int tail_overlap_t(misohip_ctx *ctx, const misohip_grid *g, const misohip_params *p,
                   const BcDev &b, const Shape &sh, R dt, const misohip_fields *cur,
                   const misohip_fields *dst, const misohip_fields *qq, const misohip_work *w,
                   int wrap, bool fuse_cfl, const double *heat_q) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  MISOHIP_TRY(hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone) return TAIL_NOT_TAKEN;
  int zch, znch;
  av_walk_split(sh, 2, ctx->tile.sweep_chunk, &zch, &znch);
  const bool pz = wrap & 4;
  TailPlan P;
  if ((!pz && b.periodic[2] != 2) || !tail_plan(sh.Kt, sh.mk, zch, znch, pz, &P))
    return TAIL_NOT_TAKEN;
  // a call that must not allocate uses the streams only once they exist
  if (!ctx->tail_side[0] && ctx->no_alloc) return TAIL_NOT_TAKEN;
  if (!ctx->tail_side[0]) {
    for (hipStream_t &st : ctx->tail_side)
      MISOHIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (hipEvent_t &e : ctx->tail_ev)
      MISOHIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  hipEvent_t *ev_xy = ctx->tail_ev;
  hipEvent_t ev_fork = ctx->tail_ev[TAIL_SLABS], ev_jb = ctx->tail_ev[TAIL_SLABS + 1];
  hipEvent_t ev_jz = ctx->tail_ev[TAIL_SLABS + 2];
  hipStream_t sb = ctx->tail_side[0], sz = ctx->tail_side[1];
  if (!pz) MISOHIP_TRY(launch_ghost_fill<R>(ctx->stream, fld<R>(dst), sh, b, 4));
  // one record for both walks, from the fork to the joins
  Timed t_(ctx, MISOHIP_KIND_AV_X);
  if (fuse_cfl) MISOHIP_TRY(launch_cfl_init(ctx->stream, ctx->d_slot));
  MISOHIP_TRY(hipEventRecord(ev_fork, ctx->stream));
  MISOHIP_TRY(hipStreamWaitEvent(sb, ev_fork, 0));
  MISOHIP_TRY(hipStreamWaitEvent(sz, ev_fork, 0));
  AvScalars xs = av_scalars<R>(p, dt, ctx->tile.av_xy_chunk);
  xs.wrap = wrap;
  xs.xy_waves = ctx->tile.av_fuse_xy == 4 ? 2 : 3;
  R *cc = static_cast<R *>(w->cc);
  for (int s = 0; s < P.ns; ++s) {
    if (P.xyn[s] > 0) {  // the last slabs of a short grid can be empty
      xs.part0 = P.xy0[s];
      xs.nparts = P.xyn[s];
      MISOHIP_TRY(launch_av_sweep_xy<R>(sb, cfld<R>(cur), fld<R>(dst), cc,
                                        static_cast<const R *>(g->dxi),
                                        static_cast<const R *>(g->dyi), sh, b, xs));
    }
    MISOHIP_TRY(hipEventRecord(ev_xy[s], sb));
  }
  AvScalars zs = av_scalars<R>(p, dt, ctx->tile.sweep_chunk);
  zs.wrap = wrap;
  CflFuse<R> cf;
  cf.dx = static_cast<const R *>(g->dx);
  cf.dy = static_cast<const R *>(g->dy);
  cf.dz = static_cast<const R *>(g->dz);
  cf.gm = (double)(R)p->gm;
  cf.cfl = (double)(R)p->cfl_number;
  cf.slot = ctx->d_slot;
  cf.heat_q = heat_q;
  cf.heat_dt = (double)dt;
  for (int s = 0; s < P.ns; ++s) {
    if (P.z_xy[s] >= 0) MISOHIP_TRY(hipStreamWaitEvent(sz, ev_xy[P.z_xy[s]], 0));
    zs.part0 = s == 0 ? 0 : P.c1[s - 1];
    zs.nparts = P.c1[s] - zs.part0;
    if (fuse_cfl)
      MISOHIP_TRY(launch_av_last_sweep_cfl<R>(sz, cfld<R>(dst), fld<R>(qq), cc,
                                              static_cast<const R *>(g->dyi),
                                              static_cast<const R *>(g->dzi), sh, zs, cf));
    else
      MISOHIP_TRY(launch_av_sweep<R>(sz, 2, cfld<R>(dst), fld<R>(qq), cc, cc,
                                     static_cast<const R *>(g->dzi), sh, zs));
  }
  MISOHIP_TRY(hipEventRecord(ev_jb, sb));
  MISOHIP_TRY(hipEventRecord(ev_jz, sz));
  MISOHIP_TRY(hipStreamWaitEvent(ctx->stream, ev_jb, 0));
  MISOHIP_TRY(hipStreamWaitEvent(ctx->stream, ev_jz, 0));
  return MISOHIP_OK;
}

template <typename R>
int update_t(misohip_ctx *ctx, const misohip_grid *g, const misohip_params *p,
             const misohip_bc *bc, const misohip_fields *qq, const misohip_work *w, R dt,
             bool fuse_cfl = false, const double *heat_q = nullptr,
             bool fill_periodic = true) {
  const R cfl_number = (R)p->cfl_number;
  const R min_dxyz = (R)g->min_dxyz;
  // divb_parameters_set (337-341)
  const R ch_divb = 0.8 * cfl_number * min_dxyz / dt;
  const R ch2 = ch_divb * ch_divb;
  const R tau = 2.0 * dt;
  const R dts[4] = {(R)(dt / 4.0), (R)(dt / 3.0), (R)(dt / 2.0), dt};
  const BcDev b = make_bc(bc);
  const Shape sh = make_shape(g);
  // In 3D, with the x and y AV sweeps as one walk, a z-march RK variant and
  // no user boundary, the RK stages, the x-y walk and the z walk read each
  // ghost cell of a periodic axis at its periodic source (the value the fill
  // would copy there): the fills between those passes leave the periodic
  // faces out (none at all in a periodic box), and the step's last fill
  // writes every face as before.
  const bool fuse_xy = g->is && g->js && g->ks && ctx->tile.av_fuse_xy && g->j_size >= 2 &&
                       !ctx->bc_fn;
  int wrap = 0;
  if (fuse_xy && ctx->tile.rk_variant >= 4 && ctx->tile.rk_variant <= 6)
    for (int a = 0; a < 3; ++a)
      if (bc->periodic[a] == 1) wrap |= 1 << a;
  // stage 1 reads qq as both orgn and argm; later stages ping-pong between
  // the two workspace field sets.
  const misohip_fields *argm = qq;
  const misohip_fields *bufs[2] = {&w->rslt, &w->argm};
  for (int s = 0; s < 4; ++s) {
    const misohip_fields *out = s == 3 ? &w->rslt : bufs[s & 1];
    if (s == 3 && argm == &w->rslt) out = &w->argm;
    const R damp = (R)std::exp(-dts[s] / tau);
    int rc = rk_stage_t<R>(ctx, g, p, qq, argm, out, dts[s], ch2, damp, wrap);
    if (rc) return rc;
    rc = ghost_t<R>(ctx, out, sh, b, 7 & ~wrap);
    if (rc) return rc;
    argm = out;
  }
  // AV: x sweep reads the post-RK state and stores cc; y and z read cc.  In
  // 2D runs the z sweep (extent one, pointwise) is applied by the y sweep to
  // each of its results before they are stored.
  // In 3D runs the x and y sweeps can run as one walk (tile choice
  // av_fuse_xy): the y walk x-sweeps every row it takes in.
  // Neither fusion runs with a user boundary: the reference applies the
  // boundary after every sweep (mhd_integrator_host.hpp:286-299), so the y
  // sweep reads the callback's ghost rows of the x-swept state and the z
  // sweep the callback's output of the y sweep.
  const bool fuse_z = g->is && g->js && !g->ks && !ctx->bc_fn;
  const misohip_fields *cur = argm;
  const misohip_fields *other = cur == &w->argm ? &w->rslt : &w->argm;
  if (fuse_xy && ctx->tile.av_fuse_xy >= 3 && ((wrap & 4) || bc->periodic[2] == 2)) {
    const int rc = tail_overlap_t<R>(ctx, g, p, b, sh, dt, cur, other, qq, w, wrap, fuse_cfl,
                                     heat_q);
    if (rc != TAIL_NOT_TAKEN) {
      if (rc) return rc;
      return ghost_t<R>(ctx, qq, sh, b, fill_periodic ? 7 : 7 & ~wrap);
    }
  }
  for (int dir = 0; dir < 3; ++dir) {
    if (dir == 1 && fuse_xy) continue;
    const misohip_fields *dst = (dir == 2 || (dir == 1 && fuse_z)) ? qq : other;
    int rc;
    const bool last = dir == 2 || (dir == 1 && fuse_z);
    if (dir == 0 && fuse_xy) {
      Timed t_(ctx, MISOHIP_KIND_AV_X);
      AvScalars sc = av_scalars<R>(p, dt, ctx->tile.av_xy_chunk);
      sc.wrap = wrap;
      sc.xy_waves = ctx->tile.av_fuse_xy == 2 || ctx->tile.av_fuse_xy == 4 ? 2 : 3;
      MISOHIP_TRY(launch_av_sweep_xy<R>(ctx->stream, cfld<R>(cur), fld<R>(dst),
                                        static_cast<R *>(w->cc), static_cast<const R *>(g->dxi),
                                        static_cast<const R *>(g->dyi), sh, b, sc));
      rc = MISOHIP_OK;
    } else if (last && fuse_cfl) {
      Timed t_(ctx, fuse_z ? MISOHIP_KIND_AV_Y : MISOHIP_KIND_AV_Z);
      CflFuse<R> cf;
      cf.dx = static_cast<const R *>(g->dx);
      cf.dy = static_cast<const R *>(g->dy);
      cf.dz = static_cast<const R *>(g->dz);
      cf.gm = (double)(R)p->gm;
      cf.cfl = (double)(R)p->cfl_number;
      cf.slot = ctx->d_slot;
      cf.heat_q = heat_q;
      cf.heat_dt = (double)dt;
      MISOHIP_TRY(launch_cfl_init(ctx->stream, ctx->d_slot));
      AvScalars sc = av_scalars<R>(p, dt, ctx->tile.sweep_chunk);
      sc.wrap = fuse_xy ? wrap : 0;
      MISOHIP_TRY(launch_av_last_sweep_cfl<R>(
          ctx->stream, cfld<R>(cur), fld<R>(dst), static_cast<const R *>(w->cc),
          static_cast<const R *>(g->dyi), static_cast<const R *>(g->dzi), sh, sc, cf));
      rc = MISOHIP_OK;
    } else if (dir == 1 && fuse_z) {
      Timed t_(ctx, MISOHIP_KIND_AV_Y);
      MISOHIP_TRY(launch_av_sweep_y_degen_z<R>(
          ctx->stream, cfld<R>(cur), fld<R>(dst), static_cast<const R *>(w->cc),
          static_cast<const R *>(g->dyi), static_cast<const R *>(g->dzi), sh,
          av_scalars<R>(p, dt, ctx->tile.sweep_chunk)));
      rc = MISOHIP_OK;
    } else {
      rc = av_sweep_t<R>(ctx, g, p, dir, dt, cur, dst, w->cc, fuse_xy ? wrap : 0);
    }
    if (rc) return rc;
    // after the x-y walk only the z walk reads the state, along z alone; the
    // next step of a run reads the periodic faces of this step's result at
    // their sources, so a run fills them after its last step only
    const int axes = last ? (fill_periodic ? 7 : 7 & ~wrap)
                          : (dir == 0 && fuse_xy && wrap) ? 4 & ~wrap : 7;
    rc = ghost_t<R>(ctx, dst, sh, b, axes);
    if (rc) return rc;
    if (dir == 1 && fuse_z) break;
    other = cur;
    cur = dst;
  }
  return MISOHIP_OK;
}

// ModelBase::run's loop body nsteps times; from the second step on, dt comes
// from the CFL reduction the previous step's last AV sweep carried out --
// except with a user boundary, whose callback runs after that sweep and may
// write interior cells, so the CFL pass then reads the state it leaves, as
// ModelBase::update does (mhd_model_base.hpp:57-62).
template <typename R>
int run_t(misohip_ctx *ctx, const misohip_grid *g, const misohip_params *p, const misohip_bc *bc,
          const misohip_fields *qq, const misohip_work *w, int nsteps, double *dts) {
  const bool can_fuse = g->is && g->js && !ctx->bc_fn;  // 3D, or 2D in x-y
  double d = 0.0;
  for (int s = 0; s < nsteps; ++s) {
    if (s == 0 || !can_fuse) {
      int rc = cfl_t<R>(ctx, g, p, qq, &d);
      if (rc) return rc;
    } else {
      MISOHIP_TRY(hipMemcpyAsync(ctx->h_slot, ctx->d_slot, sizeof(unsigned),
                                 hipMemcpyDeviceToHost, ctx->stream));
      MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
      float f;
      std::memcpy(&f, ctx->h_slot, sizeof(float));
      d = (double)(R)f;
    }
    if (dts) dts[s] = d;
    int rc = update_t<R>(ctx, g, p, bc, qq, w, (R)d, can_fuse && s + 1 < nsteps, nullptr,
                         s + 1 == nsteps);
    if (rc) return rc;
  }
  return MISOHIP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Workspace of misohip_update_noalloc: the 19 arrays of misohip_work (argm,
// rslt, cc) carved out of one file-scope device array, so that a caller that
// cannot allocate (stream capture, serving loops) can still step a grid.
// Every GPU holds its own copy of the array, at its own address, for the life
// of the process; misohip_create looks that address up once per device with
// the device current.  On the host the name `scratch` is not a device
// address: only the looked-up pointer is ever passed to a kernel.
static __device__ __align__(256) unsigned char scratch[MISOHIP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
void *g_scratch_ptr[kMaxScratchDevices];
std::mutex g_scratch_mu;

// the current device's copy of `scratch` (called with ctx->device current)
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return MISOHIP_OK;  // noalloc refuses it
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return MISOHIP_OK;
  void *p = nullptr;
  MISOHIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return MISOHIP_E_HIP;
  g_scratch_ptr[device] = p;
  return MISOHIP_OK;
}

// argm, rslt and cc of a grid in the device's scratch array, each array on a
// 256-byte boundary; MISOHIP_E_RANGE when the 19 arrays do not fit
int noalloc_work(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt, misohip_work *w) {
  if (ctx->device < 0 || ctx->device >= kMaxScratchDevices) return MISOHIP_E_RANGE;
  void *base;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[ctx->device];
  }
  if (!base) return MISOHIP_E_ARG;  // a context not made by misohip_create
  const size_t elt = dt == MISOHIP_F64 ? sizeof(double) : sizeof(float);
  const size_t n = (size_t)g->i_total * g->j_total * g->k_total;
  const size_t arr = (n * elt + 255) / 256 * 256;
  if (arr > MISOHIP_NOALLOC_SCRATCH_BYTES / 19) return MISOHIP_E_RANGE;
  unsigned char *b = static_cast<unsigned char *>(base);
  for (int m = 0; m < 9; ++m) {
    w->argm.q[m] = b + (size_t)m * arr;
    w->rslt.q[m] = b + (size_t)(9 + m) * arr;
  }
  w->cc = b + (size_t)18 * arr;
  return MISOHIP_OK;
}
}  // namespace

extern "C" {

const char *misohip_strerror(int code) {
  switch (code) {
    case MISOHIP_OK: return "success";
    case MISOHIP_E_HIP: return g_last_hip_error[0] ? g_last_hip_error : "HIP runtime error";
    case MISOHIP_E_ARG: return "invalid argument or grid shape";
    case MISOHIP_E_DTYPE: return "unknown dtype";
    case MISOHIP_E_NOMEM: return "device allocation failed";
    case MISOHIP_E_RANGE: return "size beyond the supported maximum";
    case MISOHIP_E_CALLBACK: return "the user boundary callback failed";
    default: return "unknown error";
  }
}

int misohip_create(int device, void *stream, misohip_ctx **out) {
  if (!out) return MISOHIP_E_ARG;
  *out = nullptr;
  MISOHIP_TRY(hipSetDevice(device));
  misohip_ctx *c = new (std::nothrow) misohip_ctx;
  if (!c) return MISOHIP_E_NOMEM;
  c->device = device;
  c->stream = static_cast<hipStream_t>(stream);
  c->d_slot = nullptr;
  c->h_slot = nullptr;
  hipError_t e = hipMalloc(&c->d_slot, 256);
  if (e == hipSuccess) e = hipHostMalloc(&c->h_slot, 256, hipHostMallocDefault);
  if (e != hipSuccess) {
    if (c->d_slot) (void)hipFree(c->d_slot);
    delete c;
    return hip_fail(e);
  }
  // the device's address of the no-allocation workspace, looked up once
  if (int rc = scratch_setup(device)) {
    (void)hipFree(c->d_slot);
    (void)hipHostFree(c->h_slot);
    delete c;
    return rc;
  }
  *out = c;
  return MISOHIP_OK;
}

int misohip_set_tiles(misohip_ctx *ctx, const misohip_tiles *t) {
  if (!ctx || !t) return MISOHIP_E_ARG;
  if (t->rk_pencil_bx < 1 || t->rk_pencil_by < 1 || t->sweep_chunk < 4 || t->sweep_chunk > 4096 ||
      t->rk_variant < 0 || t->rk_variant > 6 || t->rk_zchunk < 1 || t->rk_zchunk > 4096 ||
      t->av_fuse_xy < 0 || t->av_fuse_xy > 4 || t->av_xy_chunk < 4 || t->av_xy_chunk > 4096)
    return MISOHIP_E_ARG;
  ctx->tile = *t;
  return MISOHIP_OK;
}

int misohip_get_tiles(misohip_ctx *ctx, misohip_tiles *t) {
  if (!ctx || !t) return MISOHIP_E_ARG;
  *t = ctx->tile;
  return MISOHIP_OK;
}

int misohip_profile_enable(misohip_ctx *ctx, int on) {
  if (!ctx) return MISOHIP_E_ARG;
  ctx->profiling = on != 0;
  return MISOHIP_OK;
}

int misohip_profile_read(misohip_ctx *ctx, double *ms, int *launches, int nkinds) {
  if (!ctx || nkinds < 0 || (nkinds > 0 && (!ms || !launches))) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < nkinds; ++k) {
    ms[k] = 0.0;
    launches[k] = 0;
  }
  for (size_t r = 0; r < ctx->rec_kind.size(); ++r) {
    float t = 0.f;
    MISOHIP_TRY(hipEventElapsedTime(&t, ctx->rec_start[r], ctx->rec_stop[r]));
    const int k = ctx->rec_kind[r];
    if (k < nkinds) {
      ms[k] += t;
      launches[k] += 1;
    }
    ctx->pool.push_back(ctx->rec_start[r]);
    ctx->pool.push_back(ctx->rec_stop[r]);
  }
  ctx->rec_kind.clear();
  ctx->rec_start.clear();
  ctx->rec_stop.clear();
  return MISOHIP_OK;
}

int misohip_destroy(misohip_ctx *ctx) {
  if (!ctx) return MISOHIP_E_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (hipEvent_t e : ctx->pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->rec_start) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->rec_stop) (void)hipEventDestroy(e);
  if (ctx->d_slot) (void)hipFree(ctx->d_slot);
  if (ctx->h_slot) (void)hipHostFree(ctx->h_slot);
  if (ctx->rt_kap_y) (void)hipFree(ctx->rt_kap_y);
  if (ctx->rt_src_y) (void)hipFree(ctx->rt_src_y);
  if (ctx->rt_parts) (void)hipFree(ctx->rt_parts);
  if (ctx->rt_side) {
    (void)hipStreamSynchronize(ctx->rt_side);
    (void)hipStreamDestroy(ctx->rt_side);
  }
  if (ctx->rt_fork) (void)hipEventDestroy(ctx->rt_fork);
  if (ctx->rt_join) (void)hipEventDestroy(ctx->rt_join);
  for (hipStream_t st : ctx->tail_side)
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  for (hipEvent_t e : ctx->tail_ev)
    if (e) (void)hipEventDestroy(e);
  delete ctx;
  return MISOHIP_OK;
}

int misohip_set_bc_callback(misohip_ctx *ctx, misohip_bc_fn fn, void *user) {
  if (!ctx) return MISOHIP_E_ARG;
  ctx->bc_fn = fn;
  ctx->bc_user = fn ? user : nullptr;
  return MISOHIP_OK;
}

int misohip_set_source_callback(misohip_ctx *ctx, misohip_src_fn fn, void *user,
                                void *const arrays[4]) {
  if (!ctx) return MISOHIP_E_ARG;
  if (fn) {
    if (!arrays) return MISOHIP_E_ARG;
    for (int d = 0; d < 4; ++d)
      if (!arrays[d]) return MISOHIP_E_ARG;
  }
  ctx->src_fn = fn;
  ctx->src_user = fn ? user : nullptr;
  for (int d = 0; d < 4; ++d) ctx->src_arr[d] = fn ? arrays[d] : nullptr;
  return MISOHIP_OK;
}

int misohip_get_source_callback(misohip_ctx *ctx, misohip_src_fn *fn, void **user,
                                void *arrays[4]) {
  if (!ctx || !fn || !user || !arrays) return MISOHIP_E_ARG;
  *fn = ctx->src_fn;
  *user = ctx->src_user;
  for (int d = 0; d < 4; ++d) arrays[d] = ctx->src_arr[d];
  return MISOHIP_OK;
}

int misohip_get_bc_callback(misohip_ctx *ctx, misohip_bc_fn *fn, void **user) {
  if (!ctx || !fn || !user) return MISOHIP_E_ARG;
  *fn = ctx->bc_fn;
  *user = ctx->bc_user;
  return MISOHIP_OK;
}

int misohip_set_stream(misohip_ctx *ctx, void *stream) {
  if (!ctx) return MISOHIP_E_ARG;
  ctx->stream = static_cast<hipStream_t>(stream);
  return MISOHIP_OK;
}

int misohip_synchronize(misohip_ctx *ctx) {
  if (!ctx) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
  return MISOHIP_OK;
}

int misohip_alloc(misohip_ctx *ctx, size_t bytes, void **out) {
  if (!ctx || !out) return MISOHIP_E_ARG;
  *out = nullptr;
  if (bytes == 0) return MISOHIP_OK;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  MISOHIP_TRY(hipMalloc(out, bytes));
  return MISOHIP_OK;
}

int misohip_free(misohip_ctx *ctx, void *p) {
  if (!ctx) return MISOHIP_E_ARG;
  if (!p) return MISOHIP_OK;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
  MISOHIP_TRY(hipFree(p));
  return MISOHIP_OK;
}

int misohip_memcpy(misohip_ctx *ctx, void *dst, const void *src, size_t bytes, int kind) {
  if (!ctx) return MISOHIP_E_ARG;
  hipMemcpyKind k;
  switch (kind) {
    case MISOHIP_H2D: k = hipMemcpyHostToDevice; break;
    case MISOHIP_D2H: k = hipMemcpyDeviceToHost; break;
    case MISOHIP_D2D: k = hipMemcpyDeviceToDevice; break;
    default: return MISOHIP_E_ARG;
  }
  if (bytes == 0) return MISOHIP_OK;
  if (!dst || !src) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  MISOHIP_TRY(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
  return MISOHIP_OK;
}

int misohip_copy_to_device(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                           const void *host, void *dev) {
  if (!ctx || !host || !dev) return MISOHIP_E_ARG;
  if (int rc = check_grid(g)) return rc;
  if (dt != MISOHIP_F64 && dt != MISOHIP_F32) return MISOHIP_E_DTYPE;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  const size_t es = dt == MISOHIP_F64 ? 8 : 4;
  const size_t n = (size_t)g->i_total * g->j_total * g->k_total;
  void *tmp = nullptr;
  MISOHIP_TRY(hipMalloc(&tmp, n * es));
  hipError_t e = hipMemcpyAsync(tmp, host, n * es, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = dt == MISOHIP_F64
            ? launch_transpose<double>(ctx->stream, (const double *)tmp, (double *)dev,
                                       g->i_total, g->j_total, g->k_total, true)
            : launch_transpose<float>(ctx->stream, (const float *)tmp, (float *)dev, g->i_total,
                                      g->j_total, g->k_total, true);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(e);
  return MISOHIP_OK;
}

int misohip_copy_to_host(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                         const void *dev, void *host) {
  if (!ctx || !host || !dev) return MISOHIP_E_ARG;
  if (int rc = check_grid(g)) return rc;
  if (dt != MISOHIP_F64 && dt != MISOHIP_F32) return MISOHIP_E_DTYPE;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  const size_t es = dt == MISOHIP_F64 ? 8 : 4;
  const size_t n = (size_t)g->i_total * g->j_total * g->k_total;
  void *tmp = nullptr;
  MISOHIP_TRY(hipMalloc(&tmp, n * es));
  hipError_t e = dt == MISOHIP_F64
                     ? launch_transpose<double>(ctx->stream, (const double *)dev, (double *)tmp,
                                                g->i_total, g->j_total, g->k_total, false)
                     : launch_transpose<float>(ctx->stream, (const float *)dev, (float *)tmp,
                                               g->i_total, g->j_total, g->k_total, false);
  if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, n * es, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(e);
  return MISOHIP_OK;
}

#define MISOHIP_COMMON_CHECKS(ctx, g, dt)                            \
  if (!ctx) return MISOHIP_E_ARG;                                    \
  if (int rc_ = check_grid(g)) return rc_;                           \
  if (dt != MISOHIP_F64 && dt != MISOHIP_F32) return MISOHIP_E_DTYPE; \
  MISOHIP_TRY(hipSetDevice(ctx->device));

int misohip_ghost_fill(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                       const misohip_fields *qq, const misohip_bc *bc) {
  MISOHIP_COMMON_CHECKS(ctx, g, dt)
  if (check_fields(qq) || bc_invalid(ctx, bc)) return MISOHIP_E_ARG;
  const BcDev b = make_bc(bc);
  if (dt == MISOHIP_F64) return ghost_t<double>(ctx, qq, make_shape(g), b);
  return ghost_t<float>(ctx, qq, make_shape(g), b);
}

int misohip_rk_stage(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                     const misohip_params *p, const misohip_fields *orgn,
                     const misohip_fields *argm, const misohip_fields *rslt, double dt_s,
                     double ch2, double damp) {
  MISOHIP_COMMON_CHECKS(ctx, g, dt)
  if (!p || check_fields(orgn) || check_fields(argm) || check_fields(rslt)) return MISOHIP_E_ARG;
  if (src_invalid(ctx, p)) return MISOHIP_E_ARG;
  if (dt == MISOHIP_F64) return rk_stage_t<double>(ctx, g, p, orgn, argm, rslt, dt_s, ch2, damp);
  return rk_stage_t<float>(ctx, g, p, orgn, argm, rslt, (float)dt_s, (float)ch2, (float)damp);
}

int misohip_char_vel(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                     const misohip_params *p, const misohip_fields *qq, void *cc) {
  MISOHIP_COMMON_CHECKS(ctx, g, dt)
  if (!p || check_fields(qq) || !cc) return MISOHIP_E_ARG;
  const int n = g->i_total * g->j_total * g->k_total;
  if (dt == MISOHIP_F64)
    MISOHIP_TRY(launch_char_vel<double>(ctx->stream, cfld<double>(qq), (double *)cc, n,
                                        av_scalars<double>(p, 0.0)));
  else
    MISOHIP_TRY(launch_char_vel<float>(ctx->stream, cfld<float>(qq), (float *)cc, n,
                                       av_scalars<float>(p, 0.0f)));
  return MISOHIP_OK;
}

int misohip_av_sweep(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                     const misohip_params *p, int dir, double dt_step, const misohip_fields *qq,
                     const misohip_fields *rslt, const void *cc) {
  MISOHIP_COMMON_CHECKS(ctx, g, dt)
  if (!p || check_fields(qq) || check_fields(rslt) || !cc || dir < 0 || dir > 2)
    return MISOHIP_E_ARG;
  if (dt == MISOHIP_F64) return av_sweep_t<double>(ctx, g, p, dir, dt_step, qq, rslt, (void *)cc);
  return av_sweep_t<float>(ctx, g, p, dir, (float)dt_step, qq, rslt, (void *)cc);
}

int misohip_cfl(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                const misohip_params *p, const misohip_fields *qq, double *dt_out) {
  MISOHIP_COMMON_CHECKS(ctx, g, dt)
  if (!p || check_fields(qq) || !dt_out) return MISOHIP_E_ARG;
  if (dt == MISOHIP_F64) return cfl_t<double>(ctx, g, p, qq, dt_out);
  return cfl_t<float>(ctx, g, p, qq, dt_out);
}

int misohip_update(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                   const misohip_params *p, const misohip_bc *bc, const misohip_fields *qq,
                   const misohip_work *w, double dt_step) {
  MISOHIP_COMMON_CHECKS(ctx, g, dt)
  if (!p || bc_invalid(ctx, bc) || !w || check_fields(qq) || check_fields(&w->argm) ||
      check_fields(&w->rslt) ||
      !w->cc)
    return MISOHIP_E_ARG;
  if (src_invalid(ctx, p)) return MISOHIP_E_ARG;
  if (!(dt_step > 0.0)) return MISOHIP_E_ARG;
  if (dt == MISOHIP_F64) return update_t<double>(ctx, g, p, bc, qq, w, dt_step);
  return update_t<float>(ctx, g, p, bc, qq, w, (float)dt_step);
}

int misohip_update_heated(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                          const misohip_params *p, const misohip_bc *bc,
                          const misohip_fields *qq, const misohip_work *w, double dt_step,
                          const void *qrad, double *next_dt) {
  MISOHIP_COMMON_CHECKS(ctx, g, dt)
  if (!p || bc_invalid(ctx, bc) || !w || check_fields(qq) || check_fields(&w->argm) ||
      check_fields(&w->rslt) || !w->cc || !qrad || dt != MISOHIP_F64)
    return MISOHIP_E_ARG;
  if (src_invalid(ctx, p)) return MISOHIP_E_ARG;
  if (!(dt_step > 0.0)) return MISOHIP_E_ARG;
  const double *q = static_cast<const double *>(qrad);
  // the last AV sweep heats each cell it stores and reduces the CFL; not with
  // a user boundary, whose callback must see the unheated state after the
  // update and the heated one before the CFL
  if (g->is && g->js && !ctx->bc_fn) {
    int rc = update_t<double>(ctx, g, p, bc, qq, w, dt_step, true, q);
    if (rc) return rc;
    if (next_dt) {
      MISOHIP_TRY(hipMemcpyAsync(ctx->h_slot, ctx->d_slot, sizeof(unsigned),
                                 hipMemcpyDeviceToHost, ctx->stream));
      MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
      float f;
      std::memcpy(&f, ctx->h_slot, sizeof(float));
      *next_dt = (double)f;
    }
    return MISOHIP_OK;
  }
  // otherwise the same operations one after another
  int rc = update_t<double>(ctx, g, p, bc, qq, w, dt_step);
  if (rc) return rc;
  const Shape sh = make_shape(g);
  {
    Timed t_(ctx, MISOHIP_KIND_RT);
    MISOHIP_TRY(launch_rt_heat(ctx->stream, static_cast<const double *>(qq->q[0]),
                               static_cast<double *>(qq->q[7]), q, sh, dt_step));
  }
  rc = ghost_t<double>(ctx, qq, sh, make_bc(bc));
  if (rc) return rc;
  if (next_dt) return cfl_t<double>(ctx, g, p, qq, next_dt);
  return MISOHIP_OK;
}


int misohip_run(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                const misohip_params *p, const misohip_bc *bc, const misohip_fields *qq,
                const misohip_work *w, int nsteps, double *dts) {
  MISOHIP_COMMON_CHECKS(ctx, g, dt)
  if (!p || bc_invalid(ctx, bc) || !w || check_fields(qq) || check_fields(&w->argm) ||
      check_fields(&w->rslt) || !w->cc || nsteps < 0)
    return MISOHIP_E_ARG;
  if (src_invalid(ctx, p)) return MISOHIP_E_ARG;
  if (dt == MISOHIP_F64) return run_t<double>(ctx, g, p, bc, qq, w, nsteps, dts);
  return run_t<float>(ctx, g, p, bc, qq, w, nsteps, dts);
}

int misohip_step(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                 const misohip_params *p, const misohip_bc *bc, const misohip_fields *qq,
                 const misohip_work *w, double *dt_out) {
  double d = 0.0;
  int rc = misohip_cfl(ctx, g, dt, p, qq, &d);
  if (rc) return rc;
  if (dt_out) *dt_out = d;
  return misohip_update(ctx, g, dt, p, bc, qq, w, d);
}

int misohip_stream_bench(misohip_ctx *ctx, int kind, const void *src, void *dst, size_t bytes) {
  if (!ctx || kind < 0 || kind > 2 || bytes % 16) return MISOHIP_E_ARG;
  if ((kind != 2 && !src) || (kind != 1 && !dst)) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  MISOHIP_TRY(launch_stream(ctx->stream, src, dst, bytes, kind,
                            reinterpret_cast<double *>(ctx->d_slot)));
  return MISOHIP_OK;
}

int misohip_update_noalloc(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                           const misohip_params *p, const misohip_bc *bc,
                           const misohip_fields *qq, double dt_step) {
  MISOHIP_COMMON_CHECKS(ctx, g, dt)
  if (!p || bc_invalid(ctx, bc) || check_fields(qq)) return MISOHIP_E_ARG;
  if (src_invalid(ctx, p)) return MISOHIP_E_ARG;
  if (!(dt_step > 0.0)) return MISOHIP_E_ARG;
  misohip_work w;
  if (int rc = noalloc_work(ctx, g, dt, &w)) return rc;
  // profiling records events from a pool that may have to grow; this entry
  // point must not create any, so it runs untimed
  // (and takes the overlapped AV walks only on streams an earlier call made)
  const bool prof = ctx->profiling;
  ctx->profiling = false;
  ctx->no_alloc = true;
  const int rc = dt == MISOHIP_F64 ? update_t<double>(ctx, g, p, bc, qq, &w, dt_step)
                                   : update_t<float>(ctx, g, p, bc, qq, &w, (float)dt_step);
  ctx->no_alloc = false;
  ctx->profiling = prof;
  return rc;
}

}  // extern "C"
