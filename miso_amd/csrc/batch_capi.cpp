// C ABI of the batched ensemble (include/misohip.h, misohip_batch_*): many
// independent runs of uneven size advanced together, one launch per phase
// for all of them, each with its own dt kept on the device, and the whole
// step (or several) captured once in a hipGraph and replayed.
//
// No counterpart in the reference, which runs one model per process
// (miso/include/miso/mhd_model_base.hpp:89-117).  Each instance computes
// exactly what misohip_step computes for it alone: the kernels are the
// single-run kernels' block functions, the per-instance scalars are formed
// with the host path's arithmetic (batch_params_kernel), and exp(-dt_s/tau)
// is taken from a table of host-library values.
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "ctx_internal.hpp"
#include "launch.hpp"

using namespace misohip;

namespace {

enum Phase { P_RK = 0, P_GHOST, P_CFL, P_AVX, P_AVY, P_AVZ, P_COUNT };

template <typename R> struct BatchT {
  int n = 0;
  int key = 0;          // active axes, (is ? 4 : 0) | (js ? 2 : 0) | (ks ? 1 : 0)
  int mode_y = 1, mode_z = 1;
  misohip_params p{};
  BcDev bc{};
  int total[P_COUNT] = {};
  BatchInst<R> *d_inst = nullptr;
  BatchDyn<R> *d_dyn = nullptr;
  int *d_start = nullptr;  // P_COUNT arrays of n + 1 block offsets
  DampTable<R> tab{};
  std::vector<BatchInst<R>> h_inst;  // host copy of the instance table
};

}  // namespace

struct misohip_batch {
  misohip_dtype dtype;
  void *impl;  // BatchT<double> or BatchT<float>
  hipStream_t stream = nullptr;  // capture / replay stream (default priority)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int steps_per_graph = 0;
};

namespace {

Shape shape_of(const misohip_grid *g) {
  Shape s;
  s.It = g->i_total; s.Jt = g->j_total; s.Kt = g->k_total;
  s.mi = g->i_margin; s.mj = g->j_margin; s.mk = g->k_margin;
  s.is = g->is; s.js = g->js; s.ks = g->ks;
  return s;
}

// block counts of one instance per phase, matching the kernels' decoding
void phase_blocks(const Shape &sh, int chunk, int mode_y, int mode_z, long out[P_COUNT]) {
  const long ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nk = sh.Kt - 2 * sh.mk;
  if (sh.is && sh.js && !sh.ks)  // x-y runs: y-walk waves, four per block
    out[P_RK] = (((ni + RK_YWALK_OUT - 1) / RK_YWALK_OUT) *
                     ((nj + RK_YWALK_ROWS - 1) / RK_YWALK_ROWS) + 3) / 4;
  else
    out[P_RK] = ((ni + 63) / 64) * ((nj + 3) / 4) * nk;
  const long nx = 2L * sh.mi * sh.Jt * sh.Kt, ny = 2L * sh.mj * ni * sh.Kt,
             nz = 2L * sh.mk * ni * nj;
  out[P_GHOST] = (nx + ny + nz + 255) / 256;
  long c = (ni * nj * nk + 255) / 256;
  out[P_CFL] = c < 1 ? 1 : (c > 512 ? 512 : c);
  const long j_n = nj + 2 * sh.js, k_n = nk + 2 * sh.ks;
  out[P_AVX] = ((ni + 59) / 60) * ((j_n * k_n + 3) / 4);
  // walking sweeps: chunks x x-blocks x other-axis blocks (walk_waves_x)
  auto walk = [&](long n_dir, long n_oth) {
    const long wx = walk_waves_x((int)n_oth), wo = 4 / wx;
    return ((n_dir + chunk - 1) / chunk) * (((ni + 63) / 64 + wx - 1) / wx) *
           ((n_oth + wo - 1) / wo);
  };
  if (mode_y == 1)
    out[P_AVY] = walk(nj, nk);
  else
    out[P_AVY] = ((ni + 63) / 64) * ((nj + 3) / 4) * nk;
  if (mode_z == 1)
    out[P_AVZ] = walk(nk, nj);
  else
    out[P_AVZ] = ((ni + 63) / 64) * ((nj + 3) / 4) * nk;
}

template <typename R> void fill_table(DampTable<R> &t) {
  // the host path evaluates (R)std::exp(q) with q of type R
  t.exact[0] = (R)std::exp(R(-0.125));
  t.exact[1] = R(0);
  t.exact[2] = (R)std::exp(R(-0.25));
  t.exact[3] = (R)std::exp(R(-0.5));
  t.q0 = R(-1.0 / 6.0);
  R q = t.q0;
  for (int k = 0; k < 16; ++k) q = std::nextafter(q, R(0));  // 16 steps toward zero
  for (int k = 0; k < 33; ++k) {
    t.third[k] = (R)std::exp(q);
    q = std::nextafter(q, R(-1));  // one ulp larger in magnitude
  }
}

// Block offsets of every phase from the instance table, and both tables
// to the device (synchronous).
template <typename R> int upload_layout(BatchT<R> *b) {
  const int n = b->n;
  std::vector<int> start((size_t)P_COUNT * (n + 1), 0);
  long tot[P_COUNT] = {};
  for (int m = 0; m < n; ++m) {
    long blk[P_COUNT];
    phase_blocks(b->h_inst[m].sh, b->h_inst[m].chunk, b->mode_y, b->mode_z, blk);
    for (int ph = 0; ph < P_COUNT; ++ph) {
      tot[ph] += blk[ph];
      if (tot[ph] > 2147483647L) return MISOHIP_E_RANGE;
      start[(size_t)ph * (n + 1) + m + 1] = (int)tot[ph];
    }
  }
  for (int ph = 0; ph < P_COUNT; ++ph) b->total[ph] = (int)tot[ph];
  MISOHIP_TRY(hipMemcpy(b->d_inst, b->h_inst.data(), sizeof(BatchInst<R>) * n,
                        hipMemcpyHostToDevice));
  MISOHIP_TRY(hipMemcpy(b->d_start, start.data(), sizeof(int) * start.size(),
                        hipMemcpyHostToDevice));
  return MISOHIP_OK;
}

template <typename R>
int create_t(misohip_ctx *ctx, int n, const misohip_grid *grids, const misohip_fields *qq,
             const misohip_work *work, const misohip_params *p, const misohip_bc *bc,
             BatchT<R> **out) {
  *out = nullptr;
  BatchT<R> *b = new (std::nothrow) BatchT<R>;
  if (!b) return MISOHIP_E_NOMEM;
  b->n = n;
  b->p = *p;
  const int chunk = ctx->tile.sweep_chunk < 4 ? 4 : ctx->tile.sweep_chunk;
  BcDev bd{};
  for (int a = 0; a < 3; ++a) {
    if (bc->periodic[a] == 2 || bc->periodic[a] == 3) {  // fixed and user boundaries are not batched
      delete b;
      return MISOHIP_E_ARG;
    }
    bd.periodic[a] = bc->periodic[a] ? 1 : 0;
    for (int f = 0; f < 9; ++f) bd.sign[a][f] = bc->sym_sign[a][f] >= 0 ? 1.0f : -1.0f;
  }
  b->bc = bd;
  std::vector<BatchInst<R>> &inst = b->h_inst;
  inst.resize(n);
  for (int m = 0; m < n; ++m) {
    const misohip_grid *g = &grids[m];
    const int key = (g->is ? 4 : 0) | (g->js ? 2 : 0) | (g->ks ? 1 : 0);
    // the x sweep needs an active x axis and every instance the same axes
    if (!(key & 4) || !(key & 2) || (m > 0 && key != b->key) || g->i_margin != 2 ||
        g->j_margin != 2 || (g->ks && g->k_margin != 2)) {
      delete b;
      return MISOHIP_E_ARG;
    }
    b->key = key;
    b->mode_y = 1;
    b->mode_z = g->ks ? 1 : 2;
    BatchInst<R> &I = inst[m];
    for (int f = 0; f < 9; ++f) {
      I.st[0].q[f] = static_cast<R *>(qq[m].q[f]);
      I.st[1].q[f] = static_cast<R *>(work[m].argm.q[f]);
      I.st[2].q[f] = static_cast<R *>(work[m].rslt.q[f]);
      if (!qq[m].q[f] || !work[m].argm.q[f] || !work[m].rslt.q[f]) {
        delete b;
        return MISOHIP_E_ARG;
      }
    }
    I.cc = static_cast<R *>(work[m].cc);
    I.dx = static_cast<const R *>(g->dx); I.dy = static_cast<const R *>(g->dy);
    I.dz = static_cast<const R *>(g->dz); I.dxi = static_cast<const R *>(g->dxi);
    I.dyi = static_cast<const R *>(g->dyi); I.dzi = static_cast<const R *>(g->dzi);
    I.sh = shape_of(g);
    I.min_dxyz = g->min_dxyz;
    // this run's y walk alone decides (the tuner's per-run choice replaces it)
    int nch = 0;
    walk_split(g->j_size, 64L * ((g->i_size + 63) / 64) * (g->ks ? g->k_size : 1), chunk, I.chunk,
               nch);
  }
  fill_table(b->tab);
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&b->d_inst), sizeof(BatchInst<R>) * n);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&b->d_dyn), sizeof(BatchDyn<R>) * n);
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void **>(&b->d_start), sizeof(int) * P_COUNT * (n + 1));
  if (e == hipSuccess) e = hipMemset(b->d_dyn, 0, sizeof(BatchDyn<R>) * n);
  int rc = MISOHIP_OK;
  if (e == hipSuccess) rc = upload_layout(b);
  if (rc) {
    (void)hipFree(b->d_inst);
    (void)hipFree(b->d_dyn);
    (void)hipFree(b->d_start);
    delete b;
    return rc;
  }
  if (e != hipSuccess) {
    (void)hipFree(b->d_inst);
    (void)hipFree(b->d_dyn);
    (void)hipFree(b->d_start);
    delete b;
    return hip_fail(e);
  }
  *out = b;
  return MISOHIP_OK;
}

template <typename R> void destroy_t(BatchT<R> *b) {
  if (!b) return;
  (void)hipFree(b->d_inst);
  (void)hipFree(b->d_dyn);
  (void)hipFree(b->d_start);
  delete b;
}

// One step of every instance: CFL -> step scalars -> 4 RK stages -> 3 AV
// sweeps, each followed by the ghost fill (the order of misohip_step).
template <typename R> hipError_t step_t(hipStream_t st, BatchT<R> *b) {
  const int n = b->n;
  auto S = [&](int ph) { return b->d_start + (size_t)ph * (n + 1); };
  hipError_t e = launch_batch_cfl<R>(st, b->d_inst, b->d_dyn, S(P_CFL), n, b->total[P_CFL],
                                     (double)(R)b->p.gm, (double)(R)b->p.cfl_number);
  if (e != hipSuccess) return e;
  e = launch_batch_params<R>(st, b->d_inst, b->d_dyn, n, b->p.cfl_number, b->tab);
  if (e != hipSuccess) return e;
  RkScalars sc{};
  sc.gm = (R)b->p.gm; sc.grav = (R)b->p.grav;
  sc.src_kind = b->p.src_kind; sc.grav_axis = b->p.grav_axis;
  // stage ping-pong of update_t: argm source and output set per stage
  const int in_set[4] = {0, 2, 1, 2}, out_set[4] = {2, 1, 2, 1};
  for (int s = 0; s < 4; ++s) {
    e = launch_batch_rk<R>(st, b->d_inst, b->d_dyn, S(P_RK), n, b->total[P_RK], s, in_set[s],
                           out_set[s], sc, b->key);
    if (e != hipSuccess) return e;
    e = launch_batch_ghost<R>(st, b->d_inst, S(P_GHOST), n, b->total[P_GHOST], out_set[s], b->bc);
    if (e != hipSuccess) return e;
  }
  AvScalars av{};
  av.ep = (R)b->p.ep; av.fh = (R)b->p.fh; av.gm = (R)b->p.gm;
  av.cs_fac = (R)b->p.cs_fac; av.ca_fac = (R)b->p.ca_fac; av.vv_fac = (R)b->p.vv_fac;
  av.chunk = 0;  // per instance (BatchInst::chunk)
  // x: argm(1) -> rslt(2); y: rslt(2) -> argm(1); z: argm(1) -> qq(0).  2D
  // runs: the y sweep applies the degenerate z sweep and writes qq(0).
  const bool fuse_z = b->mode_z == 2;
  const int av_in[3] = {1, 2, 1}, av_out[3] = {2, fuse_z ? 0 : 1, 0};
  const int mode[3] = {0, fuse_z ? 3 : b->mode_y, b->mode_z};
  const int ph[3] = {P_AVX, P_AVY, P_AVZ};
  for (int d = 0; d < (fuse_z ? 2 : 3); ++d) {
    e = launch_batch_av<R>(st, d, mode[d], b->d_inst, b->d_dyn, S(ph[d]), n, b->total[ph[d]],
                           av_in[d], av_out[d], av);
    if (e != hipSuccess) return e;
    e = launch_batch_ghost<R>(st, b->d_inst, S(P_GHOST), n, b->total[P_GHOST], av_out[d], b->bc);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t step_any(hipStream_t st, misohip_batch *b) {
  if (b->dtype == MISOHIP_F64) return step_t<double>(st, static_cast<BatchT<double> *>(b->impl));
  return step_t<float>(st, static_cast<BatchT<float> *>(b->impl));
}

}  // namespace

extern "C" {

int misohip_batch_create(misohip_ctx *ctx, misohip_dtype dt, int n, const misohip_grid *grids,
                         const misohip_fields *qq, const misohip_work *work,
                         const misohip_params *p, const misohip_bc *bc, misohip_batch **out) {
  if (!ctx || !grids || !qq || !work || !p || !bc || !out || n < 1) return MISOHIP_E_ARG;
  if (dt != MISOHIP_F64 && dt != MISOHIP_F32) return MISOHIP_E_DTYPE;
  // sources: none or gravity (a user source callback runs per run, not per batch)
  if (p->src_kind < 0 || p->src_kind > 1 || (p->src_kind == 1 && (p->grav_axis < 0 || p->grav_axis > 2)))
    return MISOHIP_E_ARG;
  *out = nullptr;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  misohip_batch *b = new (std::nothrow) misohip_batch;
  if (!b) return MISOHIP_E_NOMEM;
  b->dtype = dt;
  int rc;
  if (dt == MISOHIP_F64) {
    BatchT<double> *t = nullptr;
    rc = create_t<double>(ctx, n, grids, qq, work, p, bc, &t);
    b->impl = t;
  } else {
    BatchT<float> *t = nullptr;
    rc = create_t<float>(ctx, n, grids, qq, work, p, bc, &t);
    b->impl = t;
  }
  if (rc) {
    delete b;
    return rc;
  }
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
  if (e != hipSuccess) {
    misohip_batch_destroy(b);
    return hip_fail(e);
  }
  *out = b;
  return MISOHIP_OK;
}

int misohip_batch_destroy(misohip_batch *b) {
  if (!b) return MISOHIP_E_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->dtype == MISOHIP_F64) destroy_t(static_cast<BatchT<double> *>(b->impl));
  else destroy_t(static_cast<BatchT<float> *>(b->impl));
  delete b;
  return MISOHIP_OK;
}

// Capture steps_per_graph batched steps into one hipGraph (kernel nodes
// only: the CFL slots are reset by a kernel, never a memset node).
int misohip_batch_capture(misohip_ctx *ctx, misohip_batch *b, int steps_per_graph) {
  if (!ctx || !b || steps_per_graph < 1 || steps_per_graph > 1000) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  if (b->exec) {
    MISOHIP_TRY(hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
  }
  if (b->graph) {
    MISOHIP_TRY(hipGraphDestroy(b->graph));
    b->graph = nullptr;
  }
  MISOHIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int s = 0; s < steps_per_graph && e == hipSuccess; ++s) e = step_any(b->stream, b);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(e);
  }
  MISOHIP_TRY(e2);
  b->graph = g;
  MISOHIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->steps_per_graph = steps_per_graph;
  return MISOHIP_OK;
}

// Advance every instance by nsteps: graph replays when a graph is captured
// (nsteps must then be a multiple of its step count), direct launches on
// the context's stream otherwise.  Ordered after earlier work on the
// context's stream and before later work on it; no host synchronisation.
int misohip_batch_step(misohip_ctx *ctx, misohip_batch *b, int nsteps) {
  if (!ctx || !b || nsteps < 0) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  if (b->exec) {
    if (nsteps % b->steps_per_graph) return MISOHIP_E_ARG;
    MISOHIP_TRY(hipEventRecord(b->ev_in, ctx->stream));
    MISOHIP_TRY(hipStreamWaitEvent(b->stream, b->ev_in, 0));
    for (int r = 0; r < nsteps / b->steps_per_graph; ++r)
      MISOHIP_TRY(hipGraphLaunch(b->exec, b->stream));
    MISOHIP_TRY(hipEventRecord(b->ev_out, b->stream));
    MISOHIP_TRY(hipStreamWaitEvent(ctx->stream, b->ev_out, 0));
    return MISOHIP_OK;
  }
  for (int s = 0; s < nsteps; ++s) MISOHIP_TRY(step_any(ctx->stream, b));
  return MISOHIP_OK;
}

// Per-run walk length of the y/z sweeps (>= 4), e.g. chosen by a tile
// autotuner for each run size; drops a captured graph.
int misohip_batch_set_chunks(misohip_ctx *ctx, misohip_batch *b, const int *chunks) {
  if (!ctx || !b || !chunks) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
  MISOHIP_TRY(hipStreamSynchronize(b->stream));
  auto apply = [&](auto *t) -> int {
    for (int m = 0; m < t->n; ++m)
      if (chunks[m] < 4 || chunks[m] > 4096) return MISOHIP_E_ARG;
    for (int m = 0; m < t->n; ++m) t->h_inst[m].chunk = chunks[m];
    return upload_layout(t);
  };
  const int rc = b->dtype == MISOHIP_F64 ? apply(static_cast<BatchT<double> *>(b->impl))
                                         : apply(static_cast<BatchT<float> *>(b->impl));
  if (rc) return rc;
  if (b->exec) {
    MISOHIP_TRY(hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
  }
  if (b->graph) {
    MISOHIP_TRY(hipGraphDestroy(b->graph));
    b->graph = nullptr;
  }
  b->steps_per_graph = 0;
  return MISOHIP_OK;
}

// Last dt, simulated time and step count of every instance (synchronises);
// any array may be NULL.  *exp_fallback (if given) is nonzero when some
// damp factor had to be evaluated on the device instead of the table.
int misohip_batch_state(misohip_ctx *ctx, misohip_batch *b, double *dt, double *time,
                        long *steps, int *exp_fallback) {
  if (!ctx || !b) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
  MISOHIP_TRY(hipStreamSynchronize(b->stream));
  int fb = 0;
  auto read = [&](auto *t) -> int {
    using D = typename std::remove_pointer<decltype(t->d_dyn)>::type;
    std::vector<D> h(t->n);
    MISOHIP_TRY(hipMemcpy(h.data(), t->d_dyn, sizeof(D) * t->n, hipMemcpyDeviceToHost));
    for (int m = 0; m < t->n; ++m) {
      if (dt) dt[m] = (double)h[m].dt;
      if (time) time[m] = h[m].time;
      if (steps) steps[m] = h[m].steps;
      fb |= h[m].exp_fallback;
    }
    return MISOHIP_OK;
  };
  int rc = b->dtype == MISOHIP_F64 ? read(static_cast<BatchT<double> *>(b->impl))
                                   : read(static_cast<BatchT<float> *>(b->impl));
  if (rc) return rc;
  if (exp_fallback) *exp_fallback = fb;
  return MISOHIP_OK;
}

}  // extern "C"
