/*
 * misohip.h -- C ABI of the MI355X (gfx950) backend for MISO's MHD time step.
 *
 * Every entry point is plain C: POD descriptors, raw device pointers, sizes
 * and scalars; no C++ or torch types cross this boundary.  Each call is
 * stream-ordered on the context's HIP stream and returns 0 on success or a
 * negative MISOHIP_E* code (see misohip_strerror).
 *
 * Device layout: a 3D field of the grid (i_total, j_total, k_total) is stored
 * x-fastest, element (i, j, k) at ((k * j_total + j) * i_total + i).  MISO's
 * host arrays are z-fastest (miso/include/miso/array3d.hpp:54-60);
 * misohip_copy_to_device / misohip_copy_to_host convert between the two.
 * Field order inside misohip_fields follows miso/include/miso/mhd_fields.hpp:
 * ro, vx, vy, vz, bx, by, bz, ei, ph.
 *
 * The entry points replace these reference interfaces:
 *   misohip_create / destroy      ExecContext<Real, backend::CUDA>
 *                                 (miso/include/miso/mhd_core.hpp:15-33)
 *   misohip_copy_to_device/host   Fields::copy_from H<->D
 *                                 (miso/include/miso/mhd_fields.hpp:83-107,148-172)
 *   misohip_alloc / free / memcpy Array3D<T, backend::CUDA> storage and copies
 *                                 (miso/include/miso/array3d.hpp:156-167,206-211,254-263)
 *   misohip_ghost_fill            Integrator::apply_boundary_condition =
 *                                 bc.apply + HaloExchanger::apply
 *                                 (miso/include/miso/mhd_integrator_host.hpp:250-255,
 *                                  miso/include/miso/mhd_halo_exchange.hpp:51-266,
 *                                  miso/include/miso/boundary_condition.hpp:108-182)
 *   misohip_rk_stage              Integrator::update_sc4
 *                                 (miso/include/miso/mhd_integrator_host.hpp:57-247,
 *                                  miso/include/miso/mhd_integrator_cuda.cuh:63-289,340-392)
 *   misohip_char_vel              ArtificialViscosity::characteristic_velocity_eval
 *                                 (miso/include/miso/mhd_artificial_viscosity_host.hpp:50-68)
 *   misohip_av_sweep              ArtificialViscosity::update(direction)
 *                                 (miso/include/miso/mhd_artificial_viscosity_host.hpp:70-348)
 *   misohip_cfl                   Integrator::cfl
 *                                 (miso/include/miso/mhd_integrator_host.hpp:303-334,
 *                                  miso/include/miso/mhd_integrator_cuda.cuh:449-478)
 *   misohip_update                Integrator::update
 *                                 (miso/include/miso/mhd_integrator_host.hpp:337-350)
 *   misohip_step                  ModelBase::update = cfl + update
 *                                 (miso/include/miso/mhd_model_base.hpp:57-62)
 *   misohip_run                   ModelBase::run's stepping loop
 *                                 (miso/include/miso/mhd_model_base.hpp:89-117)
 */
#ifndef MISOHIP_H
#define MISOHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define MISOHIP_OK 0
#define MISOHIP_E_HIP (-1)      /* a HIP runtime call failed            */
#define MISOHIP_E_ARG (-2)      /* invalid argument / shape             */
#define MISOHIP_E_DTYPE (-3)    /* unknown dtype                        */
#define MISOHIP_E_NOMEM (-4)    /* device allocation failed             */
#define MISOHIP_E_RANGE (-5)    /* size beyond a documented limit       */
#define MISOHIP_E_CALLBACK (-6) /* the user boundary callback failed    */

typedef enum misohip_dtype { MISOHIP_F64 = 0, MISOHIP_F32 = 1 } misohip_dtype;

typedef struct misohip_ctx misohip_ctx;

/* Shape and metric of one rank's grid (miso/include/miso/grid.hpp:103-141).
 * is/js/ks are 1 on an axis with more than one cell and 0 otherwise; the
 * metric arrays are device arrays of the field dtype, ghosts included. */
typedef struct misohip_grid {
  int i_size, j_size, k_size;
  int i_total, j_total, k_total;
  int i_margin, j_margin, k_margin;
  int is, js, ks;
  double min_dxyz; /* already float-narrowed, grid.hpp:241 */
  const void *dx, *dy, *dz;
  const void *dxi, *dyi, *dzi;
} misohip_grid;

/* The nine primitive fields, device pointers, x-fastest. */
typedef struct misohip_fields {
  void *q[9];
} misohip_fields;

/* Boundary description.  periodic[a] = 1: the axis is filled by the
 * one-rank halo self-exchange; 0: the mirror boundary of
 * boundary_condition.hpp:108-136 on both faces with sign sym_sign[axis][f];
 * 2: fixed, every fill sets the ghost cells of the axis from the `fixed`
 * fields (nine device fields of the grid's shape and layout; only their
 * ghost cells are read) -- the user boundary functor that writes given
 * values into the ghost zone (mhd_model_base.hpp:131-155 contract; the
 * reference applies the user functor before the halo exchange, so edges
 * shared with a periodic axis take the periodic image); 3: user, the axis'
 * ghost cells are left to the context's boundary callback
 * (misohip_set_bc_callback), which every fill calls first, and the fill maps
 * only the other axes. */
typedef struct misohip_bc {
  int periodic[3];
  int sym_sign[3][9];
  const void *fixed[9];
} misohip_bc;

/* A user boundary condition (BC::apply, run before the halo exchange,
 * mhd_integrator_host.hpp:250-255; its device form is a functor launched
 * with for_each / bc::symmetric, demo/mhd3d_magnetosphere/src/bc.hpp:39-158).
 * Every ghost fill of the context -- misohip_ghost_fill and the fills inside
 * misohip_update, misohip_step and misohip_run -- calls fn(user, fields,
 * stream) first; fn enqueues its device work on `stream` (the context's)
 * and returns 0, or nonzero to stop (the call then returns
 * MISOHIP_E_CALLBACK).  fn = NULL removes it.  A misohip_bc with an axis of
 * kind 3 needs one.  With a callback set, every sweep runs as its own pass
 * followed by a fill, and the CFL is a pass of its own over the state the
 * last fill leaves, in the reference's order (mhd_integrator_host.hpp:
 * 280-300, mhd_model_base.hpp:57-62): no x-y walk, no z sweep folded into
 * the 2D y sweep, no CFL or radiative heating folded into the last sweep, so
 * a callback may write interior cells (floors, blending near an obstacle).
 * The reference has no such entry point: there BC is a template argument. */
typedef int (*misohip_bc_fn)(void *user, const struct misohip_fields *q, void *stream);
int misohip_set_bc_callback(misohip_ctx *ctx, misohip_bc_fn fn, void *user);
/* The context's current callback and its user pointer (NULL, NULL when none
 * is set), so that a caller installing its own for one call can put the
 * previous one back.  No counterpart in the reference (see above). */
int misohip_get_bc_callback(misohip_ctx *ctx, misohip_bc_fn *fn, void **user);

/* A user source term (Source::vx, vy, vz, ei; mhd_model_base.hpp:131-155,
 * e.g. demo/mhd3d_magnetosphere/src/src.hpp:4-41).  With
 * misohip_params.src_kind = 2 every RK stage -- misohip_rk_stage and the
 * stages inside misohip_update, misohip_step, misohip_run and
 * misohip_update_heated -- first calls fn(user, argm, arrays, stream): fn
 * enqueues on `stream` (the context's) the device work that writes
 * src.vx, src.vy, src.vz and src.ei of that stage's argm state into
 * arrays[0..3] (device fields of the grid's shape and layout) at every
 * interior cell, and returns 0, or nonzero to stop (the call then returns
 * MISOHIP_E_CALLBACK).  The stage kernel adds arrays[0..2] to the momentum
 * equations inside dt and arrays[3] to the internal energy outside it, where
 * the reference evaluates the functor (mhd_integrator_host.hpp:134, 151,
 * 168, 241).  The arrays belong to the caller and must stay valid while the
 * callback is set; fn = NULL removes it.  The reference has no such entry
 * point: there Source is a template argument inlined into the stage. */
typedef int (*misohip_src_fn)(void *user, const struct misohip_fields *argm,
                              void *const arrays[4], void *stream);
int misohip_set_source_callback(misohip_ctx *ctx, misohip_src_fn fn, void *user,
                                void *const arrays[4]);
/* The context's current source callback, its user pointer and arrays (all
 * NULL when none is set), so that a caller installing its own for one call
 * can put the previous one back.  No counterpart in the reference. */
int misohip_get_source_callback(misohip_ctx *ctx, misohip_src_fn *fn, void **user,
                                void *arrays[4]);

/* Config values the step consumes (config_defaults.hpp, eos.hpp) and the
 * source term: src_kind 0 = none (EmptySourceTerm), 1 = constant gravity
 * -grav*ro along grav_axis (0/1/2), 2 = the context's source callback
 * (misohip_set_source_callback). */
typedef struct misohip_params {
  double gm;
  double cfl_number;
  double ep, fh, cs_fac, ca_fac, vv_fac;
  int src_kind;
  int grav_axis;
  double grav;
} misohip_params;

/* Workspace owned by the caller: the Integrator's qq_argm / qq_rslt and the
 * frozen AV speed cc (mhd_integrator_host.hpp:13-52). */
typedef struct misohip_work {
  misohip_fields argm;
  misohip_fields rslt;
  void *cc;
} misohip_work;

/* Context: device ordinal and the HIP stream every call is ordered on
 * (NULL = the null stream).  Creates no stream at creation; the first
 * misohip_rt_solve with both multi-plane and plane rays creates one more
 * (default priority) for the multi-plane sweeps, forked from and joined back
 * to the context's stream with events inside each iteration, and
 * misohip_destroy frees it. */
int misohip_create(int device, void *stream, misohip_ctx **out);
int misohip_destroy(misohip_ctx *ctx);
int misohip_set_stream(misohip_ctx *ctx, void *stream);
int misohip_synchronize(misohip_ctx *ctx);
const char *misohip_strerror(int code);

/* Raw device memory, for callers with no HIP binding of their own.
 * misohip_alloc: `bytes` of uninitialised device memory on the context's
 * device (NULL for 0 bytes).  misohip_free: waits for the context's stream,
 * then releases memory from misohip_alloc (NULL is a no-op).
 * misohip_memcpy: `bytes` copied as-is (no layout change), ordered on the
 * context's stream; a host buffer must stay valid until the stream has
 * reached the copy (misohip_synchronize). */
enum { MISOHIP_H2D = 1, MISOHIP_D2H = 2, MISOHIP_D2D = 3 };
int misohip_alloc(misohip_ctx *ctx, size_t bytes, void **out);
int misohip_free(misohip_ctx *ctx, void *p);
int misohip_memcpy(misohip_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);

/* MISO-order (z-fastest) host array <-> x-fastest device field.  `host`
 * holds i_total*j_total*k_total elements; synchronous. */
int misohip_copy_to_device(misohip_ctx *ctx, const misohip_grid *g,
                           misohip_dtype dt, const void *host, void *dev);
int misohip_copy_to_host(misohip_ctx *ctx, const misohip_grid *g,
                         misohip_dtype dt, const void *dev, void *host);

/* Ghost fill of all nine fields. */
int misohip_ghost_fill(misohip_ctx *ctx, const misohip_grid *g,
                       misohip_dtype dt, const misohip_fields *qq,
                       const misohip_bc *bc);

/* One RK stage: rslt (interior) from orgn (pointwise) and argm (radius-2
 * stencils).  dt_s is the stage step (dt/4, dt/3, dt/2, dt), ch2 = ch_divb^2
 * and damp = exp(-dt_s / tau_divb), both computed by the host exactly as
 * mhd_integrator_host.hpp:272,337-341 do. */
int misohip_rk_stage(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                     const misohip_params *p, const misohip_fields *orgn,
                     const misohip_fields *argm, const misohip_fields *rslt,
                     double dt_s, double ch2, double damp);

/* Characteristic AV speed cc over every cell (ghosts included). */
int misohip_char_vel(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                     const misohip_params *p, const misohip_fields *qq,
                     void *cc);

/* One directional AV sweep (dir 0/1/2): rslt (interior) from qq and the
 * frozen cc. */
int misohip_av_sweep(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                     const misohip_params *p, int dir, double dt_step,
                     const misohip_fields *qq, const misohip_fields *rslt,
                     const void *cc);

/* CFL time step over the interior; synchronises and writes *dt_out. */
int misohip_cfl(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                const misohip_params *p, const misohip_fields *qq,
                double *dt_out);

/* Integrator::update(dt): four RK stages and three AV sweeps, each followed
 * by the ghost fill; qq is advanced in place. */
int misohip_update(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                   const misohip_params *p, const misohip_bc *bc,
                   const misohip_fields *qq, const misohip_work *w,
                   double dt_step);

/* Integrator::update, then the operator-split radiative heating
 * ei += dt * qrad / ro on the interior (misohip_rt_heat) and the ghost fill,
 * and, when next_dt is not NULL, Integrator::cfl of the heated state into
 * *next_dt (host-synchronising).  With x and y active the last AV sweep
 * heats each cell as it stores it and reduces the CFL candidates from the
 * heated values, so the heating and CFL passes disappear; the results are
 * the same bits as misohip_update, misohip_rt_heat, misohip_ghost_fill and
 * misohip_cfl in sequence.  FP64 only.  No counterpart in the reference,
 * which has no radiative heating. */
int misohip_update_heated(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                          const misohip_params *p, const misohip_bc *bc,
                          const misohip_fields *qq, const misohip_work *w, double dt_step,
                          const void *qrad, double *next_dt);

/* nsteps of ModelBase::update on the same fields (the loop of
 * ModelBase::run, mhd_model_base.hpp:89-117): from the second step on, the
 * CFL reduction is carried out by the previous step's last AV sweep on the
 * values it stores (3D runs and 2D x-y runs; others reduce separately), so
 * dts[] (optional, nsteps values) equal those of nsteps misohip_step calls.
 * Where the step reads periodic ghosts at their sources (3D, x and y AV
 * sweeps as one walk, a z-march RK variant, no user boundary), the periodic
 * faces are filled after the last step only; the fields returned are the
 * same bits as those of nsteps misohip_step calls.
 * The caller must not modify qq between the steps -- use misohip_step for
 * that (e.g. to add a source between steps). */
int misohip_run(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                const misohip_params *p, const misohip_bc *bc,
                const misohip_fields *qq, const misohip_work *w, int nsteps,
                double *dts);

/* ModelBase::update: dt = cfl(qq), then update(dt).  *dt_out receives dt. */
int misohip_step(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                 const misohip_params *p, const misohip_bc *bc,
                 const misohip_fields *qq, const misohip_work *w,
                 double *dt_out);

/* Plain streaming kernels that calibrate the HBM rate the MHD kernels are
 * compared against: kind 0 copies `bytes` from src to dst, 1 reads src, 2
 * writes dst; 16-byte accesses, four in flight per lane, nontemporal stores,
 * 16384 blocks of 256 threads; bytes a multiple of 16.  Stream-ordered.  No
 * counterpart in the reference: measurement only. */
int misohip_stream_bench(misohip_ctx *ctx, int kind, const void *src, void *dst, size_t bytes);

/* Bytes of the per-device workspace of misohip_update_noalloc. */
#define MISOHIP_NOALLOC_SCRATCH_BYTES ((size_t)48 << 20)

/* misohip_update without the workspace argument and without any allocation:
 * argm, rslt and cc (19 arrays, each rounded up to 256 bytes) live in a
 * device array of MISOHIP_NOALLOC_SCRATCH_BYTES that the library holds on
 * every GPU it runs on until the process ends, so one array (all cells,
 * ghosts included) may take at most MISOHIP_NOALLOC_SCRATCH_BYTES / 19 bytes
 * after rounding: 331,104 cells in FP64 (a 64^3 interior, 68^3 with
 * ghosts), 662,208 in FP32 (an 83^3 interior); larger grids return
 * MISOHIP_E_RANGE.  No hipMalloc, hipFree or
 * event creation happens inside (launch timing is off for the call).  Calls
 * on one GPU share its array: calls that overlap on one GPU (two streams,
 * two threads) race.  Results equal misohip_update's bit for bit.  No
 * counterpart in the reference: it is for callers that cannot allocate --
 * code under HIP stream capture, tight serving loops, harnesses that forbid
 * allocation. */
int misohip_update_noalloc(misohip_ctx *ctx, const misohip_grid *g, misohip_dtype dt,
                           const misohip_params *p, const misohip_bc *bc,
                           const misohip_fields *qq, double dt_step);

/* ---------------------------------------------------------------------
 * Radiative transfer (FP64).  Replaces rt::RT::solve and its helpers
 * (miso/include/miso/rt_host.hpp:104-689, CPU-only in the reference) and
 * the quadrature (miso/include/miso/rt_quadrature.hpp:14-63).  kappa and S
 * are fields of the grid (x-fastest, ghosts included); rint and rint_old
 * hold num_rays such fields back to back.  face_kind per (axis, side) of a
 * non-periodic axis: 0 untouched, 1 incoming I = S (thermal), 2 incoming
 * I = 0, 3 incoming I = face_value (set_incoming_boundary_on_physical_faces,
 * rt_host.hpp:50-101).  num_rays 24 = Carlson A4, otherwise one +z ray.
 * x_rays_yfast 1: a ray whose dominant axis is x (with more than one cell
 * on x and on y) keeps its field of rint y-fastest, index
 * ((k * i_total + i) * j_total + j), so the planes it sweeps are contiguous;
 * misohip_rt_ray_layout names those rays.  0: every ray x-fastest.
 * sweep_planes 2..16: a ray whose shorter transverse axis (c) has at most
 * 32 lattice nodes (the x and y rays of a thin slab) advances that many
 * planes per launch; 0 or 1: one launch per plane for every ray.  Results
 * are the same bits under every setting of both. */
typedef struct misohip_rt_config {
  int num_rays;
  int periodic[3];
  int face_kind[6];
  double face_value[6];
  double tolerance;
  int max_iters;
  int x_rays_yfast;
  int sweep_planes;
} misohip_rt_config;

/* The angular quadrature the solver uses (AngularQuadrature,
 * miso/include/miso/rt_quadrature.hpp:14-63): for num_rays 24 the Carlson A4
 * set, otherwise one +z ray.  Writes the weights and direction cosines of
 * every ray (arrays of at least 24) and returns the ray count; no device
 * work. */
int misohip_rt_quadrature(int num_rays, double *w, double *mu_x, double *mu_y, double *mu_z);

/* RT::solve: iterate halo exchange, boundary, sweep until the float-
 * narrowed max change is below tolerance (rt_host.hpp:632-651).  Writes the
 * iteration count and the last max change; synchronises once per
 * iteration.  rint_old is the solver's scratch: the change is taken inside
 * the sweep, and rint_old keeps only each ray's upwind lattice faces as
 * they were at the start of the last iteration. */
int misohip_rt_solve(misohip_ctx *ctx, const misohip_grid *g, const misohip_rt_config *rt,
                     const void *kappa, const void *S, void *rint, void *rint_old,
                     int *iters, double *max_diff);

/* Bit n of *yfast_mask set: ray n's field of rint is y-fastest under this
 * grid and configuration (see x_rays_yfast); no device work besides reading
 * the grid spacings.  No counterpart in the reference (its rint is one
 * layout). */
int misohip_rt_ray_layout(misohip_ctx *ctx, const misohip_grid *g, const misohip_rt_config *rt,
                          unsigned *yfast_mask);

/* Q_rad = 4 pi kappa (J - S), J = sum_r w_r I_r, on the RT lattice (no
 * counterpart in the reference, which has no MHD coupling). */
int misohip_rt_qrad(misohip_ctx *ctx, const misohip_grid *g, const misohip_rt_config *rt,
                    const void *kappa, const void *S, const void *rint, void *qrad);

/* Grey opacity and source function of the radiative-MHD workload from the
 * fields, on every cell: T = (gm - 1) ei, kappa = kappa0 ro T^3,
 * S = s_rad T^4 (s_rad: the Stefan-Boltzmann constant over pi in code
 * units, which sets the radiative relaxation time against the gas energy). */
int misohip_rt_source(misohip_ctx *ctx, const misohip_grid *g, const misohip_fields *qq,
                      double gm, double kappa0, double s_rad, void *kappa, void *S);

/* Radiative heating as an operator-split update after Integrator::update,
 * interior cells: ei += dt * qrad / ro.  (MISO's Source::ei is added outside
 * dt, mhd_integrator_host.hpp:289-308, so a rate cannot go through it.) */
int misohip_rt_heat(misohip_ctx *ctx, const misohip_grid *g, const misohip_fields *qq,
                    const void *qrad, double dt);

/* ---------------------------------------------------------------------
 * Per-column MLP surrogate of the radiative heating (no counterpart in the
 * reference).  Network n_in -> n_hidden (x n_hidden_layers, SiLU) -> n_out,
 * weights stored [out][in] as FP32 masters with BF16 copies; GEMMs run on
 * the bf16 MFMA units with FP32 accumulation.  X is bf16 [batch][n_in],
 * Y and T fp32 [batch][n_out].  All of n_in, n_hidden, n_out and max_batch
 * must be multiples of 32.  The handle owns its device memory. */
typedef struct misohip_mlp misohip_mlp;
int misohip_mlp_create(misohip_ctx *ctx, int n_in, int n_hidden, int n_hidden_layers, int n_out,
                       int max_batch, misohip_mlp **out);
int misohip_mlp_destroy(misohip_mlp *m);
/* set layer weights (host, [out][in]) and biases; resets the Adam moments */
int misohip_mlp_set_layer(misohip_ctx *ctx, misohip_mlp *m, int layer, const float *w_host,
                          const float *b_host);
/* which: 0 weights, 1 biases, 2 weight gradients, 3 bias gradients */
int misohip_mlp_get_layer(misohip_ctx *ctx, misohip_mlp *m, int layer, int which, float *host);
/* feature and target normalisation (host arrays of n_in and n_out) */
int misohip_mlp_set_norm(misohip_ctx *ctx, misohip_mlp *m, const float *mu_in, const float *sd_in,
                         const float *mu_out, const float *sd_out);
int misohip_mlp_forward(misohip_ctx *ctx, misohip_mlp *m, const void *X, int batch, float *Y);
/* forward + MSE loss + backward into the gradient buffers */
int misohip_mlp_grad(misohip_ctx *ctx, misohip_mlp *m, const void *X, const float *T, int batch,
                     double *loss);
int misohip_mlp_adam(misohip_ctx *ctx, misohip_mlp *m, float lr, float beta1, float beta2,
                     float eps, float weight_decay);
/* column features of a radiative box (columns along z, i fastest):
 * normalised (ln ro, ln T, |B|^2) of the nz interior cells of each column
 * (n_in = 3 nz), and with n_in = 4 nz also the horizontal Laplacian of ln T
 * at a stride of 4 cells (x and y periodic); FP64 fields.  Xf (optional)
 * receives the fp32 values. */
int misohip_mlp_features(misohip_ctx *ctx, misohip_mlp *m, const misohip_grid *g,
                         const misohip_fields *qq, double gm, void *X, float *Xf);
/* Q_rad on every interior cell from the network, plus corr (optional FP64
 * field) where given; flagged columns (flags optional) take q_rt */
int misohip_mlp_qrad(misohip_ctx *ctx, misohip_mlp *m, const misohip_grid *g, const void *X,
                     float *Y, const int *flags, const void *q_rt, const void *corr,
                     void *qrad);
/* residual indicator of the last network output Y against a plain solve
 * q_rt: flag columns whose relative L2 error over the column exceeds tol;
 * corr (optional) receives q_rt - network on every interior cell, the
 * correction misohip_mlp_qrad can add until the next solve */
int misohip_mlp_residual(misohip_ctx *ctx, misohip_mlp *m, const misohip_grid *g, const float *Y,
                         const void *q_rt, int *flags, double tol, void *corr);

/* Launch shapes of the stencil kernels (no counterpart in the reference,
 * whose CUDA kernels use fixed 8x8x8 blocks, cuda_util.cuh:35-60).
 * rk_pencil_bx/by: pencil of the RK-stage block order, in blocks of 64x4
 * cells; sweep_chunk: cells each thread of the y/z AV sweeps walks.
 * Defaults {2, 8, 64}; the autotuner sets them per grid. */
typedef struct misohip_tiles {
  int rk_pencil_bx, rk_pencil_by;
  int sweep_chunk;
  int rk_variant; /* 3D RK kernel: 0 = three load phases, 1 = one per stencil
                     point, 2 = as 1 held to 4 waves per SIMD, 3 = x-y
                     neighbourhood staged in LDS, 4 = z-march (a block walks
                     a 64 x 4 column through rk_zchunk planes, z neighbours
                     in registers and LDS, x-y neighbourhood in LDS), 5 = the
                     z-march over 64 x 8 columns (512 threads), 6 = the
                     z-march over 64 x 12 columns (768 threads, three waves
                     per SIMD) for FP32; FP64 runs 6 as 5 */
  int rk_zchunk;  /* planes per block of the z-march kernels, 1..4096 */
  int av_fuse_xy; /* 3D runs: 1 = the x and y AV sweeps as one walk that
                     x-sweeps every row it takes in (the x sweep's ghost fill
                     drops out), FP64 code for three waves per SIMD; 2 = the
                     same walk compiled for two waves per SIMD (no registers
                     parked in scratch); 3, 4 = 1, 2 with the z walk
                     overlapped with the x-y walk where z is periodic or
                     fixed (the z walk's chunks in up to sixteen slabs on
                     one more stream of the context, the x-y walk's slabs
                     on another, each z slab once the x-y walk has
                     produced the planes it reads); 0 = separate x and y
                     sweeps.
                     Same results. */
  int av_xy_chunk; /* rows walked per wave by that walk, 4..4096 */
} misohip_tiles;
int misohip_set_tiles(misohip_ctx *ctx, const misohip_tiles *t);
int misohip_get_tiles(misohip_ctx *ctx, misohip_tiles *t);

/* Kernel timing with HIP events recorded on the context's stream around every
 * launch the library makes (no counterpart in the reference, which has no
 * timers).  Kinds: see MISOHIP_KIND_*.  misohip_profile_read synchronises,
 * writes the summed milliseconds and launch counts of the first `nkinds`
 * kinds and clears the record. */
#define MISOHIP_KIND_RK_STAGE 0
#define MISOHIP_KIND_AV_X 1
#define MISOHIP_KIND_AV_Y 2
#define MISOHIP_KIND_AV_Z 3
#define MISOHIP_KIND_GHOST 4
#define MISOHIP_KIND_CFL 5
#define MISOHIP_KIND_RT 6
#define MISOHIP_KIND_MLP 7
#define MISOHIP_KIND_COUNT 8
int misohip_profile_enable(misohip_ctx *ctx, int on);
int misohip_profile_read(misohip_ctx *ctx, double *ms, int *launches, int nkinds);

/* ---------------------------------------------------------------------
 * Batched ensemble (no counterpart in the reference, which runs one model
 * per process, miso/include/miso/mhd_model_base.hpp:89-117).  n independent
 * runs of one dtype and the same active axes (x and y active; z active in
 * all or none), margin 2, each with its own grid, fields, workspace and dt,
 * advanced together: one launch per phase covers every run, the step
 * scalars (dt, ch2, damp) stay on the device, and a captured hipGraph
 * replays whole steps.  Every run computes exactly what misohip_step
 * computes for it alone.  The handle records the pointers it is given; the
 * caller keeps those buffers alive until misohip_batch_destroy.  The handle
 * owns one stream (default priority) for capture and replay, ordered
 * against the context's stream with events. */
typedef struct misohip_batch misohip_batch;
int misohip_batch_create(misohip_ctx *ctx, misohip_dtype dt, int n, const misohip_grid *grids,
                         const misohip_fields *qq, const misohip_work *work,
                         const misohip_params *p, const misohip_bc *bc, misohip_batch **out);
int misohip_batch_destroy(misohip_batch *b);
/* Capture steps_per_graph steps (1..1000) into one graph; later
 * misohip_batch_step calls replay it. */
int misohip_batch_capture(misohip_ctx *ctx, misohip_batch *b, int steps_per_graph);
/* Advance every run by nsteps (a multiple of the captured step count when a
 * graph exists); asynchronous. */
int misohip_batch_step(misohip_ctx *ctx, misohip_batch *b, int nsteps);
/* Per-run walk length (cells per thread, 4..4096) of the y/z AV sweeps,
 * as a tile autotuner picks it per run size; drops a captured graph. */
int misohip_batch_set_chunks(misohip_ctx *ctx, misohip_batch *b, const int *chunks);
/* Last dt, simulated time and step count of each run (synchronises); any
 * output may be NULL.  *exp_fallback != 0 reports that a GLM damping factor
 * was evaluated on the device rather than taken from the host table. */
int misohip_batch_state(misohip_ctx *ctx, misohip_batch *b, double *dt, double *time,
                        long *steps, int *exp_fallback);

/* ---------------------------------------------------------------------
 * Uniform-table interpolation over device arrays: UniformTableInterpolator1D
 * and 2D (miso/include/miso/table_interpolator.hpp:22-166) with their
 * interpolate() over flat or 3D arrays (count elements).  The table has n
 * (n0 x n1, row-major, the second index contiguous) entries spanning
 * [x_min, x_max] uniformly; outside the range the nearest entry is taken.
 * Element type dt for table, x and y.  Asynchronous on the context's stream. */
int misohip_table1d(misohip_ctx *ctx, misohip_dtype dt, const void *table, int n, double x_min,
                    double x_max, const void *x, void *y, long count);
int misohip_table2d(misohip_ctx *ctx, misohip_dtype dt, const void *table, int n0, int n1,
                    double x0_min, double x0_max, double x1_min, double x1_max, const void *x0,
                    const void *x1, void *y, long count);

#ifdef __cplusplus
}
#endif
#endif /* MISOHIP_H */
