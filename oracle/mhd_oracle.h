/*
 * CPU oracle for the MISO MHD time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of MISO's `backend::Host` path (the reference
 * CPU implementation) used as the correctness checker of the HIP backend.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in the product (miso_amd/, include/) links or calls it.
 *
 * Layout: every 3D array uses MISO's own order, (i, j, k) with k fastest:
 * index = (i * j_total + j) * k_total + k  (miso/include/miso/array3d.hpp:54-60).
 * A field set is 9 consecutive arrays in the order ro, vx, vy, vz, bx, by,
 * bz, ei, ph (miso/include/miso/mhd_fields.hpp:13-34).
 *
 * Every function exists twice: `_f64` (Real = double) and `_f32`
 * (Real = float), exactly like the reference's templates instantiated with
 * the two Real types.
 */
#ifndef MISO_ORACLE_H
#define MISO_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Integer shape of a MISO grid (miso/include/miso/grid.hpp:34-100,213-242). */
typedef struct og_grid {
  int i_size, j_size, k_size;
  int i_total, j_total, k_total;
  int i_margin, j_margin, k_margin;
  int is, js, ks;
} og_grid;

/* Run parameters: the config keys the step consumes plus the boundary
 * description (miso/include/miso/config_defaults.hpp, eos.hpp,
 * boundary_condition.hpp).  `periodic[a]` = 1 selects the halo
 * self-exchange on axis a; 0 the mirror boundary of
 * boundary_condition.hpp:108-136 on both sides with per-field sign
 * `sym_sign[a][f]` (+1 / -1); 2 sets the axis' ghost cells from `fixed`
 * (a user boundary functor writing given ghost values; it runs first, then
 * the mirrors, then the halo exchange).  `src_kind` 1 = constant gravity along
 * `grav_axis` (the force -g*ro of demo/hd2d_rayleigh_taylor/src/main.cpp:109-112);
 * 2 = the point-mass gravity of demo/mhd3d_magnetosphere/src/src.hpp:18-34
 * (g = `grav`) with a heating src.ei = src_heat * ro.
 */
typedef struct og_params {
  double gm, cfl_number, ep, fh, cs_fac, ca_fac, vv_fac;
  int periodic[3];
  int sym_sign[3][9];
  int src_kind;
  int grav_axis;
  double grav;
  double src_heat;
  const void *fixed; /* periodic[a] == 2: 9 fields (the q layout) whose ghost
                        cells the fill copies in */
} og_params;

/* Metric arrays are packed as x[It] y[Jt] z[Kt] dx[It] dy[Jt] dz[Kt]
 * dxi[It] dyi[Jt] dzi[Kt]; og_grid_build_* fills them and returns min_dxyz. */
#define OG_DECL(R, S)                                                          \
  int og_grid_build_##S(int i_size, int j_size, int k_size, int margin,        \
                        double x_min, double x_max, double y_min,              \
                        double y_max, double z_min, double z_max,              \
                        og_grid *g, R *metric, R *min_dxyz);                   \
  void og_ghost_fill_##S(const og_grid *g, const og_params *p, R *q);          \
  void og_rk_stage_##S(const og_grid *g, const R *metric, const og_params *p,  \
                       R dt, R ch2, R tau, const R *orgn, const R *argm,       \
                       R *rslt, R *work4);                                     \
  void og_char_vel_##S(const og_grid *g, const og_params *p, const R *q,       \
                       R *cc, R *cs);                                          \
  void og_av_sweep_##S(const og_grid *g, const R *metric, const og_params *p,  \
                       int dir, R dt, const R *q, R *rslt, const R *cc);       \
  R og_cfl_##S(const og_grid *g, const R *metric, const og_params *p,          \
               const R *q, R *cs);                                             \
  void og_update_##S(const og_grid *g, const R *metric, R min_dxyz,            \
                     const og_params *p, R dt, R *q, R *work);                 \
  void og_run_##S(const og_grid *g, const R *metric, R min_dxyz,               \
                  const og_params *p, int nsteps, R *q, R *work, double *dts); \
  R og_dqq_eval_##S(R qq_dw, R qq_cn, R qq_up, R ep);                          \
  R og_flux_core_##S(R qq_dw, R qq_up, R dqq_dw, R dqq_up, R cc, R fh);

OG_DECL(double, f64)
OG_DECL(float, f32)

/* Scratch needed by og_update_* / og_run_*: 9N (argm) + 9N (rslt) + 6N. */
size_t og_work_elems(const og_grid *g);

/* Radiative transfer (rt_oracle.c), Real = double. */
int og_rt_quadrature(int num_rays, double *w, double *mx, double *my, double *mz);
int og_rt_solve(const og_grid *g, const double *dxyz, int num_rays, const double *kap,
                const double *S, double *I, const int *periodic, const int *face_kind,
                const double *face_value, double tolerance, int max_iters, double *max_diff);
void og_rt_qrad(const og_grid *g, int num_rays, const double *kap, const double *S,
                const double *I, double *qrad);

/* OpenMP threads used by the oracle (0 = leave the runtime default). */
void og_set_threads(int n);
int og_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
