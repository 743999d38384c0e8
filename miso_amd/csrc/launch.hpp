// Kernel launchers shared between the C ABI (capi.cpp) and the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace misohip {

// Scalars of one RK stage (Integrator::update_sc4 arguments and members).
struct RkScalars {
  double dt, ch2, damp, gm, grav;
  int src_kind, grav_axis;
  int pencil_bx, pencil_by;  // pencil shape in blocks of 64 x 4 cells
};

// Scalars of one AV sweep (ArtificialViscosity members plus dt).
struct AvScalars {
  double dt, ep, fh, gm, cs_fac, ca_fac, vv_fac;
  int chunk;  // cells walked per thread by the y/z sweeps
};

// Boundary description on the device: periodic flag and mirror signs.
struct BcDev {
  int periodic[3];
  float sign[3][9];
};

// one ray of the sweep, everything the plane kernel needs
struct RtRay {
  int start;              // linear index (within the ray's block) of the first plane's (lo_b, lo_c) node
  int da, db, dc;         // linear step of the node along a (signed), b, c
  int ua, ub, uc;         // upwind offsets (linear) along a, b, c (0 if degenerate)
  int nb, nc;             // transverse node counts
  int nplanes;
  double w00, w10, w01, w11;  // bilinear weights
  double ds;
  double wq;              // quadrature weight
};

struct RtRays {
  RtRay r[24];
};

hipError_t launch_rt_sweep_plane(hipStream_t, double *, const double *, const double *,
                                 const RtRays &, int, int, long, int);
hipError_t launch_rt_halo(hipStream_t, double *, int, int, int, int, int, int, unsigned, long,
                          int);
hipError_t launch_rt_face(hipStream_t, double *, const double *, int, int, int, int, int, int,
                          int, int, int, double, unsigned, long, int);
hipError_t launch_rt_maxdiff(hipStream_t, const double *, const double *, const int[3],
                             const int[3], int, int, long, int, unsigned *);
hipError_t launch_rt_qrad(hipStream_t, const double *, const double *, const double *, double *,
                          const int[3], const int[3], int, int, long, int, const RtRays &);

hipError_t launch_rt_source(hipStream_t, const double *, const double *, double *, double *, long,
                            double, double, double);
hipError_t launch_rt_heat(hipStream_t, const double *, double *, const double *, const Shape &,
                          double);

template <typename R>
hipError_t launch_rk_stage(hipStream_t, const Fld<const R> &, const Fld<const R> &,
                           const Fld<R> &, const R *, const R *, const R *, const Shape &,
                           const RkScalars &);
template <typename R>
hipError_t launch_av_sweep(hipStream_t, int, const Fld<const R> &, const Fld<R> &, R *, const R *,
                           const R *, const Shape &, const AvScalars &);
template <typename R>
hipError_t launch_char_vel(hipStream_t, const Fld<const R> &, R *, int, const AvScalars &);
template <typename R>
hipError_t launch_ghost_fill(hipStream_t, const Fld<R> &, const Shape &, const BcDev &);
template <typename R>
hipError_t launch_cfl(hipStream_t, const Fld<const R> &, const R *, const R *, const R *,
                      const Shape &, double, double, unsigned *);
template <typename R>
hipError_t launch_transpose(hipStream_t, const R *, R *, int, int, int, bool);

}  // namespace misohip
