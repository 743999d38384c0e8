// Kernel launchers shared between the C ABI (capi.cpp) and the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace misohip {

// Scalars of one RK stage (Integrator::update_sc4 arguments and members).
struct RkScalars {
  double dt, ch2, damp, gm, grav;
  int src_kind, grav_axis;
};

// Scalars of one AV sweep (ArtificialViscosity members plus dt).
struct AvScalars {
  double dt, ep, fh, gm, cs_fac, ca_fac, vv_fac;
};

// Boundary description on the device: periodic flag and mirror signs.
struct BcDev {
  int periodic[3];
  float sign[3][9];
};

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
