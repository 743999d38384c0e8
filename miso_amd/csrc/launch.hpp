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
  int variant;  // 3D kernel: 0 three load phases, 1 one per point, 2 = 1 at 4 waves/SIMD, 3 LDS tile,
                // 4 z-march over 64 x 4 columns, 5 over 64 x 8
  int zchunk;   // planes walked per block by the z-march kernel
  // z-march kernel: axes (bit 0 x, 1 y, 2 z) whose ghost cells it does not
  // read but maps to their periodic source cells, so that no ghost fill
  // needs to run before it on those axes
  int wrap = 0;
  // src_kind 2: src.vx, src.vy, src.vz, src.ei of the stage's argm (fields
  // of the grid's shape, the run's Real)
  const void *src_arr[4];
};

// Scalars of one AV sweep (ArtificialViscosity members plus dt).
struct AvScalars {
  double dt, ep, fh, gm, cs_fac, ca_fac, vv_fac;
  int chunk;  // cells walked per thread by the y/z sweeps
  // x-y walk and z walk: axes (bit 0 x, 1 y, 2 z) whose ghost cells they
  // map to periodic source cells instead of reading them
  int wrap = 0;
  int xy_waves = 3;  // FP64 x-y walk: waves per SIMD it is compiled for (3 or 2)
  // a launch over part of the grid: the z walk's chunks [part0, part0 +
  // nparts), the x-y walk's planes part0, part0 + 1, ... (nparts of them; a
  // range past the top z ring plane continues at the bottom one); 0: all
  int part0 = 0, nparts = 0;
};

// Boundary description on the device: axis kind (0 mirror, 1 periodic,
// 2 fixed from the `fixed` fields), mirror signs, fixed-value fields.
struct BcDev {
  int periodic[3];
  float sign[3][9];
  const void *fixed[9];
};

// one ray of the sweep, everything the plane kernel needs
struct RtRay {
  int start;              // linear index (within the ray's block) of the first plane's (lo_b, lo_c) node
  int da, db, dc;         // linear step of the node along a (signed), b, c
  int ua, ub, uc;         // upwind offsets (linear) along a, b, c (0 if degenerate)
  int nb, nc;             // transverse node counts
  int sb, sc;             // upwind step along b, c in nodes (+1, -1, or 0 if degenerate)
  int nplanes;
  double w00, w10, w01, w11;  // bilinear weights
  double ds;
  double wq;              // quadrature weight
  int id;                 // the ray's slot in the intensity array
  int tr;                 // 1: intensities (and kappa, S) y-fastest, see rt.hip
};

struct RtRays {
  RtRay r[24];
};

// lattice (lo, n per axis), ray directions and layouts for rt_faces_kernel
struct RtFaces {
  int lo[3], n[3];
  unsigned pos[3];  // bit r: ray r goes + along the axis
  unsigned tr;      // bit r: ray r is kept y-fastest
  int It, Jt;
};
constexpr int RT_PARTS = 64;  // change-reduction slots, 32 words apart

hipError_t launch_rt_sweep_plane(hipStream_t, double *, const double *, const double *,
                                 const double *, const double *, const RtRays &, int, int, long,
                                 int, unsigned *);
hipError_t launch_rt_transpose_xy(hipStream_t, const double *, double *, int, int, int);
// several planes per launch for rays whose c axis has at most RT_MULTI_ROWS - 1 nodes
constexpr int RT_MULTI_COLS = 32;  // b nodes a workgroup owns
constexpr int RT_MULTI_PLANES = 16;
constexpr int RT_MULTI_ROWS = 33;
hipError_t launch_rt_sweep_multi(hipStream_t, double *, const double *, const double *,
                                 const double *, const double *, const RtRays &, int, int, int,
                                 long, int, unsigned *);
hipError_t launch_rt_halo(hipStream_t, double *, int, int, int, int, int, int, unsigned, unsigned,
                          long, int);
hipError_t launch_rt_face(hipStream_t, double *, const double *, int, int, int, int, int, int,
                          int, int, int, double, unsigned, long, int);
hipError_t launch_rt_faces(hipStream_t, const double *, double *, bool, const RtFaces &, long, int,
                           unsigned *);
hipError_t launch_rt_parts_reduce(hipStream_t, unsigned *, unsigned *, bool);
hipError_t launch_rt_qrad(hipStream_t, const double *, const double *, const double *, double *,
                          const int[3], const int[3], int, int, long, int, unsigned,
                          const RtRays &);
hipError_t launch_rt_source(hipStream_t, const double *, const double *, double *, double *, long,
                            double, double, double);
hipError_t launch_rt_heat(hipStream_t, const double *, double *, const double *, const Shape &,
                          double);

// The next step's CFL reduction carried out by the last AV sweep of a step:
// metric, the run's constants and the reduction slot.
template <typename R> struct CflFuse {
  const R *dx, *dy, *dz;
  double gm, cfl;
  unsigned *slot;
  // optional radiative heating applied to each stored cell first (FP64
  // only): ei = ei + heat_dt * heat_q / ro, the operator-split heating
  const double *heat_q = nullptr;
  double heat_dt = 0.0;
};
hipError_t launch_cfl_init(hipStream_t, unsigned *);
// plain streaming kernels (calib.hip): kind 0 copy, 1 read, 2 write
hipError_t launch_stream(hipStream_t, const void *, void *, size_t, int, double *);

// Waves of a walking-sweep block laid along x: all four when the other
// transverse axis has fewer than four cells (2D runs), else one (the four
// waves then take four rows of that axis).
__host__ __device__ inline int walk_waves_x(int n_other) { return n_other < 4 ? 4 : 1; }

// Chunks of a walk over n_dir cells: the requested length (at least 4),
// halved while the launch would have fewer than 2^18 threads (a small grid
// walked in long chunks leaves the chip idle), then evened out.  Every
// chunk recomputes its first slopes and flux from the fields, so the
// results do not depend on the split.
inline void walk_split(int n_dir, long threads_per_chunk, int chunk_req, int &chunk,
                       int &nchunks) {
  int c = chunk_req < 4 ? 4 : chunk_req;
  while (c > 4 && (long)((n_dir + c - 1) / c) * threads_per_chunk < (1L << 18)) c = (c + 1) / 2;
  if (c < 4) c = 4;
  nchunks = (n_dir + c - 1) / c;
  chunk = (n_dir + nchunks - 1) / nchunks;
}

// ---------------------------------------------------------------- batches
// One instance of a batched launch: its three field sets (0 = qq, 1 = argm,
// 2 = rslt), frozen AV speed, metric and shape.  Stored in device memory;
// a block finds its instance by a binary search over per-phase block
// offsets.
template <typename R> struct BatchInst {
  Fld<R> st[3];
  R *cc;
  const R *dx, *dy, *dz, *dxi, *dyi, *dzi;
  Shape sh;
  double min_dxyz;
  int chunk;  // cells walked per thread by the y/z sweeps of this instance
};

// Per-instance step scalars, produced on the device from the CFL result.
template <typename R> struct BatchDyn {
  unsigned cfl_bits;  // float bits of the CFL minimum
  int exp_fallback;   // set if damp had to be evaluated on the device
  R dt;
  R dts[4];
  R ch2;
  R damp[4];
  double time;        // simulated time of the instance
  long steps;
};

// exp(-dt_s/tau) exactly as the host library evaluates it, tabulated for
// the quotients the four stages can produce (stage fractions 1/4, 1/2 and 1
// give -1/8, -1/4, -1/2 exactly; 1/3 gives a value within a few ulps of
// -1/6).
template <typename R> struct DampTable {
  R exact[4];      // stages 0, 2, 3 (index by stage; [1] unused)
  R third[33];     // stage 1: quotient bits q0 - 16 .. q0 + 16
  R q0;            // the quotient nearest -1/6
};

__device__ __forceinline__ int batch_find(const int *__restrict__ start, int n, int b) {
  int lo = 0, hi = n - 1;  // start[0] = 0 <= b < start[n]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The batched RK stage of runs with x and y axes and no z axis walks y:
// each wave owns RK_YWALK_OUT x outputs over RK_YWALK_ROWS rows, four waves
// per block (batch_rk_ywalk_kernel; the batch's block counts use these).
constexpr int RK_YWALK_OUT = 60;
constexpr int RK_YWALK_ROWS = 16;
// a single x-y run walks y only when that makes at least this many waves
// (eight per CU); smaller runs keep the one-plane kernel
constexpr int RK_YWALK_MIN_WAVES = 2048;

template <typename R>
hipError_t launch_batch_rk(hipStream_t, const BatchInst<R> *, const BatchDyn<R> *, const int *,
                           int, int, int, int, int, const RkScalars &, int);
template <typename R>
hipError_t launch_batch_ghost(hipStream_t, const BatchInst<R> *, const int *, int, int, int,
                              const BcDev &);
template <typename R>
hipError_t launch_batch_cfl(hipStream_t, const BatchInst<R> *, BatchDyn<R> *, const int *, int,
                            int, double, double);
template <typename R>
hipError_t launch_batch_params(hipStream_t, const BatchInst<R> *, BatchDyn<R> *, int, double,
                               const DampTable<R> &);
template <typename R>
hipError_t launch_batch_av(hipStream_t, int, int, const BatchInst<R> *, const BatchDyn<R> *,
                           const int *, int, int, int, int, const AvScalars &);

template <typename R>
hipError_t launch_rk_stage(hipStream_t, const Fld<const R> &, const Fld<const R> &,
                           const Fld<R> &, const R *, const R *, const R *, const Shape &,
                           const RkScalars &);
template <typename R>
hipError_t launch_av_sweep(hipStream_t, int, const Fld<const R> &, const Fld<R> &, R *, const R *,
                           const R *, const Shape &, const AvScalars &);
template <typename R>
hipError_t launch_av_sweep_y_degen_z(hipStream_t, const Fld<const R> &, const Fld<R> &,
                                     const R *, const R *, const R *, const Shape &,
                                     const AvScalars &);
// the last AV sweep of a step (z walk in 3D, y walk with the degenerate z
// sweep in 2D) also reducing the next step's CFL candidates into cf.slot
template <typename R>
hipError_t launch_av_last_sweep_cfl(hipStream_t, const Fld<const R> &, const Fld<R> &,
                                    const R *, const R *, const R *, const Shape &,
                                    const AvScalars &, const CflFuse<R> &);
template <typename R>
hipError_t launch_av_sweep_xy(hipStream_t, const Fld<const R> &, const Fld<R> &, R *, const R *,
                              const R *, const Shape &, const BcDev &, const AvScalars &);
// the chunks a y (dir 1) or z (dir 2) walk of this shape is cut into
void av_walk_split(const Shape &sh, int dir, int chunk_req, int *chunk, int *nchunks);
template <typename R>
hipError_t launch_char_vel(hipStream_t, const Fld<const R> &, R *, int, const AvScalars &);
template <typename R>
hipError_t launch_table1d(hipStream_t, const R *, int, R, R, const R *, R *, long);
template <typename R>
hipError_t launch_table2d(hipStream_t, const R *, int, int, R, R, R, R, const R *, const R *, R *,
                          long);
template <typename R>
// axes: bit 0 x, 1 y, 2 z -- the faces to fill (edges of an axis left out
// keep their values)
hipError_t launch_ghost_fill(hipStream_t, const Fld<R> &, const Shape &, const BcDev &,
                             int axes = 7);
template <typename R>
hipError_t launch_cfl(hipStream_t, const Fld<const R> &, const R *, const R *, const R *,
                      const Shape &, double, double, unsigned *);
template <typename R>
hipError_t launch_transpose(hipStream_t, const R *, R *, int, int, int, bool);

}  // namespace misohip
