// miso_hip_device.hpp -- device-side helpers for user code of the MI355X
// backend: the backend::HIP forms of MISO's for_each (execution.hpp:56-100)
// and bc::symmetric (boundary_condition.hpp:54-182), with views of the
// device fields and grid, for boundary functors of kind USER
// (miso::hip::Boundary::user in miso_hip.hpp), and make_source, which turns a
// MISO-style Source struct into a UserSource.  A magnetosphere-style
// boundary (demo/mhd3d_magnetosphere/src/bc.hpp:39-158) is a functor that
// launches these on the stream it is given.  HIP-only: compile the user's
// translation unit with hipcc for gfx950.
//
//   miso::hip::Boundary bc(Boundary::PERIODIC, Boundary::PERIODIC, Boundary::USER);
//   bc.user([&](const misohip_fields &q, hipStream_t s) {
//     miso::hip::FieldsView<double> f(q, *grid.desc());
//     miso::hip::GridView<double> g(grid);
//     miso::hip::symmetric(miso::backend::HIP{}, f.ro, g, Sign::Pos, Direction::Z,
//                          Side::INNER, s);
//     ...
//   });
#pragma once

#include <hip/hip_runtime.h>

#include "miso_hip.hpp"

#define MISO_HIP_LAMBDA [=] __host__ __device__

namespace miso {
namespace hip {

/// Half-open index range (Range1D, execution.hpp:16-25).
struct Range1D {
  int begin, end;
  __host__ __device__ int size() const { return end > begin ? end - begin : 0; }
};

/// Product of three ranges (Range3D, execution.hpp:33-53); the first index
/// runs fastest over the threads here, matching the x-fastest fields.
struct Range3D {
  Range1D range0, range1, range2;
  __host__ __device__ long size() const {
    return (long)range0.size() * range1.size() * range2.size();
  }
};

template <class F> __global__ void for_each_kernel(Range1D r, F f) {
  const int n = r.size();
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    f(r.begin + t);
}

template <class F> __global__ void for_each_kernel(Range3D r, F f) {
  const long n = r.size();
  const int n0 = r.range0.size(), n1 = r.range1.size();
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const int i0 = r.range0.begin + (int)(t % n0);
    const long u = t / n0;
    const int i1 = r.range1.begin + (int)(u % n1);
    const int i2 = r.range2.begin + (int)(u / n1);
    f(i0, i1, i2);
  }
}

inline unsigned for_each_blocks(long n) {
  const long b = (n + 255) / 256;
  return (unsigned)(b < 65536 ? b : 65536);
}

/// Parallel for over a 1D range on `stream` (for_each(backend::CUDA, ...)).
template <class F>
void for_each(backend::HIP, Range1D range, F f, hipStream_t stream = nullptr) {
  if (range.size() == 0) return;
  hipLaunchKernelGGL(for_each_kernel<F>, dim3(for_each_blocks(range.size())), dim3(256), 0,
                     stream, range, f);
  MISO_HIP_RT_CHECK(hipGetLastError());
}

/// Parallel for over a 3D range on `stream`: f(i0, i1, i2).
template <class F>
void for_each(backend::HIP, Range3D range, F f, hipStream_t stream = nullptr) {
  if (range.size() == 0) return;
  hipLaunchKernelGGL(for_each_kernel<F>, dim3(for_each_blocks(range.size())), dim3(256), 0,
                     stream, range, f);
  MISO_HIP_RT_CHECK(hipGetLastError());
}

/// One device field with MISO's (i, j, k) indexing over the x-fastest
/// layout: element (i, j, k) at (k * j_total + j) * i_total + i.
template <typename Real> struct Array3DView {
  Real *p;
  int i_total, j_total, k_total;
  __host__ __device__ Real &operator()(int i, int j, int k) const {
    return p[((long)k * j_total + j) * i_total + i];
  }
};

/// The nine fields of a misohip_fields (FieldsView, mhd_fields.hpp).
template <typename Real> struct FieldsView {
  Array3DView<Real> ro, vx, vy, vz, bx, by, bz, ei, ph;
  FieldsView(const misohip_fields &q, const misohip_grid &g) {
    Array3DView<Real> *v[9] = {&ro, &vx, &vy, &vz, &bx, &by, &bz, &ei, &ph};
    for (int f = 0; f < 9; ++f)
      *v[f] = Array3DView<Real>{static_cast<Real *>(q.q[f]), g.i_total, g.j_total, g.k_total};
  }
  __host__ __device__ Array3DView<Real> operator[](int f) const {
    switch (f) {
      case 0: return ro;
      case 1: return vx;
      case 2: return vy;
      case 3: return vz;
      case 4: return bx;
      case 5: return by;
      case 6: return bz;
      case 7: return ei;
      default: return ph;
    }
  }
};

/// Shape, coordinates and metric of the grid on the device (GridView,
/// grid.hpp:103-141).
template <typename Real> struct GridView {
  int i_total, j_total, k_total;
  int is, js, ks;
  int i_margin, j_margin, k_margin;
  const Real *x, *y, *z;
  const Real *dx, *dy, *dz;
  const Real *dxi, *dyi, *dzi;
  Real min_dxyz;
  explicit GridView(const Grid<Real> &grid) {
    const misohip_grid &g = *grid.desc();
    i_total = g.i_total; j_total = g.j_total; k_total = g.k_total;
    is = g.is; js = g.js; ks = g.ks;
    i_margin = g.i_margin; j_margin = g.j_margin; k_margin = g.k_margin;
    x = grid.x(); y = grid.y(); z = grid.z();
    dx = static_cast<const Real *>(g.dx);
    dy = static_cast<const Real *>(g.dy);
    dz = static_cast<const Real *>(g.dz);
    dxi = static_cast<const Real *>(g.dxi);
    dyi = static_cast<const Real *>(g.dyi);
    dzi = static_cast<const Real *>(g.dzi);
    min_dxyz = Real(g.min_dxyz);
  }
};

enum class Sign { Pos, Neg };
enum class Direction { X, Y, Z };
enum class Side { INNER, OUTER };

/// Mirror of ghost layer g (0 .. margin-1) across a face: the ghost index
/// and the interior index it copies (boundary_condition.hpp:77-90).
__host__ __device__ inline void symmetric_pair(int g, int total, int margin, Side side,
                                               int &ghost, int &target) {
  if (side == Side::INNER) {
    ghost = g;
    target = 2 * margin - 1 - g;
  } else {
    ghost = total - margin + g;
    target = total - margin - 1 - g;
  }
}

template <typename Real>
__global__ void symmetric_kernel(Array3DView<Real> a, int margin, Real sign, int dir, Side side) {
  const int n0 = dir == 0 ? margin : a.i_total;
  const int n1 = dir == 1 ? margin : a.j_total;
  const int n2 = dir == 2 ? margin : a.k_total;
  const long n = (long)n0 * n1 * n2;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    int c[3] = {(int)(t % n0), (int)((t / n0) % n1), (int)(t / ((long)n0 * n1))};
    int gi = c[dir], ti = c[dir];
    symmetric_pair(c[dir], dir == 0 ? a.i_total : dir == 1 ? a.j_total : a.k_total, margin, side,
                   gi, ti);
    int s[3] = {c[0], c[1], c[2]};
    c[dir] = gi;
    s[dir] = ti;
    a(c[0], c[1], c[2]) = sign * a(s[0], s[1], s[2]);
  }
}

/// bc::symmetric (boundary_condition.hpp:108-182): the ghost layers of one
/// face of `arr` set to +-its mirror image, over the full extent of the two
/// other axes, on `stream`.
template <typename Real>
void symmetric(backend::HIP, Array3DView<Real> arr, const GridView<Real> &grid, Sign sign,
               Direction direction, Side side, hipStream_t stream = nullptr) {
  const int dir = direction == Direction::X ? 0 : direction == Direction::Y ? 1 : 2;
  const int margin = dir == 0 ? grid.i_margin : dir == 1 ? grid.j_margin : grid.k_margin;
  if (margin == 0) return;
  const long n = (long)margin * (dir == 0 ? 1 : arr.i_total) * (dir == 1 ? 1 : arr.j_total) *
                 (dir == 2 ? 1 : arr.k_total);
  hipLaunchKernelGGL(symmetric_kernel<Real>, dim3(for_each_blocks(n)), dim3(256), 0, stream, arr,
                     margin, sign == Sign::Pos ? Real(1) : Real(-1), dir, side);
  MISO_HIP_RT_CHECK(hipGetLastError());
}

/// A MISO Source struct -- vx, vy, vz, ei(FieldsView<const Real>, i, j, k)
/// as __host__ __device__ members (mhd_model_base.hpp:131-155,
/// demo/mhd3d_magnetosphere/src/src.hpp:4-41) -- as a UserSource: before
/// every RK stage a kernel evaluates the four members on the stage's argm at
/// every interior cell and stores them for the stage to add, where the
/// reference inlines the calls (mhd_integrator_host.hpp:134,151,168,241).
///
///   struct PointMass { GridView<double> g; double gg; __host__ __device__
///     double vx(FieldsView<const double> q, int i, int j, int k) const {...} ... };
///   auto src = miso::hip::make_source(mhd.grid, PointMass{GridView<double>(mhd.grid), 0.3});
///   mhd.run(eos, bc, src, nsteps);
template <typename Real, class Src> UserSource<Real> make_source(const Grid<Real> &grid, Src src) {
  const misohip_grid g = *grid.desc();
  return UserSource<Real>(grid, [g, src](const misohip_fields &argm, Real *const out[4],
                                         hipStream_t stream) {
    const FieldsView<const Real> q(argm, g);
    const Array3DView<Real> sx{out[0], g.i_total, g.j_total, g.k_total};
    const Array3DView<Real> sy{out[1], g.i_total, g.j_total, g.k_total};
    const Array3DView<Real> sz{out[2], g.i_total, g.j_total, g.k_total};
    const Array3DView<Real> se{out[3], g.i_total, g.j_total, g.k_total};
    const Range3D inner{{g.i_margin, g.i_total - g.i_margin},
                        {g.j_margin, g.j_total - g.j_margin},
                        {g.k_margin, g.k_total - g.k_margin}};
    for_each(backend::HIP{}, inner,
             [=] __device__(int i, int j, int k) {
               sx(i, j, k) = src.vx(q, i, j, k);
               sy(i, j, k) = src.vy(q, i, j, k);
               sz(i, j, k) = src.vz(q, i, j, k);
               se(i, j, k) = src.ei(q, i, j, k);
             },
             stream);
  });
}

}  // namespace hip
}  // namespace miso
