// Uniform-table interpolation on the device: MISO's UniformTableInterpolator1D
// and 2D (miso/include/miso/table_interpolator.hpp:11-166), whose interpolate()
// runs element by element over arrays through for_each.  One thread per
// element, grid-stride; the arithmetic is the reference's, in the table's
// precision: the index clamped to the table, the integer part truncated and
// clamped to n-2, the linear (bilinear) weights in the same order.
#include "launch.hpp"

namespace misohip {

template <typename R> __device__ __forceinline__ R clampv(R v, R lo, R hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

template <typename R>
__device__ __forceinline__ void table_index(R x, R x_min, R dxi, int n, int &i, R &t) {
  R f = (x - x_min) * dxi;
  f = clampv<R>(f, R(0), R(n - 1));
  i = clampv<int>(static_cast<int>(f), 0, n - 2);
  t = f - R(i);
}

template <typename R>
__global__ __launch_bounds__(256) void table1d_kernel(const R *__restrict__ table, int n,
                                                      R x_min, R dxi, const R *__restrict__ x,
                                                      R *__restrict__ y, long count) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
       e += (long)gridDim.x * blockDim.x) {
    int i;
    R t;
    table_index<R>(x[e], x_min, dxi, n, i, t);
    y[e] = (R(1) - t) * table[i] + t * table[i + 1];
  }
}

template <typename R>
__global__ __launch_bounds__(256) void table2d_kernel(const R *__restrict__ table, int n0, int n1,
                                                      R x0_min, R dxi0, R x1_min, R dxi1,
                                                      const R *__restrict__ x0,
                                                      const R *__restrict__ x1,
                                                      R *__restrict__ y, long count) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
       e += (long)gridDim.x * blockDim.x) {
    int i, j;
    R t, u;
    table_index<R>(x0[e], x0_min, dxi0, n0, i, t);
    table_index<R>(x1[e], x1_min, dxi1, n1, j, u);
    const R ti = R(1) - t, ui = R(1) - u;
    const int ij00 = i * n1 + j, ij01 = ij00 + 1, ij10 = ij00 + n1, ij11 = ij10 + 1;
    y[e] = ti * ui * table[ij00] + ti * u * table[ij01] + t * ui * table[ij10] +
           t * u * table[ij11];
  }
}

static unsigned table_blocks(long count) {
  const long b = (count + 255) / 256;
  return (unsigned)(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

template <typename R>
hipError_t launch_table1d(hipStream_t st, const R *table, int n, R x_min, R dxi, const R *x, R *y,
                          long count) {
  hipLaunchKernelGGL(table1d_kernel<R>, dim3(table_blocks(count)), dim3(256), 0, st, table, n,
                     x_min, dxi, x, y, count);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_table2d(hipStream_t st, const R *table, int n0, int n1, R x0_min, R dxi0,
                          R x1_min, R dxi1, const R *x0, const R *x1, R *y, long count) {
  hipLaunchKernelGGL(table2d_kernel<R>, dim3(table_blocks(count)), dim3(256), 0, st, table, n0,
                     n1, x0_min, dxi0, x1_min, dxi1, x0, x1, y, count);
  return hipGetLastError();
}

template hipError_t launch_table1d<double>(hipStream_t, const double *, int, double, double,
                                           const double *, double *, long);
template hipError_t launch_table1d<float>(hipStream_t, const float *, int, float, float,
                                          const float *, float *, long);
template hipError_t launch_table2d<double>(hipStream_t, const double *, int, int, double, double,
                                           double, double, const double *, const double *,
                                           double *, long);
template hipError_t launch_table2d<float>(hipStream_t, const float *, int, int, float, float,
                                          float, float, const float *, const float *, float *,
                                          long);

}  // namespace misohip
