// Driver that exposes pieces of the REFERENCE's own headers through a C ABI,
// so the oracle's restatement can be checked against the real thing.  Only
// the self-contained headers are used (limiter.hpp + utility.hpp +
// cuda_compat.hpp + constants.hpp, rt_quadrature.hpp, table_interpolator.hpp
// with its array headers); the integrator
// headers need yaml-cpp and MPI and are not buildable here.  Compiled from
// /root/reference by oracle/Makefile into oracle/_ref/ -- test
// infrastructure only, never shipped.
#include <miso/constants.hpp>
#include <miso/limiter.hpp>
#include <miso/rt_quadrature.hpp>
#include <miso/table_interpolator.hpp>
#include <miso/utility.hpp>

extern "C" {
double ref_dqq_eval_f64(double a, double b, double c, double ep) {
  return miso::limiter::dqq_eval<double>(a, b, c, ep);
}
float ref_dqq_eval_f32(float a, float b, float c, float ep) {
  return miso::limiter::dqq_eval<float>(a, b, c, ep);
}
double ref_flux_core_f64(double a, double b, double da, double db, double cc, double fh) {
  return miso::limiter::flux_core<double>(a, b, da, db, cc, fh);
}
float ref_flux_core_f32(float a, float b, float da, float db, float cc, float fh) {
  return miso::limiter::flux_core<float>(a, b, da, db, cc, fh);
}
double ref_min2_f64(double a, double b) { return miso::util::min2<double>(a, b); }
double ref_max3_f64(double a, double b, double c) { return miso::util::max3<double>(a, b, c); }
double ref_min3_f64(double a, double b, double c) { return miso::util::min3<double>(a, b, c); }
double ref_const_f64(int which) {
  switch (which) {
    case 0: return miso::inv12<double>;
    case 1: return miso::pii4<double>;
    case 2: return miso::pii8<double>;
    default: return miso::pi<double>;
  }
}
float ref_const_f32(int which) {
  switch (which) {
    case 0: return miso::inv12<float>;
    case 1: return miso::pii4<float>;
    case 2: return miso::pii8<float>;
    default: return miso::pi<float>;
  }
}
// Carlson A4 quadrature: writes num_rays * 4 values (w, mu_x, mu_y, mu_z)
int ref_quadrature_f64(int num_rays, double *out) {
  miso::rt::AngularQuadrature<double> q(num_rays);
  for (int r = 0; r < q.num_rays; ++r) {
    out[4 * r + 0] = q.weights[r];
    out[4 * r + 1] = q.mu_x[r];
    out[4 * r + 2] = q.mu_y[r];
    out[4 * r + 3] = q.mu_z[r];
  }
  return q.num_rays;
}
// uniform-table interpolation, element by element
void ref_table1d_f64(const double *t, int n, double x_min, double x_max, const double *x,
                     double *y, int count) {
  const double dx = (x_max - x_min) / static_cast<double>(n - 1), dxi = 1.0 / dx;
  for (int i = 0; i < count; ++i) y[i] = miso::interpolate_uniform_table1d<double>(t, x_min, dxi, x[i], n);
}
void ref_table1d_f32(const float *t, int n, float x_min, float x_max, const float *x, float *y,
                     int count) {
  const float dx = (x_max - x_min) / static_cast<float>(n - 1), dxi = 1.0f / dx;
  for (int i = 0; i < count; ++i) y[i] = miso::interpolate_uniform_table1d<float>(t, x_min, dxi, x[i], n);
}
void ref_table2d_f64(const double *t, int n0, int n1, double a0, double b0, double a1, double b1,
                     const double *x0, const double *x1, double *y, int count) {
  const double d0 = (b0 - a0) / static_cast<double>(n0 - 1), d1 = (b1 - a1) / static_cast<double>(n1 - 1);
  const double i0 = 1.0 / d0, i1 = 1.0 / d1;
  for (int i = 0; i < count; ++i)
    y[i] = miso::interpolate_uniform_table2d<double>(t, a0, i0, a1, i1, x0[i], x1[i], n0, n1);
}
void ref_table2d_f32(const float *t, int n0, int n1, float a0, float b0, float a1, float b1,
                     const float *x0, const float *x1, float *y, int count) {
  const float d0 = (b0 - a0) / static_cast<float>(n0 - 1), d1 = (b1 - a1) / static_cast<float>(n1 - 1);
  const float i0 = 1.0f / d0, i1 = 1.0f / d1;
  for (int i = 0; i < count; ++i)
    y[i] = miso::interpolate_uniform_table2d<float>(t, a0, i0, a1, i1, x0[i], x1[i], n0, n1);
}
}
