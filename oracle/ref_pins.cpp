// Driver that exposes pieces of the REFERENCE's own headers through a C ABI,
// so the oracle's restatement can be checked against the real thing.  Only
// the self-contained headers are used (limiter.hpp + utility.hpp +
// cuda_compat.hpp + constants.hpp, rt_quadrature.hpp); the integrator
// headers need yaml-cpp and MPI and are not buildable here.  Compiled from
// /root/reference by oracle/Makefile into oracle/_ref/ -- test
// infrastructure only, never shipped.
#include <miso/constants.hpp>
#include <miso/limiter.hpp>
#include <miso/rt_quadrature.hpp>
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
}
