// Tests of the C++ host layer (include/miso_hip.hpp), built into a shared
// library that pytest loads with ctypes, so every check runs inside the
// test process (no child program is started from a process that has
// touched the GPU).
//
//   cpp_test_hd1d_shock_tube   the reference's end-to-end test,
//                              miso/tests/serial/test_hd1d_shock_tube.cpp:20-81:
//                              Sod along x, y and z (FP32, 128 cells, gamma
//                              1.4, mirror faces, cfl 0.5, run to t = 0.2),
//                              L1 error of ro and vx below 5e-2 against the
//                              exact Riemann solution, x/y/z agreement 1e-5
//   cpp_run_steps              N steps of a given state through the C++
//                              classes, for bitwise comparison with the oracle
//   cpp_boundary_scope         a C++ call runs exactly its own Boundary's
//                              functor (none when it has none) and leaves a
//                              callback set through the C ABI in place
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "miso_hip.hpp"

using namespace miso;

namespace {

// Exact solution of the Riemann problem of the Euler equations (Toro,
// "Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4).
struct ExactRiemann {
  double gm, rl, ul, pl, rr, ur, pr, cl, cr, ps, us;
  ExactRiemann(double gm_, double rl_, double ul_, double pl_, double rr_, double ur_, double pr_)
      : gm(gm_), rl(rl_), ul(ul_), pl(pl_), rr(rr_), ur(ur_), pr(pr_) {
    cl = std::sqrt(gm * pl / rl);
    cr = std::sqrt(gm * pr / rr);
    double p = 0.5 * (pl + pr);
    for (int it = 0; it < 100; ++it) {
      double fl, dfl, fr, dfr;
      side(p, rl, pl, cl, fl, dfl);
      side(p, rr, pr, cr, fr, dfr);
      const double f = fl + fr + (ur - ul);
      const double pn = std::fmax(1e-12, p - f / (dfl + dfr));
      if (std::fabs(pn - p) < 1e-15 * (pn + p)) { p = pn; break; }
      p = pn;
    }
    ps = p;
    double fl, dfl, fr, dfr;
    side(ps, rl, pl, cl, fl, dfl);
    side(ps, rr, pr, cr, fr, dfr);
    us = 0.5 * (ul + ur) + 0.5 * (fr - fl);
  }
  void side(double p, double rk, double pk, double ck, double &f, double &df) const {
    if (p > pk) {  // shock
      const double A = 2.0 / ((gm + 1.0) * rk), B = (gm - 1.0) / (gm + 1.0) * pk;
      const double q = std::sqrt(A / (p + B));
      f = (p - pk) * q;
      df = q * (1.0 - 0.5 * (p - pk) / (p + B));
    } else {  // rarefaction
      const double e = (gm - 1.0) / (2.0 * gm);
      f = 2.0 * ck / (gm - 1.0) * (std::pow(p / pk, e) - 1.0);
      df = 1.0 / (rk * ck) * std::pow(p / pk, -(gm + 1.0) / (2.0 * gm));
    }
  }
  // density and velocity at x/t = s
  void sample(double s, double &r, double &u) const {
    const double g1 = (gm - 1.0) / (gm + 1.0);
    if (s <= us) {
      if (ps > pl) {
        const double sl = ul - cl * std::sqrt((gm + 1.0) / (2.0 * gm) * ps / pl + (gm - 1.0) / (2.0 * gm));
        if (s <= sl) { r = rl; u = ul; }
        else { r = rl * (ps / pl + g1) / (g1 * ps / pl + 1.0); u = us; }
      } else {
        const double shl = ul - cl;
        const double csl = cl * std::pow(ps / pl, (gm - 1.0) / (2.0 * gm));
        const double stl = us - csl;
        if (s <= shl) { r = rl; u = ul; }
        else if (s > stl) { r = rl * std::pow(ps / pl, 1.0 / gm); u = us; }
        else {
          const double c = 2.0 / (gm + 1.0) * (cl + 0.5 * (gm - 1.0) * (ul - s));
          r = rl * std::pow(c / cl, 2.0 / (gm - 1.0));
          u = 2.0 / (gm + 1.0) * (cl + 0.5 * (gm - 1.0) * ul + s);
        }
      }
    } else {
      if (ps > pr) {
        const double sr = ur + cr * std::sqrt((gm + 1.0) / (2.0 * gm) * ps / pr + (gm - 1.0) / (2.0 * gm));
        if (s >= sr) { r = rr; u = ur; }
        else { r = rr * (ps / pr + g1) / (g1 * ps / pr + 1.0); u = us; }
      } else {
        const double shr = ur + cr;
        const double csr = cr * std::pow(ps / pr, (gm - 1.0) / (2.0 * gm));
        const double str = us + csr;
        if (s >= shr) { r = rr; u = ur; }
        else if (s <= str) { r = rr * std::pow(ps / pr, 1.0 / gm); u = us; }
        else {
          const double c = 2.0 / (gm + 1.0) * (cr - 0.5 * (gm - 1.0) * (ur - s));
          r = rr * std::pow(c / cr, 2.0 / (gm - 1.0));
          u = 2.0 / (gm + 1.0) * (-cr + 0.5 * (gm - 1.0) * ur + s);
        }
      }
    }
  }
};

using Real = float;

struct SodRun {
  std::vector<Real> ro, vn, s;  // interior density, normal velocity, coordinate
  double time = 0.0;
};

// One shock tube along `axis` through the MISO-shaped classes.
SodRun run_sod(int axis) {
  const double gm = 1.4, rol = 1.0, prl = 1.0, ror = 0.125, prr = 0.1;
  const int n = 128;
  int sz[3] = {1, 1, 1};
  sz[axis] = n;
  hip::HostGrid<Real> hg(sz[0], sz[1], sz[2], 2, 0.f, 1.f, 0.f, 1.f, 0.f, 1.f);
  hip::ExecContext<Real> ctx(0);
  hip::MHDConfig cfg;  // config_defaults.hpp: cfl 0.5, AV defaults
  hip::MHD<Real> mhd(cfg, hg, ctx);
  hip::IdealEOS eos{gm};
  hip::Boundary bc(hip::Boundary::MIRROR, hip::Boundary::MIRROR, hip::Boundary::MIRROR);
  hip::NoSource src;
  auto ic = [&](hip::HostFields<Real> &q, const hip::HostGrid<Real> &g) {
    for (int i = 0; i < g.i_total(); ++i)
      for (int j = 0; j < g.j_total(); ++j)
        for (int k = 0; k < g.k_total(); ++k) {
          const Real xyz = axis == 0 ? g.x()[i] : axis == 1 ? g.y()[j] : g.z()[k];
          const bool left = xyz < Real(0.5);
          q(hip::RO, i, j, k) = left ? Real(rol) : Real(ror);
          q(hip::EI, i, j, k) = Real((left ? prl : prr) / (gm - 1.0) / (left ? rol : ror));
        }
  };
  mhd.apply_initial_condition(ic, bc);
  double time = 0.0;
  const double tend = 0.2;
  while (time < tend) {  // ModelBase::run / update (mhd_model_base.hpp:57-62,89-117)
    const Real dt = mhd.cfl(eos);
    mhd.update(dt, eos, bc, src);
    time += double(Real(dt));
  }
  hip::HostFields<Real> out(hg.i_total(), hg.j_total(), hg.k_total());
  mhd.qq.copy_to(ctx, out);
  SodRun r;
  r.time = time;
  const auto &coord = axis == 0 ? hg.x() : axis == 1 ? hg.y() : hg.z();
  const int m = 2;
  for (int a = m; a < m + n; ++a) {
    const int i = axis == 0 ? a : 0, j = axis == 1 ? a : 0, k = axis == 2 ? a : 0;
    r.ro.push_back(out(hip::RO, i, j, k));
    r.vn.push_back(out(hip::VX + axis, i, j, k));
    r.s.push_back(coord[a]);
  }
  return r;
}

}  // namespace

// the radiative step's C++ entry compiles for the types a MISO model uses
template double hip::Integrator<double>::update_heated<hip::IdealEOS, hip::Boundary, hip::Gravity>(
    double, const hip::IdealEOS &, const hip::Boundary &, const hip::Gravity &,
    hip::Fields<double> &, const void *);

extern "C" {

// Returns 0 when every check passes; msg receives a one-line summary.
int cpp_test_hd1d_shock_tube(char *msg, int len) {
  SodRun x = run_sod(0), y = run_sod(1), z = run_sod(2);
  ExactRiemann ex(1.4, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1);
  double dro = 0, dvx = 0, sro = 0, svx = 0;
  for (std::size_t a = 0; a < x.ro.size(); ++a) {
    double r, u;
    ex.sample((double(x.s[a]) - 0.5) / x.time, r, u);
    dro += std::fabs(x.ro[a] - r);
    dvx += std::fabs(x.vn[a] - u);
    sro += x.ro[a];
    svx += x.vn[a];
  }
  double worst = 0.0;
  for (std::size_t a = 0; a < x.ro.size(); ++a) {
    const double d1 = std::fabs(double(x.ro[a]) - y.ro[a]) / (1.0 + std::fmax(std::fabs(x.ro[a]), std::fabs(y.ro[a])));
    const double d2 = std::fabs(double(x.ro[a]) - z.ro[a]) / (1.0 + std::fmax(std::fabs(x.ro[a]), std::fabs(z.ro[a])));
    worst = std::fmax(worst, std::fmax(d1, d2));
  }
  const bool ok = dro / sro < 5e-2 && dvx / svx < 5e-2 && worst <= 1e-5 && x.time == y.time &&
                  x.time == z.time;
  std::snprintf(msg, len, "t=%.6f L1(ro)=%.4e L1(vx)=%.4e xyz=%.3e p*=%.6f", x.time, dro / sro,
                dvx / svx, worst, ex.ps);
  return ok ? 0 : 1;
}

// Runs nsteps of ModelBase::update on a state given in MISO order (9 fields
// of i_total*j_total*k_total values, ghosts included) through HostGrid,
// MHD and Integrator; writes the final state and every dt.
int cpp_run_steps(int fp64, int ni, int nj, int nk, const double *bounds, const int *kinds,
                  const int *signs, double gm, double cfl_number, int src_kind, int grav_axis,
                  double grav, int nsteps, const void *q_in, void *q_out, double *dts) {
  auto run = [&](auto zero) -> int {
    using R = decltype(zero);
    hip::HostGrid<R> hg(ni, nj, nk, 2, R(bounds[0]), R(bounds[1]), R(bounds[2]), R(bounds[3]),
                        R(bounds[4]), R(bounds[5]));
    hip::ExecContext<R> ctx(0);
    hip::MHDConfig cfg;
    cfg.cfl_number = cfl_number;
    hip::MHD<R> mhd(cfg, hg, ctx);
    hip::Boundary bc{static_cast<hip::Boundary::Kind>(kinds[0]),
                     static_cast<hip::Boundary::Kind>(kinds[1]),
                     static_cast<hip::Boundary::Kind>(kinds[2])};
    for (int a = 0; a < 3; ++a)
      for (int f = 0; f < 9; ++f) bc.sign(a, f, signs[a * 9 + f]);
    hip::IdealEOS eos{gm};
    const std::size_t N = hg.n_total();
    const R *src_q = static_cast<const R *>(q_in);
    auto ic = [&](hip::HostFields<R> &q, const hip::HostGrid<R> &) {
      for (int f = 0; f < 9; ++f) std::memcpy(q.q[f].data(), src_q + f * N, sizeof(R) * N);
    };
    // a fixed axis takes its ghost values from the initial state
    hip::Fields<R> fixed(mhd.grid);
    if (kinds[0] == 2 || kinds[1] == 2 || kinds[2] == 2) {
      hip::HostFields<R> h(hg.i_total(), hg.j_total(), hg.k_total());
      ic(h, hg);
      fixed.copy_from(ctx, h);
      bc.fixed_values(fixed);
    }
    mhd.apply_initial_condition(ic, bc);
    if (src_kind == 1) {  // step by step: cfl, then update
      for (int s = 0; s < nsteps; ++s) {
        hip::Gravity g{grav_axis, grav};
        const R dt = mhd.cfl(eos);
        mhd.update(dt, eos, bc, g);
        dts[s] = double(dt);
      }
    } else {  // the whole loop in one call
      const std::vector<double> d = mhd.run(eos, bc, hip::NoSource{}, nsteps);
      for (int s = 0; s < nsteps; ++s) dts[s] = d[s];
    }
    hip::HostFields<R> out(hg.i_total(), hg.j_total(), hg.k_total());
    mhd.qq.copy_to(ctx, out);
    R *dst = static_cast<R *>(q_out);
    for (int f = 0; f < 9; ++f) std::memcpy(dst + f * N, out.q[f].data(), sizeof(R) * N);
    return 0;
  };
  try {
    return fp64 ? run(double(0)) : run(float(0));
  } catch (const std::exception &e) {
    std::fprintf(stderr, "cpp_run_steps: %s\n", e.what());
    return 2;
  }
}

namespace {
int count_calls(void *user, const misohip_fields *, void *) {
  ++*static_cast<int *>(user);
  return 0;
}
}  // namespace

// A callback installed through the C ABI; C++ fills and steps with a plain
// Boundary must not run it, with a user Boundary must run the functor
// instead, and both must leave it installed.  Returns 0 when all hold; msg
// names the first check that failed.
int cpp_boundary_scope(char *msg, int len) {
  auto fail = [&](const char *what) {
    std::snprintf(msg, len, "%s", what);
    return 1;
  };
  try {
    hip::HostGrid<double> hg(8, 8, 8, 2, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0);
    hip::ExecContext<double> ctx(0);
    hip::MHDConfig cfg;
    hip::MHD<double> mhd(cfg, hg, ctx);
    hip::IdealEOS eos{5.0 / 3.0};
    auto ic = [](hip::HostFields<double> &q, const hip::HostGrid<double> &) {
      for (std::size_t n = 0; n < q.q[0].size(); ++n) {
        q.q[0][n] = 1.0 + 0.01 * double(n % 7);
        q.q[7][n] = 1.5;
      }
    };
    int c_calls = 0;
    if (misohip_set_bc_callback(ctx.handle(), &count_calls, &c_calls)) return fail("set");
    hip::Boundary plain{hip::Boundary::PERIODIC, hip::Boundary::PERIODIC,
                        hip::Boundary::PERIODIC};
    mhd.apply_initial_condition(ic, plain);
    const double dt = mhd.cfl(eos);
    mhd.update(dt, eos, plain, hip::NoSource{});
    ctx.synchronize();
    if (c_calls != 0) return fail("a plain Boundary ran the C-level callback");
    int user_calls = 0;
    hip::Boundary user = plain;
    user.user([&](const misohip_fields &, hipStream_t) { ++user_calls; });
    mhd.update(dt, eos, user, hip::NoSource{});
    ctx.synchronize();
    if (c_calls != 0) return fail("a user Boundary ran the C-level callback");
    if (user_calls != 7) return fail("the user functor did not run before each of 7 fills");
    misohip_bc_fn fn = nullptr;
    void *u = nullptr;
    if (misohip_get_bc_callback(ctx.handle(), &fn, &u) || fn != &count_calls || u != &c_calls)
      return fail("the C-level callback was not restored");
    hip::Fields<double> &qq = mhd.qq;
    misohip_bc d = plain.d;
    if (misohip_ghost_fill(ctx.handle(), mhd.grid.desc(), MISOHIP_F64, qq.desc(), &d))
      return fail("ghost fill");
    ctx.synchronize();
    if (c_calls != 1) return fail("the restored callback did not run on a C-level fill");
    misohip_set_bc_callback(ctx.handle(), nullptr, nullptr);
    std::snprintf(msg, len, "ok");
    return 0;
  } catch (const std::exception &e) {
    return fail(e.what());
  }
}

// RT::solve through miso::hip::RT on kappa and S given in MISO order (24
// rays); writes the iteration count, every ray's intensities in MISO order
// and Q_rad (MISO order, computed on the device).
int cpp_rt_solve(int ni, int nj, int nk, const double *bounds, const double *kappa,
                 const double *S, const int *periodic, const int *face, const double *value,
                 double tol, int max_iters, int yfast, int sweep_planes, int *iters,
                 double *I_out, double *q_out) {
  try {
    hip::HostGrid<double> hg(ni, nj, nk, 2, bounds[0], bounds[1], bounds[2], bounds[3],
                             bounds[4], bounds[5]);
    hip::ExecContext<double> ctx(0);
    hip::Grid<double> grid(hg);
    hip::RT rt(ctx, grid, 24, yfast != 0);
    rt.sweep_planes = sweep_planes;
    const std::size_t N = hg.n_total();
    rt.set_fields(std::vector<double>(kappa, kappa + N), std::vector<double>(S, S + N));
    hip::RtBoundary bc;
    for (int a = 0; a < 3; ++a) bc.periodic[a] = periodic[a];
    for (int f = 0; f < 6; ++f) {
      bc.face[f] = face[f];
      bc.value[f] = value[f];
    }
    *iters = rt.solve(tol, max_iters, bc);
    for (int n = 0; n < rt.num_rays; ++n) {
      const std::vector<double> I = rt.intensity(n);
      std::memcpy(I_out + N * n, I.data(), sizeof(double) * N);
    }
    hip::DeviceArray<double> q(N);
    rt.qrad(q);
    MISO_HIP_CHECK(misohip_copy_to_host(ctx.handle(), grid.desc(), MISOHIP_F64, q.data(), q_out));
    ctx.synchronize();
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "cpp_rt_solve: %s\n", e.what());
    return 2;
  }
}

// RT::save then RT::load into a second RT: after a solve on the given
// kappa, S (24 rays, x and y periodic, thermal bottom, vacuum top), the file
// is written to `path` and read back; returns 0 when every ray's
// intensities and the fields come back equal, 1 otherwise.
int cpp_rt_save_load(int ni, int nj, int nk, const double *kappa, const double *S,
                     const char *path) {
  try {
    hip::HostGrid<double> hg(ni, nj, nk, 2, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5);
    hip::ExecContext<double> ctx(0);
    hip::Grid<double> grid(hg);
    hip::RT a(ctx, grid, 24);
    const std::size_t N = hg.n_total();
    a.set_fields(std::vector<double>(kappa, kappa + N), std::vector<double>(S, S + N));
    a.solve(1e-8, 100, hip::RtBoundary{});
    a.save(path);
    hip::RT b(ctx, grid, 24);
    b.load(path);
    for (int n = 0; n < 24; ++n)
      if (a.intensity(n) != b.intensity(n)) return 1;
    if (a.weights != b.weights || a.mu_x != b.mu_x) return 1;
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "cpp_rt_save_load: %s\n", e.what());
    return 2;
  }
}

}  // extern "C"
