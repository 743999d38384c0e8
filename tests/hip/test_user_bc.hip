// Tests of user boundary functors (Boundary::USER, miso_hip_device.hpp):
// the same run with a boundary descriptor and with a user functor that
// writes the same ghost values through for_each / symmetric.  Built into
// tests/hip/_build/libmiso_hip_user_bc.so and called from
// tests/test_user_bc.py.
#include <cstring>
#include <vector>

#include "miso_hip_device.hpp"

using namespace miso;
namespace mh = miso::hip;

namespace {

// nsteps of MHD::run on a state given in MISO order (9 fields with ghosts)
template <typename Setup>
int run_case(int ni, int nj, int nk, const double *q_in, int nsteps, Setup setup, double *q_out,
             double *dts) {
  mh::HostGrid<double> hg(ni, nj, nk, 2, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0);
  mh::ExecContext<double> ctx(0);
  mh::MHDConfig cfg;
  cfg.cfl_number = 0.4;
  mh::MHD<double> mhd(cfg, hg, ctx);
  const std::size_t N = hg.n_total();
  auto ic = [&](mh::HostFields<double> &q, const mh::HostGrid<double> &) {
    for (int f = 0; f < 9; ++f) std::memcpy(q.q[f].data(), q_in + f * N, sizeof(double) * N);
  };
  mh::Fields<double> reservoir(mhd.grid);
  {
    mh::HostFields<double> h(hg.i_total(), hg.j_total(), hg.k_total());
    ic(h, hg);
    reservoir.copy_from(ctx, h);
  }
  mh::Boundary bc = setup(mhd, reservoir);
  mhd.apply_initial_condition(ic, bc);
  const std::vector<double> d = mhd.run(mh::IdealEOS{5.0 / 3.0}, bc, mh::NoSource{}, nsteps);
  for (int s = 0; s < nsteps; ++s) dts[s] = d[s];
  mh::HostFields<double> out(hg.i_total(), hg.j_total(), hg.k_total());
  mhd.qq.copy_to(ctx, out);
  for (int f = 0; f < 9; ++f) std::memcpy(q_out + f * N, out.q[f].data(), sizeof(double) * N);
  return 0;
}

}  // namespace

extern "C" {

// z mirrored with per-field signs (vz, bz odd), x and y periodic: by the
// descriptor (kind MIRROR) and by a USER functor calling symmetric on both
// z faces of every field.
int hip_user_bc_mirror_z(int ni, int nj, int nk, const double *q_in, int nsteps, double *q_desc,
                         double *q_user, double *dts_desc, double *dts_user) {
  try {
    const int odd[9] = {0, 0, 0, 1, 0, 0, 1, 0, 0};
    run_case(ni, nj, nk, q_in, nsteps,
             [&](mh::MHD<double> &, mh::Fields<double> &) {
               mh::Boundary bc(mh::Boundary::PERIODIC, mh::Boundary::PERIODIC,
                               mh::Boundary::MIRROR);
               for (int f = 0; f < 9; ++f) bc.sign(2, f, odd[f] ? -1 : 1);
               return bc;
             },
             q_desc, dts_desc);
    run_case(ni, nj, nk, q_in, nsteps,
             [&](mh::MHD<double> &m, mh::Fields<double> &) {
               mh::Boundary bc(mh::Boundary::PERIODIC, mh::Boundary::PERIODIC,
                               mh::Boundary::USER);
               const mh::Grid<double> *grid = &m.grid;
               bc.user([grid, odd](const misohip_fields &q, hipStream_t s) {
                 mh::FieldsView<double> v(q, *grid->desc());
                 mh::GridView<double> g(*grid);
                 for (int f = 0; f < 9; ++f)
                   for (mh::Side side : {mh::Side::INNER, mh::Side::OUTER})
                     mh::symmetric(backend::HIP{}, v[f], g, odd[f] ? mh::Sign::Neg : mh::Sign::Pos,
                                   mh::Direction::Z, side, s);
               });
               return bc;
             },
             q_user, dts_user);
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "hip_user_bc_mirror_z: %s\n", e.what());
    return 2;
  }
}

// x ghost cells held at the initial state (a prescribed inflow/outflow),
// y and z periodic: by the descriptor (kind FIXED, reading a reservoir) and
// by a USER functor that copies the reservoir's x ghost cells with for_each.
int hip_user_bc_fixed_x(int ni, int nj, int nk, const double *q_in, int nsteps, double *q_desc,
                        double *q_user, double *dts_desc, double *dts_user) {
  try {
    run_case(ni, nj, nk, q_in, nsteps,
             [&](mh::MHD<double> &, mh::Fields<double> &res) {
               mh::Boundary bc(mh::Boundary::FIXED, mh::Boundary::PERIODIC,
                               mh::Boundary::PERIODIC);
               bc.fixed_values(res);
               return bc;
             },
             q_desc, dts_desc);
    run_case(ni, nj, nk, q_in, nsteps,
             [&](mh::MHD<double> &m, mh::Fields<double> &res) {
               mh::Boundary bc(mh::Boundary::USER, mh::Boundary::PERIODIC,
                               mh::Boundary::PERIODIC);
               const mh::Grid<double> *grid = &m.grid;
               const mh::Fields<double> *r = &res;
               bc.user([grid, r](const misohip_fields &q, hipStream_t s) {
                 const misohip_grid &g = *grid->desc();
                 mh::FieldsView<double> v(q, g), w(*r->desc(), g);
                 const int m = g.i_margin, it = g.i_total;
                 // the 2m ghost columns of x over every j, k: i = 0..m-1 and
                 // i_total-m .. i_total-1
                 mh::for_each(backend::HIP{},
                              mh::Range3D{{0, 2 * m}, {0, g.j_total}, {0, g.k_total}},
                              [=] __device__(int a, int j, int k) {
                                const int i = a < m ? a : it - 2 * m + a;
                                for (int f = 0; f < 9; ++f) v[f](i, j, k) = w[f](i, j, k);
                              },
                              s);
               });
               return bc;
             },
             q_user, dts_user);
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "hip_user_bc_fixed_x: %s\n", e.what());
    return 2;
  }
}

}  // extern "C"

// ---------------------------------------------------------------- sources
namespace {
// the point-mass gravity of demo/mhd3d_magnetosphere/src/src.hpp:18-34 with
// a heating src.ei = h ro, written as a MISO Source struct
struct PointMass {
  mh::GridView<double> grid;
  double g_grav, heat;
  __host__ __device__ double r3(int i, int j, int k) const {
    const double r = sqrt(grid.x[i] * grid.x[i] + grid.y[j] * grid.y[j] + grid.z[k] * grid.z[k]);
    return r * r * r;
  }
  __host__ __device__ double vx(mh::FieldsView<const double> qq, int i, int j, int k) const {
    return -qq.ro(i, j, k) * g_grav * grid.x[i] / r3(i, j, k);
  }
  __host__ __device__ double vy(mh::FieldsView<const double> qq, int i, int j, int k) const {
    return -qq.ro(i, j, k) * g_grav * grid.y[j] / r3(i, j, k);
  }
  __host__ __device__ double vz(mh::FieldsView<const double> qq, int i, int j, int k) const {
    return -qq.ro(i, j, k) * g_grav * grid.z[k] / r3(i, j, k);
  }
  __host__ __device__ double ei(mh::FieldsView<const double> qq, int i, int j, int k) const {
    return heat * qq.ro(i, j, k);
  }
};
}  // namespace

extern "C" {

// nsteps of MHD::run (use_run) or of cfl + update on a state in MISO order,
// periodic, with the point-mass Source through make_source; bounds[6] are the
// box's extents.
int hip_user_source_point_mass(int ni, int nj, int nk, const double *bounds, const double *q_in,
                               int nsteps, int use_run, double g_grav, double heat,
                               double *q_out, double *dts) {
  try {
    mh::HostGrid<double> hg(ni, nj, nk, 2, bounds[0], bounds[1], bounds[2], bounds[3], bounds[4],
                            bounds[5]);
    mh::ExecContext<double> ctx(0);
    mh::MHDConfig cfg;
    cfg.cfl_number = 0.4;
    mh::MHD<double> mhd(cfg, hg, ctx);
    const std::size_t N = hg.n_total();
    auto ic = [&](mh::HostFields<double> &q, const mh::HostGrid<double> &) {
      for (int f = 0; f < 9; ++f) std::memcpy(q.q[f].data(), q_in + f * N, sizeof(double) * N);
    };
    mh::Boundary bc;
    mhd.apply_initial_condition(ic, bc);
    auto src = mh::make_source(mhd.grid, PointMass{mh::GridView<double>(mhd.grid), g_grav, heat});
    mh::IdealEOS eos{5.0 / 3.0};
    if (use_run) {
      const std::vector<double> d = mhd.run(eos, bc, src, nsteps);
      for (int s = 0; s < nsteps; ++s) dts[s] = d[s];
    } else {
      for (int s = 0; s < nsteps; ++s) {
        const double dt = mhd.cfl(eos);
        mhd.update(dt, eos, bc, src);
        dts[s] = dt;
      }
    }
    mh::HostFields<double> out(hg.i_total(), hg.j_total(), hg.k_total());
    mhd.qq.copy_to(ctx, out);
    for (int f = 0; f < 9; ++f) std::memcpy(q_out + f * N, out.q[f].data(), sizeof(double) * N);
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "hip_user_source_point_mass: %s\n", e.what());
    return 2;
  }
}

}  // extern "C"
