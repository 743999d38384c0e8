// miso_hip.hpp -- C++17 host layer of the MI355X backend, over misohip.h.
//
// The classes below carry MISO's names, constructor arguments and method
// signatures for the MHD time step so that a MISO model can run on the HIP
// backend by changing its Backend type; MISO's own backend::HIP
// specialisations would forward to them:
//
//   miso::hip::ExecContext<Real>     ExecContext<Real, backend::CUDA>
//                                    (miso/include/miso/mhd_core.hpp:15-33)
//   miso::hip::HostGrid<Real>        Grid<Real, backend::Host> for one rank
//                                    (miso/include/miso/grid.hpp:22-99,147-242)
//   miso::hip::Grid<Real>            Grid<Real, backend::CUDA> (device metric)
//   miso::hip::HostFields<Real>      Fields<Real, backend::Host> (z-fastest)
//                                    (miso/include/miso/mhd_fields.hpp:13-54)
//   miso::hip::Fields<Real>          Fields<Real, backend::CUDA> (x-fastest)
//   miso::hip::Integrator<Real>      Integrator<Real, Backend>::update / cfl /
//                                    apply_boundary_condition
//                                    (miso/include/miso/mhd_integrator_host.hpp:13-357)
//   miso::hip::MHD<Real>             mhd::MHD<Real, Backend> (miso/include/miso/mhd.hpp:15-50)
//   miso::hip::RT                    rt::RT<double> (miso/include/miso/rt_host.hpp:104-689),
//                                    CPU-only in the reference
//
// User boundary and source functors cannot cross a C ABI as template
// arguments; they are descriptors -- Boundary (per axis periodic / mirror
// with per-field signs / fixed ghosts) and Gravity / NoSource -- or device
// functors run through the context's callbacks: Boundary::user (a functor
// run before every ghost fill) and UserSource (the four source terms of each
// RK stage's argm written into arrays the stage adds; make_source in
// miso_hip_device.hpp builds one from a MISO Source struct).  Errors follow the reference's
// convention (miso/include/miso/cuda_util.cuh:17-33): a failed backend call
// prints file, line and message and aborts; invalid configuration throws
// std::runtime_error (miso/include/miso/grid.hpp:244-276).
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "misohip.h"

namespace miso {
namespace backend {
/// Tag of the MI355X backend (beside MISO's Host and CUDA, backend.hpp:8,12).
struct HIP {};
}  // namespace backend

namespace hip {

inline void check(int rc, const char *file, int line) {
  if (rc != MISOHIP_OK) {
    std::fprintf(stderr, "MISO HIP Error: %s %s %d\n", misohip_strerror(rc), file, line);
    std::fflush(stderr);
    std::abort();
  }
}
#define MISO_HIP_CHECK(call) ::miso::hip::check((call), __FILE__, __LINE__)

inline void check_hip(hipError_t e, const char *file, int line) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "HIP Error: %s %s %d\n", hipGetErrorString(e), file, line);
    std::fflush(stderr);
    std::abort();
  }
}
#define MISO_HIP_RT_CHECK(call) ::miso::hip::check_hip((call), __FILE__, __LINE__)

template <typename Real> constexpr misohip_dtype dtype_of() {
  static_assert(std::is_same<Real, double>::value || std::is_same<Real, float>::value,
                "Real must be double or float");
  return std::is_same<Real, double>::value ? MISOHIP_F64 : MISOHIP_F32;
}

/// Device buffer, move-only (as Array3D<T, CUDA>, array3d.hpp:284-303).
template <typename T> class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(std::size_t n) : n_(n) {
    MISO_HIP_RT_CHECK(hipMalloc(reinterpret_cast<void **>(&p_), sizeof(T) * (n ? n : 1)));
    MISO_HIP_RT_CHECK(hipMemset(p_, 0, sizeof(T) * (n ? n : 1)));
  }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceArray &operator=(DeviceArray &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_;
      o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  ~DeviceArray() { release(); }
  T *data() const { return p_; }
  std::size_t size() const { return n_; }
  void upload(const T *h, std::size_t n) {
    MISO_HIP_RT_CHECK(hipMemcpy(p_, h, sizeof(T) * n, hipMemcpyHostToDevice));
  }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
  }
  T *p_ = nullptr;
  std::size_t n_ = 0;
};

/// Execution context: device ordinal and the HIP stream of every call.
template <typename Real> class ExecContext {
 public:
  explicit ExecContext(int device = 0, hipStream_t stream = nullptr) {
    MISO_HIP_RT_CHECK(hipSetDevice(device));
    MISO_HIP_CHECK(misohip_create(device, stream, &ctx_));
  }
  ExecContext(const ExecContext &) = delete;
  ExecContext &operator=(const ExecContext &) = delete;
  ~ExecContext() {
    if (ctx_) misohip_destroy(ctx_);
  }
  misohip_ctx *handle() const { return ctx_; }
  void synchronize() const { MISO_HIP_CHECK(misohip_synchronize(ctx_)); }

 private:
  misohip_ctx *ctx_ = nullptr;
};

/// One axis of a uniform grid (AxisGrid + UniformAxisProvider, grid.hpp:20-99).
template <typename Real> struct AxisGrid {
  int size = 0, total = 0, margin = 0, stride = 0;
  std::vector<Real> s, ds, dsi;
  AxisGrid() = default;
  AxisGrid(int size_, int margin_, Real min_value, Real max_value) {
    if (size_ <= 0 || margin_ < 1) throw std::runtime_error("Invalid grid size or margin");
    if (!(min_value < max_value)) throw std::runtime_error("Invalid grid range");
    size = size_;
    stride = size > 1 ? 1 : 0;
    margin = margin_ * stride;
    total = size + 2 * margin;
    s.resize(total);
    ds.resize(total);
    dsi.resize(total);
    const Real space = (max_value - min_value) / Real(size);
    if (size == 1) {
      s[0] = min_value + (Real(0) + Real(0.5)) * space;
      ds[0] = Real(1.e30);
      dsi[0] = Real(0.0);
      return;
    }
    for (int i = 0; i < total; ++i) {
      const int ig = i - margin;
      s[i] = min_value + (Real(ig) + Real(0.5)) * space;
      ds[i] = space;
      dsi[i] = Real(1.0 / ds[i]);
    }
  }
};

/// Host grid of one rank (the single-rank case of grid.hpp:147-242).
template <typename Real> struct HostGrid {
  AxisGrid<Real> xg, yg, zg;
  Real x_min, x_max, y_min, y_max, z_min, z_max;
  Real min_dxyz;
  HostGrid(int i_size, int j_size, int k_size, int margin, Real x_min_, Real x_max_,
           Real y_min_, Real y_max_, Real z_min_, Real z_max_)
      : xg(i_size, margin, x_min_, x_max_), yg(j_size, margin, y_min_, y_max_),
        zg(k_size, margin, z_min_, z_max_), x_min(x_min_), x_max(x_max_), y_min(y_min_),
        y_max(y_max_), z_min(z_min_), z_max(z_max_) {
    // util::min3 narrows through float (utility.hpp:91-124)
    const float m = std::fmin(std::fmin(float(xg.ds[0]), float(yg.ds[0])), float(zg.ds[0]));
    min_dxyz = Real(m);
  }
  int i_size() const { return xg.size; }
  int j_size() const { return yg.size; }
  int k_size() const { return zg.size; }
  int i_total() const { return xg.total; }
  int j_total() const { return yg.total; }
  int k_total() const { return zg.total; }
  int i_margin() const { return xg.margin; }
  int j_margin() const { return yg.margin; }
  int k_margin() const { return zg.margin; }
  const std::vector<Real> &x() const { return xg.s; }
  const std::vector<Real> &y() const { return yg.s; }
  const std::vector<Real> &z() const { return zg.s; }
  std::size_t n_total() const {
    return std::size_t(xg.total) * std::size_t(yg.total) * std::size_t(zg.total);
  }
};

/// Device grid: the descriptor plus the device metric arrays.
template <typename Real> class Grid {
 public:
  explicit Grid(const HostGrid<Real> &h)
      : host_(h), x_(h.xg.total), y_(h.yg.total), z_(h.zg.total), dx_(h.xg.total),
        dy_(h.yg.total), dz_(h.zg.total), dxi_(h.xg.total), dyi_(h.yg.total),
        dzi_(h.zg.total) {
    x_.upload(h.xg.s.data(), h.xg.total);
    y_.upload(h.yg.s.data(), h.yg.total);
    z_.upload(h.zg.s.data(), h.zg.total);
    dx_.upload(h.xg.ds.data(), h.xg.total);
    dy_.upload(h.yg.ds.data(), h.yg.total);
    dz_.upload(h.zg.ds.data(), h.zg.total);
    dxi_.upload(h.xg.dsi.data(), h.xg.total);
    dyi_.upload(h.yg.dsi.data(), h.yg.total);
    dzi_.upload(h.zg.dsi.data(), h.zg.total);
    d_.i_size = h.i_size(); d_.j_size = h.j_size(); d_.k_size = h.k_size();
    d_.i_total = h.i_total(); d_.j_total = h.j_total(); d_.k_total = h.k_total();
    d_.i_margin = h.i_margin(); d_.j_margin = h.j_margin(); d_.k_margin = h.k_margin();
    d_.is = h.xg.stride; d_.js = h.yg.stride; d_.ks = h.zg.stride;
    d_.min_dxyz = double(h.min_dxyz);
    d_.dx = dx_.data(); d_.dy = dy_.data(); d_.dz = dz_.data();
    d_.dxi = dxi_.data(); d_.dyi = dyi_.data(); d_.dzi = dzi_.data();
  }
  const misohip_grid *desc() const { return &d_; }
  const HostGrid<Real> &host() const { return host_; }
  /// cell-centre coordinates on the device, ghosts included
  const Real *x() const { return x_.data(); }
  const Real *y() const { return y_.data(); }
  const Real *z() const { return z_.data(); }

 private:
  HostGrid<Real> host_;
  DeviceArray<Real> x_, y_, z_, dx_, dy_, dz_, dxi_, dyi_, dzi_;
  misohip_grid d_{};
};

enum Field { RO = 0, VX, VY, VZ, BX, BY, BZ, EI, PH, N_FIELDS };

/// Host fields in MISO order: (i, j, k) at (i * Jt + j) * Kt + k.
template <typename Real> struct HostFields {
  int i_total, j_total, k_total;
  std::array<std::vector<Real>, N_FIELDS> q;
  HostFields(int it, int jt, int kt) : i_total(it), j_total(jt), k_total(kt) {
    for (auto &f : q) f.assign(std::size_t(it) * jt * kt, Real(0));
  }
  Real &operator()(int f, int i, int j, int k) {
    return q[f][(std::size_t(i) * j_total + j) * k_total + k];
  }
  Real operator()(int f, int i, int j, int k) const {
    return q[f][(std::size_t(i) * j_total + j) * k_total + k];
  }
};

/// Device fields: nine x-fastest arrays.
template <typename Real> class Fields {
 public:
  explicit Fields(const Grid<Real> &g) : grid_(&g) {
    const std::size_t n = g.host().n_total();
    for (int f = 0; f < N_FIELDS; ++f) {
      a_[f] = DeviceArray<Real>(n);
      d_.q[f] = a_[f].data();
    }
  }
  const misohip_fields *desc() const { return &d_; }
  Real *field(int f) const { return a_[f].data(); }

  /// Fields<HIP>::copy_from(Fields<Host>) (mhd_fields.hpp:83-107)
  void copy_from(const ExecContext<Real> &ctx, const HostFields<Real> &h) {
    for (int f = 0; f < N_FIELDS; ++f)
      MISO_HIP_CHECK(misohip_copy_to_device(ctx.handle(), grid_->desc(), dtype_of<Real>(),
                                            h.q[f].data(), a_[f].data()));
  }
  /// Fields<Host>::copy_from(Fields<HIP>) (mhd_fields.hpp:148-172)
  void copy_to(const ExecContext<Real> &ctx, HostFields<Real> &h) const {
    for (int f = 0; f < N_FIELDS; ++f)
      MISO_HIP_CHECK(misohip_copy_to_host(ctx.handle(), grid_->desc(), dtype_of<Real>(),
                                          a_[f].data(), h.q[f].data()));
  }

 private:
  const Grid<Real> *grid_;
  std::array<DeviceArray<Real>, N_FIELDS> a_;
  misohip_fields d_{};
};

/// Boundary descriptor: per axis periodic (the halo self-exchange), mirror
/// with a sign per field (bc::symmetric, boundary_condition.hpp:108-136),
/// fixed (every fill copies the axis' ghost cells from given fields: a user
/// boundary functor writing set values, mhd_model_base.hpp:131-155) or user
/// (a functor run before every fill, BC::apply: see user() and
/// miso_hip_device.hpp for for_each / symmetric on the device).
struct Boundary {
  enum Kind { MIRROR = 0, PERIODIC = 1, FIXED = 2, USER = 3 };
  /// the user functor: called with the fields to fill and the stream
  using UserFn = std::function<void(const misohip_fields &, hipStream_t)>;
  misohip_bc d{};
  std::shared_ptr<UserFn> fn;
  Boundary(Kind x = PERIODIC, Kind y = PERIODIC, Kind z = PERIODIC) {
    d.periodic[0] = x; d.periodic[1] = y; d.periodic[2] = z;
    for (auto &a : d.sym_sign)
      for (int &s : a) s = 1;
  }
  /// sign of field f when mirrored across the faces of axis a (+1 / -1)
  Boundary &sign(int axis, int f, int s) {
    d.sym_sign[axis][f] = s >= 0 ? 1 : -1;
    return *this;
  }
  /// the fields whose ghost cells a FIXED axis takes (kept alive by the caller)
  template <typename F> Boundary &fixed_values(const F &fields) {
    for (int f = 0; f < 9; ++f) d.fixed[f] = fields.field(f);
    return *this;
  }
  /// the functor of the USER axes (BC::apply): f(const misohip_fields &,
  /// hipStream_t) enqueues its kernels on the stream it is given
  template <typename F> Boundary &user(F f) {
    fn = std::make_shared<UserFn>(std::move(f));
    return *this;
  }
  /// C callback over fn (misohip_set_bc_callback); exceptions stop the step
  static int call(void *user, const misohip_fields *q, void *stream) {
    try {
      (*static_cast<const UserFn *>(user))(*q, static_cast<hipStream_t>(stream));
      return 0;
    } catch (...) {
      return 1;
    }
  }
};

/// Registers a boundary's user functor on the context for one call.
class BoundaryScope {
 public:
  BoundaryScope(misohip_ctx *ctx, const Boundary &bc) : ctx_(ctx) {
    MISO_HIP_CHECK(misohip_get_bc_callback(ctx_, &prev_fn_, &prev_user_));
    // exactly this call's boundary: the functor, or none when bc has none
    MISO_HIP_CHECK(misohip_set_bc_callback(ctx_, bc.fn ? &Boundary::call : nullptr,
                                           bc.fn ? bc.fn.get() : nullptr));
  }
  BoundaryScope(const BoundaryScope &) = delete;
  BoundaryScope &operator=(const BoundaryScope &) = delete;
  ~BoundaryScope() { misohip_set_bc_callback(ctx_, prev_fn_, prev_user_); }

 private:
  misohip_ctx *ctx_;
  misohip_bc_fn prev_fn_ = nullptr;
  void *prev_user_ = nullptr;
};

/// Source descriptors (the user Source functor's two forms here).
struct NoSource {
  void fill(misohip_params &p) const { p.src_kind = 0; p.grav_axis = 2; p.grav = 0.0; }
};
/// constant gravity: src.v[axis] = -g ro (demo/hd2d_rayleigh_taylor/src/main.cpp:109-112)
struct Gravity {
  int axis;
  double g;
  void fill(misohip_params &p) const { p.src_kind = 1; p.grav_axis = axis; p.grav = g; }
};

/// A user source term (Source::vx, vy, vz, ei; mhd_model_base.hpp:131-155):
/// a functor that writes the four terms of a stage's argm state into four
/// device arrays of the grid's shape, run before every RK stage through the
/// context's source callback (misohip_set_source_callback); the stage adds
/// them where the reference inlines the functor.  make_source
/// (miso_hip_device.hpp) builds one from a MISO-style Source struct.
template <typename Real> struct UserSource {
  using Fn = std::function<void(const misohip_fields &argm, Real *const out[4], hipStream_t)>;
  std::shared_ptr<Fn> fn;
  std::shared_ptr<std::array<DeviceArray<Real>, 4>> arrays;
  UserSource(const Grid<Real> &grid, Fn f)
      : fn(std::make_shared<Fn>(std::move(f))),
        arrays(std::make_shared<std::array<DeviceArray<Real>, 4>>()) {
    for (auto &a : *arrays) a = DeviceArray<Real>(grid.host().n_total());
  }
  void fill(misohip_params &p) const { p.src_kind = 2; p.grav_axis = 2; p.grav = 0.0; }
  /// C callback over fn; exceptions stop the step
  static int call(void *user, const misohip_fields *argm, void *const out[4], void *stream) {
    try {
      const UserSource *self = static_cast<const UserSource *>(user);
      Real *o[4] = {static_cast<Real *>(out[0]), static_cast<Real *>(out[1]),
                    static_cast<Real *>(out[2]), static_cast<Real *>(out[3])};
      (*self->fn)(*argm, o, static_cast<hipStream_t>(stream));
      return 0;
    } catch (...) {
      return 1;
    }
  }
};

/// Registers a user source on the context for one call and puts back the
/// callback that was there; the descriptor sources need nothing.
class SourceScope {
 public:
  SourceScope(misohip_ctx *, const NoSource &) {}
  SourceScope(misohip_ctx *, const Gravity &) {}
  template <typename Real> SourceScope(misohip_ctx *ctx, const UserSource<Real> &src) : ctx_(ctx) {
    MISO_HIP_CHECK(misohip_get_source_callback(ctx_, &prev_fn_, &prev_user_, prev_arr_));
    void *arr[4];
    for (int d = 0; d < 4; ++d) arr[d] = (*src.arrays)[d].data();
    MISO_HIP_CHECK(misohip_set_source_callback(ctx_, &UserSource<Real>::call,
                                               const_cast<UserSource<Real> *>(&src), arr));
  }
  SourceScope(const SourceScope &) = delete;
  SourceScope &operator=(const SourceScope &) = delete;
  ~SourceScope() {
    if (ctx_) misohip_set_source_callback(ctx_, prev_fn_, prev_user_, prev_arr_);
  }

 private:
  misohip_ctx *ctx_ = nullptr;
  misohip_src_fn prev_fn_ = nullptr;
  void *prev_user_ = nullptr;
  void *prev_arr_[4] = {nullptr, nullptr, nullptr, nullptr};
};

/// The ideal-gas EOS the step uses (eos::IdealEOS, eos.hpp:11-43).
struct IdealEOS {
  double gm;
};

/// Config keys the step reads (config_defaults.hpp:4-32).
struct MHDConfig {
  double cfl_number = 0.5;
  double ep = 1.0, fh = 1.0, cs_fac = 1.0, ca_fac = 1.0, vv_fac = 1.0;
};

/// Integrator: cfl, update, apply_boundary_condition
/// (mhd_integrator_host.hpp:13-357).  Owns qq_argm, qq_rslt and cc.
template <typename Real> class Integrator {
 public:
  Integrator(const MHDConfig &cfg, const Grid<Real> &grid, ExecContext<Real> &ctx)
      : cfg_(cfg), grid_(grid), ctx_(ctx), argm_(grid), rslt_(grid), cc_(grid.host().n_total()) {
    if (cfg.ep < 0 || cfg.fh < 0)
      throw std::runtime_error("artificial viscosity ep and fh must be >= 0");
    w_.argm = *argm_.desc();
    w_.rslt = *rslt_.desc();
    w_.cc = cc_.data();
  }

  template <typename BC> void apply_boundary_condition(const BC &bc, Fields<Real> &qq) {
    BoundaryScope scope(ctx_.handle(), bc);
    MISO_HIP_CHECK(misohip_ghost_fill(ctx_.handle(), grid_.desc(), dtype_of<Real>(), qq.desc(),
                                      &bc.d));
  }

  template <typename EOS> Real cfl(const Fields<Real> &qq, const EOS &eos) {
    misohip_params p = params(eos, NoSource{});
    double dt = 0.0;
    MISO_HIP_CHECK(misohip_cfl(ctx_.handle(), grid_.desc(), dtype_of<Real>(), &p, qq.desc(), &dt));
    return Real(dt);
  }

  template <typename EOS, typename BC, typename Source>
  void update(Real dt, const EOS &eos, const BC &bc, const Source &src, Fields<Real> &qq) {
    misohip_params p = params(eos, src);
    BoundaryScope scope(ctx_.handle(), bc);
    SourceScope sscope(ctx_.handle(), src);
    MISO_HIP_CHECK(misohip_update(ctx_.handle(), grid_.desc(), dtype_of<Real>(), &p, &bc.d,
                                  qq.desc(), &w_, double(dt)));
  }

  /// update, the radiative heating ei += dt * qrad / ro (qrad: a device
  /// field of the grid) and the ghost fill; returns the next step's dt (the
  /// CFL of the heated state).  FP64; with x and y active the heating and
  /// the CFL reduction run inside the last AV sweep.
  template <typename EOS, typename BC, typename Source>
  double update_heated(Real dt, const EOS &eos, const BC &bc, const Source &src, Fields<Real> &qq,
                       const void *qrad) {
    static_assert(std::is_same<Real, double>::value, "the radiative heating is FP64");
    misohip_params p = params(eos, src);
    double next = 0.0;
    BoundaryScope scope(ctx_.handle(), bc);
    SourceScope sscope(ctx_.handle(), src);
    MISO_HIP_CHECK(misohip_update_heated(ctx_.handle(), grid_.desc(), dtype_of<Real>(), &p,
                                         &bc.d, qq.desc(), &w_, double(dt), qrad, &next));
    return next;
  }

  /// cfl + update in one call with one host synchronisation
  /// (ModelBase::update, mhd_model_base.hpp:57-62); returns dt.
  template <typename EOS, typename BC, typename Source>
  Real step(const EOS &eos, const BC &bc, const Source &src, Fields<Real> &qq) {
    misohip_params p = params(eos, src);
    double dt = 0.0;
    BoundaryScope scope(ctx_.handle(), bc);
    SourceScope sscope(ctx_.handle(), src);
    MISO_HIP_CHECK(misohip_step(ctx_.handle(), grid_.desc(), dtype_of<Real>(), &p, &bc.d,
                                qq.desc(), &w_, &dt));
    return Real(dt);
  }

  /// nsteps of step() in one call (the loop of ModelBase::run,
  /// mhd_model_base.hpp:89-117); the dt of every step, the same values.
  template <typename EOS, typename BC, typename Source>
  std::vector<double> run(const EOS &eos, const BC &bc, const Source &src, Fields<Real> &qq,
                          int nsteps) {
    misohip_params p = params(eos, src);
    std::vector<double> dts(nsteps > 0 ? nsteps : 0);
    BoundaryScope scope(ctx_.handle(), bc);
    SourceScope sscope(ctx_.handle(), src);
    MISO_HIP_CHECK(misohip_run(ctx_.handle(), grid_.desc(), dtype_of<Real>(), &p, &bc.d,
                               qq.desc(), &w_, nsteps, dts.data()));
    return dts;
  }

 private:
  template <typename EOS, typename Source> misohip_params params(const EOS &eos, const Source &src) {
    misohip_params p{};
    p.gm = eos.gm;
    p.cfl_number = cfg_.cfl_number;
    p.ep = cfg_.ep; p.fh = cfg_.fh;
    p.cs_fac = cfg_.cs_fac; p.ca_fac = cfg_.ca_fac; p.vv_fac = cfg_.vv_fac;
    src.fill(p);
    return p;
  }
  MHDConfig cfg_;
  const Grid<Real> &grid_;
  ExecContext<Real> &ctx_;
  Fields<Real> argm_, rslt_;
  DeviceArray<Real> cc_;
  misohip_work w_{};
};

/// mhd::MHD: grid, fields and integrator (mhd.hpp:15-50).
template <typename Real> struct MHD {
  Grid<Real> grid;
  Fields<Real> qq;
  ExecContext<Real> &exec_ctx;
  Integrator<Real> integrator;

  MHD(const MHDConfig &cfg, const HostGrid<Real> &grid_h, ExecContext<Real> &ctx)
      : grid(grid_h), qq(grid), exec_ctx(ctx), integrator(cfg, grid, ctx) {}

  /// ic(HostFields&, const HostGrid&) fills the host fields, then H2D + BC
  template <typename IC, typename BC> void apply_initial_condition(const IC &ic, const BC &bc) {
    const HostGrid<Real> &h = grid.host();
    HostFields<Real> qq_h(h.i_total(), h.j_total(), h.k_total());
    ic(qq_h, h);
    qq.copy_from(exec_ctx, qq_h);
    integrator.apply_boundary_condition(bc, qq);
  }
  template <typename EOS> Real cfl(const EOS &eos) { return integrator.cfl(qq, eos); }
  template <typename EOS, typename BC, typename Source>
  void update(Real dt, const EOS &eos, const BC &bc, const Source &src) {
    integrator.update(dt, eos, bc, src, qq);
  }
  /// nsteps of cfl + update (ModelBase::run's loop); the dt of every step
  template <typename EOS, typename BC, typename Source>
  std::vector<double> run(const EOS &eos, const BC &bc, const Source &src, int nsteps) {
    return integrator.run(eos, bc, src, qq, nsteps);
  }
};

/// The RT boundary as descriptors (the reference passes a functor that calls
/// set_incoming_boundary_on_physical_faces, rt_host.hpp:50-101): per axis
/// periodic (the halo self-exchange, rt_host.hpp:229-452), otherwise per
/// face (x inner, x outer, y inner, ...) untouched, thermal I = S, vacuum
/// I = 0 or a constant value on the incoming rays.
struct RtBoundary {
  enum Kind { UNTOUCHED = 0, THERMAL = 1, VACUUM = 2, VALUE = 3 };
  std::array<int, 3> periodic{{1, 1, 0}};
  std::array<int, 6> face{{UNTOUCHED, UNTOUCHED, UNTOUCHED, UNTOUCHED, THERMAL, VACUUM}};
  std::array<double, 6> value{{0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
};

/// rt::RT<double> on the device: rint and rint_old (num_rays fields of the
/// grid), src_func and abs_coeff (x-fastest fields), the angular quadrature
/// on the host and the lattice bounds ib0..kb1 (rt_host.hpp:127-154).  With
/// x_rays_yfast (default; fixed for the object's life) the rays whose
/// dominant axis is x keep their field of rint y-fastest; intensity(),
/// save() and load() move any ray in MISO's order.
class RT {
 public:
  RT(ExecContext<double> &ctx, const Grid<double> &grid, int num_rays, bool x_rays_yfast = true)
      : ctx_(&ctx), grid_(&grid), n_(grid.host().n_total()), yfast_(x_rays_yfast) {
    weights.resize(24);
    mu_x.resize(24);
    mu_y.resize(24);
    mu_z.resize(24);
    const int n = misohip_rt_quadrature(num_rays, weights.data(), mu_x.data(), mu_y.data(),
                                        mu_z.data());
    if (n < 1) MISO_HIP_CHECK(n);
    this->num_rays = n;
    for (auto *v : {&weights, &mu_x, &mu_y, &mu_z}) v->resize(n);
    rint = DeviceArray<double>(n_ * n);
    rint_old = DeviceArray<double>(n_ * n);
    src_func = DeviceArray<double>(n_);
    abs_coeff = DeviceArray<double>(n_);
    const HostGrid<double> &h = grid.host();
    const int m[3] = {h.i_margin(), h.j_margin(), h.k_margin()};
    const int st[3] = {h.xg.stride, h.yg.stride, h.zg.stride};
    const int sz[3] = {h.i_size(), h.j_size(), h.k_size()};
    int b0[3], b1[3];
    for (int a = 0; a < 3; ++a) {
      b0[a] = m[a] - st[a];
      b1[a] = b0[a] + (sz[a] > 1 ? sz[a] : 0);
    }
    ib0 = b0[0]; ib1 = b1[0]; jb0 = b0[1]; jb1 = b1[1]; kb0 = b0[2]; kb1 = b1[2];
    const misohip_rt_config c = config(0.0, 1, RtBoundary{});
    MISO_HIP_CHECK(misohip_rt_ray_layout(ctx_->handle(), grid_->desc(), &c, &yfast_mask_));
  }

  /// kappa and S from host fields in MISO order ((i * Jt + j) * Kt + k)
  void set_fields(const std::vector<double> &kappa, const std::vector<double> &S) {
    if (kappa.size() != n_ || S.size() != n_) throw std::runtime_error("RT field size");
    MISO_HIP_CHECK(misohip_copy_to_device(ctx_->handle(), grid_->desc(), MISOHIP_F64,
                                          kappa.data(), abs_coeff.data()));
    MISO_HIP_CHECK(misohip_copy_to_device(ctx_->handle(), grid_->desc(), MISOHIP_F64, S.data(),
                                          src_func.data()));
  }

  /// RT::solve (rt_host.hpp:632-651): iterate until the float-narrowed max
  /// change is below tolerance; returns the iteration count and prints the
  /// reference's message when max_iters are used up without converging.
  int solve(double tolerance, int max_iters, const RtBoundary &bc) {
    const misohip_rt_config c = config(tolerance, max_iters, bc);
    int it = 0;
    double md = 0.0;
    MISO_HIP_CHECK(misohip_rt_solve(ctx_->handle(), grid_->desc(), &c, abs_coeff.data(),
                                    src_func.data(), rint.data(), rint_old.data(), &it, &md));
    last_diff = md;
    if (!(md < tolerance)) std::printf("  RT did not converge in %d iterations.\n", max_iters);
    return it;
  }

  /// Q_rad = 4 pi kappa (J - S) on the lattice into a device field of the grid
  void qrad(DeviceArray<double> &out) const {
    if (out.size() < n_) throw std::runtime_error("RT qrad field size");
    const misohip_rt_config c = config(0.0, 1, RtBoundary{});
    MISO_HIP_CHECK(misohip_rt_qrad(ctx_->handle(), grid_->desc(), &c, abs_coeff.data(),
                                   src_func.data(), rint.data(), out.data()));
  }

  /// Ray n's intensities in MISO order ((i * Jt + j) * Kt + k), host copy
  std::vector<double> intensity(int n) const {
    std::vector<double> dev(n_), out(n_);
    ctx_->synchronize();
    MISO_HIP_RT_CHECK(hipMemcpy(dev.data(), rint.data() + n_ * std::size_t(n),
                                sizeof(double) * n_, hipMemcpyDeviceToHost));
    reorder(n, dev.data(), out.data(), true);
    return out;
  }
  /// Ray n's intensities from MISO order into the solver's layout
  void set_intensity(int n, const std::vector<double> &miso) {
    if (miso.size() != n_) throw std::runtime_error("RT intensity size");
    std::vector<double> dev(n_);
    reorder(n, miso.data(), dev.data(), false);
    MISO_HIP_RT_CHECK(hipMemcpy(rint.data() + n_ * std::size_t(n), dev.data(),
                                sizeof(double) * n_, hipMemcpyHostToDevice));
  }

  /// RT::save (rt_host.hpp:173-199): int ray count, weights, mu_x, mu_y,
  /// mu_z, then src_func, abs_coeff and rint, all in MISO order.
  void save(const std::string &file_path) const {
    std::FILE *f = std::fopen(file_path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot open " + file_path);
    std::fwrite(&num_rays, sizeof(int), 1, f);
    for (const auto *v : {&weights, &mu_x, &mu_y, &mu_z})
      std::fwrite(v->data(), sizeof(double), v->size(), f);
    for (const auto *a : {&src_func, &abs_coeff}) {
      std::vector<double> h(n_);
      MISO_HIP_CHECK(misohip_copy_to_host(ctx_->handle(), grid_->desc(), MISOHIP_F64, a->data(),
                                          h.data()));
      ctx_->synchronize();
      std::fwrite(h.data(), sizeof(double), n_, f);
    }
    for (int n = 0; n < num_rays; ++n) {
      const std::vector<double> h = intensity(n);
      std::fwrite(h.data(), sizeof(double), n_, f);
    }
    std::fclose(f);
  }
  /// RT::load (rt_host.hpp:201-227): the ray count must match.
  void load(const std::string &file_path) {
    std::FILE *f = std::fopen(file_path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + file_path);
    auto rd = [&](void *p, std::size_t sz, std::size_t cnt) {
      if (std::fread(p, sz, cnt, f) != cnt) {
        std::fclose(f);
        throw std::runtime_error("short RT file " + file_path);
      }
    };
    int n = 0;
    rd(&n, sizeof(int), 1);
    if (n != num_rays) {
      std::fclose(f);
      throw std::runtime_error("RT file ray count does not match");
    }
    for (auto *v : {&weights, &mu_x, &mu_y, &mu_z}) rd(v->data(), sizeof(double), v->size());
    std::vector<double> s(n_), k(n_), I(n_);
    rd(s.data(), sizeof(double), n_);
    rd(k.data(), sizeof(double), n_);
    set_fields(k, s);
    for (int r = 0; r < num_rays; ++r) {
      rd(I.data(), sizeof(double), n_);
      set_intensity(r, I);
    }
    std::fclose(f);
  }

  DeviceArray<double> rint, rint_old, src_func, abs_coeff;
  std::vector<double> weights, mu_x, mu_y, mu_z;
  int num_rays = 0;
  int ib0 = 0, ib1 = 0, jb0 = 0, jb1 = 0, kb0 = 0, kb1 = 0;
  int sweep_planes = 16;
  double last_diff = 0.0;

 private:
  // ray n from the solver's layout to MISO order (to_miso), or back
  void reorder(int n, const double *src, double *dst, bool to_miso) const {
    const HostGrid<double> &h = grid_->host();
    const std::size_t it = h.i_total(), jt = h.j_total(), kt = h.k_total();
    const bool yf = (yfast_mask_ >> n) & 1u;
    for (std::size_t i = 0; i < it; ++i)
      for (std::size_t j = 0; j < jt; ++j)
        for (std::size_t k = 0; k < kt; ++k) {
          const std::size_t d = yf ? (k * it + i) * jt + j : (k * jt + j) * it + i;
          const std::size_t m = (i * jt + j) * kt + k;
          if (to_miso) dst[m] = src[d];
          else dst[d] = src[m];
        }
  }
  misohip_rt_config config(double tolerance, int max_iters, const RtBoundary &bc) const {
    misohip_rt_config c{};
    c.num_rays = num_rays;
    for (int a = 0; a < 3; ++a) c.periodic[a] = bc.periodic[a];
    for (int f = 0; f < 6; ++f) {
      c.face_kind[f] = bc.face[f];
      c.face_value[f] = bc.value[f];
    }
    c.tolerance = tolerance;
    c.max_iters = max_iters;
    c.x_rays_yfast = yfast_ ? 1 : 0;
    c.sweep_planes = sweep_planes;
    return c;
  }
  ExecContext<double> *ctx_;
  const Grid<double> *grid_;
  std::size_t n_;
  bool yfast_;
  unsigned yfast_mask_ = 0u;
};

}  // namespace hip
}  // namespace miso
