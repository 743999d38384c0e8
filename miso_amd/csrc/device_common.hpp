// Device-side building blocks shared by the MHD kernels (gfx950).
//
// Arithmetic helpers reproduce MISO's host semantics bit for bit in FP64:
// min/max narrow their operands to float (miso/include/miso/utility.hpp:91-124),
// constants are rounded to Real once (miso/include/miso/constants.hpp:7-16).
// The library is compiled with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace misohip {

template <typename R> struct Const {
  static constexpr R inv12 = static_cast<R>(1.0 / 12.0);
  static constexpr R pii4 = static_cast<R>(1.0 / (4.0 * 3.14159265358979323846));
  static constexpr R pii8 = static_cast<R>(1.0 / (8.0 * 3.14159265358979323846));
};

// float-narrowing min/max, as the reference's util::min2/max2/min3/max3
template <typename R> __device__ __forceinline__ R nmin2(R a, R b) {
  return static_cast<R>(fminf(static_cast<float>(a), static_cast<float>(b)));
}
template <typename R> __device__ __forceinline__ R nmax2(R a, R b) {
  return static_cast<R>(fmaxf(static_cast<float>(a), static_cast<float>(b)));
}
template <typename R> __device__ __forceinline__ R nmin3(R a, R b, R c) {
  return static_cast<R>(
      fminf(fminf(static_cast<float>(a), static_cast<float>(b)), static_cast<float>(c)));
}
template <typename R> __device__ __forceinline__ R nmax3(R a, R b, R c) {
  return static_cast<R>(
      fmaxf(fmaxf(static_cast<float>(a), static_cast<float>(b)), static_cast<float>(c)));
}

__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double dfabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float dfabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double dcopysign(double x, double y) { return __builtin_copysign(x, y); }
__device__ __forceinline__ float dcopysign(float x, float y) { return __builtin_copysignf(x, y); }
// a*b + c rounded once, in the storage type (__builtin_fma alone would take
// float operands to double: FP64-rate instructions and two conversions)
__device__ __forceinline__ double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float dfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// x in a register of its own: the register a load wrote is free from here
// on, so a loop can load the next value into it without keeping the old one
// alive (a register copy at the loop's end would have to wait for the load)
__device__ __forceinline__ double reg_copy(double x) {
  double y;
  asm volatile("v_mov_b64 %0, %1" : "=v"(y) : "v"(x));
  return y;
}
__device__ __forceinline__ float reg_copy(float x) {
  float y;
  asm volatile("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
  return y;
}

// Whole-wave shifts by one lane as DPP moves (VALU, no LDS round trip):
// wave_shr:1 (0x138) gives lane i the value of lane i-1, wave_shl:1 (0x130)
// that of lane i+1; the end lanes, which no caller stores, get 0.
template <int CTRL> __device__ __forceinline__ int dpp_move(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(dpp_move<CTRL>(__float_as_int(v)));
}
template <int CTRL> __device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_move<CTRL>((int)(unsigned)(b & 0xffffffffLL));
  const unsigned hi = (unsigned)dpp_move<CTRL>((int)(unsigned)((unsigned long long)b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename R>
__device__ __forceinline__ R lane_up(R v) {  // value of lane - 1
  return dpp_move<0x138>(v);
}
template <typename R>
__device__ __forceinline__ R lane_dn(R v) {  // value of lane + 1
  return dpp_move<0x130>(v);
}

// Workgroups are dispatched round-robin over the 8 XCDs (linear block L runs
// on XCD L % 8, and each XCD takes its blocks in increasing L).  Mapping L to
// the virtual block V below gives every XCD one contiguous range of V, i.e. a
// contiguous slab of the grid, so stencil halos are re-read from its own L2.
__device__ __forceinline__ int xcd_virtual_block(int L, int nb) {
  const int x = L & 7, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + (L >> 3);
}

// Pencil order of a 3D grid of blocks (nbx, nby, nbz): the x-y block plane
// is cut into pencils of pbx x pby blocks; inside a pencil the block plane is
// fastest and z next, so the z-neighbour planes of a block were touched a few
// hundred blocks earlier on the same XCD and are still in its L2.  Combined
// with xcd_virtual_block every XCD sweeps whole pencils.
struct Pencils {
  int nbx, nby, nbz, pbx, pby, npx, nb;
  __host__ bool init(int nbx_, int nby_, int nbz_, int pbx_, int pby_) {
    nbx = nbx_; nby = nby_; nbz = nbz_;
    pbx = pbx_ < 1 ? 1 : (pbx_ > nbx ? nbx : pbx_);
    pby = pby_ < 1 ? 1 : (pby_ > nby ? nby : pby_);
    npx = (nbx + pbx - 1) / pbx;
    const int npy = (nby + pby - 1) / pby;
    const long n = (long)npx * npy * pbx * pby * nbz;
    if (n > 2147483647L || n <= 0) return false;
    nb = (int)n;
    return true;
  }
  // false for the padding blocks of partial edge pencils
  __device__ bool decode(int V, int &bx, int &by, int &bz) const {
    const int per = pbx * pby;
    const int in = V % per, r = V / per;
    bz = r % nbz;
    const int pid = r / nbz;
    bx = (pid % npx) * pbx + in % pbx;
    by = (pid / npx) * pby + in / pbx;
    return bx < nbx && by < nby;
  }
};

// kind: 0 mirror, 1 periodic, 2 fixed (the ghost cell keeps its coordinate
// and takes its value from the fixed fields)
__device__ __forceinline__ int map_coord(int x, int total, int m, int kind, bool &mirrored) {
  mirrored = false;
  if (m == 0 || kind >= 2) return x;  // fixed (2) and user (3) axes keep their ghosts
  const int n = total - 2 * m;
  if (x < m) {
    if (kind == 1) return x + n;
    mirrored = true;
    return 2 * m - 1 - x;
  }
  if (x >= total - m) {
    if (kind == 1) return x - n;
    mirrored = true;
    return 2 * (total - m) - 1 - x;
  }
  return x;
}

// Periodic source coordinate of x on an axis of `total` cells with margin m
// when `on` (a ghost coordinate maps to the interior cell that the periodic
// ghost fill copies into it); x itself otherwise.
__device__ __forceinline__ int wrap_coord(int x, int total, int m, bool on) {
  if (!on) return x;
  const int n = total - 2 * m;
  return x < m ? x + n : (x >= total - m ? x - n : x);
}

// Blocks of the z-march kernel: (nbx, nby) columns of 64 x 4 cells, each cut
// into nzc chunks of kz planes; x fastest, then y, then the chunk, so that
// with xcd_virtual_block every XCD walks compact x-y regions of one slab.
// The requested chunk is halved (not below 4 planes) while there would be
// fewer blocks than the 256 CUs (a small grid in long chunks would leave
// most of the chip idle); every chunk loads its own first planes, so the
// results do not depend on the split.
struct March {
  int nbx, nby, nzc, kz, nb;
  __host__ bool init(int nbx_, int nby_, int nk, int kz_) {
    if (kz_ < 1 || nk < 1) return false;
    nbx = nbx_; nby = nby_;
    while (kz_ > 4 && (long)nbx * nby * ((nk + kz_ - 1) / kz_) < 256)
      kz_ = (kz_ + 1) / 2 < 4 ? 4 : (kz_ + 1) / 2;
    nzc = (nk + kz_ - 1) / kz_;
    kz = (nk + nzc - 1) / nzc;  // the chunks as even as the count allows
    const long n = (long)nbx * nby * nzc;
    if (n > 2147483647L || n <= 0) return false;
    nb = (int)n;
    return true;
  }
};

// Field set: nine device arrays in MISO's order ro vx vy vz bx by bz ei ph.
template <typename T> struct Fld {
  T *q[9];
};

// Shape of one rank's grid; device arrays are x-fastest:
// (i, j, k) -> (k * Jt + j) * It + i.
struct Shape {
  int It, Jt, Kt;
  int mi, mj, mk;
  int is, js, ks;
  __host__ __device__ int idx(int i, int j, int k) const { return (k * Jt + j) * It + i; }
  __host__ __device__ int stride(int d) const {
    return d == 0 ? is : d == 1 ? js * It : ks * It * Jt;
  }
};

// Generalised minmod slope, limiter.hpp:10-22.  The two terms of the final
// sum, min(0, max) and max(0, min) with min <= max, are float-narrowed and
// at least one of them is a zero, so their sum is exact in float: it is
// formed there and widened once (the same value, zeros' signs included, as
// the sum of the two widened terms).
template <typename R>
__device__ __forceinline__ R dqq_eval(R qq_dw, R qq_cn, R qq_up, R ep) {
  const R dqq_dw = qq_cn - qq_dw;
  const R dqq_up = qq_up - qq_cn;
  const R dqq_cn = (qq_up - qq_dw) * R(0.5);
  const float a = static_cast<float>(ep * dqq_dw), b = static_cast<float>(ep * dqq_up),
              c = static_cast<float>(dqq_cn);
  const float dqq_max = fmaxf(fmaxf(a, b), c), dqq_min = fminf(fminf(a, b), c);
  return static_cast<R>(fminf(0.0f, dqq_max) + fmaxf(0.0f, dqq_min));
}

// Limited interface flux, limiter.hpp:24-37.
template <typename R>
__device__ __forceinline__ R flux_core(R qq_dw, R qq_up, R dqq_dw, R dqq_up, R cc, R fh) {
  // 0.5*d is exact, so q + 0.5*d is one correctly rounded fma
  const R qql = dfma(R(0.5), dqq_dw, qq_dw);
  const R qqr = dfma(R(-0.5), dqq_up, qq_up);
  R dqq = qq_up - qq_dw;
  dqq = dcopysign(nmax2(dfabs(dqq), R(1e-20)), dqq);
  const R ra = nmin2(R(1.0), (qqr - qql) / dqq);
  const R pp = (R(0.5) + dcopysign(R(0.5), ra)) * nmax2(R(0.0), R(1.0) + fh * (ra - R(1.0)));
  return R(-0.5) * cc * pp * (qqr - qql);
}

// CFL candidate of one cell (Integrator::cfl, mhd_integrator_host.hpp:303-334):
// cfl * min(dx, dy, dz) / (cs + |v| + ca), narrowed to float.
template <typename R>
__device__ __forceinline__ float cfl_candidate_v(R ro, R vx, R vy, R vz, R bx, R by, R bz, R ei,
                                                 R gm, R cfl, R hmin) {
  const R cs = dsqrt(gm * (gm - R(1)) * ei);
  const R vv = dsqrt(+vx * vx + vy * vy + vz * vz);
  const R ca = dsqrt((+bx * bx + by * by + bz * bz) / ro * Const<R>::pii4);
  const R total_vel = cs + vv + ca;
  return static_cast<float>(cfl * hmin / total_vel);
}

// minimum over the 64 lanes of a wave (all lanes must be active)
__device__ __forceinline__ unsigned wave_umin(unsigned v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = min(v, (unsigned)__shfl_xor((int)v, o, 64));
  return v;
}

}  // namespace misohip
