// One Runge-Kutta stage of MISO's 4th-order central-difference MHD scheme
// (Integrator::update_sc4, miso/include/miso/mhd_integrator_host.hpp:57-247;
// the CUDA backend splits it over ten kernels, mhd_integrator_cuda.cuh:16-289).
//
// One fused kernel computes all nine fields of a cell.  Lanes run along x,
// the contiguous device axis, so every load of a wave is one coalesced
// segment; the y/z neighbours of a row are re-read from L1/L2, and the
// XCD-aware block order keeps each XCD on one contiguous z-slab.
//
// Bit parity with the reference CPU path: each stencil evaluates
//   ((((-q[+2]) + 8 q[+1]) - 8 q[-1]) + q[-2]) * inv12 * dxi
// on the per-point products (a*b)*c exactly as space_centered_4th does
// (mhd_integrator.hpp:14-57): since -a and 8a are exact, (-a)*b*c and
// (8a)*b*c equal -(a*b*c) and 8(a*b*c) bit for bit, so forming the product
// once per point keeps every rounding of the reference; for the same reason
// acc + 8 t is one fma(8, t, acc).  The equations' sums keep the reference's
// term order; they are accumulated direction by direction (x, then y, then
// z) as running partial sums, which is the same order of operations.
#include "launch.hpp"

namespace misohip {

// Quantities differentiated along direction D at one point (17 of them):
//  0 ro*v_D   1 (ro*vx)*v_D  2 (ro*vy)*v_D  3 (ro*vz)*v_D   4 pr   5 bb
//  6 bx*b_D   7 by*b_D       8 bz*b_D       9..12 induction products
// 13 ph      14 b_D         15 ht*v_D      16 vb*b_D
// Induction products (the a*b operands of the induction terms along D):
//  D=x: vx*by vy*bx vx*bz vz*bx;  D=y: vy*bx vx*by vy*bz vz*by;
//  D=z: vz*bx vx*bz vz*by vy*bz.
// The seven products of a cell that every direction's terms use:
//  0..2 ro*vx, ro*vy, ro*vz   3 pr   4 bb   5 ht   6 vb
// (eos.hpp:30 for pr = (gm - 1) * ro * ei; mhd_integrator_host.hpp:69-89).
template <typename R>
__device__ __forceinline__ void cell_products(const R f[9], R gm1, R p[7]) {
  const R ro = f[0], vx = f[1], vy = f[2], vz = f[3];
  const R bx = f[4], by = f[5], bz = f[6], ei = f[7];
  p[0] = ro * vx;
  p[1] = ro * vy;
  p[2] = ro * vz;
  p[3] = gm1 * ro * ei;
  p[4] = bx * bx + by * by + bz * bz;
  p[5] = ro * ei + p[3] + p[4] * Const<R>::pii4 + R(0.5) * ro * (vx * vx + vy * vy + vz * vz);
  p[6] = vx * bx + vy * by + vz * bz;
}

// The 17 terms along D from a cell's fields (ro and ei are not read) and
// its products.
template <typename R, int D>
__device__ __forceinline__ void point_terms_p(const R f[9], const R p[7], R o[17]) {
  const R vx = f[1], vy = f[2], vz = f[3];
  const R bx = f[4], by = f[5], bz = f[6], ph = f[8];
  const R vd = D == 0 ? vx : (D == 1 ? vy : vz);
  const R bd = D == 0 ? bx : (D == 1 ? by : bz);
  o[0] = p[D];
  o[1] = p[0] * vd;
  o[2] = p[1] * vd;
  o[3] = p[2] * vd;
  o[4] = p[3];
  o[5] = p[4];
  o[6] = bx * bd;
  o[7] = by * bd;
  o[8] = bz * bd;
  if (D == 0) {
    o[9] = vx * by; o[10] = vy * bx; o[11] = vx * bz; o[12] = vz * bx;
  } else if (D == 1) {
    o[9] = vy * bx; o[10] = vx * by; o[11] = vy * bz; o[12] = vz * by;
  } else {
    o[9] = vz * bx; o[10] = vx * bz; o[11] = vz * by; o[12] = vy * bz;
  }
  o[13] = ph;
  o[14] = bd;
  o[15] = p[5] * vd;
  o[16] = p[6] * bd;
}

template <typename R, int D>
__device__ __forceinline__ void point_terms(const R f[9], R gm1, R o[17]) {
  R p[7];
  cell_products(f, gm1, p);
  point_terms_p<R, D>(f, p, o);
}

// Register fence: the 17 running sums are materialised and `idx` becomes
// opaque, so loads addressed from it cannot be hoisted above this point.
template <typename R> __device__ __forceinline__ void fence17(int &idx, R S[17]) {
  asm volatile("" : "+v"(idx), "+v"(S[0]), "+v"(S[1]), "+v"(S[2]), "+v"(S[3]), "+v"(S[4]),
               "+v"(S[5]), "+v"(S[6]), "+v"(S[7]), "+v"(S[8]), "+v"(S[9]), "+v"(S[10]),
               "+v"(S[11]), "+v"(S[12]), "+v"(S[13]), "+v"(S[14]), "+v"(S[15]), "+v"(S[16]));
}

// All 17 derivatives along D at linear index c (stride s, inverse spacing dd).
// SPLIT: fence after each stencil point (9 loads in flight per wave instead
// of 36, fewer registers, more waves per SIMD); the arithmetic is the same.
// Field accessors: global memory (the nine arrays) or an LDS tile holding
// the nine fields of a block's x-y neighbourhood.
template <typename R> struct GlobalAcc {
  const Fld<const R> &a;
  __device__ __forceinline__ R operator()(int m, int i) const { return a.q[m][i]; }
};
constexpr int RK_TX = 68, RK_TY = 8, RK_TILE = RK_TX * RK_TY;  // 64 x 4 cells + 2-cell halo
template <typename R, int TSIZE = RK_TILE> struct TileAcc {
  const R *t;
  __device__ __forceinline__ R operator()(int m, int i) const { return t[m * TSIZE + i]; }
};

template <typename R, int D, int SPLIT, typename Acc>
__device__ __forceinline__ void dir_derivs(const Acc &a, int c, int s, R dd, R gm1, R S[17]) {
  R f[9], t[17];
#pragma unroll
  for (int m = 0; m < 9; ++m) f[m] = a(m, c + 2 * s);
  point_terms<R, D>(f, gm1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = -t[n];
  if (SPLIT) fence17(c, S);
#pragma unroll
  for (int m = 0; m < 9; ++m) f[m] = a(m, c + s);
  point_terms<R, D>(f, gm1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = dfma(R(8), t[n], S[n]);
  if (SPLIT) fence17(c, S);
#pragma unroll
  for (int m = 0; m < 9; ++m) f[m] = a(m, c - s);
  point_terms<R, D>(f, gm1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = dfma(R(-8), t[n], S[n]);
  if (SPLIT) fence17(c, S);
#pragma unroll
  for (int m = 0; m < 9; ++m) f[m] = a(m, c - 2 * s);
  point_terms<R, D>(f, gm1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = (S[n] + t[n]) * Const<R>::inv12 * dd;
}

template <typename R, int D, bool ACTIVE, int SPLIT = 0, typename Acc = GlobalAcc<R>>
__device__ __forceinline__ void derivs_or_zero(const Acc &a, int c, int s, R dd, R gm1,
                                               R S[17]) {
  if (ACTIVE) {
    dir_derivs<R, D, SPLIT>(a, c, s, dd, gm1, S);
  } else {
#pragma unroll
    for (int n = 0; n < 17; ++n) S[n] = R(0);
  }
}

// The stage update of one cell from its direction sums (the tail of
// update_sc4, mhd_integrator_host.hpp:113-243); o: orgn's nine fields at the
// cell; gro: argm's density there (gravity source).
template <typename R>
__device__ __forceinline__ void rk_update(const R o[9], const Fld<R> &rslt, int c, R gro,
                                          const RkScalars &sc, const R dt, const R ch2,
                                          const R damp, R a_ro, R a_vx, R d4x, R d5x, R t_vx,
                                          R a_vy, R d4y, R d5y, R t_vy, R a_vz, R d4z, R d5z,
                                          R t_vz, R a_bx, R a_by, R a_bz, R a_ph, R a_ei,
                                          R t_ei) {
  const R pii4 = Const<R>::pii4, pii8 = Const<R>::pii8;
  const R o_ro = o[0], o_vx = o[1], o_vy = o[2], o_vz = o[3];
  const R o_bx = o[4], o_by = o[5], o_bz = o[6];
  const R o_ei = o[7], o_ph = o[8];

  // source functor (gravity reads argm's density, mhd_integrator_host.hpp:134;
  // a user source arrives as the four arrays its callback wrote)
  R src[3] = {R(0), R(0), R(0)};
  R src_ei = R(0);
  if (sc.src_kind == 1) src[sc.grav_axis] = -R(sc.grav) * gro;
  if (sc.src_kind == 2) {
#pragma unroll
    for (int d = 0; d < 3; ++d) src[d] = static_cast<const R *>(sc.src_arr[d])[c];
    src_ei = static_cast<const R *>(sc.src_arr[3])[c];
  }

  // continuity
  const R ro = o_ro + dt * a_ro;
  // equations of motion: (ro v)_0 + dt*(flux - grad p - grad bb/8pi + tension/4pi + src)
  const R vx = (o_ro * o_vx + dt * (a_vx - d4x - d5x + pii4 * t_vx + src[0])) / ro;
  const R vy = (o_ro * o_vy + dt * (a_vy - d4y - d5y + pii4 * t_vy + src[1])) / ro;
  const R vz = (o_ro * o_vz + dt * (a_vz - d4z - d5z + pii4 * t_vz + src[2])) / ro;
  // induction
  const R bx = o_bx + dt * a_bx;
  const R by = o_by + dt * a_by;
  const R bz = o_bz + dt * a_bz;
  // GLM divergence cleaning; damp = exp(-dt/tau) from the host
  const R ph = (o_ph + dt * ch2 * a_ph) * damp;
  // energy: total energy of orgn, flux divergence, recovery of ei
  const double eo_ro = o_ro;
  const double Et = +eo_ro * double(o_ei) +
                    0.5 * eo_ro * (+double(o_vx) * o_vx + double(o_vy) * o_vy + double(o_vz) * o_vz) +
                    double(pii8) * (+double(o_bx) * o_bx + double(o_by) * o_by + double(o_bz) * o_bz);
  const R eflux = dt * (a_ei + pii4 * t_ei);
  const double dro = ro, dvx = vx, dvy = vy, dvz = vz, dbx = bx, dby = by, dbz = bz;
  const double ei = (Et + double(eflux) - 0.5 * dro * (+dvx * dvx + dvy * dvy + dvz * dvz) -
                     double(pii8) * (+dbx * dbx + dby * dby + dbz * dbz) + double(src_ei)) /
                    dro;

  rslt.q[0][c] = ro;
  rslt.q[1][c] = vx;
  rslt.q[2][c] = vy;
  rslt.q[3][c] = vz;
  rslt.q[4][c] = bx;
  rslt.q[5][c] = by;
  rslt.q[6][c] = bz;
  rslt.q[7][c] = R(ei);
  rslt.q[8][c] = ph;
}

// Energy recovery is carried in double for both storage types ("pressure
// recovery from total energy" is the cancellation-prone step).
// One 64 x 4 block of cells (x-tile blk_x, y-rows 4*blk_y.., plane blk_z);
// the stage step dt, ch2 and damp come as arguments so that a batched launch
// can take them per instance from device memory.
// TILE: the x and y derivatives read the block's 68 x 8 neighbourhood of
// argm from LDS, loaded once by the whole block (3D only: both halos exist).
template <typename R, bool HX, bool HY, bool HZ, int SPLIT = 0, bool TILE = false>
__device__ __forceinline__ void rk_block(const Fld<const R> &orgn, const Fld<const R> &argm,
                                         const Fld<R> &rslt, const R *__restrict__ dxi,
                                         const R *__restrict__ dyi, const R *__restrict__ dzi,
                                         const Shape &sh, const RkScalars &sc, const R dt,
                                         const R ch2, const R damp, int blk_x, int blk_y,
                                         int blk_z) {
  const int i = blk_x * 64 + (threadIdx.x & 63) + sh.mi;
  const int j = blk_y * 4 + (threadIdx.x >> 6) + sh.mj;
  const int k = blk_z + sh.mk;
  R *tile = nullptr;
  if (TILE) {
    __shared__ R lds[9 * RK_TILE];
    tile = lds;
    const int i0 = blk_x * 64 + sh.mi - 2, j0 = blk_y * 4 + sh.mj - 2;
    for (int t = threadIdx.x; t < RK_TILE; t += 256) {
      const int gi = i0 + t % RK_TX, gj = j0 + t / RK_TX;
      if (gi < sh.It && gj < sh.Jt) {
        const int gc = sh.idx(gi, gj, k);
#pragma unroll
        for (int m = 0; m < 9; ++m) lds[m * RK_TILE + t] = argm.q[m][gc];
      }
    }
    __syncthreads();
  }
  if (i >= sh.It - sh.mi || j >= sh.Jt - sh.mj) return;
  const int c = sh.idx(i, j, k);
  const int tc = ((threadIdx.x >> 6) + 2) * RK_TX + (threadIdx.x & 63) + 2;
  const R gm1 = R(sc.gm) - R(1);
  const R pii8 = Const<R>::pii8;
  R S[17];

  // ---- x: start every running sum
  if (TILE)
    derivs_or_zero<R, 0, HX, SPLIT>(TileAcc<R>{tile}, tc, 1, HX ? dxi[i] : R(0), gm1, S);
  else
    derivs_or_zero<R, 0, HX, SPLIT>(GlobalAcc<R>{argm}, c, 1, HX ? dxi[i] : R(0), gm1, S);
  R a_ro = -S[0];
  R a_vx = -S[1], d4x = S[4], d5x = S[5] * pii8, t_vx = S[6];
  R a_vy = -S[2], t_vy = S[7];
  R a_vz = -S[3], t_vz = S[8];
  R a_by = -S[9], k_by = S[10];
  R a_bz = -S[11], k_bz = S[12];
  R ph_bx = S[13];
  R a_ph = -S[14];
  R a_ei = -S[15], t_ei = S[16];
  // Phase fence: the y loads' index depends (opaquely) on an x result, so
  // the compiler cannot hoist them above the x phase; each phase keeps 36
  // loads in flight instead of 108, which keeps the wave count per SIMD up.
  int cy = TILE ? tc : c;
  asm volatile("" : "+v"(cy), "+v"(a_ro), "+v"(a_vx), "+v"(d4x), "+v"(d5x), "+v"(t_vx),
               "+v"(a_vy), "+v"(t_vy), "+v"(a_vz), "+v"(t_vz), "+v"(a_by), "+v"(k_by),
               "+v"(a_bz), "+v"(k_bz), "+v"(ph_bx), "+v"(a_ph), "+v"(a_ei), "+v"(t_ei));

  // ---- y
  if (TILE)
    derivs_or_zero<R, 1, HY, SPLIT>(TileAcc<R>{tile}, cy, RK_TX, HY ? dyi[j] : R(0), gm1, S);
  else
    derivs_or_zero<R, 1, HY, SPLIT>(GlobalAcc<R>{argm}, cy, sh.It, HY ? dyi[j] : R(0), gm1, S);
  a_ro = a_ro - S[0];
  a_vx = a_vx - S[1];
  t_vx = t_vx + S[6];
  a_vy = a_vy - S[2];
  const R d4y = S[4], d5y = S[5] * pii8;
  t_vy = t_vy + S[7];
  a_vz = a_vz - S[3];
  t_vz = t_vz + S[8];
  R a_bx = -S[9];
  const R k_bx = S[10];
  const R ph_by = S[13];
  a_bz = a_bz - S[11] + k_bz + S[12];
  a_ph = a_ph - S[14];
  a_ei = a_ei - S[15];
  t_ei = t_ei + S[16];
  int cz = c;
  R d4y_ = d4y, d5y_ = d5y, k_bx_ = k_bx, ph_by_ = ph_by;
  asm volatile("" : "+v"(cz), "+v"(a_ro), "+v"(a_vx), "+v"(d4x), "+v"(d5x), "+v"(t_vx),
               "+v"(a_vy), "+v"(t_vy), "+v"(d4y_), "+v"(d5y_), "+v"(a_vz), "+v"(t_vz),
               "+v"(a_bx), "+v"(k_bx_), "+v"(a_by), "+v"(k_by), "+v"(ph_by_), "+v"(ph_bx),
               "+v"(a_bz), "+v"(a_ph), "+v"(a_ei), "+v"(t_ei));

  // ---- z
  derivs_or_zero<R, 2, HZ, SPLIT>(GlobalAcc<R>{argm}, cz, sh.It * sh.Jt, HZ ? dzi[k] : R(0), gm1,
                                  S);
  a_ro = a_ro - S[0];
  a_vx = a_vx - S[1];
  t_vx = t_vx + S[6];
  a_vy = a_vy - S[2];
  t_vy = t_vy + S[7];
  a_vz = a_vz - S[3];
  const R d4z = S[4], d5z = S[5] * pii8;
  t_vz = t_vz + S[8];
  a_bx = a_bx - S[9] + k_bx_ + S[10] - ph_bx;
  a_by = a_by - S[11] + k_by + S[12] - ph_by_;
  a_bz = a_bz - S[13];
  a_ph = a_ph - S[14];
  a_ei = a_ei - S[15];
  t_ei = t_ei + S[16];

  R o[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) o[m] = orgn.q[m][c];
  const R gro = sc.src_kind == 1 ? argm.q[0][c] : R(0);
  rk_update<R>(o, rslt, c, gro, sc, dt, ch2, damp, a_ro, a_vx, d4x, d5x, t_vx, a_vy, d4y_,
               d5y_, t_vy, a_vz, d4z, d5z, t_vz, a_bx, a_by, a_bz, a_ph, a_ei, t_ei);
}

template <typename R, bool HX, bool HY, bool HZ, int SPLIT>
__global__ __launch_bounds__(256) void rk_stage_kernel(Fld<const R> orgn, Fld<const R> argm,
                                                       Fld<R> rslt, const R *__restrict__ dxi,
                                                       const R *__restrict__ dyi,
                                                       const R *__restrict__ dzi, Shape sh,
                                                       RkScalars sc, Pencils pc) {
  int blk_x, blk_y, blk_z;
  if (!pc.decode(xcd_virtual_block(blockIdx.x, pc.nb), blk_x, blk_y, blk_z)) return;
  rk_block<R, HX, HY, HZ, SPLIT>(orgn, argm, rslt, dxi, dyi, dzi, sh, sc, R(sc.dt), R(sc.ch2),
                          R(sc.damp), blk_x, blk_y, blk_z);
}

// The per-point-fence variant held to 128 registers (4 waves per SIMD; the
// compiler spills a few dwords to scratch).
template <typename R, bool HX, bool HY, bool HZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void
rk_stage_kernel_w4(Fld<const R> orgn, Fld<const R> argm, Fld<R> rslt, const R *__restrict__ dxi,
                   const R *__restrict__ dyi, const R *__restrict__ dzi, Shape sh, RkScalars sc,
                   Pencils pc) {
  int blk_x, blk_y, blk_z;
  if (!pc.decode(xcd_virtual_block(blockIdx.x, pc.nb), blk_x, blk_y, blk_z)) return;
  rk_block<R, HX, HY, HZ, 1>(orgn, argm, rslt, dxi, dyi, dzi, sh, sc, R(sc.dt), R(sc.ch2),
                             R(sc.damp), blk_x, blk_y, blk_z);
}

// The x-y neighbourhood staged in LDS (3D).
template <typename R>
__global__ __launch_bounds__(256) void rk_stage_kernel_tile(Fld<const R> orgn, Fld<const R> argm,
                                                            Fld<R> rslt, const R *__restrict__ dxi,
                                                            const R *__restrict__ dyi,
                                                            const R *__restrict__ dzi, Shape sh,
                                                            RkScalars sc, Pencils pc) {
  int blk_x, blk_y, blk_z;
  if (!pc.decode(xcd_virtual_block(blockIdx.x, pc.nb), blk_x, blk_y, blk_z)) return;
  rk_block<R, true, true, true, 0, true>(orgn, argm, rslt, dxi, dyi, dzi, sh, sc, R(sc.dt),
                                         R(sc.ch2), R(sc.damp), blk_x, blk_y, blk_z);
}

// ------------------------------------------------------------ z-march (3D)
// A block owns one 64 x TY column of cells (TY = 4 or 8 rows, 64*TY
// threads) and walks it up through `kz` planes.  The z stencil reads no
// global memory: each thread keeps its own cell's fields at planes k-2 and
// k-1 in registers and at planes k+1 and k+2 in two thread-private LDS slots
// (FP64 has no room for all four in registers at two waves per SIMD); plane
// k is the centre of the 68 x (TY+4) x-y tile in LDS, which is rebuilt for
// every plane from the slot of plane k+1 and from halo loads issued during
// the previous plane's update; plane k+3 is loaded while plane k is
// computed.  Global loads per cell: 9 (new plane) + 9 * halo/(64 TY) (halo:
// 288 cells for TY = 4, 304 for TY = 8) + 9 (orgn), against 64 for the
// one-plane-per-block kernels.  The arithmetic is rk_block's, term for term.

// Tile position of halo slot s of a 64 x TY block's 68 x (TY+4) tile: the
// two rows below, the two rows above, then columns 0, 1, 66, 67 of the TY
// centre rows.
template <int TY>
__device__ __forceinline__ int rk_halo_pos(int s) {
  if (s < 2 * RK_TX) return s;
  if (s < 4 * RK_TX) return (TY + 2) * RK_TX + (s - 2 * RK_TX);
  const int q = s - 4 * RK_TX, col = q & 3;
  return (2 + (q >> 2)) * RK_TX + (col < 2 ? col : 64 + col);
}

// dir_derivs with the four stencil points' fields given.
template <typename R, int D>
__device__ __forceinline__ void derivs_pts(const R (&p2)[9], const R (&p1)[9], const R (&m1)[9],
                                           const R (&m2)[9], R dd, R gm1, R S[17]) {
  R t[17];
  point_terms<R, D>(p2, gm1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = -t[n];
  point_terms<R, D>(p1, gm1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = dfma(R(8), t[n], S[n]);
  point_terms<R, D>(m1, gm1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = dfma(R(-8), t[n], S[n]);
  point_terms<R, D>(m2, gm1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = (S[n] + t[n]) * Const<R>::inv12 * dd;
}

// derivs_pts with each point's products given beside its fields.
template <typename R, int D>
__device__ __forceinline__ void derivs_pts_p(const R (&f2)[9], const R (&q2)[7],
                                             const R (&f1)[9], const R (&q1)[7],
                                             const R (&g1)[9], const R (&r1)[7],
                                             const R (&g2)[9], const R (&r2)[7], R dd, R S[17]) {
  R t[17];
  point_terms_p<R, D>(f2, q2, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = -t[n];
  point_terms_p<R, D>(f1, q1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = dfma(R(8), t[n], S[n]);
  point_terms_p<R, D>(g1, r1, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = dfma(R(-8), t[n], S[n]);
  point_terms_p<R, D>(g2, r2, t);
#pragma unroll
  for (int n = 0; n < 17; ++n) S[n] = (S[n] + t[n]) * Const<R>::inv12 * dd;
}

// dir_derivs over an LDS tile of fields (tile) and one of the cells'
// products (ptile), both TS cells per field.
template <typename R, int D, int TS>
__device__ __forceinline__ void dir_derivs_tp(const R *tile, const R *ptile, int c, int s,
                                              R dd, R S[17]) {
  R f[4][9], q[4][7];
  const int at[4] = {c + 2 * s, c + s, c - s, c - 2 * s};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int m = 0; m < 9; ++m) f[a][m] = (m == 0 || m == 7) ? R(0) : tile[m * TS + at[a]];
#pragma unroll
    for (int n = 0; n < 7; ++n) q[a][n] = ptile[n * TS + at[a]];
  }
  derivs_pts_p<R, D>(f[0], q[0], f[1], q[1], f[2], q[2], f[3], q[3], dd, S);
}

// SAME: orgn is argm (the first stage); its values are the tile centre.
template <typename R, int TY, bool SAME>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(TY == 12 ? 3 : 2,
                                                                        TY == 12 ? 3 : 2))) void
rk_stage_kernel_march(Fld<const R> orgn, Fld<const R> argm, Fld<R> rslt,
                      const R *__restrict__ dxi, const R *__restrict__ dyi,
                      const R *__restrict__ dzi, Shape sh, RkScalars sc, March mc) {
  constexpr int NT = 64 * TY;                 // threads
  constexpr int TS = RK_TX * (TY + 4);        // tile cells
  constexpr int NH = TS - NT;                 // halo cells
  constexpr bool TWO = NH > NT;               // some threads own two halo cells
  // DB (the 8- and 12-row columns): two x-y tiles.  While the block works
  // on plane k's tile, each thread writes its share of plane k+1's into the
  // other one -- its own cell from its z slot, its halo cell as soon as the
  // y phase is done (the halo loads were issued at the top of the
  // iteration, so they do not stay in registers across the z phase and the
  // update, where they would spill, and a scratch access waits for every
  // global load before it, prefetches included) -- so one block barrier per
  // plane is enough, and one slot (plane k+1, then k+2) per thread.  The
  // 4-row column (registers too few to issue the halo early, LDS too small
  // for nothing) keeps one tile, two slots and two barriers per plane.
  constexpr bool DB = TY >= 8;
  __shared__ R tiles[(DB ? 2 : 1) * 9 * TS];
  int tb = 0;  // offset of plane k's tile in `tiles`
  __shared__ R slot[(DB ? 1 : 2) * 9 * NT];  // [plane parity][field][thread]
  // PRE (FP32, 8-row column): each cell's seven products (cell_products)
  // are formed once, when its values enter the tile or the thread's z
  // window, and kept beside them, in a second pair of LDS tiles and in
  // registers; the derivatives read them instead of forming them at each of
  // the cell's 12 stencil uses (the same operations on the same values, so
  // the same bits)
  constexpr bool PRE = DB && sizeof(R) == 4 && TY == 8 && !TWO;
  __shared__ R ptiles[PRE ? 2 * 7 * TS : 1];
  const int V = xcd_virtual_block(blockIdx.x, mc.nb);
  const int bx = V % mc.nbx, rr = V / mc.nbx, by = rr % mc.nby, zc = rr / mc.nby;
  const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
  const int i = bx * 64 + tx + sh.mi, j = by * TY + ty + sh.mj;
  const int kb = zc * mc.kz + sh.mk;
  const int ke = min(kb + mc.kz, sh.Kt - sh.mk);
  const bool valid = i < sh.It - sh.mi && j < sh.Jt - sh.mj;
  const int plane = sh.It * sh.Jt;
  // in-plane offsets; a position outside the array reads cell (0, 0) of the
  // plane instead, whose value no valid cell uses.  A thread on a ghost
  // column or row of a wrapped axis (the last tile of an axis whose size is
  // not a multiple of the tile's) loads its periodic source: its values are
  // the x or y halo of the valid cells next to it
  const bool wx = sc.wrap & 1, wy = sc.wrap & 2, wz = sc.wrap & 4;
  const int pc = (i < sh.It && j < sh.Jt) ? wrap_coord(i, sh.It, sh.mi, wx) +
                                                sh.It * wrap_coord(j, sh.Jt, sh.mj, wy)
                                          : 0;
  const int i0 = bx * 64 + sh.mi - 2, j0 = by * TY + sh.mj - 2;
  const bool has0 = tid < NH;
  const bool has1 = TWO && tid + NT < NH;
  const int p0 = rk_halo_pos<TY>(has0 ? tid : 0), p1 = rk_halo_pos<TY>(has1 ? tid + NT : 0);
  // a halo cell on a ghost column or row of a wrapped axis is read at its
  // periodic source (the value the ghost fill would have copied there)
  const int g0i = wrap_coord(i0 + p0 % RK_TX, sh.It, sh.mi, wx);
  const int g0j = wrap_coord(j0 + p0 / RK_TX, sh.Jt, sh.mj, wy);
  const int g1i = wrap_coord(i0 + p1 % RK_TX, sh.It, sh.mi, wx);
  const int g1j = wrap_coord(j0 + p1 / RK_TX, sh.Jt, sh.mj, wy);
  const int h0 = (g0i < sh.It && g0j < sh.Jt) ? g0i + sh.It * g0j : 0;
  const int h1 = (g1i < sh.It && g1j < sh.Jt) ? g1i + sh.It * g1j : 0;
  const int tc = (ty + 2) * RK_TX + tx + 2;
  const R gm1 = R(sc.gm) - R(1);
  const R pii8 = Const<R>::pii8;
  const R ddx = dxi[valid ? i : sh.mi], ddy = dyi[valid ? j : sh.mj];
  const R dt = R(sc.dt), ch2 = R(sc.ch2), damp = R(sc.damp);
  auto sl = [&](int k, int m) -> R & {
    return slot[((DB ? 0 : (k & 1)) * 9 + m) * NT + tid];
  };

  // plane q's own cell (a ghost plane of a wrapped z axis: its source)
  auto zcell = [&](int q) { return pc + wrap_coord(q, sh.Kt, sh.mk, wz) * plane; };
  R zm2[9], zm1[9];  // own cell at planes k-2, k-1
  R zn[9];           // own cell at plane k+2, stored to its slot before the z phase
  // PRE: products of the own cell at planes k-2, k-1, k (the tile centre),
  // k+1 and k+2
  R pm2[PRE ? 7 : 1], pm1[PRE ? 7 : 1], pc0[PRE ? 7 : 1], pp1[PRE ? 7 : 1], pp2[PRE ? 7 : 1];
  auto put_products = [&](const R (&f)[9], int at, R (&p)[7]) {
    cell_products(f, gm1, p);
#pragma unroll
    for (int n = 0; n < 7; ++n) ptiles[n * TS + at] = p[n];
  };
  {
    const int c = pc + kb * plane, h = kb * plane;
    const int cm2 = zcell(kb - 2), cm1 = zcell(kb - 1), cp1 = zcell(kb + 1), cp2 = zcell(kb + 2);
    R fc[9], fh[9], fp[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      zm2[m] = argm.q[m][cm2];
      zm1[m] = argm.q[m][cm1];
      fp[m] = argm.q[m][cp1];
      sl(kb + 1, m) = fp[m];
      zn[m] = argm.q[m][cp2];
      fc[m] = argm.q[m][c];
      tiles[m * TS + tc] = fc[m];
      fh[m] = has0 ? argm.q[m][h + h0] : R(0);
      if (has0) tiles[m * TS + p0] = fh[m];
      if (has1) tiles[m * TS + p1] = argm.q[m][h + h1];
    }
    if constexpr (PRE) {
      cell_products(zm2, gm1, pm2);
      cell_products(zm1, gm1, pm1);
      cell_products(fp, gm1, pp1);
      put_products(fc, tc, pc0);
      if (has0) {
        R ph[7];
        put_products(fh, p0, ph);
      }
    }
  }
  __syncthreads();

  for (int k = kb; k < ke; ++k) {
    const bool more = k + 1 < ke;  // uniform over the block
    // plane k+1's halo, issued first: it lands during the x and y phases
    // and goes into the next tile before the z phase (the 4-row tile loads
    // it before the update instead)
    R ha[9], hb[9];
    if (DB && more) {
      const int h = (k + 1) * plane;
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        ha[m] = has0 ? argm.q[m][h + h0] : R(0);
        hb[m] = has1 ? argm.q[m][h + h1] : R(0);
      }
    }
    // plane k's orgn values, issued next so that their latency hides
    // behind the x, y and z derivatives (not in the FP64 4-row tile, which
    // would spill more registers)
    constexpr bool EARLY = !SAME && (TY >= 8 || sizeof(R) < 8);
    R og[9];
    if (EARLY && valid) {
#pragma unroll
      for (int m = 0; m < 9; ++m) og[m] = orgn.q[m][pc + k * plane];
    }
    R S[17];
    // ---- x (LDS tile)
    const R *lds = tiles + tb;
    const R *plds = ptiles + (PRE ? tb / 9 * 7 : 0);
    if constexpr (PRE)
      dir_derivs_tp<R, 0, TS>(lds, plds, tc, 1, ddx, S);
    else
      dir_derivs<R, 0, 1>(TileAcc<R, TS>{lds}, tc, 1, ddx, gm1, S);
    R a_ro = -S[0];
    R a_vx = -S[1], d4x = S[4], d5x = S[5] * pii8, t_vx = S[6];
    R a_vy = -S[2], t_vy = S[7];
    R a_vz = -S[3], t_vz = S[8];
    R a_by = -S[9], k_by = S[10];
    R a_bz = -S[11], k_bz = S[12];
    R ph_bx = S[13];
    R a_ph = -S[14];
    R a_ei = -S[15], t_ei = S[16];
    int cy = tc;
    asm volatile("" : "+v"(cy), "+v"(a_ro), "+v"(a_vx), "+v"(d4x), "+v"(d5x), "+v"(t_vx),
                 "+v"(a_vy), "+v"(t_vy), "+v"(a_vz), "+v"(t_vz), "+v"(a_by), "+v"(k_by),
                 "+v"(a_bz), "+v"(k_bz), "+v"(ph_bx), "+v"(a_ph), "+v"(a_ei), "+v"(t_ei));
    // ---- y (LDS tile)
    if constexpr (PRE)
      dir_derivs_tp<R, 1, TS>(lds, plds, cy, RK_TX, ddy, S);
    else
      dir_derivs<R, 1, 1>(TileAcc<R, TS>{lds}, cy, RK_TX, ddy, gm1, S);
    a_ro = a_ro - S[0];
    a_vx = a_vx - S[1];
    t_vx = t_vx + S[6];
    a_vy = a_vy - S[2];
    const R d4y = S[4], d5y = S[5] * pii8;
    t_vy = t_vy + S[7];
    a_vz = a_vz - S[3];
    t_vz = t_vz + S[8];
    R a_bx = -S[9];
    const R k_bx = S[10];
    const R ph_by = S[13];
    a_bz = a_bz - S[11] + k_bz + S[12];
    a_ph = a_ph - S[14];
    a_ei = a_ei - S[15];
    t_ei = t_ei + S[16];
    // ---- z (slots and registers)
    const int nb = DB ? 9 * TS - tb : 0;  // plane k+1's tile
    if constexpr (DB) {
      // plane k+1 from the slot into the next tile's centre, plane k+2
      // (loaded during the previous plane) into the slot, plane k+1's halo
      // into the next tile
      R zp1[9];
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        zp1[m] = sl(k + 1, m);
        tiles[nb + m * TS + tc] = zp1[m];
        sl(k + 2, m) = zn[m];
        if (more && has0) tiles[nb + m * TS + p0] = ha[m];
      }
      if constexpr (PRE) {
        const int pnb = nb / 9 * 7;
        cell_products(zn, gm1, pp2);
        R ph[7];
        if (more && has0) cell_products(ha, gm1, ph);
#pragma unroll
        for (int n = 0; n < 7; ++n) {
          ptiles[pnb + n * TS + tc] = pp1[n];
          if (more && has0) ptiles[pnb + n * TS + p0] = ph[n];
        }
        derivs_pts_p<R, 2>(zn, pp2, zp1, pp1, zm1, pm1, zm2, pm2, dzi[k], S);
      } else {
        derivs_pts<R, 2>(zn, zp1, zm1, zm2, dzi[k], gm1, S);
      }
    } else {
#pragma unroll
      for (int m = 0; m < 9; ++m) sl(k + 2, m) = zn[m];
      R zp2[9], zp1[9];
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        zp2[m] = sl(k + 2, m);
        zp1[m] = sl(k + 1, m);
      }
      derivs_pts<R, 2>(zp2, zp1, zm1, zm2, dzi[k], gm1, S);
    }
    a_ro = a_ro - S[0];
    a_vx = a_vx - S[1];
    t_vx = t_vx + S[6];
    a_vy = a_vy - S[2];
    t_vy = t_vy + S[7];
    a_vz = a_vz - S[3];
    const R d4z = S[4], d5z = S[5] * pii8;
    t_vz = t_vz + S[8];
    a_bx = a_bx - S[9] + k_bx + S[10] - ph_bx;
    a_by = a_by - S[11] + k_by + S[12] - ph_by;
    a_bz = a_bz - S[13];
    a_ph = a_ph - S[14];
    a_ei = a_ei - S[15];
    t_ei = t_ei + S[16];
    if (!DB && more) {
      const int h = (k + 1) * plane;
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        ha[m] = has0 ? argm.q[m][h + h0] : R(0);
        hb[m] = has1 ? argm.q[m][h + h1] : R(0);
      }
    }
    if (valid) {
      // plane k's own values are the tile centre (orgn of the first stage)
      R o[9];
#pragma unroll
      for (int m = 0; m < 9; ++m)
        o[m] = SAME ? lds[m * TS + tc] : EARLY ? og[m] : orgn.q[m][pc + k * plane];
      rk_update<R>(o, rslt, pc + k * plane, SAME ? o[0] : lds[tc], sc, dt, ch2, damp, a_ro,
                   a_vx, d4x, d5x, t_vx, a_vy, d4y, d5y, t_vy, a_vz, d4z, d5z, t_vz, a_bx, a_by,
                   a_bz, a_ph, a_ei, t_ei);
    }
    // plane k+3, used in the next plane's z phase
    if (more && k + 3 < sh.Kt) {
      const int cn = zcell(k + 3);
#pragma unroll
      for (int m = 0; m < 9; ++m) zn[m] = argm.q[m][cn];
    }
    if constexpr (DB) {
      if (more) {
        // plane k's centre (which only this thread wrote) moves to the
        // register window, and plane k+1's tile becomes the current one
#pragma unroll
        for (int m = 0; m < 9; ++m) {
          zm2[m] = zm1[m];
          zm1[m] = tiles[tb + m * TS + tc];
        }
        if constexpr (PRE) {
#pragma unroll
          for (int n = 0; n < 7; ++n) {
            pm2[n] = pm1[n];
            pm1[n] = pc0[n];
            pc0[n] = pp1[n];
            pp1[n] = pp2[n];
          }
        }
        tb = nb;
      }
      __syncthreads();  // plane k+1's tile is complete and plane k's is free
    } else {
      __syncthreads();  // every wave is done with plane k's tile
      if (more) {
        // the tile of plane k+1: own centre from the slot (plane k's centre,
        // which only this thread wrote, moves to the register window first),
        // halo from the loads issued before the update
#pragma unroll
        for (int m = 0; m < 9; ++m) {
          zm2[m] = zm1[m];
          zm1[m] = tiles[m * TS + tc];
          tiles[m * TS + tc] = sl(k + 1, m);
          if (has0) tiles[m * TS + p0] = ha[m];
          if (has1) tiles[m * TS + p1] = hb[m];
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------ y walk (2D)
// Runs with x and y axes and no z axis: a wave owns 64 consecutive x cells
// -- 60 outputs and two halo cells at either end -- and walks `rows` rows
// up y from interior row y0.  A lane forms its own cell's x terms once and
// takes its neighbours' by DPP lane shifts; the y stencil reads a register
// window of five rows of fields, the next row loaded a step ahead.  Each
// cell is read once per stage instead of once per stencil point, and the
// wave needs no LDS and no barrier.  The arithmetic of a cell is
// rk_block's, term for term and in the same order (the z sums of a run
// without a z axis are rk_block's zeros), so the bits are the same.
// (RK_YWALK_OUT outputs and RK_YWALK_ROWS rows per wave: launch.hpp)

template <typename R>
__device__ __forceinline__ void rk_ywalk_wave(const Fld<const R> &orgn, const Fld<const R> &argm,
                                              const Fld<R> &rslt, const R *__restrict__ dxi,
                                              const R *__restrict__ dyi, const Shape &sh,
                                              const RkScalars &sc, const R dt, const R ch2,
                                              const R damp, int strip, int y0, int rows) {
  const int lane = threadIdx.x & 63;
  const int xi = strip * RK_YWALK_OUT + lane - 2;  // interior x of the lane's cell
  // lanes past the array read its last column: no output lane uses them
  const int i = min(xi + sh.mi, sh.It - 1);
  const bool out = lane >= 2 && lane < 62 && xi < sh.It - 2 * sh.mi;
  const R gm1 = R(sc.gm) - R(1);
  const R pii8 = Const<R>::pii8;
  const R ddx = dxi[i];
  const int j0 = y0 + sh.mj;  // array row of the first output row
  R w[5][9];                  // rows j-2 .. j+2 of the lane's column
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    const int c = sh.idx(i, j0 - 2 + r, 0);
#pragma unroll
    for (int m = 0; m < 9; ++m) w[r][m] = argm.q[m][c];
  }
  R nx[9] = {};  // row j+3
  if (rows > 1) {
    const int c = sh.idx(i, j0 + 3, 0);
#pragma unroll
    for (int m = 0; m < 9; ++m) nx[m] = argm.q[m][c];
  }
  for (int t = 0; t < rows; ++t) {
    const int j = j0 + t;
    const int c = sh.idx(i, j, 0);
    R o[9];
    if (out) {
#pragma unroll
      for (int m = 0; m < 9; ++m) o[m] = orgn.q[m][c];
    }
    R S[17];
    // ---- x: the lane's own terms, its neighbours' by lane shifts
    {
      R tx[17];
      point_terms<R, 0>(w[2], gm1, tx);
#pragma unroll
      for (int n = 0; n < 17; ++n) {
        const R p1 = lane_dn(tx[n]), p2 = lane_dn(p1);
        const R m1 = lane_up(tx[n]), m2 = lane_up(m1);
        R a = -p2;
        a = dfma(R(8), p1, a);
        a = dfma(R(-8), m1, a);
        S[n] = (a + m2) * Const<R>::inv12 * ddx;
      }
    }
    R a_ro = -S[0];
    R a_vx = -S[1], d4x = S[4], d5x = S[5] * pii8, t_vx = S[6];
    R a_vy = -S[2], t_vy = S[7];
    R a_vz = -S[3], t_vz = S[8];
    R a_by = -S[9], k_by = S[10];
    R a_bz = -S[11], k_bz = S[12];
    R ph_bx = S[13];
    R a_ph = -S[14];
    R a_ei = -S[15], t_ei = S[16];
    // ---- y: the register window
    derivs_pts<R, 1>(w[4], w[3], w[1], w[0], dyi[j], gm1, S);
    a_ro = a_ro - S[0];
    a_vx = a_vx - S[1];
    t_vx = t_vx + S[6];
    a_vy = a_vy - S[2];
    const R d4y = S[4], d5y = S[5] * pii8;
    t_vy = t_vy + S[7];
    a_vz = a_vz - S[3];
    t_vz = t_vz + S[8];
    R a_bx = -S[9];
    const R k_bx = S[10];
    const R ph_by = S[13];
    a_bz = a_bz - S[11] + k_bz + S[12];
    a_ph = a_ph - S[14];
    a_ei = a_ei - S[15];
    t_ei = t_ei + S[16];
    // ---- z: no axis (rk_block's zero sums)
#pragma unroll
    for (int n = 0; n < 17; ++n) S[n] = R(0);
    a_ro = a_ro - S[0];
    a_vx = a_vx - S[1];
    t_vx = t_vx + S[6];
    a_vy = a_vy - S[2];
    t_vy = t_vy + S[7];
    a_vz = a_vz - S[3];
    const R d4z = S[4], d5z = S[5] * pii8;
    t_vz = t_vz + S[8];
    a_bx = a_bx - S[9] + k_bx + S[10] - ph_bx;
    a_by = a_by - S[11] + k_by + S[12] - ph_by;
    a_bz = a_bz - S[13];
    a_ph = a_ph - S[14];
    a_ei = a_ei - S[15];
    t_ei = t_ei + S[16];
    if (out) {
      const R gro = sc.src_kind == 1 ? w[2][0] : R(0);
      rk_update<R>(o, rslt, c, gro, sc, dt, ch2, damp, a_ro, a_vx, d4x, d5x, t_vx, a_vy, d4y,
                   d5y, t_vy, a_vz, d4z, d5z, t_vz, a_bx, a_by, a_bz, a_ph, a_ei, t_ei);
    }
    // the window moves up a row; row j+4 is loaded for the step after next
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      w[0][m] = w[1][m];
      w[1][m] = w[2][m];
      w[2][m] = w[3][m];
      w[3][m] = w[4][m];
      w[4][m] = nx[m];
    }
    if (t + 2 < rows) {  // row j+4 is inside the array when j+2 is an output row
      const int cn = sh.idx(i, j + 4, 0);
#pragma unroll
      for (int m = 0; m < 9; ++m) nx[m] = argm.q[m][cn];
    }
  }
}

// One x-y run (the z-march variants' choice on a grid without a z axis):
// four waves per block, each a (strip, row chunk) of the run.
template <typename R>
__global__ __launch_bounds__(256) void rk_stage_kernel_ywalk(Fld<const R> orgn, Fld<const R> argm,
                                                             Fld<R> rslt,
                                                             const R *__restrict__ dxi,
                                                             const R *__restrict__ dyi, Shape sh,
                                                             RkScalars sc, int nb) {
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj;
  const int nstrip = (ni + RK_YWALK_OUT - 1) / RK_YWALK_OUT;
  const int nch = (nj + RK_YWALK_ROWS - 1) / RK_YWALK_ROWS;
  const int W = xcd_virtual_block(blockIdx.x, nb) * 4 + (threadIdx.x >> 6);
  if (W >= nstrip * nch) return;  // uniform over the wave
  const int strip = W % nstrip, ch = W / nstrip;
  const int y0 = ch * RK_YWALK_ROWS;
  rk_ywalk_wave<R>(orgn, argm, rslt, dxi, dyi, sh, sc, R(sc.dt), R(sc.ch2), R(sc.damp), strip,
                   y0, min(RK_YWALK_ROWS, nj - y0));
}

template <typename R, bool HX, bool HY, bool HZ>
static hipError_t launch_rk_t(hipStream_t st, const Fld<const R> &o, const Fld<const R> &a,
                              const Fld<R> &r, const R *dxi, const R *dyi, const R *dzi,
                              const Shape &sh, const RkScalars &sc) {
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nk = sh.Kt - 2 * sh.mk;
  if (HX && HY && HZ && (sc.variant >= 4 && sc.variant <= 6)) {
    // variant 6: the 12-row column at three waves per SIMD, for FP32 (an
    // FP64 stage needs more than the 170 registers that allows: it runs the
    // 8-row column)
    const int ty = sc.variant == 4 ? 4 : (sc.variant == 6 && sizeof(R) == 4) ? 12 : 8;
    March mc;
    if (sh.mi < 2 || sh.mj < 2 || sh.mk < 2 ||
        !mc.init((ni + 63) / 64, (nj + ty - 1) / ty, nk, sc.zchunk))
      return hipErrorInvalidValue;
    bool same = true;
    for (int m = 0; m < 9; ++m) same = same && o.q[m] == a.q[m];
    if (ty == 4 && same)
      hipLaunchKernelGGL((rk_stage_kernel_march<R, 4, true>), dim3(mc.nb), dim3(256), 0, st, o,
                         a, r, dxi, dyi, dzi, sh, sc, mc);
    else if (ty == 4)
      hipLaunchKernelGGL((rk_stage_kernel_march<R, 4, false>), dim3(mc.nb), dim3(256), 0, st, o,
                         a, r, dxi, dyi, dzi, sh, sc, mc);
    else if (ty == 12) {
      if constexpr (sizeof(R) == 4) {
        if (same)
          hipLaunchKernelGGL((rk_stage_kernel_march<R, 12, true>), dim3(mc.nb), dim3(768), 0, st,
                             o, a, r, dxi, dyi, dzi, sh, sc, mc);
        else
          hipLaunchKernelGGL((rk_stage_kernel_march<R, 12, false>), dim3(mc.nb), dim3(768), 0, st,
                             o, a, r, dxi, dyi, dzi, sh, sc, mc);
      }
    } else if (same)
      hipLaunchKernelGGL((rk_stage_kernel_march<R, 8, true>), dim3(mc.nb), dim3(512), 0, st, o,
                         a, r, dxi, dyi, dzi, sh, sc, mc);
    else
      hipLaunchKernelGGL((rk_stage_kernel_march<R, 8, false>), dim3(mc.nb), dim3(512), 0, st, o,
                         a, r, dxi, dyi, dzi, sh, sc, mc);
    return hipGetLastError();
  }
  const long yw_waves = (long)((ni + RK_YWALK_OUT - 1) / RK_YWALK_OUT) *
                        ((nj + RK_YWALK_ROWS - 1) / RK_YWALK_ROWS);
  if (HX && HY && !HZ && sc.variant >= 4 && sc.variant <= 6 && yw_waves >= RK_YWALK_MIN_WAVES) {
    // an x-y run large enough to fill the chip with y-walk waves (a smaller
    // one runs the one-plane kernel below: 256^2 is 80 waves)
    if (sh.mi < 2 || sh.mj < 2) return hipErrorInvalidValue;
    const long nb = (yw_waves + 3) / 4;
    if (nb > 2147483647L) return hipErrorInvalidValue;
    hipLaunchKernelGGL((rk_stage_kernel_ywalk<R>), dim3((unsigned)nb), dim3(256), 0, st, o, a, r,
                       dxi, dyi, sh, sc, (int)nb);
    return hipGetLastError();
  }
  Pencils pc;
  if (!pc.init((ni + 63) / 64, (nj + 3) / 4, nk, sc.pencil_bx, sc.pencil_by))
    return hipErrorInvalidValue;
  if (HX && HY && HZ && sc.variant == 1)
    hipLaunchKernelGGL((rk_stage_kernel<R, HX, HY, HZ, 1>), dim3(pc.nb), dim3(256), 0, st, o, a,
                       r, dxi, dyi, dzi, sh, sc, pc);
  else if (HX && HY && HZ && sc.variant == 2)
    hipLaunchKernelGGL((rk_stage_kernel_w4<R, HX, HY, HZ>), dim3(pc.nb), dim3(256), 0, st, o, a,
                       r, dxi, dyi, dzi, sh, sc, pc);
  else if (HX && HY && HZ && sc.variant == 3)
    hipLaunchKernelGGL((rk_stage_kernel_tile<R>), dim3(pc.nb), dim3(256), 0, st, o, a, r, dxi,
                       dyi, dzi, sh, sc, pc);
  else
    hipLaunchKernelGGL((rk_stage_kernel<R, HX, HY, HZ, 0>), dim3(pc.nb), dim3(256), 0, st, o, a,
                       r, dxi, dyi, dzi, sh, sc, pc);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_rk_stage(hipStream_t st, const Fld<const R> &o, const Fld<const R> &a,
                           const Fld<R> &r, const R *dxi, const R *dyi, const R *dzi,
                           const Shape &sh, const RkScalars &sc) {
  const int key = (sh.is ? 4 : 0) | (sh.js ? 2 : 0) | (sh.ks ? 1 : 0);
  switch (key) {
    case 7: return launch_rk_t<R, true, true, true>(st, o, a, r, dxi, dyi, dzi, sh, sc);
    case 6: return launch_rk_t<R, true, true, false>(st, o, a, r, dxi, dyi, dzi, sh, sc);
    case 5: return launch_rk_t<R, true, false, true>(st, o, a, r, dxi, dyi, dzi, sh, sc);
    case 3: return launch_rk_t<R, false, true, true>(st, o, a, r, dxi, dyi, dzi, sh, sc);
    case 4: return launch_rk_t<R, true, false, false>(st, o, a, r, dxi, dyi, dzi, sh, sc);
    case 2: return launch_rk_t<R, false, true, false>(st, o, a, r, dxi, dyi, dzi, sh, sc);
    case 1: return launch_rk_t<R, false, false, true>(st, o, a, r, dxi, dyi, dzi, sh, sc);
    default: return launch_rk_t<R, false, false, false>(st, o, a, r, dxi, dyi, dzi, sh, sc);
  }
}

// ---------------------------------------------------------------- batches
template <typename R> __device__ __forceinline__ Fld<const R> as_const(const Fld<R> &f) {
  Fld<const R> o;
#pragma unroll
  for (int m = 0; m < 9; ++m) o.q[m] = f.q[m];
  return o;
}

// One RK stage of every instance of a batch in one launch: orgn = qq, argm =
// field set in_set, rslt = out_set; the stage step, ch2 and damp are the
// instance's own (device memory).  Blocks run 64 x 4 cells, x fastest.
// The batch's x-y runs: four waves per block, each a (strip, row chunk) of
// one run; a run has ceil(ni / RK_YWALK_OUT) strips and ceil(nj / RK_YWALK_ROWS) chunks.
template <typename R>
__global__ __launch_bounds__(256) void batch_rk_ywalk_kernel(const BatchInst<R> *__restrict__ inst,
                                                             const BatchDyn<R> *__restrict__ dyn,
                                                             const int *__restrict__ start, int n,
                                                             int total, int stage, int in_set,
                                                             int out_set, RkScalars sc) {
  const int V = xcd_virtual_block(blockIdx.x, total);
  const int m = batch_find(start, n, V);
  const BatchInst<R> &I = inst[m];
  const Shape sh = I.sh;
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj;
  const int nstrip = (ni + RK_YWALK_OUT - 1) / RK_YWALK_OUT;
  const int nch = (nj + RK_YWALK_ROWS - 1) / RK_YWALK_ROWS;
  const int W = (V - start[m]) * 4 + (threadIdx.x >> 6);
  if (W >= nstrip * nch) return;  // uniform over the wave
  const int strip = W % nstrip, ch = W / nstrip;
  const int y0 = ch * RK_YWALK_ROWS;
  const BatchDyn<R> &D = dyn[m];
  rk_ywalk_wave<R>(as_const(I.st[0]), as_const(I.st[in_set]), I.st[out_set], I.dxi, I.dyi, sh,
                   sc, D.dts[stage], D.ch2, D.damp[stage], strip, y0,
                   min(RK_YWALK_ROWS, nj - y0));
}

template <typename R, bool HX, bool HY, bool HZ>
__global__ __launch_bounds__(256) void batch_rk_kernel(const BatchInst<R> *__restrict__ inst,
                                                       const BatchDyn<R> *__restrict__ dyn,
                                                       const int *__restrict__ start, int n,
                                                       int total, int stage, int in_set,
                                                       int out_set, RkScalars sc) {
  const int V = xcd_virtual_block(blockIdx.x, total);
  const int m = batch_find(start, n, V);
  const int b = V - start[m];
  const BatchInst<R> &I = inst[m];
  const Shape sh = I.sh;
  const int nbx = (sh.It - 2 * sh.mi + 63) / 64, nby = (sh.Jt - 2 * sh.mj + 3) / 4;
  const int blk_x = b % nbx, r = b / nbx;
  const BatchDyn<R> &D = dyn[m];
  rk_block<R, HX, HY, HZ>(as_const(I.st[0]), as_const(I.st[in_set]), I.st[out_set], I.dxi, I.dyi,
                          I.dzi, sh, sc, D.dts[stage], D.ch2, D.damp[stage], blk_x, r % nby,
                          r / nby);
}

template <typename R, bool HX, bool HY, bool HZ>
static hipError_t launch_batch_rk_t(hipStream_t st, const BatchInst<R> *inst,
                                    const BatchDyn<R> *dyn, const int *start, int n, int total,
                                    int stage, int in_set, int out_set, const RkScalars &sc) {
  hipLaunchKernelGGL((batch_rk_kernel<R, HX, HY, HZ>), dim3(total), dim3(256), 0, st, inst, dyn,
                     start, n, total, stage, in_set, out_set, sc);
  return hipGetLastError();
}

// key = (is ? 4 : 0) | (js ? 2 : 0) | (ks ? 1 : 0), the same for every instance
template <typename R>
hipError_t launch_batch_rk(hipStream_t st, const BatchInst<R> *inst, const BatchDyn<R> *dyn,
                           const int *start, int n, int total, int stage, int in_set,
                           int out_set, const RkScalars &sc, int key) {
  switch (key) {
    case 7: return launch_batch_rk_t<R, true, true, true>(st, inst, dyn, start, n, total, stage, in_set, out_set, sc);
    case 6:  // x-y runs: the y walk (its own block count, batch_rk_blocks)
      hipLaunchKernelGGL((batch_rk_ywalk_kernel<R>), dim3(total), dim3(256), 0, st, inst, dyn,
                         start, n, total, stage, in_set, out_set, sc);
      return hipGetLastError();
    case 5: return launch_batch_rk_t<R, true, false, true>(st, inst, dyn, start, n, total, stage, in_set, out_set, sc);
    case 3: return launch_batch_rk_t<R, false, true, true>(st, inst, dyn, start, n, total, stage, in_set, out_set, sc);
    case 4: return launch_batch_rk_t<R, true, false, false>(st, inst, dyn, start, n, total, stage, in_set, out_set, sc);
    case 2: return launch_batch_rk_t<R, false, true, false>(st, inst, dyn, start, n, total, stage, in_set, out_set, sc);
    case 1: return launch_batch_rk_t<R, false, false, true>(st, inst, dyn, start, n, total, stage, in_set, out_set, sc);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_batch_rk<double>(hipStream_t, const BatchInst<double> *,
                                            const BatchDyn<double> *, const int *, int, int, int,
                                            int, int, const RkScalars &, int);
template hipError_t launch_batch_rk<float>(hipStream_t, const BatchInst<float> *,
                                           const BatchDyn<float> *, const int *, int, int, int,
                                           int, int, const RkScalars &, int);

template hipError_t launch_rk_stage<double>(hipStream_t, const Fld<const double> &,
                                            const Fld<const double> &, const Fld<double> &,
                                            const double *, const double *, const double *,
                                            const Shape &, const RkScalars &);
template hipError_t launch_rk_stage<float>(hipStream_t, const Fld<const float> &,
                                           const Fld<const float> &, const Fld<float> &,
                                           const float *, const float *, const float *,
                                           const Shape &, const RkScalars &);

}  // namespace misohip
