// In-device ghost fill, CFL reduction and host<->device layout transposes.
//
// Ghost fill replaces Integrator::apply_boundary_condition for one rank on
// one GPU (miso/include/miso/mhd_integrator_host.hpp:250-255): the user
// mirror boundary on physical faces (boundary_condition.hpp:77-136) and the
// MPI halo self-exchange on periodic axes (mhd_halo_exchange.hpp:51-266)
// become one pass in which every ghost cell reads its source cell directly.
// Each coordinate is mapped on its own (periodic wrap or mirror), so faces
// carry exactly the values the reference's exchange produces and edges and
// corners are well defined (the reference leaves them stale; no interior
// stencil reads them).  A fixed axis (a user boundary functor that leaves
// the faces alone, like EmptyBoundaryCondition, mhd_model_base.hpp:125)
// keeps its ghost values: the coordinate maps to itself.
//
// CFL: Integrator::cfl (mhd_integrator_host.hpp:303-334) narrows every
// candidate to float through util::min2, so dt = (float)min(candidates).
// Candidates are positive, so the IEEE bit pattern of the float orders like
// an unsigned integer: a wavefront shuffle-min on the bits, an LDS block min,
// then one unsigned atomicMin per block.
#include "launch.hpp"

namespace misohip {


// Flattened over the shell: x slabs (all j, k), then y slabs (interior i,
// all k), then z slabs (interior i, j).  The three sets partition the ghosts.
template <typename R>
__device__ __forceinline__ void ghost_cell(const Fld<R> &qq, const Shape &sh, const BcDev &bc,
                                           long n_x, long n_y, long t) {
  int i, j, k;
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj;
  if (t < n_x) {
    // x slabs: 2*mi columns x Jt x Kt, i fastest
    const int w = 2 * sh.mi;
    const int a = (int)(t % w);
    const long r = t / w;
    j = (int)(r % sh.Jt);
    k = (int)(r / sh.Jt);
    i = a < sh.mi ? a : sh.It - 2 * sh.mi + a;
  } else if (t < n_x + n_y) {
    const long u = t - n_x;
    i = sh.mi + (int)(u % ni);
    const long r = u / ni;
    const int a = (int)(r % (2 * sh.mj));
    k = (int)(r / (2 * sh.mj));
    j = a < sh.mj ? a : sh.Jt - 2 * sh.mj + a;
  } else {
    const long u = t - n_x - n_y;
    i = sh.mi + (int)(u % ni);
    const long r = u / ni;
    j = sh.mj + (int)(r % nj);
    const int a = (int)(r / nj);
    k = a < sh.mk ? a : sh.Kt - 2 * sh.mk + a;
  }
  bool mx, my, mz;
  const int si = map_coord(i, sh.It, sh.mi, bc.periodic[0], mx);
  const int sj = map_coord(j, sh.Jt, sh.mj, bc.periodic[1], my);
  const int sk = map_coord(k, sh.Kt, sh.mk, bc.periodic[2], mz);
  const int dst = sh.idx(i, j, k), src = sh.idx(si, sj, sk);
  // a ghost coordinate of a fixed axis reads the fixed fields
  const bool from_fixed = (bc.periodic[0] == 2 && (i < sh.mi || i >= sh.It - sh.mi)) ||
                          (bc.periodic[1] == 2 && (j < sh.mj || j >= sh.Jt - sh.mj)) ||
                          (bc.periodic[2] == 2 && (k < sh.mk || k >= sh.Kt - sh.mk));
#pragma unroll
  for (int f = 0; f < 9; ++f) {
    R v = from_fixed ? static_cast<const R *>(bc.fixed[f])[src] : qq.q[f][src];
    if (mx) v = R(bc.sign[0][f]) * v;
    if (my) v = R(bc.sign[1][f]) * v;
    if (mz) v = R(bc.sign[2][f]) * v;
    qq.q[f][dst] = v;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void ghost_fill_kernel(Fld<R> qq, Shape sh, BcDev bc, long n_x,
                                                         long n_y, long n_tot) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tot) return;
  ghost_cell<R>(qq, sh, bc, n_x, n_y, t);
}

__host__ __device__ inline void ghost_counts(const Shape &sh, long &n_x, long &n_y, long &n_tot) {
  const long ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj;
  n_x = 2L * sh.mi * sh.Jt * sh.Kt;
  n_y = 2L * sh.mj * ni * sh.Kt;
  n_tot = n_x + n_y + 2L * sh.mk * ni * nj;
}

template <typename R>
hipError_t launch_ghost_fill(hipStream_t st, const Fld<R> &q, const Shape &sh, const BcDev &bc,
                             int axes) {
  const long ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj;
  // an axis left out contributes no slab; the others keep their extents, so
  // every cell a filled face holds is still written exactly once
  const long n_x = (axes & 1) ? 2L * sh.mi * sh.Jt * sh.Kt : 0;
  const long n_y = (axes & 2) ? 2L * sh.mj * ni * sh.Kt : 0;
  const long n_z = (axes & 4) ? 2L * sh.mk * ni * nj : 0;
  const long n_tot = n_x + n_y + n_z;
  if (n_tot == 0) return hipSuccess;
  const long blocks = (n_tot + 255) / 256;
  hipLaunchKernelGGL((ghost_fill_kernel<R>), dim3((unsigned)blocks), dim3(256), 0, st, q, sh, bc,
                     n_x, n_y, n_tot);
  return hipGetLastError();
}

// ---------------------------------------------------------------- CFL
__global__ void cfl_init_kernel(unsigned *slot) {
  if (threadIdx.x == 0) slot[0] = __float_as_uint(1.e10f);
}

template <typename R>
__device__ __forceinline__ float cfl_candidate(const Fld<const R> &q, int c, R gm, R cfl, R hmin) {
  return cfl_candidate_v(q.q[0][c], q.q[1][c], q.q[2][c], q.q[3][c], q.q[4][c], q.q[5][c],
                         q.q[6][c], q.q[7][c], gm, cfl, hmin);
}

// Grid-stride over the interior; one atomic per block.
template <typename R>
__global__ __launch_bounds__(256) void cfl_kernel(Fld<const R> q, const R *__restrict__ dx,
                                                  const R *__restrict__ dy,
                                                  const R *__restrict__ dz, Shape sh, double gm,
                                                  double cfl, unsigned *slot) {
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nk = sh.Kt - 2 * sh.mk;
  const long n = (long)ni * nj * nk;
  unsigned best = __float_as_uint(1.e10f);
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const int i = sh.mi + (int)(t % ni);
    const long r = t / ni;
    const int j = sh.mj + (int)(r % nj);
    const int k = sh.mk + (int)(r / nj);
    const R hmin = nmin3(dx[i], dy[j], dz[k]);
    const float cand = cfl_candidate(q, sh.idx(i, j, k), R(gm), R(cfl), hmin);
    best = min(best, __float_as_uint(cand));
  }
  best = wave_umin(best);
  __shared__ unsigned red[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned b = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = min(b, red[w]);
    atomicMin(slot, b);
  }
}

hipError_t launch_cfl_init(hipStream_t st, unsigned *slot) {
  hipLaunchKernelGGL(cfl_init_kernel, dim3(1), dim3(64), 0, st, slot);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_cfl(hipStream_t st, const Fld<const R> &q, const R *dx, const R *dy,
                      const R *dz, const Shape &sh, double gm, double cfl, unsigned *slot) {
  hipLaunchKernelGGL(cfl_init_kernel, dim3(1), dim3(64), 0, st, slot);
  const long n = (long)(sh.It - 2 * sh.mi) * (sh.Jt - 2 * sh.mj) * (sh.Kt - 2 * sh.mk);
  long blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((cfl_kernel<R>), dim3((unsigned)blocks), dim3(256), 0, st, q, dx, dy, dz, sh,
                     gm, cfl, slot);
  return hipGetLastError();
}

// ---------------------------------------------------------------- batches
// Ghost fill of field set `set` of every instance; 256 ghost cells per block.
template <typename R>
__global__ __launch_bounds__(256) void batch_ghost_kernel(const BatchInst<R> *__restrict__ inst,
                                                          const int *__restrict__ start, int n,
                                                          int set, BcDev bc) {
  const int m = batch_find(start, n, blockIdx.x);
  const BatchInst<R> &I = inst[m];
  const Shape sh = I.sh;
  long n_x, n_y, n_tot;
  ghost_counts(sh, n_x, n_y, n_tot);
  const long t = (long)(blockIdx.x - start[m]) * blockDim.x + threadIdx.x;
  if (t >= n_tot) return;
  ghost_cell<R>(I.st[set], sh, bc, n_x, n_y, t);
}

__global__ void batch_cfl_init_kernel(unsigned *bits, int stride_words, int n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) bits[(long)m * stride_words] = __float_as_uint(1.e10f);
}

// CFL of every instance of the qq set: the blocks of an instance stride over
// its interior; one atomicMin per block into the instance's slot.
template <typename R>
__global__ __launch_bounds__(256) void batch_cfl_kernel(const BatchInst<R> *__restrict__ inst,
                                                        BatchDyn<R> *__restrict__ dyn,
                                                        const int *__restrict__ start, int n,
                                                        double gm, double cfl) {
  const int m = batch_find(start, n, blockIdx.x);
  const BatchInst<R> &I = inst[m];
  const Shape sh = I.sh;
  const Fld<const R> q = [&] {
    Fld<const R> o;
#pragma unroll
    for (int f = 0; f < 9; ++f) o.q[f] = I.st[0].q[f];
    return o;
  }();
  const int nblk = start[m + 1] - start[m];
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nk = sh.Kt - 2 * sh.mk;
  const long ncell = (long)ni * nj * nk;
  unsigned best = __float_as_uint(1.e10f);
  for (long t = (long)(blockIdx.x - start[m]) * blockDim.x + threadIdx.x; t < ncell;
       t += (long)nblk * blockDim.x) {
    const int i = sh.mi + (int)(t % ni);
    const long r = t / ni;
    const int j = sh.mj + (int)(r % nj);
    const int k = sh.mk + (int)(r / nj);
    const R hmin = nmin3(I.dx[i], I.dy[j], I.dz[k]);
    best = min(best, __float_as_uint(cfl_candidate(q, sh.idx(i, j, k), R(gm), R(cfl), hmin)));
  }
  best = wave_umin(best);
  __shared__ unsigned red[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned b = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = min(b, red[w]);
    atomicMin(&dyn[m].cfl_bits, b);
  }
}

template <typename R> __device__ __forceinline__ long long key_bits(R x) {
  if (sizeof(R) == 8) return (long long)__double_as_longlong(double(x));
  return (long long)__float_as_uint(float(x));
}

// Step scalars of every instance from its CFL result, with the host's
// arithmetic: dt = (Real)(float)min, ch = 0.8 cfl min_dxyz / dt evaluated in
// double then rounded, tau = 2 dt, stage steps dt/4, dt/3, dt/2, dt, and
// damp = exp(-dt_s / tau) from the table of host-library values.
template <typename R>
__global__ void batch_params_kernel(const BatchInst<R> *__restrict__ inst,
                                    BatchDyn<R> *__restrict__ dyn, int n, double cfl_number,
                                    DampTable<R> tab) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  BatchDyn<R> &D = dyn[m];
  const R dt = R(__uint_as_float(D.cfl_bits));
  const R cfl = R(cfl_number), hmin = R(inst[m].min_dxyz);
  const R ch_divb = R(0.8 * double(cfl) * double(hmin) / double(dt));
  const R tau = R(2.0 * double(dt));
  const R dts[4] = {R(double(dt) / 4.0), R(double(dt) / 3.0), R(double(dt) / 2.0), dt};
  D.dt = dt;
  D.ch2 = ch_divb * ch_divb;
  const R exact_q[4] = {R(-0.125), R(0), R(-0.25), R(-0.5)};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    D.dts[s] = dts[s];
    const R q = -dts[s] / tau;
    R v;
    bool found = false;
    if (s != 1) {
      if (q == exact_q[s]) { v = tab.exact[s]; found = true; }
    } else {
      const long long off = key_bits(q) - key_bits(tab.q0);
      if (off >= -16 && off <= 16) { v = tab.third[off + 16]; found = true; }
    }
    if (!found) {
      v = R(exp(double(q)));
      D.exp_fallback = 1;
    }
    D.damp[s] = v;
  }
  D.time += double(dt);
  D.steps += 1;
}

template <typename R>
hipError_t launch_batch_ghost(hipStream_t st, const BatchInst<R> *inst, const int *start, int n,
                              int total, int set, const BcDev &bc) {
  hipLaunchKernelGGL((batch_ghost_kernel<R>), dim3(total), dim3(256), 0, st, inst, start, n, set,
                     bc);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_batch_cfl(hipStream_t st, const BatchInst<R> *inst, BatchDyn<R> *dyn,
                            const int *start, int n, int total, double gm, double cfl) {
  static_assert(sizeof(BatchDyn<R>) % sizeof(unsigned) == 0, "slot stride");
  hipLaunchKernelGGL(batch_cfl_init_kernel, dim3((n + 63) / 64), dim3(64), 0, st,
                     &dyn[0].cfl_bits, (int)(sizeof(BatchDyn<R>) / sizeof(unsigned)), n);
  hipLaunchKernelGGL((batch_cfl_kernel<R>), dim3(total), dim3(256), 0, st, inst, dyn, start, n,
                     gm, cfl);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_batch_params(hipStream_t st, const BatchInst<R> *inst, BatchDyn<R> *dyn, int n,
                               double cfl_number, const DampTable<R> &tab) {
  hipLaunchKernelGGL((batch_params_kernel<R>), dim3((n + 63) / 64), dim3(64), 0, st, inst, dyn, n,
                     cfl_number, tab);
  return hipGetLastError();
}

#define MISOHIP_BATCH_INST(R)                                                                     \
  template hipError_t launch_batch_ghost<R>(hipStream_t, const BatchInst<R> *, const int *, int, \
                                            int, int, const BcDev &);                             \
  template hipError_t launch_batch_cfl<R>(hipStream_t, const BatchInst<R> *, BatchDyn<R> *,      \
                                          const int *, int, int, double, double);                 \
  template hipError_t launch_batch_params<R>(hipStream_t, const BatchInst<R> *, BatchDyn<R> *,   \
                                             int, double, const DampTable<R> &);
MISOHIP_BATCH_INST(double)
MISOHIP_BATCH_INST(float)
#undef MISOHIP_BATCH_INST

// ---------------------------------------------------------- transposes
// MISO order (i, j, k) k-fastest  <->  device order x-fastest.  For each j
// this is a 2D transpose of an It x Kt matrix, tiled through LDS.
template <typename R, bool TO_DEVICE>
__global__ __launch_bounds__(256) void transpose_kernel(const R *__restrict__ src,
                                                        R *__restrict__ dst, int It, int Jt,
                                                        int Kt) {
  __shared__ R tile[32][33];
  const int j = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  if (TO_DEVICE) {
    // read along k (contiguous in MISO order), write along i
    const int k0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
    for (int r = ty; r < 32; r += 8) {
      const int i = i0 + r, k = k0 + tx;
      if (i < It && k < Kt) tile[r][tx] = src[((long)i * Jt + j) * Kt + k];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
      const int k = k0 + r, i = i0 + tx;
      if (i < It && k < Kt) dst[((long)k * Jt + j) * It + i] = tile[tx][r];
    }
  } else {
    const int i0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    for (int r = ty; r < 32; r += 8) {
      const int k = k0 + r, i = i0 + tx;
      if (i < It && k < Kt) tile[r][tx] = src[((long)k * Jt + j) * It + i];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
      const int i = i0 + r, k = k0 + tx;
      if (i < It && k < Kt) dst[((long)i * Jt + j) * Kt + k] = tile[tx][r];
    }
  }
}

template <typename R>
hipError_t launch_transpose(hipStream_t st, const R *src, R *dst, int It, int Jt, int Kt,
                            bool to_device) {
  if (to_device) {
    dim3 grid((Kt + 31) / 32, (It + 31) / 32, Jt);
    hipLaunchKernelGGL((transpose_kernel<R, true>), grid, dim3(256), 0, st, src, dst, It, Jt, Kt);
  } else {
    dim3 grid((It + 31) / 32, (Kt + 31) / 32, Jt);
    hipLaunchKernelGGL((transpose_kernel<R, false>), grid, dim3(256), 0, st, src, dst, It, Jt, Kt);
  }
  return hipGetLastError();
}

#define MISOHIP_INST(R)                                                                           \
  template hipError_t launch_ghost_fill<R>(hipStream_t, const Fld<R> &, const Shape &,           \
                                           const BcDev &, int);                                   \
  template hipError_t launch_cfl<R>(hipStream_t, const Fld<const R> &, const R *, const R *,     \
                                    const R *, const Shape &, double, double, unsigned *);        \
  template hipError_t launch_transpose<R>(hipStream_t, const R *, R *, int, int, int, bool);
MISOHIP_INST(double)
MISOHIP_INST(float)
#undef MISOHIP_INST

}  // namespace misohip
