// Grey short-characteristics radiative transfer on the GPU (FP64), the
// counterpart of MISO's CPU-only rt::RT (miso/include/miso/rt_host.hpp).
//
// The sweep along a ray's dominant axis is sequential plane by plane, but
// within a plane every node is independent (its four upwind nodes all lie in
// the previous plane), and the 24 rays are independent.  So one launch
// advances all rays by one plane: rays x transverse nodes of that plane.
// Per-ray constants (dominant axis, sweep order, bilinear weights, ds) are
// computed on the host with the reference's expressions, float narrowing
// included (rt_host.hpp:461-508).
//
// Device layout: rint[(n * Kt + k) * Jt + j) * It + i] (x fastest); kappa
// and S as the MHD fields.  The RT lattice is ib0 = margin - is ...
// ib1 = ib0 + size on every axis (rt_host.hpp:127-154).
// A ray whose dominant axis is x sweeps planes of fixed i, whose nodes lie
// It apart in that layout; such rays ("transposed" rays, R.tr) keep their
// intensities y-fastest instead, rint[(n * Kt + k) * It + i) * Jt + j], and
// read kappa and S from copies in the same layout, so their planes are
// contiguous as well.  Per node the arithmetic does not change.
#include "launch.hpp"

namespace misohip {

__device__ __forceinline__ double rt_integrate(double Iu, double Su, double Sd, double dtau) {
  const double eps_tau = 1.e-4;
  const double exp_dtau = exp(-dtau);
  const double u0 = 1.0 - exp_dtau;
  const double u1 = dtau < eps_tau ? 0.5 * dtau : 1.0 - (u0 / dtau);
  return Iu * exp_dtau + (u0 - u1) * Su + u1 * Sd;
}

// plane p of every ray of `rays` (blockIdx.y = entry, R.id = its slot)
__global__ __launch_bounds__(256) void rt_sweep_plane_kernel(double *__restrict__ I,
                                                             const double *__restrict__ kap_x,
                                                             const double *__restrict__ S_x,
                                                             const double *__restrict__ kap_y,
                                                             const double *__restrict__ S_y,
                                                             RtRays rays, int p, long ray_stride) {
  const RtRay &R = rays.r[blockIdx.y];
  if (p >= R.nplanes) return;
  const double *__restrict__ kap = R.tr ? kap_y : kap_x;
  const double *__restrict__ S = R.tr ? S_y : S_x;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.nb * R.nc) return;
  const int ib = t % R.nb, ic = t / R.nb;
  const int c = R.start + p * R.da + ib * R.db + ic * R.dc;  // cell index (kappa, S)
  const int c00 = c - R.ua;
  const int c10 = c00 - R.ub, c01 = c00 - R.uc, c11 = c00 - R.ub - R.uc;
  double *Ir = I + ray_stride * R.id;
  const double iu = R.w00 * Ir[c00] + R.w10 * Ir[c10] + R.w01 * Ir[c01] + R.w11 * Ir[c11];
  const double su = R.w00 * S[c00] + R.w10 * S[c10] + R.w01 * S[c01] + R.w11 * S[c11];
  const double ku = R.w00 * kap[c00] + R.w10 * kap[c10] + R.w01 * kap[c01] + R.w11 * kap[c11];
  const double dtau = 0.5 * (ku + kap[c]) * R.ds;
  Ir[c] = rt_integrate(iu, su, S[c], dtau);
}

// periodic halo self-exchange along `axis` over the full transverse extent
// (rt_host.hpp:236-452): + rays copy node b1 into b0, - rays b0 into b1.
// rays outside ray_mask are left alone (the other layout's launch does them)
__global__ __launch_bounds__(256) void rt_halo_kernel(double *__restrict__ I, int axis, int b0,
                                                      int b1, int It, int Jt, int Kt,
                                                      unsigned pos_mask, unsigned ray_mask,
                                                      long ray_stride, int nrays) {
  const int n1 = axis == 0 ? Jt : It, n2 = axis == 2 ? Jt : Kt;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n1 * n2 * nrays) return;
  const int u = (int)(t % n1);
  const long r = t / n1;
  const int v = (int)(r % n2), n = (int)(r / n2);
  if (!((ray_mask >> n) & 1u)) return;
  int i, j, k;
  if (axis == 0) { j = u; k = v; i = 0; }
  else if (axis == 1) { i = u; k = v; j = 0; }
  else { i = u; j = v; k = 0; }
  const bool pos = (pos_mask >> n) & 1u;
  int s[3] = {i, j, k}, d[3] = {i, j, k};
  s[axis] = pos ? b1 : b0;
  d[axis] = pos ? b0 : b1;
  double *Ir = I + ray_stride * n;
  Ir[(d[2] * Jt + d[1]) * It + d[0]] = Ir[(s[2] * Jt + s[1]) * It + s[0]];
}

// incoming boundary on one physical face (rt_host.hpp:50-89): nodes of the
// face layer over the lattice's transverse range, rays in `in_mask`.
__global__ __launch_bounds__(256) void rt_face_kernel(double *__restrict__ I,
                                                      const double *__restrict__ S, int axis,
                                                      int layer, int lo0, int lo1, int n0, int n1,
                                                      int It, int Jt, int kind, double value,
                                                      unsigned in_mask, long ray_stride,
                                                      int nrays) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n0 * n1 * nrays) return;
  const int u = lo0 + (int)(t % n0);
  const long r = t / n0;
  const int v = lo1 + (int)(r % n1), n = (int)(r / n1);
  if (!((in_mask >> n) & 1u)) return;
  int i, j, k;
  if (axis == 0) { i = layer; j = u; k = v; }
  else if (axis == 1) { j = layer; i = u; k = v; }
  else { k = layer; i = u; j = v; }
  const int c = (k * Jt + j) * It + i;
  I[ray_stride * n + c] = kind == 1 ? S[c] : (kind == 2 ? 0.0 : value);
}

// max over lattice nodes of |I - old|, float-narrowed as util::max2 does
// (rt_host.hpp:612-629); positive floats order like their bit patterns.
__global__ __launch_bounds__(256) void rt_maxdiff_kernel(const double *__restrict__ I,
                                                         const double *__restrict__ old, int lo_i,
                                                         int lo_j, int lo_k, int ni, int nj,
                                                         int nk, int It, int Jt, long ray_stride,
                                                         int nrays, unsigned ray_mask,
                                                         unsigned *slot) {
  const long n_all = (long)ni * nj * nk * nrays;
  unsigned best = 0u;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n_all;
       t += (long)gridDim.x * blockDim.x) {
    const int i = lo_i + (int)(t % ni);
    long r = t / ni;
    const int j = lo_j + (int)(r % nj);
    r /= nj;
    const int k = lo_k + (int)(r % nk);
    const int n = (int)(r / nk);
    if (!((ray_mask >> n) & 1u)) continue;
    const long c = ray_stride * n + (long)(k * Jt + j) * It + i;
    const float d = (float)fabs(I[c] - old[c]);
    best = max(best, __float_as_uint(d));
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) best = max(best, (unsigned)__shfl_xor((int)best, o, 64));
  __shared__ unsigned red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned b = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = max(b, red[w]);
    atomicMax(slot, b);
  }
}

__global__ void rt_slot_clear_kernel(unsigned *slot) {
  if (threadIdx.x == 0) slot[0] = 0u;
}

// Q_rad = 4 pi kappa (J - S), J = sum_r w_r I_r in ray order, lattice nodes
// (a ray in tr_mask holds its intensities y-fastest)
__global__ __launch_bounds__(256) void rt_qrad_kernel(const double *__restrict__ I,
                                                      const double *__restrict__ kap,
                                                      const double *__restrict__ S,
                                                      double *__restrict__ q, int lo_i, int lo_j,
                                                      int lo_k, int ni, int nj, int nk, int It,
                                                      int Jt, long ray_stride, int nrays,
                                                      unsigned tr_mask, RtRays w) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ni * nj * nk) return;
  const int i = lo_i + (int)(t % ni);
  const long r = t / ni;
  const int j = lo_j + (int)(r % nj), k = lo_k + (int)(r / nj);
  const int c = (k * Jt + j) * It + i, ct = (k * It + i) * Jt + j;
  double J = 0.0;
  for (int n = 0; n < nrays; ++n)
    J += w.r[n].wq * I[ray_stride * n + (((tr_mask >> n) & 1u) ? ct : c)];
  q[c] = 4.0 * 3.14159265358979323846 * kap[c] * (J - S[c]);
}

hipError_t launch_rt_sweep_plane(hipStream_t st, double *I, const double *kap, const double *S,
                                 const double *kap_y, const double *S_y, const RtRays &rays,
                                 int nrays, int p, long ray_stride, int max_nodes) {
  dim3 grid((max_nodes + 255) / 256, nrays);
  hipLaunchKernelGGL(rt_sweep_plane_kernel, grid, dim3(256), 0, st, I, kap, S, kap_y, S_y, rays,
                     p, ray_stride);
  return hipGetLastError();
}

// (Kt, Jt, It) x-fastest -> (Kt, It, Jt) y-fastest, 32 x 32 tiles through LDS
__global__ __launch_bounds__(256) void rt_transpose_xy_kernel(const double *__restrict__ src,
                                                              double *__restrict__ dst, int It,
                                                              int Jt, int tiles_i) {
  __shared__ double tile[32][33];
  const int ti = blockIdx.x % tiles_i, tj = blockIdx.x / tiles_i, k = blockIdx.y;
  const long plane = (long)It * Jt * k;
  for (int r = threadIdx.x >> 5; r < 32; r += 8) {
    const int i = ti * 32 + (threadIdx.x & 31), j = tj * 32 + r;
    if (i < It && j < Jt) tile[r][threadIdx.x & 31] = src[plane + (long)j * It + i];
  }
  __syncthreads();
  for (int r = threadIdx.x >> 5; r < 32; r += 8) {
    const int j = tj * 32 + (threadIdx.x & 31), i = ti * 32 + r;
    if (i < It && j < Jt) dst[plane + (long)i * Jt + j] = tile[threadIdx.x & 31][r];
  }
}

hipError_t launch_rt_transpose_xy(hipStream_t st, const double *src, double *dst, int It, int Jt,
                                  int Kt) {
  const int ti = (It + 31) / 32, tj = (Jt + 31) / 32;
  hipLaunchKernelGGL(rt_transpose_xy_kernel, dim3(ti * tj, Kt), dim3(256), 0, st, src, dst, It,
                     Jt, ti);
  return hipGetLastError();
}

hipError_t launch_rt_halo(hipStream_t st, double *I, int axis, int b0, int b1, int It, int Jt,
                          int Kt, unsigned pos_mask, unsigned ray_mask, long ray_stride,
                          int nrays) {
  const int n1 = axis == 0 ? Jt : It, n2 = axis == 2 ? Jt : Kt;
  const long n = (long)n1 * n2 * nrays;
  hipLaunchKernelGGL(rt_halo_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, I, axis,
                     b0, b1, It, Jt, Kt, pos_mask, ray_mask, ray_stride, nrays);
  return hipGetLastError();
}

hipError_t launch_rt_face(hipStream_t st, double *I, const double *S, int axis, int layer,
                          int lo0, int lo1, int n0, int n1, int It, int Jt, int kind, double value,
                          unsigned in_mask, long ray_stride, int nrays) {
  const long n = (long)n0 * n1 * nrays;
  hipLaunchKernelGGL(rt_face_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, I, S,
                     axis, layer, lo0, lo1, n0, n1, It, Jt, kind, value, in_mask, ray_stride,
                     nrays);
  return hipGetLastError();
}

hipError_t launch_rt_maxdiff(hipStream_t st, const double *I, const double *old, const int lo[3],
                             const int n[3], int It, int Jt, long ray_stride, int nrays,
                             unsigned ray_mask, bool clear, unsigned *slot) {
  if (clear) hipLaunchKernelGGL(rt_slot_clear_kernel, dim3(1), dim3(64), 0, st, slot);
  hipLaunchKernelGGL(rt_maxdiff_kernel, dim3(2048), dim3(256), 0, st, I, old, lo[0], lo[1], lo[2],
                     n[0], n[1], n[2], It, Jt, ray_stride, nrays, ray_mask, slot);
  return hipGetLastError();
}

hipError_t launch_rt_qrad(hipStream_t st, const double *I, const double *kap, const double *S,
                          double *q, const int lo[3], const int n[3], int It, int Jt,
                          long ray_stride, int nrays, unsigned tr_mask, const RtRays &w) {
  const long cnt = (long)n[0] * n[1] * n[2];
  hipLaunchKernelGGL(rt_qrad_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, I,
                     kap, S, q, lo[0], lo[1], lo[2], n[0], n[1], n[2], It, Jt, ray_stride, nrays,
                     tr_mask, w);
  return hipGetLastError();
}

}  // namespace misohip

namespace misohip {

// T = (gm-1) ei, kappa = kappa0 ro T^3, S = s_rad T^4 on every cell
__global__ __launch_bounds__(256) void rt_source_kernel(const double *__restrict__ ro,
                                                        const double *__restrict__ ei,
                                                        double *__restrict__ kap,
                                                        double *__restrict__ S, long n, double gm,
                                                        double kappa0, double s_rad) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double T = (gm - 1.0) * ei[c];
  const double T2 = T * T;
  kap[c] = kappa0 * ro[c] * (T2 * T);
  S[c] = s_rad * (T2 * T2);
}

// ei += dt * q / ro on the interior
__global__ __launch_bounds__(256) void rt_heat_kernel(const double *__restrict__ ro,
                                                      double *__restrict__ ei,
                                                      const double *__restrict__ q, Shape sh,
                                                      double dt) {
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nk = sh.Kt - 2 * sh.mk;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ni * nj * nk) return;
  const int i = sh.mi + (int)(t % ni);
  const long r = t / ni;
  const int j = sh.mj + (int)(r % nj), k = sh.mk + (int)(r / nj);
  const int c = sh.idx(i, j, k);
  ei[c] = ei[c] + dt * q[c] / ro[c];
}

hipError_t launch_rt_source(hipStream_t st, const double *ro, const double *ei, double *kap,
                            double *S, long n, double gm, double kappa0, double s_rad) {
  hipLaunchKernelGGL(rt_source_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ro, ei,
                     kap, S, n, gm, kappa0, s_rad);
  return hipGetLastError();
}

hipError_t launch_rt_heat(hipStream_t st, const double *ro, double *ei, const double *q,
                          const Shape &sh, double dt) {
  const long n = (long)(sh.It - 2 * sh.mi) * (sh.Jt - 2 * sh.mj) * (sh.Kt - 2 * sh.mk);
  hipLaunchKernelGGL(rt_heat_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ro, ei,
                     q, sh, dt);
  return hipGetLastError();
}

}  // namespace misohip
