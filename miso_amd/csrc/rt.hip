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

// one node from its four upwind neighbours (00: same transverse position,
// 10: one step back along b, 01: along c, 11: both) and its own kappa, S
__device__ __forceinline__ double rt_node(const RtRay &R, double I00, double I10, double I01,
                                          double I11, double S00, double S10, double S01,
                                          double S11, double K00, double K10, double K01,
                                          double K11, double kc, double sc) {
  const double iu = R.w00 * I00 + R.w10 * I10 + R.w01 * I01 + R.w11 * I11;
  const double su = R.w00 * S00 + R.w10 * S10 + R.w01 * S01 + R.w11 * S11;
  const double ku = R.w00 * K00 + R.w10 * K10 + R.w01 * K01 + R.w11 * K11;
  const double dtau = 0.5 * (ku + kc) * R.ds;
  return rt_integrate(iu, su, sc, dtau);
}

// The solve's convergence measure is the max over lattice nodes of
// |I_end - I_start| of an iteration, float-narrowed as util::max2 does
// (rt_host.hpp:612-629); positive floats order like their bit patterns.
// Every lattice node of a ray is either swept (exactly once per iteration,
// so the sweep takes the change from the value it overwrites) or on one of
// the ray's upwind faces (written only by the halo and face kernels; their
// start values are saved first and compared after).  Workgroup maxima go
// by atomicMax into one of RT_PARTS slots, 32 words apart, then one small
// kernel reduces them.
__device__ __forceinline__ unsigned rt_change(double now, double before) {
  return __float_as_uint((float)fabs(now - before));
}

__device__ __forceinline__ void rt_block_max(unsigned best, unsigned *part) {
  __shared__ unsigned red[16];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) best = max(best, (unsigned)__shfl_xor((int)best, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned b = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = max(b, red[w]);
    if (b) atomicMax(part + 32 * ((blockIdx.x + 7 * blockIdx.y) % RT_PARTS), b);
  }
}

// plane p of every ray of `rays` (blockIdx.y = entry, R.id = its slot)
__global__ __launch_bounds__(256) void rt_sweep_plane_kernel(double *__restrict__ I,
                                                             const double *__restrict__ kap_x,
                                                             const double *__restrict__ S_x,
                                                             const double *__restrict__ kap_y,
                                                             const double *__restrict__ S_y,
                                                             RtRays rays, int p, long ray_stride,
                                                             unsigned *part) {
  const RtRay &R = rays.r[blockIdx.y];
  if (p >= R.nplanes) return;  // uniform over the block
  const double *__restrict__ kap = R.tr ? kap_y : kap_x;
  const double *__restrict__ S = R.tr ? S_y : S_x;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned best = 0u;
  if (t < R.nb * R.nc) {
    const int ib = t % R.nb, ic = t / R.nb;
    const int c = R.start + p * R.da + ib * R.db + ic * R.dc;  // cell index (kappa, S)
    const int c00 = c - R.ua;
    const int c10 = c00 - R.ub, c01 = c00 - R.uc, c11 = c00 - R.ub - R.uc;
    double *Ir = I + ray_stride * R.id;
    const double before = Ir[c];
    const double now = rt_node(R, Ir[c00], Ir[c10], Ir[c01], Ir[c11], S[c00], S[c10], S[c01],
                               S[c11], kap[c00], kap[c10], kap[c01], kap[c11], kap[c], S[c]);
    Ir[c] = now;
    best = rt_change(now, before);
  }
  rt_block_max(best, part);
}

// Planes p0 .. p0 + P - 1 of every ray of `rays` in one launch, for rays
// whose c axis is short (nc < RT_MULTI_ROWS; a thin slab's x and y rays).
// Node coordinates are mirrored so that upwind is always one step back:
// u = ib (sb > 0) or nb - 1 - ib, likewise v.  A workgroup owns the
// RT_MULTI_COLS columns u in [U0, U0 + W) of a ray over the whole c range,
// and computes plane by plane the cone u in [U0 - P, U0 + W) of nodes
// that its last plane depends on; the plane is kept in LDS, column u = -1
// and row v = -1 (the upwind boundary nodes, which the sweep never writes)
// are read from memory.  Every node is computed by rt_node from the same
// values as in rt_sweep_plane_kernel, so the results are the same bits;
// cone nodes outside the owned columns are recomputed, not stored, and the
// cone's stale edge (u below U0 - P + s + 1 at step s) never reaches them.
constexpr int RT_MULTI_THREADS = 1024;

__global__ __launch_bounds__(RT_MULTI_THREADS) void rt_sweep_multi_kernel(double *__restrict__ I,
                                                             const double *__restrict__ kap_x,
                                                             const double *__restrict__ S_x,
                                                             const double *__restrict__ kap_y,
                                                             const double *__restrict__ S_y,
                                                             RtRays rays, int p0, int P,
                                                             long ray_stride, unsigned *part) {
  constexpr int W = RT_MULTI_COLS, X = RT_MULTI_COLS + RT_MULTI_PLANES, ROWS = RT_MULTI_ROWS;
  constexpr int NT = RT_MULTI_THREADS, CPT = (X * ROWS + NT - 1) / NT;
  __shared__ double Il[ROWS][X], Kl[ROWS][X], Sl[ROWS][X];
  const RtRay &R = rays.r[blockIdx.y];
  const int U0 = blockIdx.x * W;
  // uniform over the workgroup
  if (p0 >= R.nplanes || U0 >= R.nb || R.nc + 1 > ROWS || P < 1 || P > RT_MULTI_PLANES) return;
  const double *__restrict__ kap = R.tr ? kap_y : kap_x;
  const double *__restrict__ S = R.tr ? S_y : S_x;
  double *Ir = I + ray_stride * R.id;
  const int ext = W + P, ncell = ext * (R.nc + 1);
  const int pend = min(p0 + P, R.nplanes);
  // per cell: memory index at plane 0, role (0 none, 1 boundary, read every
  // plane; 2 computed; 3 computed and owned, its value before the sweep read
  // for the change; 4 the cone's far edge e = 0,
  // read for the plane before p0 only: never valid after it, and its node
  // may be written by another workgroup meanwhile), LDS position
  int off[CPT], role[CPT], ce[CPT], cr[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int idx = threadIdx.x + q * NT;
    const int e = idx % ext, r = idx / ext;
    const int u = U0 - P + e, v = r - 1;
    ce[q] = e;
    cr[q] = r;
    role[q] = 0;
    off[q] = 0;
    if (idx >= ncell || u < -1 || u >= R.nb) continue;
    const int ib = R.sb > 0 ? u : R.nb - 1 - u;
    const int ic = R.sc > 0 ? v : R.nc - 1 - v;
    off[q] = R.start + ib * R.db + ic * R.dc;
    if (u == -1 || v == -1) role[q] = 1;
    else if (e == 0) role[q] = 4;
    else role[q] = (u >= U0 && u < U0 + W) ? 3 : 2;
  }
  // the plane before p0
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (role[q]) {
      const int c = off[q] + (p0 - 1) * R.da;
      Il[cr[q]][ce[q]] = Ir[c];
      Kl[cr[q]][ce[q]] = kap[c];
      Sl[cr[q]][ce[q]] = S[c];
    }
  // kappa, S and the boundary (role 1) or before-sweep (role 3) intensity of
  // a plane's cells; loads run two planes ahead of the computation
  auto load = [&](int p, double *k, double *s, double *b) {
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      k[q] = s[q] = b[q] = 0.0;
      if (p < pend && role[q] && role[q] < 4) {
        const int c = off[q] + p * R.da;
        k[q] = kap[c];
        s[q] = S[c];
        if (role[q] == 1 || role[q] == 3) b[q] = Ir[c];
      }
    }
  };
  double kc[CPT], sc[CPT], bnd[CPT], k1[CPT], s1[CPT], b1[CPT];
  load(p0, kc, sc, bnd);
  load(p0 + 1, k1, s1, b1);
  __syncthreads();
  unsigned best = 0u;
  for (int p = p0; p < pend; ++p) {
    double kn[CPT], sn[CPT], bn[CPT];
    load(p + 2, kn, sn, bn);
    double in[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      in[q] = bnd[q];
      if (role[q] == 2 || role[q] == 3) {
        const int r = cr[q], e = ce[q];
        in[q] = rt_node(R, Il[r][e], Il[r][e - 1], Il[r - 1][e], Il[r - 1][e - 1], Sl[r][e],
                        Sl[r][e - 1], Sl[r - 1][e], Sl[r - 1][e - 1], Kl[r][e], Kl[r][e - 1],
                        Kl[r - 1][e], Kl[r - 1][e - 1], kc[q], sc[q]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (role[q] && role[q] < 4) {
        Il[cr[q]][ce[q]] = in[q];
        Kl[cr[q]][ce[q]] = kc[q];
        Sl[cr[q]][ce[q]] = sc[q];
        if (role[q] == 3) {
          best = max(best, rt_change(in[q], bnd[q]));
          Ir[off[q] + p * R.da] = in[q];
        }
      }
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      kc[q] = k1[q];
      sc[q] = s1[q];
      bnd[q] = b1[q];
      k1[q] = kn[q];
      s1[q] = sn[q];
      b1[q] = bn[q];
    }
    __syncthreads();
  }
  rt_block_max(best, part);
}

hipError_t launch_rt_sweep_multi(hipStream_t st, double *I, const double *kap, const double *S,
                                 const double *kap_y, const double *S_y, const RtRays &rays,
                                 int nrays, int p0, int P, long ray_stride, int max_nb,
                                 unsigned *part) {
  dim3 grid((max_nb + RT_MULTI_COLS - 1) / RT_MULTI_COLS, nrays);
  hipLaunchKernelGGL(rt_sweep_multi_kernel, grid, dim3(RT_MULTI_THREADS), 0, st, I, kap, S, kap_y,
                     S_y, rays, p0, P, ray_stride, part);
  return hipGetLastError();
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

// The upwind faces of every ray over the lattice (axes with more than one
// node; blockIdx.y = ray * 3 + axis): the layer x_a = lo_a (ray going +) or
// lo_a + n_a - 1, all lattice nodes of the other two axes.  save: old = I
// there; otherwise the change of I against old, into the parts.
__global__ __launch_bounds__(256) void rt_faces_kernel(const double *__restrict__ I,
                                                       double *__restrict__ old, int save,
                                                       RtFaces f, long ray_stride,
                                                       unsigned *part) {
  const int n = blockIdx.y / 3, a = blockIdx.y % 3;
  if (f.n[a] <= 1) return;  // uniform over the block
  const int t0 = a == 0 ? 1 : 0, t1 = a == 2 ? 1 : 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned best = 0u;
  if (t < f.n[t0] * f.n[t1]) {
    int x[3];
    x[a] = f.lo[a] + (((f.pos[a] >> n) & 1u) ? 0 : f.n[a] - 1);
    x[t0] = f.lo[t0] + t % f.n[t0];
    x[t1] = f.lo[t1] + t / f.n[t0];
    const long c = ray_stride * n + (((f.tr >> n) & 1u) ? ((long)x[2] * f.It + x[0]) * f.Jt + x[1]
                                                         : ((long)x[2] * f.Jt + x[1]) * f.It + x[0]);
    if (save) old[c] = I[c];
    else best = rt_change(I[c], old[c]);
  }
  if (!save) rt_block_max(best, part);
}

// slot = max of the parts; the parts are cleared for the next iteration
__global__ __launch_bounds__(64) void rt_parts_reduce_kernel(unsigned *part, unsigned *slot,
                                                             int clear_only) {
  unsigned b = 0u;
  for (int s = threadIdx.x; s < RT_PARTS; s += 64) {
    b = max(b, part[32 * s]);
    part[32 * s] = 0u;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) b = max(b, (unsigned)__shfl_xor((int)b, o, 64));
  if (threadIdx.x == 0 && !clear_only) slot[0] = b;
}

// Q_rad = 4 pi kappa (J - S), J = sum_r w_r I_r in ray order, lattice nodes.
// A block covers 32 x 32 (i, j) nodes of one plane k; a ray kept y-fastest
// (tr_mask) is read along j into LDS and taken from there, the others along
// i, so every read is coalesced and each node sums its rays in order.
__global__ __launch_bounds__(256) void rt_qrad_kernel(const double *__restrict__ I,
                                                      const double *__restrict__ kap,
                                                      const double *__restrict__ S,
                                                      double *__restrict__ q, int lo_i, int lo_j,
                                                      int lo_k, int ni, int nj, int It, int Jt,
                                                      long ray_stride, int nrays, int tiles_i,
                                                      unsigned tr_mask, RtRays w) {
  __shared__ double tile[32][33];
  const int ti = blockIdx.x % tiles_i, tj = blockIdx.x / tiles_i, k = lo_k + blockIdx.y;
  const int l = threadIdx.x & 31, r0 = threadIdx.x >> 5;
  double J[4] = {0.0, 0.0, 0.0, 0.0};
  for (int n = 0; n < nrays; ++n) {
    const double *In = I + ray_stride * n;
    const double wq = w.r[n].wq;
    if ((tr_mask >> n) & 1u) {  // uniform over the block
      __syncthreads();
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int il = r0 + 8 * m, i = ti * 32 + il, j = tj * 32 + l;
        if (i < ni && j < nj) tile[il][l] = In[((long)k * It + lo_i + i) * Jt + lo_j + j];
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < 4; ++m) J[m] += wq * tile[l][r0 + 8 * m];
    } else {
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int i = ti * 32 + l, j = tj * 32 + r0 + 8 * m;
        if (i < ni && j < nj) J[m] += wq * In[((long)k * Jt + lo_j + j) * It + lo_i + i];
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int i = ti * 32 + l, j = tj * 32 + r0 + 8 * m;
    if (i < ni && j < nj) {
      const long c = ((long)k * Jt + lo_j + j) * It + lo_i + i;
      q[c] = 4.0 * 3.14159265358979323846 * kap[c] * (J[m] - S[c]);
    }
  }
}

hipError_t launch_rt_sweep_plane(hipStream_t st, double *I, const double *kap, const double *S,
                                 const double *kap_y, const double *S_y, const RtRays &rays,
                                 int nrays, int p, long ray_stride, int max_nodes,
                                 unsigned *part) {
  dim3 grid((max_nodes + 255) / 256, nrays);
  hipLaunchKernelGGL(rt_sweep_plane_kernel, grid, dim3(256), 0, st, I, kap, S, kap_y, S_y, rays,
                     p, ray_stride, part);
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

hipError_t launch_rt_faces(hipStream_t st, const double *I, double *old, bool save,
                          const RtFaces &f, long ray_stride, int nrays, unsigned *part) {
  int m = 1;
  for (int a = 0; a < 3; ++a) {
    const int t0 = a == 0 ? 1 : 0, t1 = a == 2 ? 1 : 2;
    m = max(m, f.n[t0] * f.n[t1]);
  }
  hipLaunchKernelGGL(rt_faces_kernel, dim3((m + 255) / 256, 3 * nrays), dim3(256), 0, st, I, old,
                     save ? 1 : 0, f, ray_stride, part);
  return hipGetLastError();
}

hipError_t launch_rt_parts_reduce(hipStream_t st, unsigned *part, unsigned *slot,
                                  bool clear_only) {
  hipLaunchKernelGGL(rt_parts_reduce_kernel, dim3(1), dim3(64), 0, st, part, slot,
                     clear_only ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_rt_qrad(hipStream_t st, const double *I, const double *kap, const double *S,
                          double *q, const int lo[3], const int n[3], int It, int Jt,
                          long ray_stride, int nrays, unsigned tr_mask, const RtRays &w) {
  const int ti = (n[0] + 31) / 32, tj = (n[1] + 31) / 32;
  hipLaunchKernelGGL(rt_qrad_kernel, dim3(ti * tj, n[2]), dim3(256), 0, st, I, kap, S, q, lo[0],
                     lo[1], lo[2], n[0], n[1], It, Jt, ray_stride, nrays, ti, tr_mask, w);
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
