// Per-column MLP surrogate of the radiative heating: BF16 operands on the
// CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16), FP32 accumulation, FP32
// master weights.  No counterpart in the reference (it has no MHD-RT
// coupling at all).
//
// Network: x (3*nz features of one column) -> 256 -> 256 -> 256 -> nz,
// SiLU on the hidden layers.  Weights are stored [out][in] (row-major), so
// a forward layer is C = X * W^T and the weight gradient is dW = dY^T * X.
//
// GEMM tile: 128 x 128 outputs per 256-thread block, each wave a 64 x 64
// quadrant (2 x 2 MFMA 32x32 tiles), K staged through LDS 32 at a time.
// MFMA 32x32x16 lane layout: lane l holds A[l%32][8*(l/32) .. +7],
// B[8*(l/32) .. +7][l%32], and accumulator v holds
// D[8*(v/4) + 4*(l/32) + v%4][l%32].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.hpp"

namespace misohip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;  // LDS row pitch in bf16

// C[m][n] = sum_k A(m,k) B(k,n) over the block's K range.
//   TA = false: A stored [M][lda] (A(m,k) at m*lda+k);  true: [K][lda] (k*lda+m)
//   TB = true : B stored [N][ldb] (B(k,n) at n*ldb+k);  false: [K][ldb] (k*ldb+n)
// K must be a multiple of 32 within each split; M, N arbitrary.
// EPI 0: raw fp32 to C (+ split offset);  1: pre = acc + bias -> C, and
// bf16(silu(pre)) -> H;  2: acc + bias -> C (output layer);  3: as 1
// without C (inference: the pre-activations are kept only for training).
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const uint16_t *__restrict__ A,
                                                        const uint16_t *__restrict__ B,
                                                        float *__restrict__ C,
                                                        uint16_t *__restrict__ H,
                                                        const float *__restrict__ bias, int M,
                                                        int N, int K, int lda, int ldb, int ldc,
                                                        int k_per_split) {
  __shared__ uint16_t As[BM * LDK];
  __shared__ uint16_t Bs[BN * LDK];
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kb = blockIdx.z * k_per_split;
  const int ke = min(K, kb + k_per_split);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

  for (int k0 = kb; k0 < ke; k0 += BK) {
    // ---- stage A tile (BM x BK) as As[m][k]
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int c = tid + 256 * r;
      if (!TA) {
        const int row = c >> 2, kc = (c & 3) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (m0 + row < M) v = *reinterpret_cast<const uint4 *>(A + (long)(m0 + row) * lda + k0 + kc);
        *reinterpret_cast<uint4 *>(&As[row * LDK + kc]) = v;
      } else {
        const int kr = c >> 4, mc = (c & 15) * 8;
        uint16_t e[8];
        if (m0 + mc + 7 < M) {
          uint4 v = *reinterpret_cast<const uint4 *>(A + (long)(k0 + kr) * lda + m0 + mc);
          const uint16_t *p = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
          for (int q = 0; q < 8; ++q) e[q] = p[q];
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q)
            e[q] = (m0 + mc + q < M) ? A[(long)(k0 + kr) * lda + m0 + mc + q] : (uint16_t)0;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) As[(mc + q) * LDK + kr] = e[q];
      }
    }
    // ---- stage B tile (BK x BN) as Bs[n][k]
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int c = tid + 256 * r;
      if (TB) {
        const int col = c >> 2, kc = (c & 3) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (n0 + col < N) v = *reinterpret_cast<const uint4 *>(B + (long)(n0 + col) * ldb + k0 + kc);
        *reinterpret_cast<uint4 *>(&Bs[col * LDK + kc]) = v;
      } else {
        const int kr = c >> 4, nc = (c & 15) * 8;
        uint16_t e[8];
        if (n0 + nc + 7 < N) {
          uint4 v = *reinterpret_cast<const uint4 *>(B + (long)(k0 + kr) * ldb + n0 + nc);
          const uint16_t *p = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
          for (int q = 0; q < 8; ++q) e[q] = p[q];
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q)
            e[q] = (n0 + nc + q < N) ? B[(long)(k0 + kr) * ldb + n0 + nc + q] : (uint16_t)0;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) Bs[(nc + q) * LDK + kr] = e[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      const int kk = ks + 8 * (lane >> 5);
      bf16x8 af[2], bfv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
        af[a] = *reinterpret_cast<const bf16x8 *>(&As[(wm + 32 * a + (lane & 31)) * LDK + kk]);
#pragma unroll
      for (int b = 0; b < 2; ++b)
        bfv[b] = *reinterpret_cast<const bf16x8 *>(&Bs[(wn + 32 * b + (lane & 31)) * LDK + kk]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[a], bfv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }

  // ---- epilogue
  float *Cs = C + (EPI == 0 ? (long)blockIdx.z * M * ldc : 0);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int m = m0 + wm + 32 * a + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int n = n0 + wn + 32 * b + (lane & 31);
        if (m >= M || n >= N) continue;
        float x = acc[a][b][v];
        if (EPI >= 1) x += bias[n];
        if (EPI != 3) Cs[(long)m * ldc + n] = x;
        if (EPI == 1 || EPI == 3) {
          const float s = x / (1.f + expf(-x));
          H[(long)m * ldc + n] = f2bf(s);
        }
      }
}

// sum of split-K partials (fixed order, deterministic)
__global__ void splitk_reduce_kernel(const float *__restrict__ part, float *__restrict__ out,
                                     long n, int splits) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[(long)z * n + t];
  out[t] = s;
}

// dpre = dact * silu'(pre); dact is fp32 (loss grad) or accumulated; out bf16 + fp32
__global__ void silu_backward_kernel(const float *__restrict__ dact, const float *__restrict__ pre,
                                     uint16_t *__restrict__ dpre_bf, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float x = pre[t];
  const float sg = 1.f / (1.f + expf(-x));
  dpre_bf[t] = f2bf(dact[t] * sg * (1.f + x * (1.f - sg)));
}

// column sums db[n] = sum_m d[m][n]: 64 columns per block along the lanes,
// 16 row groups (rows g, g+16, ...) summed in a fixed order (deterministic)
__global__ __launch_bounds__(1024) void colsum_bf16_kernel(const uint16_t *__restrict__ d,
                                                           float *__restrict__ db, int M, int N) {
  __shared__ float part[16][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + lane;
  float s = 0.f;
  if (n < N)
    for (int m = g; m < M; m += 16) s += bf2f(d[(long)m * N + n]);
  part[g][lane] = s;
  __syncthreads();
  if (g == 0 && n < N) {
    float t = part[0][lane];
#pragma unroll
    for (int r = 1; r < 16; ++r) t += part[r][lane];
    db[n] = t;
  }
}

// MSE loss on a batch: dY = 2 (Y - T) / (M N) (bf16), per-block partial losses
__global__ void mse_kernel(const float *__restrict__ Y, const float *__restrict__ T,
                           uint16_t *__restrict__ dY, float *__restrict__ part, long n,
                           float scale) {
  __shared__ float red[256];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (t < n) {
    const float d = Y[t] - T[t];
    dY[t] = f2bf(2.f * d * scale);
    e = d * d;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam on fp32 master weights, then the bf16 copy used by the GEMMs
__global__ void adam_kernel(float *__restrict__ w, uint16_t *__restrict__ wb,
                            const float *__restrict__ g, float *__restrict__ m1,
                            float *__restrict__ m2, long n, float lr, float b1, float b2,
                            float eps, float c1, float c2, float wd) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gr = g[t] + wd * w[t];
  const float a = b1 * m1[t] + (1.f - b1) * gr;
  const float b = b2 * m2[t] + (1.f - b2) * gr * gr;
  m1[t] = a;
  m2[t] = b;
  const float upd = (a / c1) / (sqrtf(b / c2) + eps);
  const float nw = w[t] - lr * upd;
  w[t] = nw;
  if (wb) wb[t] = f2bf(nw);
}

__global__ void to_bf16_kernel(const float *__restrict__ x, uint16_t *__restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = f2bf(x[t]);
}

// Column features of the radiative box: for column (i, j) and cell kz of the
// interior, x[col][f*nz + kz] = (feat_f - mu[f*nz+kz]) / sd[f*nz+kz] with
// feat = (ln ro, ln T, |B|^2), T = (gm-1) ei.  Columns run i fastest.
// Column features, tiled through LDS: a block reads 64 consecutive x cells
// (one y row) x 32 heights with lanes along x (coalesced field reads),
// then writes each column's 32 consecutive heights (coalesced row writes of
// X).  nfeat == 4 adds the horizontal Laplacian of ln T at stride h (x and y
// periodic: neighbours wrap inside the interior).
constexpr int FT_X = 64, FT_Z = 32;

__global__ __launch_bounds__(256) void column_features_kernel(
    Fld<const double> q, Shape sh, double gm, const float *__restrict__ mu,
    const float *__restrict__ sd, uint16_t *__restrict__ X, float *__restrict__ Xf, int nfeat,
    int h) {
  __shared__ float tile[4][FT_Z][FT_X + 1];
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nz = sh.Kt - 2 * sh.mk;
  const int i0 = blockIdx.x * FT_X, j = blockIdx.y, kz0 = blockIdx.z * FT_Z;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = i0 + lane;
  for (int r = wv; r < FT_Z; r += 4) {
    const int kz = kz0 + r;
    if (i >= ni || kz >= nz) continue;
    const int k = sh.mk + kz;
    const int c = sh.idx(sh.mi + i, sh.mj + j, k);
    // the features end in FP32 (then BF16): logarithms in FP32 of the FP64
    // arguments, sums in FP64
    const double ro = q.q[0][c], ei = q.q[7][c];
    const double bx = q.q[4][c], by = q.q[5][c], bz = q.q[6][c];
    const double lnT = (double)logf((float)((gm - 1.0) * ei));
    tile[0][r][lane] = logf((float)ro);
    tile[1][r][lane] = (float)lnT;
    tile[2][r][lane] = (float)(bx * bx + by * by + bz * bz);
    if (nfeat == 4) {
      const int ip = (i + h) % ni, im = (i - h % ni + ni) % ni;
      const int jp = (j + h) % nj, jm = (j - h % nj + nj) % nj;
      const double s =
          (double)logf((float)((gm - 1.0) * q.q[7][sh.idx(sh.mi + ip, sh.mj + j, k)])) +
          (double)logf((float)((gm - 1.0) * q.q[7][sh.idx(sh.mi + im, sh.mj + j, k)])) +
          (double)logf((float)((gm - 1.0) * q.q[7][sh.idx(sh.mi + i, sh.mj + jp, k)])) +
          (double)logf((float)((gm - 1.0) * q.q[7][sh.idx(sh.mi + i, sh.mj + jm, k)]));
      tile[3][r][lane] = (float)(s - 4.0 * lnT);
    }
  }
  __syncthreads();
  // write: 32 lanes along the heights of one (column, feature) pair
  const int zl = threadIdx.x & (FT_Z - 1), pr = threadIdx.x / FT_Z;  // 8 pairs per pass
  const int kz = kz0 + zl;
  if (kz >= nz) return;
  for (int p = pr; p < nfeat * FT_X; p += 256 / FT_Z) {
    const int f = p / FT_X, il = p % FT_X;
    if (i0 + il >= ni) continue;
    const long col = (long)j * ni + i0 + il;
    const int fi = f * nz + kz;
    const float x = (tile[f][zl][il] - mu[fi]) / sd[fi];
    X[col * nfeat * nz + fi] = f2bf(x);
    if (Xf) Xf[col * nfeat * nz + fi] = x;
  }
}

// Q_rad field from the network output: qrad = y * sd_q[kz] + mu_q[kz]
// (+ corr, the column correction the last plain solve left).  Tiled like
// the features: Y rows (heights of a column) are read along the heights,
// the field is written along x.
__global__ __launch_bounds__(256) void column_output_kernel(
    const float *__restrict__ Y, Shape sh, const float *__restrict__ mu,
    const float *__restrict__ sd, const int *__restrict__ flag, const double *__restrict__ q_rt,
    const double *__restrict__ corr, double *__restrict__ qrad) {
  __shared__ float tile[FT_Z][FT_X + 1];
  const int ni = sh.It - 2 * sh.mi, nz = sh.Kt - 2 * sh.mk;
  const int i0 = blockIdx.x * FT_X, j = blockIdx.y, kz0 = blockIdx.z * FT_Z;
  {
    const int zl = threadIdx.x & (FT_Z - 1), pr = threadIdx.x / FT_Z;
    const int kz = kz0 + zl;
    for (int il = pr; il < FT_X; il += 256 / FT_Z)
      if (i0 + il < ni && kz < nz) tile[zl][il] = Y[((long)j * ni + i0 + il) * nz + kz];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = i0 + lane;
  if (i >= ni) return;
  const long col = (long)j * ni + i;
  const bool fallback = flag && flag[col];
  for (int r = wv; r < FT_Z; r += 4) {
    const int kz = kz0 + r;
    if (kz >= nz) break;
    const int c = sh.idx(sh.mi + i, sh.mj + j, sh.mk + kz);
    if (fallback) {
      qrad[c] = q_rt[c];  // fallback column: last plain-RT value
      continue;
    }
    const double net = (double)(tile[r][lane] * sd[kz] + mu[kz]);
    qrad[c] = corr ? net + corr[c] : net;
  }
}

// residual indicator per column: relative L2 of (mlp - rt) over the column
// above `tol` -> flag = 1; corr (optional) receives rt - mlp on every cell
__global__ void column_residual_kernel(const float *__restrict__ Y, Shape sh,
                                       const float *__restrict__ mu, const float *__restrict__ sd,
                                       const double *__restrict__ q_rt, int *__restrict__ flag,
                                       float tol, double *__restrict__ corr) {
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nz = sh.Kt - 2 * sh.mk;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (long)ni * nj) return;
  const int i = (int)(col % ni), j = (int)(col / ni);
  double num = 0.0, den = 0.0;
  for (int kz = 0; kz < nz; ++kz) {
    const int c = sh.idx(sh.mi + i, sh.mj + j, sh.mk + kz);
    const double qr = q_rt[c];
    const double qm = (double)(Y[col * nz + kz] * sd[kz] + mu[kz]);
    if (corr) corr[c] = qr - qm;
    num += (qm - qr) * (qm - qr);
    den += qr * qr;
  }
  flag[col] = (num > (double)tol * tol * den) ? 1 : 0;
}

// ------------------------------------------------------------- launchers
template <bool TA, bool TB, int EPI>
static hipError_t gemm_t(hipStream_t st, const uint16_t *A, const uint16_t *B, float *C,
                         uint16_t *H, const float *bias, int M, int N, int K, int lda, int ldb,
                         int ldc, int splits) {
  const int kps = ((K / splits + BK - 1) / BK) * BK;
  const int nz = (K + kps - 1) / kps;
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, nz);
  hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, EPI>), grid, dim3(256), 0, st, A, B, C, H, bias, M,
                     N, K, lda, ldb, ldc, kps);
  return hipGetLastError();
}

hipError_t launch_gemm_bf16(hipStream_t st, int ta, int tb, int epi, const uint16_t *A,
                            const uint16_t *B, float *C, uint16_t *H, const float *bias, int M,
                            int N, int K, int lda, int ldb, int ldc, int splits) {
  if (K % BK) return hipErrorInvalidValue;
  if (splits < 1) splits = 1;
  if (epi != 0 && splits != 1) return hipErrorInvalidValue;
  if (!ta && tb) {
    if (epi == 0) return gemm_t<false, true, 0>(st, A, B, C, H, bias, M, N, K, lda, ldb, ldc, splits);
    if (epi == 1) return gemm_t<false, true, 1>(st, A, B, C, H, bias, M, N, K, lda, ldb, ldc, 1);
    if (epi == 3) return gemm_t<false, true, 3>(st, A, B, C, H, bias, M, N, K, lda, ldb, ldc, 1);
    return gemm_t<false, true, 2>(st, A, B, C, H, bias, M, N, K, lda, ldb, ldc, 1);
  }
  if (!ta && !tb) return gemm_t<false, false, 0>(st, A, B, C, H, bias, M, N, K, lda, ldb, ldc, splits);
  if (ta && !tb) return gemm_t<true, false, 0>(st, A, B, C, H, bias, M, N, K, lda, ldb, ldc, splits);
  return gemm_t<true, true, 0>(st, A, B, C, H, bias, M, N, K, lda, ldb, ldc, splits);
}

hipError_t launch_splitk_reduce(hipStream_t st, const float *part, float *out, long n,
                                int splits) {
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     part, out, n, splits);
  return hipGetLastError();
}

hipError_t launch_silu_backward(hipStream_t st, const float *dact, const float *pre,
                                uint16_t *dpre, long n) {
  hipLaunchKernelGGL(silu_backward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     dact, pre, dpre, n);
  return hipGetLastError();
}

hipError_t launch_colsum_bf16(hipStream_t st, const uint16_t *d, float *db, int M, int N) {
  hipLaunchKernelGGL(colsum_bf16_kernel, dim3((N + 63) / 64), dim3(1024), 0, st, d, db, M, N);
  return hipGetLastError();
}

hipError_t launch_mse(hipStream_t st, const float *Y, const float *T, uint16_t *dY, float *part,
                      long n, float scale, int *nblocks) {
  const int nb = (int)((n + 255) / 256);
  *nblocks = nb;
  hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(256), 0, st, Y, T, dY, part, n, scale);
  return hipGetLastError();
}

hipError_t launch_adam(hipStream_t st, float *w, uint16_t *wb, const float *g, float *m1,
                       float *m2, long n, float lr, float b1, float b2, float eps, float c1,
                       float c2, float wd) {
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, wb, g,
                     m1, m2, n, lr, b1, b2, eps, c1, c2, wd);
  return hipGetLastError();
}

hipError_t launch_to_bf16(hipStream_t st, const float *x, uint16_t *y, long n) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y,
                     n);
  return hipGetLastError();
}

hipError_t launch_column_features(hipStream_t st, const Fld<const double> &q, const Shape &sh,
                                  double gm, const float *mu, const float *sd, uint16_t *X,
                                  float *Xf, int nfeat, int h) {
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nz = sh.Kt - 2 * sh.mk;
  if (nfeat < 3 || nfeat > 4) return hipErrorInvalidValue;
  dim3 grid((ni + FT_X - 1) / FT_X, nj, (nz + FT_Z - 1) / FT_Z);
  hipLaunchKernelGGL(column_features_kernel, grid, dim3(256), 0, st, q, sh, gm, mu, sd, X, Xf,
                     nfeat, h);
  return hipGetLastError();
}

hipError_t launch_column_output(hipStream_t st, const float *Y, const Shape &sh, const float *mu,
                                const float *sd, const int *flag, const double *q_rt,
                                const double *corr, double *qrad) {
  const int ni = sh.It - 2 * sh.mi, nj = sh.Jt - 2 * sh.mj, nz = sh.Kt - 2 * sh.mk;
  dim3 grid((ni + FT_X - 1) / FT_X, nj, (nz + FT_Z - 1) / FT_Z);
  hipLaunchKernelGGL(column_output_kernel, grid, dim3(256), 0, st, Y, sh, mu, sd, flag, q_rt, corr,
                     qrad);
  return hipGetLastError();
}

hipError_t launch_column_residual(hipStream_t st, const float *Y, const Shape &sh,
                                  const float *mu, const float *sd, const double *q_rt, int *flag,
                                  float tol, double *corr) {
  const long n = (long)(sh.It - 2 * sh.mi) * (sh.Jt - 2 * sh.mj);
  hipLaunchKernelGGL(column_residual_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     Y, sh, mu, sd, q_rt, flag, tol, corr);
  return hipGetLastError();
}

}  // namespace misohip
