// HBM rate of plain streaming kernels on this part: the ceiling the MHD
// kernels are compared against (the benchmark's --calibrate-copy).  Each
// lane moves 16 bytes per access; a grid-stride loop with four accesses in
// flight per lane; stores are nontemporal, as the MHD kernels' are.
#include "launch.hpp"

namespace misohip {

typedef double calib_v2 __attribute__((ext_vector_type(2)));

// kind 0: copy src -> dst; 1: read src (a sum that is stored only if it
// equals a value no finite input sums to); 2: write dst
__global__ __launch_bounds__(256) void stream_kernel(const calib_v2 *__restrict__ src,
                                                     calib_v2 *__restrict__ dst, long n, int kind,
                                                     double *__restrict__ sink) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long st = (long)gridDim.x * blockDim.x;
  if (kind == 0) {
    for (; i + 3 * st < n; i += 4 * st) {
      const calib_v2 a = src[i], b = src[i + st], c = src[i + 2 * st], d = src[i + 3 * st];
      __builtin_nontemporal_store(a, &dst[i]);
      __builtin_nontemporal_store(b, &dst[i + st]);
      __builtin_nontemporal_store(c, &dst[i + 2 * st]);
      __builtin_nontemporal_store(d, &dst[i + 3 * st]);
    }
    for (; i < n; i += st) __builtin_nontemporal_store(src[i], &dst[i]);
  } else if (kind == 1) {
    double s = 0.0;
    for (; i + 3 * st < n; i += 4 * st) {
      const calib_v2 a = src[i], b = src[i + st], c = src[i + 2 * st], d = src[i + 3 * st];
      s += (a.x + a.y) + (b.x + b.y) + (c.x + c.y) + (d.x + d.y);
    }
    for (; i < n; i += st) s += src[i].x + src[i].y;
    if (s == -1.2345e300) sink[0] = s;
  } else {
    const calib_v2 v = {1.0, 2.0};
    for (; i < n; i += st) __builtin_nontemporal_store(v, &dst[i]);
  }
}

hipError_t launch_stream(hipStream_t st, const void *src, void *dst, size_t bytes, int kind,
                         double *sink) {
  const long n = (long)(bytes / sizeof(calib_v2));
  hipLaunchKernelGGL(stream_kernel, dim3(16384), dim3(256), 0, st,
                     static_cast<const calib_v2 *>(src), static_cast<calib_v2 *>(dst), n, kind,
                     sink);
  return hipGetLastError();
}

}  // namespace misohip
