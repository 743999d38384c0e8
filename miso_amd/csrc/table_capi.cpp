// C ABI of the device table interpolators (misohip_table1d / misohip_table2d).
#include <hip/hip_runtime.h>

#include "ctx_internal.hpp"
#include "launch.hpp"

using namespace misohip;

namespace {

template <typename R>
int table1d_t(misohip_ctx *ctx, const void *table, int n, double x_min, double x_max,
              const void *x, void *y, long count) {
  // the constructor's spacing and its inverse, in the table's type
  const R dx = (R(x_max) - R(x_min)) / static_cast<R>(n - 1);
  const R dxi = R(1) / dx;
  MISOHIP_TRY(launch_table1d<R>(ctx->stream, static_cast<const R *>(table), n, R(x_min), dxi,
                                static_cast<const R *>(x), static_cast<R *>(y), count));
  return MISOHIP_OK;
}

template <typename R>
int table2d_t(misohip_ctx *ctx, const void *table, int n0, int n1, double x0_min, double x0_max,
              double x1_min, double x1_max, const void *x0, const void *x1, void *y,
              long count) {
  const R dx0 = (R(x0_max) - R(x0_min)) / static_cast<R>(n0 - 1);
  const R dxi0 = R(1) / dx0;
  const R dx1 = (R(x1_max) - R(x1_min)) / static_cast<R>(n1 - 1);
  const R dxi1 = R(1) / dx1;
  MISOHIP_TRY(launch_table2d<R>(ctx->stream, static_cast<const R *>(table), n0, n1, R(x0_min),
                                dxi0, R(x1_min), dxi1, static_cast<const R *>(x0),
                                static_cast<const R *>(x1), static_cast<R *>(y), count));
  return MISOHIP_OK;
}

}  // namespace

extern "C" {

int misohip_table1d(misohip_ctx *ctx, misohip_dtype dt, const void *table, int n, double x_min,
                    double x_max, const void *x, void *y, long count) {
  if (!ctx || !table || n < 2 || !(x_max > x_min) || count < 0 || (count && (!x || !y)))
    return MISOHIP_E_ARG;
  if (count == 0) return MISOHIP_OK;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  if (dt == MISOHIP_F64) return table1d_t<double>(ctx, table, n, x_min, x_max, x, y, count);
  return table1d_t<float>(ctx, table, n, x_min, x_max, x, y, count);
}

int misohip_table2d(misohip_ctx *ctx, misohip_dtype dt, const void *table, int n0, int n1,
                    double x0_min, double x0_max, double x1_min, double x1_max, const void *x0,
                    const void *x1, void *y, long count) {
  if (!ctx || !table || n0 < 2 || n1 < 2 || !(x0_max > x0_min) || !(x1_max > x1_min) ||
      count < 0 || (count && (!x0 || !x1 || !y)))
    return MISOHIP_E_ARG;
  if ((long)n0 * n1 > 2147483647L) return MISOHIP_E_RANGE;
  if (count == 0) return MISOHIP_OK;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  if (dt == MISOHIP_F64)
    return table2d_t<double>(ctx, table, n0, n1, x0_min, x0_max, x1_min, x1_max, x0, x1, y,
                             count);
  return table2d_t<float>(ctx, table, n0, n1, x0_min, x0_max, x1_min, x1_max, x0, x1, y, count);
}

}  // extern "C"
