// Internal state of a misohip context and helpers shared by the C-ABI
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../../include/misohip.h"

struct misohip_ctx {
  int device;
  hipStream_t stream;
  unsigned *d_slot;  // CFL reduction slot (device)
  unsigned *h_slot;  // pinned host words: [0] CFL, [1] RT change, bytes 64..87 RT spacings
  // timing: pairs of events recorded around launches, tagged by kind
  bool profiling = false;
  std::vector<hipEvent_t> pool;
  std::vector<int> rec_kind;
  std::vector<hipEvent_t> rec_start, rec_stop;
  misohip_tiles tile = {2, 8, 256, 5, 256, 1, 64};
  // y-fastest copies of kappa and S for the RT rays kept y-fastest (grown
  // on demand, freed with the context)
  double *rt_kap_y = nullptr, *rt_src_y = nullptr;
  size_t rt_cells = 0;
  unsigned *rt_parts = nullptr;  // the RT change-reduction slots
  // a second stream (default priority) and two events: an RT solve sweeps
  // its multi-plane rays there while the plane rays sweep on `stream`
  hipStream_t rt_side = nullptr;
  hipEvent_t rt_fork = nullptr, rt_join = nullptr;
  // two more streams (default priority) for the AV sweeps run as a pipeline
  // over z slabs (the x-y walk on the first, the z walk on the second), and
  // its events: one per slab (16), the fork and the two joins
  hipStream_t tail_side[2] = {nullptr, nullptr};
  hipEvent_t tail_ev[19] = {};
  // set for the length of a call that must not allocate or create streams
  // or events (misohip_update_noalloc)
  bool no_alloc = false;
  misohip_bc_fn bc_fn = nullptr;  // the user boundary callback and its argument
  void *bc_user = nullptr;
  misohip_src_fn src_fn = nullptr;  // the user source callback, its argument and arrays
  void *src_user = nullptr;
  void *src_arr[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace misohip_detail {
// RAII scope: records a start event on construction and a stop event on
// destruction when profiling is on.
struct Timed {
  misohip_ctx *c;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  static hipEvent_t get(misohip_ctx *c) {
    if (!c->pool.empty()) {
      hipEvent_t e = c->pool.back();
      c->pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  Timed(misohip_ctx *c_, int k) : c(c_), kind(k) {
    if (!c->profiling) return;
    a = get(c);
    b = get(c);
    if (a) (void)hipEventRecord(a, c->stream);
  }
  ~Timed() {
    if (!a || !b) return;
    (void)hipEventRecord(b, c->stream);
    c->rec_kind.push_back(kind);
    c->rec_start.push_back(a);
    c->rec_stop.push_back(b);
  }
};
}  // namespace misohip_detail
using misohip_detail::Timed;

namespace misohip_detail {
extern thread_local char g_last_hip_error[256];
int hip_fail(hipError_t e);
}  // namespace misohip_detail
using misohip_detail::hip_fail;

#define MISOHIP_TRY(expr)                      \
  do {                                         \
    hipError_t e_ = (expr);                    \
    if (e_ != hipSuccess) return hip_fail(e_); \
  } while (0)
