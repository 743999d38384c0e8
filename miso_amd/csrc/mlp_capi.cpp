// C ABI of the per-column MLP surrogate (include/misohip.h, misohip_mlp_*).
// The library owns the network state; every GEMM is the bf16 MFMA kernel of
// mlp.hip, every elementwise step a kernel of the same file.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "ctx_internal.hpp"
#include "launch.hpp"

using namespace misohip;

namespace misohip {
hipError_t launch_gemm_bf16(hipStream_t, int, int, int, const uint16_t *, const uint16_t *,
                            float *, uint16_t *, const float *, int, int, int, int, int, int, int);
hipError_t launch_splitk_reduce(hipStream_t, const float *, float *, long, int);
hipError_t launch_silu_backward(hipStream_t, const float *, const float *, uint16_t *, long);
hipError_t launch_colsum_bf16(hipStream_t, const uint16_t *, float *, int, int);
hipError_t launch_mse(hipStream_t, const float *, const float *, uint16_t *, float *, long, float,
                      int *);
hipError_t launch_adam(hipStream_t, float *, uint16_t *, const float *, float *, float *, long,
                       float, float, float, float, float, float, float);
hipError_t launch_to_bf16(hipStream_t, const float *, uint16_t *, long);
hipError_t launch_column_features(hipStream_t, const Fld<const double> &, const Shape &, double,
                                  const float *, const float *, uint16_t *, float *, int, int);
hipError_t launch_column_output(hipStream_t, const float *, const Shape &, const float *,
                                const float *, const int *, const double *, const double *,
                                double *);
hipError_t launch_column_residual(hipStream_t, const float *, const Shape &, const float *,
                                  const float *, const double *, int *, float, double *);
}  // namespace misohip

struct misohip_mlp {
  int n_in, n_hidden, n_out, n_layers;  // n_layers = hidden layers + output layer
  int max_batch;
  int splits;
  long t;  // Adam step
  std::vector<int> in, out;
  std::vector<float *> w, b, gw, gb, mw, vw, mb, vb;
  std::vector<uint16_t *> wb;
  std::vector<float *> pre;      // [max_batch][out_l] fp32
  std::vector<uint16_t *> act;   // [max_batch][out_l] bf16 (hidden layers)
  float *dact = nullptr;         // [max_batch][max width] fp32
  uint16_t *dpre = nullptr;      // [max_batch][max width] bf16
  uint16_t *dpre2 = nullptr;
  float *part = nullptr;         // split-K partials
  float *loss_part = nullptr;
  float *h_loss = nullptr;       // pinned
  float *mu_in = nullptr, *sd_in = nullptr, *mu_out = nullptr, *sd_out = nullptr;
  std::vector<void *> allocs;
};

namespace {

template <typename T> hipError_t alloc(misohip_mlp *m, T **p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, sizeof(T) * (n ? n : 1));
  if (e != hipSuccess) return e;
  e = hipMemset(q, 0, sizeof(T) * (n ? n : 1));
  m->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return e;
}

// keep_pre: store the hidden layers' pre-activations (the backward pass
// reads them; inference does not need them)
int forward(misohip_ctx *ctx, misohip_mlp *m, const uint16_t *X, int batch, float *Y,
            bool keep_pre = true) {
  const uint16_t *inp = X;
  for (int l = 0; l < m->n_layers; ++l) {
    const bool last = l == m->n_layers - 1;
    float *C = last && Y ? Y : m->pre[l];
    MISOHIP_TRY(launch_gemm_bf16(ctx->stream, 0, 1, last ? 2 : keep_pre ? 1 : 3, inp, m->wb[l], C,
                                 last ? nullptr : m->act[l], m->b[l], batch, m->out[l], m->in[l],
                                 m->in[l], m->in[l], m->out[l], 1));
    inp = m->act[l];
  }
  return MISOHIP_OK;
}

// gradients of all weights from dpre of the output layer (bf16, in m->dpre)
int backward(misohip_ctx *ctx, misohip_mlp *m, const uint16_t *X, int batch) {
  uint16_t *dcur = m->dpre, *dnext = m->dpre2;
  for (int l = m->n_layers - 1; l >= 0; --l) {
    const uint16_t *inp = l == 0 ? X : m->act[l - 1];
    const int O = m->out[l], I = m->in[l];
    // gW = dpre^T * inp  ([O][I], reduction over the batch, split over K)
    // split the batch reduction into equal multiples of 32 (or not at all)
    int splits = m->splits;
    while (splits > 1 && (batch % splits || (batch / splits) % 32)) --splits;
    if (splits > 1) {
      MISOHIP_TRY(launch_gemm_bf16(ctx->stream, 1, 0, 0, dcur, inp, m->part, nullptr, nullptr, O,
                                   I, batch, O, I, I, splits));
      MISOHIP_TRY(launch_splitk_reduce(ctx->stream, m->part, m->gw[l], (long)O * I, splits));
    } else {
      MISOHIP_TRY(launch_gemm_bf16(ctx->stream, 1, 0, 0, dcur, inp, m->gw[l], nullptr, nullptr, O,
                                   I, batch, O, I, I, 1));
    }
    MISOHIP_TRY(launch_colsum_bf16(ctx->stream, dcur, m->gb[l], batch, O));
    if (l > 0) {
      // dact = dpre * W  ([batch][I]); dpre_prev = dact * silu'(pre_prev)
      MISOHIP_TRY(launch_gemm_bf16(ctx->stream, 0, 0, 0, dcur, m->wb[l], m->dact, nullptr,
                                   nullptr, batch, I, O, O, I, I, 1));
      MISOHIP_TRY(launch_silu_backward(ctx->stream, m->dact, m->pre[l - 1], dnext, (long)batch * I));
      std::swap(dcur, dnext);
    }
  }
  return MISOHIP_OK;
}

int check_batch(misohip_mlp *m, int batch) {
  if (batch < 1 || batch > m->max_batch) return MISOHIP_E_RANGE;
  return MISOHIP_OK;
}

Shape shape_of(const misohip_grid *g) {
  Shape s;
  s.It = g->i_total; s.Jt = g->j_total; s.Kt = g->k_total;
  s.mi = g->i_margin; s.mj = g->j_margin; s.mk = g->k_margin;
  s.is = g->is; s.js = g->js; s.ks = g->ks;
  return s;
}

}  // namespace

extern "C" {

int misohip_mlp_create(misohip_ctx *ctx, int n_in, int n_hidden, int n_hidden_layers, int n_out,
                       int max_batch, misohip_mlp **out) {
  if (!ctx || !out || n_in <= 0 || n_hidden <= 0 || n_out <= 0 || n_hidden_layers < 1)
    return MISOHIP_E_ARG;
  if (n_in % 32 || n_hidden % 32 || n_out % 32 || max_batch < 32 || max_batch % 32)
    return MISOHIP_E_ARG;  // GEMM K dimensions and the split-K batch
  *out = nullptr;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  misohip_mlp *m = new (std::nothrow) misohip_mlp;
  if (!m) return MISOHIP_E_NOMEM;
  m->n_in = n_in;
  m->n_hidden = n_hidden;
  m->n_out = n_out;
  m->n_layers = n_hidden_layers + 1;
  m->max_batch = max_batch;
  m->splits = 16;
  m->t = 0;
  int maxw = n_hidden > n_out ? n_hidden : n_out;
  if (n_in > maxw) maxw = n_in;
  hipError_t e = hipSuccess;
  for (int l = 0; l < m->n_layers && e == hipSuccess; ++l) {
    const int I = l == 0 ? n_in : n_hidden;
    const int O = l == m->n_layers - 1 ? n_out : n_hidden;
    m->in.push_back(I);
    m->out.push_back(O);
    float *p[8] = {};
    uint16_t *pb = nullptr;
    const size_t nw = (size_t)O * I;
    e = alloc(m, &p[0], nw);
    if (e == hipSuccess) e = alloc(m, &p[1], O);
    if (e == hipSuccess) e = alloc(m, &p[2], nw);
    if (e == hipSuccess) e = alloc(m, &p[3], O);
    if (e == hipSuccess) e = alloc(m, &p[4], nw);
    if (e == hipSuccess) e = alloc(m, &p[5], nw);
    if (e == hipSuccess) e = alloc(m, &p[6], O);
    if (e == hipSuccess) e = alloc(m, &p[7], O);
    if (e == hipSuccess) e = alloc(m, &pb, nw);
    m->w.push_back(p[0]); m->b.push_back(p[1]); m->gw.push_back(p[2]); m->gb.push_back(p[3]);
    m->mw.push_back(p[4]); m->vw.push_back(p[5]); m->mb.push_back(p[6]); m->vb.push_back(p[7]);
    m->wb.push_back(pb);
    float *pre = nullptr;
    uint16_t *act = nullptr;
    if (e == hipSuccess) e = alloc(m, &pre, (size_t)max_batch * O);
    if (e == hipSuccess && l < m->n_layers - 1) e = alloc(m, &act, (size_t)max_batch * O);
    m->pre.push_back(pre);
    m->act.push_back(act);
  }
  if (e == hipSuccess) e = alloc(m, &m->dact, (size_t)max_batch * maxw);
  if (e == hipSuccess) e = alloc(m, &m->dpre, (size_t)max_batch * maxw);
  if (e == hipSuccess) e = alloc(m, &m->dpre2, (size_t)max_batch * maxw);
  if (e == hipSuccess) e = alloc(m, &m->part, (size_t)m->splits * maxw * maxw);
  if (e == hipSuccess) e = alloc(m, &m->loss_part, (size_t)max_batch * n_out / 256 + 1);
  if (e == hipSuccess) e = alloc(m, &m->mu_in, n_in);
  if (e == hipSuccess) e = alloc(m, &m->sd_in, n_in);
  if (e == hipSuccess) e = alloc(m, &m->mu_out, n_out);
  if (e == hipSuccess) e = alloc(m, &m->sd_out, n_out);
  if (e == hipSuccess) e = hipHostMalloc(&m->h_loss, sizeof(float) * ((size_t)max_batch * n_out / 256 + 1), hipHostMallocDefault);
  if (e != hipSuccess) {
    for (void *q : m->allocs) (void)hipFree(q);
    delete m;
    return hip_fail(e);
  }
  *out = m;
  return MISOHIP_OK;
}

int misohip_mlp_destroy(misohip_mlp *m) {
  if (!m) return MISOHIP_E_ARG;
  (void)hipDeviceSynchronize();
  for (void *q : m->allocs) (void)hipFree(q);
  if (m->h_loss) (void)hipHostFree(m->h_loss);
  delete m;
  return MISOHIP_OK;
}

int misohip_mlp_set_layer(misohip_ctx *ctx, misohip_mlp *m, int layer, const float *w_host,
                          const float *b_host) {
  if (!ctx || !m || layer < 0 || layer >= m->n_layers || !w_host || !b_host) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  const size_t nw = (size_t)m->out[layer] * m->in[layer];
  MISOHIP_TRY(hipMemcpy(m->w[layer], w_host, sizeof(float) * nw, hipMemcpyHostToDevice));
  MISOHIP_TRY(hipMemcpy(m->b[layer], b_host, sizeof(float) * m->out[layer], hipMemcpyHostToDevice));
  MISOHIP_TRY(launch_to_bf16(ctx->stream, m->w[layer], m->wb[layer], (long)nw));
  MISOHIP_TRY(hipMemsetAsync(m->mw[layer], 0, sizeof(float) * nw, ctx->stream));
  MISOHIP_TRY(hipMemsetAsync(m->vw[layer], 0, sizeof(float) * nw, ctx->stream));
  MISOHIP_TRY(hipMemsetAsync(m->mb[layer], 0, sizeof(float) * m->out[layer], ctx->stream));
  MISOHIP_TRY(hipMemsetAsync(m->vb[layer], 0, sizeof(float) * m->out[layer], ctx->stream));
  MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
  m->t = 0;
  return MISOHIP_OK;
}

// which: 0 = fp32 master weights, 1 = biases, 2 = weight gradients, 3 = bias gradients
int misohip_mlp_get_layer(misohip_ctx *ctx, misohip_mlp *m, int layer, int which, float *host) {
  if (!ctx || !m || layer < 0 || layer >= m->n_layers || !host || which < 0 || which > 3)
    return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
  const size_t nw = (size_t)m->out[layer] * m->in[layer];
  const float *src = which == 0 ? m->w[layer] : which == 1 ? m->b[layer]
                   : which == 2 ? m->gw[layer] : m->gb[layer];
  const size_t n = (which == 0 || which == 2) ? nw : (size_t)m->out[layer];
  MISOHIP_TRY(hipMemcpy(host, src, sizeof(float) * n, hipMemcpyDeviceToHost));
  return MISOHIP_OK;
}

int misohip_mlp_set_norm(misohip_ctx *ctx, misohip_mlp *m, const float *mu_in, const float *sd_in,
                         const float *mu_out, const float *sd_out) {
  if (!ctx || !m || !mu_in || !sd_in || !mu_out || !sd_out) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  MISOHIP_TRY(hipMemcpy(m->mu_in, mu_in, sizeof(float) * m->n_in, hipMemcpyHostToDevice));
  MISOHIP_TRY(hipMemcpy(m->sd_in, sd_in, sizeof(float) * m->n_in, hipMemcpyHostToDevice));
  MISOHIP_TRY(hipMemcpy(m->mu_out, mu_out, sizeof(float) * m->n_out, hipMemcpyHostToDevice));
  MISOHIP_TRY(hipMemcpy(m->sd_out, sd_out, sizeof(float) * m->n_out, hipMemcpyHostToDevice));
  return MISOHIP_OK;
}

int misohip_mlp_forward(misohip_ctx *ctx, misohip_mlp *m, const void *X, int batch, float *Y) {
  if (!ctx || !m || !X || !Y) return MISOHIP_E_ARG;
  if (int rc = check_batch(m, batch)) return rc;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  Timed t_(ctx, MISOHIP_KIND_MLP);
  return forward(ctx, m, static_cast<const uint16_t *>(X), batch, Y);
}

// loss and gradients (MSE on the batch); *loss receives the mean squared error
int misohip_mlp_grad(misohip_ctx *ctx, misohip_mlp *m, const void *X, const float *T, int batch,
                     double *loss) {
  if (!ctx || !m || !X || !T) return MISOHIP_E_ARG;
  if (int rc = check_batch(m, batch)) return rc;
  if (batch % 32) return MISOHIP_E_ARG;  // the batch is the weight-gradient GEMM's K
  MISOHIP_TRY(hipSetDevice(ctx->device));
  {
    Timed t_(ctx, MISOHIP_KIND_MLP);
    const uint16_t *Xb = static_cast<const uint16_t *>(X);
    if (int rc = forward(ctx, m, Xb, batch, nullptr)) return rc;
    const long n = (long)batch * m->n_out;
    int nb = 0;
    MISOHIP_TRY(launch_mse(ctx->stream, m->pre[m->n_layers - 1], T, m->dpre, m->loss_part, n,
                           1.f / (float)n, &nb));
    if (int rc = backward(ctx, m, Xb, batch)) return rc;
    if (loss) {
      MISOHIP_TRY(hipMemcpyAsync(m->h_loss, m->loss_part, sizeof(float) * nb,
                                 hipMemcpyDeviceToHost, ctx->stream));
    }
    if (loss) {
      MISOHIP_TRY(hipStreamSynchronize(ctx->stream));
      double s = 0.0;
      for (int i = 0; i < nb; ++i) s += m->h_loss[i];
      *loss = s / (double)n;
    }
  }
  return MISOHIP_OK;
}

// Adam update of every layer from the current gradients
int misohip_mlp_adam(misohip_ctx *ctx, misohip_mlp *m, float lr, float beta1, float beta2,
                     float eps, float weight_decay) {
  if (!ctx || !m || !(lr > 0.f)) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  Timed t_(ctx, MISOHIP_KIND_MLP);
  m->t += 1;
  const float c1 = 1.f - std::pow(beta1, (float)m->t), c2 = 1.f - std::pow(beta2, (float)m->t);
  for (int l = 0; l < m->n_layers; ++l) {
    const long nw = (long)m->out[l] * m->in[l];
    MISOHIP_TRY(launch_adam(ctx->stream, m->w[l], m->wb[l], m->gw[l], m->mw[l], m->vw[l], nw, lr,
                            beta1, beta2, eps, c1, c2, weight_decay));
    MISOHIP_TRY(launch_adam(ctx->stream, m->b[l], nullptr, m->gb[l], m->mb[l], m->vb[l],
                            m->out[l], lr, beta1, beta2, eps, c1, c2, 0.f));
  }
  return MISOHIP_OK;
}

int misohip_mlp_features(misohip_ctx *ctx, misohip_mlp *m, const misohip_grid *g,
                         const misohip_fields *qq, double gm, void *X, float *Xf) {
  if (!ctx || !m || !g || !qq || !X) return MISOHIP_E_ARG;
  const int nz = g->k_total - 2 * g->k_margin;
  const int nfeat = nz > 0 && m->n_in % nz == 0 ? m->n_in / nz : 0;
  if (nfeat != 3 && nfeat != 4) return MISOHIP_E_ARG;
  if (nfeat == 4 && (g->i_size < 2 || g->j_size < 2)) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  Fld<const double> q;
  for (int f = 0; f < 9; ++f) q.q[f] = static_cast<const double *>(qq->q[f]);
  Timed t_(ctx, MISOHIP_KIND_MLP);
  MISOHIP_TRY(launch_column_features(ctx->stream, q, shape_of(g), gm, m->mu_in, m->sd_in,
                                     static_cast<uint16_t *>(X), Xf, nfeat, 4));
  return MISOHIP_OK;
}

// Q_rad of every interior cell from the network (columns along z), plus the
// correction field corr when given; columns with flag != 0 take q_rt
// instead.  Y is scratch of (columns x n_out).
int misohip_mlp_qrad(misohip_ctx *ctx, misohip_mlp *m, const misohip_grid *g, const void *X,
                     float *Y, const int *flags, const void *q_rt, const void *corr,
                     void *qrad) {
  if (!ctx || !m || !g || !X || !Y || !qrad) return MISOHIP_E_ARG;
  const int ncol = (g->i_total - 2 * g->i_margin) * (g->j_total - 2 * g->j_margin);
  if (g->k_total - 2 * g->k_margin != m->n_out) return MISOHIP_E_ARG;
  if (flags && !q_rt) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  Timed t_(ctx, MISOHIP_KIND_MLP);
  for (int c0 = 0; c0 < ncol; c0 += m->max_batch) {
    const int nb = ncol - c0 < m->max_batch ? ncol - c0 : m->max_batch;
    if (int rc = forward(ctx, m, static_cast<const uint16_t *>(X) + (long)c0 * m->n_in, nb,
                         Y + (long)c0 * m->n_out, false))
      return rc;
  }
  MISOHIP_TRY(launch_column_output(ctx->stream, Y, shape_of(g), m->mu_out, m->sd_out, flags,
                                   static_cast<const double *>(q_rt),
                                   static_cast<const double *>(corr), static_cast<double *>(qrad)));
  return MISOHIP_OK;
}

// residual indicator: flag columns whose network Q_rad (from Y) differs from
// the plain-RT q_rt by more than tol in relative L2 over the column; corr
// (optional) receives q_rt - network on every interior cell
int misohip_mlp_residual(misohip_ctx *ctx, misohip_mlp *m, const misohip_grid *g, const float *Y,
                         const void *q_rt, int *flags, double tol, void *corr) {
  if (!ctx || !m || !g || !Y || !q_rt || !flags) return MISOHIP_E_ARG;
  MISOHIP_TRY(hipSetDevice(ctx->device));
  Timed t_(ctx, MISOHIP_KIND_MLP);
  MISOHIP_TRY(launch_column_residual(ctx->stream, Y, shape_of(g), m->mu_out, m->sd_out,
                                     static_cast<const double *>(q_rt), flags, (float)tol,
                                     static_cast<double *>(corr)));
  return MISOHIP_OK;
}

}  // extern "C"
