// C ABI of the warm-start network (include/ps_mi355x.h, ps_mlp_*): a dense
// GELU MLP whose forward pass, backward pass and Adam update run in the
// kernels of ps_mlp.hip on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/ps_mi355x.h"
#include "ps_ctx.h"
#include "ps_mlp.h"

struct ps_mlp {
  ps_ctx* ctx = nullptr;
  int n_in = 0, n_hidden = 0, n_layers = 0, n_out = 0;
  std::vector<int> kin, kout;           // per layer (hidden layers, then the output layer)
  std::vector<size_t> w_off, b_off;     // offsets into the parameter vector
  size_t n_params = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // parameters, gradients, Adam moments
  long long adam_t = 0;
  // activations for up to `cap` rows
  int cap = 0;
  float *x = nullptr, *t = nullptr, *dy = nullptr, *sq = nullptr, *da = nullptr, *db = nullptr;
  std::vector<float*> z, h;
  int* idx = nullptr;
};

namespace {

#define MLP_HIP(call)                                    \
  do {                                                   \
    if ((call) != hipSuccess) {                          \
      if (net->ctx) net->ctx->err = "HIP failure in " #call; \
      return PS_E_HIP;                                   \
    }                                                    \
  } while (0)

int ensure_rows(ps_mlp* net, int rows) {
  if (rows <= net->cap) return PS_OK;
  auto fr = [](float*& p) {
    if (p) hipFree(p);
    p = nullptr;
  };
  fr(net->x);
  fr(net->t);
  fr(net->dy);
  fr(net->sq);
  fr(net->da);
  fr(net->db);
  for (auto& p : net->z) fr(p);
  for (auto& p : net->h) fr(p);
  if (net->idx) hipFree(net->idx);
  net->idx = nullptr;
  net->cap = 0;
  const size_t r = rows;
  const int wmax = std::max(std::max(net->n_in, net->n_hidden), net->n_out);
  MLP_HIP(hipMalloc(&net->x, r * net->n_in * sizeof(float)));
  MLP_HIP(hipMalloc(&net->t, r * net->n_out * sizeof(float)));
  MLP_HIP(hipMalloc(&net->dy, r * net->n_out * sizeof(float)));
  MLP_HIP(hipMalloc(&net->sq, r * sizeof(float)));
  MLP_HIP(hipMalloc(&net->da, r * wmax * sizeof(float)));
  MLP_HIP(hipMalloc(&net->db, r * wmax * sizeof(float)));
  net->z.assign(net->n_layers + 1, nullptr);
  net->h.assign(net->n_layers, nullptr);
  for (int l = 0; l <= net->n_layers; ++l) MLP_HIP(hipMalloc(&net->z[l], r * net->kout[l] * sizeof(float)));
  for (int l = 0; l < net->n_layers; ++l) MLP_HIP(hipMalloc(&net->h[l], r * net->kout[l] * sizeof(float)));
  MLP_HIP(hipMalloc(&net->idx, r * sizeof(int)));
  net->cap = rows;
  return PS_OK;
}

// Forward pass of `rows` rows already in net->x; output in net->z[n_layers].
int forward(ps_mlp* net, int rows, bool bf16) {
  hipStream_t s = net->ctx->stream;
  const float* in = net->x;
  for (int l = 0; l <= net->n_layers; ++l) {
    // hidden layers: pre-activation z (kept for the backward pass) and its
    // GELU h from the same launch
    float* h = l < net->n_layers ? net->h[l] : nullptr;
    MLP_HIP(psm_gemm(bf16, false, false, rows, net->kout[l], net->kin[l], in, net->kin[l], net->w + net->w_off[l],
                     net->kout[l], net->w + net->b_off[l], net->z[l], net->kout[l], s, h));
    if (h) in = h;
  }
  return PS_OK;
}

// Loss and parameter gradients (into net->g) for `rows` rows in net->x /
// net->t; per-row squared errors in net->sq.
int backward(ps_mlp* net, int rows) {
  hipStream_t s = net->ctx->stream;
  int rc = forward(net, rows, true);
  if (rc) return rc;
  const int L = net->n_layers;
  MLP_HIP(psm_mse(rows, net->n_out, net->n_out, net->z[L], net->t, net->dy, net->sq, s));
  const float* dz = net->dy;  // gradient at the pre-activation of layer l
  for (int l = L; l >= 0; --l) {
    const float* in = l > 0 ? net->h[l - 1] : net->x;
    const int K = net->kin[l], N = net->kout[l];
    // dW = in^T dz, db = column sums of dz
    MLP_HIP(psm_gemm(true, true, false, K, N, rows, in, K, dz, N, nullptr, net->g + net->w_off[l], N, s));
    MLP_HIP(psm_colsum(rows, N, dz, net->g + net->b_off[l], s));
    if (l > 0) {
      // d(in) = dz W^T, then through the GELU of layer l-1
      float* dh = (dz == net->da) ? net->db : net->da;
      MLP_HIP(psm_gemm(true, false, true, rows, K, N, dz, N, net->w + net->w_off[l], N, nullptr, dh, K, s));
      MLP_HIP(psm_gelu_bwd(rows * K, dh, net->z[l - 1], dh, s));
      dz = dh;
    }
  }
  return PS_OK;
}

}  // namespace

extern "C" {

int ps_mlp_create(ps_ctx* ctx, int n_in, int n_hidden, int n_layers, int n_out, unsigned long long seed,
                  ps_mlp** out) {
  if (!ctx || !out || n_in < 1 || n_hidden < 1 || n_layers < 1 || n_out < 1) return PS_E_ARG;
  *out = nullptr;
  if (hipSetDevice(ctx->device) != hipSuccess) return PS_E_HIP;
  ps_mlp* net = new ps_mlp();
  net->ctx = ctx;
  net->n_in = n_in;
  net->n_hidden = n_hidden;
  net->n_layers = n_layers;
  net->n_out = n_out;
  size_t off = 0;
  for (int l = 0; l <= n_layers; ++l) {
    const int k = l == 0 ? n_in : n_hidden, o = l == n_layers ? n_out : n_hidden;
    net->kin.push_back(k);
    net->kout.push_back(o);
    net->w_off.push_back(off);
    off += (size_t)k * o;
    net->b_off.push_back(off);
    off += o;
  }
  net->n_params = off;
  // He-uniform weights (GELU layers), zero biases, from a seeded generator
  std::vector<float> p(off, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l <= n_layers; ++l) {
    const float lim = std::sqrt(6.0f / (float)net->kin[l]) * (l == n_layers ? 0.5f : 1.0f);
    std::uniform_real_distribution<float> U(-lim, lim);
    for (size_t i = 0; i < (size_t)net->kin[l] * net->kout[l]; ++i) p[net->w_off[l] + i] = U(rng);
  }
  const size_t bytes = off * sizeof(float);
  if (hipMalloc(&net->w, bytes) != hipSuccess || hipMalloc(&net->g, bytes) != hipSuccess ||
      hipMalloc(&net->m, bytes) != hipSuccess || hipMalloc(&net->v, bytes) != hipSuccess ||
      hipMemcpy(net->w, p.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(net->m, 0, bytes) != hipSuccess || hipMemset(net->v, 0, bytes) != hipSuccess) {
    ps_mlp_free(net);
    return PS_E_HIP;
  }
  *out = net;
  return PS_OK;
}

long long ps_mlp_n_params(const ps_mlp* net) { return net ? (long long)net->n_params : 0; }

int ps_mlp_get_params(const ps_mlp* net, float* params) {
  if (!net || !params) return PS_E_ARG;
  return hipMemcpy(params, net->w, net->n_params * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess ? PS_OK
                                                                                                     : PS_E_HIP;
}

int ps_mlp_set_params(ps_mlp* net, const float* params) {
  if (!net || !params) return PS_E_ARG;
  if (hipMemcpy(net->w, params, net->n_params * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return PS_E_HIP;
  net->adam_t = 0;
  const size_t bytes = net->n_params * sizeof(float);
  if (hipMemset(net->m, 0, bytes) != hipSuccess || hipMemset(net->v, 0, bytes) != hipSuccess) return PS_E_HIP;
  return PS_OK;
}

int ps_mlp_forward(ps_mlp* net, int rows, const float* x, float* y, int bf16) {
  if (!net || rows < 0 || (rows && (!x || !y))) return PS_E_ARG;
  if (rows == 0) return PS_OK;
  hipStream_t s = net->ctx->stream;
  const int chunk = 8192;
  for (int r0 = 0; r0 < rows; r0 += chunk) {
    const int r = std::min(chunk, rows - r0);
    int rc = ensure_rows(net, r);
    if (rc) return rc;
    MLP_HIP(hipMemcpyAsync(net->x, x + (size_t)r0 * net->n_in, (size_t)r * net->n_in * sizeof(float),
                           hipMemcpyHostToDevice, s));
    rc = forward(net, r, bf16 != 0);
    if (rc) return rc;
    MLP_HIP(hipMemcpyAsync(y + (size_t)r0 * net->n_out, net->z[net->n_layers], (size_t)r * net->n_out * sizeof(float),
                           hipMemcpyDeviceToHost, s));
  }
  MLP_HIP(hipStreamSynchronize(s));
  return PS_OK;
}

int ps_mlp_loss_grad(ps_mlp* net, int rows, const float* x, const float* t, float* loss, float* grads) {
  if (!net || rows < 1 || !x || !t) return PS_E_ARG;
  hipStream_t s = net->ctx->stream;
  int rc = ensure_rows(net, rows);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(net->x, x, (size_t)rows * net->n_in * sizeof(float), hipMemcpyHostToDevice, s));
  MLP_HIP(hipMemcpyAsync(net->t, t, (size_t)rows * net->n_out * sizeof(float), hipMemcpyHostToDevice, s));
  rc = backward(net, rows);
  if (rc) return rc;
  std::vector<float> sq(rows);
  MLP_HIP(hipMemcpyAsync(sq.data(), net->sq, rows * sizeof(float), hipMemcpyDeviceToHost, s));
  if (grads) MLP_HIP(hipMemcpyAsync(grads, net->g, net->n_params * sizeof(float), hipMemcpyDeviceToHost, s));
  MLP_HIP(hipStreamSynchronize(s));
  if (loss) {
    double tot = 0;
    for (float q : sq) tot += q;
    *loss = (float)(tot / ((double)rows * net->n_out));
  }
  return PS_OK;
}

int ps_mlp_train(ps_mlp* net, int n, const float* x, const float* t, int epochs, int batch, float lr,
                 unsigned long long seed, float* loss) {
  if (!net || n < 1 || !x || !t || epochs < 0 || batch < 1 || !(lr > 0)) return PS_E_ARG;
  hipStream_t s = net->ctx->stream;
  batch = std::min(batch, n);
  int rc = ensure_rows(net, batch);
  if (rc) return rc;
  float *X = nullptr, *T = nullptr;
  if (hipMalloc(&X, (size_t)n * net->n_in * sizeof(float)) != hipSuccess ||
      hipMalloc(&T, (size_t)n * net->n_out * sizeof(float)) != hipSuccess ||
      hipMemcpy(X, x, (size_t)n * net->n_in * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(T, t, (size_t)n * net->n_out * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    if (X) hipFree(X);
    if (T) hipFree(T);
    return PS_E_HIP;
  }
  std::mt19937_64 rng(seed);
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<float> sq(batch);
  rc = PS_OK;
  for (int e = 0; e < epochs && rc == PS_OK; ++e) {
    std::shuffle(perm.begin(), perm.end(), rng);
    double tot = 0;
    int seen = 0;
    for (int b0 = 0; b0 + batch <= n && rc == PS_OK; b0 += batch) {
      if (hipMemcpyAsync(net->idx, perm.data() + b0, batch * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
          psm_gather_rows(batch, net->n_in, net->idx, X, net->x, s) != hipSuccess ||
          psm_gather_rows(batch, net->n_out, net->idx, T, net->t, s) != hipSuccess) {
        rc = PS_E_HIP;
        break;
      }
      if ((rc = backward(net, batch)) != PS_OK) break;
      ++net->adam_t;
      const float c1 = 1.f - std::pow(0.9f, (float)net->adam_t), c2 = 1.f - std::pow(0.999f, (float)net->adam_t);
      if (psm_adam((int)net->n_params, net->w, net->g, net->m, net->v, lr, 0.9f, 0.999f, 1e-8f, c1, c2, s) !=
              hipSuccess ||
          hipMemcpyAsync(sq.data(), net->sq, batch * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        rc = PS_E_HIP;
        break;
      }
      for (float q : sq) tot += q;
      seen += batch;
    }
    if (loss) loss[e] = seen ? (float)(tot / ((double)seen * net->n_out)) : 0.f;
  }
  hipFree(X);
  hipFree(T);
  return rc;
}

int ps_mlp_free(ps_mlp* net) {
  if (!net) return PS_E_ARG;
  for (float* p : {net->w, net->g, net->m, net->v, net->x, net->t, net->dy, net->sq, net->da, net->db})
    if (p) hipFree(p);
  for (float* p : net->z)
    if (p) hipFree(p);
  for (float* p : net->h)
    if (p) hipFree(p);
  if (net->idx) hipFree(net->idx);
  delete net;
  return PS_OK;
}

}  // extern "C"
