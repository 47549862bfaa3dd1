// ehx_mlp.hip -- training kernels of the learned layout selector.
//
// The selector (ehx_selector.h) is an 8 -> 32 -> 32 -> 2 ReLU MLP whose
// weights are stored in FP16 and applied with FP32 accumulation.  Training
// keeps FP32 master weights on the host optimizer; these kernels compute the
// masked mean-squared loss and its gradient with respect to the FP16-rounded
// weights the inference path uses:
//   k_mlp_grad     one sample per lane: forward, backward, then each parameter's
//                  contribution is reduced across the wavefront (DPP shuffles)
//                  and added into an LDS accumulator, then into global memory;
//   k_mlp_predict  forward only;
//   k_set_features the selector's per-set features for every child set of a
//                  trie, computed by the same device code the trie build uses.
#include <string.h>

#include "../../include/ehx_ext.h"
#include "ehx_internal.h"
#include "ehx_selector.h"

namespace ehx {

struct MlpData {
  float *x, *y, *m;  // n x SEL_IN, n x SEL_OUT, n x SEL_OUT (mask)
  uint64_t n;
  float inv_count;   // 1 / sum(mask)
};

__device__ __forceinline__ float rnd_half(float v) { return __half2float(__float2half(v)); }

// Block = 256 threads; LDS holds the rounded weights and the gradient accumulator.
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ params, const float *__restrict__ norm,
                                                  MlpData D, float *__restrict__ grad, float *__restrict__ loss) {
  __shared__ float W[SEL_NPARAM];
  __shared__ float G[SEL_NPARAM];
  __shared__ float s_loss;
  for (int i = threadIdx.x; i < SEL_NPARAM; i += blockDim.x) {
    W[i] = rnd_half(params[i]);
    G[i] = 0.0f;
  }
  if (threadIdx.x == 0) s_loss = 0.0f;
  __syncthreads();
  const uint32_t lane = __lane_id();
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t chunks = (D.n + 63) / 64;
  float lsum = 0.0f;
  for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < chunks; c += nwaves) {
    const uint64_t i = c * 64 + lane;
    const bool ok = i < D.n;
    float x[SEL_IN], z1[SEL_H], h1[SEL_H], z2[SEL_H], h2[SEL_H], o[SEL_OUT], g[SEL_OUT];
#pragma unroll
    for (int f = 0; f < SEL_IN; f++) x[f] = ok ? (D.x[i * SEL_IN + f] - norm[f]) * norm[SEL_IN + f] : 0.0f;
#pragma unroll
    for (int j = 0; j < SEL_H; j++) {
      float a = W[SEL_OFF_B1 + j];
#pragma unroll
      for (int f = 0; f < SEL_IN; f++) a += W[SEL_OFF_W1 + j * SEL_IN + f] * x[f];
      z1[j] = a;
      h1[j] = a > 0.0f ? a : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < SEL_H; j++) {
      float a = W[SEL_OFF_B2 + j];
#pragma unroll
      for (int f = 0; f < SEL_H; f++) a += W[SEL_OFF_W2 + j * SEL_H + f] * h1[f];
      z2[j] = a;
      h2[j] = a > 0.0f ? a : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < SEL_OUT; k++) {
      float a = W[SEL_OFF_B3 + k];
#pragma unroll
      for (int f = 0; f < SEL_H; f++) a += W[SEL_OFF_W3 + k * SEL_H + f] * h2[f];
      o[k] = a;
      const float m = ok ? D.m[i * SEL_OUT + k] : 0.0f;
      const float e = ok ? (a - D.y[i * SEL_OUT + k]) : 0.0f;
      lsum += m * e * e;
      g[k] = 2.0f * m * e * D.inv_count;
    }
    // layer 3
#pragma unroll
    for (int k = 0; k < SEL_OUT; k++) {
      float gb = wave_sum(g[k]);
      if (lane == 0) atomicAdd(&G[SEL_OFF_B3 + k], gb);
      for (int j = 0; j < SEL_H; j++) {
        float v = wave_sum(g[k] * h2[j]);
        if (lane == 0) atomicAdd(&G[SEL_OFF_W3 + k * SEL_H + j], v);
      }
    }
    float g2[SEL_H];
#pragma unroll
    for (int j = 0; j < SEL_H; j++) {
      float a = 0.0f;
#pragma unroll
      for (int k = 0; k < SEL_OUT; k++) a += g[k] * W[SEL_OFF_W3 + k * SEL_H + j];
      g2[j] = z2[j] > 0.0f ? a : 0.0f;
    }
    // layer 2
    for (int j = 0; j < SEL_H; j++) {
      float gb = wave_sum(g2[j]);
      if (lane == 0) atomicAdd(&G[SEL_OFF_B2 + j], gb);
      for (int f = 0; f < SEL_H; f++) {
        float v = wave_sum(g2[j] * h1[f]);
        if (lane == 0) atomicAdd(&G[SEL_OFF_W2 + j * SEL_H + f], v);
      }
    }
    float g1[SEL_H];
#pragma unroll
    for (int f = 0; f < SEL_H; f++) {
      float a = 0.0f;
      for (int j = 0; j < SEL_H; j++) a += g2[j] * W[SEL_OFF_W2 + j * SEL_H + f];
      g1[f] = z1[f] > 0.0f ? a : 0.0f;
    }
    // layer 1
    for (int j = 0; j < SEL_H; j++) {
      float gb = wave_sum(g1[j]);
      if (lane == 0) atomicAdd(&G[SEL_OFF_B1 + j], gb);
#pragma unroll
      for (int f = 0; f < SEL_IN; f++) {
        float v = wave_sum(g1[j] * x[f]);
        if (lane == 0) atomicAdd(&G[SEL_OFF_W1 + j * SEL_IN + f], v);
      }
    }
  }
  lsum = wave_sum(lsum);
  if (lane == 0) atomicAdd(&s_loss, lsum);
  __syncthreads();
  for (int i = threadIdx.x; i < SEL_NPARAM; i += blockDim.x)
    if (G[i] != 0.0f) atomicAdd(&grad[i], G[i]);
  if (threadIdx.x == 0) atomicAdd(loss, s_loss * D.inv_count);
}

__global__ void k_mlp_predict(const __half *__restrict__ w, const float *__restrict__ norm,
                              const float *__restrict__ x, uint64_t n, float *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    selector_forward(w, norm, x + i * SEL_IN, out + i * SEL_OUT);
}

__global__ void k_to_half(const float *__restrict__ in, __half *__restrict__ out, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = __float2half(in[i]);
}

__global__ void k_set_features(DTrie t, float *__restrict__ out) {
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < t.nh; h += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t a = t.off[h], n = t.off[h + 1] - a;
    const uint32_t *v = t.vals + a;
    const uint32_t first = n ? v[0] : 0, last = n ? v[n - 1] : 0;
    const uint64_t span = n > 1 ? (uint64_t)(last - first) : 0;
    selector_features(n, span, first, last, set_work(t, v, n), out + h * SEL_IN);
  }
}

}  // namespace ehx

using namespace ehx;

struct ehx_mlp {
  ehx_ctx *ctx;
  MlpData d;
  void *mem;
};

extern "C" {

int ehx_trie_set_features(ehx_ctx *ctx, const ehx_trie *t, float *out) {
  if (!ctx || !t || !out) return fail(EHX_ERR_INVALID, "null argument");
  EHX_HIP(hipSetDevice(ctx->device));
  DevBuf &b = ctx->scratch[S_IN0];
  int rc;
  if ((rc = b.ensure(t->nh * SEL_IN * 4 + 4))) return rc;
  k_set_features<<<grid_for(t->nh, 256, 16384), 256, 0, ctx->stream>>>(view(t), b.as<float>());
  EHX_HIP(hipGetLastError());
  return copy_d2h(ctx, out, b.p, t->nh * SEL_IN * 4);
}

int ehx_mlp_create(ehx_ctx *ctx, const float *x, const float *y, const float *m, uint64_t n, ehx_mlp **out) {
  if (!ctx || !x || !y || !m || !out || n == 0) return fail(EHX_ERR_INVALID, "bad training data");
  EHX_HIP(hipSetDevice(ctx->device));
  double cnt = 0;
  for (uint64_t i = 0; i < n * SEL_OUT; i++) cnt += m[i];
  if (cnt <= 0) return fail(EHX_ERR_INVALID, "empty mask");
  ehx_mlp *h = new ehx_mlp();
  h->ctx = ctx;
  const size_t bx = n * SEL_IN * 4, by = n * SEL_OUT * 4;
  hipError_t e = hipMalloc(&h->mem, bx + 2 * by + 256);
  if (e != hipSuccess) { delete h; return fail(EHX_ERR_NOMEM, "mlp data: %s", hipGetErrorString(e)); }
  h->d.x = (float *)h->mem;
  h->d.y = (float *)((char *)h->mem + bx);
  h->d.m = (float *)((char *)h->mem + bx + by);
  h->d.n = n;
  h->d.inv_count = (float)(1.0 / cnt);
  int rc;
  if ((rc = copy_h2d(ctx, h->d.x, x, bx)) || (rc = copy_h2d(ctx, h->d.y, y, by)) || (rc = copy_h2d(ctx, h->d.m, m, by))) {
    (void)hipFree(h->mem);
    delete h;
    return rc;
  }
  *out = h;
  return EHX_OK;
}

int ehx_mlp_destroy(ehx_mlp *h) {
  if (!h) return EHX_OK;
  EHX_HIP(hipSetDevice(h->ctx->device));
  EHX_HIP(hipFree(h->mem));
  delete h;
  return EHX_OK;
}

int ehx_mlp_grad(ehx_mlp *h, const float *params, const float *norm, float *grad, float *loss) {
  if (!h || !params || !norm || !grad || !loss) return fail(EHX_ERR_INVALID, "null argument");
  ehx_ctx *ctx = h->ctx;
  EHX_HIP(hipSetDevice(ctx->device));
  DevBuf &b = ctx->scratch[S_MISC];
  int rc;
  const size_t np = SEL_NPARAM;
  if ((rc = b.ensure((2 * np + 2 * SEL_IN + 4) * 4))) return rc;
  // layout: params[np] | norm[2 * SEL_IN] | grad[np] | loss
  float *dp = b.as<float>(), *dn = dp + np, *dg = dn + 2 * SEL_IN, *dl = dg + np;
  if ((rc = copy_h2d(ctx, dp, params, np * 4)) || (rc = copy_h2d(ctx, dn, norm, 2 * SEL_IN * 4))) return rc;
  if ((rc = fill_u32(ctx, (uint32_t *)dg, 0u, np + 1))) return rc;  // grad and loss start at 0.0f
  const uint64_t chunks = (h->d.n + 63) / 64;
  k_mlp_grad<<<grid_for(chunks * 64, 256, 1024), 256, 0, ctx->stream>>>(dp, dn, h->d, dg, dl);
  EHX_HIP(hipGetLastError());
  if ((rc = copy_d2h(ctx, grad, dg, np * 4))) return rc;
  return copy_d2h(ctx, loss, dl, 4);
}

int ehx_mlp_predict(ehx_ctx *ctx, const float *params, const float *norm, const float *x, uint64_t n, float *out) {
  if (!ctx || !params || !norm || !x || !out) return fail(EHX_ERR_INVALID, "null argument");
  EHX_HIP(hipSetDevice(ctx->device));
  DevBuf &b = ctx->scratch[S_MISC], &bx = ctx->scratch[S_IN0];
  int rc;
  const size_t np = SEL_NPARAM;
  if ((rc = b.ensure((np + 2 * SEL_IN) * 4 + np * 2 + 256)) || (rc = bx.ensure(n * (SEL_IN + SEL_OUT) * 4 + 4))) return rc;
  float *dp = b.as<float>(), *dn = dp + np;
  __half *dh = (__half *)(dn + 2 * SEL_IN);
  float *dx = bx.as<float>(), *dout = dx + n * SEL_IN;
  if ((rc = copy_h2d(ctx, dp, params, np * 4)) || (rc = copy_h2d(ctx, dn, norm, 2 * SEL_IN * 4)) ||
      (rc = copy_h2d(ctx, dx, x, n * SEL_IN * 4)))
    return rc;
  k_to_half<<<8, 256, 0, ctx->stream>>>(dp, dh, (int)np);
  EHX_HIP(hipGetLastError());
  k_mlp_predict<<<grid_for(n, 256, 16384), 256, 0, ctx->stream>>>(dh, dn, dx, n, dout);
  EHX_HIP(hipGetLastError());
  return copy_d2h(ctx, out, dout, n * SEL_OUT * 4);
}

}  // extern "C"
