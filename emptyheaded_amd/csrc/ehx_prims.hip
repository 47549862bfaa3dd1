// ehx_prims.hip -- hand-written device primitives for gfx950: exclusive scans,
// stable LSD radix sort of 64-bit keys, max reduction, fills.
//
// Scan: three phases (block sums -> recursive scan of the sums -> block scan
// plus base), 256 threads x 8 items per block.  Radix sort: 8-bit digits,
// 4096-key tiles; per pass a tile histogram, one exclusive scan over the
// digit-major histogram, and a stable scatter whose in-tile ranks come from
// 64-lane ballots (8 ballots match the lanes holding the same digit).
#include <stdarg.h>
#include <stdio.h>

#include <string.h>

#include <atomic>
#include <thread>
#include <vector>

#include "ehx_internal.h"

namespace ehx {

static constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

template <class Tin, class Tout>
__global__ __launch_bounds__(256) void k_block_sums(const Tin *__restrict__ in, Tout *__restrict__ sums,
                                                    uint64_t n) {
  __shared__ Tout part[SCAN_T / 64];
  uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  Tout s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_I; i++) {
    uint64_t idx = base + (uint64_t)i * SCAN_T + threadIdx.x;
    if (idx < n) s += (Tout)in[idx];
  }
  s = wave_sum(s);
  if (__lane_id() == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    Tout t = 0;
    for (int w = 0; w < SCAN_T / 64; w++) t += part[w];
    sums[blockIdx.x] = t;
  }
}

// out[i] = base + exclusive prefix within the tile; if last tile, out[n] = total.
template <class Tin, class Tout>
__global__ __launch_bounds__(256) void k_scan_apply(const Tin *__restrict__ in, Tout *__restrict__ out,
                                                    const Tout *__restrict__ bases, uint64_t n) {
  __shared__ Tout wtot[SCAN_T / 64];
  uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  // blocked arrangement: thread t owns items base + t*SCAN_I .. +SCAN_I
  Tout v[SCAN_I];
  Tout s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_I; i++) {
    uint64_t idx = base + (uint64_t)threadIdx.x * SCAN_I + i;
    v[i] = (idx < n) ? (Tout)in[idx] : (Tout)0;
    s += v[i];
  }
  Tout incl = wave_incl_scan(s);
  const int w = threadIdx.x >> 6;
  if (__lane_id() == 63) wtot[w] = incl;
  __syncthreads();
  Tout wbase = bases ? bases[blockIdx.x] : (Tout)0;
  for (int j = 0; j < w; j++) wbase += wtot[j];
  Tout run = wbase + incl - s;
#pragma unroll
  for (int i = 0; i < SCAN_I; i++) {
    uint64_t idx = base + (uint64_t)threadIdx.x * SCAN_I + i;
    if (idx < n) out[idx] = run;
    run += v[i];
    if (idx + 1 == n) out[n] = run;
  }
}

template <class T> __global__ void k_fill(T *p, T v, uint64_t n);

template <class Tin, class Tout>
static int scan_impl(ehx_ctx *ctx, const Tin *in, Tout *out, uint64_t n, int level) {
  if (n == 0) {  // out[0] = 0 by a fill kernel (memset nodes misbehave under graph replay)
    k_fill<Tout><<<1, 64, 0, ctx->stream>>>(out, (Tout)0, 1);
    EHX_HIP(hipGetLastError());
    return EHX_OK;
  }
  uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    k_scan_apply<Tin, Tout><<<1, SCAN_T, 0, ctx->stream>>>(in, out, (const Tout *)nullptr, n);
    EHX_HIP(hipGetLastError());
    return EHX_OK;
  }
  if (level >= 4) return fail(EHX_ERR_UNSUPPORTED, "scan: input too large (%llu)", (unsigned long long)n);
  DevBuf &sb = ctx->scratch[S_SCAN0 + level];
  int rc = sb.ensure(sizeof(Tout) * (nb + 1));
  if (rc) return rc;
  Tout *sums = sb.as<Tout>();
  k_block_sums<Tin, Tout><<<(unsigned)nb, SCAN_T, 0, ctx->stream>>>(in, sums, n);
  EHX_HIP(hipGetLastError());
  rc = scan_impl<Tout, Tout>(ctx, sums, sums, nb, level + 1);  // in-place is safe: apply reads then writes own tile
  if (rc) return rc;
  k_scan_apply<Tin, Tout><<<(unsigned)nb, SCAN_T, 0, ctx->stream>>>(in, out, sums, n);
  EHX_HIP(hipGetLastError());
  return EHX_OK;
}

int scan_reserve(ehx_ctx *ctx, uint64_t n) {
  for (int level = 0; level < 4; level++) {
    const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb <= 1) return EHX_OK;
    int rc = ctx->scratch[S_SCAN0 + level].ensure(8 * (nb + 1));
    if (rc) return rc;
    n = nb;
  }
  return EHX_OK;
}

int scan_u32_to_u64(ehx_ctx *ctx, const uint32_t *in, uint64_t *out, uint64_t n) {
  return scan_impl<uint32_t, uint64_t>(ctx, in, out, n, 0);
}
int scan_u64(ehx_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n) {
  return scan_impl<uint64_t, uint64_t>(ctx, in, out, n, 0);
}
int scan_u32(ehx_ctx *ctx, const uint32_t *in, uint32_t *out, uint64_t n) {
  return scan_impl<uint32_t, uint32_t>(ctx, in, out, n, 0);
}

// ---- radix sort ---------------------------------------------------------------
static constexpr int RS_WAVES = 4, RS_ROUNDS = 16, RS_TILE = RS_WAVES * 64 * RS_ROUNDS;

__global__ __launch_bounds__(256) void k_rs_hist(const uint64_t *__restrict__ keys, uint64_t n, int shift,
                                                 uint32_t *__restrict__ hist, uint32_t ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int i = 0; i < RS_TILE / 256; i++) {
    uint64_t idx = base + (uint64_t)i * 256 + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_rs_scatter(const uint64_t *__restrict__ keys, uint64_t *__restrict__ out,
                                                    uint64_t n, int shift, const uint32_t *__restrict__ ofs,
                                                    uint32_t ntiles) {
  __shared__ uint32_t wcnt[RS_WAVES][256];
  __shared__ uint32_t tbase[256];
  const int w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < RS_WAVES * 256; i += 256) (&wcnt[0][0])[i] = 0;
  tbase[threadIdx.x] = ofs[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
  __syncthreads();
  const uint64_t sub = (uint64_t)blockIdx.x * RS_TILE + (uint64_t)w * 64 * RS_ROUNDS;
  const uint64_t below = lanes_below();
  uint64_t k[RS_ROUNDS];
  uint32_t loc[RS_ROUNDS];
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; r++) {
    uint64_t idx = sub + (uint64_t)r * 64 + __lane_id();
    bool valid = idx < n;
    k[r] = valid ? keys[idx] : 0;
    uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t bb = __ballot((d >> b) & 1u);
      m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t rank = (uint32_t)__popcll(m & below);
    uint32_t before = wcnt[w][d];
    loc[r] = before + rank;
    if (valid && rank == 0) wcnt[w][d] = before + (uint32_t)__popcll(m);
  }
  __syncthreads();
  {
    uint32_t run = 0;
    for (int j = 0; j < RS_WAVES; j++) {
      uint32_t c = wcnt[j][threadIdx.x];
      wcnt[j][threadIdx.x] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; r++) {
    uint64_t idx = sub + (uint64_t)r * 64 + __lane_id();
    if (idx < n) {
      uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
      out[(uint64_t)tbase[d] + wcnt[w][d] + loc[r]] = k[r];
    }
  }
}

int radix_sort_u64(ehx_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, int bits, int from, uint64_t **result) {
  if (result) *result = keys;
  if (n <= 1 || bits <= from) return EHX_OK;
  if (n >= (1ull << 32)) return fail(EHX_ERR_UNSUPPORTED, "radix sort: more than 2^32 keys");
  uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  uint64_t hn = (uint64_t)ntiles * 256;
  DevBuf &hb = ctx->scratch[S_HIST];
  int rc = hb.ensure(sizeof(uint32_t) * (2 * hn + 2));
  if (rc) return rc;
  uint32_t *hist = hb.as<uint32_t>(), *ofs = hist + hn + 1;
  uint64_t *src = keys, *dst = tmp;
  const int passes = (bits - from + 7) / 8;
  for (int p = 0; p < passes; p++) {
    const int shift = from + 8 * p;
    k_rs_hist<<<ntiles, 256, 0, ctx->stream>>>(src, n, shift, hist, ntiles);
    EHX_HIP(hipGetLastError());
    rc = scan_u32(ctx, hist, ofs, hn);
    if (rc) return rc;
    k_rs_scatter<<<ntiles, 256, 0, ctx->stream>>>(src, dst, n, shift, ofs, ntiles);
    EHX_HIP(hipGetLastError());
    uint64_t *t = src; src = dst; dst = t;
  }
  if (result) *result = src;
  else if (src != keys) EHX_HIP(hipMemcpyAsync(keys, src, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
  return EHX_OK;
}

// ---- reductions and fills ------------------------------------------------------
__global__ __launch_bounds__(256) void k_max_u32(const uint32_t *__restrict__ in, uint64_t n,
                                                 uint32_t *__restrict__ out) {
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    m = max(m, in[i]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor(m, d, 64));
  if (__lane_id() == 0) atomicMax(out, m);
}

int reduce_max_u32(ehx_ctx *ctx, const uint32_t *in, uint64_t n, uint32_t *host_out) {
  DevBuf &b = ctx->scratch[S_MISC];
  int rc = b.ensure(64);
  if (rc) return rc;
  rc = fill_u32(ctx, b.as<uint32_t>(), 0, 1);
  if (rc) return rc;
  if (n) {
    k_max_u32<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(in, n, b.as<uint32_t>());
    EHX_HIP(hipGetLastError());
  }
  return copy_d2h(ctx, host_out, b.p, 4);
}

template <class T> __global__ void k_fill(T *p, T v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int fill_u32(ehx_ctx *ctx, uint32_t *p, uint32_t v, uint64_t n) {
  if (!n) return EHX_OK;
  k_fill<uint32_t><<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(p, v, n);
  EHX_HIP(hipGetLastError());
  return EHX_OK;
}
int fill_u64(ehx_ctx *ctx, uint64_t *p, uint64_t v, uint64_t n) {
  if (!n) return EHX_OK;
  k_fill<uint64_t><<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(p, v, n);
  EHX_HIP(hipGetLastError());
  return EHX_OK;
}

__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// A plain copy of known size: reads and writes `bytes` bytes once each (the
// reference point for scaling HBM traffic counters).
int copy_calibrate(ehx_ctx *ctx, uint64_t bytes, double *ms) {
  bytes &= ~(uint64_t)15;
  DevBuf &a = ctx->scratch[S_SORT_A], &b = ctx->scratch[S_SORT_B];
  int rc;
  if ((rc = a.ensure(bytes)) || (rc = b.ensure(bytes))) return rc;
  if ((rc = fill_u64(ctx, a.as<uint64_t>(), 0x0123456789ABCDEFull, bytes / 8))) return rc;
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K0], ctx->stream));
  k_copy16<<<grid_for(bytes / 16, 256, 8192), 256, 0, ctx->stream>>>(a.as<uint4>(), b.as<uint4>(), bytes / 16);
  EHX_HIP(hipGetLastError());
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K1], ctx->stream));
  EHX_HIP(hipEventSynchronize(ctx->ev[EHX_EV_K1]));
  float t = 0;
  EHX_HIP(hipEventElapsedTime(&t, ctx->ev[EHX_EV_K0], ctx->ev[EHX_EV_K1]));
  if (ms) *ms = t;
  return EHX_OK;
}

int copy_d2h(ehx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!bytes) return EHX_OK;
  EHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  EHX_HIP(hipStreamSynchronize(ctx->stream));
  return EHX_OK;
}
int copy_h2d(ehx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!bytes) return EHX_OK;
  EHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  EHX_HIP(hipStreamSynchronize(ctx->stream));
  return EHX_OK;
}

int upload_h2d(ehx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (bytes < 2 * UPLOAD_CHUNK) {
    if (bytes) EHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return EHX_OK;
  }
  if (!ctx->pin[0][0]) {
    for (int t = 0; t < UPLOAD_THREADS; t++)
      for (int b = 0; b < 2; b++) {
        hipError_t e = hipHostMalloc(&ctx->pin[t][b], UPLOAD_CHUNK, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->pin_ev[t][b], hipEventDisableTiming);
        if (e != hipSuccess) {
          (void)hipGetLastError();
          upload_release(ctx);  // no pinned memory: the runtime stages the copy itself
          EHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
          return EHX_OK;
        }
      }
  }
  // thread t copies bytes [lo_t, hi_t) through its two buffers in turn, waiting
  // for a buffer's previous copy (this call's or an earlier one's) before refilling it
  const size_t per = (bytes + UPLOAD_THREADS - 1) / UPLOAD_THREADS;
  std::vector<hipError_t> err(UPLOAD_THREADS, hipSuccess);
  auto work = [&](int t) {
    const size_t lo = std::min(bytes, per * t), hi = std::min(bytes, lo + per);
    err[t] = hipSetDevice(ctx->device);  // a new thread starts on device 0
    int b = 0;
    for (size_t o = lo; o < hi && err[t] == hipSuccess; o += UPLOAD_CHUNK, b ^= 1) {
      const size_t n = std::min(UPLOAD_CHUNK, hi - o);
      hipError_t e = hipEventSynchronize(ctx->pin_ev[t][b]);
      if (e == hipSuccess) {
        memcpy(ctx->pin[t][b], (const char *)src + o, n);
        e = hipMemcpyAsync((char *)dst + o, ctx->pin[t][b], n, hipMemcpyHostToDevice, ctx->stream);
      }
      if (e == hipSuccess) e = hipEventRecord(ctx->pin_ev[t][b], ctx->stream);
      err[t] = e;
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < UPLOAD_THREADS; t++) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  for (hipError_t e : err)
    if (e != hipSuccess) return fail(EHX_ERR_DEVICE, "upload: %s", hipGetErrorString(e));
  return EHX_OK;
}

void upload_release(ehx_ctx *ctx) {
  for (int t = 0; t < UPLOAD_THREADS; t++)
    for (int b = 0; b < 2; b++) {
      if (ctx->pin_ev[t][b]) {
        (void)hipEventSynchronize(ctx->pin_ev[t][b]);
        (void)hipEventDestroy(ctx->pin_ev[t][b]);
      }
      if (ctx->pin[t][b]) (void)hipHostFree(ctx->pin[t][b]);
      ctx->pin_ev[t][b] = nullptr;
      ctx->pin[t][b] = nullptr;
    }
}

// Bumped whenever a context buffer is freed or moves: a captured graph that
// names buffer addresses is replayed only while the count it was captured at
// holds.
static std::atomic<uint64_t> g_devbuf_gen{0};
uint64_t devbuf_generation() { return g_devbuf_gen.load(); }

int DevBuf::ensure(size_t bytes) {
  if (bytes <= cap) return EHX_OK;
  g_devbuf_gen.fetch_add(1);
  if (p) {
    hipError_t e = hipFree(p);
    if (e != hipSuccess) return fail(EHX_ERR_DEVICE, "hipFree: %s", hipGetErrorString(e));
    p = nullptr;
    cap = 0;
  }
  size_t want = bytes + bytes / 4 + 256;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    return fail(EHX_ERR_NOMEM, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
  }
  cap = want;
  return EHX_OK;
}

void DevBuf::release() {
  if (p) g_devbuf_gen.fetch_add(1);
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

}  // namespace ehx
