// ehx_tuned.hip -- degree-binned persistent last level (tuned and learned forms).
//
// Pipeline (all device-side, so the whole level can be captured in a hipGraph):
//   k_classify     estimate the work of each prefix (sampled sum of child-set
//                  sizes over S(p)) and bin it by log2(work), largest first;
//   k_bucket_scan  bin offsets and queue length; resets the queue cursor;
//   k_enqueue      scatter prefix ids into the binned queue;
//   k_tuned        persistent wavefronts pull `grab` prefixes per atomic, stage
//                  S(p) in LDS (bitmap or hash) and stream each child(x) through it.
// Results are exact: the count is the same intersection the plain form
// computes, summed in uint64 (order-independent).
#include <stdio.h>
#include <string.h>

#include "ehx_tuned.h"

namespace ehx {

static TuneParams g_params = {4096, 64, 9, 2, 0, 0};
TuneParams &tune_params() { return g_params; }

bool tuned_applicable(const ehx_plan *p, TunedPlan *tp) {
  const int K = p->num_attrs;
  if (K < 3) return false;
  memset(tp, 0, sizeof(*tp));
  tp->K = K;
  int xcount = 0;
  for (int i = 0; i < p->num_ops[K - 1]; i++) {
    const ehx_op &o = p->ops[K - 1][i];
    if (o.parent == K - 2) { tp->x_rel = o.rel; xcount++; }
    else if (o.parent >= 0) { tp->s_rel[tp->ns] = o.rel; tp->s_parent[tp->ns] = o.parent; tp->ns++; }
    else return false;
  }
  if (xcount != 1 || tp->ns < 1) return false;
  bool used[EHX_MAX_OPS] = {false};
  int matched = 0;
  for (int i = 0; i < p->num_ops[K - 2]; i++) {
    const ehx_op &o = p->ops[K - 2][i];
    if (o.parent < 0) { tp->h_rel[tp->nhf++] = o.rel; continue; }
    bool found = false;
    for (int s = 0; s < tp->ns && !found; s++)
      if (!used[s] && tp->s_rel[s] == o.rel && tp->s_parent[s] == o.parent) { used[s] = true; found = true; }
    if (!found) return false;
    matched++;
  }
  return matched == tp->ns;
}

static constexpr uint32_t EMPTY = 0xFFFFFFFFu;
static constexpr uint32_t NBUCKETS = 64, BUCKET_NONE = 63;

__device__ __forceinline__ uint32_t hash32(uint32_t x, uint32_t shift) { return (x * 0x9E3779B1u) >> shift; }

__device__ __forceinline__ bool resolve_s(const TunedPlan &P, const TunedArgs &A, uint64_t p, DSet *S) {
  uint32_t pv[EHX_MAX_ATTRS];
  for (int j = 0; j <= P.K - 3; j++) pv[j] = A.prefix[j][p];
  for (int i = 0; i < P.ns; i++) {
    const DTrie &t = A.T.t[P.s_rel[i]];
    int32_t h = node_of(t, pv[P.s_parent[i]]);
    if (h < 0) return false;
    S[i] = set_of(t, (uint32_t)h);
    if (S[i].n == 0) return false;
  }
  return true;
}

__device__ __forceinline__ bool head_ok(const TunedPlan &P, const TunedArgs &A, uint32_t x) {
  for (int i = 0; i < P.nhf; i++)
    if (node_of(A.T.t[P.h_rel[i]], x) < 0) return false;
  return true;
}

__global__ void k_classify(TunedPlan P, TunedArgs A, uint32_t *__restrict__ bucket, unsigned *__restrict__ hist) {
  __shared__ unsigned h[NBUCKETS];
  if (threadIdx.x < NBUCKETS) h[threadIdx.x] = 0;
  __syncthreads();
  const DTrie &tx = A.T.t[P.x_rel];
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < A.nprefix;
       p += (uint64_t)gridDim.x * blockDim.x) {
    DSet S[EHX_MAX_OPS];
    uint32_t b = BUCKET_NONE;
    if (resolve_s(P, A, p, S)) {
      int mi = 0;
      for (int i = 1; i < P.ns; i++)
        if (S[i].n < S[mi].n) mi = i;
      const DSet D = S[mi];
      // sample up to 32 elements of the smallest S set
      const uint32_t ns = D.n < 32 ? D.n : 32;
      uint64_t w = 0;
#pragma unroll 8
      for (uint32_t k = 0; k < ns; k++) {
        uint32_t x = D.v[(uint64_t)k * D.n / ns];
        int32_t nx = node_of(tx, x);
        if (nx >= 0) w += tx.off[nx + 1] - tx.off[nx];
      }
      w = w * D.n / (ns ? ns : 1) + D.n;
      if (P.ns > 1) w = w / 2 + 1;
      uint32_t lg = 63 - __clzll(w | 1);
      b = 31 - (lg > 31 ? 31 : lg);
    }
    bucket[p] = b;
    atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  if (threadIdx.x < NBUCKETS && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// ctl layout: [0..63] hist, [64..127] bucket base / cursor, [128] queue length, [129] pull cursor
__global__ void k_bucket_scan(unsigned *__restrict__ ctl) {
  const uint32_t lane = threadIdx.x;
  unsigned c = (lane < BUCKET_NONE) ? ctl[lane] : 0u;
  unsigned incl = wave_incl_scan(c);
  ctl[64 + lane] = incl - c;
  if (lane == 63) { ctl[128] = incl; ctl[129] = 0; }
}

__global__ void k_enqueue(const uint32_t *__restrict__ bucket, uint64_t n, unsigned *__restrict__ ctl,
                          uint32_t *__restrict__ queue) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t b = bucket[p];
    if (b != BUCKET_NONE) queue[atomicAdd(&ctl[64 + b], 1u)] = (uint32_t)p;
  }
}

template <int SLOTS, int VCAP>
__device__ uint64_t tuned_prefix(const TunedPlan &P, const TunedArgs &A, uint32_t p, uint32_t *tab, uint32_t *sv,
                                 uint32_t *cst, uint32_t *cof, int big_deg, int bitmap_all) {
  const uint32_t lane = __lane_id();
  DSet S[EHX_MAX_OPS];
  if (!resolve_s(P, A, p, S)) return 0;
  int mi = 0;
  for (int i = 1; i < P.ns; i++)
    if (S[i].n <= S[mi].n) mi = i;
  const uint32_t *src = nullptr;
  uint32_t sn = 0;
  if (P.ns == 1) {
    src = S[0].v;
    sn = S[0].n;
  } else if (VCAP > 0 && S[mi].n <= (uint32_t)VCAP) {
    int alg;
    sn = wave_intersect<true>(S, P.ns, sv, &alg);
    __syncthreads();
    if (sn == 0) return 0;
    src = sv;
  }
  // staging decision (wave-uniform)
  int mode = 2;  // 0 bitmap, 1 hash, 2 probe the S sets in HBM
  uint32_t bm_base = 0, bm_words = 0, hshift = 0, hmask = 0;
  if (src) {
    const uint32_t first = src[0], last = src[sn - 1];
    const uint32_t w32 = (last >> 5) - (first >> 5) + 1;
    const bool is_bitset = (P.ns == 1) ? (S[0].lay == EHX_RANGE_BITSET)
                                       : (sn > 2 && (uint64_t)(last - first) <= 128ull * sn);
    if (w32 <= (uint32_t)SLOTS && (is_bitset || bitmap_all)) mode = 0;
    else if (sn <= (uint32_t)SLOTS / 2 && last != EMPTY) mode = 1;
    if (mode == 0) {
      bm_base = first >> 5;
      bm_words = w32;
      for (uint32_t i = lane; i < w32; i += 64) tab[i] = 0;
      __syncthreads();
      if (P.ns == 1 && S[0].lay == EHX_RANGE_BITSET) {
        // copy the range_bitset words; 32-bit word i of the bitmap covers
        // values [(bm_base + i) * 32, +32)
        for (uint32_t k = lane; k < S[0].W; k += 64) {
          uint64_t w = S[0].w[k];
          uint32_t i0 = (S[0].start + k) * 2u - bm_base;  // wraps to 0xFFFFFFFF when the low half precedes first
          if (i0 < w32) tab[i0] = (uint32_t)w;
          if (i0 + 1u < w32) tab[i0 + 1u] = (uint32_t)(w >> 32);
        }
      } else {
        for (uint32_t i = lane; i < sn; i += 64) {
          uint32_t x = src[i];
          atomicOr(&tab[(x >> 5) - bm_base], 1u << (x & 31));
        }
      }
    } else if (mode == 1) {
      uint32_t H = 64, lg = 6;
      while (H < 2 * sn) { H <<= 1; lg++; }
      hshift = 32 - lg;
      hmask = H - 1;
      for (uint32_t i = lane; i < H; i += 64) tab[i] = EMPTY;
      __syncthreads();
      for (uint32_t i = lane; i < sn; i += 64) {
        uint32_t x = src[i];
        uint32_t h = hash32(x, hshift);
        while (atomicCAS(&tab[h], EMPTY, x) != EMPTY) h = (h + 1) & hmask;
      }
    }
  }
  __syncthreads();

  auto probe = [&](uint32_t c) -> uint32_t {
    if (mode == 0) {
      uint32_t w = (c >> 5) - bm_base;
      return (w < bm_words) ? ((tab[w] >> (c & 31)) & 1u) : 0u;
    }
    if (mode == 1) {
      uint32_t h = hash32(c, hshift);
      for (;;) {
        uint32_t v = tab[h];
        if (v == c) return 1u;
        if (v == EMPTY) return 0u;
        h = (h + 1) & hmask;
      }
    }
    for (int i = 0; i < P.ns; i++)
      if (!set_contains(S[i], c)) return 0u;
    return 1u;
  };

  const DTrie &tx = A.T.t[P.x_rel];
  const uint32_t *xsrc = src ? src : S[mi].v;
  const uint32_t xn = src ? sn : S[mi].n;
  uint64_t hits = 0;
  for (uint32_t base = 0; base < xn; base += 64) {
    const uint32_t idx = base + lane;
    bool ok = idx < xn;
    const uint32_t x = ok ? xsrc[idx] : 0u;
    if (ok && !src)
      for (int i = 0; i < P.ns && ok; i++)
        if (i != mi) ok = set_contains(S[i], x);
    if (ok) ok = head_ok(P, A, x);
    const int32_t nx = ok ? node_of(tx, x) : -1;
    uint32_t o = 0, d = 0;
    if (nx >= 0) {
      o = tx.off[nx];
      d = tx.off[nx + 1] - o;
    }
    uint64_t big = __ballot(d >= (uint32_t)big_deg);
    while (big) {
      const int l = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const uint32_t bo = __shfl(o, l, 64), bd = __shfl(d, l, 64);
      for (uint32_t e = lane; e < bd; e += 64) hits += probe(tx.vals[bo + e]);
    }
    const uint32_t ds = (d < (uint32_t)big_deg) ? d : 0u;
    const uint32_t incl = wave_incl_scan(ds);
    const uint32_t T = __shfl(incl, 63, 64);
    if (T) {
      cst[lane] = incl - ds;
      cof[lane] = o;
      __syncthreads();
      for (uint32_t e = lane; e < T; e += 64) {
        uint32_t lo = 0, hi = 64;
        while (hi - lo > 1) {
          uint32_t mid = (lo + hi) >> 1;
          if (cst[mid] <= e) lo = mid; else hi = mid;
        }
        hits += probe(tx.vals[cof[lo] + (e - cst[lo])]);
      }
      __syncthreads();
    }
  }
  __syncthreads();
  return hits;
}

template <int SLOTS, int VCAP>
__global__ __launch_bounds__(64) void k_tuned(TunedPlan P, TunedArgs A, const uint32_t *__restrict__ queue,
                                              unsigned *__restrict__ ctl, unsigned long long *__restrict__ acc,
                                              int big_deg, int grab, int bitmap_all) {
  __shared__ uint32_t tab[SLOTS];
  __shared__ uint32_t sv[VCAP > 0 ? VCAP : 1];
  __shared__ uint32_t cst[64], cof[64];
  const uint32_t qlen = ctl[128];
  unsigned long long cnt = 0;
  for (;;) {
    uint32_t q0 = 0;
    if (__lane_id() == 0) q0 = atomicAdd(&ctl[129], (unsigned)grab);
    q0 = __shfl(q0, 0, 64);
    if (q0 >= qlen) break;
    const uint32_t q1 = min(q0 + (uint32_t)grab, qlen);
    for (uint32_t qi = q0; qi < q1; qi++)
      cnt += tuned_prefix<SLOTS, VCAP>(P, A, queue[qi], tab, sv, cst, cof, big_deg, bitmap_all);
  }
  cnt = wave_sum(cnt);
  if (__lane_id() == 0) atomicAdd(acc, cnt);
}

template <int SLOTS>
static void launch_tuned(hipStream_t st, unsigned grid, const TunedPlan &P, const TunedArgs &A, const uint32_t *queue,
                         unsigned *ctl, unsigned long long *acc, const TuneParams &tp) {
  if (P.ns == 1)
    k_tuned<SLOTS, 0><<<grid, 64, 0, st>>>(P, A, queue, ctl, acc, tp.big_deg, tp.grab, tp.bitmap_all);
  else
    k_tuned<SLOTS, 2048><<<grid, 64, 0, st>>>(P, A, queue, ctl, acc, tp.big_deg, tp.grab, tp.bitmap_all);
}

// One cached graph per context: the key is every value baked into the launches.
struct GraphCache {
  bool valid = false;
  int device = -1;
  hipStream_t stream = nullptr;
  unsigned char key[sizeof(TunedPlan) + sizeof(TunedArgs) + sizeof(TuneParams) + 64];
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
};
static GraphCache g_cache;

static int enqueue_level(ehx_ctx *ctx, const TunedPlan &P, const TunedArgs &A, unsigned long long *acc,
                         uint32_t *bucket, unsigned *ctl, uint32_t *queue, const TuneParams &tp) {
  hipStream_t st = ctx->stream;
  int rc;
  if ((rc = fill_u32(ctx, ctl, 0u, 130))) return rc;
  k_classify<<<grid_for(A.nprefix, 256, 16384), 256, 0, st>>>(P, A, bucket, ctl);
  EHX_HIP(hipGetLastError());
  k_bucket_scan<<<1, 64, 0, st>>>(ctl);
  EHX_HIP(hipGetLastError());
  k_enqueue<<<grid_for(A.nprefix, 256, 16384), 256, 0, st>>>(bucket, A.nprefix, ctl, queue);
  EHX_HIP(hipGetLastError());
  unsigned grid = (unsigned)ctx->num_cus * (unsigned)(tp.waves_per_cu > 0 ? tp.waves_per_cu : 1);
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K0], st));
  switch (tp.lds_slots) {
    case 2048: launch_tuned<2048>(st, grid, P, A, queue, ctl, acc, tp); break;
    case 8192: launch_tuned<8192>(st, grid, P, A, queue, ctl, acc, tp); break;
    default: launch_tuned<4096>(st, grid, P, A, queue, ctl, acc, tp); break;
  }
  EHX_HIP(hipGetLastError());
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K1], st));
  return EHX_OK;
}

int tuned_run(ehx_ctx *ctx, const TunedPlan &P, const TunedArgs &A, int form, unsigned long long *acc) {
  (void)form;
  if (A.nprefix == 0) return EHX_OK;
  if (A.nprefix >= (1ull << 32)) return fail(EHX_ERR_UNSUPPORTED, "tuned level: more than 2^32 prefixes");
  const TuneParams tp = g_params;
  DevBuf &bb = ctx->scratch[S_QUEUE], &cb = ctx->scratch[S_MISC];
  int rc;
  if ((rc = bb.ensure(A.nprefix * 8 + 64)) || (rc = cb.ensure(256 * 4))) return rc;
  uint32_t *bucket = bb.as<uint32_t>(), *queue = bucket + A.nprefix;
  unsigned *ctl = cb.as<unsigned>();
  if (!tp.use_graph) return enqueue_level(ctx, P, A, acc, bucket, ctl, queue, tp);

  unsigned char key[sizeof(g_cache.key)];
  memset(key, 0, sizeof(key));
  size_t o = 0;
  memcpy(key + o, &P, sizeof(P)); o += sizeof(P);
  memcpy(key + o, &A, sizeof(A)); o += sizeof(A);
  memcpy(key + o, &tp, sizeof(tp)); o += sizeof(tp);
  void *ptrs[4] = {(void *)acc, (void *)bucket, (void *)ctl, (void *)ctx->stream};
  memcpy(key + o, ptrs, sizeof(ptrs));
  if (!(g_cache.valid && g_cache.device == ctx->device && memcmp(g_cache.key, key, sizeof(key)) == 0)) {
    if (g_cache.exec) { (void)hipGraphExecDestroy(g_cache.exec); g_cache.exec = nullptr; }
    if (g_cache.graph) { (void)hipGraphDestroy(g_cache.graph); g_cache.graph = nullptr; }
    g_cache.valid = false;
    EHX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_level(ctx, P, A, acc, bucket, ctl, queue, tp);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc || e != hipSuccess) {
      // capture refused (e.g. an event record): run this level eagerly from now on
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      g_params.use_graph = 0;
      return enqueue_level(ctx, P, A, acc, bucket, ctl, queue, g_params);
    }
    e = hipGraphInstantiate(&g_cache.exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) { (void)hipGraphDestroy(g); return fail(EHX_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e)); }
    g_cache.graph = g;
    memcpy(g_cache.key, key, sizeof(key));
    g_cache.device = ctx->device;
    g_cache.valid = true;
  }
  EHX_HIP(hipGraphLaunch(g_cache.exec, ctx->stream));
  return EHX_OK;
}

void tuned_forget_graphs() {
  if (g_cache.exec) (void)hipGraphExecDestroy(g_cache.exec);
  if (g_cache.graph) (void)hipGraphDestroy(g_cache.graph);
  g_cache = GraphCache();
}

}  // namespace ehx
