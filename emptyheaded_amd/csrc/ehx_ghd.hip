// Top bag of a two-bag plan: the sum over the tuples (a, x) of a relation of
// wa[a] * wb[x], where wa and wb are per-value annotations that lower bags
// produced (the generated annotated fold, query_compiler/src/main/scala/codegen/
// query/EHGenerator.scala:399-434, over the intermediate bag tries of
// EHGenerator.scala:711-726).  LollipopAgg = sum over Edge(a,x) of T(a) (with
// wb = 1) and BarbellAgg = sum over Edge(a,x) of T(a) T(x), T(v) = COUNT(*) of
// the triangle bag grouped by its first attribute.
#include <hip/hip_runtime.h>

#include "ehx_internal.h"
#include "ehx_join_common.h"

namespace ehx {

// A wavefront per head: its lanes sum wb over the child set, lane 0 multiplies
// by wa of the head.  Values at or beyond na / nb weigh 0.  Exact modulo 2^64.
__global__ __launch_bounds__(256) void k_weighted_pairs(DTrie t, const int64_t *__restrict__ wa, uint64_t na,
                                                        const int64_t *__restrict__ wb, uint64_t nb,
                                                        unsigned long long *__restrict__ acc) {
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long total = 0;
  for (uint64_t h = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; h < t.nh; h += nwaves) {
    const uint32_t a = t.head[h];
    const unsigned long long w = a < na ? (unsigned long long)wa[a] : 0ull;
    if (!w) continue;  // wave-uniform
    const uint32_t o = t.off[h], n = t.off[h + 1] - o;
    unsigned long long s = 0;
    for (uint32_t k = __lane_id(); k < n; k += 64) {
      const uint32_t x = t.vals[o + k];
      if (x < nb) s += (unsigned long long)wb[x];
    }
    total += w * s;  // lanes' partial sums times the same weight add up to w * sum
  }
  total = wave_sum(total);
  if (__lane_id() == 0 && total) atomicAdd(acc, total);
}

int weighted_pairs(ehx_ctx *ctx, const ehx_trie *t, const int64_t *wa, uint64_t na, const int64_t *wb, uint64_t nb,
                   int64_t *result) {
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int rc;
  DevBuf &ba = ctx->scratch[S_IN0], &bb = ctx->scratch[S_IN1], &ac = ctx->scratch[S_ACC];
  if ((rc = ba.ensure(na * 8 + 8)) || (rc = bb.ensure(nb * 8 + 8)) || (rc = ac.ensure(64))) return rc;
  if (na) EHX_HIP(hipMemcpyAsync(ba.p, wa, na * 8, hipMemcpyHostToDevice, st));
  if (nb) EHX_HIP(hipMemcpyAsync(bb.p, wb, nb * 8, hipMemcpyHostToDevice, st));
  if ((rc = fill_u64(ctx, ac.as<uint64_t>(), 0, 1))) return rc;
  k_weighted_pairs<<<grid_for(t->nh * 64, 256, 16384), 256, 0, st>>>(view(t), ba.as<int64_t>(), na, bb.as<int64_t>(),
                                                                     nb, ac.as<unsigned long long>());
  EHX_HIP(hipGetLastError());
  uint64_t r = 0;
  if ((rc = copy_d2h(ctx, &r, ac.p, 8))) return rc;
  *result = (int64_t)r;
  return EHX_OK;
}

// Rows (x, y) of trie t with x < y, compacted into lo0/lo1 (any order); cnt[0]
// counts them, cnt[1] the rows with x > y, cnt[2] the rows with x == y.  One
// wavefront per head, its lanes over the child set 64 at a time.
__global__ __launch_bounds__(256) void k_orient(DTrie t, uint32_t *__restrict__ lo0, uint32_t *__restrict__ lo1,
                                                unsigned *__restrict__ cnt) {
  const uint32_t lane = __lane_id();
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  unsigned n_hi = 0, n_eq = 0;
  for (uint64_t h = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; h < t.nh; h += nwaves) {
    const uint32_t x = t.head[h], o = t.off[h], n = t.off[h + 1] - o;
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
      const uint32_t k = k0 + lane;
      const uint32_t y = k < n ? t.vals[o + k] : 0u;
      const bool lo = k < n && x < y;
      n_hi += (k < n && x > y) ? 1u : 0u;
      n_eq += (k < n && x == y) ? 1u : 0u;
      const uint64_t m = __ballot(lo);
      unsigned base = 0;
      if (lane == 0 && m) base = atomicAdd(&cnt[0], (unsigned)__popcll(m));
      base = __shfl(base, 0, 64);
      if (lo) {
        const uint32_t p = base + wave_popc_below(m);
        lo0[p] = x;
        lo1[p] = y;
      }
    }
  }
  n_hi = wave_sum(n_hi);
  n_eq = wave_sum(n_eq);
  if (lane == 0) {
    if (n_hi) atomicAdd(&cnt[1], n_hi);
    if (n_eq) atomicAdd(&cnt[2], n_eq);
  }
}

int trie_orient(ehx_ctx *ctx, const ehx_trie *t, int policy, ehx_trie **lower, ehx_trie **reverse,
                uint64_t *counts) {
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int rc;
  const uint64_t m = t->m;
  if (m >= (1ull << 32)) return fail(EHX_ERR_UNSUPPORTED, "more than 2^32 rows");
  DevBuf &b0 = ctx->scratch[S_F0], &b1 = ctx->scratch[S_F1], &bc = ctx->scratch[S_SEGO];
  if ((rc = b0.ensure(m * 4 + 4)) || (rc = b1.ensure(m * 4 + 4)) || (rc = bc.ensure(64))) return rc;
  if ((rc = fill_u32(ctx, bc.as<uint32_t>(), 0u, 4))) return rc;
  if (m) {
    k_orient<<<grid_for(t->nh * 64, 256, 16384), 256, 0, st>>>(view(t), b0.as<uint32_t>(), b1.as<uint32_t>(),
                                                               bc.as<unsigned>());
    EHX_HIP(hipGetLastError());
  }
  uint32_t c[3] = {0, 0, 0};
  if ((rc = copy_d2h(ctx, c, bc.p, 12))) return rc;
  if (counts) {
    counts[0] = c[0];
    counts[1] = c[1];
    counts[2] = c[2];
  }
  *lower = *reverse = nullptr;
  if (c[0] == 0) return EHX_OK;  // no rows with x < y: no tries
  if ((rc = ehx_trie_build(ctx, b0.as<uint32_t>(), b1.as<uint32_t>(), c[0], policy, 1, lower))) return rc;
  if ((rc = ehx_trie_build(ctx, b1.as<uint32_t>(), b0.as<uint32_t>(), c[0], policy, 1, reverse))) {
    ehx_trie_destroy(*lower);
    *lower = nullptr;
    return rc;
  }
  return EHX_OK;
}

}  // namespace ehx
