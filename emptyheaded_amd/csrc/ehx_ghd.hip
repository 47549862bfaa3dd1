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

// Orienting trie t: the rows (x, y) with x < y are, for each head x, the
// suffix of x's sorted child set after the values <= x.  k_orient_count finds
// that suffix per head by binary search (lcnt[h] = its length) and counts the
// rows with x > y (cnt[1]) and x == y (cnt[2]), one atomic per block.
__global__ __launch_bounds__(256) void k_orient_count(DTrie t, uint32_t *__restrict__ lcnt, unsigned *__restrict__ cnt) {
  __shared__ unsigned red[2][4];
  unsigned n_hi = 0, n_eq = 0;
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < t.nh; h += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t x = t.head[h], o = t.off[h], n = t.off[h + 1] - o;
    uint32_t lo = 0, hi = n;  // first index whose value is >= x
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (t.vals[o + mid] < x) lo = mid + 1;
      else hi = mid;
    }
    const uint32_t eq = (lo < n && t.vals[o + lo] == x) ? 1u : 0u;
    n_hi += lo;
    n_eq += eq;
    lcnt[h] = n - lo - eq;
  }
  n_hi = wave_sum(n_hi);
  n_eq = wave_sum(n_eq);
  const uint32_t w = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    red[0][w] = n_hi;
    red[1][w] = n_eq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const unsigned e = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    if (a) atomicAdd(&cnt[1], a);
    if (e) atomicAdd(&cnt[2], e);
  }
}

// The rows with x < y into lo0/lo1 at their place in (x, y) order: pos (nh + 1
// entries) is the exclusive scan of lcnt.  Each wave fills ORIENT_CHUNK output
// rows; a lane finds the head of its row by binary search over pos restricted
// to the heads the chunk spans (none when one head, a hub, fills the chunk).
constexpr uint32_t ORIENT_CHUNK = 1024;
// The last index i in [lo, hi] with pos[i] <= r, given pos[lo] <= r and pos
// non-decreasing.
__device__ __forceinline__ uint32_t last_at_most(const uint32_t *__restrict__ pos, uint32_t lo, uint32_t hi,
                                                 uint32_t r) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (pos[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
__global__ __launch_bounds__(256) void k_orient_copy(DTrie t, const uint32_t *__restrict__ lcnt,
                                                     const uint32_t *__restrict__ pos, uint32_t total,
                                                     uint32_t *__restrict__ lo0, uint32_t *__restrict__ lo1) {
  const uint32_t lane = __lane_id();
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t nchunks = ((uint64_t)total + ORIENT_CHUNK - 1) / ORIENT_CHUNK;
  for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < nchunks; c += nwaves) {
    const uint32_t r0 = (uint32_t)(c * ORIENT_CHUNK);
    const uint32_t r1 = (uint32_t)((uint64_t)total < (c + 1) * ORIENT_CHUNK ? total : (c + 1) * ORIENT_CHUNK);
    // the heads of rows r0 and r1 - 1 (pos[0] = 0 <= r0, pos[nh] = total > r1 - 1)
    const uint32_t hs = last_at_most(pos, 0, (uint32_t)t.nh, r0);
    const uint32_t he = last_at_most(pos, hs, (uint32_t)t.nh, r1 - 1);
    for (uint32_t r = r0 + lane; r < r1; r += 64) {
      const uint32_t lo = last_at_most(pos, hs, he, r);
      const uint32_t s = t.off[lo + 1] - lcnt[lo];
      lo0[r] = t.head[lo];
      lo1[r] = t.vals[s + (r - pos[lo])];
    }
  }
}

// Symmetry check over the rows (x, y), x < y, that k_orient_copy wrote: a row
// whose reverse (y, x) is absent from t (y is no head, or x is not in y's
// child set) is a miss; cnt[3] += misses, one atomic per block.  With as many
// rows x > y as x < y and no miss, every row x > y is the reverse of one x < y
// row, so the relation is symmetric.
__global__ __launch_bounds__(256) void k_orient_check(DTrie t, const uint32_t *__restrict__ lo0,
                                                      const uint32_t *__restrict__ lo1, uint32_t total,
                                                      unsigned *__restrict__ cnt) {
  __shared__ unsigned red[4];
  unsigned miss = 0;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < total; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t x = lo0[r], y = lo1[r];
    const int32_t ny = node_of(t, y);
    if (ny < 0) {
      ++miss;
      continue;
    }
    const uint32_t o = t.off[ny], e = t.off[ny + 1];
    const uint32_t i = lower_bound(t.vals, o, e, x);
    miss += (i < e && t.vals[i] == x) ? 0u : 1u;
  }
  miss = wave_sum(miss);
  if (__lane_id() == 0) red[threadIdx.x >> 6] = miss;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned a = red[0] + red[1] + red[2] + red[3];
    if (a) atomicAdd(&cnt[3], a);
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
  uint32_t c[4] = {0, 0, 0, 0};
  if (m) {
    const uint64_t nh = t->nh;
    DevBuf &bl = ctx->scratch[S_F2], &bp = ctx->scratch[S_F3];
    if ((rc = bl.ensure(nh * 4 + 4)) || (rc = bp.ensure(nh * 4 + 8))) return rc;
    k_orient_count<<<grid_for(nh, 256, 2048), 256, 0, st>>>(view(t), bl.as<uint32_t>(), bc.as<unsigned>());
    EHX_HIP(hipGetLastError());
    if ((rc = scan_u32(ctx, bl.as<uint32_t>(), bp.as<uint32_t>(), nh))) return rc;
    if ((rc = copy_d2h(ctx, &c[0], bp.as<uint32_t>() + nh, 4))) return rc;
    if (c[0])
      k_orient_copy<<<grid_for(((uint64_t)c[0] + ORIENT_CHUNK - 1) / ORIENT_CHUNK * 64, 256, 16384), 256, 0, st>>>(
          view(t), bl.as<uint32_t>(), bp.as<uint32_t>(), c[0], b0.as<uint32_t>(), b1.as<uint32_t>());
    EHX_HIP(hipGetLastError());
    if (c[0])
      k_orient_check<<<grid_for(c[0], 256, 16384), 256, 0, st>>>(view(t), b0.as<uint32_t>(), b1.as<uint32_t>(), c[0],
                                                                bc.as<unsigned>());
    EHX_HIP(hipGetLastError());
    if ((rc = copy_d2h(ctx, &c[1], bc.as<unsigned>() + 1, 12))) return rc;
  }
  if (counts)
    for (int i = 0; i < 4; ++i) counts[i] = c[i];
  if (lower) *lower = nullptr;
  if (reverse) *reverse = nullptr;
  // the halves are built only on request and only for a symmetric relation
  // without loops (the one case the oriented triangle passes are exact for)
  if (!lower || !reverse || c[0] == 0 || c[1] != c[0] || c[2] != 0 || c[3] != 0) return EHX_OK;
  if ((rc = ehx_trie_build(ctx, b0.as<uint32_t>(), b1.as<uint32_t>(), c[0], policy, 1, lower))) return rc;
  if ((rc = ehx_trie_build(ctx, b1.as<uint32_t>(), b0.as<uint32_t>(), c[0], policy, 1, reverse))) {
    ehx_trie_destroy(*lower);
    *lower = nullptr;
    return rc;
  }
  return EHX_OK;
}

}  // namespace ehx
