// Unweighted single-source closure (the reference's tc::unweighted_single_source,
// storage_engine/include/TransitiveClosure.hpp:12-123, which recursive queries
// such as SSSP(x;y)*[c=0] :- Edge(w,x),SSSP(w),y<-[1+MIN(w;1)] run): a frontier
// expanded level by level over the child sets of a trie, each vertex taking the
// level at which it is first reached (the start: 0).  Where the reference marks
// a visited bitset without atomics and lets a vertex enter the frontier twice,
// here one compare-and-swap on the distance claims it.
#include <hip/hip_runtime.h>

#include "ehx_internal.h"
#include "ehx_join_common.h"

namespace ehx {

constexpr uint32_t UNREACHED = 0xFFFFFFFFu;

__global__ void k_set_u32(uint32_t *__restrict__ p, uint64_t i, uint32_t v) { p[i] = v; }

// One wavefront per frontier vertex; its lanes stream the child set 64 at a
// time and append the children they claim to the next frontier (one atomic per
// wave per step).
__global__ __launch_bounds__(256) void k_bfs_level(DTrie t, const uint32_t *__restrict__ fr, uint32_t fn,
                                                   uint32_t *__restrict__ dist, uint64_t nd, uint32_t level,
                                                   uint32_t *__restrict__ nxt, unsigned *__restrict__ ncnt) {
  const uint32_t lane = __lane_id();
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < fn; i += nwaves) {
    const int32_t h = node_of(t, fr[i]);
    if (h < 0) continue;  // no child set (wave-uniform)
    const uint32_t o = t.off[h], n = t.off[h + 1] - o;
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
      const uint32_t k = k0 + lane;
      bool claim = false;
      uint32_t x = 0;
      if (k < n) {
        x = t.vals[o + k];
        if (x < nd && dist[x] == UNREACHED) claim = atomicCAS(&dist[x], UNREACHED, level) == UNREACHED;
      }
      const uint64_t m = __ballot(claim);
      unsigned base = 0;
      if (lane == 0 && m) base = atomicAdd(ncnt, (unsigned)__popcll(m));
      base = __shfl(base, 0, 64);
      if (claim) nxt[base + wave_popc_below(m)] = x;
    }
  }
}

// Top-down step for a frontier of at most WIDE_FRONTIER vertices (the start,
// typically a hub): every wave of the grid takes a strided share of each
// frontier vertex's child set, so a large set is streamed by the whole chip
// rather than by one wave.  Claims and appends as in k_bfs_level.
constexpr uint32_t WIDE_FRONTIER = 32;
__global__ __launch_bounds__(256) void k_bfs_level_wide(DTrie t, const uint32_t *__restrict__ fr, uint32_t fn,
                                                        uint32_t *__restrict__ dist, uint64_t nd, uint32_t level,
                                                        uint32_t *__restrict__ nxt, unsigned *__restrict__ ncnt) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint32_t i = 0; i < fn; i++) {
    const int32_t h = node_of(t, fr[i]);
    if (h < 0) continue;  // no child set (grid-uniform)
    const uint32_t o = t.off[h], n = t.off[h + 1] - o;
    for (uint64_t k0 = wave * 64; k0 < n; k0 += nwaves * 64) {
      const uint64_t k = k0 + lane;
      bool claim = false;
      uint32_t x = 0;
      if (k < n) {
        x = t.vals[o + k];
        if (x < nd && dist[x] == UNREACHED) claim = atomicCAS(&dist[x], UNREACHED, level) == UNREACHED;
      }
      const uint64_t m = __ballot(claim);
      unsigned base = 0;
      if (lane == 0 && m) base = atomicAdd(ncnt, (unsigned)__popcll(m));
      base = __shfl(base, 0, 64);
      if (claim) nxt[base + wave_popc_below(m)] = x;
    }
  }
}

// Bottom-up step (direction-optimizing BFS, for large frontiers): one lane per
// unreached vertex v < nd scans its in-neighbours (v's child set in the reverse
// trie) for one at the frontier's level and stops at the first; v then takes
// level + 1 and joins the next frontier.  A vertex claimed in this step holds
// level + 1, never `level`, so plain stores suffice.
__global__ __launch_bounds__(256) void k_bfs_bottom_up(DTrie r, uint32_t *__restrict__ dist, uint64_t nd,
                                                       uint32_t level, uint32_t *__restrict__ nxt,
                                                       unsigned *__restrict__ ncnt) {
  const uint32_t lane = __lane_id();
  for (uint64_t v0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ull; v0 < nd;
       v0 += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = v0 + lane;
    bool found = false;
    if (v < nd && dist[v] == UNREACHED) {
      const int32_t h = node_of(r, (uint32_t)v);
      if (h >= 0)
        for (uint32_t k = r.off[h], e = r.off[h + 1]; k < e; k++) {
          const uint32_t u = r.vals[k];
          if (u < nd && dist[u] == level) {
            found = true;
            break;
          }
        }
      if (found) dist[v] = level + 1;
    }
    const uint64_t m = __ballot(found);
    unsigned base = 0;
    if (lane == 0 && m) base = atomicAdd(ncnt, (unsigned)__popcll(m));
    base = __shfl(base, 0, 64);
    if (found) nxt[base + wave_popc_below(m)] = (uint32_t)v;
  }
}

// Top-down steps, and bottom-up ones over the reverse trie `rev` while the
// frontier holds at least nd / bottom_up vertices (bottom_up > 0).
int sssp_directed(ehx_ctx *ctx, const ehx_trie *t, const ehx_trie *rev, uint32_t start, int iterations,
                  uint32_t *dist_out, uint64_t nd, uint32_t *levels_out, int bottom_up) {
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int rc;
  if (nd == 0 || start >= nd) return fail(EHX_ERR_INVALID, "start %u outside the %llu distances", start,
                                          (unsigned long long)nd);
  if (nd >= (1ull << 32)) return fail(EHX_ERR_UNSUPPORTED, "more than 2^32 values");
  DevBuf &db = ctx->scratch[S_SORT_A], &fa = ctx->scratch[S_SORT_B], &fb = ctx->scratch[S_POS],
         &cb = ctx->scratch[S_MISC];
  if ((rc = db.ensure(nd * 4)) || (rc = fa.ensure(nd * 4 + 4)) || (rc = fb.ensure(nd * 4 + 4)) || (rc = cb.ensure(64)))
    return rc;
  uint32_t *dist = db.as<uint32_t>(), *cur = fa.as<uint32_t>(), *nxt = fb.as<uint32_t>();
  unsigned *cnt = cb.as<unsigned>();
  if ((rc = fill_u32(ctx, dist, UNREACHED, nd))) return rc;
  k_set_u32<<<1, 1, 0, st>>>(dist, start, 0u);
  k_set_u32<<<1, 1, 0, st>>>(cur, 0, start);
  EHX_HIP(hipGetLastError());
  const DTrie dv = view(t);
  const DTrie rv = rev ? view(rev) : dv;
  const uint64_t switch_at = bottom_up > 0 ? nd / (uint64_t)bottom_up : ~0ull;
  uint32_t fn = 1, level = 0;
  while (fn && (iterations < 0 || (int)level < iterations)) {
    if ((rc = fill_u32(ctx, cnt, 0u, 1))) return rc;
    if (rev && fn >= switch_at)
      k_bfs_bottom_up<<<grid_for(nd, 256, 16384), 256, 0, st>>>(rv, dist, nd, level, nxt, cnt);
    else if (fn <= WIDE_FRONTIER)
      k_bfs_level_wide<<<(unsigned)ctx->num_cus * 8, 256, 0, st>>>(dv, cur, fn, dist, nd, level + 1, nxt, cnt);
    else
      k_bfs_level<<<grid_for((uint64_t)fn * 64, 256, 16384), 256, 0, st>>>(dv, cur, fn, dist, nd, level + 1, nxt,
                                                                             cnt);
    EHX_HIP(hipGetLastError());
    if ((rc = copy_d2h(ctx, &fn, cnt, 4))) return rc;
    uint32_t *tmp = cur;
    cur = nxt;
    nxt = tmp;
    level++;
  }
  if (levels_out) *levels_out = level;
  return copy_d2h(ctx, dist_out, dist, nd * 4);
}

int sssp_unweighted(ehx_ctx *ctx, const ehx_trie *t, uint32_t start, int iterations, uint32_t *dist_out, uint64_t nd,
                    uint32_t *levels_out) {
  return sssp_directed(ctx, t, nullptr, start, iterations, dist_out, nd, levels_out, 0);
}

}  // namespace ehx
