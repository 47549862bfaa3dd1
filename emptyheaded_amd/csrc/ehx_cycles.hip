// ehx_cycles.hip -- undirected 4-cycle COUNT over a symmetric relation.
//
// The reference engine has no 4-cycle query, so the count is defined here: the
// number of 4-cycles of the undirected graph, each cycle once.  A cycle is
// charged to its largest vertex a and to the vertex c opposite a.  With N⁻(a)
// the neighbors of a below a,
//     cycles = sum_a sum_{c < a} C(|N⁻(a) ∩ N(c)|, 2),
// and |N⁻(a) ∩ N(c)| is the number of wedges a-b-c with b in N⁻(a), c in N(b),
// c < a.  The work per a is its wedge bound sum_{b in N⁻(a)} deg(b); with
// vertices numbered by ascending degree that is the O(m sqrt(m)) wedge bound.
//
// Two bins by wedge bound:
//   <= C4_WAVE_MAX  one wavefront per a; endpoints counted in an LDS hash table
//                   of >= 2x the bound slots (load <= 1/2, so it cannot fill);
//   larger          one workgroup per a (queue); endpoints counted in dense LDS
//                   counters over the window [lowest endpoint, a) when it has at
//                   most C4_DENSE ids, else in a global-memory hash table in
//                   hash-partitioned passes.
#include <stdio.h>
#include <string.h>

#include "ehx_join_common.h"

namespace ehx {

constexpr uint32_t C4_SLOTS = 1024;            // LDS hash slots per wavefront
constexpr uint32_t C4_WAVE_MAX = C4_SLOTS / 2;  // wedge bound handled by one wavefront
constexpr uint32_t C4_DENSE = 8192;            // dense LDS counters per workgroup
constexpr uint32_t C4_GSLOTS = 1u << 19;       // global table slots per workgroup
constexpr uint32_t C4_EMPTY = 0xFFFFFFFFu;
// ctl words
enum { C4_BIG_COUNT = 0, C4_BIG_NEXT = 1, C4_ERR = 2, C4_CTL_WORDS = 4 };

__device__ __forceinline__ uint32_t c4_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ uint64_t pairs_of(uint32_t k) { return (uint64_t)k * (k - (k ? 1u : 0u)) / 2; }

// The relation linked to itself (LINK): for the element at position e of
// vals, lnk[e] = its child list {off, length, ...} and hnt[e] = how many values
// of that list lie below the head value of the set holding e.  For b at
// position e of a's set, N(b)'s endpoints below a are then vals[lnk[e].x, +
// hnt[e]): two loads at e instead of a value -> node lookup and a binary
// search, each a dependent load.
struct C4Link {
  const uint4 *lnk;
  const uint32_t *hnt;
};

// Wedge bound of a over its lower neighbors b = vals[b0..b1), summed by the
// wave: the endpoints below a (LINK), or deg(b).
template <bool LINK>
__device__ __forceinline__ uint64_t c4_wedge_bound(const DTrie &S, const C4Link &L, uint32_t b0, uint32_t b1) {
  uint64_t t = 0;
  for (uint32_t bi = b0 + __lane_id(); bi < b1; bi += 64) {
    if (LINK) {
      t += L.hnt[bi];
    } else {
      const int32_t hb = node_of(S, S.vals[bi]);
      if (hb >= 0) t += S.off[hb + 1] - S.off[hb];
    }
  }
  return wave_sum(t);
}

// Calls f(c) for every wedge endpoint c < a of the lower neighbors vals[b0..b1)
// (b1 - b0 <= 64).  Lane l owns b = vals[b0 + l] and N(b) cut at a (a binary
// search: the lists are sorted, so the endpoints below a are a prefix).  The
// lists are streamed by groups of G lanes, 64/G lists at a time, G sized to
// about four endpoints per lane per list (adjacent b's have similar degrees),
// with four loads in flight per lane.
template <bool LINK, class F>
__device__ __forceinline__ void c4_wedges64(const DTrie &S, const C4Link &L, uint32_t a, uint32_t b0, uint32_t b1,
                                            F f) {
  const uint32_t lane = __lane_id();
  const uint32_t bi = b0 + lane;
  uint32_t p0 = 0, len = 0;
  if (bi < b1) {
    if (LINK) {
      p0 = L.lnk[bi].x;
      len = L.hnt[bi];
    } else {
      const int32_t hb = node_of(S, S.vals[bi]);
      if (hb >= 0) {
        p0 = S.off[hb];
        len = lower_bound(S.vals, p0, S.off[hb + 1], a) - p0;
      }
    }
  }
  const uint32_t nl = b1 - b0;
  const uint32_t avg = wave_sum(len) / nl;
  uint32_t slg = 2;
  while (slg < 6 && (4u << slg) < avg) slg++;
  const uint32_t G = 1u << slg, qt = lane >> slg, ql = lane & (G - 1u), lists = 64u >> slg;
  for (uint32_t r0 = 0; r0 < nl; r0 += lists) {
    const uint32_t k = r0 + qt;
    // shuffles on every lane: a source lane outside the active set would read as 0
    const uint32_t po = __shfl(p0, k < nl ? k : 0, 64);
    const uint32_t lk = __shfl(len, k < nl ? k : 0, 64);
    const uint32_t ln = k < nl ? lk : 0u;
    for (uint32_t i = ql; i < ln; i += 4 * G) {
      const uint32_t c0 = S.vals[po + i];
      const uint32_t c1 = i + G < ln ? S.vals[po + i + G] : C4_EMPTY;
      const uint32_t c2 = i + 2 * G < ln ? S.vals[po + i + 2 * G] : C4_EMPTY;
      const uint32_t c3 = i + 3 * G < ln ? S.vals[po + i + 3 * G] : C4_EMPTY;
      f(c0);  // every loaded endpoint is below a
      if (c1 != C4_EMPTY) f(c1);
      if (c2 != C4_EMPTY) f(c2);
      if (c3 != C4_EMPTY) f(c3);
    }
  }
}

// Open-addressing insert (count + 1) into a table of mask + 1 slots.
__device__ __forceinline__ bool c4_insert(uint32_t *K, uint32_t *Cn, uint32_t mask, uint32_t c) {
  uint32_t h = c4_hash(c) & mask;
  for (uint32_t probe = 0; probe <= mask; probe++) {
    const uint32_t old = atomicCAS(&K[h], C4_EMPTY, c);
    if (old == C4_EMPTY || old == c) {
      atomicAdd(&Cn[h], 1u);
      return true;
    }
    h = (h + 1) & mask;
  }
  return false;
}

// One wavefront per vertex a with a small wedge bound; larger ones are queued.
template <bool LINK>
__global__ __launch_bounds__(256) void k_c4_wave(DTrie S, C4Link L, uint64_t begin, uint64_t end, SegOut seg,
                                                 unsigned long long *__restrict__ acc, uint32_t *__restrict__ big,
                                                 uint32_t *__restrict__ ctl) {
  __shared__ uint32_t keys[4][C4_SLOTS];
  __shared__ uint32_t cnts[4][C4_SLOTS];
  const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
  uint32_t *K = keys[w], *Cn = cnts[w];
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long total = 0;
  bool lost = false;
  for (uint64_t i = begin + (((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6); i < end; i += nwaves) {
    const uint32_t a = S.head[i];
    const uint32_t o0 = S.off[i];
    const uint32_t lb = lower_bound(S.vals, o0, S.off[i + 1], a);
    if (lb == o0) continue;
    const uint64_t bound = c4_wedge_bound<LINK>(S, L, o0, lb);
    if (bound < 2) continue;
    if (bound > C4_WAVE_MAX) {
      if (lane == 0) big[atomicAdd(&ctl[C4_BIG_COUNT], 1u)] = (uint32_t)i;
      continue;
    }
    uint32_t ts = 64;
    while (ts < 2 * bound) ts <<= 1;  // <= C4_SLOTS: load stays <= 1/2
    for (uint32_t j = lane; j < ts; j += 64) {
      K[j] = C4_EMPTY;
      Cn[j] = 0;
    }
    wave_sync();
    for (uint32_t b0 = o0; b0 < lb; b0 += 64)
      c4_wedges64<LINK>(S, L, a, b0, min(b0 + 64, lb),
                        [&](uint32_t c) { lost |= !c4_insert(K, Cn, ts - 1, c); });
    wave_sync();  // every lane's counts are in before they are summed
    uint64_t s = 0;
    for (uint32_t j = lane; j < ts; j += 64) s += pairs_of(Cn[j]);
    s = wave_sum(s);
    if (lane == 0) {
      total += s;
      if (seg.out) seg_add(seg, a, s);
    }
    wave_sync();
  }
  if (lost) atomicOr(&ctl[C4_ERR], 1u);
  if (lane == 0 && total) atomicAdd(acc, total);
}

__device__ __forceinline__ uint64_t block_sum(uint64_t v, unsigned long long *red) {
  v = wave_sum(v);
  __syncthreads();
  if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__device__ __forceinline__ uint32_t block_min(uint32_t v, unsigned long long *red) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = min(v, (uint32_t)__shfl_xor(v, d, 64));
  __syncthreads();
  if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (uint32_t)min(min(red[0], red[1]), min(red[2], red[3]));
}

// One workgroup per queued vertex.
template <bool LINK>
__global__ __launch_bounds__(256) void k_c4_block(DTrie S, C4Link L, const uint32_t *__restrict__ big, uint32_t *__restrict__ ctl,
                                                  uint32_t *__restrict__ gkeys, uint32_t *__restrict__ gcnts,
                                                  SegOut seg, unsigned long long *__restrict__ acc) {
  __shared__ uint32_t H[C4_DENSE];
  __shared__ uint32_t s_item;
  __shared__ unsigned long long red[4];
  const uint32_t w = threadIdx.x >> 6;
  uint32_t *K = gkeys + (uint64_t)blockIdx.x * C4_GSLOTS, *Cn = gcnts + (uint64_t)blockIdx.x * C4_GSLOTS;
  const uint32_t n_big = ctl[C4_BIG_COUNT];
  unsigned long long total = 0;
  bool lost = false;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&ctl[C4_BIG_NEXT], 1u);
    __syncthreads();
    const uint32_t qi = s_item;
    __syncthreads();
    if (qi >= n_big) break;
    const uint32_t i = big[qi];
    const uint32_t a = S.head[i];
    const uint32_t o0 = S.off[i];
    const uint32_t lb = lower_bound(S.vals, o0, S.off[i + 1], a);
    uint64_t part = 0;
    uint32_t lowest = C4_EMPTY;
    for (uint32_t bi = o0 + threadIdx.x; bi < lb; bi += 256) {
      if (LINK) {
        const uint32_t h = L.hnt[bi];
        if (h) {
          part += h;
          lowest = min(lowest, S.vals[L.lnk[bi].x]);
        }
      } else {
        const int32_t hb = node_of(S, S.vals[bi]);
        if (hb >= 0 && S.off[hb + 1] > S.off[hb]) {
          part += S.off[hb + 1] - S.off[hb];
          lowest = min(lowest, S.vals[S.off[hb]]);
        }
      }
    }
    const uint64_t bound = block_sum(part, red);
    lowest = block_min(lowest, red);
    uint64_t s = 0;
    if (lowest < a && a - lowest <= C4_DENSE) {
      // dense counters over the endpoint window [lowest, a)
      const uint32_t R = a - lowest;
      for (uint32_t j = threadIdx.x; j < R; j += 256) H[j] = 0;
      __syncthreads();
      for (uint32_t b0 = o0 + 64 * w; b0 < lb; b0 += 256)
        c4_wedges64<LINK>(S, L, a, b0, min(b0 + 64, lb),
                          [&](uint32_t c) { atomicAdd(&H[c - lowest], 1u); });
      __syncthreads();
      uint64_t t = 0;
      for (uint32_t j = threadIdx.x; j < R; j += 256) t += pairs_of(H[j]);
      s = block_sum(t, red);
    } else if (lowest < a) {
      // global hash table; passes hold about bound / P endpoints each at load <= 1/2
      const uint32_t P = (uint32_t)max((2 * bound + C4_GSLOTS - 1) / C4_GSLOTS, (uint64_t)1);
      const uint64_t per = (2 * bound + P - 1) / P;
      uint32_t ts = 256;
      while (ts < C4_GSLOTS && ts < per) ts <<= 1;
      for (uint32_t p = 0; p < P; p++) {
        for (uint32_t j = threadIdx.x; j < ts; j += 256) {
          K[j] = C4_EMPTY;
          Cn[j] = 0;
        }
        __threadfence();
        __syncthreads();
        for (uint32_t b0 = o0 + 64 * w; b0 < lb; b0 += 256)
          c4_wedges64<LINK>(S, L, a, b0, min(b0 + 64, lb), [&](uint32_t c) {
            if (P == 1 || c4_hash(c ^ 0x5BD1E995u) % P == p) lost |= !c4_insert(K, Cn, ts - 1, c);
          });
        __threadfence();
        __syncthreads();
        uint64_t t = 0;
        for (uint32_t j = threadIdx.x; j < ts; j += 256)
          t += pairs_of(__hip_atomic_load(&Cn[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        s += block_sum(t, red);
      }
    }
    __syncthreads();  // H and the table are reused by the next item
    if (threadIdx.x == 0) {
      total += s;
      if (seg.out) seg_add(seg, a, s);
    }
  }
  if (lost) atomicOr(&ctl[C4_ERR], 1u);
  if (threadIdx.x == 0 && total) atomicAdd(acc, total);
}

int cycle4_count(ehx_ctx *ctx, const ehx_trie *t, uint64_t begin, uint64_t end, const SegOut *seg, int64_t *result,
                 ehx_join_stats *stats) {
  if (!ctx || !t || !result) return fail(EHX_ERR_INVALID, "null argument");
  if (t->device != ctx->device) return fail(EHX_ERR_INVALID, "trie lives on another device");
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (end > t->nh) end = t->nh;
  if (begin > end) begin = end;
  const uint64_t n = end - begin;
  if (end >= (1ull << 32)) return fail(EHX_ERR_UNSUPPORTED, "4-cycle count: more than 2^32 head values");
  ehx_join_stats S;
  memset(&S, 0, sizeof(S));
  S.frontier[0] = n;
  DevBuf &ac = ctx->scratch[S_ACC], &cb = ctx->scratch[S_MISC], &ob = ctx->scratch[S_QUEUE];
  int rc;
  if ((rc = ac.ensure(64)) || (rc = cb.ensure(256 * 4)) || (rc = ob.ensure(n * 4 + 4))) return rc;
  if ((rc = fill_u64(ctx, ac.as<uint64_t>(), 0, 8)) || (rc = fill_u32(ctx, cb.as<uint32_t>(), 0, C4_CTL_WORDS))) return rc;
  const DTrie D = view(t);
  const SegOut so = seg ? *seg : SegOut{};
  // the relation's child pointers into itself with their hints (built once and
  // cached on the trie, a derived index like its child lookup); without hints
  // (none kept, or no room) the kernels look children up and search instead
  C4Link L = {nullptr, nullptr};
  if ((rc = trie_link(ctx, const_cast<ehx_trie *>(t), t, &L.lnk, &L.hnt))) return rc;
  const bool link = L.lnk && L.hnt;
  EHX_HIP(hipEventRecord(ctx->ev[0], st));
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K0], st));
  if (n) {
    const unsigned g = grid_for(n * 64, 256, (unsigned)ctx->num_cus * 8);
    if (link)
      k_c4_wave<true><<<g, 256, 0, st>>>(D, L, begin, end, so, ac.as<unsigned long long>(), ob.as<uint32_t>(),
                                         cb.as<uint32_t>());
    else
      k_c4_wave<false><<<g, 256, 0, st>>>(D, L, begin, end, so, ac.as<unsigned long long>(), ob.as<uint32_t>(),
                                          cb.as<uint32_t>());
    EHX_HIP(hipGetLastError());
  }
  uint32_t ctl[C4_CTL_WORDS];
  if ((rc = copy_d2h(ctx, ctl, cb.p, sizeof(ctl)))) return rc;
  if (ctl[C4_ERR]) return fail(EHX_ERR_DEVICE, "4-cycle count: endpoint table overflow");
  S.frontier[1] = ctl[C4_BIG_COUNT];
  if (ctl[C4_BIG_COUNT]) {
    const uint32_t gmax = 4u * (uint32_t)ctx->num_cus;  // four 32 KB-LDS workgroups per CU
    const unsigned grid = ctl[C4_BIG_COUNT] < gmax ? ctl[C4_BIG_COUNT] : gmax;
    DevBuf &kb = ctx->scratch[S_SORT_A], &nb = ctx->scratch[S_SORT_B];
    if ((rc = kb.ensure((uint64_t)grid * C4_GSLOTS * 4)) || (rc = nb.ensure((uint64_t)grid * C4_GSLOTS * 4))) return rc;
    if (link)
      k_c4_block<true><<<grid, 256, 0, st>>>(D, L, ob.as<uint32_t>(), cb.as<uint32_t>(), kb.as<uint32_t>(),
                                             nb.as<uint32_t>(), so, ac.as<unsigned long long>());
    else
      k_c4_block<false><<<grid, 256, 0, st>>>(D, L, ob.as<uint32_t>(), cb.as<uint32_t>(), kb.as<uint32_t>(),
                                              nb.as<uint32_t>(), so, ac.as<unsigned long long>());
    EHX_HIP(hipGetLastError());
    if ((rc = copy_d2h(ctx, ctl, cb.p, sizeof(ctl)))) return rc;
    if (ctl[C4_ERR]) return fail(EHX_ERR_DEVICE, "4-cycle count: endpoint table overflow");
  }
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K1], st));
  unsigned long long accv;
  if ((rc = copy_d2h(ctx, &accv, ac.p, 8))) return rc;
  float ms = 0;
  if (hipEventElapsedTime(&ms, ctx->ev[EHX_EV_K0], ctx->ev[EHX_EV_K1]) == hipSuccess) S.main_kernel_ms = ms;
  S.total_ms = S.main_kernel_ms;
  S.form_used = EHX_FORM_TUNED;
  *result = (int64_t)accv;
  if (stats) *stats = S;
  return EHX_OK;
}

int seg_begin(ehx_ctx *ctx, const uint32_t *bounds, uint64_t nseg, SegOut *so);
int seg_end(ehx_ctx *ctx, const SegOut &so, int64_t *counts);

int cycle4_count_segments(ehx_ctx *ctx, const ehx_trie *t, const uint32_t *bounds, uint64_t nseg, int64_t *counts,
                          int64_t *result, ehx_join_stats *stats) {
  if (!ctx) return fail(EHX_ERR_INVALID, "null ctx");
  EHX_HIP(hipSetDevice(ctx->device));
  SegOut so;
  int rc;
  if ((rc = seg_begin(ctx, bounds, nseg, &so))) return rc;
  if ((rc = cycle4_count(ctx, t, 0, UINT64_MAX, &so, result, stats))) return rc;
  return seg_end(ctx, so, counts);
}

}  // namespace ehx
