// ehx_join_common.h -- wavefront-level set intersections shared by the join
// kernels.  Every routine is called by all 64 lanes of a wave with wave-uniform
// arguments and returns the wave-uniform result count.  With EMIT, the sorted
// intersection is written to out[0..count) by 64-lane ballot + prefix
// compaction.
//
// Algorithms follow the reference's per-pair choice
// (ops/intersection/set.hpp:53-91, uinteger.hpp:1209-1227):
//   uint x uint, max/min <= 32   -> merge (merge-path split across the 64 lanes)
//   uint x uint, max/min > 32    -> gallop (each lane binary-searches a rare key)
//   uint x bitset                -> probe (hetero.hpp:29-60)
//   bitset x bitset              -> AND + popcount over the word overlap (bitset.hpp:619-661)
//   3+ sets                      -> the smallest set drives, the others are probed
#pragma once

#include "ehx_internal.h"

namespace ehx {

enum { ALG_MERGE = 0, ALG_GALLOP = 1, ALG_PROBE = 2, ALG_AND = 3, ALG_EMPTY = 4 };

__device__ __forceinline__ uint32_t wave_popc_below(uint64_t m) { return (uint32_t)__popcll(m & lanes_below()); }

// Number of A elements among the first d of the merge of A and B (A first on ties).
__device__ __forceinline__ uint32_t merge_split(const uint32_t *A, uint32_t na, const uint32_t *B, uint32_t nb,
                                                uint64_t d) {
  uint32_t lo = (d > nb) ? (uint32_t)(d - nb) : 0u;
  uint32_t hi = (d < na) ? (uint32_t)d : na;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (A[mid] <= B[d - 1 - mid]) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <bool EMIT>
__device__ uint32_t wave_merge(const uint32_t *A, uint32_t na, const uint32_t *B, uint32_t nb, uint32_t *out) {
  const uint64_t total = (uint64_t)na + nb;
  const uint32_t lane = __lane_id();
  const uint64_t d0 = total * lane / 64, d1 = total * (lane + 1) / 64;
  const uint32_t i0 = merge_split(A, na, B, nb, d0);
  const uint32_t i1 = merge_split(A, na, B, nb, d1);
  uint32_t j = (uint32_t)(d0 - i0);
  uint32_t c = 0;
  for (uint32_t i = i0; i < i1; i++) {
    uint32_t a = A[i];
    while (j < nb && B[j] < a) j++;
    if (j < nb && B[j] == a) c++;
  }
  uint32_t incl = wave_incl_scan(c);
  uint32_t tot = __shfl(incl, 63, 64);
  if (EMIT && c) {
    uint32_t o = incl - c;
    j = (uint32_t)(d0 - i0);
    for (uint32_t i = i0; i < i1; i++) {
      uint32_t a = A[i];
      while (j < nb && B[j] < a) j++;
      if (j < nb && B[j] == a) out[o++] = a;
    }
  }
  return tot;
}

template <bool EMIT>
__device__ uint32_t wave_gallop(const uint32_t *R, uint32_t nr, const uint32_t *F, uint32_t nf, uint32_t *out) {
  uint32_t lo = 0, cnt = 0;
  for (uint32_t base = 0; base < nr; base += 64) {
    uint32_t idx = base + __lane_id();
    bool valid = idx < nr;
    uint32_t x = valid ? R[idx] : 0xFFFFFFFFu;
    uint32_t pos = valid ? lower_bound(F, lo, nf, x) : nf;
    bool hit = valid && pos < nf && F[pos] == x;
    uint64_t m = __ballot(hit);
    if (EMIT && hit) out[cnt + wave_popc_below(m)] = x;
    cnt += (uint32_t)__popcll(m);
    lo = __shfl(pos, 63, 64);
    if (lo >= nf) break;
  }
  return cnt;
}

template <bool EMIT>
__device__ uint32_t wave_probe(const DSet &U, const DSet &B, uint32_t *out) {
  uint32_t cnt = 0;
  const uint32_t wend = B.start + B.W;
  for (uint32_t base = 0; base < U.n; base += 64) {
    if ((U.v[base] >> 6) >= wend) break;  // past the bitset window (hetero.hpp:49-51)
    uint32_t idx = base + __lane_id();
    bool valid = idx < U.n;
    uint32_t x = valid ? U.v[idx] : 0u;
    bool hit = valid && bit_test(B, x);
    uint64_t m = __ballot(hit);
    if (EMIT && hit) out[cnt + wave_popc_below(m)] = x;
    cnt += (uint32_t)__popcll(m);
  }
  return cnt;
}

template <bool EMIT>
__device__ uint32_t wave_and(const DSet &A, const DSet &B, uint32_t *out) {
  const uint32_t s = max(A.start, B.start);
  const uint32_t e = min(A.start + A.W, B.start + B.W);
  uint32_t cnt = 0;
  for (uint32_t base = s; base < e; base += 64) {
    uint32_t w = base + __lane_id();
    uint64_t r = (w < e) ? (A.w[w - A.start] & B.w[w - B.start]) : 0ull;
    uint32_t c = (uint32_t)__popcll(r);
    uint32_t incl = wave_incl_scan(c);
    if (EMIT && c) {
      uint32_t o = cnt + incl - c;
      while (r) {
        int b = __builtin_ctzll(r);
        out[o++] = 64u * w + (uint32_t)b;
        r &= r - 1;
      }
    }
    cnt += __shfl(incl, 63, 64);
  }
  return cnt;
}

// The smallest set drives; every other set is probed (bit test or binary search).
template <bool EMIT>
__device__ uint32_t wave_nway(const DSet *S, int n, uint32_t *out) {
  int mi = 0;
  for (int i = 1; i < n; i++)
    if (S[i].n <= S[mi].n) mi = i;
  const DSet D = S[mi];
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < D.n; base += 64) {
    uint32_t idx = base + __lane_id();
    bool hit = idx < D.n;
    uint32_t x = hit ? D.v[idx] : 0u;
    for (int i = 0; i < n && hit; i++)
      if (i != mi) hit = set_contains(S[i], x);
    uint64_t m = __ballot(hit);
    if (EMIT && hit) out[cnt + wave_popc_below(m)] = x;
    cnt += (uint32_t)__popcll(m);
  }
  return cnt;
}

// Learned uint x uint choice: ptab[s * PAIR_RATIO_BINS + r] for a smaller set
// of 2^s .. 2^(s+1) - 1 values and a larger/smaller ratio of 2^r .. 2^(r+1) - 1
// (bins past the last are clamped into it): 0 merge, 1 gallop, 2 the
// reference's ratio rule.
constexpr uint32_t PAIR_SMALL_BINS = EHX_PAIR_SMALL_BINS, PAIR_RATIO_BINS = EHX_PAIR_RATIO_BINS;

__device__ __forceinline__ bool pair_gallop(const uint8_t *ptab, uint32_t rn, uint32_t fn) {
  const bool rule = (uint64_t)fn > 32ull * rn;  // uinteger.hpp:1209-1227
  if (!ptab) return rule;
  const uint32_t sb = min(31u - (uint32_t)__builtin_clz(rn), PAIR_SMALL_BINS - 1);
  const uint32_t rb = min(31u - (uint32_t)__builtin_clz(fn / rn), PAIR_RATIO_BINS - 1);
  const uint32_t c = ptab[sb * PAIR_RATIO_BINS + rb];
  return c < 2u ? c == 1u : rule;
}

// Candidate set of one tuple: the intersection of n resolved sets.  With
// `pick`, a bitset pair whose smaller set holds fewer values than the pair's
// overlapping words is probed value by value instead of AND-ed word by word
// (same result, less traffic for sparse windows); without it the algorithms are
// exactly the reference's per-pair choice.  ptab (learned form, when a pair
// table is loaded): the uint x uint choice of merge or gallop (pair_gallop).
template <bool EMIT>
__device__ __forceinline__ uint32_t wave_intersect(const DSet *S, int n, uint32_t *out, int *alg, bool pick = false,
                                                   const uint8_t *ptab = nullptr) {
  for (int i = 0; i < n; i++)
    if (S[i].n == 0) { *alg = ALG_EMPTY; return 0; }
  if (n == 1) {
    *alg = ALG_MERGE;
    if (EMIT)
      for (uint32_t i = __lane_id(); i < S[0].n; i += 64) out[i] = S[0].v[i];
    return S[0].n;
  }
  if (n == 2) {
    const DSet &A = S[0], &B = S[1];
    if (A.lay == EHX_RANGE_BITSET && B.lay == EHX_RANGE_BITSET) {
      if (pick) {
        const uint32_t s0 = max(A.start, B.start), e0 = min(A.start + A.W, B.start + B.W);
        const uint32_t ov = e0 > s0 ? e0 - s0 : 0u;
        if (min(A.n, B.n) < ov) {
          *alg = ALG_PROBE;
          return (A.n <= B.n) ? wave_probe<EMIT>(A, B, out) : wave_probe<EMIT>(B, A, out);
        }
      }
      *alg = ALG_AND;
      return wave_and<EMIT>(A, B, out);
    }
    if (A.lay == EHX_RANGE_BITSET) { *alg = ALG_PROBE; return wave_probe<EMIT>(B, A, out); }
    if (B.lay == EHX_RANGE_BITSET) { *alg = ALG_PROBE; return wave_probe<EMIT>(A, B, out); }
    const DSet &r = (A.n > B.n) ? B : A;
    const DSet &f = (A.n > B.n) ? A : B;
    if (pair_gallop(ptab, r.n, f.n)) { *alg = ALG_GALLOP; return wave_gallop<EMIT>(r.v, r.n, f.v, f.n, out); }
    *alg = ALG_MERGE;
    return wave_merge<EMIT>(r.v, r.n, f.v, f.n, out);
  }
  *alg = ALG_PROBE;
  return wave_nway<EMIT>(S, n, out);
}

// Orders the LDS accesses of one wavefront's lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// ---- plan on the device --------------------------------------------------------
struct LevelDesc {
  int32_t nops;
  int32_t rel[EHX_MAX_OPS];
  int32_t parent[EHX_MAX_OPS];
  uint32_t value[EHX_MAX_OPS];  // the constant of an EHX_PARENT_CONST operand
};

struct Tries {
  DTrie t[EHX_MAX_RELS];
};

// Resolve the operand sets of a level for a tuple; false if a child is absent.
__device__ __forceinline__ bool resolve(const Tries &T, const LevelDesc &L, const uint32_t *tuple, DSet *S) {
  for (int i = 0; i < L.nops; i++) {
    const DTrie &t = T.t[L.rel[i]];
    if (L.parent[i] >= 0) {
      int32_t h = node_of(t, tuple[L.parent[i]]);
      if (h < 0) return false;  // NULL TrieBlock: empty (TrieBuilder.cpp:249-254)
      S[i] = set_of(t, (uint32_t)h);
    } else if (L.parent[i] == EHX_PARENT_HEAD) {
      S[i] = set_of(t, t.nh);
    } else {  // EHX_PARENT_CONST: a selection
      int32_t h = node_of(t, L.value[i]);
      if (h < 0) return false;
      S[i] = set_of(t, (uint32_t)h);
    }
  }
  return true;
}

// Per-segment COUNT output: a result tuple whose first attribute value v lies in
// [bounds[s], bounds[s+1]) adds to out[s].  Many independent queries packed into
// one trie (disjoint value ranges) get their own counts from one launch.
struct SegOut {
  const uint32_t *bounds;   // nseg + 1 ascending level-0 values (device)
  unsigned long long *out;  // nseg counters (device); null when no segments were asked for
  uint32_t nseg;
  uint32_t pad;
};

__device__ __forceinline__ int32_t seg_of(const SegOut &s, uint32_t v) {
  if (v < s.bounds[0] || v >= s.bounds[s.nseg]) return -1;
  uint32_t lo = 0, hi = s.nseg;  // last s with bounds[s] <= v
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s.bounds[mid] <= v) lo = mid;
    else hi = mid;
  }
  return (int32_t)lo;
}

__device__ __forceinline__ void seg_add(const SegOut &s, uint32_t v, unsigned long long c) {
  if (!c) return;
  const int32_t i = seg_of(s, v);
  if (i >= 0) atomicAdd(&s.out[i], c);
}

}  // namespace ehx
