// ehx_tuned.hip -- degree-binned persistent last level (tuned and learned forms).
//
// Pipeline (all device-side, so the whole level can be captured in a hipGraph):
//   k_classify     estimate the work of each prefix (sampled sum of child-set
//                  sizes over S(p)) and bin it: a "block" bin for prefixes whose
//                  work or staged set is large, a "wave" bin for the rest, each
//                  ordered by log2(work), largest first;
//   k_bucket_scan  bin offsets, queue length, block-bin size; resets cursors;
//   k_enqueue      scatter prefix ids into the binned queue (one global atomic
//                  per bin per workgroup);
//   k_tuned        persistent 4-wave workgroups.  Phase 1: the workgroup takes
//                  block-bin prefixes one at a time, stages S(p) once in LDS and
//                  its four waves split the child sets.  Phase 2: each wave takes
//                  wave-bin prefixes on its own with a quarter of the LDS.
// Staging: S(p) as an LDS bitmap when it is a bitset whose word window fits,
// else as an open-addressing hash table behind an LDS Bloom-filter bitmap (most
// probes miss and stop at the filter), else probed in HBM.  Child sets shorter
// than big_deg are expanded into an LDS owner window so every lane streams one
// element per step; longer ones are streamed by the whole wave.
// Results are exact: the count is the same intersection the plain form
// computes, summed in uint64 (order-independent).
#include <stdio.h>
#include <string.h>

#include "ehx_tuned.h"

namespace ehx {

static TuneParams g_params = {4096, 32, 24, 2, 0, 0, 4096, 16, 0, 1 << 21, 0};
TuneParams &tune_params() { return g_params; }

bool tuned_applicable(const ehx_plan *p, TunedPlan *tp) {
  const int K = p->num_attrs;
  if (K < 3) return false;
  for (int k = K - 2; k < K; k++)  // selections in the fused levels run in the plain form
    for (int i = 0; i < p->num_ops[k]; i++)
      if (p->ops[k][i].parent == EHX_PARENT_CONST) return false;
  memset(tp, 0, sizeof(*tp));
  tp->K = K;
  int xcount = 0;
  for (int i = 0; i < p->num_ops[K - 1]; i++) {
    const ehx_op &o = p->ops[K - 1][i];
    if (o.parent == K - 2) { tp->x_rel = o.rel; xcount++; }
    else if (o.parent >= 0) { tp->s_rel[tp->ns] = o.rel; tp->s_parent[tp->ns] = o.parent; tp->ns++; }
    else return false;
  }
  if (xcount != 1 || tp->ns < 1 || tp->ns > TUNED_MAX_S) return false;
  bool used[EHX_MAX_OPS] = {false};
  int matched = 0, other = 0;
  for (int i = 0; i < p->num_ops[K - 2]; i++) {
    const ehx_op &o = p->ops[K - 2][i];
    if (o.parent < 0) {
      if (o.rel != tp->x_rel) tp->h_rel[tp->nhf++] = o.rel;  // x's own head is implied by its child lookup
      continue;
    }
    bool found = false;
    for (int s = 0; s < tp->ns && !found; s++)
      if (!used[s] && tp->s_rel[s] == o.rel && tp->s_parent[s] == o.parent) { used[s] = true; found = true; }
    if (found) {
      matched++;
    } else {
      tp->d_rel = o.rel;
      tp->d_parent = o.parent;
      other++;
    }
  }
  if (tp->nhf > TUNED_MAX_H) return false;
  if (other == 0) return matched == tp->ns;
  // one driver set, and no S set among level K-2's operands
  tp->nd = 1;
  return other == 1 && matched == 0;
}

static constexpr uint32_t EMPTY = 0xFFFFFFFFu;
static constexpr uint32_t NBUCKETS = 64, BUCKET_NONE = 63, WAVE_BIN = 32;
// A workgroup-bin prefix of estimated work w is split into up to SPLIT_MAX
// parts of about split_work each, over disjoint ranges of E(p) (each part
// stages S(p) itself); at most SPLIT_EXTRA queue entries are added per level.
static constexpr uint32_t SPLIT_MAX = 64, SPLIT_EXTRA = 1u << 16;
// Single-S prefixes whose estimated work is below tiny_work and whose S(p) holds
// at most TINY_S values go to the tiny bin (bucket TINY_BUCKET, the end of the
// queue), which k_tiny runs up to 64 prefixes per wavefront at a time.
static constexpr uint32_t TINY_BUCKET = 62, TINY_S = 128;
static constexpr uint32_t TINY_U = 4;       // elements per lane per step in k_tiny
static constexpr uint32_t TINY_BATCH = 16;  // queue entries a wavefront takes at once (<= 64)
static constexpr int OWN = 256;  // owner window per wave

void tuned_bind(const TunedPlan &P, const Tries &T, TunedArgs &A) {
  for (int i = 0; i < P.ns && i < TUNED_MAX_S; i++) {
    A.ts[i] = T.t[P.s_rel[i]];
    A.sp[i] = A.prefix[P.s_parent[i]];
  }
  // a single staged set takes its staging layout from the learned selector
  // when its trie has one; every other use keeps the density rule's layout
  if (P.ns == 1 && A.ts[0].stage) A.ts[0].layout = A.ts[0].stage;
  A.tx = T.t[P.x_rel];
  if (P.nd) {
    A.td = T.t[P.d_rel];
    A.dp = A.prefix[P.d_parent];
  }
  A.nh = P.nhf < TUNED_MAX_H ? P.nhf : TUNED_MAX_H;
  for (int i = 0; i < A.nh; i++) A.th[i] = T.t[P.h_rel[i]];
}

// S[i] for a runtime i, as unrolled selects (a runtime index would put S in scratch).
template <int MAXS>
__device__ __forceinline__ DSet sel(const DSet *S, int i) {
  DSet r = S[0];
#pragma unroll
  for (int k = 1; k < MAXS; k++)
    if (i == k) r = S[k];
  return r;
}

template <int MAXS>
__device__ __forceinline__ bool resolve_s(const TunedPlan &P, const TunedArgs &A, uint64_t p, DSet *S) {
#pragma unroll
  for (int i = 0; i < MAXS; i++) {
    if (i >= P.ns) break;
    const int32_t h = node_of(A.ts[i], A.sp[i][p]);
    if (h < 0) return false;
    S[i] = set_of(A.ts[i], (uint32_t)h);
    if (S[i].n == 0) return false;
  }
  return true;
}

// The driver set of a prefix (plans with nd = 1); true with nothing to do otherwise.
__device__ __forceinline__ bool resolve_d(const TunedPlan &P, const TunedArgs &A, uint64_t p, DSet &D) {
  if (!P.nd) return true;
  const int32_t h = node_of(A.td, A.dp[p]);
  if (h < 0) return false;
  D = set_of(A.td, (uint32_t)h);
  return D.n > 0;
}

__device__ __forceinline__ bool head_ok(const TunedPlan &P, const TunedArgs &A, uint32_t x) {
#pragma unroll
  for (int i = 0; i < TUNED_MAX_H; i++)
    if (i < A.nh && node_of(A.th[i], x) < 0) return false;
  return true;
}

// Wave capacity of the hash table for a given LDS slot budget.
__host__ __device__ constexpr uint32_t wave_hash_cap(int slots) { return (uint32_t)slots / 8; }

// 32-bit words of the bitmap staging of a set with values in [first, last]:
// it starts at the 64-bit word holding `first` (an even 32-bit word, so a
// 64-bit bitset word maps onto one aligned 64-bit LDS word) and has an even
// length.  The staging needs one more, zero, word after it.
__host__ __device__ __forceinline__ uint32_t bitmap_words(uint32_t first, uint32_t last) {
  return (((last >> 5) - ((first >> 6) << 1)) | 1u) + 1u;
}

// bucket[p] = bin | parts << 6; hist[b] counts queue entries (parts)
// MAXS: S sets resolved per prefix (1 for single-S plans: no runtime-indexed array)
template <int MAXS>
__global__ void k_classify(TunedPlan P, TunedArgs A, uint32_t *__restrict__ bucket, unsigned *__restrict__ hist,
                           int slots, int big_work, uint4 *__restrict__ pdesc, uint32_t split_work,
                           uint32_t tiny_work) {
  __shared__ unsigned h[NBUCKETS];
  if (threadIdx.x < NBUCKETS) h[threadIdx.x] = 0;
  __syncthreads();
  const DTrie &tx = A.tx;
  const uint64_t np = A.nprefix_dev ? *A.nprefix_dev : A.nprefix;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (uint64_t)gridDim.x * blockDim.x) {
    DSet S[MAXS];
    DSet Dd;
    uint32_t b = BUCKET_NONE, parts = 1;
    if (resolve_s<MAXS>(P, A, p, S) && resolve_d(P, A, p, Dd)) {
      int mi = 0;
      for (int i = 1; i < P.ns && i < MAXS; i++)
        if (S[i].n < S[mi].n) mi = i;
      const DSet D = MAXS == 1 ? S[0] : S[mi];
      const DSet X = P.nd ? Dd : D;  // the set x runs over
      uint64_t w = 0;
      if (pdesc && A.xlink && A.xwork) {  // the set's child-list lengths, summed when it was linked
        const int32_t hx = P.nd ? node_of(A.td, A.dp[p]) : node_of(A.ts[0], A.sp[0][p]);
        w = (hx >= 0 ? A.xwork[hx] : 0ull) + X.n + (P.nd ? D.n : 0u);
      } else {
        const uint32_t ns = X.n < 32 ? X.n : 32;  // sample up to 32 of its elements
        if (pdesc && A.xlink) {  // child pointers: one independent load per sample
          const uint4 *xl = A.xlink + (X.v - (P.nd ? A.td.vals : A.ts[0].vals));
#pragma unroll 8
          for (uint32_t k = 0; k < ns; k++) w += xl[(uint64_t)k * X.n / ns].y;
        } else {
#pragma unroll 8
          for (uint32_t k = 0; k < ns; k++) {
            uint32_t x = X.v[(uint64_t)k * X.n / ns];
            int32_t nx = node_of(tx, x);
            if (nx >= 0) w += tx.off[nx + 1] - tx.off[nx];
          }
        }
        w = w * X.n / (ns ? ns : 1) + X.n + (P.nd ? D.n : 0u);
      }
      if (P.ns > 1) w = w / 2 + 1;
      // the wave bin needs S(p) to fit a quarter of the LDS, as a hash table or as a bitmap
      bool fits_wave = D.n <= wave_hash_cap(slots);
      if (!fits_wave && P.ns == 1 && D.n)  // a wave can stage it as a bitmap
        fits_wave = bitmap_words(D.v[0], D.v[D.n - 1]) < (uint32_t)slots / 4;
      const bool big = w >= (uint64_t)big_work || !fits_wave;
      uint32_t lg = 63 - __clzll(w | 1);
      b = (big ? 0u : WAVE_BIN) + (31 - (lg > 31 ? 31 : lg));
      if (!big && b > TINY_BUCKET - 1) b = TINY_BUCKET - 1;  // the last wave bucket holds w < 4
      if (pdesc && !big && !A.rec && w < tiny_work && D.n <= TINY_S && D.v[D.n - 1] != EMPTY) b = TINY_BUCKET;
      if (pdesc) {  // single-S plans (S[0] = D here); the third and fourth words locate E(p)
        pdesc[2 * p] = make_uint4((uint32_t)(D.v - A.ts[0].vals), D.n,
                                  (uint32_t)(X.v - (P.nd ? A.td.vals : A.ts[0].vals)), X.n);
        pdesc[2 * p + 1] = make_uint4(D.v[0], D.v[D.n - 1], (uint32_t)(D.w - A.ts[0].words), D.lay);
        if (big && split_work && !A.rec && w > split_work) {
          uint64_t want = (w + split_work - 1) / split_work;
          const uint32_t by_x = X.n / 64u;  // at least one full step of a workgroup per part
          if (want > by_x) want = by_x;
          if (want > SPLIT_MAX) want = SPLIT_MAX;
          if (want > 1) {
            const unsigned got = atomicAdd(&hist[132], (unsigned)want - 1u);  // the split budget
            if (got + (unsigned)want - 1u <= SPLIT_EXTRA) parts = (uint32_t)want;
          }
        }
      }
    }
    bucket[p] = b | (parts << 6);
    atomicAdd(&h[b], parts);
  }
  __syncthreads();
  if (threadIdx.x < NBUCKETS && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// ctl: [0..63] hist, [64..127] bin base / cursor, [128] queue length,
//      [129] block-bin cursor, [130] wave-bin cursor, [131] block-bin size,
//      [132] queue entries added by splits, [133] tiny-bin start
__global__ void k_bucket_scan(unsigned *__restrict__ ctl) {
  const uint32_t lane = threadIdx.x;
  unsigned c = (lane < BUCKET_NONE) ? ctl[lane] : 0u;
  unsigned incl = wave_incl_scan(c);
  ctl[64 + lane] = incl - c;
  if (lane == WAVE_BIN - 1) ctl[131] = incl;
  if (lane == TINY_BUCKET) ctl[133] = incl - c;  // where the tiny bin starts
  if (lane == 63) { ctl[128] = incl; ctl[129] = 0; ctl[130] = 0; }
}

__global__ __launch_bounds__(256) void k_enqueue(const uint32_t *__restrict__ bucket, uint64_t n_host,
                                                 const uint64_t *__restrict__ n_dev, unsigned *__restrict__ ctl,
                                                 uint32_t *__restrict__ queue, const uint4 *__restrict__ pdesc,
                                                 uint4 *__restrict__ qdesc) {
  __shared__ unsigned cnt[NBUCKETS], base[NBUCKETS];
  const uint64_t n = n_dev ? *n_dev : n_host;
  for (uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x; t0 < n; t0 += (uint64_t)gridDim.x * blockDim.x) {
    if (threadIdx.x < NBUCKETS) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t p = t0 + threadIdx.x;
    const uint32_t bp = (p < n) ? bucket[p] : BUCKET_NONE;
    const uint32_t b = bp & 63u, parts = bp >> 6;
    unsigned local = 0;
    if (b != BUCKET_NONE) local = atomicAdd(&cnt[b], parts);
    __syncthreads();
    if (threadIdx.x < NBUCKETS && cnt[threadIdx.x]) base[threadIdx.x] = atomicAdd(&ctl[64 + threadIdx.x], cnt[threadIdx.x]);
    __syncthreads();
    if (b != BUCKET_NONE) {
      const uint32_t at = base[b] + local;
      uint4 d0 = make_uint4(0, 0, 0, 0), d1 = d0;
      if (pdesc) {
        d0 = pdesc[2 * p];
        d1 = pdesc[2 * p + 1];
      }
      for (uint32_t k = 0; k < parts; k++) {  // part k takes E(p)[xn * k / parts, xn * (k + 1) / parts)
        queue[at + k] = (uint32_t)p;
        if (pdesc) {
          const uint32_t lo = (uint32_t)((uint64_t)d0.w * k / parts), hi = (uint32_t)((uint64_t)d0.w * (k + 1) / parts);
          qdesc[2ull * (at + k)] = make_uint4(d0.x, d0.y, d0.z + lo, hi - lo);
          qdesc[2ull * (at + k) + 1] = d1;
        }
      }
    }
    __syncthreads();
  }
}

// Staging of one S(p) in LDS.  mode 0 bitmap, 1 hash + filter, 2 probe in HBM.
struct Stage {
  int mode;
  uint32_t bm_base, bm_words;
  uint32_t hshift, hmask, fshift;
  const uint32_t *tab, *flt;
};

__device__ __forceinline__ uint32_t hash_t(uint32_t c, uint32_t shift) { return (c * 0x9E3779B1u) >> shift; }
__device__ __forceinline__ uint32_t hash_f(uint32_t c, uint32_t shift) { return (c * 0x85EBCA77u) >> shift; }

template <int MODE, int MAXS>
__device__ __forceinline__ uint32_t probe_m(const Stage &s, const DSet *S, int ns, uint32_t c) {
  if (MODE == 0) {  // branch-free: an index past the window clamps to the zero word tab[bm_words]
    const uint32_t w = min((c >> 5) - s.bm_base, s.bm_words);
    return __builtin_amdgcn_ubfe(s.tab[w], c, 1u);  // bit c & 31 of the word
  }
  if (MODE == 1) {
    const uint32_t f = hash_f(c, s.fshift);
    if (!((s.flt[f >> 5] >> (f & 31)) & 1u)) return 0u;
    uint32_t h = hash_t(c, s.hshift);
    for (;;) {  // EMPTY first: a probed 0xFFFFFFFF (never staged) must not match a free slot
      const uint32_t v = s.tab[h];
      if (v == EMPTY) return 0u;
      if (v == c) return 1u;
      h = (h + 1) & s.hmask;
    }
  }
#pragma unroll
  for (int i = 0; i < MAXS; i++)
    if (i < ns && !set_contains(S[i], c)) return 0u;
  return 1u;
}

// Bitmap probe of a value held relative to the bitmap's first value
// (r = c - 32 * bm_base): one subtraction fewer per probe than probe_m<0>.
// Bit c & 31 is bit r & 31, since the base is a multiple of 32; a value below
// the window wraps to a large r and, like one above it, reads the zero word.
__device__ __forceinline__ uint32_t probe_rebased(const Stage &s, uint32_t r) {
  const uint32_t w = min(r >> 5, s.bm_words);
  return __builtin_amdgcn_ubfe(s.tab[w], r, 1u);
}

// Probes the 15 values of one gap-coded chunk (a, b): the first value, then
// the 14 reached by its 16-bit gaps; `last` receives the probe of the 15th.
// The last chunk of a list is padded with zero gaps, so its padding repeats
// the list's last value: the caller takes `last` once per padding slot off that
// chunk's count.
template <int MODE, int MAXS>
__device__ __forceinline__ uint32_t probe_chunk(const Stage &st, const DSet *S, int ns, const uint4 a, const uint4 b,
                                                uint32_t &last) {
  if (MODE == 0) {  // the running sum of the gaps is kept relative to the bitmap
    uint32_t r = a.x - (st.bm_base << 5);
    uint32_t h = probe_rebased(st, r), p = 0;
    const uint32_t w[7] = {a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int j = 0; j < 7; j++) {
      r += w[j] & 0xFFFFu;
      h += probe_rebased(st, r);
      r += w[j] >> 16;
      p = probe_rebased(st, r);
      h += p;
    }
    last = p;
    return h;
  }
  uint32_t v = a.x;
  uint32_t h = probe_m<MODE, MAXS>(st, S, ns, v), p = 0;
  if (MODE != 2) {
    const uint32_t w[7] = {a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int j = 0; j < 7; j++) {
      v += w[j] & 0xFFFFu;
      h += probe_m<MODE, MAXS>(st, S, ns, v);
      v += w[j] >> 16;
      p = probe_m<MODE, MAXS>(st, S, ns, v);
      h += p;
    }
  } else {  // probes of sets in HBM (rare): a rolled loop keeps the code small
#pragma unroll 1
    for (uint32_t j = 0; j < 7; j++) {
      const uint32_t wv = j == 0 ? a.y : j == 1 ? a.z : j == 2 ? a.w : j == 3 ? b.x : j == 4 ? b.y : j == 5 ? b.z : b.w;
      v += wv & 0xFFFFu;
      h += probe_m<MODE, MAXS>(st, S, ns, v);
      v += wv >> 16;
      p = probe_m<MODE, MAXS>(st, S, ns, v);
      h += p;
    }
  }
  last = p;
  return h;
}

// The child list of one x of E(p) held by a lane: values vals[o, o + d), and
// its gap-coded chunks pk[ko, ko + kn) when streamed from that copy; or, when
// kn carries NAV_WORDS, its bitset words (ko: word base, kn's low bits: first
// word index), which a bitmap staging reads instead of the values.
struct XList {
  uint32_t o, d, ko, kn;
};

// A wave takes xw elements of E(p) per step: 64 on its own; inside a
// workgroup 16, so the four waves share even a small E(p) evenly, or 64 when
// E(p) holds at least XW_WIDE_MIN elements (a step's fixed costs -- the
// ballots, the LDS rank tables, the short lists' window -- then serve four
// times as many lists).
constexpr uint32_t XW_WIDE_MIN = 256;
template <bool BLOCK>
__device__ __forceinline__ uint32_t x_width(uint32_t xn) { return (BLOCK && xn < XW_WIDE_MIN) ? 16u : 64u; }

// The lane's x at position base + lane of E(p), and its child list; xl: E(p)'s
// child pointers (A.xlink at E(p)'s offset), or null; xh: their hints
// (A.xhint at E(p)'s offset) when the values below the parent value cannot be
// in S(p), or null.  A trimmed list starts at its first value not below the
// parent value (at that value's chunk when streamed gap-coded).
template <bool BLOCK, int MAXS>
__device__ __forceinline__ XList x_list(const TunedPlan &P, const TunedArgs &A, const DSet *S, const uint32_t *xsrc,
                                        uint32_t xn, uint32_t base, uint32_t xw, bool filter_s, int mi, int big_deg,
                                        int gap, const uint4 *xl, const uint32_t *xh) {
  const DTrie &tx = A.tx;
  const uint32_t lane = __lane_id(), idx = base + lane;
  bool ok = lane < xw && idx < xn;
  XList r = {0u, 0u, 0u, 0u};
  if (xl && !filter_s && A.nh == 0) {  // one 16-byte load, the value itself not needed
    if (ok) {
      const uint4 nv = xl[idx];
      uint32_t cut = xh ? xh[idx] : 0u;
      r.o = nv.x;
      r.d = nv.y;
      if (nv.w & NAV_WORDS) {  // bitset words: the values, if read, are trimmed exactly
        r.ko = nv.z;
        r.kn = nv.w;
      } else if (gap && r.d >= (uint32_t)big_deg && nv.w) {
        const uint32_t k = cut / PK_VALS;  // whole chunks below the parent value
        r.ko = nv.z + k;
        r.kn = nv.w - k;
        cut = k * PK_VALS;
      }
      r.o += cut;
      r.d -= cut;
    }
    return r;
  }
  const uint32_t x = ok ? xsrc[idx] : 0u;
  if (ok && filter_s) {
#pragma unroll
    for (int i = 0; i < MAXS; i++)
      if (i < P.ns && i != mi && ok) ok = set_contains(S[i], x);
  }
  if (ok) ok = head_ok(P, A, x);
  if (ok && tx.nav) {  // one 16-byte lookup by value
    if (x < tx.nav_len) {
      const uint4 nv = tx.nav[x];
      r.o = nv.x;
      r.d = nv.y;
      if ((nv.w & NAV_WORDS) || (gap && r.d >= (uint32_t)big_deg)) {
        r.ko = nv.z;
        r.kn = nv.w;
      }
    }
  } else if (ok) {
    const int32_t nx = node_of(tx, x);
    if (nx >= 0) {
      r.o = tx.off[nx];
      r.d = tx.off[nx + 1] - r.o;
      if (gap && r.d >= (uint32_t)big_deg && r.d >= PK_MIN) {
        r.ko = tx.pk_off[nx];
        r.kn = tx.pk_off[nx + 1] - r.ko;
      }
    }
  }
  return r;
}

// Streams child(x) for every x of E(p) through the staged S(p); MODE is the
// staging (0 bitmap, 1 hash + filter, 2 sets probed in HBM).  With `gap`, a
// long list that the trie holds gap-coded is read from that copy (about 2.1
// bytes per value instead of 4).  `first` is the caller's x_list of the wave's
// first step, issued before the staging so that its loads overlap it.
template <int MODE, bool BLOCK, int MAXS, int DEEP>
__device__ __forceinline__ uint64_t stream_children(const TunedPlan &P, const TunedArgs &A, const Stage &st,
                                                    const DSet *S, const uint32_t *xsrc, uint32_t xn, bool filter_s,
                                                    int mi, uint32_t *own, int big_deg, uint32_t slg, int gap,
                                                    const XList first, const uint4 *xl, const uint32_t *xh) {
  const DTrie &tx = A.tx;
  const uint32_t lane = __lane_id();
  const uint32_t wave = BLOCK ? (threadIdx.x >> 6) : 0u;
  const uint32_t nwaves = BLOCK ? (blockDim.x >> 6) : 1u;
  uint64_t hits = 0;
  const uint32_t XW = x_width<BLOCK>(xn);
  // several loads in flight per lane for the LDS stagings; sets probed in HBM
  // (rare) stay one quad (chunk) per step, which keeps the kernel's code small
  constexpr bool UNROLL = MODE != 2;
  for (uint32_t base = wave * XW; base < xn; base += nwaves * XW) {
    XList xi = first;
    if (base != wave * XW) xi = x_list<BLOCK, MAXS>(P, A, S, xsrc, xn, base, XW, filter_s, mi, big_deg, gap, xl, xh);
    const uint32_t o = xi.o, d = xi.d, ko = xi.ko, kn = xi.kn;
    const uint32_t G = 1u << slg, qt = lane >> slg, ql = lane & (G - 1u), lists = 64u >> slg;
    // long child sets held as bitset words, against a bitmap staging: 64 / G
    // at a time, one per group of G lanes, each lane ANDing 64-bit words of the
    // list with the staged bitmap's words over their common window (64 values
    // per word, popcount of the AND)
    const bool wl = MODE == 0 && d >= (uint32_t)big_deg && d && (kn & NAV_WORDS);
    if (MODE == 0) {
      const uint64_t wm = __ballot(wl);
      if (wm) {
        const uint32_t nwl = (uint32_t)__popcll(wm);
        if (wl) own[wave_popc_below(wm)] = lane;
        wave_sync();
        const uint32_t s_lo = st.bm_base >> 1, s_hi = s_lo + (st.bm_words >> 1) - 1u;  // in 64-bit words
        const uint2 *tab2 = reinterpret_cast<const uint2 *>(st.tab);                   // bm_base is even
        for (uint32_t r0 = 0; r0 < nwl; r0 += lists) {
          const int myl = (r0 + qt < nwl) ? (int)own[r0 + qt] : -1;
          const int src = myl < 0 ? 0 : myl;
          const uint32_t so = __shfl(o, src, 64), sd = __shfl(d, src, 64);
          const uint32_t sko = __shfl(ko, src, 64), skn = __shfl(kn, src, 64);
          uint32_t h32 = 0;
          if (myl >= 0) {
            // the list's words cover [first value >> 6, last value >> 6]
            const uint32_t lo = max(skn & ~NAV_WORDS, s_lo), hi = min(tx.vals[so + sd - 1] >> 6, s_hi);
            const uint32_t nk = hi >= lo ? hi - lo + 1u : 0u;
            const uint64_t *wp = tx.words + (uint32_t)(sko + lo);
            const uint2 *tp = tab2 + (lo - s_lo);
            uint32_t k = ql;
            for (; DEEP && k + 3 * G < nk; k += 4 * G) {  // four words in flight
              const uint64_t g0 = wp[k], g1 = wp[k + G], g2 = wp[k + 2 * G], g3 = wp[k + 3 * G];
              const uint2 t0 = tp[k], t1 = tp[k + G], t2 = tp[k + 2 * G], t3 = tp[k + 3 * G];
              h32 += __popc((uint32_t)g0 & t0.x) + __popc((uint32_t)(g0 >> 32) & t0.y) +
                     __popc((uint32_t)g1 & t1.x) + __popc((uint32_t)(g1 >> 32) & t1.y) +
                     __popc((uint32_t)g2 & t2.x) + __popc((uint32_t)(g2 >> 32) & t2.y) +
                     __popc((uint32_t)g3 & t3.x) + __popc((uint32_t)(g3 >> 32) & t3.y);
            }
            for (; k + G < nk; k += 2 * G) {  // two words in flight
              const uint64_t g0 = wp[k], g1 = wp[k + G];
              const uint2 t0 = tp[k], t1 = tp[k + G];
              h32 += __popc((uint32_t)g0 & t0.x) + __popc((uint32_t)(g0 >> 32) & t0.y) +
                     __popc((uint32_t)g1 & t1.x) + __popc((uint32_t)(g1 >> 32) & t1.y);
            }
            if (k < nk) {
              const uint64_t g0 = wp[k];
              const uint2 t0 = tp[k];
              h32 += __popc((uint32_t)g0 & t0.x) + __popc((uint32_t)(g0 >> 32) & t0.y);
            }
          }
          hits += h32;
        }
        wave_sync();  // own[] is reused below
      }
    }
    // long gap-coded child sets (single-S plans only: the 4-clique's kernel keeps
    // its registers for the materialized S(p)): 64 / G at a time, one per group
    // of G = 2^slg lanes, one chunk of PK_VALS values per lane per step, two in
    // flight (the x's of one prefix are adjacent ids, so the lengths in a round
    // are alike).  Rounds of coded lists only, so no group waits on another
    // group's other kind of list.  Every chunk is probed whole, the zero-padded
    // last one too: its padding repeats the list's last value, which is probed
    // `pad` times more than it occurs and taken off again by the lane that took
    // that chunk.
    const bool cl = MAXS == 1 && d >= (uint32_t)big_deg && !wl && kn && !(kn & NAV_WORDS);
    if (MAXS == 1) {
      const uint64_t cm = __ballot(cl);
      if (cm) {
        const uint32_t ncl = (uint32_t)__popcll(cm);
        if (cl) own[wave_popc_below(cm)] = lane;
        wave_sync();
        for (uint32_t r0 = 0; r0 < ncl; r0 += lists) {
          const int myl = (r0 + qt < ncl) ? (int)own[r0 + qt] : -1;
          const int src = myl < 0 ? 0 : myl;
          const uint32_t sd = __shfl(d, src, 64), sko = __shfl(ko, src, 64);
          const uint32_t bd = myl < 0 ? 0u : sd;
          const uint4 *q = tx.pk + 2ull * sko;
          const uint32_t nk = (bd + PK_VALS - 1) / PK_VALS, klast = nk - 1u, pad = nk * PK_VALS - bd;
          uint32_t h32 = 0, corr = 0, c = ql;
          // two chunks (30 values) in flight per step while some lane of the wave
          // has two left; a lane with one left takes it in the same step (its
          // second chunk a masked repeat), so a one-chunk step runs only when no
          // lane has two left
          for (; UNROLL; c += 2 * G) {
            const bool two = c + G < nk;
            if (!__ballot(two)) break;
            if (c < nk) {
              const uint32_t c2 = two ? c + G : c;
              const uint4 a0 = q[2 * c], b0 = q[2 * c + 1], a1 = q[2 * c2], b1 = q[2 * c2 + 1];
              uint32_t l0, l1;
              const uint32_t p0 = probe_chunk<MODE, MAXS>(st, S, P.ns, a0, b0, l0);
              const uint32_t p1 = probe_chunk<MODE, MAXS>(st, S, P.ns, a1, b1, l1);
              h32 += p0 + (two ? p1 : 0u);
              if (c == klast) corr = l0;
              if (two && c2 == klast) corr = l1;
            }
          }
          for (; c < nk; c += G) {
            uint32_t l0;
            h32 += probe_chunk<MODE, MAXS>(st, S, P.ns, q[2 * c], q[2 * c + 1], l0);
            if (c == klast) corr = l0;
          }
          hits += h32 - pad * corr;  // only the lane that took the last chunk has corr != 0
        }
        wave_sync();  // own[] is reused below
      }
    }
    // other long child sets (not gap-coded): 64 / G at a time, one per group,
    // streamed with 128-bit loads, two quads per lane in flight per step
    const uint64_t big = __ballot(d >= (uint32_t)big_deg && !wl && !cl);
    const uint32_t nbig = (uint32_t)__popcll(big);
    // own[k] = the lane holding the k-th long set (the owner window is free until below)
    if (d >= (uint32_t)big_deg && !wl && !cl) own[wave_popc_below(big)] = lane;
    wave_sync();
    for (uint32_t r0 = 0; r0 < nbig; r0 += lists) {
      const int myl = (r0 + qt < nbig) ? (int)own[r0 + qt] : -1;
      const int src = myl < 0 ? 0 : myl;
      const uint32_t so = __shfl(o, src, 64), sd = __shfl(d, src, 64);
      const uint32_t bo = so, bd = myl < 0 ? 0u : sd;
      uint32_t h32 = 0;  // this round's hits of the lane (< 2^32)
      const uint32_t head = min((4u - (bo & 3u)) & 3u, bd);
      for (uint32_t j = ql; j < head; j += G) h32 += probe_m<MODE, MAXS>(st, S, P.ns, tx.vals[bo + j]);
      const uint4 *q = reinterpret_cast<const uint4 *>(tx.vals + bo + head);  // 16-byte aligned
      const uint32_t n4 = (bd - head) >> 2;
      uint32_t i = ql;
      for (; UNROLL && DEEP && i + 3 * G < n4; i += 4 * G) {  // four quads in flight
        const uint4 v0 = q[i], v1 = q[i + G], v2 = q[i + 2 * G], v3 = q[i + 3 * G];
        h32 += probe_m<MODE, MAXS>(st, S, P.ns, v0.x) + probe_m<MODE, MAXS>(st, S, P.ns, v0.y) +
               probe_m<MODE, MAXS>(st, S, P.ns, v0.z) + probe_m<MODE, MAXS>(st, S, P.ns, v0.w) +
               probe_m<MODE, MAXS>(st, S, P.ns, v1.x) + probe_m<MODE, MAXS>(st, S, P.ns, v1.y) +
               probe_m<MODE, MAXS>(st, S, P.ns, v1.z) + probe_m<MODE, MAXS>(st, S, P.ns, v1.w) +
               probe_m<MODE, MAXS>(st, S, P.ns, v2.x) + probe_m<MODE, MAXS>(st, S, P.ns, v2.y) +
               probe_m<MODE, MAXS>(st, S, P.ns, v2.z) + probe_m<MODE, MAXS>(st, S, P.ns, v2.w) +
               probe_m<MODE, MAXS>(st, S, P.ns, v3.x) + probe_m<MODE, MAXS>(st, S, P.ns, v3.y) +
               probe_m<MODE, MAXS>(st, S, P.ns, v3.z) + probe_m<MODE, MAXS>(st, S, P.ns, v3.w);
      }
      // two quads per step while some lane has two left; a lane with one left
      // takes it in the same step (its second quad a masked repeat)
      for (; UNROLL; i += 2 * G) {
        const bool two = i + G < n4;
        if (!__ballot(two)) break;
        if (i < n4) {
          const uint4 v0 = q[i], v1 = q[two ? i + G : i];
          const uint32_t p0 = probe_m<MODE, MAXS>(st, S, P.ns, v0.x) + probe_m<MODE, MAXS>(st, S, P.ns, v0.y) +
                              probe_m<MODE, MAXS>(st, S, P.ns, v0.z) + probe_m<MODE, MAXS>(st, S, P.ns, v0.w);
          const uint32_t p1 = probe_m<MODE, MAXS>(st, S, P.ns, v1.x) + probe_m<MODE, MAXS>(st, S, P.ns, v1.y) +
                              probe_m<MODE, MAXS>(st, S, P.ns, v1.z) + probe_m<MODE, MAXS>(st, S, P.ns, v1.w);
          h32 += p0 + (two ? p1 : 0u);
        }
      }
      for (; i < n4; i += G) {
        const uint4 v0 = q[i];
        h32 += probe_m<MODE, MAXS>(st, S, P.ns, v0.x) + probe_m<MODE, MAXS>(st, S, P.ns, v0.y) +
               probe_m<MODE, MAXS>(st, S, P.ns, v0.z) + probe_m<MODE, MAXS>(st, S, P.ns, v0.w);
      }
      const uint32_t done = head + 4 * n4;
      for (uint32_t j = ql; j < bd - done; j += G) h32 += probe_m<MODE, MAXS>(st, S, P.ns, tx.vals[bo + done + j]);
      hits += h32;
    }
    wave_sync();  // the owner window below reuses own[]
    // short ones: expand owners into the LDS window, then one element per lane per step
    const uint32_t ds = (d < (uint32_t)big_deg) ? d : 0u;
    const uint32_t incl = wave_incl_scan(ds);
    const uint32_t T = __shfl(incl, 63, 64);
    const uint32_t s0 = incl - ds;
    const uint32_t rebase = o - s0;  // element e of the window lives at vals[rebase + e] (mod 2^32)
    for (uint32_t r0 = 0; r0 < T; r0 += OWN) {
      const uint32_t lo = s0 > r0 ? s0 : r0;
      const uint32_t hi = (s0 + ds < r0 + OWN) ? s0 + ds : r0 + OWN;
      for (uint32_t i = lo; i < hi; i++) own[i - r0] = rebase;
      wave_sync();
      const uint32_t lim = (T - r0 < (uint32_t)OWN) ? T - r0 : (uint32_t)OWN;
      uint32_t h32 = 0;
      for (uint32_t e = lane; e < lim; e += 64) h32 += probe_m<MODE, MAXS>(st, S, P.ns, tx.vals[own[e] + r0 + e]);
      hits += h32;
      wave_sync();
    }
  }
  return hits;
}

// One prefix, processed by a group: the whole 4-wave workgroup (BLOCK) or one
// wave.  tab/flt/sv are the group's LDS; own is the calling wave's window.
// dsc0/dsc1: the prefix's descriptor (A.desc, always present for single-S plans).
template <bool BLOCK, int MAXS, int DEEP>
__device__ __forceinline__ uint64_t tuned_prefix(const TunedPlan &P, const TunedArgs &A, uint32_t p, const uint4 dsc0,
                                 const uint4 dsc1, uint32_t *tab, uint32_t tab_slots,
                                 uint32_t *flt, uint32_t flt_words, uint32_t *sv, uint32_t sv_cap, uint32_t *own,
                                 int big_deg, int bitmap_all, uint32_t slg, int gap, int &mode_used) {
  const uint32_t lane = __lane_id();
  const uint32_t gtid = BLOCK ? threadIdx.x : lane;
  const uint32_t gsize = BLOCK ? blockDim.x : 64u;
  const uint32_t wave = BLOCK ? (threadIdx.x >> 6) : 0u;
  const uint32_t nwaves = BLOCK ? (blockDim.x >> 6) : 1u;
  auto gsync = [&]() {
    if (BLOCK) __syncthreads();
    else wave_sync();
  };
  // the launch picks MAXS = 1 or 2 only for plans with exactly that many S sets
  const int NS = (MAXS <= 2) ? MAXS : P.ns;
  DSet S[MAXS];
  DSet Dd;
  constexpr bool use_desc = MAXS == 1;  // tuned_run gives single-S plans descriptors
  if (use_desc) {
    S[0].v = A.ts[0].vals + dsc0.x;
    S[0].n = dsc0.y;
    S[0].w = A.ts[0].words + dsc1.z;
    S[0].lay = dsc1.w;
    S[0].start = dsc1.x >> 6;
    S[0].W = (dsc1.w == EHX_RANGE_BITSET) ? (dsc1.y >> 6) - (dsc1.x >> 6) + 1u : 0u;
    // E(p) (or this part of it): in D's trie for driver plans, else in S's own
    Dd.v = (P.nd ? A.td.vals : A.ts[0].vals) + dsc0.z;
    Dd.n = dsc0.w;
  } else if (!resolve_s<MAXS>(P, A, p, S) || !resolve_d(P, A, p, Dd)) {
    return 0;
  }
  int mi = 0;
#pragma unroll
  for (int i = 1; i < MAXS; i++)
    if (i < P.ns && S[i].n <= sel<MAXS>(S, mi).n) mi = i;
  const uint32_t *src = nullptr;
  uint32_t sn = 0;
  if (P.ns == 1) {
    src = S[0].v;
    sn = S[0].n;
  } else if (MAXS > 1) {
    uint32_t mn = S[0].n;
#pragma unroll
    for (int i = 1; i < MAXS; i++)
      if (i < P.ns && S[i].n < mn) mn = S[i].n;
    if (mn <= sv_cap) {
      // materialize S(p) = S_1 ∩ .. ∩ S_s (in a workgroup, by its first wave)
      int alg;
      if (BLOCK) {
        __shared__ uint32_t s_sn;
        if (wave == 0) {
          uint32_t c = wave_intersect<true>(S, NS, sv, &alg);
          if (lane == 0) s_sn = c;
        }
        __syncthreads();
        sn = s_sn;
        __syncthreads();
      } else {
        sn = wave_intersect<true>(S, NS, sv, &alg);
        wave_sync();
      }
      if (sn == 0) return 0;
      src = sv;
    }
  }
  // one S set: E(p) is known now, so the first step's child lookups are issued
  // ahead of the staging
  XList first = {0u, 0u, 0u, 0u};
  const bool x_in_d = P.nd || use_desc;  // E(p) is Dd
  const uint4 *xl = (use_desc && A.xlink) ? A.xlink + dsc0.z : nullptr;
  const uint32_t *xh = nullptr;
  if (xl && A.xhint) {
    // E(p)'s parent value y (its set is child(y)); child values below y are
    // below S(p)'s first value when that is at least y
    const uint32_t y = __builtin_amdgcn_readfirstlane((P.nd ? A.dp : A.sp[0])[p]);
    if (dsc1.x >= y) xh = A.xhint + dsc0.z;
  }
  if (MAXS == 1) {
    const uint32_t *xs = x_in_d ? Dd.v : src;
    const uint32_t xn1 = x_in_d ? Dd.n : sn;
    const uint32_t xw = x_width<BLOCK>(xn1);
    first = x_list<BLOCK, MAXS>(P, A, S, xs, xn1, wave * xw, xw, false, 0, big_deg, gap, xl, xh);
  }
  // staging decision (group-uniform)
  Stage st;
  st.mode = 2;
  mode_used = 2;
  st.tab = tab;
  st.flt = flt;
  st.bm_base = st.bm_words = st.hshift = st.hmask = st.fshift = 0;
  if (src) {
    const uint32_t first = use_desc ? dsc1.x : src[0], last = use_desc ? dsc1.y : src[sn - 1];
    const uint32_t w32 = bitmap_words(first, last);
    const bool is_bitset = (P.ns == 1) ? (S[0].lay == EHX_RANGE_BITSET)
                                       : (sn > 2 && (uint64_t)(last - first) <= 128ull * sn);
    const bool bm_fits = w32 < tab_slots, hash_fits = sn <= tab_slots / 2 && last != EMPTY;  // bitmap + zero word
    if (A.force_mode == 0) st.mode = bm_fits ? 0 : (hash_fits ? 1 : 2);
    else if (A.force_mode == 1) st.mode = hash_fits ? 1 : (bm_fits ? 0 : 2);
    else if (bm_fits && (is_bitset || bitmap_all)) st.mode = 0;
    else if (hash_fits) st.mode = 1;
    else if (bm_fits) st.mode = 0;  // a bitmap still beats probing the sets in HBM
    if (st.mode == 0) {
      st.bm_base = (first >> 6) << 1;
      st.bm_words = w32;
      for (uint32_t i = gtid; i <= w32; i += gsize) tab[i] = 0;  // tab[w32]: the zero word
      gsync();
      if (P.ns == 1 && S[0].lay == EHX_RANGE_BITSET) {
        // copy range_bitset words; 32-bit word i of the bitmap covers values [(bm_base + i) * 32, +32)
        for (uint32_t k = gtid; k < S[0].W; k += gsize) {
          const uint64_t wv = S[0].w[k];
          const uint32_t i0 = (S[0].start + k) * 2u - st.bm_base;  // the base is S's first 64-bit word
          if (i0 < w32) tab[i0] = (uint32_t)wv;
          if (i0 + 1u < w32) tab[i0 + 1u] = (uint32_t)(wv >> 32);
        }
      } else {
        for (uint32_t i = gtid; i < sn; i += gsize) {
          const uint32_t x = src[i];
          atomicOr(&tab[(x >> 5) - st.bm_base], 1u << (x & 31));
        }
      }
    } else if (st.mode == 1) {
      uint32_t H = 64, lg = 6;
      while (H < 2 * sn) { H <<= 1; lg++; }
      st.hshift = 32 - lg;
      st.hmask = H - 1;
      // the whole filter area, even for small sets: a false positive sends its
      // lane into the table loop and stalls all 64 lanes, so the filter is
      // kept as sparse as the LDS allows
      const uint32_t fw = flt_words;
      uint32_t lgf = 5;
      while ((1u << lgf) < fw * 32) lgf++;
      st.fshift = 32 - lgf;
      for (uint32_t i = gtid; i < H; i += gsize) tab[i] = EMPTY;
      for (uint32_t i = gtid; i < fw; i += gsize) flt[i] = 0;
      gsync();
      for (uint32_t i = gtid; i < sn; i += gsize) {
        const uint32_t x = src[i];
        uint32_t h = hash_t(x, st.hshift);
        while (atomicCAS(&tab[h], EMPTY, x) != EMPTY) h = (h + 1) & st.hmask;
        const uint32_t f = hash_f(x, st.fshift);
        atomicOr(&flt[f >> 5], 1u << (f & 31));
      }
    }
  }
  gsync();
  mode_used = st.mode;

  const DSet Smi = sel<MAXS>(S, mi);
  const uint32_t *xsrc = x_in_d ? Dd.v : (src ? src : Smi.v);
  const uint32_t xn = x_in_d ? Dd.n : (src ? sn : Smi.n);
  const bool filter_s = !P.nd && !src;  // x runs over one S set: test it against the others
  if (MAXS > 1) {
    const uint32_t xw = x_width<BLOCK>(xn);
    first = x_list<BLOCK, MAXS>(P, A, S, xsrc, xn, wave * xw, xw, filter_s, mi, big_deg, gap, nullptr, nullptr);
  }
  uint64_t hits;
  // one streaming loop per staging mode: no per-probe branch on the mode
  if (st.mode == 0)
    hits = stream_children<0, BLOCK, MAXS, DEEP>(P, A, st, S, xsrc, xn, filter_s, mi, own, big_deg, slg, gap,
                                                       first, xl, xh);
  else if (st.mode == 1)
    hits = stream_children<1, BLOCK, MAXS, DEEP>(P, A, st, S, xsrc, xn, filter_s, mi, own, big_deg, slg, gap,
                                                       first, xl, xh);
  else
    hits = stream_children<2, BLOCK, MAXS, DEEP>(P, A, st, S, xsrc, xn, filter_s, mi, own, big_deg, slg, gap,
                                                       first, xl, xh);
  gsync();
  return hits;
}

// ---- tiny prefixes, many per wavefront ---------------------------------------
// Most prefixes of a skewed graph carry almost no work (on R-MAT-22, middle
// vertex first, 85 % of them hold 1 % of the work) but took a full round of
// staging and child lookups each in k_tuned.  Here a wavefront takes TINY_BATCH
// tiny-bin entries at once: each prefix's S(p) becomes its own small
// open-addressing table in the wave's LDS, and the x's of all of them, then the
// child lists of those x's, are flattened and streamed TINY_W elements per step
// (four consecutive elements per lane).  An element's owner (its prefix, or its
// x) comes from an LDS owner map: each run marks its first element, and a
// max-scan carries the marks forward, so no step waits on a chain of searches.
__device__ __forceinline__ uint32_t tiny_hash(uint32_t c) { return (c * 0x9E3779B1u) >> 20; }

constexpr uint32_t TINY_W = 256;  // flattened elements per step: 4 per lane

// Owners of elements [r0, r0 + TINY_W) of consecutive runs (run ids ascend with
// their starts; each lane holds RPL runs {start, length, id}): out[u] = id of
// the run holding element r0 + 4 * lane + u (past the last run: the last id).
// own is TINY_W words of the wave's LDS.
template <int RPL>
__device__ __forceinline__ void tiny_owners(uint32_t *own, uint32_t r0, const uint32_t (&st)[RPL],
                                            const uint32_t (&ln)[RPL], const uint32_t (&id)[RPL],
                                            uint32_t (&out)[4]) {
  const uint32_t lane = __lane_id();
  reinterpret_cast<uint4 *>(own)[lane] = make_uint4(0, 0, 0, 0);  // 0: no run starts here
  wave_sync();
#pragma unroll
  for (int u = 0; u < RPL; u++) {
    if (!ln[u]) continue;
    if (st[u] >= r0 && st[u] - r0 < TINY_W) own[st[u] - r0] = id[u] + 1;
    else if (st[u] < r0 && st[u] + ln[u] > r0) own[0] = id[u] + 1;  // the run that crosses r0
  }
  wave_sync();
  uint4 v = reinterpret_cast<uint4 *>(own)[lane];
  wave_sync();  // own is free again
  v.y = max(v.y, v.x);
  v.z = max(v.z, v.y);
  v.w = max(v.w, v.z);
  uint32_t incl = v.w;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl = max(incl, y);
  }
  uint32_t carry = __shfl_up(incl, 1, 64);
  if (lane == 0) carry = 0;
  out[0] = max(v.x, carry) - 1u;
  out[1] = max(v.y, carry) - 1u;
  out[2] = max(v.z, carry) - 1u;
  out[3] = max(v.w, carry) - 1u;
}

// x's child list {off, length}: from E(p)'s child pointers when there are any
// (past the values below E(p)'s parent value when `trim`, as in x_list), else
// by value.
__device__ __forceinline__ uint2 tiny_child(const TunedPlan &P, const TunedArgs &A, const uint32_t *xvals,
                                            uint32_t pos, bool trim) {
  if (A.xlink && A.nh == 0) {
    const uint4 nv = A.xlink[pos];
    const uint32_t cut = trim ? A.xhint[pos] : 0u;
    return make_uint2(nv.x + cut, nv.y - cut);
  }
  const uint32_t x = xvals[pos];
  if (!head_ok(P, A, x)) return make_uint2(0, 0);
  const DTrie &tx = A.tx;
  if (tx.nav) {
    if (x >= tx.nav_len) return make_uint2(0, 0);
    const uint4 nv = tx.nav[x];
    return make_uint2(nv.x, nv.y);
  }
  const int32_t nx = node_of(tx, x);
  if (nx < 0) return make_uint2(0, 0);
  const uint32_t o = tx.off[nx];
  return make_uint2(o, tx.off[nx + 1] - o);
}

template <int SLOTS>
__global__ __launch_bounds__(256) void k_tiny(TunedPlan P, TunedArgs A, const uint32_t *__restrict__ queue,
                                              unsigned *__restrict__ ctl, unsigned long long *__restrict__ acc) {
  constexpr uint32_t WS = SLOTS / 4;  // table words per wave (>= 512: a 256-slot table always fits)
  __shared__ uint32_t tab_all[4][WS];
  __shared__ __attribute__((aligned(16))) uint32_t own_all[4][TINY_W];  // read as uint4
  // per prefix of the batch: S offset, first staged element, table base, table
  // mask, E offset, first x, hits; per x of a step: list offset, first element, prefix
  __shared__ uint32_t meta_all[4][7][64];
  __shared__ uint32_t run_all[4][3][TINY_W];
  const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
  uint32_t *tab = tab_all[w], *own = own_all[w];
  uint32_t *mSo = meta_all[w][0], *mSs = meta_all[w][1], *mTb = meta_all[w][2], *mMk = meta_all[w][3];
  uint32_t *mEo = meta_all[w][4], *mXs = meta_all[w][5], *mHit = meta_all[w][6];
  uint32_t *cO = run_all[w][0], *cS = run_all[w][1], *cK = run_all[w][2];
  const uint32_t qt = ctl[133], qlen = ctl[128];
  const bool segs = A.seg.out != nullptr;
  const uint32_t *svals = A.ts[0].vals, *xvals = P.nd ? A.td.vals : A.ts[0].vals, *cvals = A.tx.vals;
  uint64_t hits = 0;
  // batches are alike, so they are dealt out by wave number (one shared atomic
  // cursor, hit by every wave at once when the kernel starts, cost more than the
  // tiny bin's work)
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t bi = blockIdx.x * (blockDim.x >> 6) + w;; bi += nwaves) {
    const uint64_t q064 = (uint64_t)qt + (uint64_t)bi * TINY_BATCH;
    if (q064 >= qlen) break;
    const uint32_t q0 = (uint32_t)q064;
    const uint32_t nb = min(TINY_BATCH, qlen - q0);
    uint4 d0 = make_uint4(0, 0, 0, 0);
    uint32_t p = 0, trim = 0;
    if (lane < nb) {
      d0 = A.desc[2ull * (q0 + lane)];  // {S offset, |S|, E offset, |E|}
      p = queue[q0 + lane];
      // child lists skip their values below E(p)'s parent value when S(p)'s
      // first value is not below it
      if (A.xhint && A.xlink && A.nh == 0 && d0.y)
        trim = A.desc[2ull * (q0 + lane) + 1].x >= (P.nd ? A.dp : A.sp[0])[p] ? 1u : 0u;
    }
    uint32_t sz = 0;  // this prefix's table: a power of two >= 4/3 |S| (load <= 3/4)
    if (lane < nb) {
      sz = 4;
      while (3 * sz < 4 * d0.y) sz <<= 1;
    }
    for (uint32_t start = 0; start < nb;) {
      // the batch's prefixes [start, end) whose tables fit the wave's LDS together
      const uint32_t mine = (lane >= start) ? sz : 0u;
      const uint32_t tincl = wave_incl_scan(mine);
      const bool in = lane >= start && lane < nb && tincl <= WS;
      const uint32_t end = start + (uint32_t)__popcll(__ballot(in));
      const uint32_t tsum = __shfl(tincl, (int)end - 1, 64);
      const uint32_t sn = in ? d0.y : 0u, xn = in ? d0.w : 0u;
      const uint32_t sincl = wave_incl_scan(sn), xincl = wave_incl_scan(xn);
      const uint32_t stot = __shfl(sincl, 63, 64), xtot = __shfl(xincl, 63, 64);
      mSo[lane] = d0.x;
      mSs[lane] = sincl - sn;
      mTb[lane] = tincl - mine;
      mMk[lane] = (sz - 1) | (trim << 31);  // bit 31: the prefix's child lists are trimmed
      mEo[lane] = d0.z;
      mXs[lane] = xincl - xn;
      mHit[lane] = 0;
      for (uint32_t i = lane; i < tsum; i += 64) tab[i] = EMPTY;
      wave_sync();
      // stage every S(p) of the sub-batch in its own table
      const uint32_t sst[1] = {sincl - sn}, sln[1] = {sn}, sid[1] = {lane};
      for (uint32_t r0 = 0; r0 < stot; r0 += TINY_W) {
        uint32_t k[4], c[4];
        tiny_owners<1>(own, r0, sst, sln, sid, k);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const uint32_t e = r0 + 4 * lane + u;
          if (e < stot) c[u] = svals[mSo[k[u]] + e - mSs[k[u]]];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const uint32_t e = r0 + 4 * lane + u;
          if (e >= stot) continue;
          const uint32_t base = mTb[k[u]], mask = mMk[k[u]] & 0x7FFFFFFFu;
          uint32_t h = tiny_hash(c[u]) & mask;
          while (atomicCAS(&tab[base + h], EMPTY, c[u]) != EMPTY) h = (h + 1) & mask;
        }
      }
      wave_sync();
      // the sub-batch's x's, TINY_W per step, then their child lists' elements
      uint32_t h32 = 0;
      const uint32_t xst[1] = {xincl - xn}, xln[1] = {xn}, xid[1] = {lane};
      for (uint32_t x0 = 0; x0 < xtot; x0 += TINY_W) {
        uint32_t k[4], o[4], dl[4];
        tiny_owners<1>(own, x0, xst, xln, xid, k);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const uint32_t t = x0 + 4 * lane + u;
          o[u] = dl[u] = 0;
          if (t < xtot) {
            const uint2 cl = tiny_child(P, A, xvals, mEo[k[u]] + t - mXs[k[u]], (mMk[k[u]] >> 31) != 0u);
            o[u] = cl.x;
            dl[u] = cl.y;
          }
        }
        // runs of the step's x's: run 4 * lane + u, in x order
        const uint32_t l4 = dl[0] + dl[1] + dl[2] + dl[3];
        const uint32_t lincl = wave_incl_scan(l4);
        const uint32_t ctot = __shfl(lincl, 63, 64);
        uint32_t cst[4], cid[4];
        cst[0] = lincl - l4;
        cst[1] = cst[0] + dl[0];
        cst[2] = cst[1] + dl[1];
        cst[3] = cst[2] + dl[2];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          cid[u] = 4 * lane + u;
          cO[cid[u]] = o[u];
          cS[cid[u]] = cst[u];
          cK[cid[u]] = k[u];
        }
        wave_sync();
        for (uint32_t e0 = 0; e0 < ctot; e0 += TINY_W) {
          uint32_t r[4], c[4];
          tiny_owners<4>(own, e0, cst, dl, cid, r);
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const uint32_t e = e0 + 4 * lane + u;
            if (e < ctot) c[u] = cvals[cO[r[u]] + e - cS[r[u]]];
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const uint32_t e = e0 + 4 * lane + u;
            if (e >= ctot) continue;
            const uint32_t kk = cK[r[u]], base = mTb[kk], mask = mMk[kk] & 0x7FFFFFFFu;
            uint32_t h = tiny_hash(c[u]) & mask, hit = 0;
            for (;;) {  // EMPTY first: 0xFFFFFFFF is never staged
              const uint32_t v = tab[base + h];
              if (v == EMPTY) break;
              if (v == c[u]) { hit = 1; break; }
              h = (h + 1) & mask;
            }
            h32 += hit;
            if (segs && hit) atomicAdd(&mHit[kk], 1u);
          }
        }
        wave_sync();  // cO, cS, cK are the next step's
      }
      hits += h32;
      if (segs && in && mHit[lane]) seg_add(A.seg, A.prefix[0][p], mHit[lane]);
      wave_sync();  // the tables and meta are the next sub-batch's
      start = end;
    }
  }
  hits = wave_sum(hits);
  if (lane == 0) atomicAdd(acc, (unsigned long long)hits);
}

// A wave-uniform uint4 moved to scalar registers.
__device__ __forceinline__ uint4 uniform4(const uint4 v) {
  return make_uint4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                    __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
}

// Single-S plans ask for 6 wavefronts per SIMD: 80 VGPRs, with some spilled to
// scratch, which still beats 4 or 5 wavefronts with every value in registers (the
// level is bound by memory latency, and more resident waves hide more of it).
// Two-S plans (4-clique) ask for 5 (96 VGPRs): left free, the allocator takes
// 129 and three waves, 40 % slower.
template <int SLOTS, int MAXS, int DEEP>
__global__ __launch_bounds__(256, MAXS == 1 ? 6 : (MAXS == 2 ? 5 : 2)) void k_tuned(TunedPlan P, TunedArgs A, const uint32_t *__restrict__ queue,
                                               unsigned *__restrict__ ctl, unsigned long long *__restrict__ acc,
                                               int big_deg, int grab, int bitmap_all, uint32_t slg, int gap) {
  constexpr uint32_t WS = SLOTS / 4;                 // per-wave table slots
  constexpr uint32_t FWB = SLOTS / 4, FWW = FWB / 4;  // filter words per block / wave
  constexpr uint32_t VC = (MAXS > 1) ? WS : 1u;      // materialized S capacity per wave
  __shared__ __align__(16) uint32_t tab[SLOTS];  // 64-bit aligned bitmap words
  __shared__ uint32_t flt[FWB];
  __shared__ uint32_t own[4][OWN];
  __shared__ uint32_t sv[4 * VC];
  __shared__ uint32_t s_item;
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t qlen = ctl[133], nbig = ctl[131];  // the tiny bin after qlen is k_tiny's
  uint64_t hits = 0;
  // phase 1: block bin, one prefix per workgroup.  Each workgroup's first item
  // and each wave's first grab are dealt out by number: a shared atomic cursor
  // hit by every workgroup at once serializes the kernel's start.
  for (bool first = true;; first = false) {
    if (threadIdx.x == 0) s_item = first ? blockIdx.x : gridDim.x + atomicAdd(&ctl[129], 1u);
    __syncthreads();
    const uint32_t qi = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (qi >= nbig) break;
    const uint64_t t0 = A.rec ? __builtin_amdgcn_s_memtime() : 0;
    int mode = -1;
    uint4 d0 = make_uint4(0, 0, 0, 0), d1 = d0;
    if (MAXS == 1) {
      d0 = uniform4(A.desc[2ull * qi]);
      d1 = uniform4(A.desc[2ull * qi + 1]);
    }
    const uint64_t h = tuned_prefix<true, MAXS, DEEP>(P, A, queue[qi], d0, d1, tab, SLOTS, flt, FWB, sv, 4 * VC, own[w],
                                                      big_deg, bitmap_all, slg, gap, mode);
    hits += h;
    if (A.seg.out) {
      const uint64_t hw = wave_sum(h);
      if (__lane_id() == 0) seg_add(A.seg, A.prefix[0][queue[qi]], hw);
    }
    if (A.rec && threadIdx.x == 0 && mode >= 0)
      A.rec[queue[qi]] = (__builtin_amdgcn_s_memtime() - t0) | ((unsigned long long)mode << 62);
  }
  // phase 2: wave bin, each wave on its own
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6), wid = blockIdx.x * (blockDim.x >> 6) + w;
  for (bool first = true;; first = false) {
    uint64_t q064 = nbig;
    if (first) {
      q064 += (uint64_t)wid * (uint32_t)grab;
    } else {
      uint32_t q0 = 0;
      if (__lane_id() == 0) q0 = atomicAdd(&ctl[130], (unsigned)grab);
      q064 += (uint64_t)nwaves * (uint32_t)grab + __builtin_amdgcn_readfirstlane(q0);  // lane 0's: all lanes are active
    }
    if (q064 >= qlen) break;
    const uint32_t q0 = (uint32_t)q064;
    const uint32_t q1 = min(q0 + (uint32_t)grab, qlen);
    // the next entry's descriptor is loaded while this one runs
    uint4 n0 = make_uint4(0, 0, 0, 0), n1 = n0;
    if (MAXS == 1) {
      n0 = A.desc[2ull * q0];
      n1 = A.desc[2ull * q0 + 1];
    }
    for (uint32_t qi = q0; qi < q1; qi++) {
      const uint4 d0 = uniform4(n0), d1 = uniform4(n1);
      if (MAXS == 1 && qi + 1 < q1) {
        n0 = A.desc[2ull * (qi + 1)];
        n1 = A.desc[2ull * (qi + 1) + 1];
      }
      const uint64_t t0 = A.rec ? __builtin_amdgcn_s_memtime() : 0;
      int mode = -1;
      const uint64_t h = tuned_prefix<false, MAXS, DEEP>(P, A, queue[qi], d0, d1, tab + w * WS, WS, flt + w * FWW, FWW,
                                                         sv + w * VC, VC, own[w], big_deg, bitmap_all, slg, gap, mode);
      hits += h;
      if (A.seg.out) {
        const uint64_t hw = wave_sum(h);
        if (__lane_id() == 0) seg_add(A.seg, A.prefix[0][queue[qi]], hw);
      }
      if (A.rec && __lane_id() == 0 && mode >= 0)
        A.rec[queue[qi]] = (__builtin_amdgcn_s_memtime() - t0) | ((unsigned long long)mode << 62);
    }
  }
  hits = wave_sum(hits);
  if (__lane_id() == 0) atomicAdd(acc, (unsigned long long)hits);
}

template <int SLOTS>
static void launch_tuned(hipStream_t st, unsigned grid, const TunedPlan &P, const TunedArgs &A, const uint32_t *queue,
                         unsigned *ctl, unsigned long long *acc, const TuneParams &tp) {
  const uint32_t slg = 31u - (uint32_t)__builtin_clz((unsigned)tp.stream_lanes);  // a power of two in [2, 64]
  const int gap = (tp.gap_coded && A.tx.pk) ? 1 : 0;
  // four quads in flight per lane pay off for narrow lane groups (many short lists); wider
  // groups keep two, and the registers they would cost
#define EHX_TUNED_ARGS P, A, queue, ctl, acc, tp.big_deg, tp.grab, tp.bitmap_all, slg, gap
  if (P.ns == 1 && slg <= 2)
    k_tuned<SLOTS, 1, 1><<<grid, 256, 0, st>>>(EHX_TUNED_ARGS);
  else if (P.ns == 1)
    k_tuned<SLOTS, 1, 0><<<grid, 256, 0, st>>>(EHX_TUNED_ARGS);
  else if (P.ns == 2)
    k_tuned<SLOTS, 2, 0><<<grid, 256, 0, st>>>(EHX_TUNED_ARGS);
  else
    k_tuned<SLOTS, TUNED_MAX_S, 0><<<grid, 256, 0, st>>>(EHX_TUNED_ARGS);
#undef EHX_TUNED_ARGS
}

// k_tiny deals its batches out by wave number, so its grid is what stays resident.
template <int SLOTS>
static void launch_tiny(ehx_ctx *ctx, const TunedPlan &P, const TunedArgs &A, const uint32_t *queue, unsigned *ctl,
                        unsigned long long *acc) {
  static int per_cu[3][64] = {};  // by SLOTS (2048, 4096, 8192) and device
  const int dev = ctx->device & 63, si = SLOTS == 2048 ? 0 : (SLOTS == 4096 ? 1 : 2);
  if (!per_cu[si][dev]) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tiny<SLOTS>, 256, 0) != hipSuccess || nb < 1) nb = 1;
    per_cu[si][dev] = nb;
  }
  k_tiny<SLOTS><<<(unsigned)(ctx->num_cus * per_cu[si][dev]), 256, 0, ctx->stream>>>(P, A, queue, ctl, acc);
}

// One cached graph per process: the key is every value baked into the launches.
struct GraphCache {
  bool valid = false;
  int device = -1;
  unsigned char key[sizeof(TunedPlan) + sizeof(TunedArgs) + sizeof(TuneParams) + 64];
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
};
static GraphCache g_cache;

// events: time the persistent kernel (eager launches); a captured level is timed
// around its graph launch instead, since event records are not graph nodes here
static int enqueue_level(ehx_ctx *ctx, const TunedPlan &P, const TunedArgs &A, unsigned long long *acc,
                         uint32_t *bucket, unsigned *ctl, uint32_t *queue, const TuneParams &tp, bool events = true) {
  hipStream_t st = ctx->stream;
  int rc;
  if ((rc = fill_u32(ctx, ctl, 0u, 134))) return rc;
  // descriptors by prefix after the queue-ordered ones (A.desc)
  uint4 *pdesc = A.desc ? const_cast<uint4 *>(A.desc) + 2 * (A.nprefix + SPLIT_EXTRA) : nullptr;
  if (P.ns == 1)
    k_classify<1><<<grid_for(A.nprefix, 256, 16384), 256, 0, st>>>(P, A, bucket, ctl, tp.lds_slots, tp.big_work, pdesc,
                                                                    (uint32_t)tp.split_work, (uint32_t)tp.tiny_work);
  else
    k_classify<TUNED_MAX_S><<<grid_for(A.nprefix, 256, 16384), 256, 0, st>>>(
        P, A, bucket, ctl, tp.lds_slots, tp.big_work, pdesc, (uint32_t)tp.split_work, (uint32_t)tp.tiny_work);
  EHX_HIP(hipGetLastError());
  k_bucket_scan<<<1, 64, 0, st>>>(ctl);
  EHX_HIP(hipGetLastError());
  k_enqueue<<<grid_for(A.nprefix, 256, 8192), 256, 0, st>>>(bucket, A.nprefix, A.nprefix_dev, ctl, queue, pdesc,
                                                            const_cast<uint4 *>(A.desc));
  EHX_HIP(hipGetLastError());
  const unsigned per_cu = (unsigned)(tp.waves_per_cu > 4 ? tp.waves_per_cu / 4 : 1);
  const unsigned grid = (unsigned)ctx->num_cus * per_cu;
  if (events) EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K0], st));
  switch (tp.lds_slots) {
    case 2048: launch_tuned<2048>(st, grid, P, A, queue, ctl, acc, tp); break;
    case 8192: launch_tuned<8192>(st, grid, P, A, queue, ctl, acc, tp); break;
    default: launch_tuned<4096>(st, grid, P, A, queue, ctl, acc, tp); break;
  }
  EHX_HIP(hipGetLastError());
  if (events) EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K1], st));
  if (A.desc && tp.tiny_work > 0) {  // the tiny bin (empty unless single-S and tiny_work > 0)
    // (on a second stream it displaced the persistent kernel's workgroups: slower)
    switch (tp.lds_slots) {
      case 2048: launch_tiny<2048>(ctx, P, A, queue, ctl, acc); break;
      case 8192: launch_tiny<8192>(ctx, P, A, queue, ctl, acc); break;
      default: launch_tiny<4096>(ctx, P, A, queue, ctl, acc); break;
    }
    EHX_HIP(hipGetLastError());
  }
  return EHX_OK;
}

int tuned_reserve(ehx_ctx *ctx, const TunedPlan &P, uint64_t nprefix) {
  DevBuf &bb = ctx->scratch[S_QUEUE], &cb = ctx->scratch[S_MISC], &db = ctx->scratch[S_DESC];
  int rc;
  // bucket[nprefix], then the queue (split parts add up to SPLIT_EXTRA entries)
  if ((rc = bb.ensure((nprefix * 2 + SPLIT_EXTRA) * 4 + 64)) || (rc = cb.ensure(256 * 4))) return rc;
  // descriptors: queue order, then by prefix (32 bytes each)
  if (P.ns == 1 && (rc = db.ensure((nprefix * 2 + SPLIT_EXTRA) * 32))) return rc;
  return EHX_OK;
}

int tuned_run(ehx_ctx *ctx, const TunedPlan &P, const TunedArgs &A_in, int form, unsigned long long *acc,
              bool allow_graph) {
  (void)form;
  if (A_in.nprefix == 0) return EHX_OK;
  if (A_in.nprefix >= (1ull << 32)) return fail(EHX_ERR_UNSUPPORTED, "tuned level: more than 2^32 prefixes");
  const TuneParams tp = g_params;
  int rc;
  if ((rc = tuned_reserve(ctx, P, A_in.nprefix))) return rc;
  DevBuf &bb = ctx->scratch[S_QUEUE], &cb = ctx->scratch[S_MISC], &db = ctx->scratch[S_DESC];
  TunedArgs A = A_in;
  A.desc = P.ns == 1 ? db.as<uint4>() : nullptr;
  uint32_t *bucket = bb.as<uint32_t>(), *queue = bucket + A.nprefix;
  unsigned *ctl = cb.as<unsigned>();
  if (!tp.use_graph || !allow_graph) return enqueue_level(ctx, P, A, acc, bucket, ctl, queue, tp, allow_graph);

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
    rc = enqueue_level(ctx, P, A, acc, bucket, ctl, queue, tp, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc || e != hipSuccess) {
      // capture refused: run this level eagerly from now on
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
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K0], ctx->stream));  // the whole captured level
  EHX_HIP(hipGraphLaunch(g_cache.exec, ctx->stream));
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K1], ctx->stream));
  return EHX_OK;
}

void tuned_forget_graphs() {
  if (g_cache.exec) (void)hipGraphExecDestroy(g_cache.exec);
  if (g_cache.graph) (void)hipGraphDestroy(g_cache.graph);
  g_cache = GraphCache();
}

}  // namespace ehx
