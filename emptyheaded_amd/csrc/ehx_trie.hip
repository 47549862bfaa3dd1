// ehx_trie.hip -- device trie construction for one binary relation.
//
// Restates Trie<A,M>::Trie(path, max_set_sizes, attr_in, annotation)
// (src/Trie.cpp:380-510) for HBM:
//   1. pack each row as (c0 << b1) | c1 and radix sort (SortColumns order,
//      Trie.cpp:173-186 / tbb::parallel_sort :402-404);
//   2. drop repeated rows (encode_tail, :291-302) with flag + exclusive scan;
//   3. split into head values and child offsets (produce_ranges, :192-224) with
//      a second flag + scan;
//   4. per-set layout (hybrid::get_type, hybrid.hpp:103-111, or a forced or
//      learned policy), word counts, scan -> bitset pool offsets;
//   5. fill range_bitset words and per-word ranks (range_bitset::build,
//      range_bitset.hpp:123-157);
//   6. dense value -> node array when the head is dense (TrieBlock::nextSize /
//      get_next_block, TrieBlock.hpp:26-35,86-111).
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <mutex>

#include "ehx_internal.h"
#include "ehx_selector.h"

namespace ehx {

// keys[i] = (c0[i] << b1) | c1[i]; order[0] = 1 when some row's key exceeds
// the next row's, order[1] = 1 when some row's c1 exceeds the next row's.
__global__ void k_pack(const uint32_t *__restrict__ c0, const uint32_t *__restrict__ c1, uint64_t n, int b1,
                       uint64_t *__restrict__ keys, unsigned *__restrict__ order) {
  bool down = false, down1 = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = ((uint64_t)c0[i] << b1) | (uint64_t)c1[i];
    keys[i] = k;
    if (i + 1 < n) {
      const uint32_t x1 = c1[i + 1];
      down |= k > (((uint64_t)c0[i + 1] << b1) | (uint64_t)x1);
      down1 |= c1[i] > x1;
    }
  }
  // a flag already set is not set again: unordered rows would otherwise send
  // one atomic per wavefront to the same word
  const volatile unsigned *seen = order;
  if (__any(down) && __lane_id() == 0 && !seen[0]) atomicOr(&order[0], 1u);
  if (__any(down1) && __lane_id() == 0 && !seen[1]) atomicOr(&order[1], 1u);
}

// flag[i] = 1 when keys[i] starts a new run of (keys >> shift).
__global__ void k_run_flags(const uint64_t *__restrict__ keys, uint64_t n, int shift, uint32_t *__restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    flags[i] = (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? 1u : 0u;
}

__global__ void k_compact_keys(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ flags,
                               const uint64_t *__restrict__ pos, uint64_t n, uint64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (flags[i]) out[pos[i]] = keys[i];
}

__global__ void k_split(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ hflags,
                        const uint64_t *__restrict__ hpos, uint64_t m, int b1, uint32_t *__restrict__ head,
                        uint32_t *__restrict__ off, uint32_t *__restrict__ vals, uint64_t nh) {
  const uint64_t mask = (b1 >= 64) ? ~0ull : ((1ull << b1) - 1);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    vals[i] = (uint32_t)(k & mask);
    if (hflags[i]) {
      uint64_t h = hpos[i];
      head[h] = (uint32_t)(k >> b1);
      off[h] = (uint32_t)i;
    }
    if (i == 0) off[nh] = (uint32_t)m;
  }
}

// Layout and bitset word count per set; set nh is the head set.
// With the learned policy the density rule stays in `layout` and the
// selector's choice goes to `stage` (the staging layout of the tuned level);
// a set gets bitset words when either says range_bitset.
__global__ void k_layouts(DTrie t, int policy, SelectorWeights sel, uint8_t *__restrict__ layout,
                          uint8_t *__restrict__ stage, uint32_t *__restrict__ bs_start,
                          uint32_t *__restrict__ wcount) {
  const uint32_t nh = t.nh;
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h <= nh; h += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t *v;
    uint32_t n;
    if (h == nh) { v = t.head; n = nh; }
    else { v = t.vals + t.off[h]; n = t.off[h + 1] - t.off[h]; }
    uint32_t first = n ? v[0] : 0, last = n ? v[n - 1] : 0;
    uint64_t span = (n > 1) ? (uint64_t)(last - first) : 0;
    int lay, sl;
    if (n == 0) lay = EHX_UINTEGER;
    else if (policy == EHX_POLICY_ALL_UINT) lay = EHX_UINTEGER;
    else if (policy == EHX_POLICY_ALL_BITSET) lay = EHX_RANGE_BITSET;
    else {
      // common::is_sparse (common.hpp:56-67): sparse iff len <= 2 or span/len > 128;
      // for 32-bit values the double test equals span > 128*len.
      lay = (n <= 2 || span > 128ull * n) ? EHX_UINTEGER : EHX_RANGE_BITSET;
    }
    sl = lay;
    if (stage && n && h < nh) sl = selector_layout(sel, n, span, first, last, set_work(t, v, n), lay);
    layout[h] = (uint8_t)lay;
    if (stage) stage[h] = (uint8_t)sl;
    bs_start[h] = first >> 6;
    wcount[h] = ((lay == EHX_RANGE_BITSET || sl == EHX_RANGE_BITSET) && n) ? ((last >> 6) - (first >> 6) + 1) : 0;
  }
}

// One lane per 64-bit word of the whole pool (the head set's window can be tens
// of thousands of words, so sets are not given a wave each): the word's set is
// the last h with bs_off[h] <= j; its elements are v[lo, hi) with
// lo = lower_bound(64 * word); rank[word] = lo.  No atomics.
__global__ __launch_bounds__(256) void k_fill_words(DTrie t, uint64_t nwords, uint64_t *__restrict__ words,
                                                    uint32_t *__restrict__ rank) {
  const uint32_t nh = t.nh;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nwords;
       j += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nh + 1;  // sets 0..nh; bs_off has nh + 2 entries
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((uint64_t)t.bs_off[mid] <= j) lo = mid; else hi = mid;
    }
    const DSet s = set_of(t, lo);
    const uint32_t i = (uint32_t)(j - t.bs_off[lo]);
    const uint64_t lo_val = 64ull * (s.start + i);
    uint32_t a = 0, b = s.n;
    while (a < b) {
      const uint32_t mid = (a + b) >> 1;
      if ((uint64_t)s.v[mid] < lo_val) a = mid + 1; else b = mid;
    }
    uint64_t wv = 0;
    for (uint32_t k = a; k < s.n && (uint64_t)s.v[k] < lo_val + 64; k++) wv |= 1ull << (s.v[k] & 63);
    words[j] = wv;
    rank[j] = a;
  }
}

// Gap-coded lists: a child list of at least PK_MIN values whose consecutive
// gaps all fit 16 bits is copied into ceil(n / PK_VALS) 32-byte chunks; chunk
// c holds value c * PK_VALS as a uint32, then the 16-bit gaps to the next 14
// values (zeros past the end).  One lane per value (the split's run flags and
// list numbers say where a value's list starts and which list it is in), so
// hub lists of 10^5 values cost no more than many short ones.

// wide[h] = 1 when list h has a gap of 2^16 or more.  list(k) = pos[k + 1] - 1.
__global__ void k_pk_wide(const uint32_t *__restrict__ vals, const uint32_t *__restrict__ starts,
                          const uint64_t *__restrict__ pos, uint64_t m, uint8_t *__restrict__ wide) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x)
    if (!starts[k] && vals[k] - vals[k - 1] >= 65536u) wide[pos[k + 1] - 1] = 1;
}

__global__ void k_pk_count(DTrie t, const uint8_t *__restrict__ wide, uint32_t *__restrict__ cnt) {
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < t.nh; h += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t n = t.off[h + 1] - t.off[h];
    cnt[h] = (n >= PK_MIN && !wide[h]) ? (n + PK_VALS - 1) / PK_VALS : 0u;
  }
}

// pk16 is the chunk pool as 16-bit words (zeroed first): 16 per chunk, the
// first value in words 0-1, gap j (j = 1..14) in word j + 1.
__global__ void k_pk_fill(DTrie t, uint64_t m, const uint64_t *__restrict__ pos, uint16_t *__restrict__ pk16) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t h = pos[k + 1] - 1;
    const uint32_t c0 = t.pk_off[h];
    if (t.pk_off[h + 1] == c0) continue;
    const uint32_t r = (uint32_t)(k - t.off[h]), j = r % PK_VALS;
    uint16_t *q = pk16 + 16ull * (c0 + r / PK_VALS);
    if (j == 0) {
      const uint32_t v = t.vals[k];
      q[0] = (uint16_t)v;
      q[1] = (uint16_t)(v >> 16);
    } else {
      q[j + 1] = (uint16_t)(t.vals[k] - t.vals[k - 1]);
    }
  }
}

__global__ void k_nav(DTrie t, uint4 *__restrict__ nav) {
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < t.nh; h += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t a = t.off[h], n = t.off[h + 1] - a, wb = t.bs_off[h], nw = t.bs_off[h + 1] - wb;
    if (nw && n >= PK_MIN && NAV_WORDS_DENSITY * nw <= n) {  // read as bitset words
      nav[t.head[h]] = make_uint4(a, n, wb - t.bs_start[h], NAV_WORDS | t.bs_start[h]);
    } else {
      const uint32_t c = t.pk_off[h];
      nav[t.head[h]] = make_uint4(a, n, c, t.pk_off[h + 1] - c);
    }
  }
}

// Child pointers: arr[e] = dst's nav entry of element e's value.  With keys
// (the transposed hint build), also keys[e] = (e << 32) | value, and mx[0]
// takes the largest value (a stable sort on the low bits then keeps e
// ascending per value).
template <bool KEYS>
__global__ void k_link(const uint32_t *__restrict__ vals, uint64_t m, DTrie dst, uint4 *__restrict__ arr,
                       uint64_t *__restrict__ keys, unsigned *__restrict__ mx) {
  uint32_t v = 0;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t x = vals[e];
    arr[e] = (x < dst.nav_len) ? dst.nav[x] : make_uint4(0, 0, 0, 0);
    if (KEYS) {
      keys[e] = (e << 32) | x;
      v = max(v, x);
    }
  }
  if (KEYS) {
    for (int d = 32; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, 64));
    if (__lane_id() == 0) atomicMax(mx, v);
  }
}

// Sampled test of whether src is dst's reverse ordering, and how far apart the
// child sets that one set's search reads lie.  Sample s is element e = s * m / ns
// = (y, x): found[0] += 1 when y is in x's child set in dst; when e + 1 is in
// the same set, found[1] += 1 and gap[0] += the distance (in rows of dst,
// capped at LINK_GAP_CAP) between the child sets of e and e + 1.
constexpr uint32_t LINK_GAP_CAP = 1u << 24;
__global__ void k_link_probe(DTrie src, uint64_t m, DTrie dst, uint32_t ns, unsigned *__restrict__ found,
                             unsigned long long *__restrict__ gap) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const uint32_t e = (uint32_t)((uint64_t)s * m / ns);
  // the set holding e: the last h with off[h] <= e
  uint32_t lo = 0, hi = src.nh;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (src.off[mid] <= e) lo = mid;
    else hi = mid;
  }
  const uint32_t y = src.head[lo], x = src.vals[e];
  const uint4 c = x < dst.nav_len ? dst.nav[x] : make_uint4(0, 0, 0, 0);
  const uint32_t i = lower_bound(dst.vals, c.x, c.x + c.y, y);
  if (i < c.x + c.y && dst.vals[i] == y) atomicAdd(&found[0], 1u);
  if (e + 1 < src.off[lo + 1]) {
    const uint32_t x2 = src.vals[e + 1];
    const uint32_t o2 = x2 < dst.nav_len ? dst.nav[x2].x : 0u;
    atomicAdd(&found[1], 1u);
    atomicAdd(gap, (unsigned long long)min(o2 > c.x ? o2 - c.x : c.x - o2, LINK_GAP_CAP));
  }
}

// Hints from the keys sorted by value.  Sorted position k holds src's element
// e = (y, x); when dst's element k lies in x's child set and its value v has e
// in its own set of src (so y = v), dst holds y at rank k - (x's child-set
// offset) and that rank is the hint.  Every other element gets LINK_UNSET,
// which k_link_hint replaces by a search.
constexpr uint32_t LINK_UNSET = 0xFFFFFFFFu;
__global__ void k_link_cand(const uint64_t *__restrict__ keys, uint64_t m, DTrie src, DTrie dst,
                            uint32_t *__restrict__ hint) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = keys[k];
    const uint32_t x = (uint32_t)key, e = (uint32_t)(key >> 32);
    uint32_t r = LINK_UNSET;
    if (x < dst.nav_len) {
      const uint4 nx = dst.nav[x];
      const uint32_t rr = (uint32_t)k - nx.x;
      if (rr < nx.y) {
        const uint32_t v = dst.vals[k];
        if (v < src.nav_len) {
          const uint4 nv = src.nav[v];
          if (e - nv.x < nv.y) r = rr;
        }
      }
    }
    hint[e] = r;
  }
}

// One wavefront per set h of src.  hint (when not null): hint[e] = number of
// values of element e's child set in dst (arr[e]) below the set's value y,
// kept where the transposed build already found it (cand) and binary-searched
// elsewhere; any[0] = 1 when some hint is nonzero.  work (when not null):
// work[h] = the sum of the set's child-set lengths (whole lengths, not less
// the hints: binned by the trimmed lengths, the R-MAT-22 middle-vertex query
// ran 5.45 ms at its best knobs against 5.11 ms).
__global__ void k_link_hint(DTrie src, const uint32_t *__restrict__ dvals, const uint4 *__restrict__ arr,
                            uint32_t *__restrict__ hint, unsigned *__restrict__ any, int cand,
                            uint64_t *__restrict__ work) {
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  bool some = false;
  for (uint64_t h = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; h < src.nh; h += nwaves) {
    const uint32_t y = src.head[h], a1 = src.off[h + 1];
    unsigned long long w = 0;
    for (uint32_t e = src.off[h] + __lane_id(); e < a1; e += 64) {
      const uint4 c = arr[e];
      w += c.y;
      if (!hint) continue;
      uint32_t r = cand ? hint[e] : LINK_UNSET;
      if (r == LINK_UNSET) {
        uint32_t lo = 0, hi = c.y;
        if (hi && dvals[c.x] >= y) hi = 0;  // the whole set lies at or above y (the common case of
                                            // a relation linked to itself in ascending orientation)
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (dvals[c.x + mid] < y) lo = mid + 1;
          else hi = mid;
        }
        hint[e] = r = lo;
      }
      some |= r != 0;
    }
    if (work) {
      w = wave_sum(w);
      if (__lane_id() == 0) work[h] = w;
    }
  }
  if (hint && __any(some) && __lane_id() == 0) atomicOr(any, 1u);
}

__global__ void k_dense(const uint32_t *__restrict__ head, uint64_t nh, int32_t *__restrict__ dense) {
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += (uint64_t)gridDim.x * blockDim.x)
    dense[head[h]] = (int32_t)h;
}

// Per-set statistics for ehx_trie_info.
__global__ void k_stats(DTrie t, unsigned long long *__restrict__ acc) {
  unsigned long long nb = 0, el = 0, ns = 0;
  uint32_t mx = 0;
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < t.nh; h += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t n = t.off[h + 1] - t.off[h];
    mx = max(mx, n);
    if (t.layout[h] == EHX_RANGE_BITSET) { nb++; el += n; }
    if ((t.stage ? t.stage[h] : t.layout[h]) == EHX_RANGE_BITSET) ns++;
  }
  nb = wave_sum(nb);
  el = wave_sum(el);
  ns = wave_sum(ns);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, d, 64));
  if (__lane_id() == 0) {
    atomicAdd(&acc[0], nb);
    atomicAdd(&acc[1], el);
    atomicMax(&acc[2], (unsigned long long)mx);
    atomicAdd(&acc[3], ns);
  }
}

static int bits_of(uint32_t v) {
  int b = 0;
  while (b < 32 && (v >> b)) b++;
  return b;
}

int trie_destroy(ehx_trie *t);

static std::mutex g_link_mu;
static std::atomic<uint64_t> g_next_uid{1};

static void free_trie_arrays(ehx_trie *t) {
  if (t->block) (void)hipFree(t->block);
  t->block = nullptr;
}

int trie_build(ehx_ctx *ctx, const uint32_t *c0, const uint32_t *c1, uint64_t rows, int policy, int on_device,
               ehx_trie **out) {
  if (!ctx || !out) return fail(EHX_ERR_INVALID, "ehx_trie_build: null argument");
  if (rows == 0) return fail(EHX_ERR_INVALID, "ehx_trie_build: empty relation (Trie.cpp:394 asserts num_rows != 0)");
  if (rows >= (1ull << 32)) return fail(EHX_ERR_UNSUPPORTED, "ehx_trie_build: at most 2^32-1 rows");
  if (policy < EHX_POLICY_RULE || policy > EHX_POLICY_LEARNED) return fail(EHX_ERR_INVALID, "bad layout policy %d", policy);
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  EHX_HIP(hipEventRecord(ctx->ev[0], st));
  const uint32_t *d0 = c0, *d1 = c1;
  int rc;
  if (!on_device) {
    DevBuf &i0 = ctx->scratch[S_IN0], &i1 = ctx->scratch[S_IN1];
    if ((rc = i0.ensure(rows * 4)) || (rc = i1.ensure(rows * 4))) return rc;
    if ((rc = upload_h2d(ctx, i0.p, c0, rows * 4)) || (rc = upload_h2d(ctx, i1.p, c1, rows * 4))) return rc;
    d0 = i0.as<uint32_t>();
    d1 = i1.as<uint32_t>();
  }
  uint32_t max0 = 0, max1 = 0;
  if ((rc = reduce_max_u32(ctx, d0, rows, &max0)) || (rc = reduce_max_u32(ctx, d1, rows, &max1))) return rc;
  const int b0 = bits_of(max0), b1 = bits_of(max1);
  DevBuf &ka = ctx->scratch[S_SORT_A], &kb = ctx->scratch[S_SORT_B];
  DevBuf &fl = ctx->scratch[S_FLAGS], &ps = ctx->scratch[S_POS];
  if ((rc = ka.ensure(rows * 8)) || (rc = kb.ensure(rows * 8)) || (rc = fl.ensure(rows * 4)) ||
      (rc = ps.ensure((rows + 1) * 8)))
    return rc;
  uint64_t *keys = ka.as<uint64_t>(), *tmp = kb.as<uint64_t>();
  uint32_t *flags = fl.as<uint32_t>();
  uint64_t *pos = ps.as<uint64_t>();
  const unsigned G = grid_for(rows, 256, 16384);
  // rows already in order need no sort; rows whose second column never
  // decreases need only the passes over the first column's bits (a stable sort
  // on them keeps the second column ascending within each first-column value)
  unsigned *order = reinterpret_cast<unsigned *>(pos);
  if ((rc = fill_u32(ctx, order, 0u, 2))) return rc;
  k_pack<<<G, 256, 0, st>>>(d0, d1, rows, b1, keys, order);
  EHX_HIP(hipGetLastError());
  unsigned ord[2] = {1, 1};
  if ((rc = copy_d2h(ctx, ord, order, 8))) return rc;
  if (ord[0] && (rc = radix_sort_u64(ctx, keys, tmp, rows, b0 + b1, ord[1] ? 0 : b1))) return rc;
  // unique rows
  k_run_flags<<<G, 256, 0, st>>>(keys, rows, 0, flags);
  EHX_HIP(hipGetLastError());
  if ((rc = scan_u32_to_u64(ctx, flags, pos, rows))) return rc;
  uint64_t m = 0;
  if ((rc = copy_d2h(ctx, &m, pos + rows, 8))) return rc;
  k_compact_keys<<<G, 256, 0, st>>>(keys, flags, pos, rows, tmp);
  EHX_HIP(hipGetLastError());
  // head runs
  k_run_flags<<<G, 256, 0, st>>>(tmp, m, b1, flags);
  EHX_HIP(hipGetLastError());
  if ((rc = scan_u32_to_u64(ctx, flags, pos, m))) return rc;
  uint64_t nh = 0;
  if ((rc = copy_d2h(ctx, &nh, pos + m, 8))) return rc;

  ehx_trie *t = new ehx_trie();
  t->device = ctx->device;
  t->uid = g_next_uid++;
  t->m = m;
  t->nh = nh;
  // dense navigation array when it costs at most ~64 entries per head plus 4 MiB
  uint32_t maxhead = 0;
  // the last head value is the largest (sorted); fetched after the split
  // sizes: head nh, off nh+1, vals m, layout nh+1, bs_off nh+2, bs_start nh+1
  size_t o_head = 0;
  size_t o_off = o_head + ((nh * 4 + 255) & ~255ull);
  size_t o_vals = o_off + (((nh + 1) * 4 + 255) & ~255ull);
  size_t o_lay = o_vals + ((m * 4 + 255) & ~255ull);
  size_t o_bsoff = o_lay + (((nh + 1) + 255) & ~255ull);
  size_t o_bsst = o_bsoff + (((nh + 2) * 4 + 255) & ~255ull);
  size_t o_stage = o_bsst + (((nh + 1) * 4 + 255) & ~255ull);
  size_t total = o_stage + (((nh + 1) + 255) & ~255ull);
  hipError_t e = hipMalloc(&t->block, total);
  if (e != hipSuccess) {
    delete t;
    return fail(EHX_ERR_NOMEM, "trie arrays (%zu bytes): %s", total, hipGetErrorString(e));
  }
  uint8_t *B = (uint8_t *)t->block;
  t->head = (uint32_t *)(B + o_head);
  t->off = (uint32_t *)(B + o_off);
  t->vals = (uint32_t *)(B + o_vals);
  t->layout = B + o_lay;
  t->bs_off = (uint32_t *)(B + o_bsoff);
  t->bs_start = (uint32_t *)(B + o_bsst);
  t->stage = nullptr;  // set below for a learned build with selector weights loaded
  k_split<<<grid_for(m, 256, 16384), 256, 0, st>>>(tmp, flags, pos, m, b1, t->head, t->off, t->vals, nh);
  e = hipGetLastError();
  if (e != hipSuccess) { free_trie_arrays(t); delete t; return fail(EHX_ERR_DEVICE, "split: %s", hipGetErrorString(e)); }
  if ((rc = copy_d2h(ctx, &maxhead, t->head + (nh - 1), 4))) { free_trie_arrays(t); delete t; return rc; }

  // layouts, bitset pool
  DevBuf &wc = ctx->scratch[S_COUNTS];
  if ((rc = wc.ensure((nh + 2) * 4))) { free_trie_arrays(t); delete t; return rc; }
  SelectorWeights sel = selector_current();
  if (policy == EHX_POLICY_LEARNED && sel.valid) t->stage = B + o_stage;
  DTrie dv = view(t);
  dv.layout = t->layout;
  k_layouts<<<grid_for(nh + 1, 256, 16384), 256, 0, st>>>(dv, policy, sel, t->layout, t->stage, t->bs_start,
                                                          wc.as<uint32_t>());
  if ((rc = scan_u32(ctx, wc.as<uint32_t>(), t->bs_off, nh + 1))) { free_trie_arrays(t); delete t; return rc; }
  uint32_t W = 0;
  if ((rc = copy_d2h(ctx, &W, t->bs_off + (nh + 1), 4))) { free_trie_arrays(t); delete t; return rc; }
  uint64_t dense_len = (uint64_t)maxhead + 1;
  bool dense = dense_len <= 64ull * nh + (1ull << 20);
  size_t wbytes = ((uint64_t)W * 8 + 255) & ~255ull, rbytes = ((uint64_t)W * 4 + 255) & ~255ull;
  size_t dbytes = dense ? ((dense_len * 4 + 255) & ~255ull) : 0;
  void *blk2 = nullptr;
  e = hipMalloc(&blk2, wbytes + rbytes + dbytes + 256);
  if (e != hipSuccess) { free_trie_arrays(t); delete t; return fail(EHX_ERR_NOMEM, "bitset pool: %s", hipGetErrorString(e)); }
  // Keep the second block alive by chaining it: store its pointer in words.
  t->words = (uint64_t *)blk2;
  t->rank = (uint32_t *)((uint8_t *)blk2 + wbytes);
  t->dense = dense ? (int32_t *)((uint8_t *)blk2 + wbytes + rbytes) : nullptr;
  t->dense_len = dense ? (uint32_t)dense_len : 0;
  dv = view(t);
  if (W) {
    k_fill_words<<<grid_for(W, 256, 16384), 256, 0, st>>>(dv, W, t->words, t->rank);
  }
  EHX_HIP(hipGetLastError());
  if (dense) {
    if ((rc = fill_u32(ctx, (uint32_t *)t->dense, 0xFFFFFFFFu, dense_len))) return rc;
    k_dense<<<grid_for(nh, 256, 16384), 256, 0, st>>>(t->head, nh, t->dense);
    EHX_HIP(hipGetLastError());
  }
  // gap-coded copy of the long child lists
  {
    DevBuf &sc = ctx->scratch[S_COUNTS], &wb = ctx->scratch[S_OFFS];
    hipError_t e1 = hipMalloc((void **)&t->pk_off, (nh + 1) * 4);
    if (e1 != hipSuccess) { trie_destroy(t); return fail(EHX_ERR_NOMEM, "chunk offsets: %s", hipGetErrorString(e1)); }
    if ((rc = sc.ensure((nh + 1) * 4)) || (rc = wb.ensure(nh + 4))) { trie_destroy(t); return rc; }
    uint8_t *wide = wb.as<uint8_t>();
    if ((rc = fill_u32(ctx, (uint32_t *)wide, 0u, (nh + 3) / 4))) { trie_destroy(t); return rc; }
    k_pk_wide<<<grid_for(m, 256, 16384), 256, 0, st>>>(t->vals, flags, pos, m, wide);
    EHX_HIP(hipGetLastError());
    k_pk_count<<<grid_for(nh, 256, 16384), 256, 0, st>>>(dv, wide, sc.as<uint32_t>());
    EHX_HIP(hipGetLastError());
    if ((rc = scan_u32(ctx, sc.as<uint32_t>(), t->pk_off, nh))) { trie_destroy(t); return rc; }
    uint32_t npk = 0;  // at most m / 15 + m / 16 chunks: fits 32 bits
    if ((rc = copy_d2h(ctx, &npk, t->pk_off + nh, 4))) { trie_destroy(t); return rc; }
    t->npk = npk;
    e1 = hipMalloc((void **)&t->pk, (uint64_t)npk * 32 + 256);
    if (e1 != hipSuccess) { trie_destroy(t); return fail(EHX_ERR_NOMEM, "coded lists: %s", hipGetErrorString(e1)); }
    dv = view(t);
    if (npk) {
      if ((rc = fill_u32(ctx, (uint32_t *)t->pk, 0u, (uint64_t)npk * 8))) { trie_destroy(t); return rc; }
      k_pk_fill<<<grid_for(m, 256, 16384), 256, 0, st>>>(dv, m, pos, (uint16_t *)t->pk);
      EHX_HIP(hipGetLastError());
    }
  }
  // child lookup by value: 16 bytes per id of the head domain, when that domain is compact
  if (dense && dense_len <= 8ull * nh + (1ull << 20)) {
    hipError_t e1 = hipMalloc((void **)&t->nav, dense_len * 16);
    if (e1 != hipSuccess) { trie_destroy(t); return fail(EHX_ERR_NOMEM, "child lookup: %s", hipGetErrorString(e1)); }
    t->nav_len = (uint32_t)dense_len;
    if ((rc = fill_u32(ctx, (uint32_t *)t->nav, 0u, dense_len * 4))) { trie_destroy(t); return rc; }
    dv = view(t);
    k_nav<<<grid_for(nh, 256, 16384), 256, 0, st>>>(dv, t->nav);
    EHX_HIP(hipGetLastError());
  }
  // stats
  DevBuf &ac = ctx->scratch[S_ACC];
  if ((rc = ac.ensure(64))) return rc;
  if ((rc = fill_u64(ctx, ac.as<uint64_t>(), 0, 4))) return rc;
  k_stats<<<grid_for(nh, 256, 4096), 256, 0, st>>>(dv, ac.as<unsigned long long>());
  EHX_HIP(hipGetLastError());
  EHX_HIP(hipEventRecord(ctx->ev[1], st));
  unsigned long long acc[4] = {0, 0, 0, 0};
  if ((rc = copy_d2h(ctx, acc, ac.p, 32))) return rc;
  uint8_t hl = 0;
  if ((rc = copy_d2h(ctx, &hl, t->layout + nh, 1))) return rc;
  float ms = 0;
  EHX_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  ehx_trie_meta &I = t->info;
  I.num_rows = m;
  I.num_heads = nh;
  I.num_bitsets = acc[0];
  I.bitset_elems = acc[1];
  I.max_card = (uint32_t)acc[2];
  I.bitset_words = W;
  I.head_layout = hl;
  I.dense_nav = dense ? 1u : 0u;
  I.max_value = max0 > max1 ? max0 : max1;
  I.build_ms = ms;
  I.num_staged_bitmaps = acc[3];
  *out = t;
  return EHX_OK;
}

// Transposed hint build: taken (automatic route) for relations of at least
// LINK_SORT_MIN rows whose LINK_SAMPLES sampled elements all find their parent
// value in their child set (dst looks like src's reverse ordering), when the
// child sets that consecutive elements of one set read lie at least
// LINK_SORT_GAP rows of dst apart on average.  Searches closer together stay in
// cache and beat the sort: the ego-network batch cut from R-MAT-22, renumbered
// net by net, averages 393 rows (R->T) and 1,548 (its symmetric relation
// linked to itself), R-MAT-22 linked to its reverse 172,897, R-MAT-18 17,118.
static constexpr uint64_t LINK_SORT_MIN = 1u << 16;
static constexpr uint32_t LINK_SAMPLES = 1024;
static constexpr uint64_t LINK_SORT_GAP = 1u << 14;

// The link's own flag words (a caller may hold live data in the other scratch
// slots): [0] samples found, [1] samples with a successor, [2, 3] their gap
// sum, [4] largest value of src, [5] some hint nonzero.
enum { LF_FOUND = 0, LF_PAIRS = 1, LF_GAP = 2, LF_MAX = 4, LF_ANY = 5, LF_WORDS = 6 };

static int link_sort_route(ehx_ctx *ctx, const ehx_trie *src, const ehx_trie *dst, unsigned *fl, int route,
                           bool *sort) {
  *sort = false;
  if (route == LINK_SEARCH || !src->nav || !dst->nav || (route == LINK_AUTO && src->m < LINK_SORT_MIN)) return EHX_OK;
  k_link_probe<<<(LINK_SAMPLES + 255) / 256, 256, 0, ctx->stream>>>(view(src), src->m, view(dst), LINK_SAMPLES, fl,
                                                                  reinterpret_cast<unsigned long long *>(fl + LF_GAP));
  EHX_HIP(hipGetLastError());
  unsigned f[4] = {0, 0, 0, 0};
  int rc = copy_d2h(ctx, f, fl, sizeof(f));
  if (rc) return rc;
  const uint64_t gap = (uint64_t)f[LF_GAP] | ((uint64_t)f[LF_GAP + 1] << 32);
  if (f[LF_FOUND] != LINK_SAMPLES) return EHX_OK;
  *sort = route == LINK_SORT || (f[LF_PAIRS] && gap >= LINK_SORT_GAP * f[LF_PAIRS]);
  return EHX_OK;
}

int trie_link(ehx_ctx *ctx, ehx_trie *src, const ehx_trie *dst, const uint4 **out, const uint32_t **hint_out,
              const uint64_t **work_out, int route) {
  *out = nullptr;
  if (hint_out) *hint_out = nullptr;
  if (work_out) *work_out = nullptr;
  if (!dst->nav || !src->m) return EHX_OK;
  std::lock_guard<std::mutex> lk(g_link_mu);
  for (const auto &l : src->links)
    if (l.target == dst->uid) {
      *out = l.arr;
      if (hint_out) *hint_out = l.hint;
      if (work_out) *work_out = l.work;
      return EHX_OK;
    }
  DevBuf &fb = ctx->scratch[S_FLAGS];
  int rc;
  if ((rc = fb.ensure(64)) || (rc = fill_u32(ctx, fb.as<uint32_t>(), 0u, LF_WORDS))) return rc;
  unsigned *fl = fb.as<unsigned>();
  bool sort = false;
  if ((rc = link_sort_route(ctx, src, dst, fl, route, &sort))) return rc;
  uint4 *arr = nullptr;
  uint32_t *hint = nullptr;
  uint64_t *work = nullptr, *keys = nullptr;
  // everything allocated before the first launch that writes it; a failed
  // allocation drops what it would hold (no hints: child sets are streamed
  // whole; no work sums: the binning samples the lists; no keys: every hint
  // is searched)
  if (hipMalloc((void **)&arr, src->m * 16) != hipSuccess) {
    (void)hipGetLastError();
    return EHX_OK;  // no room: the level looks children up by value instead
  }
  if (hipMalloc((void **)&hint, src->m * 4) != hipSuccess) {
    (void)hipGetLastError();
    hint = nullptr;
  }
  if (hipMalloc((void **)&work, src->nh * 8 + 8) != hipSuccess) {
    (void)hipGetLastError();
    work = nullptr;
  }
  if (sort && hint && hipMalloc((void **)&keys, src->m * 16) != hipSuccess) {
    (void)hipGetLastError();
    keys = nullptr;
  }
  hipError_t e = hipSuccess;
  const unsigned g = grid_for(src->m, 256, 16384);
  if (keys) k_link<true><<<g, 256, 0, ctx->stream>>>(src->vals, src->m, view(dst), arr, keys, fl + LF_MAX);
  else k_link<false><<<g, 256, 0, ctx->stream>>>(src->vals, src->m, view(dst), arr, nullptr, nullptr);
  e = hipGetLastError();
  int cand = 0;
  if (e == hipSuccess && keys) {
    unsigned mx = 0;
    uint64_t *sorted = keys;
    if (!(rc = copy_d2h(ctx, &mx, fl + LF_MAX, 4)) &&
        !(rc = radix_sort_u64(ctx, keys, keys + src->m, src->m, mx ? 32 - __builtin_clz(mx) : 1, 0, &sorted))) {
      k_link_cand<<<g, 256, 0, ctx->stream>>>(sorted, src->m, view(src), view(dst), hint);
      e = hipGetLastError();
      cand = 1;
    }
  }
  if (e == hipSuccess && !rc && (hint || work)) {
    k_link_hint<<<grid_for(src->nh * 64, 256, 16384), 256, 0, ctx->stream>>>(view(src), dst->vals, arr, hint,
                                                                            fl + LF_ANY, cand, work);
    e = hipGetLastError();
  }
  unsigned any = 0;
  if (e == hipSuccess && !rc && hint) rc = copy_d2h(ctx, &any, fl + LF_ANY, 4);
  if (e == hipSuccess && !rc) e = hipStreamSynchronize(ctx->stream);  // other contexts may use it next
  if (keys) (void)hipFree(keys);
  if (e != hipSuccess || rc) {
    (void)hipFree(arr);
    if (hint) (void)hipFree(hint);
    if (work) (void)hipFree(work);
    return rc ? rc : fail(EHX_ERR_DEVICE, "trie link: %s", hipGetErrorString(e));
  }
  // hints are kept only when some child set starts below its element's parent
  // value (for a relation linked to itself in ascending orientation none does)
  if (hint && !any) {
    (void)hipFree(hint);
    hint = nullptr;
  }
  src->links.push_back({dst->uid, arr, hint, work, hint ? cand : 0});
  *out = arr;
  if (hint_out) *hint_out = hint;
  if (work_out) *work_out = work;
  return EHX_OK;
}

int trie_link_hints(ehx_ctx *ctx, ehx_trie *src, const ehx_trie *dst, int route, uint32_t *host, uint64_t n,
                    int *state) {
  if (n != src->m) return fail(EHX_ERR_INVALID, "hint buffer holds %llu values, the trie %llu rows",
                               (unsigned long long)n, (unsigned long long)src->m);
  if (route != LINK_AUTO && route != LINK_SEARCH && route != LINK_SORT)
    return fail(EHX_ERR_INVALID, "link route %d", route);
  const uint4 *arr = nullptr;
  const uint32_t *hint = nullptr;
  int rc = trie_link(ctx, src, dst, &arr, &hint, nullptr, route);
  if (rc) return rc;
  int transposed = 0;
  {
    std::lock_guard<std::mutex> lk(g_link_mu);
    for (const auto &l : src->links)
      if (l.target == dst->uid) transposed = l.transposed;
  }
  *state = !arr ? 0 : !hint ? 1 : transposed ? 3 : 2;
  if (hint) return copy_d2h(ctx, host, hint, n * 4);
  memset(host, 0, n * 4);
  return EHX_OK;
}

int trie_destroy(ehx_trie *t) {
  if (!t) return EHX_OK;
  EHX_HIP(hipSetDevice(t->device));
  for (const auto &l : t->links) {
    EHX_HIP(hipFree(l.arr));
    if (l.hint) EHX_HIP(hipFree(l.hint));
    if (l.work) EHX_HIP(hipFree(l.work));
  }
  if (t->block) EHX_HIP(hipFree(t->block));
  if (t->words) EHX_HIP(hipFree(t->words));
  if (t->pk_off) EHX_HIP(hipFree(t->pk_off));
  if (t->pk) EHX_HIP(hipFree(t->pk));
  if (t->nav) EHX_HIP(hipFree(t->nav));
  delete t;
  return EHX_OK;
}

__global__ void k_export(DTrie t, uint32_t *__restrict__ c0) {
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t h = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; h < t.nh; h += nwaves) {
    uint32_t a = t.off[h], b = t.off[h + 1], x = t.head[h];
    for (uint32_t i = a + __lane_id(); i < b; i += 64) c0[i] = x;
  }
}

int trie_export(ehx_ctx *ctx, const ehx_trie *t, uint32_t *c0, uint32_t *c1) {
  EHX_HIP(hipSetDevice(ctx->device));
  DevBuf &b = ctx->scratch[S_IN0];
  int rc = b.ensure(t->m * 4 + 4);
  if (rc) return rc;
  k_export<<<grid_for(t->nh * 64, 256, 16384), 256, 0, ctx->stream>>>(view(t), b.as<uint32_t>());
  EHX_HIP(hipGetLastError());
  if ((rc = copy_d2h(ctx, c0, b.p, t->m * 4))) return rc;
  return copy_d2h(ctx, c1, t->vals, t->m * 4);
}

__global__ void k_node_info(DTrie t, uint8_t *__restrict__ types, uint32_t *__restrict__ cards) {
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < t.nh; h += (uint64_t)gridDim.x * blockDim.x) {
    types[h] = t.layout[h];
    cards[h] = t.off[h + 1] - t.off[h];
  }
}

int trie_node_info(ehx_ctx *ctx, const ehx_trie *t, uint8_t *types, uint32_t *cards) {
  EHX_HIP(hipSetDevice(ctx->device));
  DevBuf &b = ctx->scratch[S_IN0];
  int rc = b.ensure(t->nh * 5 + 64);
  if (rc) return rc;
  uint32_t *dc = b.as<uint32_t>();
  uint8_t *dt = (uint8_t *)(dc + t->nh);
  k_node_info<<<grid_for(t->nh, 256, 16384), 256, 0, ctx->stream>>>(view(t), dt, dc);
  EHX_HIP(hipGetLastError());
  if (types && (rc = copy_d2h(ctx, types, dt, t->nh))) return rc;
  if (cards && (rc = copy_d2h(ctx, cards, dc, t->nh * 4))) return rc;
  return EHX_OK;
}

__global__ void k_lookup(DTrie t, const uint32_t *__restrict__ keys, uint64_t n, int64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = node_of(t, keys[i]);
}

int trie_lookup(ehx_ctx *ctx, const ehx_trie *t, const uint32_t *keys, uint64_t n, int64_t *nodes) {
  EHX_HIP(hipSetDevice(ctx->device));
  if (!n) return EHX_OK;
  DevBuf &a = ctx->scratch[S_IN0], &b = ctx->scratch[S_IN1];
  int rc;
  if ((rc = a.ensure(n * 4)) || (rc = b.ensure(n * 8))) return rc;
  if ((rc = copy_h2d(ctx, a.p, keys, n * 4))) return rc;
  k_lookup<<<grid_for(n, 256, 16384), 256, 0, ctx->stream>>>(view(t), a.as<uint32_t>(), n, b.as<int64_t>());
  EHX_HIP(hipGetLastError());
  return copy_d2h(ctx, nodes, b.p, n * 8);
}

}  // namespace ehx
