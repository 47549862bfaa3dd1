// ehx_internal.h -- shared declarations of libehx (HIP for gfx950).
//
// Device data layout of a two-level trie (one binary relation), struct of
// arrays in HBM:
//   head[nh]          sorted distinct first-column values
//   off[nh+1]         child range of head node h: vals[off[h] .. off[h+1])
//   vals[m]           second-column values, sorted within each child set
//   layout[nh+1]      per-set layout (EHX_UINTEGER / EHX_RANGE_BITSET); [nh] is the head set
//   bs_off[nh+2]      word range of set h in the bitset pool (empty for uint sets)
//   bs_start[nh+1]    first 64-bit word index (value >> 6) covered by set h
//   words[W], rank[W] range_bitset words and per-word rank (elements before the word)
//   dense[dense_len]  value -> head node (or -1): the NextLevel[data] array of a dense head
// Every set keeps its sorted values (for enumeration); the layout decides which
// representation intersections read.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ehx.h"
#include "../../include/ehx_ext.h"

#define EHX_MAX_RELS 8

namespace ehx {

void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define EHX_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return ::ehx::fail(EHX_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                                \
  } while (0)

// Grow-only device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes);
  void release();
  template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};
// Changes whenever any context buffer is freed or reallocated.
uint64_t devbuf_generation();

}  // namespace ehx

// Events bracketing the last level's main launch.
#define EHX_EV_K0 (2 * EHX_MAX_ATTRS + 2)
#define EHX_EV_K1 (2 * EHX_MAX_ATTRS + 3)

// Pinned host staging for large uploads (upload_h2d): UPLOAD_THREADS host
// threads, each with two UPLOAD_CHUNK-byte pinned buffers and an event per
// buffer; allocated on the first large upload, freed with the context.
constexpr int UPLOAD_THREADS = 8;
constexpr size_t UPLOAD_CHUNK = 8u << 20;

struct ehx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2 * EHX_MAX_ATTRS + 4] = {};
  ehx::DevBuf scratch[28];
  int num_cus = 256;
  void *pin[UPLOAD_THREADS][2] = {};
  hipEvent_t pin_ev[UPLOAD_THREADS][2] = {};
};

struct ehx_trie {
  int device = 0;
  uint64_t m = 0, nh = 0;
  uint32_t *head = nullptr, *off = nullptr, *vals = nullptr;
  uint8_t *layout = nullptr;
  // learned tries: the selector's layout per set, used only to stage a set as
  // the staged operand of the tuned level (bitmap when EHX_RANGE_BITSET);
  // `layout` keeps the density rule for every other use (null otherwise)
  uint8_t *stage = nullptr;
  uint32_t *bs_off = nullptr, *bs_start = nullptr, *rank = nullptr;
  uint64_t *words = nullptr;
  int32_t *dense = nullptr;
  uint32_t dense_len = 0;
  // gap-coded copy of long child lists, streamed by the tuned level: list h
  // owns chunks pk_off[h] .. pk_off[h+1] - 1 of pk (none: not coded); a chunk is
  // 32 bytes holding PK_VALS values, the first as a uint32 and then the 16-bit
  // gaps between consecutive values
  uint32_t *pk_off = nullptr;
  uint4 *pk = nullptr;
  uint64_t npk = 0;
  // child lookup by head value (dense tries with a compact head domain):
  // nav[x] = {off, length, first chunk, chunks} of x's child list, zeros when x
  // is not a head; for a list read as bitset words (NAV_WORDS set in .w),
  // {off, length, word base, NAV_WORDS | first word}
  uint4 *nav = nullptr;
  uint32_t nav_len = 0;
  // child pointers into other tries (or this one), built on first use: for
  // target trie u, arr[e] = u's nav entry of vals[e] (the reference's tries keep
  // a child pointer beside every set element; here the child may live in a
  // different relation's trie)
  // hint[e]: how many values of that child set are below the head value of
  // the set holding element e of this trie (the parent of e); null when the
  // device had no room
  // work[h]: the sum over set h of this trie of its elements' child-set
  // lengths: the tuned level's binning reads it instead of sampling
  // transposed: the hints started from candidates of one sort (the target
  // looked like this trie's reverse ordering) rather than a search each
  struct Link {
    uint64_t target;
    uint4 *arr;
    uint32_t *hint;
    uint64_t *work;
    int transposed;
  };
  std::vector<Link> links;
  uint64_t uid = 0;  // unique per process, never reused
  ehx_trie_meta info{};
  void *block = nullptr;  // single allocation holding most arrays
};

namespace ehx {

// Scratch slots of a context.
enum Slot {
  S_SCAN0 = 0, S_SCAN1, S_SCAN2, S_SCAN3,  // scan block sums (recursion levels)
  S_SORT_A, S_SORT_B, S_HIST, S_FLAGS, S_POS, S_MISC,
  S_F0, S_F1, S_F2, S_F3, S_F4, S_F5, S_F6, S_F7,  // frontier columns
  S_COUNTS, S_OFFS, S_QUEUE, S_ACC, S_IN0, S_IN1,
  S_SEGB, S_SEGO,  // segment bounds and per-segment counters
  S_DESC,          // per-prefix operand descriptors of the tuned level
  S_NUM_SLOTS
};
static_assert(S_NUM_SLOTS <= (int)(sizeof(ehx_ctx::scratch) / sizeof(ehx_ctx::scratch[0])), "context scratch slots");

// ---- device view of a trie --------------------------------------------------
// values per 32-byte gap-coded chunk: a uint32 first value and 14 16-bit gaps
constexpr uint32_t PK_VALS = 15;
// child lists shorter than this are not gap-coded
constexpr uint32_t PK_MIN = 16;
// A child list of at least PK_MIN values whose bitset window holds at most one
// 64-bit word per NAV_WORDS_DENSITY values is read by the tuned level as those
// words (AND + popcount against a staged bitmap, 64 values per word) instead
// of value by value.  Its nav entry then carries NAV_WORDS in .w, the first
// word index (first value >> 6) in .w's low bits and, in .z, the pool offset
// of its words less that index (word w of the list is words[.z + w], mod 2^32).
constexpr uint32_t NAV_WORDS = 0x80000000u;
constexpr uint32_t NAV_WORDS_DENSITY = 4;

struct DTrie {
  const uint32_t *head, *off, *vals;
  const uint8_t *layout;
  const uint8_t *stage;  // learned tries: the selector's staging layout (else null)
  const uint32_t *bs_off, *bs_start, *rank;
  const uint64_t *words;
  const int32_t *dense;
  const uint32_t *pk_off;
  const uint4 *pk;
  const uint4 *nav;
  uint32_t nav_len;
  uint32_t dense_len, nh;
};

inline DTrie view(const ehx_trie *t) {
  DTrie d;
  d.head = t->head; d.off = t->off; d.vals = t->vals; d.layout = t->layout; d.stage = t->stage;
  d.bs_off = t->bs_off; d.bs_start = t->bs_start; d.rank = t->rank; d.words = t->words;
  d.dense = t->dense; d.dense_len = t->dense_len; d.nh = (uint32_t)t->nh;
  d.pk_off = t->pk_off; d.pk = t->pk;
  d.nav = t->nav; d.nav_len = t->nav_len;
  return d;
}

// A set on the device: sorted values plus, for bitsets, the word window.
struct DSet {
  const uint32_t *v;
  const uint64_t *w;
  uint32_t n, W, start, lay;
};

__device__ __forceinline__ int32_t node_of(const DTrie &t, uint32_t x) {
  if (t.dense) return (x < t.dense_len) ? t.dense[x] : -1;
  uint32_t lo = 0, hi = t.nh;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (t.head[mid] < x) lo = mid + 1; else hi = mid;
  }
  return (lo < t.nh && t.head[lo] == x) ? (int32_t)lo : -1;
}

__device__ __forceinline__ DSet set_of(const DTrie &t, uint32_t h) {
  DSet s;
  if (h == t.nh) { s.v = t.head; s.n = t.nh; }
  else { uint32_t a = t.off[h]; s.v = t.vals + a; s.n = t.off[h + 1] - a; }
  s.lay = t.layout[h];
  uint32_t b = t.bs_off[h];
  s.W = t.bs_off[h + 1] - b;
  s.w = t.words + b;
  s.start = t.bs_start[h];
  return s;
}

__device__ __forceinline__ bool bit_test(const DSet &s, uint32_t x) {
  uint32_t wi = (x >> 6) - s.start;
  if ((x >> 6) < s.start || wi >= s.W) return false;
  return (s.w[wi] >> (x & 63)) & 1ull;
}

// Lower bound in a sorted array (first index with v[i] >= x).
__device__ __forceinline__ uint32_t lower_bound(const uint32_t *v, uint32_t lo, uint32_t hi, uint32_t x) {
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (v[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool set_contains(const DSet &s, uint32_t x) {
  if (s.lay == EHX_RANGE_BITSET) return bit_test(s, x);
  uint32_t i = lower_bound(s.v, 0, s.n, x);
  return i < s.n && s.v[i] == x;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanes_below() {
  uint32_t l = __lane_id();
  return l ? (~0ull >> (64 - l)) : 0ull;
}

template <class T> __device__ __forceinline__ T wave_incl_scan(T v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T t = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
}

template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// ---- host-side primitives (ehx_prims.hip) ------------------------------------
// Exclusive scan of n counts into out[0..n]; out[n] = total.  `level` picks the
// scratch slots used for recursion.
int scan_u32_to_u64(ehx_ctx *ctx, const uint32_t *in, uint64_t *out, uint64_t n);
// Sizes the scans' block-sum buffers for n elements (a later scan of at most n
// allocates nothing).
int scan_reserve(ehx_ctx *ctx, uint64_t n);
int scan_u64(ehx_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n);
int scan_u32(ehx_ctx *ctx, const uint32_t *in, uint32_t *out, uint64_t n);
// Stable LSD radix sort of n 64-bit keys on their low `bits` bits; keys[] holds
// the result, tmp[] has n entries.
// Keys already sorted on their low `from` bits need only the passes over bits
// [from, bits).  With `result`, the sorted keys are left in whichever of keys
// and tmp the last pass wrote, and *result points there.
int radix_sort_u64(ehx_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, int bits, int from = 0,
                   uint64_t **result = nullptr);
int reduce_max_u32(ehx_ctx *ctx, const uint32_t *in, uint64_t n, uint32_t *host_out);
int fill_u32(ehx_ctx *ctx, uint32_t *p, uint32_t v, uint64_t n);
// How a link finds its hints: chosen from a sample (LINK_AUTO), one search per
// element, or one sort of src by value whenever every sampled element finds its
// parent value in dst (hints checked against dst and searched where that fails).
enum { LINK_AUTO = 0, LINK_SEARCH = 1, LINK_SORT = 2 };
// src's child pointers into dst (built and cached on src on first use; the
// route applies to that first build); null in *out when dst has no child
// lookup by value
int trie_link(ehx_ctx *ctx, ehx_trie *src, const ehx_trie *dst, const uint4 **out,
              const uint32_t **hint_out = nullptr, const uint64_t **work_out = nullptr, int route = LINK_AUTO);
// Builds (once) the link and copies its hints for src's n = m rows to host
// (zeros without hints); *state: 0 no link, 1 no hints (all zero), 2 searched
// hints, 3 hints from the sort.
int trie_link_hints(ehx_ctx *ctx, ehx_trie *src, const ehx_trie *dst, int route, uint32_t *host, uint64_t n,
                    int *state);
// The loaded per-pair table on the current device (null: none).
const uint8_t *pair_table_current();
int pair_table_load(int device, const uint8_t *table);
int fill_u64(ehx_ctx *ctx, uint64_t *p, uint64_t v, uint64_t n);
int copy_d2h(ehx_ctx *ctx, void *dst, const void *src, size_t bytes);
int copy_h2d(ehx_ctx *ctx, void *dst, const void *src, size_t bytes);
// Enqueues the copy of `bytes` of pageable host memory to dst on the context's
// stream; returns once src may be reused (the copy may still be in flight).
// Large copies go through the context's pinned staging, several host threads
// filling it while earlier chunks are on the bus.
int upload_h2d(ehx_ctx *ctx, void *dst, const void *src, size_t bytes);
void upload_release(ehx_ctx *ctx);

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 20) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace ehx
