/*
 * ehx_oracle.c -- CPU restatement of EmptyHeaded's generic-join hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see ehx_oracle.h): the checker for the HIP engine,
 * and the "port" CPU baseline.  Never linked into the product.
 *
 * Citations are relative to the reference tree's storage_engine/ unless noted.
 */
#define _GNU_SOURCE
#include "ehx_oracle.h"

#include <errno.h>
#include <math.h>
#include <sys/stat.h>
#include <stdio.h>
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Layout rule: common.hpp:56-67 (bitset_length = 2, bitset_req = 128.0).     */
/* (double)range/len > 128.0 equals range > 128*len for 32-bit inputs.        */
/* ------------------------------------------------------------------------- */
int orc_is_sparse(uint64_t length, uint64_t range) {
  if (length > 2) return ((double)range / (double)length) > 128.0;
  return 1;
}

/* hybrid::get_type (hybrid.hpp:103-111): range = max - min of the set. */
int orc_layout_of(const uint32_t *v, uint64_t n) {
  uint64_t span = (n > 1) ? (uint64_t)(v[n - 1] - v[0]) : 0;
  return orc_is_sparse(n, span) ? ORC_UINTEGER : ORC_RANGE_BITSET;
}

static inline uint64_t word_of(uint32_t x) { return (uint64_t)(x >> 6); }

/* range_bitset bytes: [u64 start_word][W x u64 words][W x u32 rank] with
 * W = word(max) - word(min) + 1 (range_bitset.hpp:123-157). */
uint64_t orc_bytes_needed(uint64_t n, const uint32_t *v, int layout) {
  if (n == 0) return 0;
  if (layout == ORC_UINTEGER) return 4 * n;
  uint64_t w = word_of(v[n - 1]) - word_of(v[0]) + 1;
  return w * 12 + 8;
}

uint64_t orc_set_build(uint8_t *out, const uint32_t *v, uint64_t n, int layout, orc_set *s) {
  s->cardinality = (uint32_t)n;
  s->range = n ? v[n - 1] : 0; /* Set::from_array (Set.hpp:164-172) */
  s->type = (uint8_t)layout;
  s->data = out;
  if (n == 0) { s->number_of_bytes = 0; return 0; }
  if (layout == ORC_UINTEGER) {
    memcpy(out, v, 4 * n);
    s->number_of_bytes = 4 * n;
    return 4 * n;
  }
  uint64_t start = word_of(v[0]);
  uint64_t W = word_of(v[n - 1]) - start + 1;
  uint64_t *hdr = (uint64_t *)out;
  uint64_t *words = (uint64_t *)(out + 8);
  uint32_t *rank = (uint32_t *)(out + 8 + 8 * W);
  hdr[0] = start;
  memset(words, 0, 8 * W);
  /* rank[w] = number of elements in words before w.  The reference writes it
   * only for non-empty words (range_bitset.hpp:139); defining it everywhere is
   * consistent with every read, which only touches non-empty words. */
  uint64_t i = 0;
  for (uint64_t w = 0; w < W; w++) {
    rank[w] = (uint32_t)i;
    while (i < n && word_of(v[i]) - start == w) {
      words[w] |= (uint64_t)1 << (v[i] & 63);
      i++;
    }
  }
  s->number_of_bytes = W * 12 + 8;
  return s->number_of_bytes;
}

static inline uint64_t bs_words(const orc_set *s) { return (s->number_of_bytes - 8) / 12; }
static inline uint64_t bs_start(const orc_set *s) { return ((const uint64_t *)s->data)[0]; }
static inline const uint64_t *bs_w(const orc_set *s) { return (const uint64_t *)(s->data + 8); }
static inline const uint32_t *bs_rank(const orc_set *s) {
  return (const uint32_t *)(s->data + 8 + 8 * bs_words(s));
}

/* uinteger::find -> utils::binary_search (uinteger.hpp:272-281, utils.hpp:22-43);
 * range_bitset::find: bit test then popcount(word & bits-below) + rank
 * (range_bitset.hpp:443-470, sse_masks.hpp:12-21). */
int64_t orc_set_find(const orc_set *s, uint32_t key) {
  if (s->number_of_bytes == 0) return -1;
  if (s->type == ORC_UINTEGER) {
    const uint32_t *d = (const uint32_t *)s->data;
    uint64_t lo = 0, hi = s->number_of_bytes / 4;
    while (lo < hi) {
      uint64_t mid = (lo + hi) / 2;
      if (d[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < s->number_of_bytes / 4 && d[lo] == key) ? (int64_t)lo : -1;
  }
  uint64_t W = bs_words(s), start = bs_start(s), w = word_of(key);
  if (w < start || w >= start + W) return -1;
  uint64_t word = bs_w(s)[w - start];
  if (!((word >> (key & 63)) & 1)) return -1;
  uint64_t below = (key & 63) ? (word & (((uint64_t)1 << (key & 63)) - 1)) : 0;
  return (int64_t)(__builtin_popcountll(below) + bs_rank(s)[w - start]);
}

uint64_t orc_set_decode(const orc_set *s, uint32_t *out) {
  if (s->cardinality == 0 || s->number_of_bytes == 0) return 0;
  if (s->type == ORC_UINTEGER) {
    memcpy(out, s->data, 4 * (uint64_t)s->cardinality);
    return s->cardinality;
  }
  uint64_t W = bs_words(s), start = bs_start(s), k = 0;
  const uint64_t *w = bs_w(s);
  for (uint64_t i = 0; i < W; i++) {
    uint64_t x = w[i];
    while (x) {
      int b = __builtin_ctzll(x);
      out[k++] = (uint32_t)(64 * (i + start) + (uint64_t)b);
      x &= x - 1;
    }
  }
  return k;
}

/* ------------------------------------------------------------------------- */
/* uint x uint.  Both algorithms compute the same sorted intersection; the    */
/* reference picks galloping iff min != 0 and max/min > 32 (uinteger.hpp:     */
/* 1209-1227), otherwise the 4x4 SIMD merge (set_intersect_shuffle, :1121).   */
/* out may be NULL (count only).                                              */
/* ------------------------------------------------------------------------- */
uint64_t orc_uu_merge(uint32_t *out, const uint32_t *a, uint64_t na, const uint32_t *b, uint64_t nb) {
  uint64_t i = 0, j = 0, c = 0;
  if (out) {
    while (i < na && j < nb) {
      uint32_t x = a[i], y = b[j];
      if (x == y) { out[c++] = x; i++; j++; }
      else if (x < y) i++;
      else j++;
    }
  } else {
    while (i < na && j < nb) {
      uint32_t x = a[i], y = b[j];
      c += (x == y);
      i += (x <= y);
      j += (y <= x);
    }
  }
  return c;
}

/* Galloping (uinteger.hpp:768-1014 restated as exponential + binary search of
 * each rare key in the frequent set, keeping the search frontier). */
uint64_t orc_uu_gallop(uint32_t *out, const uint32_t *rare, uint64_t nr, const uint32_t *freq, uint64_t nf) {
  uint64_t lo = 0, c = 0;
  for (uint64_t i = 0; i < nr && lo < nf; i++) {
    uint32_t key = rare[i];
    if (freq[lo] < key) {
      uint64_t step = 1, hi = lo + 1;
      while (hi < nf && freq[hi] < key) { lo = hi; step <<= 1; hi = lo + step; }
      if (hi > nf) hi = nf;
      /* freq[lo] < key; answer in (lo, hi] */
      uint64_t l = lo + 1, h = hi;
      while (l < h) { uint64_t m = (l + h) / 2; if (freq[m] < key) l = m + 1; else h = m; }
      lo = l;
      if (lo >= nf) break;
    }
    if (freq[lo] == key) { if (out) out[c] = key; c++; lo++; }
  }
  return c;
}

static uint64_t uu(uint32_t *out, const orc_set *A, const orc_set *B) {
  /* rare = the smaller set (uinteger.hpp:1255-1259). */
  const orc_set *r = (A->cardinality > B->cardinality) ? B : A;
  const orc_set *f = (A->cardinality > B->cardinality) ? A : B;
  uint64_t mn = r->cardinality, mx = f->cardinality;
  const uint32_t *rd = (const uint32_t *)r->data, *fd = (const uint32_t *)f->data;
  if (mn != 0 && (double)mx / (double)mn > 32.0) return orc_uu_gallop(out, rd, mn, fd, mx);
  return orc_uu_merge(out, rd, mn, fd, mx);
}

/* uint x bitset probe (hetero.hpp:29-60): test each uint in the bitset window,
 * stop once past it. */
static uint64_t ub(uint32_t *out, const orc_set *U, const orc_set *B) {
  const uint32_t *a = (const uint32_t *)U->data;
  uint64_t W = B->number_of_bytes ? bs_words(B) : 0;
  uint64_t start = B->number_of_bytes ? bs_start(B) : 0;
  const uint64_t *w = B->number_of_bytes ? bs_w(B) : NULL;
  uint64_t c = 0;
  for (uint64_t i = 0; i < U->cardinality; i++) {
    uint32_t x = a[i];
    uint64_t wi = word_of(x);
    if (wi >= start + W) break;
    if (wi >= start && ((w[wi - start] >> (x & 63)) & 1)) { if (out) out[c] = x; c++; }
  }
  return c;
}

/* bitset x bitset (bitset.hpp:619-661 with intersect_range_block :526-567 and
 * bs_materialize_null::unpack_range :118-129): AND over the overlapping word
 * window, popcount; when materializing write the words and a running rank. */
static uint64_t bb(orc_set *C, uint8_t *cbuf, const orc_set *A, const orc_set *B) {
  uint64_t sa = bs_start(A), sb = bs_start(B), wa = bs_words(A), wb = bs_words(B);
  uint64_t start = sa > sb ? sa : sb;
  uint64_t end = (sa + wa) > (sb + wb) ? (sb + wb) : (sa + wa);
  uint64_t total = start > end ? 0 : end - start;
  const uint64_t *pa = bs_w(A) + (start - sa), *pb = bs_w(B) + (start - sb);
  uint64_t count = 0;
  if (C) {
    uint64_t *hdr = (uint64_t *)cbuf;
    uint64_t *rw = (uint64_t *)(cbuf + 8);
    uint32_t *rr = (uint32_t *)(cbuf + 8 + 8 * total);
    hdr[0] = start;
    for (uint64_t i = 0; i < total; i++) {
      uint64_t r = pa[i] & pb[i];
      rr[i] = (uint32_t)count;
      count += (uint64_t)__builtin_popcountll(r);
      rw[i] = r;
    }
    C->cardinality = (uint32_t)count;
    C->number_of_bytes = count ? total * 12 + 8 : 0;
    C->range = count ? (uint32_t)(64 * (total + start)) : 0;
    C->type = ORC_RANGE_BITSET;
    C->data = cbuf;
  } else {
    for (uint64_t i = 0; i < total; i++) count += (uint64_t)__builtin_popcountll(pa[i] & pb[i]);
  }
  return count;
}

int orc_algorithm_of(const orc_set *a, const orc_set *b) {
  if (a->cardinality == 0 || b->cardinality == 0) return 4;
  if (a->type == ORC_RANGE_BITSET && b->type == ORC_RANGE_BITSET) return 3;
  if (a->type != b->type) return 2;
  uint64_t mn = a->cardinality < b->cardinality ? a->cardinality : b->cardinality;
  uint64_t mx = a->cardinality < b->cardinality ? b->cardinality : a->cardinality;
  return ((double)mx / (double)mn > 32.0) ? 1 : 0;
}

/* ops::set_intersect(C,A,B) dispatch (set.hpp:53-91): an empty side gives an
 * empty UINTEGER set; R x U is swapped to U x R; U x U -> U, U x R -> U,
 * R x R -> R.  No re-layout of results. */
uint64_t orc_intersect(orc_set *c, uint8_t *cbuf, const orc_set *a, const orc_set *b) {
  if (a->cardinality == 0 || b->cardinality == 0) {
    c->cardinality = 0; c->number_of_bytes = 0; c->type = ORC_UINTEGER; c->range = 0; c->data = cbuf;
    return 0;
  }
  if (a->type == ORC_RANGE_BITSET && b->type == ORC_RANGE_BITSET) return bb(c, cbuf, a, b);
  uint64_t n;
  if (a->type == ORC_UINTEGER && b->type == ORC_UINTEGER) n = uu((uint32_t *)cbuf, a, b);
  else if (a->type == ORC_UINTEGER) n = ub((uint32_t *)cbuf, a, b);
  else n = ub((uint32_t *)cbuf, b, a);
  c->cardinality = (uint32_t)n;
  c->range = n ? ((uint32_t *)cbuf)[n - 1] : 0; /* unpack_materialize::range */
  c->number_of_bytes = 4 * n;
  c->type = ORC_UINTEGER;
  c->data = cbuf;
  return n;
}

/* Count-only variant (set.hpp:131-167). */
uint64_t orc_intersect_count(const orc_set *a, const orc_set *b) {
  if (a->cardinality == 0 || b->cardinality == 0) return 0;
  if (a->type == ORC_RANGE_BITSET && b->type == ORC_RANGE_BITSET) return bb(NULL, NULL, a, b);
  if (a->type == ORC_UINTEGER && b->type == ORC_UINTEGER) return uu(NULL, a, b);
  if (a->type == ORC_UINTEGER) return ub(NULL, a, b);
  return ub(NULL, b, a);
}

/* ------------------------------------------------------------------------- */
/* Trie over a binary relation (Trie.cpp:380-510): lexicographic sort of the  */
/* rows (tbb::parallel_sort with SortColumns, :173-186, :402-404), distinct   */
/* head values (produce_ranges :192-224), de-duplicated child runs            */
/* (encode_tail :291-302), one hybrid set per node (build_block :229-262).    */
/* ------------------------------------------------------------------------- */
struct orc_trie {
  uint64_t rows;          /* distinct tuples */
  uint64_t nheads;
  orc_set head;
  uint8_t *head_bytes;
  uint32_t *head_vals;    /* decoded head values */
  orc_set *child;         /* per head node */
  uint8_t *pool;          /* child set bytes */
  int64_t *dense;         /* NextLevel[data] analog when the head is dense */
  uint64_t dense_len;
  uint32_t *flat0, *flat1;
};

static void radix_sort_u64(uint64_t *k, uint64_t n) {
  uint64_t *tmp = (uint64_t *)malloc(8 * (n ? n : 1));
  uint64_t maxk = 0;
  for (uint64_t i = 0; i < n; i++) if (k[i] > maxk) maxk = k[i];
  uint64_t *src = k, *dst = tmp;
  for (int shift = 0; shift < 64 && (maxk >> shift); shift += 16) {
    uint64_t *cnt = (uint64_t *)calloc(65536, 8);
    for (uint64_t i = 0; i < n; i++) cnt[(src[i] >> shift) & 0xFFFF]++;
    uint64_t s = 0;
    for (int d = 0; d < 65536; d++) { uint64_t c = cnt[d]; cnt[d] = s; s += c; }
    for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> shift) & 0xFFFF]++] = src[i];
    free(cnt);
    uint64_t *t = src; src = dst; dst = t;
  }
  if (src != k) memcpy(k, src, 8 * n);
  free(tmp);
}

static int policy_layout(int policy, const uint32_t *v, uint64_t n) {
  if (policy == ORC_POLICY_ALL_UINT) return ORC_UINTEGER;
  if (policy == ORC_POLICY_ALL_BITSET) return n ? ORC_RANGE_BITSET : ORC_UINTEGER;
  return orc_layout_of(v, n);
}

orc_trie *orc_trie_build(const uint32_t *c0, const uint32_t *c1, uint64_t rows, int policy) {
  orc_trie *t = (orc_trie *)calloc(1, sizeof(orc_trie));
  uint64_t *key = (uint64_t *)malloc(8 * (rows ? rows : 1));
  for (uint64_t i = 0; i < rows; i++) key[i] = ((uint64_t)c0[i] << 32) | c1[i];
  radix_sort_u64(key, rows);
  uint64_t m = 0;
  for (uint64_t i = 0; i < rows; i++) if (i == 0 || key[i] != key[i - 1]) key[m++] = key[i];
  t->rows = m;
  t->flat0 = (uint32_t *)malloc(4 * (m ? m : 1));
  t->flat1 = (uint32_t *)malloc(4 * (m ? m : 1));
  for (uint64_t i = 0; i < m; i++) { t->flat0[i] = (uint32_t)(key[i] >> 32); t->flat1[i] = (uint32_t)key[i]; }
  free(key);
  /* head values and child runs */
  uint64_t nh = 0;
  for (uint64_t i = 0; i < m; i++) if (i == 0 || t->flat0[i] != t->flat0[i - 1]) nh++;
  t->nheads = nh;
  t->head_vals = (uint32_t *)malloc(4 * (nh ? nh : 1));
  uint64_t *beg = (uint64_t *)malloc(8 * (nh + 1));
  nh = 0;
  for (uint64_t i = 0; i < m; i++)
    if (i == 0 || t->flat0[i] != t->flat0[i - 1]) { t->head_vals[nh] = t->flat0[i]; beg[nh++] = i; }
  beg[nh] = m;
  int hl = policy_layout(policy, t->head_vals, nh);
  t->head_bytes = (uint8_t *)malloc(orc_bytes_needed(nh, t->head_vals, hl) + 16);
  orc_set_build(t->head_bytes, t->head_vals, nh, hl, &t->head);
  /* child sets */
  uint64_t total = 0;
  int *lay = (int *)malloc(sizeof(int) * (nh ? nh : 1));
  for (uint64_t h = 0; h < nh; h++) {
    lay[h] = policy_layout(policy, t->flat1 + beg[h], beg[h + 1] - beg[h]);
    total += (orc_bytes_needed(beg[h + 1] - beg[h], t->flat1 + beg[h], lay[h]) + 7) & ~(uint64_t)7;
  }
  t->pool = (uint8_t *)malloc(total + 16);
  t->child = (orc_set *)calloc(nh ? nh : 1, sizeof(orc_set));
  uint64_t off = 0;
  for (uint64_t h = 0; h < nh; h++) {
    uint64_t b = orc_set_build(t->pool + off, t->flat1 + beg[h], beg[h + 1] - beg[h], lay[h], &t->child[h]);
    off += (b + 7) & ~(uint64_t)7;
  }
  free(lay);
  free(beg);
  /* Dense navigation when the head is dense by (card, range = max):
   * TrieBlock::nextSize/get_next_block (TrieBlock.hpp:26-35,86-111). */
  if (nh && !orc_is_sparse(t->head.cardinality, t->head.range)) {
    t->dense_len = (uint64_t)t->head.range + 1;
    t->dense = (int64_t *)malloc(8 * t->dense_len);
    for (uint64_t i = 0; i < t->dense_len; i++) t->dense[i] = -1;
    for (uint64_t h = 0; h < nh; h++) t->dense[t->head_vals[h]] = (int64_t)h;
  }
  return t;
}

void orc_trie_free(orc_trie *t) {
  if (!t) return;
  free(t->head_bytes); free(t->head_vals); free(t->child); free(t->pool);
  free(t->dense); free(t->flat0); free(t->flat1); free(t);
}

uint64_t orc_trie_num_rows(const orc_trie *t) { return t->rows; }
void orc_trie_head(const orc_trie *t, orc_set *out) { *out = t->head; }

static const orc_set *trie_child(const orc_trie *t, uint32_t key) {
  int64_t idx;
  if (t->dense) idx = (key < t->dense_len) ? t->dense[key] : -1;
  else idx = orc_set_find(&t->head, key);
  if (idx < 0) return NULL;
  return &t->child[idx];
}

int orc_trie_child(const orc_trie *t, uint32_t key, orc_set *out) {
  const orc_set *s = trie_child(t, key);
  if (!s) return 0;
  *out = *s;
  return 1;
}

/* Trie::foreach (Trie.cpp:130-160): tuples in lexicographic order. */
uint64_t orc_trie_tuples(const orc_trie *t, uint32_t *c0, uint32_t *c1) {
  uint64_t k = 0;
  uint32_t *buf = NULL;
  uint64_t cap = 0;
  for (uint64_t h = 0; h < t->nheads; h++) {
    const orc_set *s = &t->child[h];
    if (s->cardinality > cap) { cap = s->cardinality; buf = (uint32_t *)realloc(buf, 4 * cap); }
    uint64_t n = orc_set_decode(s, buf);
    for (uint64_t i = 0; i < n; i++) { c0[k] = t->head_vals[h]; c1[k] = buf[i]; k++; }
  }
  free(buf);
  return k;
}

void orc_trie_node_info(const orc_trie *t, uint8_t *types, uint32_t *cards) {
  for (uint64_t h = 0; h < t->nheads; h++) {
    if (types) types[h] = t->child[h].type;
    if (cards) cards[h] = t->child[h].cardinality;
  }
}

/* ------------------------------------------------------------------------- */
/* Generic join, depth-first per thread as the generated code runs it.        */
/* ------------------------------------------------------------------------- */
typedef struct {
  const orc_plan *p;
  orc_trie *const *tries;
  uint64_t maxbytes, maxcard;
  uint32_t *level0;      /* decoded level-0 candidate values */
  uint64_t n0, begin, end;
  uint64_t next;         /* dynamic scheduler cursor */
  uint64_t stride;       /* visit every stride-th chunk of 128 level-0 values */
  int64_t *per_head;     /* optional per level-0 value output */
} join_ctx;

typedef struct {
  join_ctx *jc;
  int index;             /* worker number (0 = the calling thread) */
  uint8_t *buf[ORC_MAX_ATTRS][2];
  uint32_t *vals[ORC_MAX_ATTRS];
  uint32_t tuple[ORC_MAX_ATTRS];
  int64_t sum;
} join_thread;

static int gather(const join_ctx *jc, int level, const uint32_t *tuple, const orc_set **sets) {
  for (int i = 0; i < jc->p->num_ops[level]; i++) {
    orc_op op = jc->p->ops[level][i];
    const orc_trie *t = jc->tries[op.rel];
    if (op.parent == ORC_PARENT_HEAD) sets[i] = &t->head;
    else {
      sets[i] = trie_child(t, op.parent == ORC_PARENT_CONST ? op.value : tuple[op.parent]);
      if (!sets[i]) return 0; /* NULL TrieBlock -> empty (TrieBuilder.cpp:278-304) */
    }
  }
  return 1;
}

/* TrieBuilder::build_aggregated_set for 1, 2 or N sets (TrieBuilder.cpp:
 * 223-348): N-way takes the minimum-cardinality set first (ties go to the later
 * set), then chains pairwise intersections through two alternating buffers. */
static uint64_t nway(orc_set *res, uint8_t *b0, uint8_t *b1, const orc_set **sets, int n, int count_only) {
  if (n == 1) { *res = *sets[0]; return sets[0]->cardinality; }
  if (n == 2) {
    if (count_only) return orc_intersect_count(sets[0], sets[1]);
    return orc_intersect(res, b0, sets[0], sets[1]);
  }
  int mi = 0;
  uint64_t mn = sets[0]->cardinality;
  for (int i = 1; i < n; i++) {
    if (sets[i]->cardinality <= mn) { mn = sets[i]->cardinality; mi = i; }
  }
  const orc_set *rest[ORC_MAX_OPS] = {0};
  int r = 0;
  for (int i = 0; i < n; i++) if (i != mi) rest[r++] = sets[i];
  orc_set cur, nxt;
  uint8_t *bufs[2] = {b0, b1};
  int which = 0;
  orc_intersect(&cur, bufs[which], sets[mi], rest[0]);
  for (int i = 1; i < r; i++) {
    if (count_only && i == r - 1) return orc_intersect_count(rest[i], &cur);
    which ^= 1;
    orc_intersect(&nxt, bufs[which], rest[i], &cur);
    cur = nxt;
  }
  *res = cur;
  return cur.cardinality;
}

static int64_t join_rec(join_thread *th, int level) {
  const join_ctx *jc = th->jc;
  const orc_plan *p = jc->p;
  const orc_set *sets[ORC_MAX_OPS];
  if (!gather(jc, level, th->tuple, sets)) return 0;
  orc_set r;
  if (level == p->num_attrs - 1) {
    /* COUNT: annotation += 1*...*count (EHGenerator.scala:436-480). */
    return (int64_t)nway(&r, th->buf[level][0], th->buf[level][1], sets, p->num_ops[level], 1);
  }
  uint64_t n = nway(&r, th->buf[level][0], th->buf[level][1], sets, p->num_ops[level], 0);
  if (n == 0) return 0;
  uint64_t k = orc_set_decode(&r, th->vals[level]);
  int64_t acc = 0;
  for (uint64_t i = 0; i < k; i++) {
    th->tuple[level] = th->vals[level][i];
    acc += join_rec(th, level + 1);
  }
  return acc;
}

/* Thread pinning for timing runs.  The reference's pool pins worker i to CPU
 * i (src/utils/thread_pool.cpp:44-49); here worker i runs on the i-th CPU of
 * the process's affinity mask (so a CPU share given by the host is honoured),
 * wrapping around when there are more workers than CPUs.  Off by default; the
 * calling thread's own mask is restored after the join. */
static int g_pin_threads = 0;
static cpu_set_t g_allowed;
static int g_allowed_n = 0;

void orc_set_pin_threads(int on) { g_pin_threads = on; }

static void pin_to(int index) {
  if (g_allowed_n <= 0) return;
  int k = index % g_allowed_n, seen = 0;
  for (int c = 0; c < CPU_SETSIZE; c++) {
    if (!CPU_ISSET(c, &g_allowed)) continue;
    if (seen++ == k) {
      cpu_set_t one;
      CPU_ZERO(&one);
      CPU_SET(c, &one);
      pthread_setaffinity_np(pthread_self(), sizeof(one), &one);
      return;
    }
  }
}

static void *join_worker(void *arg) {
  join_thread *th = (join_thread *)arg;
  join_ctx *jc = th->jc;
  if (g_pin_threads) pin_to(th->index);
  th->sum = 0;
  for (;;) {
    /* par::for_range dynamic chunks of 128 (parallel.cpp:65-79). */
    uint64_t s = __atomic_fetch_add(&jc->next, 128, __ATOMIC_RELAXED);
    if (s >= jc->end - jc->begin) break;
    if (jc->stride > 1 && (s / 128) % jc->stride != 0) continue;
    uint64_t e = s + 128;
    if (e > jc->end - jc->begin) e = jc->end - jc->begin;
    for (uint64_t i = jc->begin + s; i < jc->begin + e; i++) {
      th->tuple[0] = jc->level0[i];
      int64_t v = (jc->p->num_attrs == 1) ? 1 : join_rec(th, 1);
      if (jc->per_head) jc->per_head[i] = v;
      th->sum += v;
    }
  }
  return NULL;
}

static int plan_ok(const orc_plan *p) {
  if (p->num_attrs < 1 || p->num_attrs > ORC_MAX_ATTRS) return 0;
  for (int k = 0; k < p->num_attrs; k++) {
    if (p->num_ops[k] < 1 || p->num_ops[k] > ORC_MAX_OPS) return 0;
    for (int i = 0; i < p->num_ops[k]; i++) {
      if (p->ops[k][i].parent >= k || p->ops[k][i].parent < ORC_PARENT_CONST) return 0;
      if (k == 0 && p->ops[k][i].parent >= 0) return 0;
    }
  }
  return 1;
}

static uint64_t max_bytes(const orc_plan *p, orc_trie *const *tries, uint64_t *maxcard) {
  uint64_t mx = 64, mc = 1;
  for (int k = 0; k < p->num_attrs; k++)
    for (int i = 0; i < p->num_ops[k]; i++) {
      const orc_trie *t = tries[p->ops[k][i].rel];
      if (t->head.number_of_bytes > mx) mx = t->head.number_of_bytes;
      for (uint64_t h = 0; h < t->nheads; h++) {
        if (t->child[h].number_of_bytes > mx) mx = t->child[h].number_of_bytes;
        if (t->child[h].cardinality > mc) mc = t->child[h].cardinality;
      }
      if (4 * (uint64_t)t->head.cardinality > mx) mx = 4 * (uint64_t)t->head.cardinality;
      if (t->head.cardinality > mc) mc = t->head.cardinality;
    }
  if (maxcard) *maxcard = mc;
  return mx;
}

/* Level 0: ParTrieBuilder::build_aggregated_set over the head sets
 * (TrieBuilder.cpp:478-578), decoded for par_foreach_aggregate. */
static uint32_t *level0_values(const orc_plan *p, orc_trie *const *tries, uint64_t mx, uint64_t *n0) {
  const orc_set *sets[ORC_MAX_OPS];
  uint32_t dummy[1] = {0};
  if (!gather(&(join_ctx){.p = p, .tries = tries}, 0, dummy, sets)) {  /* a selection of an absent value */
    *n0 = 0;
    return (uint32_t *)malloc(4);
  }
  uint8_t *b0 = (uint8_t *)malloc(mx + 64), *b1 = (uint8_t *)malloc(mx + 64);
  orc_set r;
  nway(&r, b0, b1, sets, p->num_ops[0], 0);
  uint32_t *v = (uint32_t *)malloc(4 * ((uint64_t)r.cardinality + 1));
  *n0 = orc_set_decode(&r, v);
  free(b0); free(b1);
  return v;
}

uint64_t orc_join_level0_size(const orc_plan *p, orc_trie *const *tries) {
  if (!plan_ok(p)) return 0;
  uint64_t n0;
  uint32_t *v = level0_values(p, tries, max_bytes(p, tries, NULL), &n0);
  free(v);
  return n0;
}

static int64_t run_join(const orc_plan *p, orc_trie *const *tries, int nthreads,
                        uint64_t begin, uint64_t end, int64_t *per_head, uint64_t stride) {
  if (!plan_ok(p)) return -1;
  if (nthreads < 1) nthreads = 1;
  join_ctx jc;
  memset(&jc, 0, sizeof(jc));
  jc.p = p; jc.tries = tries; jc.maxbytes = max_bytes(p, tries, &jc.maxcard);
  jc.level0 = level0_values(p, tries, jc.maxbytes, &jc.n0);
  jc.begin = begin < jc.n0 ? begin : jc.n0;
  jc.end = end < jc.n0 ? end : jc.n0;
  if (jc.end < jc.begin) jc.end = jc.begin;
  jc.per_head = per_head;
  jc.stride = stride;
  join_thread *th = (join_thread *)calloc((size_t)nthreads, sizeof(join_thread));
  pthread_t *tid = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  cpu_set_t caller_mask;
  int restore = 0;
  if (g_pin_threads && pthread_getaffinity_np(pthread_self(), sizeof(caller_mask), &caller_mask) == 0) {
    g_allowed = caller_mask;
    g_allowed_n = CPU_COUNT(&g_allowed);
    restore = 1;
  }
  for (int t = 0; t < nthreads; t++) {
    th[t].jc = &jc;
    th[t].index = t;
    for (int k = 0; k < p->num_attrs; k++) {
      th[t].buf[k][0] = (uint8_t *)malloc(jc.maxbytes + 64);
      th[t].buf[k][1] = (uint8_t *)malloc(jc.maxbytes + 64);
      th[t].vals[k] = (uint32_t *)malloc(4 * jc.maxcard + 64);
    }
  }
  for (int t = 1; t < nthreads; t++) pthread_create(&tid[t], NULL, join_worker, &th[t]);
  join_worker(&th[0]);
  for (int t = 1; t < nthreads; t++) pthread_join(tid[t], NULL);
  if (restore) pthread_setaffinity_np(pthread_self(), sizeof(caller_mask), &caller_mask);
  /* par::reducer::evaluate folds the per-thread slots (parallel.hpp:28-33). */
  int64_t total = 0;
  for (int t = 0; t < nthreads; t++) {
    total += th[t].sum;
    for (int k = 0; k < p->num_attrs; k++) { free(th[t].buf[k][0]); free(th[t].buf[k][1]); free(th[t].vals[k]); }
  }
  free(th); free(tid); free(jc.level0);
  if (p->num_attrs == 1) return (int64_t)(jc.end - jc.begin);
  return total;
}

int64_t orc_join_count(const orc_plan *p, orc_trie *const *tries, int nthreads, uint64_t begin, uint64_t end) {
  return run_join(p, tries, nthreads, begin, end, NULL, 1);
}

int64_t orc_join_count_sampled(const orc_plan *p, orc_trie *const *tries, int nthreads, uint64_t stride) {
  return run_join(p, tries, nthreads, 0, UINT64_MAX, NULL, stride);
}

int orc_join_count_per_head(const orc_plan *p, orc_trie *const *tries, int nthreads, int64_t *out) {
  return run_join(p, tries, nthreads, 0, UINT64_MAX, out, 1) < 0 ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* Materialized result: the generated code's build_set / set_level loop       */
/* (TrieBuilder.cpp:44-221,376-446) run depth-first on one thread, every      */
/* level decoded; tuples come out in lexicographic order, the order           */
/* Trie::foreach reads an output trie (Trie.cpp:78-160).                      */
/* ------------------------------------------------------------------------- */
typedef struct {
  join_thread th;
  uint32_t *cols;  /* num_attrs columns of cap entries */
  uint64_t cap, rows;
} mat_state;

static void mat_rec(mat_state *ms, int level) {
  join_thread *th = &ms->th;
  const join_ctx *jc = th->jc;
  const orc_plan *p = jc->p;
  const orc_set *sets[ORC_MAX_OPS];
  if (!gather(jc, level, th->tuple, sets)) return;
  orc_set r;
  if (nway(&r, th->buf[level][0], th->buf[level][1], sets, p->num_ops[level], 0) == 0) return;
  uint64_t k = orc_set_decode(&r, th->vals[level]);
  for (uint64_t i = 0; i < k; i++) {
    th->tuple[level] = th->vals[level][i];
    if (level == p->num_attrs - 1) {
      if (ms->rows < ms->cap)
        for (int j = 0; j < p->num_attrs; j++) ms->cols[(uint64_t)j * ms->cap + ms->rows] = th->tuple[j];
      ms->rows++;
    } else {
      mat_rec(ms, level + 1);
    }
  }
}

uint64_t orc_join_materialize(const orc_plan *p, orc_trie *const *tries, uint32_t *cols, uint64_t cap) {
  if (!plan_ok(p)) return 0;
  join_ctx jc;
  memset(&jc, 0, sizeof(jc));
  jc.p = p; jc.tries = tries; jc.maxbytes = max_bytes(p, tries, &jc.maxcard);
  jc.level0 = level0_values(p, tries, jc.maxbytes, &jc.n0);
  mat_state ms;
  memset(&ms, 0, sizeof(ms));
  ms.th.jc = &jc;
  ms.cols = cols;
  ms.cap = cap;
  for (int k = 0; k < p->num_attrs; k++) {
    ms.th.buf[k][0] = (uint8_t *)malloc(jc.maxbytes + 64);
    ms.th.buf[k][1] = (uint8_t *)malloc(jc.maxbytes + 64);
    ms.th.vals[k] = (uint32_t *)malloc(4 * jc.maxcard + 64);
  }
  for (uint64_t i = 0; i < jc.n0; i++) {
    ms.th.tuple[0] = jc.level0[i];
    if (p->num_attrs == 1) {
      if (ms.rows < cap) cols[ms.rows] = jc.level0[i];
      ms.rows++;
    } else {
      mat_rec(&ms, 1);
    }
  }
  for (int k = 0; k < p->num_attrs; k++) { free(ms.th.buf[k][0]); free(ms.th.buf[k][1]); free(ms.th.vals[k]); }
  free(jc.level0);
  return ms.rows;
}

/* ------------------------------------------------------------------------- */
/* 4-cycles of an undirected graph (no counterpart in the reference engine,  */
/* whose queries have no 4-cycle; the count is the graph-theoretic one).      */
/* Symmetric CSR: off[n+1], adj sorted per vertex.  Each cycle is charged to  */
/* its largest vertex a: for every wedge a-b-c with b < a, c < a, H[c] += 1,  */
/* then cycles(a) = sum_c H[c](H[c]-1)/2.                                     */
/* ------------------------------------------------------------------------- */
typedef struct {
  const uint64_t *off;
  const uint32_t *adj;
  uint32_t n;
  uint64_t next;
  int64_t *per_vertex;
} c4_ctx;

typedef struct {
  c4_ctx *c;
  int64_t sum;
} c4_thread;

static void *c4_worker(void *arg) {
  c4_thread *th = (c4_thread *)arg;
  c4_ctx *c = th->c;
  uint32_t *H = (uint32_t *)calloc((size_t)c->n + 1, sizeof(uint32_t));
  uint32_t *touched = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)c->n + 1));
  th->sum = 0;
  for (;;) {
    uint64_t s = __atomic_fetch_add(&c->next, 64, __ATOMIC_RELAXED);
    if (s >= c->n) break;
    uint64_t e = s + 64 < c->n ? s + 64 : c->n;
    for (uint64_t a = s; a < e; a++) {
      uint64_t nt = 0;
      int64_t cyc = 0;
      for (uint64_t i = c->off[a]; i < c->off[a + 1] && c->adj[i] < a; i++) {
        const uint32_t b = c->adj[i];
        for (uint64_t j = c->off[b]; j < c->off[b + 1] && c->adj[j] < a; j++) {
          const uint32_t x = c->adj[j];
          if (H[x]++ == 0) touched[nt++] = x;
        }
      }
      for (uint64_t t = 0; t < nt; t++) {
        const uint64_t h = H[touched[t]];
        cyc += (int64_t)(h * (h - 1) / 2);
        H[touched[t]] = 0;
      }
      if (c->per_vertex) c->per_vertex[a] = cyc;
      th->sum += cyc;
    }
  }
  free(H);
  free(touched);
  return NULL;
}

int64_t orc_cycle4_count(const uint64_t *off, const uint32_t *adj, uint32_t n, int nthreads, int64_t *per_vertex) {
  if (nthreads < 1) nthreads = 1;
  c4_ctx c = {off, adj, n, 0, per_vertex};
  c4_thread *th = (c4_thread *)calloc((size_t)nthreads, sizeof(c4_thread));
  pthread_t *tid = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int t = 0; t < nthreads; t++) th[t].c = &c;
  for (int t = 1; t < nthreads; t++) pthread_create(&tid[t], NULL, c4_worker, &th[t]);
  c4_worker(&th[0]);
  int64_t total = 0;
  for (int t = 1; t < nthreads; t++) pthread_join(tid[t], NULL);
  for (int t = 0; t < nthreads; t++) total += th[t].sum;
  free(th);
  free(tid);
  return total;
}


/* ---- recursive annotated SUM (PageRank) ------------------------------------ */
typedef struct {
  const orc_trie *t;
  const orc_set *w;
  const float *w_vals, *cur;
  float *nxt;
  double add, mul, init;
  int fused;
  uint64_t next;  /* dynamic chunks of 128 heads (src/utils/parallel.cpp:65-79) */
} rank_ctx;

static void *rank_worker(void *arg) {
  rank_ctx *c = (rank_ctx *)arg;
  const orc_trie *t = c->t;
  uint64_t cap = 0, zcap = 0;
  uint8_t *b1 = NULL, *b2 = NULL;
  uint32_t *zs = NULL;
  for (;;) {
    const uint64_t h0 = __atomic_fetch_add(&c->next, 128, __ATOMIC_RELAXED);
    if (h0 >= t->nheads) break;
    const uint64_t h1 = h0 + 128 < t->nheads ? h0 + 128 : t->nheads;
    for (uint64_t h = h0; h < h1; h++) {
      const orc_set *s = &t->child[h];
      const uint64_t need = s->number_of_bytes + t->head.number_of_bytes + c->w->number_of_bytes + 16;
      if (need > cap) {
        cap = need;
        b1 = (uint8_t *)realloc(b1, cap);
        b2 = (uint8_t *)realloc(b2, cap);
      }
      if (s->cardinality + 64 > zcap) {  /* the z set is a subset of child(x) */
        zcap = s->cardinality + 64;
        zs = (uint32_t *)realloc(zs, 4 * zcap);
      }
      /* z set: child(x) ∩ head(R) ∩ W, smallest operand first (TrieBuilder.cpp:276-348) */
      orc_set r1, r2;
      const orc_set *ops[3] = {s, &t->head, c->w};
      int mi = 0;
      for (int i = 1; i < 3; i++)
        if (ops[i]->cardinality < ops[mi]->cardinality) mi = i;
      const orc_set *first = ops[mi], *rest[2];
      for (int i = 0, k = 0; i < 3; i++)
        if (i != mi) rest[k++] = ops[i];
      orc_intersect(&r1, b1, first, rest[0]);
      orc_intersect(&r2, b2, &r1, rest[1]);
      const uint64_t nz = orc_set_decode(&r2, zs);
      float acc = 0.0f;
      for (uint64_t k = 0; k < nz; k++) {
        const uint32_t z = zs[k];
        const double rz = (double)c->cur[orc_set_find(&t->head, z)];
        const double wz = (double)c->w_vals[orc_set_find(c->w, z)];
        const float term = (float)(((double)1.0f * c->init) * rz * wz);
        acc = acc + term;
      }
      /* g++ -O3 -march=native (cython/query/setup.py:10-20) fuses add + mul*acc
         into one double FMA (C++'s default -ffp-contract=fast); unfused on request */
      c->nxt[h] = c->fused ? (float)fma(c->mul, (double)acc, c->add) : (float)(c->add + c->mul * (double)acc);
    }
  }
  free(b1);
  free(b2);
  free(zs);
  return NULL;
}

int orc_pagerank(const orc_trie *edge, const uint32_t *w_keys, const float *w_vals, uint64_t n_w, double base,
                 double add, double mul, double init, int iterations, int nthreads, int fused_finish, float *out) {
  if (nthreads < 1) nthreads = 1;
  for (uint64_t i = 1; i < n_w; i++)
    if (w_keys[i] <= w_keys[i - 1]) return -1;
  const uint64_t nh = edge->nheads;
  orc_set w;
  const int wl = orc_layout_of(w_keys, n_w);
  uint8_t *wbytes = (uint8_t *)malloc(orc_bytes_needed(n_w, w_keys, wl) + 16);
  orc_set_build(wbytes, w_keys, n_w, wl, &w);
  float *cur = (float *)malloc(4 * (nh ? nh : 1)), *nxt = (float *)malloc(4 * (nh ? nh : 1));
  for (uint64_t h = 0; h < nh; h++) cur[h] = (float)base;
  pthread_t *tid = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int it = 0; it < iterations; it++) {
    rank_ctx c = {edge, &w, w_vals, cur, nxt, add, mul, init, fused_finish, 0};
    for (int k = 1; k < nthreads; k++) pthread_create(&tid[k], NULL, rank_worker, &c);
    rank_worker(&c);
    for (int k = 1; k < nthreads; k++) pthread_join(tid[k], NULL);
    float *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  memcpy(out, cur, 4 * nh);
  free(tid);
  free(cur);
  free(nxt);
  free(wbytes);
  return 0;
}


/* ---- the reference's on-disk relation ------------------------------------- */
typedef struct {
  uint8_t *p;
  uint64_t n, cap;
} bytebuf;

static uint64_t bb_put(bytebuf *b, const void *src, uint64_t n) {
  if (b->n + n > b->cap) {
    b->cap = (b->n + n) * 2 + 4096;
    b->p = (uint8_t *)realloc(b->p, b->cap);
  }
  const uint64_t at = b->n;
  if (src) memcpy(b->p + at, src, n);
  else memset(b->p + at, 0, n);
  b->n += n;
  return at;
}

/* TrieBlock (1 byte) + Set header (24 bytes) + the set's bytes; returns the
 * block's offset and the offset just past the set. */
static uint64_t put_block(bytebuf *b, const orc_set *s, uint64_t *tail) {
  uint8_t hdr[25];
  memset(hdr, 0, sizeof(hdr));
  hdr[0] = (uint8_t)orc_is_sparse(s->cardinality, s->range);
  memcpy(hdr + 1, &s->cardinality, 4);
  memcpy(hdr + 5, &s->range, 4);
  memcpy(hdr + 9, &s->number_of_bytes, 8);
  hdr[17] = s->type;
  const uint64_t at = bb_put(b, hdr, 25);
  bb_put(b, s->data, s->number_of_bytes);
  *tail = b->n;
  return at;
}

static int write_buffer(const char *path, const bytebuf *b, int mmapped) {
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  if (!mmapped) fprintf(f, "%llu %llu ", (unsigned long long)b->n, (unsigned long long)b->n);
  if (b->n && fwrite(b->p, 1, b->n, f) != b->n) { fclose(f); return -1; }
  fclose(f);
  return 0;
}

int orc_write_refdb_annotated(const orc_trie *t, const char *dir, int nbuffers, int mmapped, int ncols,
                              const uint8_t *annos, int anno_bytes);

/* A[nextSize] after a set (TrieBlock::set_annotation, slot by position when
   sparse, by value when dense), from the rows' annotation bytes. */
static void put_annos(bytebuf *b, const orc_set *s, const uint8_t *annos, uint64_t row0, int anno_bytes) {
  const int sp = orc_is_sparse(s->cardinality, s->range);
  const uint64_t next_size = s->cardinality == 0 ? 0 : (sp ? s->cardinality : (uint64_t)s->range + 1);
  const uint64_t at = bb_put(b, NULL, (uint64_t)anno_bytes * next_size);
  uint32_t *v = (uint32_t *)malloc(4 * ((uint64_t)s->cardinality + 1));
  const uint64_t n = orc_set_decode(s, v);
  for (uint64_t i = 0; i < n; i++)
    memcpy(b->p + at + (uint64_t)anno_bytes * (sp ? i : v[i]), annos + (row0 + i) * (uint64_t)anno_bytes,
           (size_t)anno_bytes);
  free(v);
}

int orc_write_refdb(const orc_trie *t, const char *dir, int nbuffers, int mmapped) {
  return orc_write_refdb_annotated(t, dir, nbuffers, mmapped, 2, NULL, 8);
}

int orc_write_refdb_annotated(const orc_trie *t, const char *dir, int nbuffers, int mmapped, int ncols,
                              const uint8_t *annos, int anno_bytes) {
  if (nbuffers < 1 || (ncols != 1 && ncols != 2) || anno_bytes < 1 || anno_bytes > 8) return -1;
  char path[4096];
  snprintf(path, sizeof(path), "%s", dir);
  for (char *p = path + 1; *p; p++)  /* every missing parent of dir, then dir */
    if (*p == '/') {
      *p = 0;
      mkdir(path, 0777);
      *p = '/';
    }
  mkdir(path, 0777);
  snprintf(path, sizeof(path), "%s/%s", dir, mmapped ? "mmap" : "ram");
  if (mkdir(path, 0777) != 0 && errno != EEXIST) return -1;
  bytebuf head = {0}, *bufs = (bytebuf *)calloc((size_t)nbuffers, sizeof(bytebuf));
  uint64_t tail;
  put_block(&head, &t->head, &tail);
  const int sp = orc_is_sparse(t->head.cardinality, t->head.range);
  const uint64_t next_size = t->head.cardinality == 0 ? 0 : (sp ? t->head.cardinality : (uint64_t)t->head.range + 1);
  if (ncols == 1) {  /* a unary relation: the head set, then its annotations */
    if (annos) put_annos(&head, &t->head, annos, 0, anno_bytes);
  } else {
    const uint64_t next_at = bb_put(&head, NULL, 16 * next_size);
    for (uint64_t e = 0; e < next_size; e++) {  /* slots of absent values: no child */
      const int32_t none = -1;
      memcpy(head.p + next_at + 16 * e, &none, 4);
    }
    uint64_t row = 0;
    for (uint64_t h = 0; h < t->nheads; h++) {
      const int b = (int)((h / 128) % (uint64_t)nbuffers);
      uint64_t ctail;
      const uint64_t off = put_block(&bufs[b], &t->child[h], &ctail);
      if (annos) put_annos(&bufs[b], &t->child[h], annos, row, anno_bytes);
      row += t->child[h].cardinality;
      const uint64_t slot = sp ? h : t->head_vals[h];
      const int32_t idx = b;
      memcpy(head.p + next_at + 16 * slot, &idx, 4);
      memcpy(head.p + next_at + 16 * slot + 8, &off, 8);
    }
  }
  /* trieinfo.bin: annotated, annotation (void*), num_rows, num_columns,
     num_buffers, head size, per-buffer sizes (Trie.cpp:15-35) */
  snprintf(path, sizeof(path), "%s/%s/trieinfo.bin", dir, mmapped ? "mmap" : "ram");
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  const uint8_t annotated = annos != NULL;
  const uint64_t annotation = 0, rows = ncols == 1 ? t->nheads : t->rows, nc = (uint64_t)ncols,
                 nb = (uint64_t)nbuffers;
  fwrite(&annotated, 1, 1, f);
  fwrite(&annotation, (size_t)anno_bytes, 1, f);
  fwrite(&rows, 8, 1, f);
  fwrite(&nc, 8, 1, f);
  fwrite(&nb, 8, 1, f);
  fwrite(&head.n, 8, 1, f);
  for (int b = 0; b < nbuffers; b++) fwrite(&bufs[b].n, 8, 1, f);
  fclose(f);
  int rc = 0;
  snprintf(path, sizeof(path), "%s/%s/data_head.bin", dir, mmapped ? "mmap" : "ram");
  rc |= write_buffer(path, &head, mmapped);
  for (int b = 0; b < nbuffers; b++) {
    snprintf(path, sizeof(path), "%s/%s/data_%d.bin", dir, mmapped ? "mmap" : "ram", b);
    rc |= write_buffer(path, &bufs[b], mmapped);
    free(bufs[b].p);
  }
  free(bufs);
  free(head.p);
  return rc;
}
