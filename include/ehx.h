/*
 * ehx.h -- C ABI of the MI355X generic-join engine (libehx.so).
 *
 * This is the drop-in boundary beneath EmptyHeaded's templated storage-engine
 * API: plain pointers and sizes, integer status codes, no C++ or torch types.
 * The C++ mirror of the reference API (emptyheaded_amd/engine/include) and the
 * Python binding (emptyheaded_amd/__init__.py) both sit on top of it.
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference tree).
 *
 * Conventions: every function returns 0 (EHX_OK) on success or a negative
 * EHX_ERR_* code; ehx_last_error() returns a message for the calling thread.
 * Device memory is owned by ehx objects and released by the matching
 * ehx_*_destroy call.  Results are exact integers.
 */
#ifndef EHX_H
#define EHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EHX_OK 0
#define EHX_ERR_INVALID (-1)     /* bad argument or plan */
#define EHX_ERR_DEVICE (-2)      /* HIP runtime error or no GPU */
#define EHX_ERR_NOMEM (-3)       /* allocation failed */
#define EHX_ERR_UNSUPPORTED (-4) /* shape outside documented limits */

/* Set layouts: type::layout values, storage_engine/include/utils/common.hpp:77-85. */
#define EHX_RANGE_BITSET 0
#define EHX_UINTEGER 1

/* Layout policies for trie build.  RULE is the reference density rule
 * (common.hpp:56-67, hybrid.hpp:103-111); ALL_UINT / ALL_BITSET force a layout
 * (layout never changes results); LEARNED uses the selector loaded with
 * ehx_selector_load and falls back to RULE when none is loaded. */
#define EHX_POLICY_RULE 0
#define EHX_POLICY_ALL_UINT 1
#define EHX_POLICY_ALL_BITSET 2
#define EHX_POLICY_LEARNED 3

/* Executor forms. */
#define EHX_FORM_PLAIN 0   /* one wavefront per prefix tuple and level */
#define EHX_FORM_TUNED 1   /* degree-binned persistent last level, LDS-staged */
#define EHX_FORM_LEARNED 2 /* tuned, with per-prefix staging chosen by the selector */

#define EHX_MAX_ATTRS 8
#define EHX_MAX_OPS 8

typedef struct ehx_ctx ehx_ctx;
typedef struct ehx_trie ehx_trie;

/* One operand of a join level: parent = EHX_PARENT_HEAD is the head set of
 * relation `rel` (ParTrieIterator::head / TrieIterator::get_block(0)); parent =
 * j >= 0 is the child set of attribute x_j in `rel` (TrieIterator::
 * get_next_block(0, x_j) then get_block(1), TrieIterator.cpp:23-42); parent =
 * EHX_PARENT_CONST is the child set of the constant `value` in `rel`: an
 * equality selection such as Edge(a,x), x=0 over the trie ordered (x, a) (the
 * generated selection path, query_compiler/src/main/scala/codegen/
 * EHGenerator.scala:657-701; empty when `value` is not a head).  `value` is
 * ignored for the other parents. */
#define EHX_PARENT_HEAD (-1)
#define EHX_PARENT_CONST (-2)
typedef struct {
  int32_t rel;
  int32_t parent;
  uint32_t value;
} ehx_op;

/* A single-bag COUNT(*) plan over binary relations, in attribute order: the
 * information the code generator emits as nested build_aggregated_set calls
 * (query_compiler/.../EHGenerator.scala:273-838). */
typedef struct {
  int32_t num_attrs;
  int32_t num_ops[EHX_MAX_ATTRS];
  ehx_op ops[EHX_MAX_ATTRS][EHX_MAX_OPS];
} ehx_plan;

typedef struct {
  uint64_t num_rows;      /* distinct tuples (Trie::num_rows) */
  uint64_t num_heads;     /* distinct first-column values */
  uint64_t num_bitsets;   /* child sets stored as range_bitset */
  uint64_t bitset_words;  /* 64-bit words in the bitset pool */
  uint64_t bitset_elems;  /* elements held by bitset child sets */
  uint32_t max_card;      /* largest child set */
  uint32_t head_layout;   /* EHX_UINTEGER or EHX_RANGE_BITSET */
  uint32_t dense_nav;     /* 1 if child lookup is a dense array (NextLevel[data]) */
  uint32_t max_value;     /* largest value in either column */
  double build_ms;        /* device time of the build */
  uint64_t num_staged_bitmaps; /* child sets the tuned level stages as a bitmap: the learned
                                  selector's choice (EHX_POLICY_LEARNED), else num_bitsets */
} ehx_trie_meta;

typedef struct {
  double level_ms[EHX_MAX_ATTRS];        /* device time per level */
  uint64_t frontier[EHX_MAX_ATTRS];      /* tuples materialized per level */
  uint64_t pairs_by_alg[5];              /* last level: merge, gallop, probe, and, empty */
  double total_ms;                       /* device time of the whole query */
  double main_kernel_ms;                 /* device time of the last level's main launch */
  int32_t form_used;                     /* form that ran the last level */
  int32_t pad;
} ehx_join_stats;

/* ---- library / context -------------------------------------------------- */
int ehx_version(void);
const char *ehx_last_error(void);
/* Number of visible GPUs (0 on a machine without one). */
int ehx_device_count(int *count);
/* A context owns a HIP stream and grow-only scratch on one GPU (and, once it
 * has uploaded a large relation, pinned host staging).  Calls on one context
 * come from one host thread at a time; threads that work concurrently each
 * take their own context.  Replaces the reference's
 * thread_pool::initializeThreadPool (src/utils/thread_pool.cpp:116-130) as the
 * per-process execution resource. */
int ehx_ctx_create(int device, ehx_ctx **out);
int ehx_ctx_destroy(ehx_ctx *ctx);
int ehx_ctx_synchronize(ehx_ctx *ctx);

/* ---- tries --------------------------------------------------------------- */
/* Build a two-level trie from a binary relation (columns c0, c1; rows may
 * repeat, duplicates are removed).  Replaces Trie<A,M>::Trie(path, max_set_sizes,
 * attr_in, annotation) (src/Trie.cpp:380-510): device radix sort of the rows,
 * unique, segmented scans into per-level value/offset arrays, per-set layout by
 * `policy`, range_bitset pool.  c0/c1 are host pointers, or device pointers when
 * on_device != 0. */
int ehx_trie_build(ehx_ctx *ctx, const uint32_t *c0, const uint32_t *c1, uint64_t rows,
                   int policy, int on_device, ehx_trie **out);
int ehx_trie_destroy(ehx_trie *t);
int ehx_trie_info(const ehx_trie *t, ehx_trie_meta *out);
/* Trie<A,M>::foreach (src/Trie.cpp:130-160): the distinct tuples in
 * lexicographic order, copied to host arrays of num_rows entries. */
int ehx_trie_export(ehx_ctx *ctx, const ehx_trie *t, uint32_t *c0, uint32_t *c1);
/* Per head node: layout of its child set and its cardinality (Set::type and
 * Set::cardinality, storage_engine/include/trie/set/Set.hpp:19-24). */
int ehx_trie_node_info(ehx_ctx *ctx, const ehx_trie *t, uint8_t *types, uint32_t *cards);
/* Child pointers from trie src into trie dst: for every value of src's child
 * sets, its child set in dst (position, length and gap-coded chunks), built
 * once and cached on src (16 bytes per row of src) until src is destroyed.
 * The reference keeps a child pointer beside every set element inside one
 * trie (NextLevel, storage_engine/include/trie/NextLevel.hpp:4-7, filled by
 * TrieBlock::set_next_block, trie/TrieBlock.hpp:55-74); here the child may
 * live in another relation's trie (Edge_1_0 -> Edge_0_1 for a plan that
 * reaches a vertex through one ordering and reads its set in the other).  The
 * tuned executor builds the link it needs on first use; calling this first
 * moves that cost out of the first query.  *linked (may be NULL) = 1 when the
 * link exists, 0 when dst has no value-indexed child table or the device had
 * no room (queries then look children up by value). */
int ehx_trie_link(ehx_ctx *ctx, ehx_trie *src, const ehx_trie *dst, int *linked);
/* Diagnostic, no counterpart in the reference (whose child pointers need no
 * hints): builds the link as ehx_trie_link does and copies its trim hints to
 * hints_out[n], n = src's row count: hint[e] = how many values of row e's child
 * set in dst lie below row e's first-column value (zeros when the link keeps
 * none).  route, for a link not built yet: 0 as ehx_trie_link chooses, 1 one
 * search per row, 2 one sort of src by value whenever every sampled row finds
 * its first-column value in its child set (dst looks like src's reverse
 * ordering; each hint is checked against dst and searched where that fails).
 * *state: 0 no link, 1 no hints kept (all zero), 2 searched hints, 3 hints
 * from the sort. */
int ehx_trie_link_hints(ehx_ctx *ctx, ehx_trie *src, const ehx_trie *dst, int32_t route, uint32_t *hints_out,
                        uint64_t n, int *state);
/* Child lookup for a batch of keys (TrieBlock::get_next_block(data),
 * storage_engine/include/trie/TrieBlock.hpp:86-111): node index or -1. */
int ehx_trie_lookup(ehx_ctx *ctx, const ehx_trie *t, const uint32_t *keys, uint64_t n,
                    int64_t *nodes);
/* Raw device views (for C++ callers that launch their own work). */
int ehx_trie_device_arrays(const ehx_trie *t, const uint32_t **head_vals, const uint32_t **offsets,
                           const uint32_t **values);

/* ---- set intersections ---------------------------------------------------- */
/* |child(ta, keys_a[i]) ∩ child(tb, keys_b[i])| for i < n: the count-only
 * ops::set_intersect (storage_engine/include/trie/set/ops/intersection/set.hpp:
 * 131-167) with the reference's per-pair algorithm (merge, gallop when
 * max/min > 32, uint x bitset probe, bitset AND+popcount).  Host arrays. */
int ehx_intersect_count_batch(ehx_ctx *ctx, const ehx_trie *ta, const uint32_t *keys_a,
                              const ehx_trie *tb, const uint32_t *keys_b, uint64_t n,
                              uint64_t *counts);
/* The same counts without allocating: no reference counterpart.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): no hipMalloc, hipFree or
 * allocating library call; the staged keys and the counts use a fixed device
 * array of the library's, one per GPU, held until the process ends.  At most
 * EHX_NOALLOC_MAX_PAIRS pairs per call (EHX_ERR_INVALID beyond).  Calls that
 * overlap on one GPU (from several threads or contexts) share that array and
 * race; serialize them.  The context and tries are made beforehand. */
#define EHX_NOALLOC_MAX_PAIRS (1ull << 20)
int ehx_intersect_count_batch_noalloc(ehx_ctx *ctx, const ehx_trie *ta, const uint32_t *keys_a,
                                      const ehx_trie *tb, const uint32_t *keys_b, uint64_t n,
                                      uint64_t *counts);
/* Materializing batch: the sorted intersections, concatenated into `values`
 * (capacity `cap`), with offsets[n+1]; ops::set_intersect(C,A,B) (set.hpp:53-91).
 * *total receives the number of values even when it exceeds cap (then nothing
 * is written and EHX_ERR_NOMEM is returned). */
int ehx_intersect_batch(ehx_ctx *ctx, const ehx_trie *ta, const uint32_t *keys_a,
                        const ehx_trie *tb, const uint32_t *keys_b, uint64_t n,
                        uint64_t *offsets, uint32_t *values, uint64_t cap, uint64_t *total);

/* N-way intersections of trie node sets for a batch of n calls:
 * TrieBuilder::build_aggregated_set over 1, 2 or N TrieBlocks
 * (src/TrieBuilder.cpp:223-348).  Operand s of call i is node nodes[s][i] of
 * tries[s]; node == num_heads names the trie's head set.  Host arrays.  The
 * count-only form writes counts[n]; the materializing form writes the sorted
 * results concatenated into values (capacity cap) with offsets[n+1]. */
int ehx_intersect_nodes_count(ehx_ctx *ctx, int nsets, ehx_trie *const *tries, const uint32_t *const *nodes,
                              uint64_t n, uint64_t *counts);
int ehx_intersect_nodes(ehx_ctx *ctx, int nsets, ehx_trie *const *tries, const uint32_t *const *nodes, uint64_t n,
                        uint64_t *offsets, uint32_t *values, uint64_t cap, uint64_t *total);
/* Host copies of a trie's navigation arrays: head values (num_heads entries)
 * and child offsets (num_heads + 1 entries), for host-side iterators. */
int ehx_trie_host_arrays(ehx_ctx *ctx, const ehx_trie *t, uint32_t *head_vals, uint32_t *offsets);

/* ---- generic join --------------------------------------------------------- */
/* COUNT(*) of a single-bag plan: what run_<id> computes for an aggregated bag
 * (EHGenerator.scala:134-153 with ParTrieBuilder/TrieBuilder::build_aggregated_set,
 * src/TrieBuilder.cpp:242-348,478-578, folded by par::reducer,
 * storage_engine/include/utils/parallel.hpp:8-34).  tries[r] backs relation r.
 * stats may be NULL. */
int ehx_join_count(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int form,
                   int64_t *result, ehx_join_stats *stats);
/* Same, restricted to level-0 values with index in [begin, end). */
int ehx_join_count_range(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int form,
                         uint64_t begin, uint64_t end, int64_t *result, ehx_join_stats *stats);
/* Many independent queries in one launch.  The queries' relations are packed
 * into one trie with disjoint first-attribute value ranges; counts[s] receives
 * the COUNT(*) of result tuples whose level-0 value lies in
 * [bounds[s], bounds[s+1]).  It equals what a separate run of the query over
 * segment s's tuples returns (one c_run_<id> call per query,
 * query_compiler/src/main/scala/codegen/query/QueryPlan.scala:25-39, each folded
 * by par::reducer, storage_engine/include/utils/parallel.hpp:8-34).  bounds:
 * nseg + 1 ascending host values; counts: nseg host entries; *result: the sum
 * over all level-0 values. */
int ehx_join_count_segments(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int form,
                            const uint32_t *bounds, uint64_t nseg, int64_t *counts, int64_t *result,
                            ehx_join_stats *stats);

/* Materialized result: every tuple of the join, in lexicographic order (the
 * output trie that TrieBuilder::build_set / set_level / allocate_next build,
 * src/TrieBuilder.cpp:44-221,376-446, read in Trie::foreach order,
 * src/Trie.cpp:78-160).  cols[j] (host, capacity cap) receives attribute j;
 * *rows the number of tuples.  When rows > cap nothing is copied and
 * EHX_ERR_NOMEM is returned with *rows set. */
int ehx_join_materialize(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int form,
                         uint32_t *const *cols, uint64_t cap, uint64_t *rows, ehx_join_stats *stats);

/* ---- two-bag plans ----------------------------------------------------------
 * Top bag of a plan whose lower bags were run grouped by one attribute: the sum
 * over the tuples (a, x) of trie t of wa[a] * wb[x] (host arrays indexed by
 * value; values >= na / nb weigh 0), exact modulo 2^64.  Replaces the
 * generated annotated fold over the intermediate bag tries
 * (query_compiler/src/main/scala/codegen/query/EHGenerator.scala:399-434,
 * 711-726) for LollipopAgg / BarbellAgg-shaped plans. */
int ehx_weighted_pairs(ehx_ctx *ctx, const ehx_trie *t, const int64_t *wa, uint64_t na, const int64_t *wb,
                       uint64_t nb, int64_t *result);

/* Orienting a symmetric relation.  counts[4] (may be NULL) receives the
 * numbers of rows (x, y) of trie t with x < y, with x > y, with x == y, and of
 * rows x < y whose reverse (y, x) is absent from t; the relation is symmetric
 * without loops exactly when counts[0] == counts[1] and counts[2] ==
 * counts[3] == 0.  Then, when lower and reverse are both non-NULL, the rows
 * with x < y are built on the device with layout `policy` as a trie (*lower)
 * and its reverse ordering (*reverse, rows (y, x)); otherwise (or when there
 * are no rows) both stay NULL and only the counts are taken.  The two are
 * the relation's oriented halves: per-vertex triangle counts of the symmetric
 * triangle bag follow from three oriented passes (the two-bag plans' lower
 * bag, EHGenerator.scala:711-726), see emptyheaded_amd/ghd.py. */
int ehx_trie_orient(ehx_ctx *ctx, const ehx_trie *t, int policy, ehx_trie **lower, ehx_trie **reverse,
                    uint64_t *counts);

/* ---- single-source closure --------------------------------------------------
 * Unweighted single-source shortest paths over trie t (x's neighbours = its
 * child set): dist_out[v] (nd host entries) = the BFS level at which v is first
 * reached from `start` (0 for start itself), 0xFFFFFFFF if never; at most
 * `iterations` levels (< 0: until the frontier empties); *levels = levels run.
 * Replaces tc::unweighted_single_source (storage_engine/include/
 * TransitiveClosure.hpp:12-123), which recursive queries such as SSSP
 * (test/graph/travis.py:556-567) run. */
int ehx_sssp_unweighted(ehx_ctx *ctx, const ehx_trie *t, uint32_t start, int32_t iterations, uint32_t *dist_out,
                        uint64_t nd, uint32_t *levels);
/* The same distances, direction-optimizing: `rev` is the trie of the reversed
 * relation (t itself for a symmetric one); while the frontier holds at least
 * nd / bottom_up vertices (bottom_up > 0) a level is found bottom-up -- each
 * unreached vertex looks among its in-neighbours for one in the frontier --
 * else top-down as ehx_sssp_unweighted; bottom_up = 0: top-down only.  BFS
 * levels are unique, so the result does not depend on the direction taken. */
int ehx_sssp_unweighted_dir(ehx_ctx *ctx, const ehx_trie *t, const ehx_trie *rev, uint32_t start, int32_t iterations,
                            uint32_t *dist_out, uint64_t nd, uint32_t *levels, int32_t bottom_up);

/* ---- the reference's on-disk relations --------------------------------------
 * A relation the reference's loader saved (Trie<A,M>::save, src/Trie.cpp:15-35;
 * ParMemoryBuffer::save, src/utils/ParMemoryBuffer.cpp:87-99; MemoryBuffer::save,
 * src/utils/MemoryBuffer.cpp:104-114): <dir>/ram/ (or <dir>/mmap/) trieinfo.bin,
 * data_head.bin and data_<i>.bin, read back as Trie<A,M>::load does
 * (src/Trie.cpp:37-75) and walked as Trie::foreach does (src/Trie.cpp:78-160).
 * anno_bytes = sizeof(A) of the relation's annotation type (8 for void*, long
 * and double, 4 for int and float).  Host-only: no GPU is touched.
 * ehx_refdb_columns copies the rows (lexicographic) into num_columns host arrays
 * of `rows` entries and, for an annotated relation, each row's annotation bytes
 * (zero-extended to 8) into annos (may be NULL). */
typedef struct ehx_refdb ehx_refdb;
int ehx_refdb_open(const char *dir, int anno_bytes, ehx_refdb **out);
int ehx_refdb_info(const ehx_refdb *db, uint32_t *num_columns, uint64_t *rows, int32_t *annotated);
int ehx_refdb_columns(const ehx_refdb *db, uint32_t *const *cols, uint64_t *annos);
int ehx_refdb_close(ehx_refdb *db);
/* The same, built into a device trie (binary relations) with layout `policy`:
 * Trie<A,M>::load of a reference database directory. */
int ehx_trie_load_reference(ehx_ctx *ctx, const char *dir, int anno_bytes, int policy, ehx_trie **out);
/* A relation's encoded column store, <dir>/encoded.bin
 * (EncodedColumnStore::from_binary, storage_engine/include/intermediate/
 * EncodedColumnStore.hpp:116-178): rows in file order, read through the
 * ehx_refdb calls above (the first annotation column, if any, as annotations). */
int ehx_encoded_open(const char *dir, ehx_refdb **out);
/* The trie of ordering (first, second) of that store, built on the device:
 * what the reference's loader does for every ordering of a relation
 * (query_compiler/.../createDB.scala:240-275 with Trie.cpp:380-510). */
int ehx_trie_build_encoded(ehx_ctx *ctx, const char *dir, int first, int second, int policy, ehx_trie **out);

/* ---- recursive annotated SUM (PageRank) -----------------------------------
 * The reference's recursive float-annotated rule
 *   R(x;y) :- Edge(x,z),y:float<-[base].
 *   R(x;y)*[i=iterations] :- Edge(x,z),R(z),W(z),y:float<-[add+mul*SUM(z;init)].
 * (PageRank, test/graph/travis.py:585-598, with base = 1.0/N, add 0.15,
 * mul 0.85, init 1.0, 5 iterations), as the generated code computes it
 * (query_compiler/src/main/scala/codegen/query/EHGenerator.scala:398-502 and
 * the recursion loop :752-764,820-828, reading annotations through
 * TrieIterator::get_annotation, storage_engine/src/TrieIterator.cpp:44-59).
 * R's keys are the head values of trie t (Edge, x's neighbours = its child
 * set); W is a unary float relation given as sorted distinct w_keys[n_w] with
 * annotations w_vals[n_w].  Per head x, over z ascending in child(x) ∩ R ∩ W:
 * acc(float) += (float)((init * R(z)) * W(z)); then R'(x) = (float)(add + mul*acc).
 * rank_out receives num_heads floats, by head node.  EHX_FORM_PLAIN folds in
 * that order (bit-identical to the reference's fold); EHX_FORM_TUNED (and
 * LEARNED) sums each list in parallel lanes with a fixed combining order
 * (deterministic, rounded differently).  *ms (may be NULL) = device time of the
 * iterations.  Values must be < 2^31. */
typedef struct {
  double base;        /* value of the base case (1.0/N for PageRank) */
  double add, mul;    /* y = add + mul * SUM */
  double init;        /* the SUM's per-term factor (SUM(z;init)) */
  int32_t iterations; /* recursion count ([i=iterations]); 0 = the base case */
  int32_t pad;
} ehx_rank_rule;
int ehx_pagerank(ehx_ctx *ctx, const ehx_trie *t, const uint32_t *w_keys, const float *w_vals, uint64_t n_w,
                 const ehx_rank_rule *rule, int form, float *rank_out, double *ms);

/* ---- 4-cycle count -------------------------------------------------------- */
/* Number of 4-cycles (each cycle once) of the undirected graph whose edges the
 * trie `sym` holds in both directions, restricted to cycles whose largest
 * vertex is a head value with index in [begin, end).  The reference engine has
 * no 4-cycle query (its 4-cycle would be a two-bag plan); the count is the
 * graph-theoretic one: sum over a of sum over c < a of C(|N(a) ∩ N(c) ∩ [0,a)|, 2). */
int ehx_cycle4_count(ehx_ctx *ctx, const ehx_trie *sym, uint64_t begin, uint64_t end,
                     int64_t *result, ehx_join_stats *stats);
/* Per-segment form: a cycle counts for the segment holding its largest vertex. */
int ehx_cycle4_count_segments(ehx_ctx *ctx, const ehx_trie *sym, const uint32_t *bounds,
                              uint64_t nseg, int64_t *counts, int64_t *result,
                              ehx_join_stats *stats);

/* ---- synthetic inputs (host) ---------------------------------------------- */
/* Counter-based SplitMix64 generators: identical output for identical
 * arguments on any machine. */
int ehx_gen_er(uint64_t num_vertices, uint64_t samples, uint64_t seed, uint32_t *src,
               uint32_t *dst);
int ehx_gen_rmat(int scale, uint64_t samples, double a, double b, double c, uint64_t seed,
                 uint32_t *src, uint32_t *dst);
/* Zipf(s) ranks in [0, k): P(r) proportional to (r + 1)^-s (inverse CDF of a
 * SplitMix64 uniform per index). */
int ehx_gen_zipf(uint64_t n, uint32_t k, double s, uint64_t seed, uint32_t *out);
/* Undirected clean-up: drop self loops, symmetrize, de-duplicate; relabel
 * (0 keep ids, 1 ascending-degree rank with ties by id and isolated vertices
 * dropped, 2 rank among used ids as dictionary encoding does, src/Encoding.cpp:
 * 21-48); then orient (0 both directions, 1 src < dst, 2 src > dst).  Output
 * arrays are allocated by the library (free with ehx_free). */
int ehx_graph_prepare(const uint32_t *src, const uint32_t *dst, uint64_t n, int relabel,
                      int orient, uint32_t **out_src, uint32_t **out_dst, uint64_t *out_m,
                      uint32_t *out_num_vertices);
/* Ego-network batch: num_nets subgraphs cut from an undirected graph given in
 * both directions (rows sorted and unique, as ehx_graph_prepare with orient 0
 * returns).  Net q has a target of t_q = floor(min_edges * (max_edges /
 * min_edges)^u) edges, u uniform in [0,1): log-uniform sizes.  Its center is
 * drawn among vertices of degree >= t_q / 2 (the pool halves towards the
 * highest degrees after each center whose net stays below t_q, at most 64
 * draws, then the largest net drawn is kept); the center's neighbors join in
 * hash order (a radius-1 BFS), each adding its edges to the vertices already
 * in, until the net holds t_q edges.  Vertices of net q are renumbered
 * net_base[q] + rank, rank by ascending degree inside the net; edges are
 * oriented src < dst and sorted.  net_base[num_nets + 1] and
 * net_off[num_nets + 1] (edge offsets) are caller arrays; the edge arrays are
 * allocated by the library (free with ehx_free). */
int ehx_gen_egonets(const uint32_t *src, const uint32_t *dst, uint64_t m, uint32_t num_vertices,
                    uint32_t num_nets, uint32_t min_edges, uint32_t max_edges, uint64_t seed,
                    uint32_t **out_src, uint32_t **out_dst, uint64_t *out_m, uint32_t *net_base,
                    uint64_t *net_off);
void ehx_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
