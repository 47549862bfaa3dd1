/*
 * ehx_oracle.h -- CPU restatement of EmptyHeaded's storage-engine hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine in
 * emptyheaded_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product path never links or calls it.
 *
 * It restates, in plain C, the reference algorithms cited next to each entry
 * point (paths relative to the reference tree's storage_engine/):
 *   - layout rule            include/utils/common.hpp:56-67, layouts/hybrid.hpp:103-138
 *   - Set header             include/trie/set/Set.hpp:19-24,164-172
 *   - uinteger layout        layouts/uinteger.hpp:100-110,272-281; utils/utils.hpp:22-43
 *   - range_bitset layout    layouts/range_bitset.hpp:95-157,443-470; ops/sse_masks.hpp:12-21
 *   - set_intersect dispatch ops/intersection/set.hpp:53-167
 *   - uint x uint            ops/intersection/uinteger.hpp:1209-1261 (merge vs gallop rule)
 *   - uint x bitset          ops/intersection/hetero.hpp:29-82
 *   - bitset x bitset        ops/intersection/bitset.hpp:526-567,619-666
 *   - trie build             src/Trie.cpp:173-302,380-510
 *   - child lookup           include/trie/TrieBlock.hpp:26-35,86-111
 *   - N-way builder          src/TrieBuilder.cpp:242-348,478-578
 *   - dynamic scheduler      src/utils/parallel.cpp:42-79 (chunks of 128)
 *   - COUNT fold             query_compiler/.../EHGenerator.scala:399-520,616-641
 *
 * Every result on this path is an exact integer.
 */
#ifndef EHX_ORACLE_H
#define EHX_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* type::layout values (include/utils/common.hpp:77-85). */
enum { ORC_RANGE_BITSET = 0, ORC_UINTEGER = 1 };

/* Layout policies for trie build: the density rule, or forced layouts used by
 * the layout-independence checks. */
enum { ORC_POLICY_RULE = 0, ORC_POLICY_ALL_UINT = 1, ORC_POLICY_ALL_BITSET = 2 };

typedef struct {
  uint32_t cardinality;
  uint32_t range;            /* Set.range: max element (Set.hpp:166) */
  uint64_t number_of_bytes;  /* bytes of layout data */
  uint8_t type;              /* ORC_UINTEGER or ORC_RANGE_BITSET */
  const uint8_t *data;       /* layout bytes in the reference's format */
} orc_set;

int orc_is_sparse(uint64_t length, uint64_t range);
int orc_layout_of(const uint32_t *v, uint64_t n);
uint64_t orc_bytes_needed(uint64_t length, const uint32_t *v, int layout);
uint64_t orc_set_build(uint8_t *out, const uint32_t *v, uint64_t n, int layout, orc_set *s);
int64_t orc_set_find(const orc_set *s, uint32_t key);
uint64_t orc_set_decode(const orc_set *s, uint32_t *out);

/* Intersections.  orc_intersect materializes into cbuf (caller sized to the
 * larger operand's number_of_bytes + 8); orc_intersect_count only counts. */
uint64_t orc_intersect(orc_set *c, uint8_t *cbuf, const orc_set *a, const orc_set *b);
uint64_t orc_intersect_count(const orc_set *a, const orc_set *b);
/* Which algorithm the reference picks for a layout pair:
 * 0 = merge, 1 = gallop, 2 = uint x bitset probe, 3 = bitset AND, 4 = empty. */
int orc_algorithm_of(const orc_set *a, const orc_set *b);

/* Individual algorithms (sorted uint32 inputs). */
uint64_t orc_uu_merge(uint32_t *out, const uint32_t *a, uint64_t na, const uint32_t *b, uint64_t nb);
uint64_t orc_uu_gallop(uint32_t *out, const uint32_t *rare, uint64_t nr, const uint32_t *freq, uint64_t nf);

/* Two-level trie over a binary relation (columns c0, c1). */
typedef struct orc_trie orc_trie;
orc_trie *orc_trie_build(const uint32_t *c0, const uint32_t *c1, uint64_t rows, int policy);
void orc_trie_free(orc_trie *t);
uint64_t orc_trie_num_rows(const orc_trie *t);   /* distinct tuples */
void orc_trie_head(const orc_trie *t, orc_set *out);
int orc_trie_child(const orc_trie *t, uint32_t key, orc_set *out); /* 0 = absent */
uint64_t orc_trie_tuples(const orc_trie *t, uint32_t *c0, uint32_t *c1);
/* Per head node: layout type of its child set, its cardinality. */
void orc_trie_node_info(const orc_trie *t, uint8_t *types, uint32_t *cards);

/* Single-bag generic-join COUNT plan over binary relations.  Level k's
 * candidate set is the intersection of its operands: parent = -1 names the head
 * set of relation `rel`, parent = j < k the child set of x_j in `rel`, parent =
 * -2 the child set of the constant `value` in `rel` (an equality selection,
 * query_compiler/src/main/scala/codegen/EHGenerator.scala:657-701). */
#define ORC_MAX_ATTRS 8
#define ORC_MAX_OPS 8
#define ORC_PARENT_HEAD (-1)
#define ORC_PARENT_CONST (-2)
typedef struct { int32_t rel; int32_t parent; uint32_t value; } orc_op;
typedef struct {
  int32_t num_attrs;
  int32_t num_ops[ORC_MAX_ATTRS];
  orc_op ops[ORC_MAX_ATTRS][ORC_MAX_OPS];
} orc_plan;

/* COUNT(*) of the plan.  Level-0 values in [begin, end) of the level-0 set are
 * processed (end = UINT64_MAX for all), over nthreads pthreads with dynamic
 * chunks of 128.  Returns -1 on a malformed plan. */
int64_t orc_join_count(const orc_plan *p, orc_trie *const *tries, int nthreads,
                       uint64_t begin, uint64_t end);
/* Same over every stride-th chunk of 128 level-0 values (a bounded,
 * stratified sample for CPU timing). */
int64_t orc_join_count_sampled(const orc_plan *p, orc_trie *const *tries, int nthreads, uint64_t stride);
/* Pin join workers to CPUs (worker i on the i-th CPU of the caller's affinity
 * mask, as src/utils/thread_pool.cpp:44-49 pins worker i to CPU i); for
 * timing runs. */
void orc_set_pin_threads(int on);
/* Size of the level-0 candidate set of a plan. */
uint64_t orc_join_level0_size(const orc_plan *p, orc_trie *const *tries);
/* Per level-0 value counts (out has orc_join_level0_size entries). */
int orc_join_count_per_head(const orc_plan *p, orc_trie *const *tries, int nthreads,
                            int64_t *out);

/* Every result tuple in lexicographic order into cols (num_attrs columns of
 * cap entries, column j at cols + j*cap); returns the number of tuples (only
 * the first cap are written). */
uint64_t orc_join_materialize(const orc_plan *p, orc_trie *const *tries, uint32_t *cols, uint64_t cap);

/* Recursive annotated SUM (the reference's PageRank rule, test/graph/
 * travis.py:585-598): R(x) = (float)base for every head x of `edge`, then
 * `iterations` times, per head x, over z ascending in child(x) ∩ head(R) ∩ W
 * (the N-way builder, src/TrieBuilder.cpp:276-348):
 *   acc(float) += (float)((((double)1.0f * init) * R(z)) * W(z))
 * with R(z), W(z) read as TrieIterator::get_annotation does (find the key's
 * index, src/TrieIterator.cpp:44-59), then R'(x) = (float)(add + mul * acc)
 * (query_compiler/.../EHGenerator.scala:398-502) with add + mul*acc one
 * double FMA, as g++ -O3 -march=native -mavx2 compiles the generated code
 * (cython/query/setup.py:10-20; C++'s default -ffp-contract=fast), or a
 * multiply then add when fused_finish is 0.  Each term is rounded to float
 * before it is added (what that -O3 build's vectorized loop does; nothing else
 * is contracted, built with -ffp-contract=off).  W: sorted distinct keys with
 * float annotations.  out[h] = R of head node h.  0 on success. */
int orc_pagerank(const orc_trie *edge, const uint32_t *w_keys, const float *w_vals, uint64_t n_w, double base,
                 double add, double mul, double init, int iterations, int nthreads, int fused_finish, float *out);

/* Writes trie t as the reference engine saves a relation (src/Trie.cpp:15-35
 * trieinfo.bin; the TrieBlock layout of build_head / recursive_build,
 * src/Trie.cpp:226-377; MemoryBuffer::save, src/utils/MemoryBuffer.cpp:104-114,
 * under <dir>/ram/, or raw buffer files under <dir>/mmap/ when mmapped != 0).
 * Child blocks go to nbuffers thread buffers, chunks of 128 head values dealt
 * round-robin as the dynamic scheduler would.  0 on success. */
int orc_write_refdb(const orc_trie *t, const char *dir, int nbuffers, int mmapped);
/* The same with ncols 1 (the head set alone: a unary relation) or 2, and, when
 * annos is not NULL, one anno_bytes-wide annotation per row in the trie's row
 * order, written after the last level's sets as the reference places them
 * (src/Trie.cpp:357-376 for two columns, :490-503 for one); trieinfo.bin's
 * annotation field is anno_bytes wide (sizeof(A)). */
int orc_write_refdb_annotated(const orc_trie *t, const char *dir, int nbuffers, int mmapped, int ncols,
                              const uint8_t *annos, int anno_bytes);

/* 4-cycles (each once) of an undirected graph in symmetric CSR form (off[n+1],
 * adjacency sorted per vertex); per_vertex[a] (may be NULL) receives the cycles
 * whose largest vertex is a.  No reference counterpart. */
int64_t orc_cycle4_count(const uint64_t *off, const uint32_t *adj, uint32_t n, int nthreads,
                         int64_t *per_vertex);

#ifdef __cplusplus
}
#endif
#endif
