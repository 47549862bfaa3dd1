// dropin_runtime.hpp -- the non-template core behind the drop-in API
// (implemented in emptyheaded_amd/engine/src/dropin.cpp on top of libehx).
#ifndef EHX_DROPIN_RUNTIME_H
#define EHX_DROPIN_RUNTIME_H

#include "utils/common.hpp"
#include "utils/parallel.hpp"
#include "trie/TrieBlock.hpp"

namespace ehx_dropin {

// A device-resident trie over a binary relation plus its host-side head mirror.
HostTrie *build_host_trie(const uint32_t *c0, const uint32_t *c1, size_t rows);
// From a relation folder: the reference's saved image (ram/ or mmap/
// trieinfo.bin and buffer files; anno_bytes = sizeof(A)), else ehx_rows.bin.
HostTrie *load_host_trie(const std::string &path, int anno_bytes);
// An annotated relation of arity 1 or 2 (annos: one value's bytes per row,
// zero-extended; NULL: unannotated), rows in any order.
HostTrie *build_annotated_trie(int ncols, const uint32_t *c0, const uint32_t *c1, const uint64_t *annos,
                               size_t rows);
// A relation of any arity (1..8) from its columns, with one annotation per row
// (value bytes zero-extended; NULL: unannotated); rows in any order.
HostTrie *build_rows_trie(int ncols, const uint32_t *const *cols, size_t rows, const uint64_t *annos);
// Every row (row-major, lexicographic) and, when annos is not NULL, its annotation.
void host_trie_all_rows(const HostTrie *t, std::vector<uint32_t> &rows, std::vector<uint64_t> *annos);
bool host_trie_annotated(const HostTrie *t);
const std::vector<uint64_t> *host_trie_annos(const HostTrie *t);  // NULL when unannotated
int host_trie_columns(const HostTrie *t);
// TrieIterator::get_annotation: the annotation bytes of `data` in the block's set.
uint64_t annotation_of(const void *block, uint32_t data);
void save_host_trie(const HostTrie *t, const std::string &path);
void free_host_trie(HostTrie *t);
size_t host_trie_rows(const HostTrie *t);
void host_trie_tuples(const HostTrie *t, std::vector<uint32_t> &c0, std::vector<uint32_t> &c1);

// One intersection call of a builder: operand blocks (nullptr = empty).
struct Operands {
  int n;
  const void *blocks[8];
};

// The tuples of a materialized output trie, row-major and lexicographic, with
// one annotation per row (the value's bytes, zero-extended to 64 bits) when
// the trie is annotated.
struct OutRows {
  int ncols = 0;
  bool annotated = false;
  std::vector<uint32_t> rows;
  std::vector<uint64_t> annos;
  size_t num_rows() const { return ncols ? rows.size() / ncols : 0; }
};

// ParTrieBuilder::build_aggregated_set / build_set (level 0): computed at once;
// a result's values are kept in *storage (the builder's), which out points into.
size_t aggregate_head(const Operands &ops, Set<hybrid> *out, std::vector<uint32_t> *storage);
// TrieBuilder::build_aggregated_set / build_set at builder depth `depth`
// (0 = the level below the head) of a query with `num_attrs` attributes:
// recorded or replayed by the pass driver.  A materialized call (build_set)
// always yields the result's values; an aggregated call at the innermost
// depth yields its cardinality only.
size_t aggregate(size_t num_attrs, uint32_t depth, const Operands &ops, Set<hybrid> *out, bool materialize = false);
// ParTrieBuilder::par_foreach_aggregate / par_foreach_builder: runs `body`
// (tid, index, value) over the level-0 values once per deferred level plus a
// final replay; rows emitted during the replay are gathered into `out`.
void run_passes(size_t num_attrs, const Set<hybrid> *level0, size_t nthreads,
                const std::function<void(size_t, uint32_t, uint32_t)> &body, OutRows *out = nullptr);
// Output rows of the current level-0 chunk (replay only): `prefix` followed by
// each value of `leaf`, or one row with an annotation.
void emit_rows(const uint32_t *prefix, int plen, const Set<hybrid> &leaf);
void emit_row(const uint32_t *row, int ncols, uint64_t anno_bits);

// tc::unweighted_single_source: the BFS level of every value from `start`
// over the relation's child sets, computed on the GPU (direction-optimizing
// when the relation is symmetric: the trie is then its own reverse).
// dist[v] = level or 0xFFFFFFFF (unreached), for v below max(num_distinct,
// largest value + 1); at most `iterations` levels (< 0: until the frontier
// empties).  Returns the number of levels run.
uint32_t single_source_levels(HostTrie *t, uint32_t start, int iterations, size_t num_distinct,
                              std::vector<uint32_t> &dist);

[[noreturn]] void not_implemented(const char *what);

}  // namespace ehx_dropin

#endif
