// TrieBuilder.hpp -- TrieBuilder / ParTrieBuilder of the drop-in API
// (storage_engine/include/TrieBuilder.hpp:29-165, src/TrieBuilder.cpp).
//
// The generated code consumes each intersection inside host lambdas.  The
// builders here keep that contract while running every set intersection on
// the GPU: par_foreach_aggregate runs the lambdas once per builder level, each
// time recording that level's build_aggregated_set calls (they return empty
// sets) after replaying the results of the levels above, and computes the
// recorded calls in one batched launch; a final run replays everything and is
// the only one whose reducer updates count.  Only the innermost level needs
// no materialized result: its cardinality feeds arithmetic alone.
#ifndef EHX_DROPIN_TRIEBUILDER_H
#define EHX_DROPIN_TRIEBUILDER_H

#include "Trie.hpp"
#include "utils/parallel.hpp"

template <class A, class M>
struct TrieBuilder {
  Trie<A, M> *trie;
  std::vector<const Set<hybrid> *> aggregate_sets;
  std::vector<Set<hybrid>> results;
  uint32_t cur_level;
  uint32_t tmp_level;
  size_t num_attributes;

  TrieBuilder(Trie<A, M> *t_in, const size_t num_attributes_in)
      : trie(t_in), cur_level(1), tmp_level(0), num_attributes(num_attributes_in) {
    aggregate_sets.resize(num_attributes_in);
    results.resize(num_attributes_in);
  }

  size_t build_aggregated_set(const TrieBlock<hybrid, M> *s1) {
    ehx_dropin::Operands o{1, {s1}};
    return finish(ehx_dropin::aggregate(num_attributes, tmp_level, o, &results[tmp_level]));
  }
  size_t build_aggregated_set(const TrieBlock<hybrid, M> *s1, const TrieBlock<hybrid, M> *s2) {
    ehx_dropin::Operands o{2, {s1, s2}};
    return finish(ehx_dropin::aggregate(num_attributes, tmp_level, o, &results[tmp_level]));
  }
  size_t build_aggregated_set(std::vector<const TrieBlock<hybrid, M> *> *isets) {
    ehx_dropin::Operands o{(int)isets->size(), {}};
    if (o.n > 8) ehx_dropin::not_implemented("more than 8 operand sets");
    for (int i = 0; i < o.n; i++) o.blocks[i] = isets->at(i);
    return finish(ehx_dropin::aggregate(num_attributes, tmp_level, o, &results[tmp_level]));
  }

  template <typename F>
  inline void foreach_aggregate(F f) {
    const Set<hybrid> *s = aggregate_sets.at(tmp_level);
    tmp_level++;
    s->foreach(f);
    tmp_level--;
  }

  // Materializing output tries: not part of the aggregate path; the device join
  // materializes result tuples through ehx_join_materialize instead.
  size_t build_set(const size_t, const TrieBlock<hybrid, M> *) { ehx_dropin::not_implemented("TrieBuilder::build_set"); }
  size_t build_set(const size_t, const TrieBlock<hybrid, M> *, const TrieBlock<hybrid, M> *) {
    ehx_dropin::not_implemented("TrieBuilder::build_set");
  }
  size_t build_set(const size_t, std::vector<const TrieBlock<hybrid, M> *> *) {
    ehx_dropin::not_implemented("TrieBuilder::build_set");
  }
  void set_level(const uint32_t, const uint32_t) { ehx_dropin::not_implemented("TrieBuilder::set_level"); }
  void allocate_next(const size_t) { ehx_dropin::not_implemented("TrieBuilder::allocate_next"); }
  void allocate_annotation(const size_t) { ehx_dropin::not_implemented("TrieBuilder::allocate_annotation"); }
  void set_annotation(const A, const uint32_t, const uint32_t) {
    ehx_dropin::not_implemented("TrieBuilder::set_annotation");
  }
  template <typename F>
  void foreach_builder(F) {
    ehx_dropin::not_implemented("TrieBuilder::foreach_builder");
  }

 private:
  size_t finish(size_t card) {
    aggregate_sets.at(tmp_level) = &results[tmp_level];
    return card;
  }
};

template <class A, class M>
struct ParTrieBuilder {
  Trie<A, M> *trie;
  std::vector<TrieBuilder<A, M> *> builders;
  const TrieBlock<hybrid, M> *tmp_head;
  size_t num_attributes;
  Set<hybrid> head_result;

  ParTrieBuilder(Trie<A, M> *t_in, const size_t num_attributes_in)
      : trie(t_in), tmp_head(nullptr), num_attributes(num_attributes_in) {
    builders.resize(NUM_THREADS);
    for (size_t i = 0; i < NUM_THREADS; i++) builders.at(i) = new TrieBuilder<A, M>(t_in, num_attributes_in);
  }
  ~ParTrieBuilder() {
    for (auto *b : builders) delete b;
  }

  size_t build_aggregated_set(const TrieBlock<hybrid, M> *s1) {
    ehx_dropin::Operands o{1, {s1}};
    return ehx_dropin::aggregate_head(o, &head_result);
  }
  size_t build_aggregated_set(const TrieBlock<hybrid, M> *s1, const TrieBlock<hybrid, M> *s2) {
    ehx_dropin::Operands o{2, {s1, s2}};
    return ehx_dropin::aggregate_head(o, &head_result);
  }
  size_t build_aggregated_set(std::vector<const TrieBlock<hybrid, M> *> *isets) {
    ehx_dropin::Operands o{(int)isets->size(), {}};
    if (o.n > 8) ehx_dropin::not_implemented("more than 8 operand sets");
    for (int i = 0; i < o.n; i++) o.blocks[i] = isets->at(i);
    return ehx_dropin::aggregate_head(o, &head_result);
  }

  template <typename F>
  void par_foreach_aggregate(F f) {
    for (auto *b : builders) b->tmp_level = 0;
    ehx_dropin::run_passes(num_attributes, &head_result, NUM_THREADS,
                           [&](size_t tid, uint32_t d) { f(tid, d); });
  }

  size_t build_set(const TrieBlock<hybrid, M> *) { ehx_dropin::not_implemented("ParTrieBuilder::build_set"); }
  size_t build_set(const TrieBlock<hybrid, M> *, const TrieBlock<hybrid, M> *) {
    ehx_dropin::not_implemented("ParTrieBuilder::build_set");
  }
  size_t build_set(std::vector<const TrieBlock<hybrid, M> *> *) {
    ehx_dropin::not_implemented("ParTrieBuilder::build_set");
  }
  void allocate_next() { ehx_dropin::not_implemented("ParTrieBuilder::allocate_next"); }
  void allocate_annotation() { ehx_dropin::not_implemented("ParTrieBuilder::allocate_annotation"); }
  template <typename F>
  void par_foreach_builder(F) {
    ehx_dropin::not_implemented("ParTrieBuilder::par_foreach_builder");
  }
};

#endif
