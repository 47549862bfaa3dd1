// TrieBuilder.hpp -- TrieBuilder / ParTrieBuilder of the drop-in API
// (storage_engine/include/TrieBuilder.hpp:29-165, src/TrieBuilder.cpp).
//
// The generated code consumes each intersection inside host lambdas.  The
// builders here keep that contract while running every set intersection on
// the GPU: par_foreach_aggregate / par_foreach_builder run the lambdas once per
// builder level, each time recording that level's build calls (they return
// empty sets) after replaying the results of the levels above, and compute the
// recorded calls in one batched launch; a final run replays everything and is
// the only one whose reducer updates and output rows count.  Only an
// aggregated innermost level needs no materialized result: its cardinality
// feeds arithmetic alone.  Materialized output tries are gathered as rows in
// lexicographic order (what Trie::foreach reads back).
#ifndef EHX_DROPIN_TRIEBUILDER_H
#define EHX_DROPIN_TRIEBUILDER_H

#include "Trie.hpp"
#include "utils/parallel.hpp"

// One per thread, written on every build call: its own cache lines.
template <class A, class M>
struct alignas(128) TrieBuilder {
  Trie<A, M> *trie;
  ehx_dropin::line_vector<const Set<hybrid> *> aggregate_sets;
  ehx_dropin::line_vector<Set<hybrid>> results;
  uint32_t cur_level;
  uint32_t tmp_level;
  size_t num_attributes;
  // Materialized sets by output level (cur_level) and the value the enclosing
  // foreach is at on each level: a row of the output trie is a path of them.
  ehx_dropin::line_vector<Set<hybrid>> level_sets;
  ehx_dropin::line_vector<uint32_t> prefix;

  TrieBuilder(Trie<A, M> *t_in, const size_t num_attributes_in)
      : trie(t_in), cur_level(1), tmp_level(0), num_attributes(num_attributes_in) {
    aggregate_sets.resize(num_attributes_in);
    results.resize(num_attributes_in);
    level_sets.resize(num_attributes_in + 1);
    prefix.resize(num_attributes_in + 1);
  }

  // Depth of a build call below the head: materialized levels first, then the
  // aggregated ones (the attribute order puts materialized attributes first).
  uint32_t depth() const { return cur_level - 1 + tmp_level; }

  size_t build_aggregated_set(const TrieBlock<hybrid, M> *s1) {
    ehx_dropin::Operands o{1, {s1}};
    return finish(ehx_dropin::aggregate(num_attributes, depth(), o, &results[tmp_level]));
  }
  size_t build_aggregated_set(const TrieBlock<hybrid, M> *s1, const TrieBlock<hybrid, M> *s2) {
    ehx_dropin::Operands o{2, {s1, s2}};
    return finish(ehx_dropin::aggregate(num_attributes, depth(), o, &results[tmp_level]));
  }
  size_t build_aggregated_set(std::vector<const TrieBlock<hybrid, M> *> *isets) {
    ehx_dropin::Operands o{(int)isets->size(), {}};
    if (o.n > 8) ehx_dropin::not_implemented("more than 8 operand sets");
    for (int i = 0; i < o.n; i++) o.blocks[i] = isets->at(i);
    return finish(ehx_dropin::aggregate(num_attributes, depth(), o, &results[tmp_level]));
  }

  template <typename F>
  inline void foreach_aggregate(F f) {
    const Set<hybrid> *s = aggregate_sets.at(tmp_level);
    tmp_level++;
    s->foreach(f);
    tmp_level--;
  }

  // TrieBuilder::build_set (TrieBuilder.cpp:44-221): the set of output level
  // cur_level, materialized on the GPU with the level's other calls.
  size_t build_set(const size_t tid, const TrieBlock<hybrid, M> *s1) {
    (void)tid;
    ehx_dropin::Operands o{1, {s1}};
    return ehx_dropin::aggregate(num_attributes, depth(), o, &level_sets.at(cur_level), true);
  }
  size_t build_set(const size_t tid, const TrieBlock<hybrid, M> *s1, const TrieBlock<hybrid, M> *s2) {
    (void)tid;
    ehx_dropin::Operands o{2, {s1, s2}};
    return ehx_dropin::aggregate(num_attributes, depth(), o, &level_sets.at(cur_level), true);
  }
  size_t build_set(const size_t tid, std::vector<const TrieBlock<hybrid, M> *> *isets) {
    (void)tid;
    ehx_dropin::Operands o{(int)isets->size(), {}};
    if (o.n > 8) ehx_dropin::not_implemented("more than 8 operand sets");
    for (int i = 0; i < o.n; i++) o.blocks[i] = isets->at(i);
    return ehx_dropin::aggregate(num_attributes, depth(), o, &level_sets.at(cur_level), true);
  }

  // Room for child pointers / annotations (TrieBuilder.cpp:376-446): the output
  // trie is gathered as rows, so there is nothing to reserve.
  void allocate_next(const size_t) {}
  void allocate_annotation(const size_t) {}

  // TrieBuilder::set_level (TrieBuilder.cpp:391-412): hangs the set of level
  // cur_level under `data` of level cur_level-1.  When that set is the output
  // trie's last column, its paths are rows.
  void set_level(const uint32_t index, const uint32_t data) {
    (void)index;
    if (!ehx_dropin::replaying() || cur_level + 1 != trie->num_columns) return;
    prefix.at(cur_level - 1) = data;
    ehx_dropin::emit_rows(prefix.data(), (int)cur_level, level_sets.at(cur_level));
  }
  // TrieBuilder::set_annotation (TrieBuilder.cpp:428-446): the annotation of
  // `data` in the set of level cur_level-1, the output trie's last column.
  void set_annotation(const A value, const uint32_t index, const uint32_t data) {
    (void)index;
    if (!ehx_dropin::replaying() || cur_level != trie->num_columns) return;
    prefix.at(cur_level - 1) = data;
    ehx_dropin::emit_row(prefix.data(), (int)cur_level, ehx_anno_bits<A>(value));
  }

  // TrieBuilder::foreach_builder (TrieBuilder.hpp:96-115): (index, value) over
  // the set of level cur_level.
  template <typename F>
  void foreach_builder(F f) {
    const Set<hybrid> &s = level_sets.at(cur_level);
    if (s.cardinality && !s.values) ehx_dropin::not_implemented("foreach_builder over a set without values");
    const uint32_t lvl = cur_level;
    cur_level++;
    for (uint32_t i = 0; i < s.cardinality; i++) {
      prefix.at(lvl) = s.values[i];
      f(i, s.values[i]);
    }
    cur_level--;
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
  std::vector<uint32_t> head_values;  // head_result's values when computed here

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
    return ehx_dropin::aggregate_head(o, &head_result, &head_values);
  }
  size_t build_aggregated_set(const TrieBlock<hybrid, M> *s1, const TrieBlock<hybrid, M> *s2) {
    ehx_dropin::Operands o{2, {s1, s2}};
    return ehx_dropin::aggregate_head(o, &head_result, &head_values);
  }
  size_t build_aggregated_set(std::vector<const TrieBlock<hybrid, M> *> *isets) {
    ehx_dropin::Operands o{(int)isets->size(), {}};
    if (o.n > 8) ehx_dropin::not_implemented("more than 8 operand sets");
    for (int i = 0; i < o.n; i++) o.blocks[i] = isets->at(i);
    return ehx_dropin::aggregate_head(o, &head_result, &head_values);
  }

  // ParTrieBuilder::build_aggregated_equality_selection_set (TrieBuilder.cpp:
  // 461-467): the index of `data` in s1's set, as Set::find gives it (-1 absent).
  // Answered from a head set's host values; a child set's values stay on the GPU.
  size_t build_aggregated_equality_selection_set(const uint32_t data, const TrieBlock<hybrid, M> *s1) {
    if (!s1) return (size_t)-1;
    if (!s1->set.values && s1->set.cardinality)
      ehx_dropin::not_implemented("equality selection over a child set (values on the GPU)");
    return (size_t)s1->get_set()->find(data);
  }

  template <typename F>
  void par_foreach_aggregate(F f) {
    for (auto *b : builders) {
      b->tmp_level = 0;
      b->cur_level = 1;
    }
    ehx_dropin::run_passes(num_attributes, &head_result, NUM_THREADS,
                           [&](size_t tid, uint32_t, uint32_t d) { f(tid, d); });
  }

  // ParTrieBuilder::build_set (TrieBuilder.cpp:580-700): the head set of the
  // output trie.  A one-column output is that set.
  size_t build_set(const TrieBlock<hybrid, M> *s1) {
    ehx_dropin::Operands o{1, {s1}};
    return head_rows(ehx_dropin::aggregate_head(o, &head_result, &head_values));
  }
  size_t build_set(const TrieBlock<hybrid, M> *s1, const TrieBlock<hybrid, M> *s2) {
    ehx_dropin::Operands o{2, {s1, s2}};
    return head_rows(ehx_dropin::aggregate_head(o, &head_result, &head_values));
  }
  size_t build_set(std::vector<const TrieBlock<hybrid, M> *> *isets) {
    ehx_dropin::Operands o{(int)isets->size(), {}};
    if (o.n > 8) ehx_dropin::not_implemented("more than 8 operand sets");
    for (int i = 0; i < o.n; i++) o.blocks[i] = isets->at(i);
    return head_rows(ehx_dropin::aggregate_head(o, &head_result, &head_values));
  }
  void allocate_next() {}
  void allocate_annotation() {}

  // ParTrieBuilder::par_foreach_builder (TrieBuilder.hpp:159-164): (tid, index,
  // value) over the head set; the rows its builders emit form the output trie.
  template <typename F>
  void par_foreach_builder(F f) {
    for (auto *b : builders) {
      b->tmp_level = 0;
      b->cur_level = 1;
    }
    ehx_dropin::OutRows *out = trie->output_rows();
    out->rows.clear();  // the builders' rows replace a one-column head
    out->annos.clear();
    ehx_dropin::run_passes(num_attributes, &head_result, NUM_THREADS,
                           [&](size_t tid, uint32_t i, uint32_t d) {
                             builders.at(tid)->prefix.at(0) = d;
                             f(tid, i, d);
                           },
                           out);
  }

 private:
  size_t head_rows(size_t card) {
    if (trie->num_columns == 1) {
      ehx_dropin::OutRows *out = trie->output_rows();
      out->rows.assign(head_result.values, head_result.values + card);
    }
    return card;
  }
};

#endif
