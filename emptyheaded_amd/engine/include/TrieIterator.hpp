// TrieIterator.hpp -- TrieIterator / ParTrieIterator of the drop-in API
// (storage_engine/include/TrieIterator.hpp:25-60, src/TrieIterator.cpp:13-78).
#ifndef EHX_DROPIN_TRIEITERATOR_H
#define EHX_DROPIN_TRIEITERATOR_H

#include "Trie.hpp"

// One per thread, written on every child lookup: its own cache lines.
template <class A, class M>
struct alignas(128) TrieIterator {
  Trie<A, M> *trie;
  ehx_dropin::line_vector<const TrieBlock<layout, M> *> levels;
  size_t num_rows;

  TrieIterator(Trie<A, M> *t_in) : trie(t_in), num_rows(0) {
    levels.resize(trie->num_columns);
    if (trie->num_columns > 0) levels.at(0) = trie->getHead();
  }
  void get_next_block(const size_t level, const uint32_t data) {
    const TrieBlock<layout, M> *b = levels.at(level);
    levels.at(level + 1) = b ? b->get_next_block(data, trie->memoryBuffers) : nullptr;
  }
  void get_next_block(const size_t level, const uint32_t index, const uint32_t data) {
    const TrieBlock<layout, M> *b = levels.at(level);
    levels.at(level + 1) = b ? b->get_next_block(index, data, trie->memoryBuffers) : nullptr;
  }
  // TrieIterator::get_annotation (TrieIterator.cpp:44-59): the annotation of
  // `data` in the set at `level`, (A)0 when absent.
  A get_annotation(const size_t level, const uint32_t data) {
    const TrieBlock<layout, M> *b = levels.at(level);
    return b ? ehx_anno_value<A>(ehx_dropin::annotation_of(b, data)) : (A)0;
  }
  const TrieBlock<hybrid, M> *get_block(const size_t level) const { return levels.at(level); }
};

template <class A, class M>
struct ParTrieIterator {
  A annotation;
  TrieBlock<hybrid, M> *head;
  Trie<A, M> *trie;
  std::vector<TrieIterator<A, M> *> iterators;

  ParTrieIterator(Trie<A, M> *t_in) : annotation(t_in->annotation), head(t_in->getHead()), trie(t_in) {
    iterators.resize(NUM_THREADS);
    for (size_t i = 0; i < NUM_THREADS; i++) iterators.at(i) = new TrieIterator<A, M>(t_in);
  }
  ~ParTrieIterator() {
    for (auto *it : iterators) delete it;
  }
  void get_next_block(const uint32_t data) {
    head = head ? head->get_next_block(data, trie->memoryBuffers) : nullptr;
    for (size_t i = 0; i < NUM_THREADS; i++) iterators.at(i)->levels.at(0) = head;
  }
};

#endif
