// trie/TrieBlock.hpp -- TrieBlock<T,M> of the drop-in API
// (storage_engine/include/trie/TrieBlock.hpp:9-160).  A block is a handle to
// one set of a device-resident trie: the head set, or the child set of one
// head value.  Child lookup (get_next_block) is the NextLevel[data] lookup of
// the reference done on the host mirror of the head.
#ifndef EHX_DROPIN_TRIEBLOCK_H
#define EHX_DROPIN_TRIEBLOCK_H

#include "set/Set.hpp"

namespace ehx_dropin {
struct HostTrie;
// Node index of `data` in the trie's head set, or -1.
long head_node(const HostTrie *t, uint32_t data);
// The block of head node `node` (node == num_heads: the head block).
void *block_of(const HostTrie *t, uint32_t node);
uint32_t num_heads(const HostTrie *t);
// The block one level below `data` in the non-head block (t, node), or NULL.
void *child_of(const HostTrie *t, uint32_t node, uint32_t data);
bool block_contains(const HostTrie *t, uint32_t node, uint32_t data);
}  // namespace ehx_dropin

template <class T, class M>
struct TrieBlock {
  const ehx_dropin::HostTrie *trie;
  uint32_t node;
  Set<T> set;

  TrieBlock() : trie(nullptr), node(0) {}

  inline const Set<T> *get_const_set() const { return &set; }
  inline Set<T> *get_set() const { return const_cast<Set<T> *>(&set); }
  inline bool is_head() const { return trie && node == ehx_dropin::num_heads(trie); }

  // TrieBlock::get_next_block(data): the child block of `data`, NULL if absent
  // (or when this block is the relation's last level).
  inline TrieBlock<T, M> *get_next_block(const uint32_t data, M *buffer) const {
    (void)buffer;
    if (!trie) return nullptr;
    if (!is_head()) return (TrieBlock<T, M> *)ehx_dropin::child_of(trie, node, data);
    long h = ehx_dropin::head_node(trie, data);
    return h < 0 ? nullptr : (TrieBlock<T, M> *)ehx_dropin::block_of(trie, (uint32_t)h);
  }
  // get_next_block(index, data): dense heads index by data, sparse by index.
  inline TrieBlock<T, M> *get_next_block(const uint32_t index, const uint32_t data, M *buffer) const {
    (void)index;
    return get_next_block(data, buffer);
  }
  inline bool contains(const uint32_t data) const { return trie && ehx_dropin::block_contains(trie, node, data); }
};

#endif
