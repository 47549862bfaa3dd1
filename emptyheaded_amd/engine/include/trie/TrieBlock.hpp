// trie/TrieBlock.hpp -- TrieBlock<T,M> of the drop-in API
// (storage_engine/include/trie/TrieBlock.hpp:9-160).  A block is a handle to
// one set of a device-resident trie: the head set, or the child set of one
// head value.  Child lookup (get_next_block) is the NextLevel[data] lookup of
// the reference done on the host mirror of the head.
#ifndef EHX_DROPIN_TRIEBLOCK_H
#define EHX_DROPIN_TRIEBLOCK_H

#include "set/Set.hpp"

namespace ehx_dropin {
// What inline head navigation reads: the first base of every HostTrie (the
// runtime's host mirror of a device trie), so a block reaches it from its
// trie pointer without a call.
struct HostNav {
  const int32_t *dense = nullptr;  // value -> head node, when the head is dense
  uint64_t dense_len = 0;
  const uint32_t *head = nullptr;  // sorted head values
  uint32_t nh = 0;
  char *blocks = nullptr;  // nh + 1 blocks by node (the last one is the head block)
};
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
  inline const ehx_dropin::HostNav *nav() const { return reinterpret_cast<const ehx_dropin::HostNav *>(trie); }
  inline bool is_head() const { return trie && node == nav()->nh; }

  // TrieBlock::get_next_block(data): the child block of `data`, NULL if absent
  // (or when this block is the relation's last level).  The head's NextLevel
  // lookup (TrieBlock.hpp:86-111) runs inline on the host mirror.
  inline TrieBlock<T, M> *get_next_block(const uint32_t data, M *buffer) const {
    (void)buffer;
    if (!trie) return nullptr;
    const ehx_dropin::HostNav *nv = nav();
    if (node != nv->nh) return (TrieBlock<T, M> *)ehx_dropin::child_of(trie, node, data);
    if (nv->dense) {  // one 4-byte load; the block itself is not read here
      if (data >= nv->dense_len) return nullptr;
      const int32_t h = nv->dense[data];
      return h < 0 ? nullptr : reinterpret_cast<TrieBlock<T, M> *>(nv->blocks) + h;
    }
    const uint32_t *p = std::lower_bound(nv->head, nv->head + nv->nh, data);
    const long h = (p != nv->head + nv->nh && *p == data) ? (long)(p - nv->head) : -1;
    return h < 0 ? nullptr : reinterpret_cast<TrieBlock<T, M> *>(nv->blocks) + h;
  }
  // get_next_block(index, data): dense heads index by data, sparse by index.
  inline TrieBlock<T, M> *get_next_block(const uint32_t index, const uint32_t data, M *buffer) const {
    (void)index;
    return get_next_block(data, buffer);
  }
  inline bool contains(const uint32_t data) const { return trie && ehx_dropin::block_contains(trie, node, data); }
};

#endif
