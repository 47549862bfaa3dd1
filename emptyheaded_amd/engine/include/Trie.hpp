// Trie.hpp -- Trie<A,M> of the drop-in API (storage_engine/include/Trie.hpp:22-67):
// same constructors, fields and methods; storage is a device trie built by
// libehx (radix sort + segmented scans in HBM) with a host mirror of the head.
#ifndef EHX_DROPIN_TRIE_H
#define EHX_DROPIN_TRIE_H

#include <iostream>
#include <string>
#include <vector>

#include "dropin_runtime.hpp"
#include "layout.hpp"
#include "utils/ParMemoryBuffer.hpp"

template <class A, class M>
struct Trie {
  bool annotated;
  size_t num_rows;
  size_t num_columns;
  M *memoryBuffers;
  std::vector<void *> encodings;
  A annotation;
  ehx_dropin::HostTrie *host;

  Trie() : annotated(false), num_rows(0), num_columns(0), memoryBuffers(nullptr), annotation((A)0), host(nullptr) {}

  // An output trie of a bag (aggregate holder), Trie.hpp:46-50.
  Trie(std::string path, size_t num_columns_in, bool annotated_in)
      : annotated(annotated_in), num_rows(0), num_columns(num_columns_in), memoryBuffers(new M(path, 2)),
        annotation((A)0), host(nullptr) {}

  // Build from encoded columns (Trie.cpp:380-510).  Binary relations.
  Trie(std::string path, std::vector<uint32_t> *max_set_sizes, std::vector<std::vector<uint32_t>> *attr_in,
       std::vector<A> *annotation_in)
      : annotated(annotation_in && !annotation_in->empty()), num_rows(0), num_columns(attr_in->size()),
        memoryBuffers(new M(path, 2)), annotation((A)0), host(nullptr) {
    (void)max_set_sizes;
    if (num_columns != 2) ehx_dropin::not_implemented("tries over relations of arity other than 2");
    assert(attr_in->at(0).size() == attr_in->at(1).size() && !attr_in->at(0).empty());
    host = ehx_dropin::build_host_trie(attr_in->at(0).data(), attr_in->at(1).data(), attr_in->at(0).size());
    num_rows = ehx_dropin::host_trie_rows(host);
  }

  ~Trie() {
    delete memoryBuffers;
    if (host) ehx_dropin::free_host_trie(host);
  }

  // Trie::foreach (Trie.cpp:130-160): every tuple in lexicographic order.
  void foreach (const std::function<void(std::vector<uint32_t> *, A)> body) {
    if (!host) return;
    std::vector<uint32_t> c0, c1, t(2);
    ehx_dropin::host_trie_tuples(host, c0, c1);
    for (size_t i = 0; i < c0.size(); i++) {
      t[0] = c0[i];
      t[1] = c1[i];
      body(&t, (A)0);
    }
  }

  void save() {
    if (host) ehx_dropin::save_host_trie(host, memoryBuffers->path);
  }

  TrieBlock<layout, M> *getHead() {
    if (!host) return nullptr;
    return (TrieBlock<layout, M> *)ehx_dropin::block_of(host, ehx_dropin::num_heads(host));
  }

  // Trie::load (Trie.cpp:37-75) from a relation folder written by save().
  static Trie<A, M> *load(std::string path) {
    Trie<A, M> *t = new Trie<A, M>();
    t->memoryBuffers = new M(path, 2);
    t->num_columns = 2;
    t->host = ehx_dropin::load_host_trie(path);
    t->num_rows = ehx_dropin::host_trie_rows(t->host);
    return t;
  }
};

#endif
