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

// An annotation value as the 64 bits the runtime keeps (zero-extended), and back.
template <class A>
inline uint64_t ehx_anno_bits(const A v) {
  uint64_t b = 0;
  memcpy(&b, &v, sizeof(A) < 8 ? sizeof(A) : 8);
  return b;
}
template <class A>
inline A ehx_anno_value(const uint64_t b) {
  A a = (A)0;
  memcpy(&a, &b, sizeof(A) < 8 ? sizeof(A) : 8);
  return a;
}

template <class A, class M>
struct Trie {
  bool annotated;
  size_t num_rows;
  size_t num_columns;
  M *memoryBuffers;
  std::vector<void *> encodings;
  A annotation;
  ehx_dropin::HostTrie *host;
  ehx_dropin::OutRows *out;  // rows of a materialized output trie

  Trie()
      : annotated(false), num_rows(0), num_columns(0), memoryBuffers(nullptr), annotation((A)0), host(nullptr),
        out(nullptr) {}

  // An output trie of a bag, Trie.hpp:46-50: the aggregate holder of a scalar
  // bag, or the rows its builders materialize.
  Trie(std::string path, size_t num_columns_in, bool annotated_in)
      : annotated(annotated_in), num_rows(0), num_columns(num_columns_in), memoryBuffers(new M(path, 2)),
        annotation((A)0), host(nullptr), out(nullptr) {}

  // Build from encoded columns (Trie.cpp:380-510): relations of one to eight
  // attributes, with one annotation per row when annotation_in is not empty.
  Trie(std::string path, std::vector<uint32_t> *max_set_sizes, std::vector<std::vector<uint32_t>> *attr_in,
       std::vector<A> *annotation_in)
      : annotated(annotation_in && !annotation_in->empty()), num_rows(0), num_columns(attr_in->size()),
        memoryBuffers(new M(path, 2)), annotation((A)0), host(nullptr), out(nullptr) {
    (void)max_set_sizes;
    const size_t rows = attr_in->at(0).size();
    std::vector<const uint32_t *> cols;
    for (size_t c = 0; c < num_columns; c++) {
      assert(attr_in->at(c).size() == rows);
      cols.push_back(attr_in->at(c).data());
    }
    std::vector<uint64_t> bits;
    if (annotated) {
      assert(annotation_in->size() == rows);
      for (const A &a : *annotation_in) bits.push_back(ehx_anno_bits<A>(a));
    }
    host = ehx_dropin::build_rows_trie((int)num_columns, cols.data(), rows, annotated ? bits.data() : nullptr);
    num_rows = ehx_dropin::host_trie_rows(host);
  }

  ~Trie() {
    delete memoryBuffers;
    if (host) ehx_dropin::free_host_trie(host);
    delete out;
  }

  // The row store the builders of this output trie fill.
  ehx_dropin::OutRows *output_rows() {
    if (!out) {
      out = new ehx_dropin::OutRows();
      out->ncols = (int)num_columns;
      out->annotated = annotated;
    }
    return out;
  }

  // Trie::foreach (Trie.cpp:78-160): every tuple in lexicographic order, with
  // its annotation when the trie is annotated.
  void foreach (const std::function<void(std::vector<uint32_t> *, A)> body) {
    if (out && out->ncols > 0) {
      const size_t nc = out->ncols, nr = out->num_rows();
      std::vector<uint32_t> t(nc);
      for (size_t r = 0; r < nr; r++) {
        for (size_t c = 0; c < nc; c++) t[c] = out->rows[r * nc + c];
        A a = (A)0;
        if (r < out->annos.size()) memcpy(&a, &out->annos[r], sizeof(A) < 8 ? sizeof(A) : 8);
        body(&t, a);
      }
      return;
    }
    if (!host) return;
    std::vector<uint32_t> rows;
    std::vector<uint64_t> an;
    ehx_dropin::host_trie_all_rows(host, rows, &an);
    const size_t nc = (size_t)ehx_dropin::host_trie_columns(host);
    std::vector<uint32_t> t(nc);
    for (size_t i = 0; i < an.size(); i++) {
      for (size_t c = 0; c < nc; c++) t[c] = rows[i * nc + c];
      body(&t, ehx_anno_value<A>(an[i]));
    }
  }

  void save() {
    if (host) ehx_dropin::save_host_trie(host, memoryBuffers->path);
  }

  TrieBlock<layout, M> *getHead() {
    if (!host && out && out->ncols >= 1 && out->num_rows()) {
      // a materialized output (with its annotations) read as the input of a
      // later bag or recursion step
      const size_t nr = out->num_rows(), nc = out->ncols;
      std::vector<std::vector<uint32_t>> c(nc, std::vector<uint32_t>(nr));
      std::vector<const uint32_t *> cols(nc);
      for (size_t k = 0; k < nc; k++) {
        for (size_t r = 0; r < nr; r++) c[k][r] = out->rows[nc * r + k];
        cols[k] = c[k].data();
      }
      const bool an = out->annotated && out->annos.size() == nr;
      host = ehx_dropin::build_rows_trie((int)nc, cols.data(), nr, an ? out->annos.data() : nullptr);
    }
    if (!host) return nullptr;
    return (TrieBlock<layout, M> *)ehx_dropin::block_of(host, ehx_dropin::num_heads(host));
  }

  // Trie::load (Trie.cpp:37-75): a relation folder the reference saved, or
  // one written by save() here.
  static Trie<A, M> *load(std::string path) {
    Trie<A, M> *t = new Trie<A, M>();
    t->memoryBuffers = new M(path, 2);
    t->host = ehx_dropin::load_host_trie(path, (int)(sizeof(A) < 8 ? sizeof(A) : 8));
    t->num_columns = (size_t)ehx_dropin::host_trie_columns(t->host);
    t->annotated = ehx_dropin::host_trie_annotated(t->host);
    t->num_rows = ehx_dropin::host_trie_rows(t->host);
    return t;
  }
};

#endif
