// query_io.hpp -- what Cython's PTrie / TrieWrapper does with a query's output
// trie (query_compiler/src/main/scala/codegen/trie/Trie.scala:167-241):
// reads every row and annotation through Trie::foreach.  Shared by the
// restated queries' C entry points.
#ifndef EHX_QUERY_IO_H
#define EHX_QUERY_IO_H

#include "emptyheaded.hpp"

template <class A>
inline double anno_value(const A a) { return (double)a; }
template <>
inline double anno_value<void *>(void *const) { return 0.0; }

// Copies up to `cap` rows (row-major, num_columns each) and, when `annos` is
// not NULL, each row's annotation as a double; returns the rows read.
template <class A, class M>
size_t export_rows(Trie<A, M> *t, uint32_t *rows, double *annos, size_t cap) {
  size_t r = 0;
  const size_t nc = t->num_columns;
  t->foreach ([&](std::vector<uint32_t> *tuple, A a) {
    if (r < cap) {
      for (size_t c = 0; c < nc; c++) rows[r * nc + c] = tuple->at(c);
      if (annos) annos[r] = anno_value<A>(a);
    }
    r++;
  });
  return r;
}

#endif
