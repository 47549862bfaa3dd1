// TransitiveClosure.hpp -- tc::unweighted_single_source of the drop-in API
// (storage_engine/include/TransitiveClosure.hpp:12-123), the call the code
// generator emits for a recursive rule whose base case selects one source and
// whose recursion takes a MIN until nothing changes (EHGenerator.scala:26-44,
// 86-117; SSSP in test/graph/travis.py:556-567).
//
// Same template parameters and arguments.  The level-by-level expansion runs
// on the GPU (ehx_sssp_unweighted[_dir]); the output trie receives, as the
// reference's does, one row per visited value (the start included) in value
// order, with annotation join^k(A(0)) for a value first reached at level k:
// the reference zeroes its visited block, so the start keeps A(0) (`init` is
// not read), and a value reached from frontier value f gets join(annotation
// of f).  With join(a) = 1 + a that is the BFS level.  output->num_rows = the
// number of visited values.
#ifndef EHX_DROPIN_TRANSITIVE_CLOSURE_H
#define EHX_DROPIN_TRANSITIVE_CLOSURE_H

#include <cassert>
#include <vector>

#include "Trie.hpp"
#include "trie/TrieBlock.hpp"

namespace tc {

template <class T, class M, class A, typename J>
void unweighted_single_source(const int iterations, const uint32_t start, const size_t num_distinct,
                              Trie<void *, M> *input, Trie<A, M> *output, const A init, J join) {
  (void)init;
  assert(input->num_columns == 2);
  input->getHead();  // a materialized input trie is built now
  std::vector<uint32_t> dist;
  const uint32_t levels = ehx_dropin::single_source_levels(input->host, start, iterations, num_distinct, dist);
  // annotation per level: A(0) at the start, then join applied once per level
  std::vector<uint64_t> level_anno(levels + 1);
  A a = (A)0;
  for (uint32_t k = 0; k <= levels; k++) {
    level_anno[k] = ehx_anno_bits<A>(a);
    a = join(a);
  }
  ehx_dropin::OutRows *out = output->output_rows();
  out->ncols = 1;
  out->annotated = true;
  out->rows.clear();
  out->annos.clear();
  for (size_t v = 0; v < dist.size(); v++) {
    if (dist[v] == 0xFFFFFFFFu) continue;
    out->rows.push_back((uint32_t)v);
    out->annos.push_back(level_anno[dist[v]]);
  }
  output->num_rows = out->num_rows();
}

}  // namespace tc

#endif
