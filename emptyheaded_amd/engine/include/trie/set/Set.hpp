// trie/set/Set.hpp -- Set<T> of the drop-in API (fields and methods of
// storage_engine/include/trie/set/Set.hpp:19-172).  The 24-byte header is kept
// field for field; the elements of a set the generated code iterates (an
// intersection result materialized on the GPU, or a head set) are reached
// through `values` instead of trailing bytes.
#ifndef EHX_DROPIN_SET_H
#define EHX_DROPIN_SET_H

#include "../../utils/common.hpp"

// Layout tags (storage_engine/include/trie/set/layouts/).
struct hybrid {};
struct uinteger {};
struct range_bitset {};

template <class T>
class Set {
 public:
  uint32_t cardinality;
  uint32_t range;
  size_t number_of_bytes;
  type::layout type;
  const uint32_t *values;  // sorted elements on the host, or nullptr when not materialized

  Set() : cardinality(0), range(0), number_of_bytes(0), type(type::NOT_VALID), values(nullptr) {}
  Set(uint32_t c, uint32_t r, size_t nb, type::layout t)
      : cardinality(c), range(r), number_of_bytes(nb), type(t), values(nullptr) {}

  template <typename F>
  void foreach(F f) const {
    assert(values != nullptr || cardinality == 0);
    for (uint32_t i = 0; i < cardinality; i++) f(values[i]);
  }
  template <typename F>
  void foreach_index(F f) const {
    assert(values != nullptr || cardinality == 0);
    for (uint32_t i = 0; i < cardinality; i++) f(i, values[i]);
  }
  // Index of key or -1 (uinteger::find / range_bitset::find give the rank).
  long find(const uint32_t key) const {
    if (!values) return -1;
    const uint32_t *p = std::lower_bound(values, values + cardinality, key);
    return (p != values + cardinality && *p == key) ? (long)(p - values) : -1;
  }
};

#endif
