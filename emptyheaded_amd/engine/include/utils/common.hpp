// utils/common.hpp -- compile-time configuration and layout tags of the
// drop-in storage-engine API (names follow storage_engine/include/utils/common.hpp).
#ifndef EHX_DROPIN_COMMON_H
#define EHX_DROPIN_COMMON_H

#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cassert>
#include <functional>
#include <iostream>
#include <new>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

// Host threads that run the generated code's per-thread lambdas
// (storage_engine/include/utils/common.hpp:32-34: -DNUM_THREADS_IN).
#ifndef NUM_THREADS_IN
#define NUM_THREADS_IN 16
#endif
static size_t NUM_THREADS = NUM_THREADS_IN;

// Padding between per-thread reducer slots (common.hpp:38).
#define PADDING 300

namespace ehx_dropin {
// Allocator of per-thread state: each allocation starts on a 128-byte boundary
// and is padded to a multiple of 128 bytes, so vectors written by different
// threads never share a cache line (or an adjacent-line prefetch pair).
template <class T>
struct LineAlloc {
  typedef T value_type;
  LineAlloc() = default;
  template <class U>
  LineAlloc(const LineAlloc<U> &) {}
  T *allocate(size_t n) {
    const size_t bytes = (n * sizeof(T) + 127) / 128 * 128;
    void *p = aligned_alloc(128, bytes ? bytes : 128);
    if (!p) throw std::bad_alloc();
    return static_cast<T *>(p);
  }
  void deallocate(T *p, size_t) { free(p); }
  template <class U>
  bool operator==(const LineAlloc<U> &) const { return true; }
  template <class U>
  bool operator!=(const LineAlloc<U> &) const { return false; }
};
template <class T>
using line_vector = std::vector<T, LineAlloc<T>>;
}  // namespace ehx_dropin

namespace common {
// Density rule (common.hpp:56-67): bitset iff more than 2 elements and
// span/cardinality <= 128; exact as an integer test for 32-bit values.
inline bool is_sparse(size_t length, size_t range) { return length <= 2 || range > 128 * length; }
}  // namespace common

namespace type {
enum layout : uint8_t { RANGE_BITSET = 0, UINTEGER = 1, HYBRID = 2, BLOCK_BITSET = 3, BLOCK = 4, NOT_VALID = 8 };
enum primitive : uint8_t { BOOL = 0, UINT32 = 1, UINT64 = 2, STRING = 3 };
}  // namespace type

#endif
