// utils/common.hpp -- compile-time configuration and layout tags of the
// drop-in storage-engine API (names follow storage_engine/include/utils/common.hpp).
#ifndef EHX_DROPIN_COMMON_H
#define EHX_DROPIN_COMMON_H

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cassert>
#include <functional>
#include <iostream>
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
