// emptyheaded.hpp -- umbrella header of the drop-in storage-engine API.
#ifndef EHX_DROPIN_EMPTYHEADED_H
#define EHX_DROPIN_EMPTYHEADED_H
#include "Encoding.hpp"
#include "Trie.hpp"
#include "TrieBuilder.hpp"
#include "TrieIterator.hpp"
#include "utils/parallel.hpp"
#include "utils/thread_pool.hpp"
#include "utils/timer.hpp"
#endif
