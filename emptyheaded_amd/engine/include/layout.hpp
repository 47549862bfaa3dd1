// layout.hpp -- the set layout every trie of the drop-in uses (storage_engine/include/layout.hpp).
#ifndef EHX_DROPIN_LAYOUT_H
#define EHX_DROPIN_LAYOUT_H
#include "trie/TrieBlock.hpp"
typedef hybrid layout;
#endif
