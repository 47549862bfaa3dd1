// Encoding.hpp -- Encoding<T> of the drop-in API (storage_engine/include/Encoding.hpp):
// the generated code loads one per attribute type.  Relations handed to the
// drop-in are already dictionary-encoded (dense uint32 keys), so the encoding
// is the identity over [0, num_distinct).
#ifndef EHX_DROPIN_ENCODING_H
#define EHX_DROPIN_ENCODING_H

#include <fstream>
#include <string>
#include <unordered_map>

template <class T>
struct Encoding {
  size_t num_distinct;
  struct Identity {
    uint32_t at(const T &v) const { return (uint32_t)v; }
  } value_to_key;
  Encoding() : num_distinct(0) {}
  // Reads <path>/ehx_encoding.bin (one uint64: number of keys) when present.
  static Encoding<T> *from_binary(std::string path) {
    Encoding<T> *e = new Encoding<T>();
    std::ifstream in(path + "/ehx_encoding.bin", std::ios::binary);
    if (in) in.read((char *)&e->num_distinct, sizeof(uint64_t));
    return e;
  }
  T key_to_value(uint32_t k) const { return (T)k; }
};

#endif
