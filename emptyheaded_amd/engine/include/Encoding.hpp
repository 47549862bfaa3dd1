// Encoding.hpp -- Encoding<T> of the drop-in API (storage_engine/include/Encoding.hpp,
// src/Encoding.cpp:81-104): the generated code loads one per attribute type and
// maps selection constants through value_to_key.  from_binary reads the
// reference's <path>encoding.bin (size_t num_distinct, then key_to_value[k] for
// each key) when present; relations written by this drop-in are already dense
// keys, and their encoding is the identity over [0, num_distinct).
#ifndef EHX_DROPIN_ENCODING_H
#define EHX_DROPIN_ENCODING_H

#include <fstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

template <class T>
struct Encoding {
  size_t num_distinct;
  std::vector<T> key_to_value;
  struct ValueToKey {
    bool identity = true;
    std::unordered_map<T, uint32_t> map;
    uint32_t at(const T &v) const { return identity ? (uint32_t)v : map.at(v); }
  } value_to_key;
  Encoding() : num_distinct(0) {}
  static Encoding<T> *from_binary(std::string path) {
    Encoding<T> *e = new Encoding<T>();
    std::ifstream ref(path + "encoding.bin", std::ios::binary);
    if (ref) {
      uint64_t n = 0;
      ref.read((char *)&n, sizeof(n));
      e->num_distinct = n;
      e->key_to_value.resize(n);
      if (n) ref.read((char *)e->key_to_value.data(), sizeof(T) * n);
      if (!ref) throw std::runtime_error("truncated " + path + "encoding.bin");
      e->value_to_key.identity = false;
      e->value_to_key.map.reserve(n);
      for (uint32_t k = 0; k < n; k++) e->value_to_key.map.emplace(e->key_to_value[k], k);
      return e;
    }
    std::ifstream in(path + "/ehx_encoding.bin", std::ios::binary);
    if (in) in.read((char *)&e->num_distinct, sizeof(uint64_t));
    return e;
  }
  T key_to_value_at(uint32_t k) const { return key_to_value.empty() ? (T)k : key_to_value.at(k); }
};

#endif
