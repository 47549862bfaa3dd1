// Encoding.hpp -- Encoding<T> of the drop-in API (storage_engine/include/Encoding.hpp,
// src/Encoding.cpp:54-118): the generated code loads one per attribute type and
// maps selection constants through value_to_key.  from_binary reads the
// reference's <path>encoding.bin: a size_t count, then key_to_value[k] for each
// key -- sizeof(T) bytes for a fixed-size type, a size_t length and the bytes
// for std::string (Encoding.cpp:106-118).  Relations written by this drop-in's
// tools for a fixed-size type are already dense keys, and their encoding is the
// identity over [0, num_distinct).
#ifndef EHX_DROPIN_ENCODING_H
#define EHX_DROPIN_ENCODING_H

#include <fstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

template <class T>
struct Encoding {
  size_t num_distinct;
  std::vector<T> key_to_value;
  struct ValueToKey {
    bool identity = true;
    std::unordered_map<T, uint32_t> map;
    uint32_t at(const T &v) const {
      if constexpr (std::is_arithmetic<T>::value) {
        if (identity) return (uint32_t)v;
      }
      return map.at(v);  // throws std::out_of_range for an absent value, as the reference's .at does
    }
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
      if constexpr (std::is_same<T, std::string>::value) {
        for (uint64_t k = 0; k < n && ref; k++) {
          uint64_t len = 0;
          ref.read((char *)&len, sizeof(len));
          std::string s(len, '\0');
          if (len) ref.read(&s[0], (std::streamsize)len);
          e->key_to_value[k] = std::move(s);
        }
      } else {
        if (n) ref.read((char *)e->key_to_value.data(), sizeof(T) * n);
      }
      if (!ref) throw std::runtime_error("truncated " + path + "encoding.bin");
      e->value_to_key.identity = false;
      e->value_to_key.map.reserve(n);
      for (uint32_t k = 0; k < n; k++) e->value_to_key.map.emplace(e->key_to_value[k], k);
      return e;
    }
    if constexpr (!std::is_arithmetic<T>::value)
      throw std::runtime_error("no " + path + "encoding.bin");
    std::ifstream in(path + "/ehx_encoding.bin", std::ios::binary);
    if (in) in.read((char *)&e->num_distinct, sizeof(uint64_t));
    return e;
  }
  T key_to_value_at(uint32_t k) const {
    if constexpr (std::is_arithmetic<T>::value) {
      if (key_to_value.empty()) return (T)k;
    }
    return key_to_value.at(k);
  }
};

#endif
