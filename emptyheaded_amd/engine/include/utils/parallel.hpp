// utils/parallel.hpp -- par::for_range with dynamic chunks and par::reducer
// (interfaces of storage_engine/include/utils/parallel.hpp:8-34 and
// src/utils/parallel.cpp:42-79).
#ifndef EHX_DROPIN_PARALLEL_H
#define EHX_DROPIN_PARALLEL_H

#include "common.hpp"

namespace ehx_dropin {
// 0: no pass running, 1: the builders record calls instead of computing them,
// 2: the final replay.  Reducer updates are dropped while recording, so only
// the final replay folds values.
extern int g_phase;
inline bool recording() { return g_phase == 1; }
inline bool replaying() { return g_phase == 2; }
}  // namespace ehx_dropin

namespace par {

// Dynamic scheduler: NUM_THREADS threads take chunks of block_size indices.
size_t for_range(const size_t from, const size_t to, const size_t block_size,
                 std::function<void(size_t tid, size_t i)> body);
size_t for_range(const size_t from, const size_t to, std::function<void(size_t tid, size_t i)> body);

template <class T>
class reducer {
 public:
  T *elem;
  std::function<T(T, T)> f;
  reducer(T init_in, std::function<T(T, T)> f_in) : f(f_in) {
    (void)init_in;
    elem = new T[NUM_THREADS * PADDING];
    clear();
  }
  ~reducer() { delete[] elem; }
  reducer(const reducer &) = delete;
  reducer &operator=(const reducer &) = delete;
  inline void update(size_t tid, T new_val) {
    if (ehx_dropin::recording()) return;
    elem[tid * PADDING] = f(elem[tid * PADDING], new_val);
  }
  inline void clear() {
    for (size_t i = 0; i < NUM_THREADS; i++) elem[i * PADDING] = (T)0;
  }
  inline T evaluate(T init) {
    for (size_t i = 0; i < NUM_THREADS; i++) init = f(init, elem[i * PADDING]);
    return init;
  }
};

}  // namespace par

#endif
