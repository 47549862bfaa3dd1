// utils/thread_pool.hpp -- per-query execution resources.  The reference pins
// NUM_THREADS pthreads (src/utils/thread_pool.cpp:116-130); here initializing
// also opens the GPU context every TrieBuilder call batches its work onto.
#ifndef EHX_DROPIN_THREAD_POOL_H
#define EHX_DROPIN_THREAD_POOL_H

namespace thread_pool {
void initializeThreadPool();
void deleteThreadPool();
}  // namespace thread_pool

#endif
