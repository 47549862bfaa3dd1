// utils/ParMemoryBuffer.hpp -- the memory-manager template argument of the
// drop-in API.  Tries live in GPU memory owned by libehx; this type only
// carries the on-disk folder the generated code names.
#ifndef EHX_DROPIN_PARMEMORYBUFFER_H
#define EHX_DROPIN_PARMEMORYBUFFER_H

#include <string>

struct ParMemoryBuffer {
  std::string path;
  size_t num_buffers;
  ParMemoryBuffer(std::string path_in, size_t num_in) : path(path_in), num_buffers(num_in) {}
};

// The mmap-backed variant selects the same device storage.
struct ParMMapBuffer : ParMemoryBuffer {
  ParMMapBuffer(std::string path_in, size_t num_in) : ParMemoryBuffer(path_in, num_in) {}
};

#endif
