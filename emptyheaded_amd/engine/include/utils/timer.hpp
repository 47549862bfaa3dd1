// utils/timer.hpp -- wall-clock timers printed as "Time[<label>]: x s"
// (storage_engine/include/utils/timer.hpp:7-21).
#ifndef EHX_DROPIN_TIMER_H
#define EHX_DROPIN_TIMER_H

#include <chrono>
#include <iostream>
#include <string>

namespace timer {
inline double start_clock() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline double stop_clock(std::string label, double t0) {
  double s = start_clock() - t0;
  std::cout << "Time[" << label << "]: " << s << " s" << std::endl;
  return s;
}
}  // namespace timer

#endif
