// aslTimer.h — asl::Timer: wall and processor time (src/utilities/aslTimer.h:34-53).
#ifndef ASLTIMER_H
#define ASLTIMER_H

#include <chrono>
#include <ctime>

namespace asl {

class Timer {
 public:
  void start() {
    r0_ = std::chrono::steady_clock::now();
    c0_ = std::clock();
  }
  void stop() {
    real_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - r0_).count();
    cpu_ += double(std::clock() - c0_) / CLOCKS_PER_SEC;
  }
  void reset() { real_ = cpu_ = 0.; }
  double realTime() const { return real_; }
  double processorTime() const { return cpu_; }
  double processorLoad() const { return real_ > 0. ? cpu_ / real_ : 0.; }

 private:
  std::chrono::steady_clock::time_point r0_;
  std::clock_t c0_ = 0;
  double real_ = 0., cpu_ = 0.;
};

}  // namespace asl

// as ASL's header (src/utilities/aslTimer.h:31)
using namespace std::chrono;

#endif
