// gsparc/timer.hpp -- the reference's phase timer (include/gsparc/timer.hpp,
// src/gsparc/timer.cpp): start/stop accumulate wall time; the engine's
// device phase times (HIP events) are added with add().
#ifndef GSPARC_TIMER_HPP_
#define GSPARC_TIMER_HPP_
#include <chrono>
#include <cstdio>

namespace gsparc {

class Timer {
 public:
  Timer() = default;
  void start() {
    if (running_) {
      std::fprintf(stderr, "Timer is already running!\n");
      return;
    }
    t0_ = now();
    running_ = true;
  }
  void stop() {
    if (!running_) {
      std::fprintf(stderr, "Timer is not running!\n");
      return;
    }
    total_ += now() - t0_;
    running_ = false;
  }
  double currentElapsed() const { return running_ ? now() - t0_ : 0.0; }
  void printElapsed(const char* msg = "Elapsed time") const {
    std::printf("%s: %f s\n", msg, now() - t0_);
  }
  void printTotal(const char* msg = "Total elapsed time") const {
    std::printf("%s: %f s\n", msg, total_);
  }
  double getTotalTime() const { return total_; }
  void reset() {
    total_ = 0;
    running_ = false;
  }
  // Adds a phase time measured elsewhere (seconds).
  void add(double seconds) { total_ += seconds; }

 private:
  static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
  }
  double t0_ = 0, total_ = 0;
  bool running_ = false;
};

}  // namespace gsparc
#endif
