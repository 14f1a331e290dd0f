// Wall-clock stopwatches per kernel phase (initialize / apply / finalize), the
// timing model of the reference (src/stopwatch.hpp:7-58).
#pragma once

#include <chrono>
#include <memory>
#include <string>
#include <unordered_map>

namespace SMAX {

class Stopwatch {
    using clock = std::chrono::steady_clock;
    clock::time_point begin_{};
    long double wtime_ = 0;

  public:
    void start() { begin_ = clock::now(); }
    void stop() {
        wtime_ += std::chrono::duration<long double>(clock::now() - begin_).count();
    }
    long double check() const {
        return std::chrono::duration<long double>(clock::now() - begin_).count();
    }
    long double get_wtime() const { return wtime_; }
};

class Timers {
  public:
    std::unordered_map<std::string, std::unique_ptr<Stopwatch>> all;
    void add(const std::string &name) { all[name] = std::make_unique<Stopwatch>(); }
    Stopwatch *get(const std::string &name) {
        auto it = all.find(name);
        return it != all.end() ? it->second.get() : nullptr;
    }
};

// Runs a phase's stopwatch for one scope: started on construction, stopped on
// every way out of the scope (early error returns included).
class ScopedPhase {
    Stopwatch *sw_;

  public:
    ScopedPhase(Timers *timers, const char *phase) : sw_(timers->get(phase)) {
        if (sw_) sw_->start();
    }
    ~ScopedPhase() {
        if (sw_) sw_->stop();
    }
    ScopedPhase(const ScopedPhase &) = delete;
    ScopedPhase &operator=(const ScopedPhase &) = delete;
};

} // namespace SMAX
