// Named wall-clock timers with the reference's semantics (src/logger.h:94-232): host time
// between start and stop, no device synchronisation of its own (phases that wait for the
// GPU, such as the solves, include that wait), accumulated per event over the run.
#pragma once
#include <chrono>
#include <iomanip>
#include <map>
#include <ostream>
#include <string>

namespace cuibm {

class Logger {
  public:
    void start(const std::string &event) { tic_[event] = now(); }
    void stop(const std::string &event) { timer_[event] += now() - tic_[event]; }
    const std::map<std::string, double> &timers() const { return timer_; }
    // Logger::printAllTime (src/logger.h:187-200): one line per event, then the total
    void printAll(std::ostream &os) const {
        double total = 0;
        os << "\n";
        for (const auto &e : timer_) {
            os << std::setw(16) << std::setprecision(4) << std::fixed << e.second << "\t" << e.first << "\n";
            total += e.second;
        }
        os << "-------------------------------------\n";
        os << std::setw(16) << std::setprecision(4) << total << "\tTOTAL\n";
    }
    // Logger::writeTime (src/logger.h:216-222): "event seconds" per line
    void write(std::ostream &os) const {
        for (const auto &e : timer_) os << e.first << " " << e.second << "\n";
    }

  private:
    static double now() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }
    std::map<std::string, double> tic_, timer_;
};

// times the enclosing scope as one event
class ScopedTimer {
  public:
    ScopedTimer(Logger &l, const char *event) : l_(l), e_(event) { l_.start(e_); }
    ~ScopedTimer() { l_.stop(e_); }
    ScopedTimer(const ScopedTimer &) = delete;
    ScopedTimer &operator=(const ScopedTimer &) = delete;

  private:
    Logger &l_;
    std::string e_;
};

}  // namespace cuibm
