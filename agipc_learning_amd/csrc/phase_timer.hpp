// Wall-clock phases of the host set-up path, printed to stderr when the
// environment variable MASX_PROFILE_SETUP is set (a diagnostic; silent and
// nearly free otherwise).
#pragma once
#include <chrono>
#include <cstdio>
#include <cstdlib>

namespace masx
{
class PhaseTimer
{
  public:
    explicit PhaseTimer(const char* scope) : scope_(scope), on_(std::getenv("MASX_PROFILE_SETUP") != nullptr)
    {
        if(on_) last_ = start_ = std::chrono::steady_clock::now();
    }
    void mark(const char* phase)
    {
        if(!on_) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[masx %s] %-28s %8.3f ms\n", scope_, phase,
                     std::chrono::duration<double, std::milli>(now - last_).count());
        last_ = now;
    }
    ~PhaseTimer()
    {
        if(on_)
            std::fprintf(stderr, "[masx %s] %-28s %8.3f ms\n", scope_, "total",
                         std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start_).count());
    }

  private:
    const char* scope_;
    bool        on_;
    std::chrono::steady_clock::time_point start_, last_;
};
}  // namespace masx
