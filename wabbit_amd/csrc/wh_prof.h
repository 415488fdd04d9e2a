// wh_prof.h -- optional host-side wall-time accounting of the orchestration
// code (adapt_tree, sync_ghosts_tree).  Enabled when the environment variable
// WH_HOST_PROFILE is set at the first measured scope; the totals per scope
// name are printed to stderr when the process exits.
#pragma once

#include <chrono>

namespace wh {

bool prof_enabled();
void prof_add(const char *name, double ms);

struct ProfScope {
    const char *name;
    std::chrono::steady_clock::time_point t0;
    bool on;
    explicit ProfScope(const char *n) : name(n), on(prof_enabled())
    {
        if (on) t0 = std::chrono::steady_clock::now();
    }
    ~ProfScope()
    {
        if (on)
            prof_add(name, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
};

}  // namespace wh

#define WH_PROF_CAT2(a, b) a##b
#define WH_PROF_CAT(a, b) WH_PROF_CAT2(a, b)
#define WH_PROF(name) ::wh::ProfScope WH_PROF_CAT(wh_prof_, __LINE__)(name)
