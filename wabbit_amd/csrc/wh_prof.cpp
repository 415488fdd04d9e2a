// wh_prof.cpp -- totals of the host-side scopes of wh_prof.h.
#include "wh_prof.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>

namespace wh {
namespace {

struct Table {
    std::mutex m;
    std::map<std::string, std::pair<double, long>> t;
    ~Table()
    {
        if (t.empty()) return;
        std::fprintf(stderr, "# wabbit_hip host profile (ms total, calls, us/call)\n");
        for (auto &kv : t)
            std::fprintf(stderr, "#   %-28s %10.3f %8ld %10.2f\n", kv.first.c_str(), kv.second.first, kv.second.second,
                         1e3 * kv.second.first / (double)kv.second.second);
    }
};

Table &table()
{
    static Table T;
    return T;
}

}  // namespace

bool prof_enabled()
{
    static const bool on = std::getenv("WH_HOST_PROFILE") != nullptr;
    return on;
}

void prof_add(const char *name, double ms)
{
    Table &T = table();
    std::lock_guard<std::mutex> g(T.m);
    auto &e = T.t[name];
    e.first += ms;
    e.second += 1;
}

}  // namespace wh
