// pmp_trace.cpp -- roctx ranges with the marker library opened at run time
// (see pmp_trace.h).  The library's functions, once found, stay valid for
// the life of the process (it is never closed), so a range whose tracing is
// turned off between its push and its pop still pops.
#include "pmp_trace.h"

#include <dlfcn.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <mutex>

#include "pmp/pmp.h"

namespace pmp {
namespace {

using push_fn = int (*)(const char*);
using pop_fn = int (*)();
std::atomic<push_fn> g_push{nullptr};
std::atomic<pop_fn> g_pop{nullptr};
std::atomic<bool> g_on{false};
std::once_flag g_env_once;
std::mutex g_mu;

// rocprofv3's marker library first, then the older roctracer one.  Caller
// holds g_mu.
bool load() {
    if (g_push.load() && g_pop.load()) return true;
    static const char* const libs[] = {"librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1",
                                       "libroctx64.so.4", "/opt/rocm/lib/libroctx64.so.4"};
    for (const char* l : libs) {
        void* h = dlopen(l, RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        auto push = reinterpret_cast<push_fn>(dlsym(h, "roctxRangePushA"));
        auto pop = reinterpret_cast<pop_fn>(dlsym(h, "roctxRangePop"));
        if (push && pop) {
            g_pop = pop;
            g_push = push;
            return true;
        }
        dlclose(h);
    }
    return false;
}

void env_once() {
    std::call_once(g_env_once, [] {
        const char* e = getenv("PMP_ROCTX");
        if (e && strcmp(e, "1") == 0) {
            std::lock_guard<std::mutex> lk(g_mu);
            g_on = load();
        }
    });
}

}  // namespace

bool trace_enable(bool on) {
    env_once();
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = on && load();
    return !on || g_on.load();
}

TraceRange::TraceRange(const char* name) {
    env_once();
    on = g_on.load(std::memory_order_relaxed);
    if (on) g_push.load()(name);
}

TraceRange::~TraceRange() {
    if (on) g_pop.load()();
}

}  // namespace pmp

extern "C" int pmp_set_roctx(int32_t on) { return pmp::trace_enable(on != 0) ? PMP_OK : PMP_ENOTFOUND; }
