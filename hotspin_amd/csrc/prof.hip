// prof.hip -- optional per-kernel timing with HIP events on the launching stream.
// Off by default; when on, every instrumented launch records an event pair
// so bench code can report per-kernel durations without an external profiler.
#include <vector>

#include "hs_common.hpp"

namespace {
struct Ev {
    int cat;
    hipEvent_t a, b;
};
bool g_on = false;
std::vector<Ev> g_ev;
std::vector<hipEvent_t> g_pool;
hipEvent_t get_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}
void recycle() {
    for (auto& e : g_ev) {
        g_pool.push_back(e.a);
        g_pool.push_back(e.b);
    }
    g_ev.clear();
}
}  // namespace

bool hs::prof_on() { return g_on; }
void hs::prof_begin(int cat, hipStream_t st) {
    if (!g_on) return;
    Ev e{cat, get_event(), get_event()};
    hipEventRecord(e.a, st);
    g_ev.push_back(e);
}
void hs::prof_end(int cat, hipStream_t st) {
    if (!g_on) return;
    for (auto it = g_ev.rbegin(); it != g_ev.rend(); ++it)
        if (it->cat == cat) {
            hipEventRecord(it->b, st);
            return;
        }
}

extern "C" void hs_prof_enable(int on) {
    hipDeviceSynchronize();
    recycle();
    g_on = on != 0;
}

extern "C" int hs_prof_read(double* total_ms, long* count, int ncat) {
    hipDeviceSynchronize();
    for (int c = 0; c < ncat; c++) {
        total_ms[c] = 0;
        count[c] = 0;
    }
    for (auto& e : g_ev) {
        if (e.cat < 0 || e.cat >= ncat) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            total_ms[e.cat] += ms;
            count[e.cat] += 1;
        }
    }
    recycle();
    return 0;
}
