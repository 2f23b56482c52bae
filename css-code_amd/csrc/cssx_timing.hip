// Per-phase GPU timing, the counterpart of the reference's Do_Timings buckets
// (constants.f90:82, timings(1:16) filled around each phase in rhs_mpi.f90:324-766 and
// step_rk4.f90:343-491).  When enabled, every kernel launch of the library is bracketed by a
// pair of hipEvents recorded on the launch stream; reading the table synchronizes the device,
// folds the event pairs into per-phase totals and resets.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/cssx.h"
#include "cssx_internal.hpp"

namespace cssx {
namespace {
struct Rec {
    int phase;
    hipEvent_t a, b;
    double bytes;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_open;
std::vector<hipEvent_t> g_pool;
double g_ms[CSSX_NPHASE];
double g_bytes[CSSX_NPHASE];
long g_count[CSSX_NPHASE];

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

bool timing_on() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_on;
}

int timing_begin(int phase, hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_on) return -1;
    Rec r{phase, take_event(), take_event(), 0.0};
    if (!r.a || !r.b) return -1;
    (void)hipEventRecord(r.a, stream);
    g_open.push_back(r);
    return (int)g_open.size() - 1;
}

void timing_end(int token, hipStream_t stream, double bytes) {
    if (token < 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (token >= (int)g_open.size()) return;
    g_open[token].bytes = bytes;
    (void)hipEventRecord(g_open[token].b, stream);
}

}  // namespace cssx

using namespace cssx;

extern "C" int cssx_timing_enable(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = on != 0;
    return CSSX_OK;
}

extern "C" int cssx_timing_read(double *ms, long *count, double *bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (hipDeviceSynchronize() != hipSuccess) return CSSX_EHIP;
    for (const Rec &r : g_open) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess && r.phase >= 0 && r.phase < CSSX_NPHASE) {
            g_ms[r.phase] += t;
            g_bytes[r.phase] += r.bytes;
            g_count[r.phase] += 1;
        }
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    g_open.clear();
    for (int i = 0; i < CSSX_NPHASE; i++) {
        if (ms) ms[i] = g_ms[i];
        if (count) count[i] = g_count[i];
        if (bytes) bytes[i] = g_bytes[i];
        g_ms[i] = 0.0;
        g_bytes[i] = 0.0;
        g_count[i] = 0;
    }
    return CSSX_OK;
}
