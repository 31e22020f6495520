// Process-wide state shared by the eight variant libraries.
//
// Each variant library (libsimwave_hip_<2|3>d_<cd|vd>_<f32|f64>.so) links
// against this one, so however the libraries are loaded (ctypes loads them
// RTLD_LOCAL) the dynamic loader maps a single copy per process and every
// variant sees the same counters.
//
// The receiver-overlap lease: a forward run or plan that puts the receiver
// gather on a second stream of its own (flag 4 of the device entries) joins
// the two streams with event waits in both directions.  Streams share the
// device's few hardware queues, and when several host threads enqueue such
// two-way waits at once, a wait can land in a queue ahead of the work it
// waits for, behind another stream's wait on that queue's work: a cycle
// that never drains.  One holder per process cannot form such a cycle, so
// the lease admits one holder at a time.
#include <atomic>

#define SW_COMMON_EXPORT extern "C" __attribute__((visibility("default")))

namespace {
std::atomic<int> g_overlap_holders{0};
}

// 1 if the caller now holds the lease, 0 if another run or plan holds it.
SW_COMMON_EXPORT int simwave_hip_overlap_acquire(void)
{
    int expected = 0;
    return g_overlap_holders.compare_exchange_strong(expected, 1) ? 1 : 0;
}

SW_COMMON_EXPORT void simwave_hip_overlap_release(void)
{
    g_overlap_holders.store(0);
}

// Current holders (0 or 1).
SW_COMMON_EXPORT int simwave_hip_overlap_holders(void)
{
    return g_overlap_holders.load();
}
