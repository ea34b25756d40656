// Fork / join of independent kernels of one entry point onto an auxiliary
// stream of the caller's device, so that a latency-bound kernel and a
// bandwidth-bound one (or two small ones) overlap.  Under stream capture the
// event record / wait pair becomes a fork and a join of the graph.
//
// One auxiliary stream per device (default priority, non-blocking), created
// on the first eager call on that device and kept until the process ends;
// the two events are per host thread and device.  Without a side stream
// (creation failed, or the first call on a device comes while the caller's
// stream is capturing) every kernel runs on the caller's stream, in order.
#include <mutex>

#include "drp_kernels.h"

namespace drp {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_side_mutex;
hipStream_t g_side[kMaxDevices];
bool g_side_tried[kMaxDevices];

struct ThreadEvents {
    hipEvent_t fork = nullptr, join = nullptr;
    bool tried = false;
};
thread_local ThreadEvents t_events[kMaxDevices];

// true when `main` is being captured, or when the query itself fails (a query
// the runtime refuses during another stream's capture); a failed query's error
// is cleared so that the launches after it report only their own
bool capturing(hipStream_t main) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(main, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return cs != hipStreamCaptureStatusNone;
}

hipStream_t side_stream(int dev, hipStream_t main) {
    std::lock_guard<std::mutex> lock(g_side_mutex);
    if (g_side_tried[dev]) return g_side[dev];
    if (capturing(main)) return nullptr;
    g_side_tried[dev] = true;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        s = nullptr;
    }
    g_side[dev] = s;
    return s;
}
}  // namespace

bool fork_begin(hipStream_t main, Fork &f) {
    f.side = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (dev < 0 || dev >= kMaxDevices) return false;
    ThreadEvents &e = t_events[dev];
    if (!e.tried) {
        if (capturing(main)) return false;
        e.tried = true;
        if (hipEventCreateWithFlags(&e.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&e.join, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            e.fork = e.join = nullptr;
        }
    }
    if (!e.fork || !e.join) return false;
    hipStream_t s = side_stream(dev, main);
    if (!s) return false;
    if (hipEventRecord(e.fork, main) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (hipStreamWaitEvent(s, e.fork, 0) != hipSuccess) return false;  // recorded, not joined: an error
    f.side = s;
    f.join = e.join;
    return true;
}

int fork_end(hipStream_t main, const Fork &f) {
    if (!f.side) return 0;
    if (hipEventRecord(f.join, f.side) != hipSuccess || hipStreamWaitEvent(main, f.join, 0) != hipSuccess) return -2;
    return 0;
}

}  // namespace drp
