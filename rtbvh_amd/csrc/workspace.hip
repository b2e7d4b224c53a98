// Per-device cache of build workspaces (one block per builder kind).
#include <mutex>

#include "internal.hpp"

namespace rtbvh {

namespace {

constexpr int kMaxDevices = 64;

struct Slot {
    std::mutex mu;
    void* ptr = nullptr;
    size_t cap = 0;
    bool busy = false;
    // stream-ordered users (workspace_release_on): `done` is recorded after
    // the last one's work; pending until a later user has waited for it
    hipEvent_t done = nullptr;
    bool pending = false;
};

Slot g_slots[kMaxDevices][kWsKinds];

// The last stream-ordered user's work is complete (host wait).
void settle(Slot& s) {
    if (s.pending) {
        (void)hipEventSynchronize(s.done);
        s.pending = false;
    }
}

}  // namespace

hipError_t workspace_acquire(WorkspaceKind kind, size_t bytes, Workspace* out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = -1;
    out->ptr = nullptr;
    out->cached = false;
    out->dev = dev;
    out->kind = kind;
    if (dev >= 0 && bytes <= kWorkspaceCacheMax) {
        Slot& s = g_slots[dev][kind];
        std::lock_guard<std::mutex> lk(s.mu);
        if (!s.busy) {
            settle(s);  // host-ordered users start after every earlier user
            if (bytes > s.cap) {
                // grow: the old block is idle (not busy), so it can go
                if (s.ptr) (void)hipFree(s.ptr);
                s.ptr = nullptr;
                s.cap = 0;
                const size_t want = std::min(kWorkspaceCacheMax, bytes + bytes / 16);
                if (hipMalloc(&s.ptr, want) == hipSuccess) {
                    s.cap = want;
                } else {
                    (void)hipGetLastError();
                    s.ptr = nullptr;
                }
            }
            if (s.ptr && bytes <= s.cap) {
                s.busy = true;
                out->ptr = s.ptr;
                out->cached = true;
                return hipSuccess;
            }
        }
    }
    return hipMalloc(&out->ptr, bytes);
}

void workspace_release(Workspace* w, hipStream_t stream, bool wait) {
    if (!w->ptr) return;
    if (w->cached) {
        // the next build may reuse the block at once: its work must be done
        if (wait) {
            if (stream)
                (void)hipStreamSynchronize(stream);
            else
                (void)hipDeviceSynchronize();
        }
        Slot& s = g_slots[w->dev][w->kind];
        std::lock_guard<std::mutex> lk(s.mu);
        s.busy = false;
    } else {
        (void)hipFree(w->ptr);  // waits for the device
    }
    w->ptr = nullptr;
    w->cached = false;
}

// Stream-ordered use of a cached block: the work a caller enqueues on
// `stream` after workspace_acquire_on starts after the previous user's work
// on the block (a device-side wait on its event), and workspace_release_on
// returns the block at once, recording an event after the work enqueued on
// `stream` so far.  Neither waits on the host, except when the block grows
// (the old one is freed after the last user's work).  A request the cache
// cannot serve is allocated here and freed by the release, which waits.
hipError_t workspace_acquire_on(WorkspaceKind kind, size_t bytes, hipStream_t stream,
                                Workspace* out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = -1;
    out->ptr = nullptr;
    out->cached = false;
    out->dev = dev;
    out->kind = kind;
    if (dev >= 0 && bytes <= kWorkspaceCacheMax) {
        Slot& s = g_slots[dev][kind];
        std::lock_guard<std::mutex> lk(s.mu);
        if (!s.busy) {
            if (bytes > s.cap) {
                settle(s);
                if (s.ptr) (void)hipFree(s.ptr);
                s.ptr = nullptr;
                s.cap = 0;
                const size_t want = std::min(kWorkspaceCacheMax, bytes + bytes / 16);
                if (hipMalloc(&s.ptr, want) == hipSuccess) {
                    s.cap = want;
                } else {
                    (void)hipGetLastError();
                    s.ptr = nullptr;
                }
            }
            if (s.ptr && bytes <= s.cap) {
                if (s.pending && hipStreamWaitEvent(stream, s.done, 0) != hipSuccess) {
                    (void)hipGetLastError();
                    settle(s);
                }
                s.busy = true;
                out->ptr = s.ptr;
                out->cached = true;
                return hipSuccess;
            }
        }
    }
    return hipMalloc(&out->ptr, bytes);
}

void workspace_release_on(Workspace* w, hipStream_t stream) {
    if (!w->ptr) return;
    if (w->cached) {
        Slot& s = g_slots[w->dev][w->kind];
        std::lock_guard<std::mutex> lk(s.mu);
        if (s.done == nullptr &&
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            s.done = nullptr;
        }
        if (s.done && hipEventRecord(s.done, stream) == hipSuccess) {
            s.pending = true;
        } else {
            (void)hipGetLastError();
            (void)hipStreamSynchronize(stream);
            s.pending = false;
        }
        s.busy = false;
    } else {
        (void)hipFree(w->ptr);  // waits for the device
    }
    w->ptr = nullptr;
    w->cached = false;
}

}  // namespace rtbvh
