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
};

Slot g_slots[kMaxDevices][kWsKinds];

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

}  // namespace rtbvh
