// Pinned staging ring behind mba_host::upload_tables.
#include "upload.hpp"

#include <cstring>
#include <mutex>

namespace mba_host {
namespace {

constexpr int kSlots = 64;    // copies that may be in flight per device
constexpr int kDevices = 16;  // devices with a ring

struct Slot {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;  // ev recorded after a copy not yet known complete
};

struct Ring {
    Slot slot[kSlots];
    int next = 0;
};

Ring g_ring[kDevices];
std::mutex g_mu;

}  // namespace

int upload_tables(void* dst, const void* src, size_t bytes, hipStream_t st)
{
    if (bytes == 0) return 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevices) return -1;
    std::lock_guard<std::mutex> lock(g_mu);
    Ring& R = g_ring[dev];
    Slot& S = R.slot[R.next];
    R.next = (R.next + 1) % kSlots;
    if (S.pending) {
        if (hipEventSynchronize(S.ev) != hipSuccess) return -1;
        S.pending = false;
    }
    if (S.cap < bytes) {
        size_t cap = (size_t)64 << 10;
        while (cap < bytes) cap <<= 1;
        void* p = nullptr;
        if (hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess) return -1;
        if (S.host) (void)hipHostFree(S.host);
        S.host = p;
        S.cap = cap;
    }
    if (!S.ev && hipEventCreateWithFlags(&S.ev, hipEventDisableTiming) != hipSuccess) return -1;
    std::memcpy(S.host, src, bytes);
    if (hipMemcpyAsync(dst, S.host, bytes, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
    if (hipEventRecord(S.ev, st) != hipSuccess) return -1;
    S.pending = true;
    return 0;
}

}  // namespace mba_host
