// Device memory pool.
//
// hipFree waits for the whole device, so a request loop whose requests
// allocate and free their own buffers serialises every worker on it.  Blocks
// released here are kept and handed out again instead.  A released block is
// reused only once every library stream has finished the work enqueued before
// its release: the release records an event on each registered stream (every
// context's stream) and the block waits until all of them have completed.
// Work on other streams that touches library buffers must be complete before
// the buffer's owner is destroyed (the library's own entry points synchronise
// or stay on their context's stream).
#include <map>
#include <mutex>

#include "common.hpp"

namespace feasm {

namespace {

struct Block {
    void *p = nullptr;
    size_t bytes = 0;
    std::vector<hipEvent_t> evs;
    bool unfenced = false;  // some stream could not be fenced: freed only after a device synchronisation
};

struct Pool {
    std::mutex mu;
    std::multimap<size_t, Block> cached;  // by capacity
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> spare;
    size_t cached_bytes = 0;
};

Pool &pool(int device) {
    static std::mutex mu;
    static std::map<int, Pool *> pools;  // never destroyed: blocks may be released at exit
    std::lock_guard<std::mutex> lk(mu);
    auto &p = pools[device];
    if (!p) p = new Pool;
    return *p;
}

int current_device() {
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

// capacity classes: multiples of 1/8 of the enclosing power of two (at most
// 12.5 % slack), 256 bytes at least
size_t size_class(size_t n) {
    if (n <= 256) return 256;
    size_t p2 = 256;
    while (p2 < n) p2 <<= 1;
    const size_t step = p2 / 16;
    return (n + step - 1) / step * step;
}

bool ready(Block &b) {
    if (b.unfenced) return false;
    for (hipEvent_t e : b.evs)
        if (hipEventQuery(e) != hipSuccess) return false;
    return true;
}

void recycle_events(Pool &P, Block &b) {
    for (hipEvent_t e : b.evs) P.spare.push_back(e);
    b.evs.clear();
}

// frees every cached block whose work is done (all of them when sync)
void trim_locked(Pool &P, bool sync) {
    if (sync) (void)hipDeviceSynchronize();
    for (auto it = P.cached.begin(); it != P.cached.end();) {
        if (sync || ready(it->second)) {
            recycle_events(P, it->second);
            (void)hipFree(it->second.p);
            P.cached_bytes -= it->second.bytes;
            it = P.cached.erase(it);
        } else {
            ++it;
        }
    }
}

}  // namespace

void *pool_alloc(size_t n, size_t *capacity, int *device) {
    const size_t cap = size_class(n);
    *device = current_device();
    Pool &P = pool(*device);
    {
        std::lock_guard<std::mutex> lk(P.mu);
        for (auto it = P.cached.lower_bound(cap); it != P.cached.end() && it->first <= cap + cap / 4; ++it) {
            if (!ready(it->second)) continue;
            Block b = it->second;
            P.cached.erase(it);
            P.cached_bytes -= b.bytes;
            recycle_events(P, b);
            *capacity = b.bytes;
            return b.p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, cap);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(P.mu);
            trim_locked(P, true);
        }
        e = hipMalloc(&p, cap);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(FEASM_ENOMEM, std::string("device memory exhausted (") + std::to_string(cap) + " bytes)");
        }
    }
    *capacity = cap;
    return p;
}

// makes `device` current for the scope (events and frees belong to it)
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int device) {
        const int cur = current_device();
        if (cur != device && hipSetDevice(device) == hipSuccess) prev = cur;
    }
    ~DeviceScope() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

void pool_release(void *p, size_t capacity, int device) {
    if (!p) return;
    DeviceScope scope(device);
    Pool &P = pool(device);
    std::lock_guard<std::mutex> lk(P.mu);
    Block b;
    b.p = p;
    b.bytes = capacity;
    for (hipStream_t s : P.streams) {
        // a stream being captured into a graph cannot take the fence
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
            (void)hipGetLastError();
            b.unfenced = true;
            continue;
        }
        hipEvent_t e;
        if (!P.spare.empty()) {
            e = P.spare.back();
            P.spare.pop_back();
        } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipDeviceSynchronize();  // cannot fence: free it outright
            (void)hipFree(p);
            recycle_events(P, b);
            return;
        }
        if (hipEventRecord(e, s) != hipSuccess) {
            (void)hipGetLastError();
            b.unfenced = true;
            P.spare.push_back(e);
            continue;
        }
        b.evs.push_back(e);
    }
    P.cached.emplace(capacity, std::move(b));
    P.cached_bytes += capacity;
}

void pool_register_stream(int device, hipStream_t s) {
    Pool &P = pool(device);
    std::lock_guard<std::mutex> lk(P.mu);
    P.streams.push_back(s);
}

void pool_unregister_stream(int device, hipStream_t s) {
    DeviceScope scope(device);
    Pool &P = pool(device);
    std::lock_guard<std::mutex> lk(P.mu);
    // the stream is idle (its owner synchronised it): its fences are complete
    for (size_t i = 0; i < P.streams.size(); i++)
        if (P.streams[i] == s) {
            P.streams.erase(P.streams.begin() + i);
            break;
        }
    if (P.streams.empty()) trim_locked(P, false);
}

size_t pool_cached_bytes(int device) {
    Pool &P = pool(device);
    std::lock_guard<std::mutex> lk(P.mu);
    return P.cached_bytes;
}

}  // namespace feasm
