// Device memory pool.
//
// hipFree waits for the whole device, so a request loop whose requests
// allocate and free their own buffers serialises every worker on it.  Blocks
// released here are kept and handed out again instead, also to contexts
// created later, until an allocation fails (everything cached is then freed
// after a device synchronisation and the allocation retried) or the caller
// trims the pool (feasm_pool_trim).  A released block is
// reused only once every library stream has finished the work enqueued before
// its release: the release records an event on each registered stream (every
// context's stream) and the block waits until all of them have completed.
// Work on other streams that touches library buffers must be complete before
// the buffer's owner is destroyed (the library's own entry points synchronise
// or stay on their context's stream).
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <shared_mutex>

#include "common.hpp"

namespace feasm {

namespace {

struct Block {
    void *p = nullptr;
    size_t bytes = 0;
    std::vector<hipEvent_t> evs;
    bool unfenced = false;  // some stream could not be fenced: freed only after a device synchronisation
};

thread_local hipStream_t t_fence_stream = nullptr;
thread_local int t_fence_device = -1;
hipStream_t thread_fence_stream() { return t_fence_stream; }
int thread_fence_device() { return t_fence_device; }

struct Pool {
    std::mutex mu;           // the cache and the spare events
    std::shared_mutex smu;   // the stream list: shared while fencing, exclusive to change it
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

// a query that does not return success leaves its status as the calling
// thread's last error, where the next launch check (ours or a library's)
// would take it for its own: cleared here
void clear_query_status(hipError_t q) {
    if (q != hipSuccess && hipPeekAtLastError() == q) (void)hipGetLastError();
}

bool ready(Block &b) {
    if (b.unfenced) return false;
    for (hipEvent_t e : b.evs) {
        const hipError_t q = hipEventQuery(e);
        if (q != hipSuccess) {
            clear_query_status(q);
            return false;
        }
    }
    return true;
}

// Drops every completed fence of the cached blocks.  Called when a stream is
// unregistered (it is idle then, so every fence recorded on it is complete):
// no cached block keeps an event of a stream that is about to be destroyed,
// since querying an event whose stream is gone is not safe.
void drop_completed_fences_locked(Pool &P) {
    for (auto &kv : P.cached) {
        Block &b = kv.second;
        for (size_t i = 0; i < b.evs.size();) {
            const hipError_t q = hipEventQuery(b.evs[i]);
            if (q == hipSuccess) {
                P.spare.push_back(b.evs[i]);
                b.evs[i] = b.evs.back();
                b.evs.pop_back();
            } else {
                clear_query_status(q);
                i++;
            }
        }
    }
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
        // the smallest ready block of at least the size class, up to twice it
        for (auto it = P.cached.lower_bound(cap); it != P.cached.end() && it->first <= 2 * cap; ++it) {
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
    Block b;
    b.p = p;
    b.bytes = capacity;
    {
        // the fences are recorded outside the cache lock (other threads keep
        // allocating meanwhile); the shared lock keeps every stream
        // registered, and so alive, while it is fenced
        std::shared_lock<std::shared_mutex> sl(P.smu);
        // a thread that declared the one stream its buffers are used on fences
        // that stream only (pool_set_thread_fence); others fence every stream
        const hipStream_t one[1] = {thread_fence_stream()};
        const bool only = one[0] != nullptr && thread_fence_device() == device;
        const hipStream_t *fs = only ? one : P.streams.data();
        const size_t nfs = only ? 1 : P.streams.size();
        for (size_t k = 0; k < nfs; k++) {
            const hipStream_t s = fs[k];
            // a stream being captured into a graph cannot take the fence
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
                (void)hipGetLastError();
                b.unfenced = true;
                continue;
            }
            hipEvent_t e = nullptr;
            {
                std::lock_guard<std::mutex> lk(P.mu);
                if (!P.spare.empty()) {
                    e = P.spare.back();
                    P.spare.pop_back();
                }
            }
            if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                (void)hipDeviceSynchronize();  // cannot fence: free it outright
                (void)hipFree(p);
                std::lock_guard<std::mutex> lk(P.mu);
                recycle_events(P, b);
                return;
            }
            if (hipEventRecord(e, s) != hipSuccess) {
                (void)hipGetLastError();
                b.unfenced = true;
                std::lock_guard<std::mutex> lk(P.mu);
                P.spare.push_back(e);
                continue;
            }
            b.evs.push_back(e);
        }
    }
    std::lock_guard<std::mutex> lk(P.mu);
    P.cached.emplace(capacity, std::move(b));
    P.cached_bytes += capacity;
}

void pool_register_stream(int device, hipStream_t s) {
    Pool &P = pool(device);
    std::unique_lock<std::shared_mutex> sl(P.smu);
    std::lock_guard<std::mutex> lk(P.mu);
    P.streams.push_back(s);
}

void pool_unregister_stream(int device, hipStream_t s) {
    DeviceScope scope(device);
    Pool &P = pool(device);
    std::unique_lock<std::shared_mutex> sl(P.smu);
    std::lock_guard<std::mutex> lk(P.mu);
    // the stream is idle (its owner synchronised it): its fences are complete
    for (size_t i = 0; i < P.streams.size(); i++)
        if (P.streams[i] == s) {
            P.streams.erase(P.streams.begin() + i);
            break;
        }
    // cached blocks stay for the next contexts (feasm_pool_trim frees them),
    // without the fences recorded on this stream
    drop_completed_fences_locked(P);
}

size_t pool_trim(int device) {
    DeviceScope scope(device);
    Pool &P = pool(device);
    std::unique_lock<std::shared_mutex> sl(P.smu);
    std::lock_guard<std::mutex> lk(P.mu);
    const size_t before = P.cached_bytes;
    trim_locked(P, true);
    return before - P.cached_bytes;
}

void pool_set_thread_fence(int device, hipStream_t s) {
    t_fence_stream = s;
    t_fence_device = s ? device : -1;
}

ScopedThreadFence::ScopedThreadFence(int device, hipStream_t s) : prev_s(t_fence_stream), prev_d(t_fence_device) {
    pool_set_thread_fence(device, s);
}

ScopedThreadFence::~ScopedThreadFence() {
    t_fence_stream = prev_s;
    t_fence_device = prev_d;
}

namespace {
// pinned host buffers of read_back, one per host thread; a thread's buffer
// goes back to a free list when the thread ends (never freed: the HIP runtime
// may be gone when the last threads end)
struct Pinned {
    void *p = nullptr;
    size_t bytes = 0;
};
std::mutex g_pinned_mu;
std::vector<Pinned> g_pinned_free;
struct PinnedHolder {
    Pinned b;
    ~PinnedHolder() {
        if (!b.p) return;
        std::lock_guard<std::mutex> lk(g_pinned_mu);
        g_pinned_free.push_back(b);
    }
};
thread_local PinnedHolder t_pinned;
}  // namespace

void read_back(hipStream_t s, void *host, const void *dev, size_t bytes) {
    const ReadBack one{host, dev, bytes};
    read_back(s, &one, 1);
}

void read_back(hipStream_t s, const ReadBack *parts, int n) {
    size_t bytes = 0;
    for (int i = 0; i < n; i++) bytes += (parts[i].bytes + 15) & ~size_t(15);
    Pinned &b = t_pinned.b;
    if (b.bytes < bytes) {
        Pinned nb;
        {
            std::lock_guard<std::mutex> lk(g_pinned_mu);
            for (size_t i = 0; i < g_pinned_free.size(); i++)
                if (g_pinned_free[i].bytes >= bytes) {
                    nb = g_pinned_free[i];
                    g_pinned_free.erase(g_pinned_free.begin() + i);
                    break;
                }
            if (b.p) g_pinned_free.push_back(b);  // too small for this thread now; another may use it
        }
        if (!nb.p) {
            nb.bytes = std::max<size_t>(bytes, 4096);
            FEASM_HIP(hipHostMalloc(&nb.p, nb.bytes, hipHostMallocDefault));
        }
        b = nb;
    }
    unsigned char *q = static_cast<unsigned char *>(b.p);
    for (int i = 0; i < n; i++) {
        if (parts[i].bytes) FEASM_HIP(hipMemcpyAsync(q, parts[i].dev, parts[i].bytes, hipMemcpyDeviceToHost, s));
        q += (parts[i].bytes + 15) & ~size_t(15);
    }
    FEASM_HIP(hipStreamSynchronize(s));
    q = static_cast<unsigned char *>(b.p);
    for (int i = 0; i < n; i++) {
        if (parts[i].bytes) std::memcpy(parts[i].host, q, parts[i].bytes);
        q += (parts[i].bytes + 15) & ~size_t(15);
    }
}

size_t pool_cached_bytes(int device) {
    Pool &P = pool(device);
    std::lock_guard<std::mutex> lk(P.mu);
    return P.cached_bytes;
}

}  // namespace feasm
