// Admission and eviction bookkeeping of the assembly server (server.cpp),
// free of device calls so that it can be exercised on the host alone.
//
// A request is admitted when its device-memory estimate fits the budget
// together with what is in use; resident meshes that no request is using are
// evicted, least recently used first, to make room.  A request whose estimate
// exceeds the whole budget can never run and is refused at submission.
// The caller holds its own lock around every member function, and releases
// the device memory of evicted meshes after dropping that lock.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <map>
#include <vector>

namespace feasm {

struct ResidentEntry {
    void *mesh = nullptr;  // owned by the caller; handed back when evicted
    size_t bytes = 0;
    int users = 0;         // requests running on it (at most one)
    double last_use = 0;
};

// Req needs: bool reassembly; int64_t mesh_id; size_t est.
template <class Req>
struct Admission {
    size_t budget = 0, in_use = 0;
    int running = 0;
    std::deque<Req *> queue;
    std::map<int64_t, ResidentEntry> residents;
    std::map<int64_t, int> pending_keep;  // kept meshes still being assembled: mesh id -> DOFs per node

    bool fits_budget(size_t est) const { return est <= budget; }

    // The oldest admissible request among the first `window` queued (so that
    // a large request does not hold back small ones behind it), removed from
    // the queue.  Meshes evicted to make room are appended to `evicted`.
    Req *pick(std::vector<void *> &evicted, size_t window = 8) {
        const size_t w = std::min<size_t>(queue.size(), window);
        for (size_t k = 0; k < w; k++) {
            Req *r = queue[k];
            if (r->reassembly) {
                auto it = residents.find(r->mesh_id);
                if (it == residents.end()) {
                    if (pending_keep.count(r->mesh_id)) continue;  // its mesh is still being assembled
                    queue.erase(queue.begin() + k);                 // evicted or unknown: it will fail
                    return r;
                }
                if (it->second.users > 0) continue;  // one request per mesh at a time
            }
            const size_t need = r->est;
            if (in_use + need > budget) evict(need, evicted);
            if (in_use + need <= budget) {
                queue.erase(queue.begin() + k);
                return r;
            }
        }
        return nullptr;
    }

    // least recently used idle residents, until `need` more bytes fit
    void evict(size_t need, std::vector<void *> &evicted) {
        while (in_use + need > budget) {
            int64_t victim = -1;
            double oldest = 1e300;
            for (auto &kv : residents)
                if (kv.second.users == 0 && kv.second.last_use < oldest) {
                    oldest = kv.second.last_use;
                    victim = kv.first;
                }
            if (victim < 0) return;
            auto it = residents.find(victim);
            in_use -= it->second.bytes;
            evicted.push_back(it->second.mesh);
            residents.erase(it);
        }
    }

    void start(Req *r) {
        running++;
        if (!r->reassembly) {
            in_use += r->est;
        } else {
            auto it = residents.find(r->mesh_id);
            if (it != residents.end()) it->second.users++;
        }
    }

    // a finished request: a new mesh's estimate is released, and the mesh
    // becomes resident when `keep_mesh` is given (its estimate stays in use)
    void finish(Req *r, double now, void *keep_mesh = nullptr) {
        running--;
        if (r->reassembly) {
            auto it = residents.find(r->mesh_id);
            if (it != residents.end()) {
                it->second.users--;
                it->second.last_use = now;
            }
            return;
        }
        in_use -= r->est;
        pending_keep.erase(r->mesh_id);
        if (keep_mesh) {
            ResidentEntry e;
            e.mesh = keep_mesh;
            e.bytes = r->est;
            e.last_use = now;
            in_use += e.bytes;
            residents[r->mesh_id] = e;
        }
    }

    // a queued request withdrawn before it started; false if not queued
    bool withdraw(Req *r) {
        for (size_t k = 0; k < queue.size(); k++)
            if (queue[k] == r) {
                queue.erase(queue.begin() + k);
                if (!r->reassembly) pending_keep.erase(r->mesh_id);
                return true;
            }
        return false;
    }
};

}  // namespace feasm
