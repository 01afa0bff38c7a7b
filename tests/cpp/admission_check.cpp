// Host-only checks of the assembly server's admission and eviction
// bookkeeping (finetools.jl_amd/csrc/sched.hpp): refusal of requests larger
// than the budget, admission by memory, LRU eviction of idle residents only,
// withdrawal (cancellation) of queued requests, and re-assemblies that wait
// for their mesh.  Prints "ok" and exits 0 when every check holds.
#include <cstdio>
#include <cstdlib>

#include "sched.hpp"

struct Req {
    bool reassembly = false;
    int64_t mesh_id = -1;
    size_t est = 0;
};

#define CHECK(c)                                                     \
    do {                                                             \
        if (!(c)) {                                                  \
            std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
            std::exit(1);                                            \
        }                                                            \
    } while (0)

int main() {
    using feasm::Admission;
    int meshA = 0, meshB = 0, meshC = 0;
    {  // refusal at submission: a request larger than the whole budget
        Admission<Req> a;
        a.budget = 100;
        CHECK(a.fits_budget(100));
        CHECK(!a.fits_budget(101));
    }
    {  // admission by memory: the second request waits for the first to end
        Admission<Req> a;
        a.budget = 100;
        Req r1{false, 1, 70}, r2{false, 2, 60};
        a.queue = {&r1, &r2};
        std::vector<void *> ev;
        CHECK(a.pick(ev) == &r1);
        a.start(&r1);
        CHECK(a.in_use == 70 && a.running == 1);
        CHECK(a.pick(ev) == nullptr && ev.empty());  // 70 + 60 > 100, nothing to evict
        a.finish(&r1, 1.0);
        CHECK(a.in_use == 0 && a.running == 0);
        CHECK(a.pick(ev) == &r2);
    }
    {  // least recently used idle residents are evicted first, and only as many as needed
        Admission<Req> a;
        a.budget = 100;
        a.residents[10] = {&meshA, 40, 0, 1.0};
        a.residents[11] = {&meshB, 40, 0, 2.0};
        a.in_use = 80;
        Req r{false, 3, 50};
        a.queue = {&r};
        std::vector<void *> ev;
        CHECK(a.pick(ev) == &r);
        CHECK(ev.size() == 1 && ev[0] == &meshA);
        CHECK(a.in_use == 40 && a.residents.count(11) == 1 && a.residents.count(10) == 0);
    }
    {  // a resident in use is never evicted; a request that still does not fit waits
        Admission<Req> a;
        a.budget = 100;
        a.residents[10] = {&meshA, 40, 0, 1.0};
        a.residents[11] = {&meshB, 40, 1, 0.5};  // running a re-assembly
        a.in_use = 80;
        Req r{false, 3, 70};
        a.queue = {&r};
        std::vector<void *> ev;
        CHECK(a.pick(ev) == nullptr);
        CHECK(ev.size() == 1 && ev[0] == &meshA && a.residents.count(11) == 1);
        CHECK(a.queue.size() == 1 && a.in_use == 40);
    }
    {  // a large request at the head does not hold back a small one behind it
        Admission<Req> a;
        a.budget = 100;
        a.in_use = 50;
        a.running = 1;
        Req big{false, 4, 90}, small{false, 5, 20};
        a.queue = {&big, &small};
        std::vector<void *> ev;
        CHECK(a.pick(ev) == &small);
        CHECK(a.queue.size() == 1 && a.queue[0] == &big);
    }
    {  // withdrawal (cancellation) of queued requests; a kept mesh's pending entry goes too
        Admission<Req> a;
        a.budget = 100;
        Req r1{false, 6, 10}, r2{false, 7, 10}, r3{false, 8, 10};
        a.pending_keep[7] = 3;
        a.queue = {&r1, &r2};
        CHECK(a.withdraw(&r2));
        CHECK(a.pending_keep.count(7) == 0 && a.queue.size() == 1);
        CHECK(!a.withdraw(&r3));  // not queued
    }
    {  // re-assemblies wait for their kept mesh, then run one at a time
        Admission<Req> a;
        a.budget = 1000;
        Req mk{false, 9, 100}, re1{true, 9, 0}, re2{true, 9, 0}, lost{true, 99, 0};
        a.pending_keep[9] = 3;
        a.queue = {&re1, &mk};
        std::vector<void *> ev;
        CHECK(a.pick(ev) == &mk);  // the re-assembly is skipped while its mesh is pending
        a.start(&mk);
        CHECK(a.pick(ev) == nullptr);
        a.finish(&mk, 3.0, &meshC);  // kept: now resident, its estimate stays in use
        CHECK(a.residents.count(9) == 1 && a.in_use == 100 && a.pending_keep.count(9) == 0);
        a.queue.push_back(&re2);
        CHECK(a.pick(ev) == &re1);
        a.start(&re1);
        CHECK(a.residents[9].users == 1);
        CHECK(a.pick(ev) == nullptr);  // one request per mesh at a time
        a.finish(&re1, 4.0);
        CHECK(a.residents[9].users == 0 && a.residents[9].last_use == 4.0);
        CHECK(a.pick(ev) == &re2);
        a.queue.push_back(&lost);  // unknown mesh: handed out at once, to fail
        CHECK(a.pick(ev) == &lost);
    }
    std::printf("ok\n");
    return 0;
}
