// Assembly server: continuous batching of assembly requests on one GPU.
//
// Requests (a new mesh to assemble, or a re-assembly of a resident mesh with
// new material values) are queued as they arrive.  A fixed set of worker
// threads, each with its own HIP stream, take the oldest admissible request
// as soon as they are free: a request is admitted when its estimated device
// memory fits the server's HBM budget (resident meshes not in use are
// evicted, least recently used first, to make room; sched.hpp).  A request
// that could never fit the budget is refused at submission.  Work of
// different requests therefore overlaps on the device, and a request that
// arrives while others run starts as soon as a worker and memory free up.
// A queued request can be cancelled; a running one finishes its device work
// and is then reported cancelled, its result dropped.
//
// Results are deterministic: every kernel sums in a fixed order, so a request
// gives bitwise the same matrix alone or in any mix.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "common.hpp"
#include "fused.hpp"
#include "sched.hpp"

namespace feasm {
bool fused_supported(int npe, int nqp);
FormParams make_form_params(int form, int ndn, const double *coeff, int ncoeff);
void numeric_into_ctx(feasm_ctx *ctx, feasm_mesh *m, const FormParams &fp, int rule, int order, int precision,
                      double *nzval);
}  // namespace feasm

using namespace feasm;

extern "C" int feasm_ctx_create(int device, feasm_ctx **out);
extern "C" int feasm_ctx_destroy(feasm_ctx *ctx);

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    bool reassembly = false;
    // new mesh
    int etype = 0, npe = 0, ndn = 0;
    int64_t nelem = 0, nnodes = 0;
    std::vector<int32_t> conn0;
    std::vector<double> xyz;
    // borrowed submissions: the caller's arrays, converted by the worker
    const int64_t *bconn = nullptr;
    const double *bxyz = nullptr;
    bool keep = false;
    // both
    int64_t mesh_id = -1;
    FormParams fp;
    int rule = 0, order = 0, precision = 0;
    size_t est = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = 0;  // 0 queued, 1 running, 2 done, < 0 failed (feasm status)
    bool cancel = false;  // cancelled while running: reported cancelled when its work ends
    std::string error;
    std::unique_ptr<feasm_matrix> result;
};

// A mesh's first kTileAfter assemblies take the staged path: building the
// fused kernel's tiles (about 10-20 ms per million H8 elements on the device)
// costs more than the staged path saves on one assembly (about 2.4 ms per
// million), so tiles pay off only for meshes re-assembled several times.
constexpr int kTileAfter = 4;

// Device bytes a request may hold: pattern and values always, plus the staged
// element blocks and contribution lists or, for a resident mesh, whichever
// is larger of those and the device tile builder's peak (slots, packed blobs
// and its per-node work arrays).
size_t estimate_bytes(int npe, int ndn, int64_t nelem, int64_t nnodes, bool keep, int precision) {
    const double deg = (npe == 8) ? 27.0 : 26.0;  // upper bound of the mean node degree of block meshes
    const double nb = deg * (double)nnodes, nnz = nb * ndn * ndn;
    const double npairs = npe * (npe + 1) / 2.0;
    const double rb = (precision == FEASM_MIXED) ? 4.0 : 8.0;
    double b = 4.0 * npe * nelem * 3  // connectivity, node->element map
               + 48.0 * nnodes        // coordinates, pointers, DOFs
               + 8.0 * nb             // adjacency, block columns
               + 12.0 * nnz           // rowval + nzval
               + 8.0 * nnodes * ndn;  // colptr
    const double staged = npairs * 9.0 * rb * nelem + 4.0 * npe * npe * nelem + 8.0 * nb;  // stage, contributions
    b += keep ? std::max(staged, 1400.0 * nnodes) : staged;
    return (size_t)(1.2 * b) + (8u << 20);
}

}  // namespace

struct feasm_server {
    int device = 0;
    std::vector<std::thread> workers;
    std::vector<feasm_ctx *> ctxs;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    Admission<Request> adm;  // queue, residents and the memory budget, under mu
    std::map<int64_t, std::unique_ptr<Request>> requests;
    int64_t next_id = 1, next_mesh = 1;
    bool stop = false;
    int64_t completed = 0;
    std::string last_error;
    // per worker, in seconds: idle (waiting for an admissible request),
    // converting and staging new meshes' arrays, waiting at the end of a
    // request for its device work, and in run_one overall; and the requests
    // run (FEASM_SERVER_TIMING=1 prints them to stderr when the server ends)
    struct WorkerTimes {
        double idle = 0, convert = 0, sync = 0, run = 0;
        int64_t n = 0;
    };
    std::vector<WorkerTimes> times;

    // evicted meshes are released outside the lock: no request uses them
    // (every request synchronises its stream before it finishes) and the
    // memory pool defers reuse of their blocks past the streams' pending work
    static void release_meshes(std::vector<void *> &meshes) {
        for (void *p : meshes) delete static_cast<feasm_mesh *>(p);
        meshes.clear();
    }

    // Pinned host memory for the mesh uploads (a copy from pageable memory
    // goes through the runtime's staging path).  The buffers are shared by
    // every worker of every server in the process: a request takes the
    // smallest free one that fits and gives it back when its copies have
    // completed, so a large mesh on any worker reuses the buffer of an
    // earlier large one (pinning tens of megabytes takes milliseconds).  New
    // buffers are sized to a power of two; they are kept until the process
    // ends (the HIP runtime may be gone by then, so they are never freed).
    struct StagePool {
        std::mutex mu;
        std::multimap<size_t, void *> free;  // by capacity
    };
    static StagePool &stage_pool() {
        static StagePool *p = new StagePool;
        return *p;
    }
    struct Stage {
        void *p = nullptr;
        size_t bytes = 0;
        void *get(size_t n) {
            put();
            StagePool &sp = stage_pool();
            {
                std::lock_guard<std::mutex> lk(sp.mu);
                auto it = sp.free.lower_bound(n);
                if (it != sp.free.end()) {
                    bytes = it->first;
                    p = it->second;
                    sp.free.erase(it);
                    return p;
                }
            }
            size_t cap = 1 << 20;
            while (cap < n) cap <<= 1;
            FEASM_HIP(hipHostMalloc(&p, cap, hipHostMallocDefault));
            bytes = cap;
            return p;
        }
        void put() {  // the copies from the buffer have completed
            if (!p) return;
            StagePool &sp = stage_pool();
            std::lock_guard<std::mutex> lk(sp.mu);
            sp.free.emplace(bytes, p);
            p = nullptr;
            bytes = 0;
        }
        ~Stage() { put(); }
    };

    // Helper threads that stage large uploads, counted process-wide so that
    // busy workers do not start more threads than the host has cores for.
    static std::atomic<int> &helpers_busy() {
        static std::atomic<int> n{0};
        return n;
    }
    static int reserve_helpers(int want) {
        static const int cap = std::max(2, (int)std::thread::hardware_concurrency() / 2);
        std::atomic<int> &busy = helpers_busy();
        int got = 0;
        while (got < want) {
            int cur = busy.load();
            if (cur >= cap) break;
            if (busy.compare_exchange_weak(cur, cur + 1)) got++;
        }
        return got;
    }
    static void release_helpers(int n) {
        if (n) helpers_busy() -= n;
    }

    void work(int w) {
        feasm_ctx *ctx = ctxs[w];
        (void)hipSetDevice(device);
        // every buffer this worker releases was last used on its own stream
        // (a resident mesh's requests synchronise before they finish), so the
        // memory pool fences that stream alone
        pool_set_thread_fence(device, ctx->stream);
        Stage stage;
        for (;;) {
            Request *r = nullptr;
            std::vector<void *> evicted;
            bool quit = false;
            const double t_wait = now_s();
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (stop) {
                        quit = true;
                        break;
                    }
                    r = adm.pick(evicted);
                    if (r || !evicted.empty()) break;
                    cv_work.wait(lk);
                }
                if (r) {
                    r->status = 1;
                    r->t_start = now_s();
                    adm.start(r);
                }
            }
            const double t_picked = now_s();
            times[w].idle += t_picked - t_wait;
            release_meshes(evicted);
            if (quit) return;
            if (!r) continue;
            run_one(ctx, r, stage, times[w]);  // publishes the outcome and the bookkeeping under one lock
            stage.put();                       // its uploads have completed
            times[w].run += now_s() - t_picked;
            times[w].n++;
            cv_work.notify_all();
            cv_done.notify_all();
        }
    }

    // outcome of a request whose work has ended; called with mu held
    void publish_locked(Request *r, int status, const std::string &err, std::unique_ptr<feasm_matrix> M,
                        feasm_mesh *keep_mesh) {
        if (r->cancel) {
            status = FEASM_ECANCELLED;
            M.reset();
            if (keep_mesh) {  // a cancelled request keeps nothing resident
                delete keep_mesh;
                keep_mesh = nullptr;
            }
        }
        r->status = status;
        r->error = (status == FEASM_ECANCELLED) ? std::string("request cancelled") : err;
        r->result = std::move(M);
        r->t_done = now_s();
        adm.finish(r, r->t_done, keep_mesh);
        completed++;
        r->conn0.clear();
        r->conn0.shrink_to_fit();
        r->xyz.clear();
        r->xyz.shrink_to_fit();
    }

    void run_one(feasm_ctx *ctx, Request *r, Stage &stage, WorkerTimes &wt) {
        std::unique_ptr<feasm_mesh> owned;
        try {
            feasm_mesh *m = nullptr;
            if (r->reassembly) {
                std::lock_guard<std::mutex> lk(mu);
                auto it = adm.residents.find(r->mesh_id);
                FEASM_REQUIRE(it != adm.residents.end(), FEASM_ESTATE, "resident mesh was evicted or never existed");
                m = static_cast<feasm_mesh *>(it->second.mesh);
                // one request at a time runs on a resident mesh: its pattern,
                // contribution and tile builders use this worker's context
                m->ctx = ctx;
            } else {
                owned.reset(new feasm_mesh);
                m = owned.get();
                m->ctx = ctx;
                m->etype = r->etype;
                m->npe = r->npe;
                m->nelem = r->nelem;
                m->nnodes = r->nnodes;
                // the arrays go through the worker's pinned staging memory;
                // borrowed ones are converted and checked here, in the worker
                const double t_conv = now_s();
                const size_t cb = (size_t)(r->nelem * r->npe) * 4, xb = (size_t)r->nnodes * 24;
                unsigned char *st = static_cast<unsigned char *>(stage.get(((cb + 15) & ~size_t(15)) + xb + 16));
                int32_t *sc = reinterpret_cast<int32_t *>(st);
                double *sx = reinterpret_cast<double *>(st + ((cb + 15) & ~size_t(15)));
                m->conn = dev_alloc(std::max<size_t>(cb, 4));
                m->xyz = dev_alloc(std::max<size_t>(xb, 8));
                // staged and copied in pieces of 2 MiB: each piece's copy runs
                // while the next one is staged (a copy issued before the range
                // check fails only fills the mesh's own buffer); the pieces of
                // a large mesh are shared out to helper threads
                constexpr int64_t kPiece = 1 << 19;  // 4-byte entries, 2 MiB
                const int64_t nc = r->nelem * r->npe, nn = r->nnodes;
                const int64_t *bc = r->bconn;
                const int32_t *oc = r->conn0.data();
                const double *x0 = r->bxyz ? r->bxyz : r->xyz.data();
                const int64_t ncp = (nc + kPiece - 1) / kPiece, nxp = (3 * nn + kPiece / 2 - 1) / (kPiece / 2);
                int32_t *dconn = dp<int32_t>(m->conn);
                double *dxyz = dp<double>(m->xyz);
                hipStream_t cs = ctx->stream;
                // piece p < ncp: connectivity (returns its range-check flag); then coordinates
                auto piece = [&](int64_t p) -> int64_t {
                    if (p < ncp) {
                        const int64_t i0 = p * kPiece, i1 = std::min(nc, i0 + kPiece);
                        int64_t bad = 0;
                        if (bc) {
                            // converted and range-checked in one branch-free pass
                            for (int64_t i = i0; i < i1; i++) {
                                const int64_t v = bc[i];
                                bad |= (v < 1) | (v > nn);
                                sc[i] = (int32_t)(v - 1);
                            }
                        } else {
                            std::memcpy(sc + i0, oc + i0, (size_t)(i1 - i0) * 4);
                        }
                        FEASM_HIP(hipMemcpyAsync(dconn + i0, sc + i0, (size_t)(i1 - i0) * 4, hipMemcpyHostToDevice, cs));
                        return bad;
                    }
                    const int64_t i0 = (p - ncp) * (kPiece / 2), i1 = std::min(3 * nn, i0 + kPiece / 2);
                    std::memcpy(sx + i0, x0 + i0, (size_t)(i1 - i0) * 8);
                    FEASM_HIP(hipMemcpyAsync(dxyz + i0, sx + i0, (size_t)(i1 - i0) * 8, hipMemcpyHostToDevice, cs));
                    return 0;
                };
                const int64_t np = ncp + nxp;
                // up to three helper threads from 16 MiB on, as many as the
                // process-wide helper budget has free (half the host's cores)
                const int nh = np >= 8 ? reserve_helpers(3) : 0;
                struct HelperRelease {
                    int n;
                    ~HelperRelease() { release_helpers(n); }
                } helper_release{nh};
                int64_t bad = 0;
                if (nh == 0) {
                    for (int64_t p = 0; p < np; p++) bad |= piece(p);
                } else {
                    // contiguous ranges of pieces, one per thread (the worker takes the first)
                    int64_t hbad[4] = {0, 0, 0, 0};
                    int herr[4] = {0, 0, 0, 0};
                    std::string hmsg[4];
                    const int dev = device;
                    auto run = [&](int t) {
                        try {
                            if (t > 0) FEASM_HIP(hipSetDevice(dev));
                            const int64_t p0 = np * t / (nh + 1), p1 = np * (t + 1) / (nh + 1);
                            for (int64_t p = p0; p < p1; p++) hbad[t] |= piece(p);
                        } catch (const feasm::Error &e) {
                            herr[t] = e.code;
                            hmsg[t] = e.what();
                        } catch (const std::exception &e) {
                            herr[t] = FEASM_EINVAL;
                            hmsg[t] = e.what();
                        }
                    };
                    std::vector<std::thread> th;
                    try {
                        for (int t = 1; t <= nh; t++) th.emplace_back(run, t);
                    } catch (...) {  // no thread: the worker stages the helpers' ranges as well
                    }
                    run(0);
                    for (int t = (int)th.size() + 1; t <= nh; t++) run(t);
                    for (auto &x : th) x.join();
                    for (int t = 0; t <= nh; t++) {
                        if (herr[t]) throw Error(herr[t], hmsg[t]);
                        bad |= hbad[t];
                    }
                }
                FEASM_REQUIRE(!bad, FEASM_EINVAL, "connectivity refers to a missing node");
                r->bconn = nullptr;
                r->bxyz = nullptr;
                wt.convert += now_s() - t_conv;
                m->ndn = r->ndn;
                m->nalldofs = r->nnodes * r->ndn;
                m->nodemajor = true;
                // the first assembly takes the staged path: the adjacency pass
                // writes its contribution lists too; its read-back waits for
                // the uploads
                m->pattern = build_pattern(m, m->ndn, true);
            }
            std::unique_ptr<feasm_matrix> M(new feasm_matrix);
            M->ctx = ctx;
            M->m = M->n = m->pattern->n;
            M->nnz = m->pattern->nnz;
            M->colptr = m->pattern->colptr;
            M->rowval = m->pattern->rowval;
            M->pattern = m->pattern;
            M->form_class = m->ndn;
            M->nzval = dev_alloc(std::max<size_t>(M->nnz * 8, 8));
            // staged for a mesh's first kTileAfter assemblies (all on the
            // device); later ones build the tiles once (on the device) and
            // drop the staged buffers -- when the element type and rule have a
            // fused kernel at all
            const bool fusable = fused_supported(m->npe, device_qtable(ctx, m->etype, r->rule, r->order).nqp);
            m->numeric_path = (m->assemblies < kTileAfter || !fusable) ? 1 : 0;
            if (m->assemblies == kTileAfter && fusable) {
                m->stage.reset();
                m->pattern->cptr.reset();
                m->pattern->contrib.reset();
            }
            m->assemblies++;
            numeric_into_ctx(ctx, m, r->fp, r->rule, r->order, r->precision, dp<double>(M->nzval));
            const double t_sync = now_s();
            FEASM_HIP(hipStreamSynchronize(ctx->stream));
            wt.sync += now_s() - t_sync;
            std::lock_guard<std::mutex> lk(mu);
            publish_locked(r, 2, std::string(), std::move(M), r->keep ? owned.release() : nullptr);
        } catch (const feasm::Error &e) {
            (void)hipStreamSynchronize(ctx->stream);  // uploads from the staging buffer may still run
            std::lock_guard<std::mutex> lk(mu);
            publish_locked(r, e.code, e.what(), nullptr, nullptr);
        } catch (const std::exception &e) {
            (void)hipStreamSynchronize(ctx->stream);
            std::lock_guard<std::mutex> lk(mu);
            publish_locked(r, FEASM_EINVAL, e.what(), nullptr, nullptr);
        }
    }

    int64_t enqueue(std::unique_ptr<Request> r) {
        std::lock_guard<std::mutex> lk(mu);
        r->id = next_id++;
        r->t_submit = now_s();
        Request *p = r.get();
        requests[p->id] = std::move(r);
        adm.queue.push_back(p);
        cv_work.notify_one();
        return p->id;
    }
};

namespace {
thread_local std::string g_srv_error;
int sfail(feasm_server *s, int code, const char *what) {
    g_srv_error = what;
    if (s) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->last_error = what;
    }
    return code;
}
#define S_TRY try {
#define S_CATCH(sp)                                  \
    }                                                \
    catch (const feasm::Error &e) {                  \
        return sfail((sp), e.code, e.what());        \
    }                                                \
    catch (const std::exception &e) {                \
        return sfail((sp), FEASM_EINVAL, e.what());  \
    }                                                \
    return FEASM_OK;
}  // namespace

extern "C" int feasm_server_create(int device, int workers, double hbm_fraction, feasm_server **out) {
    S_TRY
    FEASM_REQUIRE(out && workers >= 1 && workers <= 16, FEASM_EINVAL, "1..16 workers");
    FEASM_REQUIRE(hbm_fraction > 0 && hbm_fraction <= 0.95, FEASM_EINVAL, "hbm_fraction in (0, 0.95]");
    std::unique_ptr<feasm_server> s(new feasm_server);
    s->device = device;
    FEASM_HIP(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    FEASM_HIP(hipMemGetInfo(&fr, &tot));
    // blocks the library's memory pool holds for reuse are free to the
    // server's requests too
    s->adm.budget = (size_t)(hbm_fraction * (double)(fr + pool_cached_bytes(device)));
    for (int w = 0; w < workers; w++) {
        feasm_ctx *c = nullptr;
        const int rc = feasm_ctx_create(device, &c);
        FEASM_REQUIRE(rc == FEASM_OK, rc, "could not create a worker context");
        s->ctxs.push_back(c);
    }
    feasm_server *sp = s.get();
    s->times.resize(workers);
    for (int w = 0; w < workers; w++) s->workers.emplace_back([sp, w] { sp->work(w); });
    *out = s.release();
    S_CATCH(nullptr)
}

static int submit_mesh(feasm_server *s, bool borrowed, int elem_type, int64_t nelem, int npe, const int64_t *conn,
                       int64_t nnodes, const double *xyz_rowmajor, int ndn, int form, int rule, int order,
                       const double *coeff, int ncoeff, int precision, int keep_resident, int64_t *request_id,
                       int64_t *mesh_id) {
    S_TRY
    FEASM_REQUIRE(s && conn && xyz_rowmajor && request_id, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(npe == (elem_type == FEASM_H8 ? 8 : 10) && (elem_type == FEASM_H8 || elem_type == FEASM_T10),
                  FEASM_EINVAL, "element type / nodes per element");
    FEASM_REQUIRE(nelem > 0 && nnodes > 0 && nnodes < (1ll << 31) && nelem < (1ll << 25), FEASM_EUNSUPPORTED,
                  "mesh size out of range");
    FEASM_REQUIRE(precision == FEASM_FP64 || precision == FEASM_MIXED, FEASM_EINVAL, "unknown precision");
    std::unique_ptr<Request> r(new Request);
    r->etype = elem_type;
    r->npe = npe;
    r->ndn = ndn;
    r->nelem = nelem;
    r->nnodes = nnodes;
    if (borrowed) {
        r->bconn = conn;
        r->bxyz = xyz_rowmajor;
    } else {
        r->conn0.resize(nelem * npe);
        for (int64_t i = 0; i < nelem * npe; i++) {
            FEASM_REQUIRE(conn[i] >= 1 && conn[i] <= nnodes, FEASM_EINVAL, "connectivity refers to a missing node");
            r->conn0[i] = (int32_t)(conn[i] - 1);
        }
        r->xyz.assign(xyz_rowmajor, xyz_rowmajor + 3 * nnodes);
    }
    r->fp = make_form_params(form, ndn, coeff, ncoeff);
    r->rule = rule;
    r->order = order;
    r->precision = precision;
    r->keep = keep_resident != 0;
    r->est = estimate_bytes(npe, ndn, nelem, nnodes, keep_resident != 0, precision);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        // a request larger than the whole budget would wait for ever
        if (!s->adm.fits_budget(r->est))
            throw Error(FEASM_ENOMEM, "admission refused: the request needs about " +
                                          std::to_string(r->est >> 20) + " MiB of device memory, the server's budget is " +
                                          std::to_string(s->adm.budget >> 20) + " MiB");
        r->mesh_id = s->next_mesh++;
        if (r->keep) s->adm.pending_keep[r->mesh_id] = ndn;
    }
    if (mesh_id) *mesh_id = r->mesh_id;
    *request_id = s->enqueue(std::move(r));
    S_CATCH(s)
}

extern "C" int feasm_server_submit_mesh(feasm_server *s, int elem_type, int64_t nelem, int npe, const int64_t *conn,
                                        int64_t nnodes, const double *xyz_rowmajor, int ndn, int form, int rule,
                                        int order, const double *coeff, int ncoeff, int precision,
                                        int keep_resident, int64_t *request_id, int64_t *mesh_id) {
    return submit_mesh(s, false, elem_type, nelem, npe, conn, nnodes, xyz_rowmajor, ndn, form, rule, order, coeff,
                       ncoeff, precision, keep_resident, request_id, mesh_id);
}

extern "C" int feasm_server_submit_mesh_borrowed(feasm_server *s, int elem_type, int64_t nelem, int npe,
                                                 const int64_t *conn, int64_t nnodes, const double *xyz_rowmajor,
                                                 int ndn, int form, int rule, int order, const double *coeff,
                                                 int ncoeff, int precision, int keep_resident, int64_t *request_id,
                                                 int64_t *mesh_id) {
    return submit_mesh(s, true, elem_type, nelem, npe, conn, nnodes, xyz_rowmajor, ndn, form, rule, order, coeff,
                       ncoeff, precision, keep_resident, request_id, mesh_id);
}

extern "C" int feasm_server_submit_reassembly(feasm_server *s, int64_t mesh_id, int form, int rule, int order,
                                              const double *coeff, int ncoeff, int precision, int64_t *request_id) {
    S_TRY
    FEASM_REQUIRE(s && request_id, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(precision == FEASM_FP64 || precision == FEASM_MIXED, FEASM_EINVAL, "unknown precision");
    int ndn = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->adm.residents.find(mesh_id);
        if (it != s->adm.residents.end()) {
            ndn = static_cast<feasm_mesh *>(it->second.mesh)->ndn;
        } else {
            auto pk = s->adm.pending_keep.find(mesh_id);
            FEASM_REQUIRE(pk != s->adm.pending_keep.end(), FEASM_ESTATE, "no such resident mesh (not kept or evicted)");
            ndn = pk->second;
        }
    }
    std::unique_ptr<Request> r(new Request);
    r->reassembly = true;
    r->mesh_id = mesh_id;
    r->fp = make_form_params(form, ndn, coeff, ncoeff);
    r->rule = rule;
    r->order = order;
    r->precision = precision;
    r->est = 0;
    *request_id = s->enqueue(std::move(r));
    S_CATCH(s)
}

// Wait for a request; status: 2 done, < 0 failed (the feasm status code);
// latency = completion - submission, queue = start - submission (ms).
extern "C" int feasm_server_wait(feasm_server *s, int64_t request_id, double timeout_s, int *status,
                                 double *latency_ms, double *queue_ms) {
    S_TRY
    FEASM_REQUIRE(s, FEASM_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(request_id);
    FEASM_REQUIRE(it != s->requests.end(), FEASM_EINVAL, "unknown request");
    Request *r = it->second.get();
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    while (r->status == 0 || r->status == 1) {
        if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) break;
    }
    if (status) *status = r->status;
    if (latency_ms) *latency_ms = (r->status == 2 || r->status < 0) ? 1e3 * (r->t_done - r->t_submit) : -1;
    if (queue_ms) *queue_ms = r->t_start > 0 ? 1e3 * (r->t_start - r->t_submit) : -1;
    if (r->status < 0) {
        s->last_error = r->error;
        g_srv_error = r->error;
    }
    S_CATCH(s)
}

// The result matrix of a finished request (owned by the server until
// feasm_server_release).
extern "C" int feasm_server_result(feasm_server *s, int64_t request_id, feasm_matrix **out) {
    S_TRY
    FEASM_REQUIRE(s && out, FEASM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(request_id);
    FEASM_REQUIRE(it != s->requests.end(), FEASM_EINVAL, "unknown request");
    FEASM_REQUIRE(it->second->status == 2, FEASM_ESTATE, it->second->status < 0 ? it->second->error : "not done");
    *out = it->second->result.get();
    S_CATCH(s)
}

extern "C" int feasm_server_release(feasm_server *s, int64_t request_id) {
    S_TRY
    FEASM_REQUIRE(s, FEASM_EINVAL, "null server");
    std::unique_ptr<Request> dead;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->requests.find(request_id);
        FEASM_REQUIRE(it != s->requests.end(), FEASM_EINVAL, "unknown request");
        FEASM_REQUIRE(it->second->status == 2 || it->second->status < 0, FEASM_ESTATE, "request still pending");
        dead = std::move(it->second);
        s->requests.erase(it);
    }
    // freed outside the lock; the result's buffers were last used on its
    // worker's stream (the request synchronised it, and operations on the
    // matrix run on its context's stream), so only that stream is fenced
    if (dead->result && dead->result->ctx) {
        ScopedThreadFence fence(s->device, dead->result->ctx->stream);
        dead.reset();
    }
    S_CATCH(s)
}

// Cancel a request: a queued one is withdrawn at once, a running one
// finishes its device work and is then reported cancelled (its result
// dropped, its mesh not kept); either way feasm_server_wait gives status
// FEASM_ECANCELLED.  A request that already ended is FEASM_ESTATE.
extern "C" int feasm_server_cancel(feasm_server *s, int64_t request_id) {
    S_TRY
    FEASM_REQUIRE(s, FEASM_EINVAL, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->requests.find(request_id);
        FEASM_REQUIRE(it != s->requests.end(), FEASM_EINVAL, "unknown request");
        Request *r = it->second.get();
        FEASM_REQUIRE(r->status == 0 || r->status == 1, FEASM_ESTATE, "request already finished");
        if (r->status == 0) {
            FEASM_REQUIRE(s->adm.withdraw(r), FEASM_ESTATE, "request not in the queue");
            r->status = FEASM_ECANCELLED;
            r->error = "request cancelled";
            r->t_done = now_s();
            s->completed++;
        } else {
            r->cancel = true;
        }
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
    S_CATCH(s)
}

extern "C" int feasm_server_stats(feasm_server *s, int64_t *queued, int64_t *running, int64_t *completed,
                                  int64_t *residents, int64_t *bytes_in_use, int64_t *budget) {
    S_TRY
    FEASM_REQUIRE(s, FEASM_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = (int64_t)s->adm.queue.size();
    if (running) *running = s->adm.running;
    if (completed) *completed = s->completed;
    if (residents) *residents = (int64_t)s->adm.residents.size();
    if (bytes_in_use) *bytes_in_use = (int64_t)s->adm.in_use;
    if (budget) *budget = (int64_t)s->adm.budget;
    S_CATCH(s)
}

extern "C" const char *feasm_server_last_error(feasm_server *s) {
    return s ? s->last_error.c_str() : g_srv_error.c_str();
}

extern "C" int feasm_server_destroy(feasm_server *s) {
    if (!s) return FEASM_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    if (std::getenv("FEASM_SERVER_TIMING"))
        for (size_t w = 0; w < s->times.size(); w++)
            std::fprintf(stderr,
                         "feasm server worker %zu: %lld requests, idle %.1f ms, run %.1f ms (convert %.1f ms, "
                         "final wait %.1f ms)\n",
                         w, (long long)s->times[w].n, 1e3 * s->times[w].idle, 1e3 * s->times[w].run,
                         1e3 * s->times[w].convert, 1e3 * s->times[w].sync);
    s->requests.clear();
    for (auto &kv : s->adm.residents) delete static_cast<feasm_mesh *>(kv.second.mesh);
    s->adm.residents.clear();
    for (auto *c : s->ctxs) feasm_ctx_destroy(c);
    delete s;
    return FEASM_OK;
}
