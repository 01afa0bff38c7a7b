// Batched assembly of many meshes in one persistent launch, optionally
// captured in a hipGraph, with a small autotuner for the tile size.
//
// Every mesh keeps its own resident pattern and result matrix; the batch
// concatenates the meshes' tile lists (grouped by element type and
// arithmetic) and runs one persistent fused kernel per group.  Replays of the
// captured graph rewrite every value (nothing is accumulated across runs), so
// no clearing -- and no memset node -- is needed inside the graph.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <fstream>
#include <thread>
#include <sstream>

#include "common.hpp"
#include "fused.hpp"

namespace feasm {
bool fused_supported(int npe, int nqp);
FormParams make_form_params(int form, int ndn, const double *coeff, int ncoeff);
}  // namespace feasm

using namespace feasm;

struct feasm_batch {
    CtxRef ctx;
    struct Item {
        feasm_mesh *mesh;
        FormParams fp;
        int rule, order, precision;
        std::unique_ptr<feasm_matrix> mat;
    };
    std::vector<Item> items;
    std::vector<BatchLaunch> launches;
    int max_nodes = 0, threads = 0;
    size_t budget = 0;
    bool graph_on = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~feasm_batch() {
        drop_graph();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

namespace {

thread_local std::string g_batch_error;

int fail(feasm_ctx *ctx, int code, const char *what) {
    g_batch_error = what;
    if (ctx) ctx->last_error = what;
    return code;
}

#define B_TRY try {
#define B_CATCH(ctxp)                                      \
    }                                                      \
    catch (const feasm::Error &e) {                        \
        return fail((ctxp), e.code, e.what());             \
    }                                                      \
    catch (const std::exception &e) {                      \
        return fail((ctxp), FEASM_EINVAL, e.what());       \
    }                                                      \
    return FEASM_OK;

TileShape shape_for(feasm_batch *b, const feasm_batch::Item &it) {
    const DevQTable &qt = device_qtable(b->ctx, it.mesh->etype, it.rule, it.order);
    TileShape s = tile_shape_for(b->ctx, it.mesh->npe, qt.nqp, it.fp.form, it.precision);
    if (b->max_nodes > 0) s.max_nodes = std::min(255, b->max_nodes);
    if (b->budget > 0) s.budget = std::min(s.budget, b->budget);
    if (b->threads == 256 || b->threads == 512 || b->threads == 1024) s.threads = b->threads;
    return s;
}

void prepare(feasm_batch *b) {
    feasm_ctx *ctx = b->ctx;
    b->drop_graph();
    b->launches.clear();
    // one launch per (element type, rule size, precision, kernel form, DOFs
    // per node, rule): the workgroups load the rule's tables once
    std::map<std::tuple<int, int, int, int, int, int, int>, std::vector<BatchMesh>> groups;
    for (auto &it : b->items) {
        feasm_mesh *m = it.mesh;
        if (!m->pattern || m->pattern->ndn != m->ndn) m->pattern = build_pattern(m, m->ndn);
    }
    // tiles of the meshes that lack them, built concurrently on the device
    // (one host thread per mesh at a time, each with a stream of its own)
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &it : b->items) FEASM_HIP(hipStreamSynchronize(it.mesh->ctx->stream));
    {
        std::vector<std::pair<feasm_mesh *, TileShape>> todo;
        for (auto &it : b->items) {
            const TileShape s = shape_for(b, it);
            bool dup = false;
            for (auto &t : todo) dup |= (t.first == it.mesh && t.second.key() == s.key());
            if (!dup && !it.mesh->tiles.count(s.key())) todo.emplace_back(it.mesh, s);
        }
        std::vector<std::shared_ptr<TileSet>> built(todo.size());
        std::vector<std::string> errs(todo.size());
        std::vector<int> codes(todo.size(), 0);
        std::atomic<size_t> next{0};
        const int nth = (int)std::max<size_t>(1, std::min<size_t>(todo.size(), std::max(1u, std::thread::hardware_concurrency())));
        std::vector<std::thread> pool;
        for (int w = 0; w < nth; w++)
            pool.emplace_back([&]() {
                const bool dev_ok = hipSetDevice(ctx->device) == hipSuccess;
                hipStream_t st = nullptr;
                const bool st_ok = dev_ok && hipStreamCreate(&st) == hipSuccess;
                for (size_t i; (i = next++) < todo.size();) {
                    try {
                        FEASM_REQUIRE(st_ok, FEASM_EHIP, "no device stream in a tile-building thread");
                        built[i] = build_tiles(todo[i].first, todo[i].second, st);
                    } catch (const Error &e) {
                        codes[i] = e.code;
                        errs[i] = e.what();
                    } catch (const std::exception &e) {
                        codes[i] = FEASM_ENOMEM;
                        errs[i] = e.what();
                    }
                }
                if (st_ok) {  // the builder leaves its stream idle (and holds no pool event on it)
                    (void)hipStreamSynchronize(st);
                    (void)hipStreamDestroy(st);
                }
            });
        for (auto &th : pool) th.join();
        for (size_t i = 0; i < todo.size(); i++) {
            if (codes[i]) throw Error(codes[i], errs[i]);
            todo[i].first->tiles.emplace(todo[i].second.key(), built[i]);
        }
    }
    for (auto &it : b->items) {
        feasm_mesh *m = it.mesh;
        const TileShape s = shape_for(b, it);
        auto tt = m->tiles.find(s.key());
        if (!it.mat) {
            it.mat.reset(new feasm_matrix);
            it.mat->ctx = ctx;
            it.mat->m = it.mat->n = m->pattern->n;
            it.mat->nnz = m->pattern->nnz;
            it.mat->colptr = m->pattern->colptr;
            it.mat->rowval = m->pattern->rowval;
            it.mat->pattern = m->pattern;
            it.mat->form_class = m->ndn;
            it.mat->nzval = dev_alloc(std::max<size_t>(it.mat->nnz * 8, 8));
        }
        BatchMesh bm{m, tt->second, it.fp, it.rule, it.order, dp<double>(it.mat->nzval)};
        groups[std::make_tuple(m->npe, s.nqp, s.rbytes, fused_kernel_form(it.fp), it.fp.ndn, it.rule, it.order)]
            .push_back(bm);
    }
    for (auto &kv : groups) {
        BatchLaunch L;
        L.npe = std::get<0>(kv.first);
        L.nqp = std::get<1>(kv.first);
        L.rbytes = std::get<2>(kv.first);
        L.kform = std::get<3>(kv.first);
        L.ndn = std::get<4>(kv.first);
        prepare_batch_launch(ctx, L, kv.second);
        b->launches.push_back(std::move(L));
    }
}

void run_direct(feasm_batch *b) {
    for (auto &L : b->launches) launch_batch(b->ctx, L);
}

void capture(feasm_batch *b) {
    feasm_ctx *ctx = b->ctx;
    b->drop_graph();
    FEASM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    try {
        run_direct(b);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    FEASM_HIP(hipStreamEndCapture(ctx->stream, &b->graph));
    FEASM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

double time_runs(feasm_batch *b, int reps) {
    feasm_ctx *ctx = b->ctx;
    FEASM_HIP(hipEventRecord(b->ev0, ctx->stream));
    for (int r = 0; r < reps; r++) run_direct(b);
    FEASM_HIP(hipEventRecord(b->ev1, ctx->stream));
    FEASM_HIP(hipEventSynchronize(b->ev1));
    float ms = 0;
    FEASM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    return ms / reps;
}

}  // namespace

extern "C" int feasm_batch_create(feasm_ctx *ctx, feasm_batch **out) {
    B_TRY
    FEASM_REQUIRE(ctx && out, FEASM_EINVAL, "null argument");
    auto *b = new feasm_batch;
    b->ctx = ctx;
    FEASM_HIP(hipEventCreate(&b->ev0));
    FEASM_HIP(hipEventCreate(&b->ev1));
    *out = b;
    B_CATCH(ctx)
}

extern "C" int feasm_batch_add(feasm_batch *b, feasm_mesh *mesh, int form, int rule, int order, const double *coeff,
                               int ncoeff, int precision) {
    B_TRY
    FEASM_REQUIRE(b && mesh, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(mesh->ctx == b->ctx, FEASM_EINVAL, "mesh belongs to another context");
    FEASM_REQUIRE(mesh->ndn > 0 && mesh->nodemajor, FEASM_ESTATE, "batched meshes need the node-major numbering");
    FEASM_REQUIRE(precision == FEASM_FP64 || precision == FEASM_MIXED, FEASM_EINVAL, "unknown precision");
    const DevQTable &qt = device_qtable(b->ctx, mesh->etype, rule, order);
    FEASM_REQUIRE(fused_supported(mesh->npe, qt.nqp), FEASM_EUNSUPPORTED,
                  "batched launch covers H8 2x2x2 Gauss and T10 4-point rules");
    feasm_batch::Item it;
    it.mesh = mesh;
    it.fp = make_form_params(form, mesh->ndn, coeff, ncoeff);
    it.rule = rule;
    it.order = order;
    it.precision = precision;
    b->items.push_back(std::move(it));
    b->launches.clear();
    b->drop_graph();
    B_CATCH(b ? b->ctx : nullptr)
}

extern "C" int feasm_batch_prepare(feasm_batch *b, int use_graph, int max_nodes, int budget_kib, int threads) {
    B_TRY
    FEASM_REQUIRE(b, FEASM_EINVAL, "null batch");
    FEASM_REQUIRE(threads == 0 || threads == 256 || threads == 512 || threads == 1024, FEASM_EINVAL,
                  "threads: 0, 256, 512 or 1024");
    b->max_nodes = max_nodes;
    b->threads = threads;
    b->budget = budget_kib > 0 ? (size_t)budget_kib * 1024 : 0;
    prepare(b);
    b->graph_on = use_graph != 0;
    if (b->graph_on) capture(b);
    FEASM_HIP(hipStreamSynchronize(b->ctx->stream));
    B_CATCH(b ? b->ctx : nullptr)
}

extern "C" int feasm_batch_run(feasm_batch *b) {
    B_TRY
    FEASM_REQUIRE(b, FEASM_EINVAL, "null batch");
    if (b->launches.empty() && !b->items.empty()) prepare(b);
    feasm_ctx *ctx = b->ctx;
    FEASM_HIP(hipEventRecord(b->ev0, ctx->stream));
    if (b->graph_on && b->exec)
        FEASM_HIP(hipGraphLaunch(b->exec, ctx->stream));
    else
        run_direct(b);
    FEASM_HIP(hipEventRecord(b->ev1, ctx->stream));
    B_CATCH(b ? b->ctx : nullptr)
}

extern "C" int feasm_batch_last_ms(feasm_batch *b, double *ms) {
    B_TRY
    FEASM_REQUIRE(b && ms, FEASM_EINVAL, "null argument");
    FEASM_HIP(hipEventSynchronize(b->ev1));
    float t = 0;
    FEASM_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = t;
    B_CATCH(b ? b->ctx : nullptr)
}

extern "C" int feasm_batch_matrix(feasm_batch *b, int64_t i, feasm_matrix **out) {
    B_TRY
    FEASM_REQUIRE(b && out, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(i >= 0 && i < (int64_t)b->items.size(), FEASM_EINVAL, "no such batch entry");
    FEASM_REQUIRE(b->items[i].mat, FEASM_ESTATE, "batch not prepared");
    *out = b->items[i].mat.get();
    B_CATCH(b ? b->ctx : nullptr)
}

extern "C" int feasm_batch_stats(feasm_batch *b, int64_t *nmeshes, int64_t *nelem, int64_t *nnz, int64_t *ntiles) {
    B_TRY
    FEASM_REQUIRE(b, FEASM_EINVAL, "null batch");
    int64_t ne = 0, z = 0, t = 0;
    for (auto &it : b->items) {
        ne += it.mesh->nelem;
        if (it.mat) z += it.mat->nnz;
    }
    for (auto &L : b->launches) t += L.ntiles;
    if (nmeshes) *nmeshes = (int64_t)b->items.size();
    if (nelem) *nelem = ne;
    if (nnz) *nnz = z;
    if (ntiles) *ntiles = t;
    B_CATCH(b ? b->ctx : nullptr)
}

// Autotuner: tries tile sizes (owned nodes per tile) and LDS budgets per
// workgroup (one or two workgroups per CU), times `reps` direct runs of each
// after one warm-up, keeps the fastest.  Results are appended to a small text
// database (one line per key: "key max_nodes budget_kib ms") when db_path is
// given, and read back from it on later calls for the same key.
extern "C" int feasm_batch_autotune(feasm_batch *b, const char *db_path, const char *key, int reps, int *best_nodes,
                                    int *best_budget_kib, int *best_threads, double *best_ms) {
    B_TRY
    FEASM_REQUIRE(b && !b->items.empty(), FEASM_ESTATE, "empty batch");
    if (reps < 1) reps = 3;
    if (db_path && key) {
        std::ifstream in(db_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ss(line);
            std::string k;
            int n = 0, kb = 0, th = 0;
            double ms = 0;
            if ((ss >> k >> n >> kb >> th >> ms) && k == key) {
                if (best_nodes) *best_nodes = n;
                if (best_budget_kib) *best_budget_kib = kb;
                if (best_threads) *best_threads = th;
                if (best_ms) *best_ms = ms;
                b->max_nodes = n;
                b->budget = (size_t)kb * 1024;
                b->threads = th;
                prepare(b);
                if (b->graph_on) capture(b);
                FEASM_HIP(hipStreamSynchronize(b->ctx->stream));
                return FEASM_OK;
            }
        }
    }
    const int nodes_c[] = {32, 64};
    const int kib_c[] = {160, 80};
    const int thr_c[] = {512, 1024};
    double best = 1e300;
    int bn = 64, bk = 160, bt = 1024;
    for (int th : thr_c)
        for (int kb : kib_c)
            for (int n : nodes_c) {
                b->max_nodes = n;
                b->budget = (size_t)kb * 1024;
                b->threads = th;
                prepare(b);
                run_direct(b);  // warm-up
                const double ms = time_runs(b, reps);
                if (ms < best) {
                    best = ms;
                    bn = n;
                    bk = kb;
                    bt = th;
                }
            }
    b->max_nodes = bn;
    b->budget = (size_t)bk * 1024;
    b->threads = bt;
    prepare(b);
    if (b->graph_on) capture(b);
    FEASM_HIP(hipStreamSynchronize(b->ctx->stream));
    if (db_path && key) {
        std::ofstream o(db_path, std::ios::app);
        if (o) o << key << " " << bn << " " << bk << " " << bt << " " << best << "\n";
    }
    if (best_nodes) *best_nodes = bn;
    if (best_budget_kib) *best_budget_kib = bk;
    if (best_threads) *best_threads = bt;
    if (best_ms) *best_ms = best;
    B_CATCH(b ? b->ctx : nullptr)
}

extern "C" int feasm_batch_destroy(feasm_batch *b) {
    if (!b) return FEASM_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return FEASM_OK;
}
