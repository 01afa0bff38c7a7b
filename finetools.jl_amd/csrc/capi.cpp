// C ABI of libfeasm.so (declared in include/feasm.h).
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "fused.hpp"

using namespace feasm;

namespace feasm {
bool fused_supported(int npe, int nqp);
void numeric_into_ctx(feasm_ctx *ctx, feasm_mesh *m, const FormParams &fp, int rule, int order, int precision,
                      double *nzval);
FormParams make_form_params(int form, int ndn, const double *coeff, int ncoeff);
void gen_h8block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, int64_t *conn, double *xyz);
int64_t gen_t10block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, double amp_v, double amp_m,
                     uint64_t seed, int64_t *conn10, std::vector<double> &xyz10);
void perturb_nodes(int64_t nnodes, double *xyz, const double lo[3], const double hi[3], double amp, uint64_t seed);
}  // namespace feasm

namespace {

thread_local std::string g_last_error;

int fail(feasm_ctx *ctx, int code, const char *what) {
    g_last_error = what;
    if (ctx) ctx->last_error = what;
    return code;
}

#define API_TRY try {
#define API_CATCH(ctxp)                                              \
    }                                                                \
    catch (const feasm::Error &e) {                                  \
        return fail((ctxp), e.code, e.what());                       \
    }                                                                \
    catch (const std::bad_alloc &) {                                 \
        return fail((ctxp), FEASM_ENOMEM, "host out of memory");     \
    }                                                                \
    catch (const std::exception &e) {                                \
        return fail((ctxp), FEASM_EINVAL, e.what());                 \
    }                                                                \
    return FEASM_OK;

struct Timer {
    feasm_ctx *ctx;
    explicit Timer(feasm_ctx *c) : ctx(c) {}
    void mark(int i) {
        if (ctx->timing) FEASM_HIP(hipEventRecord(ctx->ring_event(i), ctx->stream));
    }
    void done() {
        if (ctx->timing) ctx->ring_advance();
    }
};

}  // namespace

FormParams feasm::make_form_params(int form, int ndn, const double *coeff, int ncoeff) {
    FormParams fp;
    std::memset(&fp, 0, sizeof(fp));
    fp.form = form;
    fp.ndn = ndn;
    FEASM_REQUIRE(coeff != nullptr || ncoeff == 0, FEASM_EINVAL, "null coefficient array");
    switch (form) {
        case FEASM_FORM_DIFFUSION_ISO:
            FEASM_REQUIRE(ncoeff == 1, FEASM_EINVAL, "isotropic diffusion takes one coefficient");
            FEASM_REQUIRE(ndn == 1, FEASM_EINVAL, "diffusion needs 1 DOF per node");
            break;
        case FEASM_FORM_DIFFUSION_MAT:
            FEASM_REQUIRE(ncoeff == 9, FEASM_EINVAL, "matrix diffusion takes a 3x3 coefficient");
            FEASM_REQUIRE(ndn == 1, FEASM_EINVAL, "diffusion needs 1 DOF per node");
            break;
        case FEASM_FORM_LIN_ELASTIC:
            FEASM_REQUIRE(ncoeff == 36, FEASM_EINVAL, "linear elasticity takes a 6x6 material matrix");
            FEASM_REQUIRE(ndn == 3, FEASM_EINVAL, "Wrong dimensions");  // blmat! (DeforModelRedModule.jl:457)
            break;
        case FEASM_FORM_DOT:
            FEASM_REQUIRE(ndn >= 1 && ndn <= 3, FEASM_EUNSUPPORTED, "dot form supports 1..3 DOFs per node");
            FEASM_REQUIRE(ncoeff == ndn * ndn, FEASM_EINVAL, "dot form takes an ndn x ndn coefficient");
            break;
        default: throw Error(FEASM_EINVAL, "unknown form");
    }
    // the reference keeps the upper triangle of each element matrix
    // (add_btdb_ut_only!/add_gkgt_ut_only! + complete_lt!,
    // src/MatrixUtilityModule.jl:120-216), which for a non-symmetric D or
    // kappa is not the matrix the node-block kernels form: such coefficients
    // are refused
    if (form == FEASM_FORM_DIFFUSION_MAT || form == FEASM_FORM_LIN_ELASTIC) {
        const int d = form == FEASM_FORM_LIN_ELASTIC ? 6 : 3;
        for (int i = 0; i < d; i++)
            for (int j = i + 1; j < d; j++)
                FEASM_REQUIRE(coeff[i * d + j] == coeff[j * d + i], FEASM_EUNSUPPORTED,
                              "the material matrix must be symmetric");
    }
    for (int i = 0; i < ncoeff; i++) fp.coef[i] = coeff[i];
    return fp;
}

namespace {

void check_rule(const feasm_mesh *m, int rule) {
    if (m->etype == FEASM_H8)
        FEASM_REQUIRE(rule == FEASM_RULE_GAUSS, FEASM_EUNSUPPORTED, "H8 elements take a Gauss rule");
    else
        FEASM_REQUIRE(rule == FEASM_RULE_TET, FEASM_EUNSUPPORTED, "T10 elements take a tetrahedral rule");
}

// the DOF checks of assemble! (src/AssemblyModule.jl:265-273) for a general
// numbering, in the order the reference meets them: element by element,
// columns then rows
void check_dofs_like_assemble(const feasm_mesh *m, int ndn, int64_t n, const std::vector<int64_t> &dofs1) {
    std::vector<int32_t> conn(m->nelem * m->npe);
    FEASM_HIP(hipMemcpy(conn.data(), m->conn->p, conn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    // rows == cols == the element's DOFs: the first column is checked, then
    // every row of that column; later columns repeat rows already checked
    for (int64_t e = 0; e < m->nelem; e++) {
        const int64_t d0 = dofs1[(int64_t)conn[e * m->npe] * ndn];
        if (d0 < 1) throw Error(FEASM_EDOF, "Column degree of freedom < 1");
        if (d0 > n) throw Error(FEASM_EDOF, "Column degree of freedom > size");
        for (int a = 0; a < m->npe; a++)
            for (int k = 0; k < ndn; k++) {
                const int64_t d = dofs1[(int64_t)conn[e * m->npe + a] * ndn + k];
                if (d < 1) throw Error(FEASM_EDOF, "Row degree of freedom < 1");
                if (d > n) throw Error(FEASM_EDOF, "Row degree of freedom > size");
            }
    }
}

// staged: the staged path will integrate (the mesh asks for it, or the rule
// has no fused kernel), so the adjacency sort also writes its contribution
// lists (a fused-path mesh does not need them)
void ensure_pattern(feasm_mesh *m, bool staged = false) {
    if (!m->pattern || m->pattern->ndn != m->ndn) m->pattern = build_pattern(m, m->ndn, staged);
}

void *ensure_stage(feasm_mesh *m, int form, int precision) {
    const size_t bytes = (size_t)m->nelem * npairs(m->npe) * stage_width(form) *
                         (precision == FEASM_MIXED ? sizeof(float) : sizeof(double));
    if (!m->stage || m->stage->bytes < bytes) {
        m->stage.reset();
        m->stage = dev_alloc(std::max<size_t>(bytes, 256));
    }
    return m->stage->p;
}

}  // namespace

TileShape feasm::mesh_tile_shape(feasm_mesh *m, int nqp, int form, int precision) {
    TileShape shape = tile_shape_for(m->ctx, m->npe, nqp, form, precision);
    if (m->tile_nodes > 0) shape.max_nodes = std::min(255, m->tile_nodes);
    if (m->tile_kib > 0) shape.budget = std::min(shape.budget, (size_t)m->tile_kib * 1024);
    if (m->tile_threads == 256 || m->tile_threads == 512 || m->tile_threads == 1024) shape.threads = m->tile_threads;
    return shape;
}

void feasm::numeric_into_ctx(feasm_ctx *ctx, feasm_mesh *m, const FormParams &fp, int rule, int order,
                             int precision, double *nzval) {
    const DevQTable &qt = device_qtable(ctx, m->etype, rule, order);
    if (m->nelem == 0 || m->pattern->nnz == 0) return;  // no element: an empty matrix, nothing to integrate
    const bool fused = m->numeric_path != 1 && fused_supported(m->npe, qt.nqp);
    FEASM_REQUIRE(fused || m->numeric_path != 2, FEASM_EUNSUPPORTED,
                  "no fused kernel for this element type and rule");
    std::shared_ptr<TileSet> tiles;
    if (fused) {
        const TileShape shape = mesh_tile_shape(m, qt.nqp, fp.form, precision);
        auto it = m->tiles.find(shape.key());
        if (it == m->tiles.end()) {
            try {
                tiles = build_tiles(m, shape);
            } catch (const Error &e) {
                // meshes beyond the tile index widths take the staged path
                // unless the fused one was asked for (a null entry remembers it)
                if (e.code != FEASM_EUNSUPPORTED || m->numeric_path == 2) throw;
            }
            m->tiles.emplace(shape.key(), tiles);
        } else {
            tiles = it->second;
        }
    }
    if (tiles) {
        if (m->pattern->bmult) m->pattern->bmult.reset();  // the staged path's contribution counts: not needed
        Timer tm(ctx);
        tm.mark(0);
        launch_fused(ctx, m, *m->pattern, *tiles, qt, fp, precision, nzval);
        tm.mark(1);
        tm.mark(2);
        tm.done();
        return;
    }
    build_contributions(m, *m->pattern);
    void *stage = ensure_stage(m, fp.form, precision);
    Timer tm(ctx);
    tm.mark(0);
    launch_element_stage(ctx, m, qt, fp, precision, stage);
    tm.mark(1);
    launch_gather(ctx, m, *m->pattern, fp, precision, stage, nzval);
    tm.mark(2);
    tm.done();
}

namespace {

void numeric_into(feasm_mesh *m, const FormParams &fp, int rule, int order, int precision, double *nzval) {
    numeric_into_ctx(m->ctx, m, fp, rule, order, precision, nzval);
}

feasm_matrix *coo_path(feasm_mesh *m, const FormParams &fp, int rule, int order, int precision) {
    feasm_ctx *ctx = m->ctx;
    const DevQTable &qt = device_qtable(ctx, m->etype, rule, order);
    const int64_t kd = (int64_t)m->ndn * m->npe;
    const int64_t nt = m->nelem * kd * kd;
    DevPtr I = dev_alloc(nt * sizeof(int32_t)), J = dev_alloc(nt * sizeof(int32_t)),
           V = dev_alloc(nt * sizeof(double));
    launch_element_coo(ctx, m, qt, fp, precision, dp<int32_t>(I), dp<int32_t>(J), dp<double>(V));
    return sparse_device(ctx, nt, dp<int32_t>(I), dp<int32_t>(J), dp<double>(V), m->nalldofs, m->nalldofs);
}

}  // namespace

void feasm_ctx::timing_configure(int slots) {
    FEASM_HIP(hipStreamSynchronize(stream));
    for (auto e : ring) (void)hipEventDestroy(e);
    ring.clear();
    ring_slots = slots;
    ring_next = ring_used = 0;
    timing = slots > 0;
    ring.resize(3 * (size_t)slots);
    for (auto &e : ring) FEASM_HIP(hipEventCreate(&e));
}

void *feasm_ctx::scratch_get(size_t bytes) {
    if (!scratch || scratch->bytes < bytes) {
        if (scratch) FEASM_HIP(hipStreamSynchronize(stream));
        scratch.reset();
        scratch = dev_alloc(std::max<size_t>(bytes, 1 << 20));
    }
    return scratch->p;
}

// ======================================================================= context
extern "C" int feasm_ctx_create(int device, feasm_ctx **out) {
    API_TRY
    FEASM_REQUIRE(out, FEASM_EINVAL, "null output pointer");
    int ndev = 0;
    FEASM_HIP(hipGetDeviceCount(&ndev));
    FEASM_REQUIRE(device >= 0 && device < ndev, FEASM_EINVAL, "no such HIP device");
    FEASM_HIP(hipSetDevice(device));
    std::unique_ptr<feasm_ctx> c(new feasm_ctx);
    c->device = device;
    FEASM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    pool_register_stream(device, c->stream);
    *out = c.release();
    API_CATCH(nullptr)
}

extern "C" int feasm_pool_trim(int device, int64_t *freed_bytes) {
    API_TRY
    int ndev = 0;
    FEASM_HIP(hipGetDeviceCount(&ndev));
    FEASM_REQUIRE(device >= 0 && device < ndev, FEASM_EINVAL, "no such HIP device");
    const size_t f = pool_trim(device);
    if (freed_bytes) *freed_bytes = (int64_t)f;
    API_CATCH(nullptr)
}

void feasm_ctx_release(feasm_ctx *ctx) {
    if (--ctx->refs > 0) return;
    (void)hipStreamSynchronize(ctx->stream);
    ctx->qtables.clear();
    ctx->scratch.reset();
    for (auto e : ctx->ring) (void)hipEventDestroy(e);
    pool_unregister_stream(ctx->device, ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int feasm_ctx_destroy(feasm_ctx *ctx) {
    if (!ctx) return FEASM_OK;
    feasm_ctx_release(ctx);
    return FEASM_OK;
}

extern "C" const char *feasm_last_error(const feasm_ctx *ctx) {
    return ctx ? ctx->last_error.c_str() : g_last_error.c_str();
}

extern "C" void *feasm_ctx_stream(feasm_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" int feasm_ctx_synchronize(feasm_ctx *ctx) {
    API_TRY
    FEASM_REQUIRE(ctx, FEASM_EINVAL, "null context");
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

extern "C" int feasm_mem_info(feasm_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes) {
    API_TRY
    size_t f = 0, t = 0;
    FEASM_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    API_CATCH(ctx)
}

extern "C" int feasm_ctx_set_timing(feasm_ctx *ctx, int slots) {
    API_TRY
    FEASM_REQUIRE(ctx, FEASM_EINVAL, "null context");
    FEASM_REQUIRE(slots >= 0 && slots <= 4096, FEASM_EINVAL, "0..4096 timing slots");
    ctx->timing_configure(slots);
    API_CATCH(ctx)
}

extern "C" int feasm_timings(feasm_ctx *ctx, double *out, int max_calls, int *ncalls) {
    API_TRY
    FEASM_REQUIRE(ctx && out && ncalls, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(ctx->timing, FEASM_ESTATE, "timing is off");
    const int n = std::min(ctx->ring_used, max_calls);
    // oldest first
    const int first = (ctx->ring_next - ctx->ring_used + ctx->ring_slots) % ctx->ring_slots;
    const int skip = ctx->ring_used - n;
    for (int i = 0; i < n; i++) {
        const int s = (first + skip + i) % ctx->ring_slots;
        hipEvent_t *ev = &ctx->ring[3 * (size_t)s];
        FEASM_HIP(hipEventSynchronize(ev[2]));
        float a = 0, b = 0;
        FEASM_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        FEASM_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
        out[3 * i + 0] = a;
        out[3 * i + 1] = b;
        out[3 * i + 2] = (double)a + (double)b;
    }
    *ncalls = n;
    API_CATCH(ctx)
}

extern "C" int feasm_timings_reset(feasm_ctx *ctx) {
    if (!ctx) return fail(nullptr, FEASM_EINVAL, "null context");
    ctx->ring_next = ctx->ring_used = 0;
    return FEASM_OK;
}

// ======================================================================= meshes
static feasm_mesh *new_mesh(feasm_ctx *ctx, int etype, int64_t nelem, int npe, int64_t nnodes) {
    FEASM_REQUIRE(ctx, FEASM_EINVAL, "null context");
    FEASM_REQUIRE(etype == FEASM_H8 || etype == FEASM_T10, FEASM_EINVAL, "unknown element type");
    FEASM_REQUIRE(npe == (etype == FEASM_H8 ? 8 : 10), FEASM_EINVAL, "nodes per element does not match the type");
    FEASM_REQUIRE(nelem >= 0 && nnodes >= 0, FEASM_EINVAL, "negative size");
    FEASM_REQUIRE(nnodes < (1ll << 31), FEASM_EUNSUPPORTED, "more than 2^31 nodes");
    FEASM_REQUIRE(nelem < (1ll << 25), FEASM_EUNSUPPORTED, "more than 2^25 elements in one mesh");
    auto *m = new feasm_mesh;
    m->ctx = ctx;
    m->etype = etype;
    m->npe = npe;
    m->nelem = nelem;
    m->nnodes = nnodes;
    return m;
}

extern "C" int feasm_mesh_create(feasm_ctx *ctx, int elem_type, int64_t nelem, int npe, const int64_t *conn,
                                 int64_t nnodes, const double *xyz_colmajor, feasm_mesh **out) {
    API_TRY
    FEASM_REQUIRE(out && (conn || nelem == 0) && (xyz_colmajor || nnodes == 0), FEASM_EINVAL, "null argument");
    std::unique_ptr<feasm_mesh> m(new_mesh(ctx, elem_type, nelem, npe, nnodes));
    std::vector<int32_t> c0(nelem * npe);
    for (int64_t i = 0; i < nelem * npe; i++) {
        FEASM_REQUIRE(conn[i] >= 1 && conn[i] <= nnodes, FEASM_EINVAL, "connectivity refers to a missing node");
        c0[i] = (int32_t)(conn[i] - 1);
    }
    std::vector<double> x(3 * nnodes);
    for (int64_t n = 0; n < nnodes; n++)
        for (int k = 0; k < 3; k++) x[3 * n + k] = xyz_colmajor[n + nnodes * k];
    m->conn = dev_alloc(std::max<size_t>(c0.size() * sizeof(int32_t), 4));
    m->xyz = dev_alloc(std::max<size_t>(x.size() * sizeof(double), 8));
    FEASM_HIP(hipMemcpy(m->conn->p, c0.data(), c0.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    FEASM_HIP(hipMemcpy(m->xyz->p, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice));
    *out = m.release();
    API_CATCH(ctx)
}

extern "C" int feasm_mesh_create_device(feasm_ctx *ctx, int elem_type, int64_t nelem, int npe,
                                        const int32_t *d_conn0, int64_t nnodes, const double *d_xyz,
                                        feasm_mesh **out) {
    API_TRY
    FEASM_REQUIRE(out && d_conn0 && d_xyz, FEASM_EINVAL, "null argument");
    std::unique_ptr<feasm_mesh> m(new_mesh(ctx, elem_type, nelem, npe, nnodes));
    m->conn = dev_alloc(std::max<size_t>(nelem * npe * sizeof(int32_t), 4));
    m->xyz = dev_alloc(std::max<size_t>(nnodes * 3 * sizeof(double), 8));
    FEASM_HIP(hipMemcpyAsync(m->conn->p, d_conn0, nelem * npe * sizeof(int32_t), hipMemcpyDeviceToDevice,
                             ctx->stream));
    FEASM_HIP(hipMemcpyAsync(m->xyz->p, d_xyz, nnodes * 3 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    // the host path checks every entry; here a kernel does (an entry out of
    // range would send the pattern kernels outside their arrays)
    FEASM_REQUIRE(conn_in_range(ctx, dp<int32_t>(m->conn), nelem * npe, nnodes), FEASM_EINVAL,
                  "connectivity refers to a missing node");
    *out = m.release();
    API_CATCH(ctx)
}

extern "C" int feasm_mesh_destroy(feasm_mesh *mesh) {
    if (!mesh) return FEASM_OK;
    (void)hipStreamSynchronize(mesh->ctx->stream);
    delete mesh;
    return FEASM_OK;
}

extern "C" int feasm_mesh_set_xyz(feasm_mesh *m, const double *xyz_rowmajor) {
    API_TRY
    FEASM_REQUIRE(m && xyz_rowmajor, FEASM_EINVAL, "null argument");
    // ordered after the assemblies already queued on the context stream
    // (feasm_bilform_into returns before its kernels finish)
    FEASM_HIP(hipMemcpyAsync(m->xyz->p, xyz_rowmajor, m->nnodes * 3 * sizeof(double), hipMemcpyHostToDevice,
                             m->ctx->stream));
    for (auto &kv : m->tiles)  // the fused kernel's tiles carry copies of the coordinates
        if (kv.second) refresh_tile_coords(m, *kv.second);
    FEASM_HIP(hipStreamSynchronize(m->ctx->stream));
    API_CATCH(m ? m->ctx : nullptr)
}

static void set_dofs(feasm_mesh *m, int ndn, const std::vector<int64_t> &rowmajor1, int64_t nalldofs, bool nm) {
    std::vector<int32_t> d0(rowmajor1.size());
    for (size_t i = 0; i < d0.size(); i++) d0[i] = (int32_t)(rowmajor1[i] - 1);
    m->dofs = dev_alloc(std::max<size_t>(d0.size() * sizeof(int32_t), 4));
    FEASM_HIP(hipMemcpy(m->dofs->p, d0.data(), d0.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (m->ndn != ndn) m->pattern.reset();
    m->ndn = ndn;
    m->nalldofs = nalldofs;
    m->nodemajor = nm;
    std::vector<bool> seen(std::max<int64_t>(nalldofs, 0), false);
    m->dofs_unique = true;
    for (int32_t d : d0)
        if (d >= 0 && d < nalldofs) {
            if (seen[d]) m->dofs_unique = false;
            seen[d] = true;
        }
}

extern "C" int feasm_dofs_set(feasm_mesh *m, int ndn, const int64_t *dofnums_colmajor, int64_t nalldofs) {
    API_TRY
    FEASM_REQUIRE(m && dofnums_colmajor, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(ndn >= 1 && ndn <= 3, FEASM_EUNSUPPORTED, "1 to 3 DOFs per node");
    FEASM_REQUIRE(nalldofs >= 0 && nalldofs < (1ll << 31), FEASM_EUNSUPPORTED, "more than 2^31 DOFs");
    std::vector<int64_t> rm(m->nnodes * ndn);
    bool nm = (nalldofs == m->nnodes * ndn);
    for (int64_t n = 0; n < m->nnodes; n++)
        for (int k = 0; k < ndn; k++) {
            const int64_t d = dofnums_colmajor[n + m->nnodes * k];
            rm[n * ndn + k] = d;
            if (d != n * ndn + k + 1) nm = false;
        }
    // checked before the mesh changes: a rejected numbering leaves the mesh
    // as it was
    if (!nm) check_dofs_like_assemble(m, ndn, nalldofs, rm);
    set_dofs(m, ndn, rm, nalldofs, nm);
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_dofs_set_nodemajor(feasm_mesh *m, int ndn) {
    API_TRY
    FEASM_REQUIRE(m, FEASM_EINVAL, "null mesh");
    FEASM_REQUIRE(ndn >= 1 && ndn <= 3, FEASM_EUNSUPPORTED, "1 to 3 DOFs per node");
    std::vector<int64_t> rm(m->nnodes * ndn);
    for (int64_t i = 0; i < (int64_t)rm.size(); i++) rm[i] = i + 1;
    set_dofs(m, ndn, rm, m->nnodes * ndn, true);
    API_CATCH(m ? m->ctx : nullptr)
}

// ======================================================================= load vectors
namespace {

void linform_into(feasm_mesh *m, int rule, int order, const double *force, int nforce, double *F) {
    FEASM_REQUIRE(m && force && F, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(m->ndn > 0, FEASM_ESTATE, "set the DOF numbering first");
    FEASM_REQUIRE(nforce == m->ndn, FEASM_EINVAL, "the force needs one value per DOF of a node");
    FEASM_REQUIRE(m->dofs_unique, FEASM_EUNSUPPORTED, "load vectors need distinct DOF numbers per (node, component)");
    check_rule(m, rule);
    feasm_ctx *ctx = m->ctx;
    if (!m->n2e_ptr) build_node_to_element(m);  // the gather walks each node's elements
    const DevQTable &qt = device_qtable(ctx, m->etype, rule, order);
    double *stage = static_cast<double *>(ctx->scratch_get((size_t)m->nelem * m->npe * m->ndn * sizeof(double)));
    Timer tm(ctx);
    tm.mark(0);
    launch_linform_dot(ctx, m, qt, force, stage, F);
    tm.mark(1);
    tm.mark(2);
    tm.done();
}

}  // namespace

extern "C" int feasm_linform_dot_device(feasm_mesh *m, int rule, int order, const double *force, int nforce,
                                        double *F_device) {
    API_TRY
    linform_into(m, rule, order, force, nforce, F_device);
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_linform_dot(feasm_mesh *m, int rule, int order, const double *force, int nforce,
                                 double *F_host) {
    API_TRY
    FEASM_REQUIRE(m && F_host, FEASM_EINVAL, "null argument");
    DevPtr F = dev_alloc(std::max<size_t>((size_t)m->nalldofs * sizeof(double), 8));
    linform_into(m, rule, order, force, nforce, dp<double>(F));
    FEASM_HIP(hipMemcpyAsync(F_host, F->p, (size_t)m->nalldofs * sizeof(double), hipMemcpyDeviceToHost,
                             m->ctx->stream));
    FEASM_HIP(hipStreamSynchronize(m->ctx->stream));
    API_CATCH(m ? m->ctx : nullptr)
}

// ======================================================================= dropzeros
extern "C" int feasm_matrix_dropzeros(feasm_matrix *mat, feasm_matrix **out) {
    API_TRY
    FEASM_REQUIRE(mat && out, FEASM_EINVAL, "null argument");
    *out = drop_zeros(mat->ctx, mat);
    API_CATCH(mat ? mat->ctx.p : nullptr)
}

// ======================================================================= free-free block
extern "C" int feasm_matrix_free_block(feasm_matrix *mat, const uint8_t *is_free, feasm_matrix **out) {
    API_TRY
    FEASM_REQUIRE(mat && is_free && out, FEASM_EINVAL, "null argument");
    *out = principal_submatrix(mat->ctx, mat, is_free);
    API_CATCH(mat ? mat->ctx.p : nullptr)
}

extern "C" int feasm_matrix_ff_block(feasm_matrix *mat, int64_t row_nfreedofs, int64_t col_nfreedofs,
                                     feasm_matrix **out) {
    API_TRY
    FEASM_REQUIRE(mat && out, FEASM_EINVAL, "null argument");
    *out = leading_block(mat->ctx, mat, row_nfreedofs, col_nfreedofs);
    API_CATCH(mat ? mat->ctx.p : nullptr)
}

// ======================================================================= lumped mass
extern "C" int feasm_bilform_dot_lumped(feasm_mesh *m, int rule, int order, const double *coeff, int ncoeff,
                                        feasm_matrix **out) {
    API_TRY
    FEASM_REQUIRE(m && coeff && out, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(m->ndn > 0, FEASM_ESTATE, "set the DOF numbering first");
    FEASM_REQUIRE(ncoeff == m->ndn * m->ndn, FEASM_EINVAL, "the coefficient is an ndn x ndn matrix");
    FEASM_REQUIRE(m->dofs_unique, FEASM_EUNSUPPORTED, "lumped mass needs distinct DOF numbers per (node, component)");
    check_rule(m, rule);
    feasm_ctx *ctx = m->ctx;
    if (!m->n2e_ptr) build_node_to_element(m);
    const DevQTable &qt = device_qtable(ctx, m->etype, rule, order);
    Timer tm(ctx);
    tm.mark(0);
    feasm_matrix *M = hrz_lumped(ctx, m, qt, coeff);
    tm.mark(1);
    tm.mark(2);
    tm.done();
    *out = M;
    API_CATCH(m ? m->ctx : nullptr)
}

// ======================================================================= bilinear forms
extern "C" int feasm_pattern_build(feasm_mesh *m) {
    API_TRY
    FEASM_REQUIRE(m && m->ndn > 0, FEASM_ESTATE, "set the DOF numbering first");
    FEASM_REQUIRE(m->nodemajor, FEASM_EUNSUPPORTED, "resident pattern needs the node-major numbering");
    m->pattern.reset();
    ensure_pattern(m);
    FEASM_HIP(hipStreamSynchronize(m->ctx->stream));  // the pattern is complete when the call returns
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_mesh_set_numeric_path(feasm_mesh *m, int path) {
    API_TRY
    FEASM_REQUIRE(m, FEASM_EINVAL, "null mesh");
    FEASM_REQUIRE(path >= 0 && path <= 2, FEASM_EINVAL, "path: 0 automatic, 1 staged, 2 fused");
    m->numeric_path = path;
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_mesh_set_tile_params(feasm_mesh *m, int max_nodes, int budget_kib, int threads) {
    API_TRY
    FEASM_REQUIRE(m, FEASM_EINVAL, "null mesh");
    FEASM_REQUIRE(max_nodes >= 0 && max_nodes <= 255 && budget_kib >= 0 &&
                      (threads == 0 || threads == 256 || threads == 512 || threads == 1024),
                  FEASM_EINVAL, "max_nodes 0..255, budget_kib >= 0, threads 0/256/512/1024");
    m->tile_nodes = max_nodes;
    m->tile_kib = budget_kib;
    m->tile_threads = threads;
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_mesh_tile_info(feasm_mesh *m, int64_t *ntiles, int64_t *nhalo) {
    API_TRY
    FEASM_REQUIRE(m, FEASM_EINVAL, "null mesh");
    int64_t t = 0, h = 0;
    for (auto &kv : m->tiles) {
        if (!kv.second) continue;
        t += kv.second->ntiles;
        h += kv.second->nhalo;
    }
    if (ntiles) *ntiles = t;
    if (nhalo) *nhalo = h;
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_tiles_compare(feasm_mesh *m, int form, int precision, int64_t *ntiles, int64_t *first_diff,
                                   double *device_ms) {
    API_TRY
    FEASM_REQUIRE(m && ntiles && first_diff && device_ms, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(m->pattern, FEASM_ESTATE, "no pattern built");
    FEASM_REQUIRE(form >= FEASM_FORM_DIFFUSION_ISO && form <= FEASM_FORM_DOT, FEASM_EINVAL, "unknown form");
    FEASM_REQUIRE(precision == FEASM_FP64 || precision == FEASM_MIXED, FEASM_EINVAL, "unknown precision");
    FEASM_REQUIRE(m->npe == 8 || m->npe == 10, FEASM_EUNSUPPORTED, "no fused kernel for this element type");
    const TileShape shape = mesh_tile_shape(m, m->npe == 8 ? 8 : 4, form, precision);
    *first_diff = compare_tile_builders(m, shape, ntiles, device_ms);
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_pattern_info(feasm_mesh *m, int64_t *nnz, int64_t *nblocks) {
    API_TRY
    FEASM_REQUIRE(m && m->pattern, FEASM_ESTATE, "no pattern built");
    if (nnz) *nnz = m->pattern->nnz;
    if (nblocks) *nblocks = m->pattern->nblocks;
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_bilform(feasm_mesh *m, int form, int rule, int order, const double *coeff, int ncoeff,
                             int precision, feasm_matrix **out) {
    API_TRY
    FEASM_REQUIRE(m && out, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(m->ndn > 0, FEASM_ESTATE, "set the DOF numbering first");
    FEASM_REQUIRE(precision == FEASM_FP64 || precision == FEASM_MIXED, FEASM_EINVAL, "unknown precision");
    check_rule(m, rule);
    const FormParams fp = make_form_params(form, m->ndn, coeff, ncoeff);
    if (!m->nodemajor) {
        *out = coo_path(m, fp, rule, order, precision);
        return FEASM_OK;
    }
    ensure_pattern(m, m->numeric_path == 1 ||
                          !fused_supported(m->npe, device_qtable(m->ctx, m->etype, rule, order).nqp));
    std::unique_ptr<feasm_matrix> M(new feasm_matrix);
    M->ctx = m->ctx;
    M->m = M->n = m->pattern->n;
    M->nnz = m->pattern->nnz;
    M->colptr = m->pattern->colptr;
    M->rowval = m->pattern->rowval;
    M->pattern = m->pattern;
    M->form_class = m->ndn;
    M->nzval = dev_alloc(std::max<size_t>(M->nnz * sizeof(double), 8));
    numeric_into(m, fp, rule, order, precision, dp<double>(M->nzval));
    FEASM_HIP(hipStreamSynchronize(m->ctx->stream));
    *out = M.release();
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_bilform_into(feasm_mesh *m, int form, int rule, int order, const double *coeff, int ncoeff,
                                  int precision, feasm_matrix *mat) {
    API_TRY
    FEASM_REQUIRE(m && mat, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(m->nodemajor && m->pattern && mat->pattern == m->pattern, FEASM_ESTATE,
                  "matrix does not hold this mesh's resident pattern");
    FEASM_REQUIRE(precision == FEASM_FP64 || precision == FEASM_MIXED, FEASM_EINVAL, "unknown precision");
    check_rule(m, rule);
    const FormParams fp = make_form_params(form, m->ndn, coeff, ncoeff);
    numeric_into(m, fp, rule, order, precision, dp<double>(mat->nzval));
    API_CATCH(m ? m->ctx : nullptr)
}

extern "C" int feasm_bilform_coo(feasm_mesh *m, int form, int rule, int order, const double *coeff, int ncoeff,
                                 int precision, feasm_matrix **out) {
    API_TRY
    FEASM_REQUIRE(m && out, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(m->ndn > 0, FEASM_ESTATE, "set the DOF numbering first");
    FEASM_REQUIRE(precision == FEASM_FP64 || precision == FEASM_MIXED, FEASM_EINVAL, "unknown precision");
    check_rule(m, rule);
    const FormParams fp = make_form_params(form, m->ndn, coeff, ncoeff);
    *out = coo_path(m, fp, rule, order, precision);
    API_CATCH(m ? m->ctx : nullptr)
}

// ======================================================================= matrices
extern "C" int feasm_matrix_dims(const feasm_matrix *mat, int64_t *m, int64_t *n, int64_t *nnz) {
    if (!mat) return fail(nullptr, FEASM_EINVAL, "null matrix");
    if (m) *m = mat->m;
    if (n) *n = mat->n;
    if (nnz) *nnz = mat->nnz;
    return FEASM_OK;
}

extern "C" int feasm_matrix_copy_csc(feasm_matrix *mat, int base, int64_t *colptr, int64_t *rowval, double *nzval) {
    API_TRY
    FEASM_REQUIRE(mat && colptr && (rowval || mat->nnz == 0) && (nzval || mat->nnz == 0), FEASM_EINVAL,
                  "null argument");
    FEASM_REQUIRE(base == 0 || base == 1, FEASM_EINVAL, "base must be 0 or 1");
    hipStream_t s = mat->ctx->stream;
    FEASM_HIP(hipMemcpyAsync(colptr, mat->colptr->p, (mat->n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    std::vector<int32_t> rv(mat->nnz);
    if (mat->nnz) {
        FEASM_HIP(hipMemcpyAsync(rv.data(), mat->rowval->p, mat->nnz * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        FEASM_HIP(hipMemcpyAsync(nzval, mat->nzval->p, mat->nnz * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    FEASM_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i <= mat->n; i++) colptr[i] += base;
    for (int64_t i = 0; i < mat->nnz; i++) rowval[i] = (int64_t)rv[i] + base;
    API_CATCH(mat ? mat->ctx : nullptr)
}

extern "C" int feasm_matrix_device_csc(feasm_matrix *mat, const int64_t **colptr, const int32_t **rowval,
                                       double **nzval) {
    if (!mat) return fail(nullptr, FEASM_EINVAL, "null matrix");
    if (colptr) *colptr = dp<int64_t>(mat->colptr);
    if (rowval) *rowval = dp<int32_t>(mat->rowval);
    if (nzval) *nzval = dp<double>(mat->nzval);
    return FEASM_OK;
}

extern "C" int feasm_matrix_destroy(feasm_matrix *mat) {
    if (!mat) return FEASM_OK;
    (void)hipStreamSynchronize(mat->ctx->stream);
    delete mat;
    return FEASM_OK;
}

// ======================================================================= generic assemblers
// The assembler protocol of src/AssemblyModule.jl (startassembly!, assemble!,
// makematrix!) for its four sparse-matrix assemblers: SysmatAssemblerSparse
// (:88-329), SysmatAssemblerSparseSymm (:342-583), SysmatAssemblerSparseDiag
// (:599-794) and SysmatAssemblerSparseHRZLumpingSymm (:943-1141).  Triples
// are buffered on the host exactly as the reference buffers them (one
// assemble! call per element matrix arrives from host code), with the same
// growth rule and error texts, and moved to the device once, at makematrix!,
// where the device sparse() builds the CSC.
struct feasm_coo {
    CtxRef ctx;
    int kind = FEASM_ASM_SPARSE;
    int64_t buffer_length = 0, buffer_pointer = 0;  // 1-based pointer as the reference
    int64_t row_nalldofs = 0, col_nalldofs = 0;
    bool nomatrixresult = false;
    std::vector<int32_t> I, J;
    std::vector<double> V;
};

extern "C" int feasm_coo_create_kind(feasm_ctx *ctx, int kind, int nomatrixresult, feasm_coo **out) {
    API_TRY
    FEASM_REQUIRE(out, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(kind >= FEASM_ASM_SPARSE && kind <= FEASM_ASM_HRZ, FEASM_EINVAL, "unknown assembler kind");
    auto *c = new feasm_coo;
    c->ctx = ctx;
    c->kind = kind;
    c->nomatrixresult = nomatrixresult != 0;
    *out = c;
    API_CATCH(ctx)
}

extern "C" int feasm_coo_create(feasm_ctx *ctx, int nomatrixresult, feasm_coo **out) {
    return feasm_coo_create_kind(ctx, FEASM_ASM_SPARSE, nomatrixresult, out);
}

// startassembly! (src/AssemblyModule.jl:209-238, :424-454, :671-699,
// :1031-1061): the buffers are sized only while the pointer is below 1 (a
// fresh assembler); afterwards they and the matrix dimensions stay as they
// are.  expectedntriples (:54-59, :357-362, :617-622, :961-966).
extern "C" int feasm_coo_start(feasm_coo *c, int64_t elem_nrows, int64_t elem_ncols, int64_t n_elem_mats,
                               int64_t row_nalldofs, int64_t col_nalldofs) {
    API_TRY
    FEASM_REQUIRE(c, FEASM_EINVAL, "null assembler");
    FEASM_REQUIRE(elem_nrows >= 0 && elem_ncols >= 0 && n_elem_mats >= 0, FEASM_EINVAL, "negative size");
    FEASM_REQUIRE(row_nalldofs < (1ll << 31) && col_nalldofs < (1ll << 31), FEASM_EUNSUPPORTED,
                  "more than 2^31 DOFs");
    if (c->kind == FEASM_ASM_SYMM)
        FEASM_REQUIRE(elem_nrows == elem_ncols, FEASM_ESIZE,
                      "Symmetric sparse matrix is assumed to be assembled from square matrices");
    if (c->kind == FEASM_ASM_HRZ)
        FEASM_REQUIRE(elem_nrows == elem_ncols, FEASM_ESIZE,
                      "Diagonal sparse matrix is assumed to be assembled from square matrices");
    int64_t expected = elem_nrows * elem_ncols * n_elem_mats;
    if (c->kind == FEASM_ASM_SYMM) expected = (elem_nrows * elem_ncols + elem_nrows) / 2 * n_elem_mats;
    if (c->kind == FEASM_ASM_DIAG || c->kind == FEASM_ASM_HRZ)
        expected = std::max(elem_nrows, elem_ncols) * n_elem_mats;
    if (c->buffer_pointer < 1) {
        c->buffer_length = expected;
        c->I.resize(c->buffer_length);
        c->J.resize(c->buffer_length);
        c->V.resize(c->buffer_length);
        c->buffer_pointer = 1;
        if (c->kind != FEASM_ASM_SPARSE)
            FEASM_REQUIRE(row_nalldofs == col_nalldofs, FEASM_EINVAL, "Row and column info do not agree");
        c->row_nalldofs = row_nalldofs;
        c->col_nalldofs = col_nalldofs;
    }
    API_CATCH(c ? c->ctx : nullptr)
}

namespace {

// _resize_buffers (src/AssemblyModule.jl:61-76)
void coo_grow(feasm_coo *c, int64_t chunk) {
    c->buffer_length += chunk;
    c->I.resize(c->buffer_length);
    c->J.resize(c->buffer_length);
    c->V.resize(c->buffer_length);
}

void check_col(const feasm_coo *c, int64_t dj) {
    if (dj < 1) throw Error(FEASM_EDOF, "Column degree of freedom < 1");
    if (dj > c->col_nalldofs) throw Error(FEASM_EDOF, "Column degree of freedom > size");
}

void check_row(const feasm_coo *c, int64_t di) {
    if (di < 1) throw Error(FEASM_EDOF, "Row degree of freedom < 1");
    if (di > c->row_nalldofs) throw Error(FEASM_EDOF, "Row degree of freedom > size");
}

// assemble!(::SysmatAssemblerSparse) (:250-282): every entry, column-major
void append_sparse(feasm_coo *c, const double *mat, const int64_t *rows, int64_t nrows, const int64_t *cols,
                   int64_t ncols) {
    int64_t p = c->buffer_pointer;
    if (c->buffer_length < p + ncols * nrows - 1) coo_grow(c, ncols * nrows * 1000);
    for (int64_t j = 0; j < ncols; j++) {
        check_col(c, cols[j]);
        for (int64_t i = 0; i < nrows; i++) check_row(c, rows[i]);
    }
    for (int64_t j = 0; j < ncols; j++)
        for (int64_t i = 0; i < nrows; i++) {
            c->V[p - 1] = mat[i + nrows * j];
            c->I[p - 1] = (int32_t)(rows[i] - 1);
            c->J[p - 1] = (int32_t)(cols[j] - 1);
            p++;
        }
    c->buffer_pointer = p;
}

// assemble!(::SysmatAssemblerSparseSymm) (:500-549): the lower triangle of
// the element matrix in local numbering, column by column
void append_symm(feasm_coo *c, const double *mat, const int64_t *rows, int64_t n, const int64_t *cols) {
    int64_t p = c->buffer_pointer;
    // the reference's growth test compares with a half-integer bound
    if ((double)c->buffer_length < (double)p + (double)(n * n + n) / 2.0 - 1.0) coo_grow(c, n * n * 1000);
    for (int64_t j = 0; j < n; j++) {
        check_col(c, cols[j]);
        for (int64_t i = j; i < n; i++) check_row(c, rows[i]);
    }
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = j; i < n; i++) {
            c->V[p - 1] = mat[i + n * j];
            c->I[p - 1] = (int32_t)(rows[i] - 1);
            c->J[p - 1] = (int32_t)(cols[j] - 1);
            p++;
        }
    c->buffer_pointer = p;
}

// assemble! of SysmatAssemblerSparseDiag (:734-760) and
// SysmatAssemblerSparseHRZLumpingSymm (:1085-1114): one triple per column,
// (dj, dj) with the diagonal entry -- for HRZ scaled by the element's total
// over its diagonal total, em2 / dem2, both summed as the reference writes
// them (column sums added in column order; diagonal in order)
void append_diag(feasm_coo *c, const double *mat, const int64_t *cols, int64_t n, bool hrz) {
    int64_t p = c->buffer_pointer;
    if (c->buffer_length < p + n * n - 1) coo_grow(c, n * n * 1000);
    double ffactor = 1.0;
    if (hrz) {
        double em2 = 0.0, dem2 = 0.0;
        for (int64_t j = 0; j < n; j++) {
            double cs = 0.0;
            for (int64_t i = 0; i < n; i++) cs += mat[i + n * j];
            em2 += cs;
        }
        for (int64_t i = 0; i < n; i++) dem2 += mat[i + n * i];
        ffactor = em2 / dem2;
    }
    for (int64_t j = 0; j < n; j++) check_col(c, cols[j]);
    for (int64_t j = 0; j < n; j++) {
        c->V[p - 1] = hrz ? mat[j + n * j] * ffactor : mat[j + n * j];
        c->I[p - 1] = (int32_t)(cols[j] - 1);
        c->J[p - 1] = (int32_t)(cols[j] - 1);
        p++;
    }
    c->buffer_pointer = p;
}

void coo_append(feasm_coo *c, const double *mat, int64_t mat_nrows, int64_t mat_ncols, const int64_t *rows,
                int64_t nrows, const int64_t *cols, int64_t ncols) {
    if (c->kind == FEASM_ASM_SPARSE) {
        FEASM_REQUIRE(mat_nrows == nrows && mat_ncols == ncols, FEASM_ESIZE, "Wrong size of matrix");
        append_sparse(c, mat, rows, nrows, cols, ncols);
        return;
    }
    FEASM_REQUIRE(nrows == ncols, FEASM_ESIZE, "Size mismatch");
    FEASM_REQUIRE(mat_nrows == nrows && mat_ncols == ncols, FEASM_ESIZE, "Size mismatch");
    if (c->kind == FEASM_ASM_SYMM)
        append_symm(c, mat, rows, nrows, cols);
    else
        append_diag(c, mat, cols, ncols, c->kind == FEASM_ASM_HRZ);
}

}  // namespace

extern "C" int feasm_coo_assemble(feasm_coo *c, const double *mat, int64_t mat_nrows, int64_t mat_ncols,
                                  const int64_t *rows, int64_t nrows, const int64_t *cols, int64_t ncols) {
    API_TRY
    FEASM_REQUIRE(c, FEASM_EINVAL, "null assembler");
    FEASM_REQUIRE(c->buffer_pointer >= 1, FEASM_ESTATE, "startassembly! has not been called");
    coo_append(c, mat, mat_nrows, mat_ncols, rows, nrows, cols, ncols);
    API_CATCH(c ? c->ctx : nullptr)
}

extern "C" int feasm_coo_assemble_many(feasm_coo *c, int64_t nmats, const double *mats, const int64_t *rows,
                                       int64_t nrows, const int64_t *cols, int64_t ncols) {
    API_TRY
    FEASM_REQUIRE(c, FEASM_EINVAL, "null assembler");
    FEASM_REQUIRE(c->buffer_pointer >= 1, FEASM_ESTATE, "startassembly! has not been called");
    for (int64_t e = 0; e < nmats; e++)
        coo_append(c, mats + e * nrows * ncols, nrows, ncols, rows + e * nrows, nrows, cols + e * ncols, ncols);
    API_CATCH(c ? c->ctx : nullptr)
}

extern "C" int feasm_coo_set_nomatrixresult(feasm_coo *c, int flag) {
    if (!c) return fail(nullptr, FEASM_EINVAL, "null assembler");
    c->nomatrixresult = flag != 0;
    return FEASM_OK;
}

extern "C" int64_t feasm_coo_ntriples(const feasm_coo *c) { return c ? std::max<int64_t>(c->buffer_pointer - 1, 0) : -1; }

// makematrix! (:304-329, :551-583, :779-794, :1125-1141): sparse() of the
// buffered triples; for the symmetric assembler then S + transpose(S) with
// the diagonal halved (symm.hip).
extern "C" int feasm_coo_makematrix(feasm_coo *c, feasm_matrix **out) {
    API_TRY
    FEASM_REQUIRE(c && out, FEASM_EINVAL, "null argument");
    *out = nullptr;
    if (c->nomatrixresult) return FEASM_OK;  // triples kept, pointer not reset (:309-317)
    const int64_t nt = std::max<int64_t>(c->buffer_pointer - 1, 0);
    feasm_ctx *ctx = c->ctx;
    FEASM_REQUIRE(ctx, FEASM_ESTATE, "assembler has no device context");
    DevPtr I = dev_alloc(std::max<size_t>(nt * sizeof(int32_t), 4)), J = dev_alloc(std::max<size_t>(nt * 4, 4)),
           V = dev_alloc(std::max<size_t>(nt * sizeof(double), 8));
    if (nt) {
        FEASM_HIP(hipMemcpy(I->p, c->I.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice));
        FEASM_HIP(hipMemcpy(J->p, c->J.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice));
        FEASM_HIP(hipMemcpy(V->p, c->V.data(), nt * sizeof(double), hipMemcpyHostToDevice));
    }
    std::unique_ptr<feasm_matrix> S(
        sparse_device(ctx, nt, dp<int32_t>(I), dp<int32_t>(J), dp<double>(V), c->row_nalldofs, c->col_nalldofs));
    if (c->kind == FEASM_ASM_SYMM) S.reset(symm_plus_transpose(ctx, S.get()));
    *out = S.release();
    c->buffer_pointer = 1;
    API_CATCH(c ? c->ctx : nullptr)
}

extern "C" int feasm_coo_destroy(feasm_coo *c) {
    delete c;
    return FEASM_OK;
}

extern "C" int feasm_sparse_device(feasm_ctx *ctx, int64_t ntrip, const int32_t *d_I, const int32_t *d_J,
                                   const double *d_V, int64_t m, int64_t n, feasm_matrix **out) {
    API_TRY
    FEASM_REQUIRE(ctx && out, FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(m < (1ll << 31) && n < (1ll << 31), FEASM_EUNSUPPORTED, "more than 2^31 rows or columns");
    *out = sparse_device(ctx, ntrip, d_I, d_J, d_V, m, n);
    API_CATCH(ctx)
}

extern "C" int feasm_sparse(feasm_ctx *ctx, int64_t ntrip, const int64_t *I, const int64_t *J, const double *V,
                            int64_t m, int64_t n, feasm_matrix **out) {
    API_TRY
    FEASM_REQUIRE(ctx && out && (ntrip == 0 || (I && J && V)), FEASM_EINVAL, "null argument");
    FEASM_REQUIRE(m >= 0 && n >= 0 && m < (1ll << 31) && n < (1ll << 31), FEASM_EUNSUPPORTED,
                  "more than 2^31 rows or columns");
    std::vector<int32_t> i0(ntrip), j0(ntrip);
    for (int64_t t = 0; t < ntrip; t++) {
        FEASM_REQUIRE(I[t] >= 1 && I[t] <= m, FEASM_EDOF, "row index out of range");
        FEASM_REQUIRE(J[t] >= 1 && J[t] <= n, FEASM_EDOF, "column index out of range");
        i0[t] = (int32_t)(I[t] - 1);
        j0[t] = (int32_t)(J[t] - 1);
    }
    DevPtr dI = dev_alloc(std::max<size_t>(ntrip * 4, 4)), dJ = dev_alloc(std::max<size_t>(ntrip * 4, 4)),
           dV = dev_alloc(std::max<size_t>(ntrip * 8, 8));
    if (ntrip) {
        FEASM_HIP(hipMemcpy(dI->p, i0.data(), ntrip * 4, hipMemcpyHostToDevice));
        FEASM_HIP(hipMemcpy(dJ->p, j0.data(), ntrip * 4, hipMemcpyHostToDevice));
        FEASM_HIP(hipMemcpy(dV->p, V, ntrip * 8, hipMemcpyHostToDevice));
    }
    *out = sparse_device(ctx, ntrip, dp<int32_t>(dI), dp<int32_t>(dJ), dp<double>(dV), m, n);
    API_CATCH(ctx)
}

// ======================================================================= generators
extern "C" int feasm_gen_h8block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, int64_t *nelem,
                                 int64_t *nnodes, int64_t *conn, double *xyz) {
    API_TRY
    FEASM_REQUIRE(nL > 0 && nW > 0 && nH > 0, FEASM_EINVAL, "block needs at least one cell per direction");
    const int64_t ne = nL * nW * nH, nn = (nL + 1) * (nW + 1) * (nH + 1);
    if (nelem) *nelem = ne;
    if (nnodes) *nnodes = nn;
    if (conn && xyz) gen_h8block(L, W, H, nL, nW, nH, conn, xyz);
    API_CATCH(nullptr)
}

extern "C" int feasm_gen_t10block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH,
                                  double amp_vertex, double amp_mid, uint64_t seed, int64_t *nelem, int64_t *nnodes,
                                  int64_t *conn, double *xyz) {
    API_TRY
    FEASM_REQUIRE(nL > 0 && nW > 0 && nH > 0, FEASM_EINVAL, "block needs at least one cell per direction");
    const int64_t ne = 6 * nL * nW * nH;
    // nodes: vertices + distinct edges of the six-tet split (closed form)
    const int64_t nx = nL, ny = nW, nz = nH;
    const int64_t nv = (nx + 1) * (ny + 1) * (nz + 1);
    const int64_t ned = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz  // axis
                        + nx * ny * (nz + 1) + nx * (ny + 1) * nz + (nx + 1) * ny * nz               // face diagonals
                        + nx * ny * nz;                                                             // body diagonal
    if (nelem) *nelem = ne;
    if (nnodes) *nnodes = nv + ned;
    if (conn && xyz) {
        std::vector<double> x;
        const int64_t ntot = gen_t10block(L, W, H, nL, nW, nH, amp_vertex, amp_mid, seed, conn, x);
        FEASM_REQUIRE(ntot == nv + ned, FEASM_EINVAL, "internal node count mismatch");
        std::memcpy(xyz, x.data(), x.size() * sizeof(double));
    }
    API_CATCH(nullptr)
}

extern "C" int feasm_gen_perturb(int64_t nnodes, double *xyz, const double lo[3], const double hi[3], double amp,
                                 uint64_t seed) {
    API_TRY
    FEASM_REQUIRE(xyz && lo && hi, FEASM_EINVAL, "null argument");
    perturb_nodes(nnodes, xyz, lo, hi, amp, seed);
    API_CATCH(nullptr)
}
