// opx_capi.cpp -- C ABI over the gfx950 leapfrog kernels (declared in include/opx.h).
//
// The context owns all device memory. Host buffers cross the ABI in the
// reference's per-cell element order and are converted here to the device's
// structure-of-arrays layout (and to the context precision).
#include "opx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "opx_ctx.h"
#include "opx_kernels.h"


namespace {

int fail(opx_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

int hip_check(opx_ctx* c, hipError_t e, const char* what)
{
    if (e == hipSuccess) return OPX_OK;
    return fail(c, OPX_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define OPX_TRY(expr, what)                               \
    do {                                                  \
        int _rc = hip_check(ctx, (expr), (what));         \
        if (_rc != OPX_OK) return _rc;                    \
    } while (0)

opx::Lattice lattice(const opx_ctx* c, double at)
{
    opx::Lattice L;
    L.n = (uint32_t)c->n;
    L.d0 = (uint32_t)c->dims[0];
    L.d1 = (uint32_t)c->dims[1];
    L.d2 = (uint32_t)c->dims[2];
    L.s0 = L.d1 * L.d2;
    L.s1 = L.d2;
    L.at = at;
    for (int i = 0; i < 3; i++) {
        L.uf[i] = -(c->g * c->as[i]) * c->at;
        L.area[i] = 1.0 / std::pow(c->as[i], 2);
        L.as[i] = c->as[i];
    }
    for (int d = 0; d < 3; d++) {
        const int j1 = d == 0 ? 1 : 0, j2 = d == 2 ? 1 : 2;
        L.cell_ratio[d] = L.area[j2] / L.area[j1];
        L.cell_scale[d] = at * L.area[j1];
    }
    L.g = c->g;
    L.force_global = c->force_global ? 1u : 0u;
    L.j_slabs = (uint32_t)c->j_slabs.size();
    for (size_t i = 0; i < c->j_slabs.size(); i++) {
        L.j_lo[i] = (uint32_t)c->j_slabs[i].first;
        L.j_hi[i] = (uint32_t)c->j_slabs[i].second;
    }
    return L;
}

// J may now be nonzero on x0 planes [lo, hi) as well: the slab list stays sorted and disjoint
// (touching or overlapping ranges merged), and over kMaxCurrentSlabs ranges the two with the
// smallest gap between them are joined
void extend_current_planes(opx_ctx* c, int lo, int hi)
{
    lo = std::max(lo, 0);
    hi = std::min(hi, c->dims[0]);
    if (lo >= hi) return;
    auto& v = c->j_slabs;
    v.push_back({lo, hi});
    std::sort(v.begin(), v.end());
    std::vector<std::pair<int, int>> m;
    for (const auto& r : v) {
        if (!m.empty() && r.first <= m.back().second) m.back().second = std::max(m.back().second, r.second);
        else m.push_back(r);
    }
    while ((int)m.size() > opx::kMaxCurrentSlabs) {
        size_t best = 0;
        for (size_t i = 1; i + 1 < m.size(); i++)
            if (m[i + 1].first - m[i].second < m[best + 1].first - m[best].second) best = i;
        m[best].second = m[best + 1].second;
        m.erase(m.begin() + best + 1);
    }
    v = m;
    c->j_nonzero = true;
}

// elements of a link / algebra / charge-density field in the device block layout (opx::fpos)
size_t link_count(const opx_ctx* c) { return (size_t)opx::padded_cells((uint64_t)c->n) * 3 * c->S; }

// Reals stored per link on the device: FP32 SU(3) links keep their first two rows
// (real parts then imaginary parts of u00..u12); the third row is rebuilt on load.
int stored_reals(int nc, int prec) { return nc == 2 ? 4 : (prec == OPX_FP64 ? 18 : 12); }
// element index (reference storage) of stored value k
int stored_slot(int S, int k) { return S == 12 ? (k < 6 ? k : k + 3) : k; }

// third row of an SU(3) link in reference storage from its first two rows
void su3_rebuild_third_row(double* u)
{
    const double* r = u;
    const double* im = u + 9;
    for (int j = 0; j < 3; j++) {
        const int a = (j + 1) % 3, b = (j + 2) % 3;
        const double cr = (r[a] * r[3 + b] - im[a] * im[3 + b]) - (r[b] * r[3 + a] - im[b] * im[3 + a]);
        const double ci = (r[a] * im[3 + b] + im[a] * r[3 + b]) - (r[b] * im[3 + a] + im[b] * r[3 + a]);
        u[6 + j] = cr;
        u[9 + 6 + j] = -ci;
    }
}

// identity links (physics/grid/Cell.java:63-68) in the device layout
std::vector<double> identity_links(const opx_ctx* c)
{
    std::vector<double> dev(link_count(c), 0.0);
    for (int d = 0; d < 3; d++)
        for (int k = 0; k < c->S; k++) {
            const int e = stored_slot(c->S, k);
            const bool diag = c->nc == 2 ? e == 0 : (e == 0 || e == 4 || e == 8);
            if (!diag) continue;
            for (int64_t s = 0; s < c->n; s++) dev[opx::fpos(d * c->S + k, s, 3 * c->S, c->tsz)] = 1.0;
        }
    return dev;
}
size_t alg_count(const opx_ctx* c) { return (size_t)opx::padded_cells((uint64_t)c->n) * 3 * c->AD; }
size_t rho_count(const opx_ctx* c) { return (size_t)opx::padded_cells((uint64_t)c->n) * c->AD; }

// SU(3) Hermitian storage <-> Gell-Mann coefficients
// (math/SU3AlgebraElement.java:93-129: get(j) and the incremental set(j))
void su3_storage_to_coeffs(const double* v, double* c)
{
    const double s3 = std::sqrt(3.0);
    c[0] = 2 * v[1];
    c[1] = 2 * v[3];
    c[2] = v[0] - v[4];
    c[3] = 2 * v[2];
    c[4] = 2 * v[6];
    c[5] = 2 * v[5];
    c[6] = 2 * v[7];
    c[7] = (v[0] + v[4] - 2 * v[8]) / s3;
}

void su3_coeffs_to_storage(const double* c, double* v)
{
    for (int k = 0; k < 9; k++) v[k] = 0.0;
    double cur[8];
    const double s3 = std::sqrt(3.0);
    for (int j = 0; j < 8; j++) {
        su3_storage_to_coeffs(v, cur);
        double diff = (c[j] - cur[j]) / 2;
        switch (j) {
        case 0: v[1] += diff; break;
        case 1: v[3] += diff; break;
        case 2: v[0] += diff; v[4] -= diff; break;
        case 3: v[2] += diff; break;
        case 4: v[6] += diff; break;
        case 5: v[5] += diff; break;
        case 6: v[7] += diff; break;
        case 7: v[0] += diff / s3; v[4] += diff / s3; v[8] -= 2 * diff / s3; break;
        }
    }
}

template <typename T>
int copy_to_device(opx_ctx* ctx, void* dst, const std::vector<double>& src)
{
    std::vector<T> tmp(src.size());
    for (size_t i = 0; i < src.size(); i++) tmp[i] = (T)src[i];
    OPX_TRY(hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
    OPX_TRY(hipStreamSynchronize(ctx->stream), "upload sync");
    return OPX_OK;
}

template <typename T>
int copy_from_device(opx_ctx* ctx, const void* src, std::vector<double>& dst)
{
    std::vector<T> tmp(dst.size());
    OPX_TRY(hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(T), hipMemcpyDeviceToHost, ctx->stream), "download");
    OPX_TRY(hipStreamSynchronize(ctx->stream), "download sync");
    for (size_t i = 0; i < dst.size(); i++) dst[i] = (double)tmp[i];
    return OPX_OK;
}

int to_device(opx_ctx* ctx, void* dst, const std::vector<double>& vals)
{
    return ctx->prec == OPX_FP64 ? copy_to_device<double>(ctx, dst, vals) : copy_to_device<float>(ctx, dst, vals);
}

int from_device(opx_ctx* ctx, const void* src, std::vector<double>& vals)
{
    return ctx->prec == OPX_FP64 ? copy_from_device<double>(ctx, src, vals) : copy_from_device<float>(ctx, src, vals);
}

void* field_ptr(opx_ctx* c, int field)
{
    switch (field) {
    case OPX_FIELD_U: return c->U;
    case OPX_FIELD_UNEXT: return c->Unext;
    case OPX_FIELD_E: return c->E;
    case OPX_FIELD_J: return c->J;
    case OPX_FIELD_RHO: return c->rho;
    default: return nullptr;
    }
}

int check_box(opx_ctx* ctx, const int* lo, const int* hi)
{
    if (!lo || !hi) return fail(ctx, OPX_E_ARG, "region bounds are null");
    return OPX_OK;
}

int materialize_rho(opx_ctx* ctx);

opx::ParticleArrays p_view(const opx_ctx* c)
{
    return opx::ParticleArrays{c->p_count, c->p_dir, c->p_pos0, c->p_pos1, c->p_vel, c->p_Q0, c->p_Q1};
}

uint32_t auto_chunks(const opx_ctx* c, const opx::Lattice& L, int tile)
{
    const uint32_t bpc = opx::stream_blocks_per_chunk(L, tile);
    uint32_t k = (uint32_t)((2 * c->num_cus + bpc / 2) / (bpc ? bpc : 1));
    if (k < 1) k = 1;
    if (k > L.d0) k = L.d0;
    return k;
}

// x0 chunks of the tile form: about three blocks per CU
uint32_t auto_tile_chunks(const opx_ctx* c, const opx::Lattice& L)
{
    const opx::TileGeom g = opx::tile_geom(c->prec, L, 1);
    const uint32_t tiles = g.tiles1 * g.tiles2;
    uint32_t k = (uint32_t)((3 * c->num_cus + tiles - 1) / tiles);
    if (k < 1) k = 1;
    if (k > L.d0) k = L.d0;
    return k;
}

// The fused E-field/link update with the configured kernel form. With reunit, a form that can
// reproject the new links onto the group as it stores them (the cell form) does so and sets
// *fused; otherwise the caller reprojects them in a pass of its own.
hipError_t launch_update(opx_ctx* c, const opx::Lattice& L, const void* U, void* Unext, void* E,
                         bool reunit = false, bool* fused = nullptr)
{
    if (fused) *fused = false;
    if (c->kernel == OPX_KERNEL_STREAM && opx::stream_supported(c->nc, c->prec, L, c->tile)) {
        const uint32_t k = c->nchunks > 0 ? (uint32_t)c->nchunks : auto_chunks(c, L, c->tile);
        return opx::launch_stream_update(c->nc, c->prec, L, c->active, c->tile, k, U, Unext, E, c->J, c->j_nonzero,
                                         c->stream);
    }
    if (c->kernel == OPX_KERNEL_STREAM) return hipErrorInvalidConfiguration;
    if (c->kernel == OPX_KERNEL_TILE) {
        if (!opx::tile_supported(c->nc)) return hipErrorInvalidConfiguration;
        // OPX_OPT_TILE selects the store policy of this form: 0 non-temporal, 1 cached
        return opx::launch_tile_update(c->nc, c->prec, L, c->active,
                                       c->nchunks > 0 ? (uint32_t)c->nchunks : auto_tile_chunks(c, L), c->tile == 0,
                                       U, Unext, E, c->J, c->j_nonzero, c->stream);
    }
    if (c->kernel == OPX_KERNEL_PLANE)
        return opx::launch_plane_update(c->nc, c->prec, L, c->active, c->tile, U, Unext, E, c->J, c->j_nonzero,
                                        c->stream);
    if (c->kernel == OPX_KERNEL_PLAQ) {
        if (!c->P) {
            hipError_t e = hipMalloc(&c->P, alg_count(c) * c->tsz);
            if (e != hipSuccess) return e;
        }
        return opx::launch_plaq_update(c->nc, c->prec, L, c->active, U, c->P, Unext, E, c->J, c->j_nonzero, c->stream);
    }
    if (c->kernel == OPX_KERNEL_CELL) {
        if (fused) *fused = reunit;
        return opx::launch_cell_update(c->nc, c->prec, L, c->active, c->tile, U, Unext, E, c->J, c->j_nonzero,
                                       reunit && fused, c->stream);
    }
    return opx::launch_combined_update(c->nc, c->prec, L, c->active, U, Unext, E, c->J, c->j_nonzero, c->stream);
}

// ImplicitTYMSolver.step (physics/fields/ImplicitTYMSolver.java:35-58): the implicit grid is
// a copy of E (and of U and Unext where an active region leaves cells untouched); ImplicitBegin
// is the fused explicit update into the copies; each iteration swaps the implicit links and
// recomputes the active ones (ImplicitStep + ImplicitDamping); ImplicitEnd copies E and Unext
// of the active links back.
hipError_t implicit_step(opx_ctx* c, const opx::Lattice& L)
{
    const size_t lb = link_count(c) * c->tsz, ab = alg_count(c) * c->tsz;
    hipError_t e = hipSuccess;
    if (!c->iU && (e = hipMalloc(&c->iU, lb)) != hipSuccess) return e;
    if (!c->iUn && (e = hipMalloc(&c->iUn, lb)) != hipSuccess) return e;
    if (!c->iE && (e = hipMalloc(&c->iE, ab)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(c->iE, c->E, ab, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess) return e;
    if (!c->active.all) {
        if ((e = hipMemcpyAsync(c->iU, c->U, lb, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(c->iUn, c->Unext, lb, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess) return e;
    }
    if ((e = launch_update(c, L, c->U, c->iUn, c->iE)) != hipSuccess) return e;
    for (int it = 0; it < c->implicit_iterations; it++) {
        std::swap(c->iU, c->iUn);
        e = opx::launch_implicit_iteration(c->nc, c->prec, L, c->active, c->U, c->Unext, c->iU, c->E, c->J,
                                           c->j_nonzero, c->iE, c->iUn, c->implicit_damping, c->stream);
        if (e != hipSuccess) return e;
    }
    return opx::launch_implicit_end(c->nc, c->prec, L, c->active, c->iE, c->iUn, c->E, c->Unext, c->stream);
}

// FieldSolver.step with the context's field solver; reunit and fused as in launch_update
hipError_t field_step(opx_ctx* c, const opx::Lattice& L, bool reunit = false, bool* fused = nullptr)
{
    if (fused) *fused = false;
    if (c->solver == OPX_SOLVER_IMPLICIT) return implicit_step(c, L);
    return launch_update(c, L, c->U, c->Unext, c->E, reunit, fused);
}

}  // namespace

namespace opx_host {
int fail(opx_ctx* c, int code, const std::string& msg) { return ::fail(c, code, msg); }
int hip_check(opx_ctx* c, hipError_t e, const char* what) { return ::hip_check(c, e, what); }
opx::Lattice lattice(const opx_ctx* c, double at) { return ::lattice(c, at); }
void storage_to_coeffs(int nc, const double* v, double* c)
{
    if (nc == 2) {
        c[0] = v[0]; c[1] = v[1]; c[2] = v[2];
    } else {
        su3_storage_to_coeffs(v, c);
    }
}
void coeffs_to_storage(int nc, const double* c, double* v)
{
    if (nc == 2) {
        v[0] = c[0]; v[1] = c[1]; v[2] = c[2];
    } else {
        su3_coeffs_to_storage(c, v);
    }
}
hipError_t launch_update(opx_ctx* c, const opx::Lattice& L, const void* U, void* Unext, void* E)
{
    return ::launch_update(c, L, U, Unext, E, false, nullptr);
}
}  // namespace opx_host

extern "C" {

int opx_create(int num_colors, int num_dims, const int* num_cells, const double* spacing, double time_step,
               double coupling, int precision, opx_ctx** out)
{
    if (!out || !num_cells || !spacing) return OPX_E_ARG;
    *out = nullptr;
    if (num_colors != 2 && num_colors != 3) return OPX_E_UNSUPPORTED;
    if (num_dims != 3) return OPX_E_UNSUPPORTED;
    if (precision != OPX_FP64 && precision != OPX_FP32) return OPX_E_UNSUPPORTED;
    int64_t n = 1;
    for (int i = 0; i < 3; i++) {
        if (num_cells[i] < 1) return OPX_E_ARG;
        if (!(spacing[i] > 0)) return OPX_E_ARG;
        n *= num_cells[i];
    }
    if (n > (int64_t)0x7fffffff) return OPX_E_UNSUPPORTED;
    opx_ctx* ctx = new opx_ctx();
    ctx->nc = num_colors;
    ctx->prec = precision;
    for (int i = 0; i < 3; i++) {
        ctx->dims[i] = num_cells[i];
        ctx->as[i] = spacing[i];
    }
    ctx->at = time_step;
    ctx->g = coupling;
    ctx->n = n;
    ctx->G = num_colors == 2 ? 4 : 18;
    ctx->S = stored_reals(num_colors, precision);
    ctx->AD = num_colors == 2 ? 3 : 8;
    ctx->AS = num_colors == 2 ? 3 : 9;
    ctx->tsz = precision == OPX_FP64 ? 8 : 4;
    int rc = hip_check(ctx, hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking), "stream");
    if (rc == OPX_OK) {
        int dev = 0, cus = 0;
        rc = hip_check(ctx, hipGetDevice(&dev), "current device");
        ctx->device = dev;
        if (rc == OPX_OK && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            cus > 0)
            ctx->num_cus = cus;
    }
    if (rc == OPX_OK) {
        ctx->stream = ctx->own_stream;
        rc = hip_check(ctx, hipMalloc(&ctx->U, link_count(ctx) * ctx->tsz), "alloc U");
    }
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&ctx->Unext, link_count(ctx) * ctx->tsz), "alloc Unext");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&ctx->E, alg_count(ctx) * ctx->tsz), "alloc E");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&ctx->J, alg_count(ctx) * ctx->tsz), "alloc J");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&ctx->rho, rho_count(ctx) * ctx->tsz), "alloc rho");
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMalloc(&ctx->partials, opx::measure_partials_size((uint32_t)n) * sizeof(double)),
                       "alloc partials");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&ctx->dout, 64 * sizeof(double)), "alloc out");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&ctx->counters, 4 * sizeof(unsigned long long)), "alloc counters");
    if (rc == OPX_OK) {
        // identities (physics/grid/Cell.java:63-68) and zero fields
        const std::vector<double> ident = identity_links(ctx);
        rc = to_device(ctx, ctx->U, ident);
        if (rc == OPX_OK) rc = to_device(ctx, ctx->Unext, ident);
    }
    if (rc == OPX_OK) rc = hip_check(ctx, hipMemsetAsync(ctx->E, 0, alg_count(ctx) * ctx->tsz, ctx->stream), "zero E");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMemsetAsync(ctx->J, 0, alg_count(ctx) * ctx->tsz, ctx->stream), "zero J");
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMemsetAsync(ctx->rho, 0, rho_count(ctx) * ctx->tsz, ctx->stream), "zero rho");
    if (rc == OPX_OK) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "create sync");
    if (rc != OPX_OK) {
        std::fprintf(stderr, "opx_create: %s\n", ctx->err.c_str());
        opx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return OPX_OK;
}

void opx_destroy(opx_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->batch_view) {
        // borrowed field memory belongs to the batch; only the view's own buffers are freed
        void* own[] = {ctx->partials, ctx->dout, ctx->counters, ctx->diag};
        for (void* b : own)
            if (b) (void)hipFree(b);
        delete ctx;
        return;
    }
    opx_host::adjoint_release(ctx);
    opx_host::cgc_release(ctx);
    void* bufs[] = {ctx->U, ctx->Unext, ctx->E, ctx->J, ctx->rho, ctx->P, ctx->iU, ctx->iUn, ctx->iE,
                    ctx->partials, ctx->dout, ctx->counters, ctx->diag,
                    ctx->sp_desc, ctx->sp_qoff, ctx->sp_q, ctx->p_dir, ctx->p_pos0, ctx->p_pos1, ctx->p_vel,
                    ctx->p_Q0, ctx->p_Q1, ctx->ps_buf, ctx->ps_temp, ctx->rs_dev};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int k = 0; k < opx_ctx::kSegSlots; k++) {
        if (ctx->rs_host[k]) (void)hipHostFree(ctx->rs_host[k]);
        if (ctx->rs_copied[k]) (void)hipEventDestroy(ctx->rs_copied[k]);
    }
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

const char* opx_last_error(const opx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int opx_get_info(const opx_ctx* ctx, int64_t* num_cells_total, int* num_colors, int* precision)
{
    if (!ctx) return OPX_E_ARG;
    if (num_cells_total) *num_cells_total = ctx->n;
    if (num_colors) *num_colors = ctx->nc;
    if (precision) *precision = ctx->prec;
    return OPX_OK;
}

int opx_set_stream(opx_ctx* ctx, void* hip_stream)
{
    if (!ctx) return OPX_E_ARG;
    OPX_TRY(hipStreamSynchronize(ctx->stream), "set_stream sync");
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return OPX_OK;
}

int opx_sync(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    OPX_TRY(hipStreamSynchronize(ctx->stream), "sync");
    return OPX_OK;
}

int opx_upload_field(opx_ctx* ctx, int field, const double* host, int64_t n_values)
{
    if (!ctx) return OPX_E_ARG;
    if (!host) return fail(ctx, OPX_E_ARG, "upload: null buffer");
    void* dst = field_ptr(ctx, field);
    if (!dst) return fail(ctx, OPX_E_ARG, "upload: unknown field");
    const int64_t n = ctx->n;
    std::vector<double> dev;   // device layout (opx::fpos), padding zero
    if (field == OPX_FIELD_U || field == OPX_FIELD_UNEXT) {
        if (n_values != n * 3 * ctx->G) return fail(ctx, OPX_E_ARG, "upload: link buffer size mismatch");
        dev.assign(link_count(ctx), 0.0);
        for (int64_t s = 0; s < n; s++)
            for (int d = 0; d < 3; d++)
                for (int k = 0; k < ctx->S; k++)
                    dev[opx::fpos(d * ctx->S + k, s, 3 * ctx->S, ctx->tsz)] = host[(s * 3 + d) * ctx->G + stored_slot(ctx->S, k)];
    } else {
        const int ndir = field == OPX_FIELD_RHO ? 1 : 3;
        if (n_values != n * ndir * ctx->AS) return fail(ctx, OPX_E_ARG, "upload: algebra buffer size mismatch");
        const uint32_t comps = ndir * ctx->AD;
        dev.assign((size_t)opx::padded_cells((uint64_t)n) * comps, 0.0);
        for (int64_t s = 0; s < n; s++)
            for (int d = 0; d < ndir; d++) {
                const double* v = host + (s * ndir + d) * ctx->AS;
                double c[8];
                if (ctx->nc == 2) {
                    c[0] = v[0]; c[1] = v[1]; c[2] = v[2];
                } else {
                    su3_storage_to_coeffs(v, c);
                }
                for (int k = 0; k < ctx->AD; k++) dev[opx::fpos(d * ctx->AD + k, s, comps, ctx->tsz)] = c[k];
            }
        if (field == OPX_FIELD_J) extend_current_planes(ctx, 0, ctx->dims[0]);
        if (field == OPX_FIELD_RHO) ctx->rho_pending = false;
    }
    return to_device(ctx, dst, dev);
}

int opx_download_field(opx_ctx* ctx, int field, double* host, int64_t n_values)
{
    if (!ctx) return OPX_E_ARG;
    if (!host) return fail(ctx, OPX_E_ARG, "download: null buffer");
    const void* src = field_ptr(ctx, field);
    if (!src) return fail(ctx, OPX_E_ARG, "download: unknown field");
    const int64_t n = ctx->n;
    if (field == OPX_FIELD_RHO) {
        int rc = materialize_rho(ctx);
        if (rc != OPX_OK) return rc;
    }
    if (field == OPX_FIELD_U || field == OPX_FIELD_UNEXT) {
        if (n_values != n * 3 * ctx->G) return fail(ctx, OPX_E_ARG, "download: link buffer size mismatch");
        std::vector<double> dev(link_count(ctx));
        int rc = from_device(ctx, src, dev);
        if (rc != OPX_OK) return rc;
        for (int64_t s = 0; s < n; s++)
            for (int d = 0; d < 3; d++) {
                double* u = host + (s * 3 + d) * ctx->G;
                for (int k = 0; k < ctx->S; k++) u[stored_slot(ctx->S, k)] = dev[opx::fpos(d * ctx->S + k, s, 3 * ctx->S, ctx->tsz)];
                if (ctx->S != ctx->G) su3_rebuild_third_row(u);
            }
        return OPX_OK;
    }
    const int ndir = field == OPX_FIELD_RHO ? 1 : 3;
    if (n_values != n * ndir * ctx->AS) return fail(ctx, OPX_E_ARG, "download: algebra buffer size mismatch");
    const uint32_t comps = ndir * ctx->AD;
    std::vector<double> dev((size_t)opx::padded_cells((uint64_t)n) * comps);
    int rc = from_device(ctx, src, dev);
    if (rc != OPX_OK) return rc;
    for (int64_t s = 0; s < n; s++)
        for (int d = 0; d < ndir; d++) {
            double c[8];
            for (int k = 0; k < ctx->AD; k++) c[k] = dev[opx::fpos(d * ctx->AD + k, s, comps, ctx->tsz)];
            double* v = host + (s * ndir + d) * ctx->AS;
            if (ctx->nc == 2) {
                v[0] = c[0]; v[1] = c[1]; v[2] = c[2];
            } else {
                su3_coeffs_to_storage(c, v);
            }
        }
    return OPX_OK;
}

int opx_set_links_from_coefficients(opx_ctx* ctx, int field, const double* coeffs, int64_t n_values)
{
    if (!ctx) return OPX_E_ARG;
    if (!coeffs) return fail(ctx, OPX_E_ARG, "coefficients: null buffer");
    if (field != OPX_FIELD_U && field != OPX_FIELD_UNEXT) return fail(ctx, OPX_E_ARG, "coefficients: not a link field");
    const int64_t n = ctx->n;
    if (n_values != n * 3 * ctx->AD) return fail(ctx, OPX_E_ARG, "coefficients: size mismatch");
    std::vector<double> soa((size_t)n_values);
    for (int64_t s = 0; s < n; s++)
        for (int d = 0; d < 3; d++)
            for (int k = 0; k < ctx->AD; k++) soa[((size_t)d * ctx->AD + k) * n + s] = coeffs[(s * 3 + d) * ctx->AD + k];
    double* dc = nullptr;
    OPX_TRY(hipMalloc(&dc, soa.size() * sizeof(double)), "alloc coefficients");
    int rc = hip_check(ctx, hipMemcpyAsync(dc, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
                       "upload coefficients");
    if (rc == OPX_OK)
        rc = hip_check(ctx, opx::launch_links_from_coeffs(ctx->nc, ctx->prec, (uint32_t)n, dc, field_ptr(ctx, field),
                                                          ctx->stream),
                       "links_from_coefficients");
    if (rc == OPX_OK) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "coefficients sync");
    (void)hipFree(dc);
    return rc;
}

int opx_set_efield_from_coefficients(opx_ctx* ctx, const double* coeffs, int64_t n_values)
{
    if (!ctx) return OPX_E_ARG;
    if (!coeffs) return fail(ctx, OPX_E_ARG, "coefficients: null buffer");
    const int64_t n = ctx->n;
    if (n_values != n * 3 * ctx->AD) return fail(ctx, OPX_E_ARG, "coefficients: size mismatch");
    std::vector<double> dev(alg_count(ctx), 0.0);
    for (int64_t s = 0; s < n; s++)
        for (int d = 0; d < 3; d++)
            for (int k = 0; k < ctx->AD; k++)
                dev[opx::fpos(d * ctx->AD + k, s, 3 * ctx->AD, ctx->tsz)] = coeffs[(s * 3 + d) * ctx->AD + k];
    return to_device(ctx, ctx->E, dev);
}

int opx_set_active_region(opx_ctx* ctx, const int* lo, const int* hi)
{
    if (!ctx) return OPX_E_ARG;
    // the batch kernels update every cell of every event
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "the events of a batch have no active region");
    int rc = check_box(ctx, lo, hi);
    if (rc != OPX_OK) return rc;
    for (int i = 0; i < 3; i++) {
        ctx->active.lo[i] = lo[i];
        ctx->active.hi[i] = hi[i];
    }
    ctx->active.all = 0;
    return OPX_OK;
}

int opx_set_evaluation_region(opx_ctx* ctx, const int* lo, const int* hi)
{
    if (!ctx) return OPX_E_ARG;
    int rc = check_box(ctx, lo, hi);
    if (rc != OPX_OK) return rc;
    for (int i = 0; i < 3; i++) {
        ctx->eval.lo[i] = lo[i];
        ctx->eval.hi[i] = hi[i];
    }
    ctx->eval.all = 0;
    return OPX_OK;
}

int opx_step(opx_ctx* ctx, double dt)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "step the batch, not one of its events");
    opx::Lattice L = lattice(ctx, dt);
    OPX_TRY(field_step(ctx, L), "step");
    return OPX_OK;
}

int opx_step_links(opx_ctx* ctx, double dt)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "step the batch, not one of its events");
    opx::Lattice L = lattice(ctx, dt);
    OPX_TRY(opx::launch_step_links(ctx->nc, ctx->prec, L, ctx->active, ctx->U, ctx->Unext, ctx->E, ctx->stream),
            "step_links");
    return OPX_OK;
}

int opx_set_field_solver(opx_ctx* ctx, int solver, int iterations, double damping)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "the events of a batch use the explicit solver");
    if (solver != OPX_SOLVER_FAST && solver != OPX_SOLVER_IMPLICIT) return fail(ctx, OPX_E_ARG, "field solver: unknown");
    if (iterations < 0) return fail(ctx, OPX_E_ARG, "field solver: negative iteration count");
    if (!std::isfinite(damping)) return fail(ctx, OPX_E_ARG, "field solver: damping must be finite");
    ctx->solver = solver;
    ctx->implicit_iterations = iterations;
    ctx->implicit_damping = damping;
    return OPX_OK;
}

int opx_group_pow(int num_colors, int precision, int count, const double* elements, double x, double* out)
{
    if ((num_colors != 2 && num_colors != 3) || (precision != OPX_FP64 && precision != OPX_FP32))
        return OPX_E_UNSUPPORTED;
    if (count < 0 || (count > 0 && (!elements || !out))) return OPX_E_ARG;
    if (count == 0) return OPX_OK;
    const size_t m = (size_t)count * (num_colors == 2 ? 4 : 18);
    const bool f64 = precision == OPX_FP64;
    const size_t tsz = f64 ? 8 : 4;
    std::vector<double> hd(f64 ? m : 0);
    std::vector<float> hf(f64 ? 0 : m);
    for (size_t k = 0; k < m; k++) {
        if (f64) hd[k] = elements[k];
        else hf[k] = (float)elements[k];
    }
    void* host = f64 ? (void*)hd.data() : (void*)hf.data();
    void *din = nullptr, *dout = nullptr;
    int rc = OPX_OK;
    if (hipMalloc(&din, m * tsz) != hipSuccess || hipMalloc(&dout, m * tsz) != hipSuccess) rc = OPX_E_DEVICE;
    if (rc == OPX_OK && hipMemcpy(din, host, m * tsz, hipMemcpyHostToDevice) != hipSuccess) rc = OPX_E_DEVICE;
    if (rc == OPX_OK && opx::launch_group_pow(num_colors, f64 ? 0 : 1, count, din, x, dout, nullptr) != hipSuccess)
        rc = OPX_E_DEVICE;
    if (rc == OPX_OK && hipMemcpy(host, dout, m * tsz, hipMemcpyDeviceToHost) != hipSuccess) rc = OPX_E_DEVICE;
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (rc != OPX_OK) return rc;
    for (size_t k = 0; k < m; k++) out[k] = f64 ? hd[k] : (double)hf[k];
    return OPX_OK;
}

int opx_store_fields(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "step the batch, not one of its events");
    std::swap(ctx->U, ctx->Unext);
    return OPX_OK;
}

int opx_reset_current(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    // J is zero outside its slabs: only they are cleared, all in one launch
    const uint32_t s0 = (uint32_t)ctx->dims[1] * (uint32_t)ctx->dims[2];
    opx::CellRanges cr{};
    cr.start[0] = 0;
    for (const auto& r : ctx->j_slabs) {
        if (cr.count == opx::kMaxCurrentSlabs) break;   // the slab list never holds more
        cr.first[cr.count] = (uint32_t)r.first * s0;
        cr.start[cr.count + 1] = cr.start[cr.count] + (uint32_t)(r.second - r.first) * s0;
        cr.count++;
    }
    OPX_TRY(opx::launch_zero_cell_ranges(ctx->prec, ctx->J, 3 * ctx->AD, cr, ctx->stream), "reset current");
    ctx->j_slabs.clear();
    ctx->j_nonzero = false;
    return OPX_OK;
}

int opx_reset_charge(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    OPX_TRY(hipMemsetAsync(ctx->rho, 0, rho_count(ctx) * ctx->tsz, ctx->stream), "reset charge");
    ctx->rho_pending = false;
    return OPX_OK;
}

int opx_reunitarize(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    OPX_TRY(opx::launch_reunitarize(ctx->nc, ctx->prec, (uint32_t)ctx->n, ctx->U, ctx->stream), "reunitarize U");
    OPX_TRY(opx::launch_reunitarize(ctx->nc, ctx->prec, (uint32_t)ctx->n, ctx->Unext, ctx->stream),
            "reunitarize Unext");
    return OPX_OK;
}

int opx_leapfrog(opx_ctx* ctx, int steps, int reunit_every)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "step the batch, not one of its events");
    if (steps < 0 || reunit_every < 0) return fail(ctx, OPX_E_ARG, "leapfrog: negative count");
    opx::Lattice L = lattice(ctx, ctx->at);
    for (int s = 0; s < steps; s++) {
        std::swap(ctx->U, ctx->Unext);
        // every reunit_every-th step reprojects the new links, inside the update where the form can
        const bool reunit = reunit_every > 0 && (ctx->steps + 1) % reunit_every == 0;
        bool fused = false;
        OPX_TRY(field_step(ctx, L, reunit, &fused), "leapfrog step");
        ctx->steps++;
        if (reunit && !fused)
            OPX_TRY(opx::launch_reunitarize(ctx->nc, ctx->prec, (uint32_t)ctx->n, ctx->Unext, ctx->stream),
                    "leapfrog reunitarize");
    }
    return OPX_OK;
}

int opx_measure(opx_ctx* ctx, double* out)
{
    if (!ctx) return OPX_E_ARG;
    if (!out) return fail(ctx, OPX_E_ARG, "measure: null output");
    int rc = materialize_rho(ctx);
    if (rc != OPX_OK) return rc;
    opx::Lattice L = lattice(ctx, ctx->at);
    OPX_TRY(opx::launch_measure(ctx->nc, ctx->prec, L, ctx->eval, ctx->U, ctx->Unext, ctx->E, ctx->rho, ctx->partials,
                                ctx->dout, ctx->stream),
            "measure");
    OPX_TRY(hipMemcpyAsync(out, ctx->dout, OPX_M_COUNT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
            "measure copy");
    OPX_TRY(hipStreamSynchronize(ctx->stream), "measure sync");
    return OPX_OK;
}

// the per-slice output buffer of the projected diagnostics, allocated on the first call and
// kept: 4 values per slice along the longest axis
static int diag_buffer(opx_ctx* ctx)
{
    if (ctx->diag) return OPX_OK;
    const int nd = std::max(ctx->dims[0], std::max(ctx->dims[1], ctx->dims[2]));
    return hip_check(ctx, hipMalloc(&ctx->diag, (size_t)4 * nd * sizeof(double)), "alloc projected diagnostics");
}

int opx_projected_energy_density(opx_ctx* ctx, int direction, double* eT_el, double* eT_mag, double* eL_el,
                                 double* eL_mag)
{
    if (!ctx) return OPX_E_ARG;
    if (direction < 0 || direction > 2) return fail(ctx, OPX_E_ARG, "projected energy: bad direction");
    if (!eT_el || !eT_mag || !eL_el || !eL_mag) return fail(ctx, OPX_E_ARG, "projected energy: null output");
    const int nd = ctx->dims[direction];
    int rc = diag_buffer(ctx);
    if (rc != OPX_OK) return rc;
    opx::Lattice L = lattice(ctx, ctx->at);
    rc = hip_check(ctx, opx::launch_projected_energy(ctx->nc, ctx->prec, L, ctx->eval, direction, ctx->U, ctx->Unext,
                                                     ctx->E, ctx->diag, ctx->stream),
                   "projected energy");
    std::vector<double> h((size_t)4 * nd);
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMemcpyAsync(h.data(), ctx->diag, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                                           ctx->stream),
                       "projected energy copy");
    if (rc == OPX_OK) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "projected energy sync");
    if (rc != OPX_OK) return rc;
    std::memcpy(eT_el, h.data(), nd * sizeof(double));
    std::memcpy(eT_mag, h.data() + nd, nd * sizeof(double));
    std::memcpy(eL_el, h.data() + 2 * (size_t)nd, nd * sizeof(double));
    std::memcpy(eL_mag, h.data() + 3 * (size_t)nd, nd * sizeof(double));
    return OPX_OK;
}

int opx_projected_energy_density_noalloc(opx_ctx* ctx, int direction, double* eT_el, double* eT_mag, double* eL_el,
                                         double* eL_mag)
{
    if (!ctx) return OPX_E_ARG;
    if (direction < 0 || direction > 2) return fail(ctx, OPX_E_ARG, "projected energy: bad direction");
    if (!eT_el || !eT_mag || !eL_el || !eL_mag) return fail(ctx, OPX_E_ARG, "projected energy: null output");
    const int nd = ctx->dims[direction];
    if (nd > opx::kNoAllocMaxSlices)
        return fail(ctx, OPX_E_ARG, "projected energy (no allocation): " + std::to_string(nd) +
                                        " slices along the direction, more than the static buffer's " +
                                        std::to_string(opx::kNoAllocMaxSlices));
    double* buf = nullptr;
    OPX_TRY(opx::noalloc_scratch(ctx->device, &buf), "projected energy (no allocation): static buffer address");
    opx::Lattice L = lattice(ctx, ctx->at);
    OPX_TRY(opx::launch_projected_energy(ctx->nc, ctx->prec, L, ctx->eval, direction, ctx->U, ctx->Unext, ctx->E, buf,
                                         ctx->stream),
            "projected energy");
    double* outs[4] = {eT_el, eT_mag, eL_el, eL_mag};
    for (int q = 0; q < 4; q++)
        OPX_TRY(hipMemcpyAsync(outs[q], buf + (size_t)q * nd, (size_t)nd * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream),
                "projected energy copy");
    OPX_TRY(hipStreamSynchronize(ctx->stream), "projected energy sync");
    return OPX_OK;
}

int opx_projected_poynting(opx_ctx* ctx, int direction, double* averaged, double* time_averaged)
{
    if (!ctx) return OPX_E_ARG;
    if (direction < 0 || direction > 2) return fail(ctx, OPX_E_ARG, "projected Poynting vector: bad direction");
    if (!averaged || !time_averaged) return fail(ctx, OPX_E_ARG, "projected Poynting vector: null output");
    if (ctx->nc != 2)
        return fail(ctx, OPX_E_UNSUPPORTED, "projected Poynting vector: SU(3) code has not been tested yet.");
    const int nd = ctx->dims[direction];
    int rc = diag_buffer(ctx);
    if (rc != OPX_OK) return rc;
    opx::Lattice L = lattice(ctx, ctx->at);
    rc = hip_check(ctx, opx::launch_projected_poynting(ctx->nc, ctx->prec, L, direction, ctx->U, ctx->Unext, ctx->E,
                                                       ctx->diag, ctx->stream),
                   "projected Poynting vector");
    std::vector<double> h((size_t)2 * nd);
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMemcpyAsync(h.data(), ctx->diag, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                                           ctx->stream),
                       "projected Poynting vector copy");
    if (rc == OPX_OK) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "projected Poynting vector sync");
    if (rc != OPX_OK) return rc;
    std::memcpy(averaged, h.data(), nd * sizeof(double));
    std::memcpy(time_averaged, h.data() + nd, nd * sizeof(double));
    return OPX_OK;
}

int opx_superparticles_upload(opx_ctx* ctx, int count, const int* orientation, const int* num,
                              const int* index_offset, const int* per_plane, const int* sub_lattice_shift,
                              const int* particles_per_cell, const double* charges, int64_t n_values)
{
    if (!ctx) return OPX_E_ARG;
    if (count < 0) return fail(ctx, OPX_E_ARG, "superparticles: negative count");
    if (count > 0 && (!orientation || !num || !index_offset || !per_plane || !sub_lattice_shift ||
                      !particles_per_cell || !charges))
        return fail(ctx, OPX_E_ARG, "superparticles: null array");
    int64_t total = 0;
    int maxn = 0;
    std::vector<int> desc((size_t)6 * count);
    std::vector<int64_t> qoff((size_t)count);
    for (int p = 0; p < count; p++) {
        if (num[p] < 0) return fail(ctx, OPX_E_ARG, "superparticles: negative particle number");
        if (particles_per_cell[p] < 1) return fail(ctx, OPX_E_ARG, "superparticles: particlePerCell must be >= 1");
        if (orientation[p] == 0) return fail(ctx, OPX_E_ARG, "superparticles: orientation must be +1 or -1");
        qoff[p] = total;
        total += num[p];
        if (num[p] > maxn) maxn = num[p];
        desc[0 * (size_t)count + p] = orientation[p];
        desc[1 * (size_t)count + p] = num[p];
        desc[2 * (size_t)count + p] = index_offset[p];
        desc[3 * (size_t)count + p] = per_plane[p];
        desc[4 * (size_t)count + p] = sub_lattice_shift[p];
        desc[5 * (size_t)count + p] = particles_per_cell[p];
    }
    if (n_values != total * ctx->AS) return fail(ctx, OPX_E_ARG, "superparticles: charge buffer size mismatch");
    if (ctx->sp_desc) (void)hipFree(ctx->sp_desc);
    if (ctx->sp_qoff) (void)hipFree(ctx->sp_qoff);
    if (ctx->sp_q) (void)hipFree(ctx->sp_q);
    ctx->sp_desc = nullptr;
    ctx->sp_qoff = nullptr;
    ctx->sp_q = nullptr;
    ctx->sp_count = 0;
    ctx->sp_total = 0;
    ctx->sp_max_num = 0;
    if (count == 0) return OPX_OK;
    OPX_TRY(hipMalloc(&ctx->sp_desc, desc.size() * sizeof(int)), "alloc sp desc");
    OPX_TRY(hipMalloc(&ctx->sp_qoff, qoff.size() * sizeof(int64_t)), "alloc sp qoff");
    OPX_TRY(hipMalloc(&ctx->sp_q, (size_t)(total > 0 ? total : 1) * ctx->AD * ctx->tsz), "alloc sp charges");
    OPX_TRY(hipMemcpyAsync(ctx->sp_desc, desc.data(), desc.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream),
            "upload sp desc");
    OPX_TRY(hipMemcpyAsync(ctx->sp_qoff, qoff.data(), qoff.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream),
            "upload sp qoff");
    std::vector<double> soa((size_t)total * ctx->AD);
    for (int64_t q = 0; q < total; q++) {
        double c[8];
        const double* v = charges + q * ctx->AS;
        if (ctx->nc == 2) {
            c[0] = v[0]; c[1] = v[1]; c[2] = v[2];
        } else {
            su3_storage_to_coeffs(v, c);
        }
        for (int k = 0; k < ctx->AD; k++) soa[(size_t)k * total + q] = c[k];
    }
    ctx->sp_count = count;
    ctx->sp_total = total;
    ctx->sp_max_num = maxn;
    ctx->sp_host = desc;
    if (total > 0) return to_device(ctx, ctx->sp_q, soa);
    return OPX_OK;
}

int opx_superparticles_count(const opx_ctx* ctx, int* count, int64_t* total_charges)
{
    if (!ctx) return OPX_E_ARG;
    if (count) *count = ctx->sp_count;
    if (total_charges) *total_charges = ctx->sp_total;
    return OPX_OK;
}

int opx_superparticles_download_descriptors(opx_ctx* ctx, int count, int* orientation, int* num, int* index_offset,
                                            int* per_plane, int* sub_lattice_shift, int* particles_per_cell)
{
    if (!ctx) return OPX_E_ARG;
    if (count != ctx->sp_count) return fail(ctx, OPX_E_ARG, "superparticles: count mismatch");
    int* out[6] = {orientation, num, index_offset, per_plane, sub_lattice_shift, particles_per_cell};
    for (int f = 0; f < 6; f++) {
        if (count > 0 && !out[f]) return fail(ctx, OPX_E_ARG, "superparticles: null array");
        for (int p = 0; p < count; p++) out[f][p] = ctx->sp_host[(size_t)f * count + p];
    }
    return OPX_OK;
}

int opx_superparticles_download_charges(opx_ctx* ctx, double* charges, int64_t n_values)
{
    if (!ctx) return OPX_E_ARG;
    if (!charges) return fail(ctx, OPX_E_ARG, "superparticles: null buffer");
    if (n_values != ctx->sp_total * ctx->AS) return fail(ctx, OPX_E_ARG, "superparticles: charge buffer size mismatch");
    if (ctx->sp_total == 0) return OPX_OK;
    std::vector<double> soa((size_t)ctx->sp_total * ctx->AD);
    int rc = from_device(ctx, ctx->sp_q, soa);
    if (rc != OPX_OK) return rc;
    for (int64_t q = 0; q < ctx->sp_total; q++) {
        double c[8];
        for (int k = 0; k < ctx->AD; k++) c[k] = soa[(size_t)k * ctx->sp_total + q];
        double* v = charges + q * ctx->AS;
        if (ctx->nc == 2) {
            v[0] = c[0]; v[1] = c[1]; v[2] = c[2];
        } else {
            su3_coeffs_to_storage(c, v);
        }
    }
    return OPX_OK;
}

int opx_particles_upload(opx_ctx* ctx, int count, const int* direction, const double* pos0, const double* vel,
                         const double* charges, int64_t n_values)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "particles belong to a single context, not a batch event");
    if (count < 0) return fail(ctx, OPX_E_ARG, "particles: negative count");
    if (count > 0 && (!direction || !pos0 || !vel || !charges)) return fail(ctx, OPX_E_ARG, "particles: null array");
    if (n_values != (int64_t)count * ctx->AS) return fail(ctx, OPX_E_ARG, "particles: charge buffer size mismatch");
    for (int k = 0; k < count; k++)
        if (direction[k] < 0 || direction[k] > 2) return fail(ctx, OPX_E_ARG, "particles: direction must be 0, 1 or 2");
    void* old[] = {ctx->p_dir, ctx->p_pos0, ctx->p_pos1, ctx->p_vel, ctx->p_Q0, ctx->p_Q1};
    for (void* b : old)
        if (b) (void)hipFree(b);
    ctx->p_dir = nullptr;
    ctx->p_pos0 = ctx->p_pos1 = ctx->p_vel = nullptr;
    ctx->p_Q0 = ctx->p_Q1 = nullptr;
    ctx->p_count = 0;
    if (count == 0) return OPX_OK;
    const size_t c = (size_t)count;
    OPX_TRY(hipMalloc(&ctx->p_dir, c * sizeof(int)), "alloc particles");
    OPX_TRY(hipMalloc(&ctx->p_pos0, 3 * c * sizeof(double)), "alloc particles");
    OPX_TRY(hipMalloc(&ctx->p_pos1, 3 * c * sizeof(double)), "alloc particles");
    OPX_TRY(hipMalloc(&ctx->p_vel, 3 * c * sizeof(double)), "alloc particles");
    OPX_TRY(hipMalloc(&ctx->p_Q0, c * ctx->AD * ctx->tsz), "alloc particles");
    OPX_TRY(hipMalloc(&ctx->p_Q1, c * ctx->AD * ctx->tsz), "alloc particles");
    std::vector<double> p0(3 * c), v(3 * c), q((size_t)ctx->AD * c);
    for (size_t k = 0; k < c; k++)
        for (int i = 0; i < 3; i++) {
            p0[(size_t)i * c + k] = pos0[3 * k + i];
            v[(size_t)i * c + k] = vel[3 * k + i];
        }
    for (size_t k = 0; k < c; k++) {
        double cf[8];
        const double* a = charges + k * ctx->AS;
        if (ctx->nc == 2) {
            cf[0] = a[0]; cf[1] = a[1]; cf[2] = a[2];
        } else {
            su3_storage_to_coeffs(a, cf);
        }
        for (int m = 0; m < ctx->AD; m++) q[(size_t)m * c + k] = cf[m];
    }
    OPX_TRY(hipMemcpyAsync(ctx->p_dir, direction, c * sizeof(int), hipMemcpyHostToDevice, ctx->stream), "upload particles");
    OPX_TRY(hipMemcpyAsync(ctx->p_pos0, p0.data(), 3 * c * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
            "upload particles");
    OPX_TRY(hipMemcpyAsync(ctx->p_vel, v.data(), 3 * c * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
            "upload particles");
    // pos1 and Q1 start at zero; the first push writes pos1 and Q1 before either is read
    OPX_TRY(hipMemsetAsync(ctx->p_pos1, 0, 3 * c * sizeof(double), ctx->stream), "clear particles");
    OPX_TRY(hipMemsetAsync(ctx->p_Q1, 0, c * ctx->AD * ctx->tsz, ctx->stream), "clear particles");
    ctx->p_count = (uint32_t)count;
    int rc = to_device(ctx, ctx->p_Q0, q);   // synchronizes the stream (host vectors stay alive until then)
    if (rc != OPX_OK) ctx->p_count = 0;
    return rc;
}

int opx_particles_download(opx_ctx* ctx, double* pos0, double* pos1, double* Q0, double* Q1, int64_t count)
{
    if (!ctx) return OPX_E_ARG;
    if (count != (int64_t)ctx->p_count) return fail(ctx, OPX_E_ARG, "particles: count mismatch");
    if (count == 0) return OPX_OK;
    const size_t c = (size_t)count;
    std::vector<double> p0(3 * c), p1(3 * c), q0((size_t)ctx->AD * c), q1((size_t)ctx->AD * c);
    OPX_TRY(hipMemcpyAsync(p0.data(), ctx->p_pos0, 3 * c * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
            "download particles");
    OPX_TRY(hipMemcpyAsync(p1.data(), ctx->p_pos1, 3 * c * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
            "download particles");
    int rc = from_device(ctx, ctx->p_Q0, q0);
    if (rc == OPX_OK) rc = from_device(ctx, ctx->p_Q1, q1);
    if (rc != OPX_OK) return rc;
    for (size_t k = 0; k < c; k++) {
        for (int i = 0; i < 3; i++) {
            if (pos0) pos0[3 * k + i] = p0[(size_t)i * c + k];
            if (pos1) pos1[3 * k + i] = p1[(size_t)i * c + k];
        }
        double a0[8], a1[8];
        for (int m = 0; m < ctx->AD; m++) {
            a0[m] = q0[(size_t)m * c + k];
            a1[m] = q1[(size_t)m * c + k];
        }
        double* o0 = Q0 ? Q0 + k * ctx->AS : nullptr;
        double* o1 = Q1 ? Q1 + k * ctx->AS : nullptr;
        if (ctx->nc == 2) {
            for (int m = 0; m < 3; m++) {
                if (o0) o0[m] = a0[m];
                if (o1) o1[m] = a1[m];
            }
        } else {
            if (o0) su3_coeffs_to_storage(a0, o0);
            if (o1) su3_coeffs_to_storage(a1, o1);
        }
    }
    return OPX_OK;
}

// the ordered deposits' work arrays for the context's particles (grown when the count grows)
static int particle_sort_work(opx_ctx* ctx, opx::ParticleSort* w)
{
    if ((uint64_t)ctx->n * 3 >= (1ull << 32))
        return fail(ctx, OPX_E_UNSUPPORTED, "particles: lattices of 2^32 / 3 cells or more");
    if (ctx->ps_capacity < ctx->p_count) {
        if (ctx->ps_buf) (void)hipFree(ctx->ps_buf);
        if (ctx->ps_temp) (void)hipFree(ctx->ps_temp);
        ctx->ps_buf = nullptr;
        ctx->ps_temp = nullptr;
        ctx->ps_capacity = 0;
        const size_t tb = opx::particle_sort_temp_bytes(ctx->p_count);
        if (tb == 0) return fail(ctx, OPX_E_DEVICE, "particles: radix sort storage query failed");
        OPX_TRY(hipMalloc(&ctx->ps_buf, (size_t)4 * ctx->p_count * sizeof(uint32_t)), "alloc particle sort");
        OPX_TRY(hipMalloc(&ctx->ps_temp, tb), "alloc particle sort");
        ctx->ps_temp_bytes = tb;
        ctx->ps_capacity = ctx->p_count;
    }
    const size_t c = ctx->ps_capacity;
    *w = opx::ParticleSort{ctx->ps_buf, ctx->ps_buf + c, ctx->ps_buf + 2 * c, ctx->ps_buf + 3 * c, ctx->ps_temp,
                           ctx->ps_temp_bytes};
    return OPX_OK;
}

int opx_particles_interpolate_chargedensity(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    int rc = materialize_rho(ctx);
    if (rc != OPX_OK) return rc;
    if (ctx->p_count == 0) return OPX_OK;
    opx::ParticleSort w;
    if ((rc = particle_sort_work(ctx, &w)) != OPX_OK) return rc;
    opx::Lattice L = lattice(ctx, ctx->at);
    OPX_TRY(opx::launch_p_chargedensity(ctx->nc, ctx->prec, L, p_view(ctx), w, ctx->rho, ctx->stream),
            "particle charge density");
    return OPX_OK;
}

int opx_particles_push(opx_ctx* ctx, double dt)
{
    if (!ctx) return OPX_E_ARG;
    opx::Lattice L = lattice(ctx, ctx->at);
    OPX_TRY(opx::launch_p_push(ctx->nc, ctx->prec, L, p_view(ctx), dt, ctx->Unext, ctx->stream), "particle push");
    return OPX_OK;
}

int opx_particles_interpolate_to_grid(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->p_count == 0) return OPX_OK;
    opx::ParticleSort w;
    int rc = particle_sort_work(ctx, &w);
    if (rc != OPX_OK) return rc;
    opx::Lattice L = lattice(ctx, ctx->at);
    OPX_TRY(opx::launch_p_to_grid(ctx->nc, ctx->prec, L, p_view(ctx), w, ctx->J, ctx->stream), "particle current");
    extend_current_planes(ctx, 0, ctx->dims[0]);
    return OPX_OK;
}

int opx_particles_reassign(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    std::swap(ctx->p_pos0, ctx->p_pos1);
    std::swap(ctx->p_Q0, ctx->p_Q1);
    return OPX_OK;
}

}  // extern "C"

namespace {

opx::SuperParticles sp_view(const opx_ctx* c)
{
    opx::SuperParticles sp;
    sp.count = c->sp_count;
    sp.max_num = c->sp_max_num;
    const int* d = c->sp_desc;
    const size_t k = (size_t)c->sp_count;
    sp.orientation = d;
    sp.num = d ? d + k : nullptr;
    sp.index_offset = d ? d + 2 * k : nullptr;
    sp.per_plane = d ? d + 3 * k : nullptr;
    sp.shift = d ? d + 4 * k : nullptr;
    sp.ppc = d ? d + 5 * k : nullptr;
    sp.q_offset = c->sp_qoff;
    sp.total = c->sp_total;
    return sp;
}

// CGCSuperParticle.getCurrentOffset / getCurrentNGPOffset on the host
// (physics/particles/CGCSuperParticle.java:71-93), used to bound the cells a
// charge-density gather has to visit.
int host_floor_div(int a, int b)
{
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}
int host_ngp_offset(int o, int off, int pp, int k, int ppc, int t)
{
    const int cur = o > 0 ? off + ((t + k) / ppc) * pp : off + host_floor_div(-t + k, ppc) * pp;
    if (o > 0) return cur + ((((t + k) % ppc) < ppc / 2) ? 0 : pp);
    return cur + ((((t - k + ppc + ppc / 2 - 1) % ppc) < ppc / 2) ? 0 : pp);
}

// The cells the super particles sit on at the deposit's time, cut into segments that
// one fixed set of particles covers (opx::CellSegments), written into the pinned
// staging buffer and copied to the device on the context stream. A particle spans at
// most as many segments as it has charges, so the lists hold at most sp_total entries.
int upload_cell_segments(opx_ctx* ctx, int t, opx::CellSegments* cs)
{
    const int cnt = ctx->sp_count;
    const int* h = ctx->sp_host.data();
    std::vector<int64_t> lo(cnt), hi(cnt), qoff(cnt);
    std::vector<int64_t> bp;
    bp.reserve(2 * (size_t)cnt);
    int64_t qo = 0;
    for (int p = 0; p < cnt; p++) {
        const int off = host_ngp_offset(h[p], h[2 * cnt + p], h[3 * cnt + p], h[4 * cnt + p], h[5 * cnt + p], t);
        qoff[p] = qo;
        qo += h[cnt + p];
        int64_t a = off, b = (int64_t)off + h[cnt + p];
        if (a < 0) a = 0;
        if (b > ctx->n) b = ctx->n;
        lo[p] = a;
        hi[p] = b;
        if (b > a) {
            bp.push_back(a);
            bp.push_back(b);
            qoff[p] -= off;   // charge of particle p on cell c: q[qoff[p] + c]
        }
    }
    std::sort(bp.begin(), bp.end());
    bp.erase(std::unique(bp.begin(), bp.end()), bp.end());
    const size_t nb = bp.size() > 0 ? bp.size() - 1 : 0;   // elementary segments [bp[s], bp[s+1])
    std::vector<uint32_t> covered(nb + 1, 0);               // particles per segment
    auto seg_of = [&](int64_t c) { return (size_t)(std::lower_bound(bp.begin(), bp.end(), c) - bp.begin()); };
    std::vector<size_t> s0(cnt), s1(cnt);
    for (int p = 0; p < cnt; p++) {
        if (hi[p] <= lo[p]) continue;
        s0[p] = seg_of(lo[p]);
        s1[p] = seg_of(hi[p]);
        for (size_t s = s0[p]; s < s1[p]; s++) covered[s]++;
    }
    // keep the covered segments; their lists in particle-list order
    std::vector<uint32_t> first, cells_prefix(1, 0), list_prefix(1, 0), kept(nb, UINT32_MAX);
    for (size_t s = 0; s < nb; s++) {
        if (covered[s] == 0) continue;
        kept[s] = (uint32_t)first.size();
        first.push_back((uint32_t)bp[s]);
        cells_prefix.push_back(cells_prefix.back() + (uint32_t)(bp[s + 1] - bp[s]));
        list_prefix.push_back(list_prefix.back() + covered[s]);
    }
    const size_t ns = first.size();
    const size_t entries = list_prefix.back();
    const size_t bytes = entries * sizeof(int64_t) + (3 * ns + 2) * sizeof(uint32_t);
    *cs = opx::CellSegments{};
    if (ns == 0) return OPX_OK;
    if (bytes > ctx->rs_capacity) {
        OPX_TRY(hipStreamSynchronize(ctx->stream), "charge density segments sync");
        if (ctx->rs_dev) (void)hipFree(ctx->rs_dev);
        ctx->rs_dev = nullptr;
        for (int k = 0; k < opx_ctx::kSegSlots; k++) {
            if (ctx->rs_host[k]) (void)hipHostFree(ctx->rs_host[k]);
            ctx->rs_host[k] = nullptr;
        }
        ctx->rs_capacity = 0;
        const size_t cap = bytes + bytes / 2 + 256;
        OPX_TRY(hipMalloc(&ctx->rs_dev, cap), "alloc charge density segments");
        for (int k = 0; k < opx_ctx::kSegSlots; k++)
            OPX_TRY(hipHostMalloc((void**)&ctx->rs_host[k], cap, hipHostMallocDefault), "alloc charge density staging");
        ctx->rs_capacity = cap;
    }
    // the slot's previous copy (kSegSlots deposits ago) must be done before it is rewritten
    const int slot = ctx->rs_slot;
    ctx->rs_slot = (slot + 1) % opx_ctx::kSegSlots;
    if (ctx->rs_copied[slot]) OPX_TRY(hipEventSynchronize(ctx->rs_copied[slot]), "charge density segments wait");
    else OPX_TRY(hipEventCreateWithFlags(&ctx->rs_copied[slot], hipEventDisableTiming), "charge density event");
    unsigned char* stage = ctx->rs_host[slot];
    int64_t* qb = (int64_t*)stage;
    uint32_t* u = (uint32_t*)(stage + entries * sizeof(int64_t));
    std::copy(first.begin(), first.end(), u);
    std::copy(cells_prefix.begin(), cells_prefix.end(), u + ns);
    std::copy(list_prefix.begin(), list_prefix.end(), u + 2 * ns + 1);
    std::vector<uint32_t> cursor(list_prefix.begin(), list_prefix.end() - 1);
    for (int p = 0; p < cnt; p++) {
        if (hi[p] <= lo[p]) continue;
        for (size_t s = s0[p]; s < s1[p]; s++) qb[cursor[kept[s]]++] = qoff[p];
    }
    OPX_TRY(hipMemcpyAsync(ctx->rs_dev, stage, bytes, hipMemcpyHostToDevice, ctx->stream),
            "upload charge density segments");
    OPX_TRY(hipEventRecord(ctx->rs_copied[slot], ctx->stream), "charge density event record");
    const uint32_t* ud = (const uint32_t*)(ctx->rs_dev + entries * sizeof(int64_t));
    cs->count = (uint32_t)ns;
    cs->cells = cells_prefix.back();
    cs->first = ud;
    cs->cells_prefix = ud + ns;
    cs->list_prefix = ud + 2 * ns + 1;
    cs->qbase = (const int64_t*)ctx->rs_dev;
    return OPX_OK;
}

int materialize_rho(opx_ctx* ctx)
{
    if (!ctx->rho_pending) return OPX_OK;
    opx::Lattice L = lattice(ctx, ctx->at);
    opx::CellSegments cs;
    int rc = upload_cell_segments(ctx, ctx->rho_pending_t, &cs);
    if (rc != OPX_OK) return rc;
    OPX_TRY(opx::launch_sp_chargedensity(ctx->nc, ctx->prec, L, ctx->sp_total, ctx->sp_q, ctx->rho, cs, ctx->stream),
            "charge density");
    ctx->rho_pending = false;
    return OPX_OK;
}

}  // namespace

extern "C" {

int opx_interpolate_to_grid(opx_ctx* ctx, int t)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->sp_count == 0) return OPX_OK;
    opx::Lattice L = lattice(ctx, ctx->at);
    OPX_TRY(opx::launch_sp_to_grid(ctx->nc, ctx->prec, L, sp_view(ctx), t, ctx->sp_q, ctx->Unext, ctx->J, ctx->stream),
            "interpolate to grid");
    // the planes the deposit can touch: every particle's current cells (the cells after the
    // ppc = 1 shift by one plane included), one plane of margin on either side
    const int cnt = ctx->sp_count;
    const int* h = ctx->sp_host.data();
    const int64_t s0 = (int64_t)ctx->dims[1] * ctx->dims[2];
    for (int p = 0; p < cnt; p++) {
        const int o = h[p], num = h[cnt + p], off = h[2 * cnt + p], pp = h[3 * cnt + p], k = h[4 * cnt + p],
                  ppc = h[5 * cnt + p];
        const int64_t cur = o > 0 ? off + (int64_t)((t + k) / ppc) * pp : off + (int64_t)host_floor_div(-t + k, ppc) * pp;
        const int64_t a = std::max<int64_t>(cur, 0), b = std::min<int64_t>(cur + num + (ppc == 1 ? pp : 0), ctx->n);
        if (a >= b) continue;
        extend_current_planes(ctx, (int)(a / s0 - 1), (int)((b + s0 - 1) / s0 + 1));
    }
    return OPX_OK;
}

int opx_interpolate_chargedensity(opx_ctx* ctx, int t)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->sp_count == 0) return OPX_OK;
    // rho is only read by measurements and downloads; the gather runs then,
    // with the charges as they are now (they change again in interpolateToGrid).
    // Deferral is exact only if nothing modifies the charges first, so it is
    // resolved immediately here.
    ctx->rho_pending = true;
    ctx->rho_pending_t = t;
    return materialize_rho(ctx);
}

int opx_simulation_initialize(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    int rc;
    if ((rc = opx_step_links(ctx, ctx->at)) != OPX_OK) return rc;
    if ((rc = opx_reset_charge(ctx)) != OPX_OK) return rc;
    if ((rc = opx_interpolate_chargedensity(ctx, 0)) != OPX_OK) return rc;
    if (ctx->p_count > 0) {
        if ((rc = opx_particles_interpolate_chargedensity(ctx)) != OPX_OK) return rc;
        if ((rc = opx_particles_push(ctx, ctx->at)) != OPX_OK) return rc;
    }
    if ((rc = opx_reset_current(ctx)) != OPX_OK) return rc;
    if ((rc = opx_interpolate_to_grid(ctx, 0)) != OPX_OK) return rc;
    if ((rc = opx_particles_interpolate_to_grid(ctx)) != OPX_OK) return rc;
    return OPX_OK;
}

int opx_simulation_step(opx_ctx* ctx, int n)
{
    if (!ctx) return OPX_E_ARG;
    if (n < 0) return fail(ctx, OPX_E_ARG, "simulation_step: negative count");
    int rc;
    for (int s = 0; s < n; s++) {
        ctx->steps++;
        const int t = ctx->steps;
        // rho is written every step and read by nothing inside a step (the solver reads J), so
        // within one call only the last step's charge density is observable: it alone is gathered
        const bool last = s == n - 1;
        if ((rc = opx_particles_reassign(ctx)) != OPX_OK) return rc;
        if ((rc = opx_store_fields(ctx)) != OPX_OK) return rc;
        {
            // every reunit_every-th step reprojects the new links, inside the update where the form can
            const bool reunit = ctx->reunit_every > 0 && t % ctx->reunit_every == 0;
            bool fused = false;
            opx::Lattice L = lattice(ctx, ctx->at);
            OPX_TRY(field_step(ctx, L, reunit, &fused), "step");
            if (reunit && !fused)
                OPX_TRY(opx::launch_reunitarize(ctx->nc, ctx->prec, (uint32_t)ctx->n, ctx->Unext, ctx->stream),
                        "step reunitarize");
        }
        if (ctx->sp_count > 0 || ctx->p_count > 0) {
            if (last) {
                if ((rc = opx_reset_charge(ctx)) != OPX_OK) return rc;
                if ((rc = opx_interpolate_chargedensity(ctx, t)) != OPX_OK) return rc;
                if (ctx->p_count > 0 && (rc = opx_particles_interpolate_chargedensity(ctx)) != OPX_OK) return rc;
            }
            if (ctx->p_count > 0 && (rc = opx_particles_push(ctx, ctx->at)) != OPX_OK) return rc;
            if ((rc = opx_reset_current(ctx)) != OPX_OK) return rc;
            if ((rc = opx_interpolate_to_grid(ctx, t)) != OPX_OK) return rc;
            if ((rc = opx_particles_interpolate_to_grid(ctx)) != OPX_OK) return rc;
        }
    }
    return OPX_OK;
}

int opx_simulation_steps_done(const opx_ctx* ctx) { return ctx ? ctx->steps : -1; }

int opx_check_health(opx_ctx* ctx, double tol, int64_t* non_finite, int64_t* non_unitary)
{
    if (!ctx) return OPX_E_ARG;
    if (!non_finite || !non_unitary) return fail(ctx, OPX_E_ARG, "health: null output");
    OPX_TRY(hipMemsetAsync(ctx->counters, 0, 4 * sizeof(unsigned long long), ctx->stream), "health reset");
    OPX_TRY(opx::launch_health(ctx->nc, ctx->prec, (uint32_t)ctx->n, ctx->U, tol, ctx->counters, ctx->stream),
            "health U");
    OPX_TRY(opx::launch_health(ctx->nc, ctx->prec, (uint32_t)ctx->n, ctx->Unext, tol, ctx->counters, ctx->stream),
            "health Unext");
    OPX_TRY(opx::launch_health_alg(ctx->nc, ctx->prec, (uint32_t)alg_count(ctx), ctx->E, ctx->counters, ctx->stream),
            "health E");
    unsigned long long h[4];
    OPX_TRY(hipMemcpyAsync(h, ctx->counters, sizeof h, hipMemcpyDeviceToHost, ctx->stream), "health copy");
    OPX_TRY(hipStreamSynchronize(ctx->stream), "health sync");
    *non_finite = (int64_t)h[0];
    *non_unitary = (int64_t)h[1];
    return (h[0] || h[1]) ? fail(ctx, OPX_E_NUMERIC, "health monitor: non-finite or non-unitary links") : OPX_OK;
}

}  // extern "C"

extern "C" {

int opx_set_option(opx_ctx* ctx, int option, int value)
{
    if (!ctx) return OPX_E_ARG;
    switch (option) {
    case OPX_OPT_KERNEL:
        if (value < OPX_KERNEL_AUTO || value > OPX_KERNEL_CELL) return fail(ctx, OPX_E_ARG, "option: bad kernel");
        if (value == OPX_KERNEL_TILE && !opx::tile_supported(ctx->nc))
            return fail(ctx, OPX_E_UNSUPPORTED, "option: the tile form is SU(2) only");
        ctx->kernel = value;
        return OPX_OK;
    case OPX_OPT_TILE:
        if (value < 0 || value >= opx::kStreamTiles) return fail(ctx, OPX_E_ARG, "option: bad tile shape");
        ctx->tile = value;
        return OPX_OK;
    case OPX_OPT_GLOBAL_POINTERS:
        if (value != 0 && value != 1) return fail(ctx, OPX_E_ARG, "option: global pointers is 0 or 1");
        ctx->force_global = value == 1;
        return OPX_OK;
    case OPX_OPT_CHUNKS:
        if (value < 0) return fail(ctx, OPX_E_ARG, "option: negative chunk count");
        ctx->nchunks = value;
        return OPX_OK;
    case OPX_OPT_REUNIT_EVERY:
        if (value < 0) return fail(ctx, OPX_E_ARG, "option: negative reprojection period");
        ctx->reunit_every = value;
        return OPX_OK;
    default:
        return fail(ctx, OPX_E_ARG, "option: unknown option");
    }
}

int opx_get_option(const opx_ctx* ctx, int option, int* value)
{
    if (!ctx || !value) return OPX_E_ARG;
    switch (option) {
    case OPX_OPT_KERNEL: *value = ctx->kernel; return OPX_OK;
    case OPX_OPT_TILE: *value = ctx->tile; return OPX_OK;
    case OPX_OPT_GLOBAL_POINTERS: *value = ctx->force_global ? 1 : 0; return OPX_OK;
    case OPX_OPT_REUNIT_EVERY: *value = ctx->reunit_every; return OPX_OK;
    case OPX_OPT_CHUNKS: {
        opx::Lattice L = lattice(ctx, ctx->at);
        *value = ctx->nchunks > 0 ? ctx->nchunks : (int)auto_chunks(ctx, L, ctx->tile);
        return OPX_OK;
    }
    default: return OPX_E_ARG;
    }
}

// the autotuner's second stage: the fastest candidates of the first, re-timed in rounds
constexpr size_t kAutotuneFinalists = 3;
constexpr int kAutotuneRounds = 4, kAutotuneRoundReps = 8;

int opx_autotune(opx_ctx* ctx, int reps, double* best_ms)
{
    if (!ctx) return OPX_E_ARG;
    if (reps < 1) return fail(ctx, OPX_E_ARG, "autotune: reps must be >= 1");
    opx::Lattice L = lattice(ctx, ctx->at);
    // candidates run on copies of E and two scratch link fields so the lattice state
    // is untouched; timed launches alternate between the scratch fields, so each
    // reads the links the previous one wrote, as consecutive leapfrog steps do
    // The scratch fields are allocated stream-ordered where the runtime can, so that releasing them
    // does not synchronize the device: a GPU left idle for a millisecond or more before sustained
    // load runs its first tens of launches at a clock that overshoots and then dips (measured up
    // to 30 % slower), so the caller's next launches should follow the autotuner's without a gap.
    void* e2 = nullptr;
    void* u2 = nullptr;
    void* u3 = nullptr;
    bool async_scratch[3] = {false, false, false};
    auto scratch_alloc = [&](void** p, size_t bytes, bool* async) -> hipError_t {
        if (hipMallocAsync(p, bytes, ctx->stream) == hipSuccess) {
            *async = true;
            return hipSuccess;
        }
        (void)hipGetLastError();
        *p = nullptr;
        return hipMalloc(p, bytes);
    };
    auto scratch_free = [&](void* p, bool async) {
        if (!p) return;
        if (async) (void)hipFreeAsync(p, ctx->stream);
        else (void)hipFree(p);
    };
    int rc = hip_check(ctx, scratch_alloc(&e2, alg_count(ctx) * ctx->tsz, &async_scratch[0]), "autotune alloc");
    if (rc == OPX_OK) rc = hip_check(ctx, scratch_alloc(&u2, link_count(ctx) * ctx->tsz, &async_scratch[1]), "autotune alloc");
    if (rc == OPX_OK) rc = hip_check(ctx, scratch_alloc(&u3, link_count(ctx) * ctx->tsz, &async_scratch[2]), "autotune alloc");
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (rc == OPX_OK) rc = hip_check(ctx, hipEventCreate(&ev0), "autotune event");
    if (rc == OPX_OK) rc = hip_check(ctx, hipEventCreate(&ev1), "autotune event");
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMemcpyAsync(e2, ctx->E, alg_count(ctx) * ctx->tsz, hipMemcpyDeviceToDevice, ctx->stream),
                       "autotune copy");
    // with an active region a launch writes the active links only, and the staples at the region's
    // edge read the rest: both scratch link fields start as the lattice's links
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMemcpyAsync(u2, ctx->Unext, link_count(ctx) * ctx->tsz, hipMemcpyDeviceToDevice,
                                           ctx->stream),
                       "autotune copy");
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMemcpyAsync(u3, ctx->U, link_count(ctx) * ctx->tsz, hipMemcpyDeviceToDevice, ctx->stream),
                       "autotune copy");
    const int saved_kernel = ctx->kernel, saved_tile = ctx->tile, saved_chunks = ctx->nchunks;
    double best = 1e30;
    int bk = OPX_KERNEL_LINK, bt = 0, bc = 0;
    struct Cand {
        int kernel, tile, chunks;
    };
    std::vector<Cand> cands;
    cands.push_back({OPX_KERNEL_LINK, 0, 0});
    cands.push_back({OPX_KERNEL_PLAQ, 0, 0});
    for (int t = 0; t < opx::kCellTiles; t++) cands.push_back({OPX_KERNEL_CELL, t, 0});
    if (opx::tile_supported(ctx->nc)) {
        const opx::TileGeom g = opx::tile_geom(ctx->prec, L, 1);
        const uint32_t tiles = g.tiles1 * g.tiles2;
        std::vector<int> ks;
        for (double f : {1.0, 2.0, 3.0, 4.0, 6.0, 8.0}) {
            int k = (int)((f * ctx->num_cus + tiles - 1) / tiles);
            if (k < 1) k = 1;
            if (k > (int)L.d0) k = (int)L.d0;
            bool dup = false;
            for (int q : ks) dup = dup || q == k;
            if (!dup) ks.push_back(k);
        }
        for (int k : ks) {
            cands.push_back({OPX_KERNEL_TILE, 0, k});
            cands.push_back({OPX_KERNEL_TILE, 1, k});
        }
    }
    for (int t = 0; t < opx::kPlaneShapes; t++)
        if (opx::plane_supported(ctx->nc, ctx->prec, L, t)) cands.push_back({OPX_KERNEL_PLANE, t, 0});
    for (int t = 0; t < opx::kStreamTiles; t++) {
        if (!opx::stream_supported(ctx->nc, ctx->prec, L, t)) continue;
        const uint32_t bpc = opx::stream_blocks_per_chunk(L, t);
        std::vector<int> ks;
        for (double f : {0.5, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0}) {
            int k = (int)(f * ctx->num_cus / (double)bpc + 0.5);
            if (k < 1) k = 1;
            if (k > (int)L.d0) k = (int)L.d0;
            bool dup = false;
            for (int q : ks) dup = dup || q == k;
            if (!dup) ks.push_back(k);
        }
        for (int k : ks) cands.push_back({OPX_KERNEL_STREAM, t, k});
    }
    // stage 1: every candidate, one untimed launch then `reps` timed back to back
    std::vector<std::pair<double, size_t>> timed;   // (ms per launch, candidate)
    for (size_t ci = 0; ci < cands.size() && rc == OPX_OK; ci++) {
        const Cand& cd = cands[ci];
        ctx->kernel = cd.kernel;
        ctx->tile = cd.tile;
        ctx->nchunks = cd.chunks;
        hipError_t e = launch_update(ctx, L, ctx->U, u2, e2);  // warm-up
        if (e != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        rc = hip_check(ctx, hipEventRecord(ev0, ctx->stream), "autotune record");
        for (int r = 0; r < reps && rc == OPX_OK; r++)
            rc = hip_check(ctx, launch_update(ctx, L, r % 2 ? u3 : u2, r % 2 ? u2 : u3, e2), "autotune launch");
        if (rc == OPX_OK) rc = hip_check(ctx, hipEventRecord(ev1, ctx->stream), "autotune record");
        if (rc == OPX_OK) rc = hip_check(ctx, hipEventSynchronize(ev1), "autotune sync");
        float ms = 0.f;
        if (rc == OPX_OK) rc = hip_check(ctx, hipEventElapsedTime(&ms, ev0, ev1), "autotune time");
        if (rc == OPX_OK) timed.push_back({ms / reps, ci});
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    // stage 2: the fastest few again, interleaved over several rounds with no host synchronization
    // in between (the GPU's clock moves over the first tens of milliseconds of sustained load, so a
    // single short sample per candidate can rank them wrongly); each keeps its fastest round
    std::sort(timed.begin(), timed.end());
    const size_t finalists = std::min<size_t>(timed.size(), kAutotuneFinalists);
    if (rc == OPX_OK && finalists > 0) {
        const int rounds = kAutotuneRounds, per = std::max(reps, kAutotuneRoundReps);
        std::vector<hipEvent_t> evs(2 * rounds * finalists, nullptr);
        for (size_t i = 0; i < evs.size() && rc == OPX_OK; i++)
            rc = hip_check(ctx, hipEventCreate(&evs[i]), "autotune event");
        for (int r = 0; r < rounds && rc == OPX_OK; r++)
            for (size_t f = 0; f < finalists && rc == OPX_OK; f++) {
                const Cand& cd = cands[timed[f].second];
                ctx->kernel = cd.kernel;
                ctx->tile = cd.tile;
                ctx->nchunks = cd.chunks;
                const size_t k = 2 * ((size_t)r * finalists + f);
                rc = hip_check(ctx, hipEventRecord(evs[k], ctx->stream), "autotune record");
                for (int q = 0; q < per && rc == OPX_OK; q++)
                    rc = hip_check(ctx, launch_update(ctx, L, q % 2 ? u3 : u2, q % 2 ? u2 : u3, e2), "autotune launch");
                if (rc == OPX_OK) rc = hip_check(ctx, hipEventRecord(evs[k + 1], ctx->stream), "autotune record");
            }
        // the scratch is released behind the last timed launch (before the wait for it)
        scratch_free(e2, async_scratch[0]);
        scratch_free(u2, async_scratch[1]);
        scratch_free(u3, async_scratch[2]);
        e2 = u2 = u3 = nullptr;
        if (rc == OPX_OK) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "autotune sync");
        for (size_t f = 0; f < finalists && rc == OPX_OK; f++) {
            double fbest = 1e30;
            for (int r = 0; r < rounds && rc == OPX_OK; r++) {
                const size_t k = 2 * ((size_t)r * finalists + f);
                float ms = 0.f;
                rc = hip_check(ctx, hipEventElapsedTime(&ms, evs[k], evs[k + 1]), "autotune time");
                fbest = std::min(fbest, (double)ms / per);
            }
            if (rc == OPX_OK && fbest < best) {
                const Cand& cd = cands[timed[f].second];
                best = fbest;
                bk = cd.kernel;
                bt = cd.tile;
                bc = cd.chunks;
            }
        }
        for (hipEvent_t e : evs)
            if (e) (void)hipEventDestroy(e);
    }
    scratch_free(e2, async_scratch[0]);
    scratch_free(u2, async_scratch[1]);
    scratch_free(u3, async_scratch[2]);
    if (rc != OPX_OK) {
        ctx->kernel = saved_kernel;
        ctx->tile = saved_tile;
        ctx->nchunks = saved_chunks;
        return rc;
    }
    ctx->kernel = bk;
    ctx->tile = bt;
    ctx->nchunks = bc;
    if (best_ms) *best_ms = best;
    return OPX_OK;
}

}  // extern "C"

extern "C" {

int opx_gauss_project(opx_ctx* ctx, double tol, int max_iter, int* iterations, double* gauss_before,
                      double* gauss_after)
{
    if (!ctx) return OPX_E_ARG;
    if (!(tol > 0) || max_iter < 1) return fail(ctx, OPX_E_ARG, "gauss_project: tol must be > 0 and max_iter >= 1");
    int rc = materialize_rho(ctx);
    if (rc != OPX_OK) return rc;
    opx::Lattice L = lattice(ctx, ctx->at);
    // the solver's double work arrays are [component][cell]: ms per-cell, ml per-link algebra values
    const uint64_t ms = (uint64_t)ctx->n * ctx->AD;
    const uint64_t ml = 3 * ms;
    double *Ed = nullptr, *F = nullptr, *lam = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr, *scr = nullptr;
    double* bufs_sz[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const size_t sizes[7] = {ml, ml, ms, ms, ms, ms, opx::gauss_dot_scratch()};
    double** ptrs[7] = {&Ed, &F, &lam, &r, &p, &Ap, &scr};
    for (int i = 0; i < 7 && rc == OPX_OK; i++) {
        rc = hip_check(ctx, hipMalloc((void**)ptrs[i], sizes[i] * sizeof(double)), "gauss alloc");
        bufs_sz[i] = *ptrs[i];
    }
    auto dot = [&](const double* a, const double* b, uint64_t m, double* out) -> int {
        int r2 = hip_check(ctx, opx::gauss_dot(m, a, b, scr, scr + 2 * 1024 + 2, ctx->stream), "gauss dot");
        if (r2 == OPX_OK)
            r2 = hip_check(ctx, hipMemcpyAsync(out, scr + 2 * 1024 + 2, sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                           "gauss dot copy");
        if (r2 == OPX_OK) r2 = hip_check(ctx, hipStreamSynchronize(ctx->stream), "gauss dot sync");
        return r2;
    };
    double rr = 0, rr0 = 0;
    int it = 0;
    if (rc == OPX_OK) rc = hip_check(ctx, opx::gauss_to_double(ctx->prec, ml, (uint32_t)ctx->n, ctx->E, Ed, ctx->stream), "gauss E");
    // r = G = D^T E - rho ; lambda = 0 ; p = r
    if (rc == OPX_OK)
        rc = hip_check(ctx, opx::gauss_apply_Dt(ctx->nc, ctx->prec, L, ctx->U, Ed, ctx->rho, r, ctx->stream), "gauss G");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMemsetAsync(lam, 0, ms * sizeof(double), ctx->stream), "gauss zero");
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMemcpyAsync(p, r, ms * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream), "gauss copy");
    if (rc == OPX_OK) rc = dot(r, r, ms, &rr);
    rr0 = rr;
    if (gauss_before) *gauss_before = rr0 / (double)ctx->n;
    while (rc == OPX_OK && it < max_iter && rr > tol * tol * rr0 && rr > 0) {
        double pAp = 0;
        rc = hip_check(ctx, opx::gauss_apply_D(ctx->nc, ctx->prec, L, ctx->U, p, F, ctx->stream), "gauss D");
        if (rc == OPX_OK)
            rc = hip_check(ctx, opx::gauss_apply_Dt(ctx->nc, ctx->prec, L, ctx->U, F, nullptr, Ap, ctx->stream), "gauss DtD");
        if (rc == OPX_OK) rc = dot(p, Ap, ms, &pAp);
        if (rc != OPX_OK || !(pAp > 0)) break;
        const double alpha = rr / pAp;
        rc = hip_check(ctx, opx::gauss_axpy(ms, alpha, p, lam, ctx->stream), "gauss axpy");
        if (rc == OPX_OK) rc = hip_check(ctx, opx::gauss_axpy(ms, -alpha, Ap, r, ctx->stream), "gauss axpy");
        double rr_new = 0;
        if (rc == OPX_OK) rc = dot(r, r, ms, &rr_new);
        if (rc == OPX_OK) rc = hip_check(ctx, opx::gauss_xpby(ms, r, rr_new / rr, p, ctx->stream), "gauss xpby");
        rr = rr_new;
        it++;
    }
    // E' = E - D lambda, then the Gauss residual of the stored field
    if (rc == OPX_OK) rc = hip_check(ctx, opx::gauss_apply_D(ctx->nc, ctx->prec, L, ctx->U, lam, F, ctx->stream), "gauss D");
    if (rc == OPX_OK) rc = hip_check(ctx, opx::gauss_sub_into(ctx->prec, ml, (uint32_t)ctx->n, Ed, F, ctx->E, ctx->stream), "gauss update");
    if (rc == OPX_OK) rc = hip_check(ctx, opx::gauss_to_double(ctx->prec, ml, (uint32_t)ctx->n, ctx->E, Ed, ctx->stream), "gauss E");
    if (rc == OPX_OK)
        rc = hip_check(ctx, opx::gauss_apply_Dt(ctx->nc, ctx->prec, L, ctx->U, Ed, ctx->rho, r, ctx->stream), "gauss G");
    double g2 = 0;
    if (rc == OPX_OK) rc = dot(r, r, ms, &g2);
    if (gauss_after) *gauss_after = g2 / (double)ctx->n;
    if (iterations) *iterations = it;
    for (int i = 0; i < 7; i++)
        if (bufs_sz[i]) (void)hipFree(bufs_sz[i]);
    return rc;
}

int opx_gauss_restore(opx_ctx* ctx, double gamma, int max_iterations, double absolute_value, int* iterations,
                      int* reason, double* violation_before, double* violation_after)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "restore the events of a batch through the batch");
    if (!std::isfinite(gamma) || !std::isfinite(absolute_value))
        return fail(ctx, OPX_E_ARG, "gauss_restore: gamma and the absolute value must be finite");
    // the reference reads one lattice spacing (Grid.getLatticeSpacing throws on uneven grids)
    if (ctx->as[0] != ctx->as[1] || ctx->as[0] != ctx->as[2])
        return fail(ctx, OPX_E_UNSUPPORTED, "gauss_restore: uneven lattice spacings (Use getGridStep(int direction).)");
    int rc = materialize_rho(ctx);
    if (rc != OPX_OK) return rc;
    opx::Lattice L = lattice(ctx, ctx->at);
    const uint64_t ms = (uint64_t)ctx->n * ctx->AD;
    const size_t eb = alg_count(ctx) * ctx->tsz, lb = link_count(ctx) * ctx->tsz;
    double *G = nullptr, *R = nullptr, *scr = nullptr;
    void *E0 = nullptr, *Un0 = nullptr;
    rc = hip_check(ctx, hipMalloc(&G, ms * sizeof(double)), "restore alloc");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&R, ms * sizeof(double)), "restore alloc");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&scr, (opx::gauss_dot_scratch() + 1) * sizeof(double)), "restore alloc");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&E0, eb), "restore alloc");
    if (rc == OPX_OK) rc = hip_check(ctx, hipMalloc(&Un0, lb), "restore alloc");
    double* total_dev = scr + opx::gauss_dot_scratch();
    const double factor = 1.0 / (ctx->at * ctx->as[0]);
    const double f2 = -0.5 * ctx->as[0] * gamma;
    // computeGaussViolation: G, and the mean violation read back to the host
    auto violation = [&](double* out) -> int {
        int r2 = hip_check(ctx, opx::restore_violation(ctx->nc, ctx->prec, L, ctx->U, ctx->Unext, ctx->rho, factor, G,
                                                       R, ctx->stream),
                           "restore violation");
        if (r2 == OPX_OK) r2 = hip_check(ctx, opx::gauss_dot(ms, R, R, scr, total_dev, ctx->stream), "restore sum");
        double v = 0;
        if (r2 == OPX_OK)
            r2 = hip_check(ctx, hipMemcpyAsync(&v, total_dev, sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                           "restore copy");
        if (r2 == OPX_OK) r2 = hip_check(ctx, hipStreamSynchronize(ctx->stream), "restore sync");
        *out = v / (double)ctx->n;
        return r2;
    };
    double total = 0, old = 0;
    int kept = 0, why = 3;
    if (rc == OPX_OK) rc = violation(&total);
    old = total;
    if (violation_before) *violation_before = total;
    // iterateRestorationAlgorithm (GaussConstraintRestoration.java:79-112)
    for (int i = 0; rc == OPX_OK && i < max_iterations; i++) {
        if (total < absolute_value) {
            why = 0;
            break;
        }
        // backupGrid: a correction changes E and, through updateLinks, Unext
        rc = hip_check(ctx, hipMemcpyAsync(E0, ctx->E, eb, hipMemcpyDeviceToDevice, ctx->stream), "restore backup");
        if (rc == OPX_OK)
            rc = hip_check(ctx, hipMemcpyAsync(Un0, ctx->Unext, lb, hipMemcpyDeviceToDevice, ctx->stream),
                           "restore backup");
        if (rc == OPX_OK)
            rc = hip_check(ctx, opx::restore_correct(ctx->nc, ctx->prec, L, ctx->U, G, f2, ctx->E, ctx->stream),
                           "restore correction");
        if (rc == OPX_OK)
            rc = hip_check(ctx, opx::launch_step_links(ctx->nc, ctx->prec, L, ctx->active, ctx->U, ctx->Unext, ctx->E,
                                                       ctx->stream),
                           "restore updateLinks");
        if (rc == OPX_OK) rc = violation(&total);
        if (rc != OPX_OK) break;
        if (old - total < 0) {
            // restoreGrid: the last correction made the violation worse
            rc = hip_check(ctx, hipMemcpyAsync(ctx->E, E0, eb, hipMemcpyDeviceToDevice, ctx->stream), "restore undo");
            if (rc == OPX_OK)
                rc = hip_check(ctx, hipMemcpyAsync(ctx->Unext, Un0, lb, hipMemcpyDeviceToDevice, ctx->stream),
                               "restore undo");
            if (rc == OPX_OK) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "restore sync");
            why = 1;
            break;
        }
        old = total;
        kept = i + 1;
        if (i == max_iterations - 1) why = 2;
    }
    if (violation_after) *violation_after = old;
    if (iterations) *iterations = kept;
    if (reason) *reason = why;
    for (void* b : {(void*)G, (void*)R, (void*)scr, E0, Un0})
        if (b) (void)hipFree(b);
    return rc;
}

int opx_unitarity_test(opx_ctx* ctx, double tol, int64_t* non_unitary, int64_t* cells, int64_t max_cells)
{
    if (!ctx) return OPX_E_ARG;
    if (!non_unitary) return fail(ctx, OPX_E_ARG, "unitarity test: null count");
    if (max_cells < 0 || (max_cells > 0 && !cells)) return fail(ctx, OPX_E_ARG, "unitarity test: bad cell list");
    const size_t nl = (size_t)ctx->n * 3;
    unsigned char* flags = nullptr;
    int rc = hip_check(ctx, hipMalloc(&flags, nl), "unitarity alloc");
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMemsetAsync(ctx->counters + 2, 0, sizeof(unsigned long long), ctx->stream),
                       "unitarity reset");
    if (rc == OPX_OK)
        rc = hip_check(ctx, opx::unitarity_flags(ctx->nc, ctx->prec, (uint32_t)ctx->n, ctx->U, ctx->Unext, tol, flags,
                                                 ctx->counters + 2, ctx->stream),
                       "unitarity test");
    unsigned long long cnt = 0;
    if (rc == OPX_OK)
        rc = hip_check(ctx, hipMemcpyAsync(&cnt, ctx->counters + 2, sizeof cnt, hipMemcpyDeviceToHost, ctx->stream),
                       "unitarity copy");
    if (rc == OPX_OK) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "unitarity sync");
    if (rc == OPX_OK && cnt > 0 && max_cells > 0) {
        // the flagged links in the reference's loop order (cell, then direction)
        std::vector<unsigned char> h(nl);
        rc = hip_check(ctx, hipMemcpy(h.data(), flags, nl, hipMemcpyDeviceToHost), "unitarity flags copy");
        int64_t k = 0;
        for (size_t i = 0; rc == OPX_OK && i < nl && k < max_cells; i++)
            if (h[i]) cells[k++] = (int64_t)(i / 3);
    }
    if (flags) (void)hipFree(flags);
    *non_unitary = (int64_t)cnt;
    return rc;
}

}  // extern "C"

// ============================================================== batched events
struct opx_batch {
    int nc = 2, prec = OPX_FP32, count = 0;
    int G = 4, S = 4, AD = 3;
    size_t tsz = 4;
    void* UA = nullptr;
    void* UB = nullptr;
    void* E = nullptr;
    uint64_t link_total = 0, alg_total = 0;
    bool small_events = true;   // every event's link field below 4 GiB (buffer loads in the cell form)
    std::vector<opx_ctx*> views;
    std::vector<opx::BatchEvent> ev;
    opx::BatchEvent* dev_ev = nullptr;
    uint32_t total_items = 0;
    uint32_t grid = 0;
    int parity = 0;       // 0: U is UA; 1: U is UB
    int steps = 0;
    int num_cus = 256;
    hipStream_t own_stream = nullptr, stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0, graph_reunit = 0, graph_parity = -1;
    int form = 0;                      // 0: per-link kernel, 1: LDS-tile kernel
    opx::TileItem* dev_items = nullptr;
    uint32_t nitems = 0;
    std::string err;
};

namespace {

int bfail(opx_batch* b, int code, const std::string& m)
{
    if (b) b->err = m;
    return code;
}

int bcheck(opx_batch* b, hipError_t e, const char* what)
{
    if (e == hipSuccess) return OPX_OK;
    return bfail(b, OPX_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define OPX_BTRY(expr, what)                          \
    do {                                              \
        int _rc = bcheck(b, (expr), (what));          \
        if (_rc != OPX_OK) return _rc;                \
    } while (0)

void batch_point_views(opx_batch* b)
{
    for (int e = 0; e < b->count; e++) {
        opx_ctx* v = b->views[e];
        char* ua = (char*)b->UA + b->ev[e].link_offset * b->tsz;
        char* ub = (char*)b->UB + b->ev[e].link_offset * b->tsz;
        v->U = b->parity ? ub : ua;
        v->Unext = b->parity ? ua : ub;
        v->E = (char*)b->E + b->ev[e].alg_offset * b->tsz;
        v->stream = b->stream;
        v->steps = b->steps;
    }
}

// one batch step: storeFields (parity flip) and the fused update of every event
int batch_step(opx_batch* b, int reunit_every)
{
    b->parity ^= 1;
    void* U = b->parity ? b->UB : b->UA;
    void* Un = b->parity ? b->UA : b->UB;
    // every reunit_every-th step reprojects the new links: inside the cell form's update, else a pass
    const bool reunit = reunit_every > 0 && (b->steps + 1) % reunit_every == 0;
    bool fused = false;
    if (b->form == 3) {
        OPX_BTRY(opx::launch_batch_cell(b->nc, b->prec, b->dev_ev, b->count, b->total_items, b->grid, b->small_events,
                                        U, Un, b->E, reunit, b->stream),
                 "batch cell update");
        fused = reunit;
    } else if (b->form >= 1)
        OPX_BTRY(opx::launch_batch_tile(b->prec, b->form == 1, b->dev_ev, b->dev_items, b->nitems, b->grid, U, Un,
                                        b->E, b->stream),
                 "batch tile update");
    else
        OPX_BTRY(opx::launch_batch_update(b->nc, b->prec, b->dev_ev, b->count, b->total_items, b->grid, U, Un, b->E,
                                          b->stream),
                 "batch update");
    b->steps++;
    if (reunit && !fused)
        OPX_BTRY(opx::launch_batch_reunitarize(b->nc, b->prec, b->dev_ev, b->count, b->total_items, b->grid, Un,
                                               b->stream),
                 "batch reunitarize");
    return OPX_OK;
}

}  // namespace

extern "C" {

int opx_batch_create(int num_colors, int count, const int* dims, const double* spacing, double time_step,
                     double coupling, int precision, opx_batch** out)
{
    if (!out || !dims || !spacing || count < 1) return OPX_E_ARG;
    *out = nullptr;
    if (num_colors != 2 && num_colors != 3) return OPX_E_UNSUPPORTED;
    if (precision != OPX_FP64 && precision != OPX_FP32) return OPX_E_UNSUPPORTED;
    opx_batch* b = new opx_batch();
    b->nc = num_colors;
    b->prec = precision;
    b->count = count;
    b->G = num_colors == 2 ? 4 : 18;
    b->S = stored_reals(num_colors, precision);
    b->AD = num_colors == 2 ? 3 : 8;
    b->tsz = precision == OPX_FP64 ? 8 : 4;
    int rc = OPX_OK;
    uint64_t lo = 0, ao = 0;
    uint32_t items = 0;
    for (int e = 0; e < count && rc == OPX_OK; e++) {
        opx_ctx* v = new opx_ctx();
        v->batch_view = true;
        v->nc = num_colors;
        v->prec = precision;
        v->G = b->G;
        v->S = b->S;
        v->AD = b->AD;
        v->AS = num_colors == 2 ? 3 : 9;
        v->tsz = b->tsz;
        int64_t n = 1;
        for (int i = 0; i < 3; i++) {
            v->dims[i] = dims[3 * e + i];
            v->as[i] = spacing[i];
            if (v->dims[i] < 1) rc = bfail(b, OPX_E_ARG, "batch: event dimensions must be >= 1");
            n *= v->dims[i];
        }
        v->n = n;
        v->at = time_step;
        v->g = coupling;
        b->views.push_back(v);
        opx::BatchEvent E;
        E.L = lattice(v, time_step);
        E.link_offset = lo;
        E.alg_offset = ao;
        E.item_start = items;
        E.pad = 0;
        b->ev.push_back(E);
        // each event's fields in the device block layout (opx::fpos), padded to whole 64-cell blocks
        const uint64_t np = opx::padded_cells((uint64_t)n);
        lo += np * 3 * b->S;
        ao += np * 3 * b->AD;
        if (np * 3 * b->S * b->tsz >= (1ull << 32)) b->small_events = false;
        const uint64_t it = 3 * (((uint64_t)n + 255) / 256);
        if ((uint64_t)items + it >= (1ull << 32)) rc = bfail(b, OPX_E_UNSUPPORTED, "batch: too many cells");
        items += (uint32_t)it;
    }
    b->link_total = lo;
    b->alg_total = ao;
    b->total_items = items;
    if (rc == OPX_OK) rc = bcheck(b, hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking), "batch stream");
    b->stream = b->own_stream;
    if (rc == OPX_OK) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            b->num_cus = cus;
        b->grid = 4 * (uint32_t)b->num_cus;
    }
    if (rc == OPX_OK) rc = bcheck(b, hipMalloc(&b->UA, lo * b->tsz), "batch alloc U");
    if (rc == OPX_OK) rc = bcheck(b, hipMalloc(&b->UB, lo * b->tsz), "batch alloc Unext");
    if (rc == OPX_OK) rc = bcheck(b, hipMalloc(&b->E, ao * b->tsz), "batch alloc E");
    if (rc == OPX_OK) rc = bcheck(b, hipMalloc(&b->dev_ev, b->ev.size() * sizeof(opx::BatchEvent)), "batch alloc events");
    if (rc == OPX_OK)
        rc = bcheck(b, hipMemcpy(b->dev_ev, b->ev.data(), b->ev.size() * sizeof(opx::BatchEvent), hipMemcpyHostToDevice),
                    "batch upload events");
    if (rc == OPX_OK && opx::tile_supported(num_colors)) {
        // tile-form work items: per event, chunks of up to 8 planes, tiles fastest
        std::vector<opx::TileItem> items_h;
        const int ty = precision == OPX_FP64 ? opx::kTileTY64 : opx::kTileTY32, tx = opx::kTileTX;
        for (int e = 0; e < count; e++) {
            const opx::Lattice& L = b->ev[e].L;
            const opx::TileGeom g = opx::tile_geom(precision, L, 1);
            const uint32_t nch = (L.d0 + 7) / 8;
            for (uint32_t ch = 0; ch < nch; ch++)
                for (uint32_t t1 = 0; t1 < g.tiles1; t1++)
                    for (uint32_t t2 = 0; t2 < g.tiles2; t2++)
                        items_h.push_back(opx::TileItem{(uint32_t)e, t1 * (uint32_t)ty, t2 * (uint32_t)tx,
                                                        (uint32_t)(((uint64_t)ch * L.d0) / nch),
                                                        (uint32_t)(((uint64_t)(ch + 1) * L.d0) / nch)});
        }
        b->nitems = (uint32_t)items_h.size();
        rc = bcheck(b, hipMalloc(&b->dev_items, items_h.size() * sizeof(opx::TileItem)), "batch alloc items");
        if (rc == OPX_OK)
            rc = bcheck(b, hipMemcpy(b->dev_items, items_h.data(), items_h.size() * sizeof(opx::TileItem),
                                     hipMemcpyHostToDevice),
                        "batch upload items");
    }
    if (rc == OPX_OK) rc = bcheck(b, hipMemset(b->E, 0, ao * b->tsz), "batch zero E");
    for (int e = 0; e < count && rc == OPX_OK; e++) {
        opx_ctx* v = b->views[e];
        v->num_cus = b->num_cus;
        rc = bcheck(b, hipMalloc(&v->partials, opx::measure_partials_size((uint32_t)v->n) * sizeof(double)), "view alloc");
        if (rc == OPX_OK) rc = bcheck(b, hipMalloc(&v->dout, 64 * sizeof(double)), "view alloc");
        if (rc == OPX_OK) rc = bcheck(b, hipMalloc(&v->counters, 4 * sizeof(unsigned long long)), "view alloc");
    }
    if (rc == OPX_OK) {
        batch_point_views(b);
        // identity links in both buffers (physics/grid/Cell.java:63-68)
        for (int e = 0; e < count && rc == OPX_OK; e++) {
            opx_ctx* v = b->views[e];
            const std::vector<double> ident = identity_links(v);
            rc = to_device(v, v->U, ident);
            if (rc == OPX_OK) rc = to_device(v, v->Unext, ident);
            if (rc != OPX_OK) b->err = v->err;
        }
    }
    if (rc != OPX_OK) {
        std::fprintf(stderr, "opx_batch_create: %s\n", b->err.c_str());
        opx_batch_destroy(b);
        return rc;
    }
    *out = b;
    return OPX_OK;
}

void opx_batch_destroy(opx_batch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (opx_ctx* v : b->views) opx_destroy(v);
    void* bufs[] = {b->UA, b->UB, b->E, b->dev_ev, b->dev_items};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
    delete b;
}

const char* opx_batch_last_error(const opx_batch* b) { return b ? b->err.c_str() : "null batch"; }

opx_ctx* opx_batch_event(opx_batch* b, int event)
{
    if (!b || event < 0 || event >= b->count) return nullptr;
    return b->views[event];
}

int opx_batch_set_stream(opx_batch* b, void* hip_stream)
{
    if (!b) return OPX_E_ARG;
    OPX_BTRY(hipStreamSynchronize(b->stream), "batch set_stream sync");
    b->stream = hip_stream ? (hipStream_t)hip_stream : b->own_stream;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    batch_point_views(b);
    return OPX_OK;
}

int opx_batch_step_links(opx_batch* b, double dt)
{
    if (!b) return OPX_E_ARG;
    for (opx_ctx* v : b->views) {
        opx::Lattice L = lattice(v, dt);
        OPX_BTRY(opx::launch_step_links(v->nc, v->prec, L, v->active, v->U, v->Unext, v->E, b->stream),
                 "batch step_links");
    }
    return OPX_OK;
}

int opx_batch_set_option(opx_batch* b, int option, int value)
{
    if (!b) return OPX_E_ARG;
    if (option == OPX_BATCH_OPT_GRID) {
        if (value < 1) return bfail(b, OPX_E_ARG, "batch option: grid must be >= 1");
        b->grid = (uint32_t)value;
    } else if (option == OPX_BATCH_OPT_FORM) {
        if (value < 0 || value > 3) return bfail(b, OPX_E_ARG, "batch option: form must be 0, 1, 2 or 3");
        if ((value == 1 || value == 2) && !b->dev_items)
            return bfail(b, OPX_E_UNSUPPORTED, "batch option: the tile form is SU(2) only");
        b->form = value;
    } else {
        return bfail(b, OPX_E_ARG, "batch option: unknown option");
    }
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return OPX_OK;
}

int opx_batch_get_option(const opx_batch* b, int option, int* value)
{
    if (!b || !value) return OPX_E_ARG;
    if (option == OPX_BATCH_OPT_GRID) *value = (int)b->grid;
    else if (option == OPX_BATCH_OPT_FORM) *value = b->form;
    else return OPX_E_ARG;
    return OPX_OK;
}

int opx_batch_leapfrog(opx_batch* b, int steps, int reunit_every, int use_graph)
{
    if (!b) return OPX_E_ARG;
    if (steps < 0 || reunit_every < 0) return bfail(b, OPX_E_ARG, "batch leapfrog: negative count");
    int rc = OPX_OK;
    // a captured period: reunit_every steps (rounded up to an even count so the
    // U / Unext roles are the same at the start of every replay)
    int period = reunit_every > 0 ? reunit_every : 2;
    if (period % 2) period *= 2;
    int done = 0;
    // leading steps until the clock is aligned with the period
    while (done < steps && (!use_graph || b->steps % period != 0) && rc == OPX_OK) {
        rc = batch_step(b, reunit_every);
        done++;
    }
    if (use_graph && rc == OPX_OK && steps - done >= period) {
        if (!b->graph || b->graph_steps != period || b->graph_reunit != reunit_every || b->graph_parity != b->parity) {
            if (b->graph) {
                (void)hipGraphExecDestroy(b->graph);
                b->graph = nullptr;
            }
            const int saved_parity = b->parity, saved_steps = b->steps;
            hipGraph_t g = nullptr;
            rc = bcheck(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "batch capture");
            for (int s = 0; s < period && rc == OPX_OK; s++) rc = batch_step(b, reunit_every);
            hipError_t ee = hipStreamEndCapture(b->stream, &g);
            if (rc == OPX_OK) rc = bcheck(b, ee, "batch end capture");
            if (rc == OPX_OK) rc = bcheck(b, hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0), "batch instantiate");
            if (g) (void)hipGraphDestroy(g);
            b->parity = saved_parity;
            b->steps = saved_steps;
            b->graph_steps = period;
            b->graph_reunit = reunit_every;
            b->graph_parity = saved_parity;
        }
        while (rc == OPX_OK && steps - done >= period) {
            rc = bcheck(b, hipGraphLaunch(b->graph, b->stream), "batch graph launch");
            b->steps += period;  // period is even: parity unchanged
            done += period;
        }
    }
    while (rc == OPX_OK && done < steps) {
        rc = batch_step(b, reunit_every);
        done++;
    }
    batch_point_views(b);
    return rc;
}

int opx_batch_measure(opx_batch* b, double* out)
{
    if (!b || !out) return OPX_E_ARG;
    for (int e = 0; e < b->count; e++) {
        int rc = opx_measure(b->views[e], out + (size_t)OPX_M_COUNT * e);
        if (rc != OPX_OK) return bfail(b, rc, b->views[e]->err);
    }
    return OPX_OK;
}

int opx_batch_sync(opx_batch* b)
{
    if (!b) return OPX_E_ARG;
    OPX_BTRY(hipStreamSynchronize(b->stream), "batch sync");
    return OPX_OK;
}

int opx_batch_autotune(opx_batch* b, int reps, double* best_ms)
{
    if (!b) return OPX_E_ARG;
    if (reps < 1) return bfail(b, OPX_E_ARG, "batch autotune: reps must be >= 1");
    void* e2 = nullptr;
    void* u2 = nullptr;
    void* u3 = nullptr;
    OPX_BTRY(hipMalloc(&e2, b->alg_total * b->tsz), "batch autotune alloc");
    int rc = bcheck(b, hipMalloc(&u2, b->link_total * b->tsz), "batch autotune alloc");
    if (rc == OPX_OK) rc = bcheck(b, hipMalloc(&u3, b->link_total * b->tsz), "batch autotune alloc");
    if (rc == OPX_OK) rc = bcheck(b, hipMemcpyAsync(e2, b->E, b->alg_total * b->tsz, hipMemcpyDeviceToDevice, b->stream), "batch autotune copy");
    hipEvent_t a0 = nullptr, a1 = nullptr;
    if (rc == OPX_OK) rc = bcheck(b, hipEventCreate(&a0), "batch autotune event");
    if (rc == OPX_OK) rc = bcheck(b, hipEventCreate(&a1), "batch autotune event");
    void* U = b->parity ? b->UB : b->UA;
    double best = 1e30;
    uint32_t bg = b->grid;
    int bf = b->form;
    struct Cand {
        int form;
        uint32_t grid;
    };
    std::vector<Cand> cands;
    for (int m : {1, 2, 3, 4, 6, 8, 12, 16, 24, 32}) cands.push_back({0, (uint32_t)(m * b->num_cus)});
    if (b->dev_items)
        for (int m : {1, 2, 3, 4, 6, 8})
            for (int f : {1, 2}) cands.push_back({f, (uint32_t)(m * b->num_cus)});
    for (int m : {1, 2, 3, 4, 6, 8, 12, 16}) cands.push_back({3, (uint32_t)(m * b->num_cus)});
    // first launch from the batch's links, timed launches alternating between two scratch fields
    auto run = [&](const Cand& c, const void* src, void* dst) {
        if (c.form == 3)
            return opx::launch_batch_cell(b->nc, b->prec, b->dev_ev, b->count, b->total_items, c.grid, b->small_events, src, dst, e2,
                                          false, b->stream);
        return c.form >= 1 ? opx::launch_batch_tile(b->prec, c.form == 1, b->dev_ev, b->dev_items, b->nitems, c.grid,
                                                    src, dst, e2, b->stream)
                           : opx::launch_batch_update(b->nc, b->prec, b->dev_ev, b->count, b->total_items, c.grid, src,
                                                      dst, e2, b->stream);
    };
    for (const Cand& c : cands) {
        if (rc != OPX_OK) break;
        const uint32_t grid = c.grid;
        rc = bcheck(b, run(c, U, u2), "batch autotune");
        if (rc == OPX_OK) rc = bcheck(b, hipEventRecord(a0, b->stream), "batch autotune record");
        for (int r = 0; r < reps && rc == OPX_OK; r++)
            rc = bcheck(b, run(c, r % 2 ? u3 : u2, r % 2 ? u2 : u3), "batch autotune");
        if (rc == OPX_OK) rc = bcheck(b, hipEventRecord(a1, b->stream), "batch autotune record");
        if (rc == OPX_OK) rc = bcheck(b, hipEventSynchronize(a1), "batch autotune sync");
        float ms = 0;
        if (rc == OPX_OK) rc = bcheck(b, hipEventElapsedTime(&ms, a0, a1), "batch autotune time");
        if (rc == OPX_OK && ms / reps < best) {
            best = ms / reps;
            bg = grid;
            bf = c.form;
        }
    }
    if (a0) (void)hipEventDestroy(a0);
    if (a1) (void)hipEventDestroy(a1);
    (void)hipFree(e2);
    if (u2) (void)hipFree(u2);
    if (u3) (void)hipFree(u3);
    if (rc != OPX_OK) return rc;
    if (bg != b->grid || bf != b->form) {
        b->grid = bg;
        b->form = bf;
        if (b->graph) {
            (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
        }
    }
    if (best_ms) *best_ms = best;
    return OPX_OK;
}

int opx_batch_get_info(const opx_batch* b, int* count, int64_t* total_cells, int* grid, int* steps_done)
{
    if (!b) return OPX_E_ARG;
    if (count) *count = b->count;
    if (total_cells) {
        int64_t t = 0;
        for (const opx_ctx* v : b->views) t += v->n;
        *total_cells = t;
    }
    if (grid) *grid = (int)b->grid;
    if (steps_done) *steps_done = b->steps;
    return OPX_OK;
}

}  // extern "C"
