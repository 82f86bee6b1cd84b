// opx_adjoint_capi.cpp -- opx_adjoint_gradient: the SU(2) leapfrog, its loss
// (lattice average of the transverse-minus-longitudinal energy density) and the
// reverse sweep with checkpoints and recomputation.
//
// Forward (as opx_step_links + opx_leapfrog): U^0 = exp(-at E^0) U_init, then per
// step E^{n+1} = E^n + at F(U^n), U^{n+1} = exp(-at E^{n+1}) U^n, and in the FP32
// form a reprojection of U^{n+1} onto SU(2) after every `reunit_every`-th step.
// (U^n, E^n) is kept at every multiple of `checkpoint_every`. The reverse sweep
// walks the segments from the last one: it recomputes the segment's states from
// its checkpoint into a (checkpoint_every + 1)-slot buffer, then for each step
// n of the segment, from the last:
//   Ebar += -at J_exp^T(Ubar U^n+),  Ubar <- exp(-at E^{n+1})^+ Ubar   (link step)
//   Ubar += at (dF/dU)^T(U^n) Ebar                                   (force)
// plus the loss's own dependence on U^{N-1} after step N-1. The reprojection
// is not transposed: it only moves a link along itself, and a cotangent along
// U never reaches Ebar because |exp(w)| = 1 for every w.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "opx.h"
#include "opx_ctx.h"
#include "opx_kernels.h"

namespace opx_host {

struct AdjointWork {
    size_t link_bytes = 0, alg_bytes = 0;
    int ck_slots = 0, seg_slots = 0;
    void* Uinit = nullptr;
    void* Ubar = nullptr;
    void* Ebar = nullptr;
    void* ckU = nullptr;    // ck_slots link fields
    void* ckE = nullptr;    // ck_slots algebra fields
    void* segU = nullptr;   // seg_slots link fields
    void* segE = nullptr;   // seg_slots algebra fields
};

void adjoint_release(opx_ctx* c)
{
    if (!c || !c->adj) return;
    AdjointWork* w = c->adj;
    void* bufs[] = {w->Uinit, w->Ubar, w->Ebar, w->ckU, w->ckE, w->segU, w->segE};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete w;
    c->adj = nullptr;
}

namespace {

#define ADJ_TRY(expr, what)                              \
    do {                                                 \
        int _rc = hip_check(ctx, (expr), (what));        \
        if (_rc != OPX_OK) return _rc;                   \
    } while (0)

int ensure_work(opx_ctx* ctx, int ck_slots, int seg_slots)
{
    const size_t cells = opx::padded_cells((uint64_t)ctx->n);   // fields in the device block layout
    const size_t lb = cells * 3 * 4 * ctx->tsz, ab = cells * 3 * 3 * ctx->tsz;
    AdjointWork* w = ctx->adj;
    if (w && w->link_bytes == lb && w->ck_slots >= ck_slots && w->seg_slots >= seg_slots) return OPX_OK;
    adjoint_release(ctx);
    w = new AdjointWork();
    ctx->adj = w;
    w->link_bytes = lb;
    w->alg_bytes = ab;
    ADJ_TRY(hipMalloc(&w->Uinit, lb), "adjoint alloc");
    ADJ_TRY(hipMalloc(&w->Ubar, lb), "adjoint alloc");
    ADJ_TRY(hipMalloc(&w->Ebar, ab), "adjoint alloc");
    ADJ_TRY(hipMalloc(&w->ckU, lb * ck_slots), "adjoint checkpoint alloc");
    ADJ_TRY(hipMalloc(&w->ckE, ab * ck_slots), "adjoint checkpoint alloc");
    ADJ_TRY(hipMalloc(&w->segU, lb * seg_slots), "adjoint segment alloc");
    ADJ_TRY(hipMalloc(&w->segE, ab * seg_slots), "adjoint segment alloc");
    w->ck_slots = ck_slots;   // set once every buffer exists, so a failed allocation is retried
    w->seg_slots = seg_slots;
    return OPX_OK;
}

void* slot(void* base, size_t bytes, int k) { return (char*)base + bytes * (size_t)k; }

}  // namespace
}  // namespace opx_host

extern "C" int opx_adjoint_gradient(opx_ctx* ctx, int steps, int reunit_every, int checkpoint_every, int direction,
                                    double* loss, double* grad, int64_t n_values)
{
    using namespace opx_host;
    if (!ctx) return OPX_E_ARG;
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "step the batch, not one of its events");
    if (ctx->nc != 2) return fail(ctx, OPX_E_UNSUPPORTED, "adjoint: SU(2) only");
    if (!ctx->active.all || !ctx->eval.all || ctx->j_nonzero)
        return fail(ctx, OPX_E_UNSUPPORTED,
                    "adjoint: pure Yang-Mills on the whole lattice only (no active or evaluation region, J = 0)");
    if (steps < 1 || reunit_every < 0 || checkpoint_every < 1)
        return fail(ctx, OPX_E_ARG, "adjoint: steps >= 1, reunit_every >= 0, checkpoint_every >= 1");
    if (direction < 0 || direction > 2) return fail(ctx, OPX_E_ARG, "adjoint: bad longitudinal direction");
    if (!loss) return fail(ctx, OPX_E_ARG, "adjoint: null loss");
    if (grad && n_values != ctx->n * 3 * 3) return fail(ctx, OPX_E_ARG, "adjoint: gradient size mismatch");
    if (3 * (uint64_t)ctx->n >= (1ull << 32)) return fail(ctx, OPX_E_UNSUPPORTED, "adjoint: lattice too large");

    const int K = checkpoint_every;
    const int nseg = (steps + K - 1) / K;
    int rc = ensure_work(ctx, grad ? nseg : 1, grad ? K + 1 : 1);
    if (rc != OPX_OK) return rc;
    AdjointWork* w = ctx->adj;
    const size_t LB = w->link_bytes, AB = w->alg_bytes;
    const opx::Lattice L = lattice(ctx, ctx->at);
    const double at = ctx->at;
    const uint32_t n = (uint32_t)ctx->n;
    hipStream_t st = ctx->stream;
    const int base = ctx->steps;

    // ---- forward, keeping checkpoints
    ADJ_TRY(hipMemcpyAsync(w->Uinit, ctx->U, LB, hipMemcpyDeviceToDevice, st), "adjoint copy");
    ADJ_TRY(hipMemcpyAsync(w->ckE, ctx->E, AB, hipMemcpyDeviceToDevice, st), "adjoint copy");
    ADJ_TRY(opx::launch_step_links(ctx->nc, ctx->prec, L, ctx->active, ctx->U, ctx->Unext, ctx->E, st), "adjoint step links");
    ADJ_TRY(hipMemcpyAsync(w->ckU, ctx->Unext, LB, hipMemcpyDeviceToDevice, st), "adjoint copy");
    for (int k = 0; k < steps; k++) {
        std::swap(ctx->U, ctx->Unext);
        ADJ_TRY(launch_update(ctx, L, ctx->U, ctx->Unext, ctx->E), "adjoint forward step");
        if (reunit_every > 0 && (base + k + 1) % reunit_every == 0)
            ADJ_TRY(opx::launch_reunitarize(ctx->nc, ctx->prec, n, ctx->Unext, st), "adjoint reunitarize");
        if (grad && (k + 1) % K == 0 && k + 1 < steps) {
            const int s = (k + 1) / K;
            ADJ_TRY(hipMemcpyAsync(slot(w->ckU, LB, s), ctx->Unext, LB, hipMemcpyDeviceToDevice, st), "adjoint copy");
            ADJ_TRY(hipMemcpyAsync(slot(w->ckE, AB, s), ctx->E, AB, hipMemcpyDeviceToDevice, st), "adjoint copy");
        }
    }
    ctx->steps = base + steps;

    // ---- loss: lattice average of e_T - e_L (ProjectedEnergyDensity components)
    {
        const int nd = ctx->dims[direction];
        std::vector<double> tel(nd), tmag(nd), lel(nd), lmag(nd);
        rc = opx_projected_energy_density(ctx, direction, tel.data(), tmag.data(), lel.data(), lmag.data());
        if (rc != OPX_OK) return rc;
        double s = 0.0;
        for (int k = 0; k < nd; k++) s += tel[k] + tmag[k] - lel[k] - lmag[k];
        *loss = s / nd;
    }
    if (!grad) return OPX_OK;

    // ---- seed: dL/dE^N and dL/dU^N
    double coefE[3], coefB[3];
    for (int j = 0; j < 3; j++) {
        const double sgn = j == direction ? -1.0 : 1.0;
        const double inv = 1.0 / ((double)ctx->n * ctx->g * ctx->g * ctx->as[j] * ctx->as[j]);
        const int k1 = (j + 1) % 3, k2 = (j + 2) % 3;
        const double f = ctx->as[j] / (ctx->as[k1] * ctx->as[k2]);
        coefE[j] = sgn * inv;
        coefB[j] = sgn * f * f * inv / 4.0;
    }
    ADJ_TRY(opx::launch_adj_seed(ctx->prec, n, ctx->E, w->Ebar, w->Ubar, coefE, st), "adjoint seed");
    ADJ_TRY(opx::launch_adj_bsq(ctx->prec, L, ctx->Unext, w->Ubar, coefB, st), "adjoint seed");

    // ---- reverse sweep over the segments
    for (int s = nseg - 1; s >= 0; s--) {
        const int n0 = s * K, len = std::min(K, steps - n0);
        ADJ_TRY(hipMemcpyAsync(w->segU, slot(w->ckU, LB, s), LB, hipMemcpyDeviceToDevice, st), "adjoint copy");
        ADJ_TRY(hipMemcpyAsync(w->segE, slot(w->ckE, AB, s), AB, hipMemcpyDeviceToDevice, st), "adjoint copy");
        for (int k = 1; k <= len; k++) {
            void* Ek = slot(w->segE, AB, k);
            ADJ_TRY(hipMemcpyAsync(Ek, slot(w->segE, AB, k - 1), AB, hipMemcpyDeviceToDevice, st), "adjoint copy");
            ADJ_TRY(launch_update(ctx, L, slot(w->segU, LB, k - 1), slot(w->segU, LB, k), Ek), "adjoint recompute");
            if (reunit_every > 0 && (base + n0 + k) % reunit_every == 0)
                ADJ_TRY(opx::launch_reunitarize(ctx->nc, ctx->prec, n, slot(w->segU, LB, k), st), "adjoint reunitarize");
        }
        for (int k = len; k >= 1; k--) {
            const void* Un = slot(w->segU, LB, k - 1);
            ADJ_TRY(opx::launch_adj_link_step(ctx->prec, n, Un, slot(w->segE, AB, k), w->Ubar, w->Ebar, at, st),
                    "adjoint link step");
            ADJ_TRY(opx::launch_adj_force(ctx->prec, L, Un, w->Ebar, w->Ubar, at, st), "adjoint force");
            if (n0 + k == steps) ADJ_TRY(opx::launch_adj_bsq(ctx->prec, L, Un, w->Ubar, coefB, st), "adjoint loss term");
        }
    }
    // the initial half step U^0 = exp(-at E^0) U_init
    ADJ_TRY(opx::launch_adj_link_step(ctx->prec, n, w->Uinit, w->ckE, w->Ubar, w->Ebar, at, st), "adjoint link step");

    // ---- dL/dE^0 to the host, per cell and direction (reference element order)
    std::vector<char> h(AB);
    ADJ_TRY(hipMemcpyAsync(h.data(), w->Ebar, AB, hipMemcpyDeviceToHost, st), "adjoint download");
    ADJ_TRY(hipStreamSynchronize(st), "adjoint sync");
    for (uint32_t c = 0; c < n; c++)
        for (int d = 0; d < 3; d++)
            for (int k = 0; k < 3; k++) {
                const size_t i = opx::fpos(d * 3 + k, c, 9, ctx->prec == OPX_FP64 ? 8u : 4u);
                grad[((size_t)c * 3 + d) * 3 + k] =
                    ctx->prec == OPX_FP64 ? ((const double*)h.data())[i] : (double)((const float*)h.data())[i];
            }
    return OPX_OK;
}
