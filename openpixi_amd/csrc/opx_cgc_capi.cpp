// opx_cgc_capi.cpp -- the CGC initial condition with the beam along x0
// (physics/initial/CGC/CGCInitialCondition.java:55-93): MVModel charges, the
// LightConePoissonSolverImprovedFull solve and the LightConeNGPSuperParticleCreator
// particles, on the device. The charge density of the next solve and the Gauss
// violation of the last one stay on the device (double, [coefficient][cell])
// between the calls; transient work arrays live for one call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "opx.h"
#include "opx_ctx.h"
#include "opx_kernels.h"

namespace opx_host {

struct CgcWork {
    double* rho = nullptr;     // charge density of the next light-cone solve
    double* gauss = nullptr;   // Gauss violation of the last light-cone solve
    double* VT = nullptr;      // [real][N1 N2] Wilson lines of the last solve on the plane behind the nucleus
};

void cgc_release(opx_ctx* c)
{
    if (!c || !c->cgc) return;
    if (c->cgc->rho) (void)hipFree(c->cgc->rho);
    if (c->cgc->gauss) (void)hipFree(c->cgc->gauss);
    if (c->cgc->VT) (void)hipFree(c->cgc->VT);
    delete c->cgc;
    c->cgc = nullptr;
}

namespace {

#define CGC_TRY(expr, what)                              \
    do {                                                 \
        int _rc = hip_check(ctx, (expr), (what));        \
        if (_rc != OPX_OK) return _rc;                   \
    } while (0)

// device buffers of one call, freed (after the stream drains) when it returns
struct Scratch {
    opx_ctx* ctx;
    std::vector<void*> bufs;
    explicit Scratch(opx_ctx* c) : ctx(c) {}
    ~Scratch()
    {
        if (bufs.empty()) return;
        (void)hipStreamSynchronize(ctx->stream);
        for (void* b : bufs) (void)hipFree(b);
    }
    template <class P>
    hipError_t alloc(P** p, size_t count)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(P));
        if (e == hipSuccess) {
            bufs.push_back(q);
            *p = (P*)q;
        }
        return e;
    }
};

int cgc_check(opx_ctx* ctx)
{
    if (ctx->batch_view) return fail(ctx, OPX_E_STATE, "CGC initial condition: not available on a batch event");
    for (int d = 0; d < 3; d++)
        if (!opx::cgc_fft_supported((uint32_t)ctx->dims[d]))
            return fail(ctx, OPX_E_UNSUPPORTED, "CGC initial condition: at most 4096 cells per axis");
    if (!ctx->cgc) ctx->cgc = new CgcWork();
    return OPX_OK;
}

int ensure(opx_ctx* ctx, double** p)
{
    if (*p) return OPX_OK;
    void* q = nullptr;
    CGC_TRY(hipMalloc(&q, (size_t)ctx->AD * ctx->n * sizeof(double)), "alloc CGC buffer");
    *p = (double*)q;
    return OPX_OK;
}

// transform of AD blocks of cells along `axes`; sign -1 forward, +1 inverse (unnormalized)
int transform(opx_ctx* ctx, Scratch& s, double2* x, std::initializer_list<int> axes, int sign)
{
    const uint64_t total = (uint64_t)ctx->AD * ctx->n;
    for (int a : axes) {
        const uint32_t len = (uint32_t)ctx->dims[a];
        if (len == 1) continue;
        const uint32_t stride = a == 0 ? (uint32_t)(ctx->dims[1] * ctx->dims[2]) : (a == 1 ? (uint32_t)ctx->dims[2] : 1u);
        double2* tw = nullptr;
        CGC_TRY(s.alloc(&tw, len), "alloc twiddles");
        CGC_TRY(opx::cgc_twiddles(len, tw, ctx->stream), "twiddles");
        CGC_TRY(opx::cgc_fft_axis(x, total, len, stride, tw, sign, ctx->stream), "transform");
    }
    return OPX_OK;
}

// [coefficient][cell] doubles on the device -> [cell][A] reference storage on the host
int download_alg(opx_ctx* ctx, const double* dev, double* host)
{
    const size_t n = (size_t)ctx->n, ad = (size_t)ctx->AD;
    std::vector<double> soa(ad * n);
    CGC_TRY(hipMemcpyAsync(soa.data(), dev, soa.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
            "download");
    CGC_TRY(hipStreamSynchronize(ctx->stream), "download sync");
    for (size_t c = 0; c < n; c++) {
        double cf[8];
        for (size_t k = 0; k < ad; k++) cf[k] = soa[k * n + c];
        coeffs_to_storage(ctx->nc, cf, host + c * ctx->AS);
    }
    return OPX_OK;
}

// the unit Gaussians of java.util.Random(seed) into dev[0, count) on the context's stream
int java_noise(opx_ctx* ctx, int64_t seed, uint64_t count, double* dev)
{
    const uint64_t scrambled = ((uint64_t)seed ^ 0x5DEECE66DULL) & ((1ULL << 48) - 1);
    const uint64_t need = (count + 1) / 2;
    // about pi/4 of the trials are accepted; more are drawn until enough are
    for (uint64_t trials = need + need / 4 + need / 32 + 4096;; trials *= 2) {
        if (trials >= (1ULL << 32)) return fail(ctx, OPX_E_UNSUPPORTED, "java.util.Random: too many values");
        void* w = nullptr;
        CGC_TRY(hipMalloc(&w, opx::java_gaussian_work(trials)), "alloc java.util.Random work");
        uint64_t accepted = 0;
        const hipError_t e = opx::java_gaussians(scrambled, count, trials, w, dev, &accepted, ctx->stream);
        (void)hipFree(w);
        CGC_TRY(e, "java.util.Random Gaussians");
        if (accepted >= need) return OPX_OK;
    }
}

int mv_charge(opx_ctx* ctx, double location, double width, double mu, double uv_cutoff, double coherence_length,
              double ir_coefficient, const double* noise, int64_t n_noise, bool seeded, int64_t seed, double* rho_out,
              int64_t n_rho)
{
    if (!ctx) return OPX_E_ARG;
    int rc = cgc_check(ctx);
    if (rc != OPX_OK) return rc;
    if (!seeded && !noise) return fail(ctx, OPX_E_ARG, "MV model: null noise buffer");
    if (!seeded && n_noise != (int64_t)ctx->AD * ctx->n)
        return fail(ctx, OPX_E_ARG, "MV model: noise must hold (colours^2 - 1) values per cell");
    if (rho_out && n_rho != ctx->n * ctx->AS) return fail(ctx, OPX_E_ARG, "MV model: charge buffer size mismatch");
    if (!(width > 0)) return fail(ctx, OPX_E_ARG, "MV model: the longitudinal width must be positive");
    if ((rc = ensure(ctx, &ctx->cgc->rho)) != OPX_OK) return rc;
    Scratch s(ctx);
    const uint64_t m = (uint64_t)ctx->AD * ctx->n;
    double* dn = nullptr;
    double2* x = nullptr;
    CGC_TRY(s.alloc(&dn, m), "alloc noise");
    CGC_TRY(s.alloc(&x, m), "alloc transform buffer");
    if (seeded) {
        if ((rc = java_noise(ctx, seed, m, dn)) != OPX_OK) return rc;
    } else {
        CGC_TRY(hipMemcpyAsync(dn, noise, m * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "upload noise");
    }
    const opx::Lattice L = lattice(ctx, ctx->at);
    const double aL = ctx->as[0], aT = ctx->as[1];
    const double norm = 1.0 / (width * std::sqrt(2 * M_PI)), i2s2 = 0.5 / (width * width);
    const double rcw = mu * ctx->g / std::sqrt(aL * aT * aT);
    const double box = ctx->dims[0] * aL;
    const double zmin = std::max(location - box / 2.0, 0.0), zmax = std::min(location + box / 2.0, box);
    const int lmin = (int)std::floor(zmin / aL), lmax = (int)std::ceil(zmax / aL);
    CGC_TRY(opx::cgc_mv_fill(L, m, dn, rcw, location, norm, i2s2, x, ctx->stream), "MV charges");
    if ((rc = transform(ctx, s, x, {0, 1, 2}, -1)) != OPX_OK) return rc;
    CGC_TRY(opx::cgc_mv_regulate(L, m, uv_cutoff * uv_cutoff, coherence_length, ir_coefficient * ir_coefficient,
                                 1.0 / (double)ctx->n, x, ctx->stream),
            "MV regulation");
    if ((rc = transform(ctx, s, x, {0, 1, 2}, +1)) != OPX_OK) return rc;
    CGC_TRY(opx::cgc_mv_store(L, ctx->active, m, lmin, lmax, x, ctx->cgc->rho, ctx->stream), "MV store");
    if (rho_out) return download_alg(ctx, ctx->cgc->rho, rho_out);
    CGC_TRY(hipStreamSynchronize(ctx->stream), "MV sync");
    return OPX_OK;
}

}  // namespace
}  // namespace opx_host

extern "C" {

int opx_cgc_mv_charge(opx_ctx* ctx, double location, double width, double mu, double uv_cutoff,
                      double coherence_length, double ir_coefficient, const double* noise, int64_t n_noise,
                      double* rho_out, int64_t n_rho)
{
    return opx_host::mv_charge(ctx, location, width, mu, uv_cutoff, coherence_length, ir_coefficient, noise, n_noise,
                               false, 0, rho_out, n_rho);
}

int opx_cgc_mv_charge_seeded(opx_ctx* ctx, double location, double width, double mu, double uv_cutoff,
                             double coherence_length, double ir_coefficient, int64_t seed, double* rho_out,
                             int64_t n_rho)
{
    return opx_host::mv_charge(ctx, location, width, mu, uv_cutoff, coherence_length, ir_coefficient, nullptr, 0,
                               true, seed, rho_out, n_rho);
}

int opx_java_gaussians(int64_t seed, int64_t count, double* out)
{
    if (count < 0 || (count > 0 && !out)) return OPX_E_ARG;
    if (count == 0) return OPX_OK;
    opx_ctx tmp;   // stream and error slot only
    void* dev = nullptr;
    if (hipMalloc(&dev, (size_t)count * sizeof(double)) != hipSuccess) return OPX_E_DEVICE;
    int rc = opx_host::java_noise(&tmp, seed, (uint64_t)count, (double*)dev);
    if (rc == OPX_OK && hipMemcpy(out, dev, (size_t)count * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = OPX_E_DEVICE;
    (void)hipFree(dev);
    return rc;
}

int opx_cgc_set_charge(opx_ctx* ctx, const double* rho, int64_t n_values)
{
    using namespace opx_host;
    if (!ctx) return OPX_E_ARG;
    int rc = cgc_check(ctx);
    if (rc != OPX_OK) return rc;
    if (!rho) return fail(ctx, OPX_E_ARG, "CGC charge: null buffer");
    if (n_values != ctx->n * ctx->AS) return fail(ctx, OPX_E_ARG, "CGC charge: buffer size mismatch");
    if ((rc = ensure(ctx, &ctx->cgc->rho)) != OPX_OK) return rc;
    const size_t n = (size_t)ctx->n;
    std::vector<double> soa((size_t)ctx->AD * n);
    for (size_t c = 0; c < n; c++) {
        double cf[8];
        storage_to_coeffs(ctx->nc, rho + c * ctx->AS, cf);
        for (int k = 0; k < ctx->AD; k++) soa[(size_t)k * n + c] = cf[k];
    }
    CGC_TRY(hipMemcpyAsync(ctx->cgc->rho, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
            "upload charge");
    CGC_TRY(hipStreamSynchronize(ctx->stream), "upload charge sync");
    return OPX_OK;
}

int opx_cgc_lightcone_solve(opx_ctx* ctx, int orientation, double* gauss_out, int64_t n_values)
{
    using namespace opx_host;
    if (!ctx) return OPX_E_ARG;
    int rc = cgc_check(ctx);
    if (rc != OPX_OK) return rc;
    if (orientation != 1 && orientation != -1) return fail(ctx, OPX_E_ARG, "light-cone solve: orientation must be +1 or -1");
    if (!ctx->cgc->rho) return fail(ctx, OPX_E_STATE, "light-cone solve: no charge density set");
    if (gauss_out && n_values != ctx->n * ctx->AS) return fail(ctx, OPX_E_ARG, "light-cone solve: buffer size mismatch");
    if ((rc = ensure(ctx, &ctx->cgc->gauss)) != OPX_OK) return rc;
    Scratch s(ctx);
    const uint64_t m = (uint64_t)ctx->AD * ctx->n;
    const size_t gn = (size_t)ctx->G * ctx->n;
    double2* x = nullptr;
    double *phi = nullptr, *Wz = nullptr, *Vm = nullptr, *Vp = nullptr;
    CGC_TRY(s.alloc(&x, m), "alloc transform buffer");
    CGC_TRY(s.alloc(&phi, m), "alloc phi");
    CGC_TRY(s.alloc(&Wz, gn), "alloc Wilson steps");
    CGC_TRY(s.alloc(&Vm, gn), "alloc Wilson lines");
    CGC_TRY(s.alloc(&Vp, gn), "alloc Wilson lines");
    const opx::Lattice L = lattice(ctx, ctx->at);
    // transverse Poisson equation on every x0 sheet
    CGC_TRY(opx::cgc_to_complex(m, ctx->cgc->rho, x, ctx->stream), "Poisson fill");
    if ((rc = transform(ctx, s, x, {1, 2}, -1)) != OPX_OK) return rc;
    CGC_TRY(opx::cgc_poisson_scale(L, m, 1.0 / ((double)ctx->dims[1] * ctx->dims[2]), x, ctx->stream), "Poisson");
    if ((rc = transform(ctx, s, x, {1, 2}, +1)) != OPX_OK) return rc;
    CGC_TRY(opx::cgc_real_part(m, x, phi, ctx->stream), "Poisson store");
    // Wilson lines at -at/2 and +at/2, the fields and the Gauss violation of the Wilson-line copy
    const int mfrac = (int)(32 * ctx->at / ctx->as[0] * 0.5);
    CGC_TRY(opx::cgc_wilson_lines(ctx->nc, L, orientation, mfrac, phi, Wz, Vm, Vp, ctx->stream), "Wilson lines");
    // the lines on the plane "far behind" the nucleus (LightConePoissonSolverImprovedFull.java:157-167)
    const size_t T = (size_t)ctx->dims[1] * ctx->dims[2];
    if (!ctx->cgc->VT) CGC_TRY(hipMalloc((void**)&ctx->cgc->VT, (size_t)ctx->G * T * sizeof(double)), "alloc VT");
    const size_t zb = orientation > 0 ? 0 : (size_t)ctx->dims[0] - 1;
    for (int k = 0; k < ctx->G; k++)
        CGC_TRY(hipMemcpyAsync(ctx->cgc->VT + k * T, Vm + (size_t)k * ctx->n + zb * T, T * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream),
                "copy VT");
    CGC_TRY(opx::cgc_links(ctx->nc, ctx->prec, L, Vm, Vp, ctx->U, ctx->Unext, ctx->E, ctx->stream), "light-cone links");
    CGC_TRY(opx::cgc_gauss(ctx->nc, ctx->prec, L, ctx->active, Vm, Vp, ctx->U, ctx->Unext, ctx->rho, ctx->cgc->gauss,
                           ctx->stream),
            "light-cone Gauss violation");
    if (gauss_out) return download_alg(ctx, ctx->cgc->gauss, gauss_out);
    CGC_TRY(hipStreamSynchronize(ctx->stream), "light-cone sync");
    return OPX_OK;
}

}  // extern "C"

namespace opx_host {
namespace {

// the planes that carry charge (LightConeNGPSuperParticleCreator.java:108-143,
// LightConeNGPParticleCreator.java:116-150): the first plane whose largest |G| exceeds 1e-9 of
// the global largest, and the last such plane below the final one
int charge_block(opx_ctx* ctx, Scratch& s, const opx::Lattice& L, int* zs_out, int* ze_out)
{
    const int Lz = ctx->dims[0];
    double* dmx = nullptr;
    CGC_TRY(s.alloc(&dmx, (size_t)Lz), "alloc plane maxima");
    CGC_TRY(opx::cgc_plane_max(ctx->nc, L, ctx->cgc->gauss, dmx, ctx->stream), "plane maxima");
    std::vector<double> mx((size_t)Lz);
    CGC_TRY(hipMemcpyAsync(mx.data(), dmx, mx.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
            "download plane maxima");
    CGC_TRY(hipStreamSynchronize(ctx->stream), "plane maxima sync");
    double gmax = 0.0;
    for (double v : mx) gmax = std::max(gmax, v);
    const double cut = gmax * 10E-10;
    int zs = 0, ze = Lz - 1;
    for (int z = 0; z < Lz; z++)
        if (mx[z] > cut) {
            zs = z;
            break;
        }
    for (int z = ze - 1; z >= 0; z--)
        if (mx[z] > cut) {
            ze = z;
            break;
        }
    *zs_out = zs;
    *ze_out = ze;
    return OPX_OK;
}

}  // namespace
}  // namespace opx_host

extern "C" {

int opx_particles_count(const opx_ctx* ctx, int64_t* count)
{
    if (!ctx || !count) return OPX_E_ARG;
    *count = ctx->p_count;
    return OPX_OK;
}

int opx_cgc_create_particles(opx_ctx* ctx, int orientation, int64_t* added, int* z_start, int* z_end)
{
    using namespace opx_host;
    if (!ctx) return OPX_E_ARG;
    int rc = cgc_check(ctx);
    if (rc != OPX_OK) return rc;
    if (orientation != 1 && orientation != -1) return fail(ctx, OPX_E_ARG, "particles: orientation must be +1 or -1");
    if (!ctx->cgc->gauss) return fail(ctx, OPX_E_STATE, "particles: no light-cone solve yet");
    const int ppc = (int)std::floor(ctx->as[0] / ctx->at + 0.5);
    if (ppc < 1) return fail(ctx, OPX_E_ARG, "particles: particlesPerCell = round(aL / at) must be >= 1");
    const opx::Lattice L = lattice(ctx, ctx->at);
    const int64_t nT = (int64_t)ctx->dims[1] * ctx->dims[2];
    Scratch s(ctx);
    int zs = 0, ze = 0;
    if ((rc = charge_block(ctx, s, L, &zs, &ze)) != OPX_OK) return rc;
    if (z_start) *z_start = zs;
    if (z_end) *z_end = ze;
    opx::SpBlock b;
    b.zs = zs;
    b.bw = ze - zs + 1;
    b.ppc = ppc;
    b.ws = 0;
    b.nl = 0;
    const uint64_t old = ctx->p_count, fresh = (uint64_t)b.bw * nT * ppc, total = old + fresh;
    if (total >= (1ull << 32)) return fail(ctx, OPX_E_UNSUPPORTED, "particles: too many particles");
    const size_t ad = (size_t)ctx->AD, tsz = ctx->tsz;
    // new arrays: the existing particles' rows copied over, the new ones appended
    int* dir = nullptr;
    double *p0 = nullptr, *p1 = nullptr, *v = nullptr;
    void *q0 = nullptr, *q1 = nullptr;
    auto drop = [&]() {
        void* bufs[] = {dir, p0, p1, v, q0, q1};
        for (void* x : bufs)
            if (x) (void)hipFree(x);
    };
    hipError_t e = hipMalloc((void**)&dir, total * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&p0, 3 * total * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&p1, 3 * total * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&v, 3 * total * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&q0, ad * total * tsz);
    if (e == hipSuccess) e = hipMalloc(&q1, ad * total * tsz);
    if (old > 0) {
        if (e == hipSuccess)
            e = hipMemcpyAsync(dir, ctx->p_dir, old * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream);
        const double* src[3] = {ctx->p_pos0, ctx->p_pos1, ctx->p_vel};
        double* dst[3] = {p0, p1, v};
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 3 && e == hipSuccess; k++)
                e = hipMemcpyAsync(dst[a] + k * total, src[a] + k * old, old * sizeof(double), hipMemcpyDeviceToDevice,
                                   ctx->stream);
        for (size_t k = 0; k < ad && e == hipSuccess; k++)
            e = hipMemcpyAsync((char*)q0 + k * total * tsz, (const char*)ctx->p_Q0 + k * old * tsz, old * tsz,
                               hipMemcpyDeviceToDevice, ctx->stream);
    }
    double* work = nullptr;
    if (e == hipSuccess) e = s.alloc(&work, opx::cgc_sp_work(b, (uint32_t)nT, ctx->AD));
    if (e == hipSuccess)
        e = opx::cgc_particles(ctx->nc, ctx->prec, L, b, orientation, ctx->cgc->gauss, work, dir, p0, p1, v, q0, q1,
                               total, old, ctx->stream);
    // every particle's Q1 is its Q0 (LightConeNGPParticleCreator.java:229-233)
    if (e == hipSuccess) e = hipMemcpyAsync(q1, q0, ad * total * tsz, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(ctx->stream);
        drop();
        return hip_check(ctx, e, "particles");
    }
    void* olds[] = {ctx->p_dir, ctx->p_pos0, ctx->p_pos1, ctx->p_vel, ctx->p_Q0, ctx->p_Q1};
    for (void* x : olds)
        if (x) (void)hipFree(x);
    ctx->p_dir = dir;
    ctx->p_pos0 = p0;
    ctx->p_pos1 = p1;
    ctx->p_vel = v;
    ctx->p_Q0 = q0;
    ctx->p_Q1 = q1;
    ctx->p_count = (uint32_t)total;
    if (added) *added = (int64_t)fresh;
    return OPX_OK;
}

int opx_cgc_create_superparticles(opx_ctx* ctx, int orientation, int num_threads, int* added, int* z_start,
                                  int* block_width)
{
    using namespace opx_host;
    if (!ctx) return OPX_E_ARG;
    int rc = cgc_check(ctx);
    if (rc != OPX_OK) return rc;
    if (orientation != 1 && orientation != -1) return fail(ctx, OPX_E_ARG, "super particles: orientation must be +1 or -1");
    if (num_threads < 1) return fail(ctx, OPX_E_ARG, "super particles: numberOfThreads must be >= 1");
    if (!ctx->cgc->gauss) return fail(ctx, OPX_E_STATE, "super particles: no light-cone solve yet");
    const int ppc = (int)std::floor(ctx->as[0] / ctx->at + 0.5);
    if (ppc < 1) return fail(ctx, OPX_E_ARG, "super particles: particlesPerCell = round(aL / at) must be >= 1");
    const opx::Lattice L = lattice(ctx, ctx->at);
    const int64_t nT = (int64_t)ctx->dims[1] * ctx->dims[2];
    Scratch s(ctx);
    int zs = 0, ze = 0;
    if ((rc = charge_block(ctx, s, L, &zs, &ze)) != OPX_OK) return rc;
    const int bw = ze - zs;
    if (z_start) *z_start = zs;
    if (block_width) *block_width = bw;
    if (bw < 1) return fail(ctx, OPX_E_STATE, "super particles: the charge block is empty");
    opx::SpBlock b;
    b.zs = zs;
    b.bw = bw;
    b.ppc = ppc;
    const int sub = std::min(num_threads, bw);
    b.ws = bw / sub;
    b.nl = bw - sub * b.ws;
    // descriptors: the existing particles, then subdivision after subdivision, ppc particles each
    const int old_count = ctx->sp_count, new_count = sub * ppc, count = old_count + new_count;
    const int64_t old_total = ctx->sp_total, new_total = (int64_t)bw * nT * ppc, total = old_total + new_total;
    std::vector<int> desc((size_t)6 * count);
    for (int f = 0; f < 6; f++)
        for (int p = 0; p < old_count; p++) desc[(size_t)f * count + p] = ctx->sp_host[(size_t)f * old_count + p];
    int64_t offset = (int64_t)zs * nT;
    int maxn = ctx->sp_max_num;
    for (int j = 0; j < sub; j++) {
        const int64_t num = (j < b.nl ? b.ws + 1 : b.ws) * nT;
        for (int k = 0; k < ppc; k++) {
            const int p = old_count + j * ppc + k;
            desc[0 * (size_t)count + p] = orientation;
            desc[1 * (size_t)count + p] = (int)num;
            desc[2 * (size_t)count + p] = (int)offset;
            desc[3 * (size_t)count + p] = (int)nT;
            desc[4 * (size_t)count + p] = k;
            desc[5 * (size_t)count + p] = ppc;
        }
        maxn = std::max<int>(maxn, (int)num);
        offset += num;
    }
    std::vector<int64_t> qoff((size_t)count);
    int64_t acc = 0;
    for (int p = 0; p < count; p++) {
        qoff[p] = acc;
        acc += desc[(size_t)count + p];
    }
    // new device arrays; the existing charges move over coefficient by coefficient
    int* ddesc = nullptr;
    int64_t* dqoff = nullptr;
    void* dq = nullptr;
    CGC_TRY(hipMalloc((void**)&ddesc, desc.size() * sizeof(int)), "alloc sp desc");
    auto drop = [&]() {
        if (ddesc) (void)hipFree(ddesc);
        if (dqoff) (void)hipFree(dqoff);
        if (dq) (void)hipFree(dq);
    };
    hipError_t e = hipMalloc((void**)&dqoff, qoff.size() * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&dq, (size_t)total * ctx->AD * ctx->tsz);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ddesc, desc.data(), desc.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dqoff, qoff.data(), qoff.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream);
    for (int k = 0; e == hipSuccess && old_total > 0 && k < ctx->AD; k++)
        e = hipMemcpyAsync((char*)dq + (size_t)k * total * ctx->tsz, (const char*)ctx->sp_q + (size_t)k * old_total * ctx->tsz,
                           (size_t)old_total * ctx->tsz, hipMemcpyDeviceToDevice, ctx->stream);
    double* work = nullptr;
    if (e == hipSuccess) e = s.alloc(&work, opx::cgc_sp_work(b, (uint32_t)nT, ctx->AD));
    if (e == hipSuccess)
        e = opx::cgc_sp_charges(ctx->nc, ctx->prec, L, b, ctx->cgc->gauss, work, dq, total, old_total, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(ctx->stream);
        drop();
        return hip_check(ctx, e, "super particles");
    }
    if (ctx->sp_desc) (void)hipFree(ctx->sp_desc);
    if (ctx->sp_qoff) (void)hipFree(ctx->sp_qoff);
    if (ctx->sp_q) (void)hipFree(ctx->sp_q);
    ctx->sp_desc = ddesc;
    ctx->sp_qoff = dqoff;
    ctx->sp_q = dq;
    ctx->sp_count = count;
    ctx->sp_total = total;
    ctx->sp_max_num = maxn;
    ctx->sp_host = desc;
    if (added) *added = new_count;
    return OPX_OK;
}

int opx_cgc_download(opx_ctx* ctx, int which, double* out, int64_t n_values)
{
    using namespace opx_host;
    if (!ctx) return OPX_E_ARG;
    if (which != OPX_CGC_CHARGE && which != OPX_CGC_GAUSS) return fail(ctx, OPX_E_ARG, "CGC download: unknown array");
    if (!out) return fail(ctx, OPX_E_ARG, "CGC download: null buffer");
    if (n_values != ctx->n * ctx->AS) return fail(ctx, OPX_E_ARG, "CGC download: buffer size mismatch");
    const double* src = !ctx->cgc ? nullptr : (which == OPX_CGC_CHARGE ? ctx->cgc->rho : ctx->cgc->gauss);
    if (!src) return fail(ctx, OPX_E_STATE, "CGC download: not computed");
    return download_alg(ctx, src, out);
}

int opx_cgc_wilson_tadpole(opx_ctx* ctx, double* value)
{
    using namespace opx_host;
    if (!ctx) return OPX_E_ARG;
    if (!value) return fail(ctx, OPX_E_ARG, "tadpole: null output");
    if (!ctx->cgc || !ctx->cgc->VT) return fail(ctx, OPX_E_STATE, "tadpole: no light-cone solve yet");
    const uint32_t T = (uint32_t)(ctx->dims[1] * ctx->dims[2]);
    Scratch s(ctx);
    double* d = nullptr;
    CGC_TRY(s.alloc(&d, 1), "alloc tadpole");
    CGC_TRY(opx::cgc_tadpole(ctx->nc, T, ctx->cgc->VT, d, ctx->stream), "tadpole");
    double sum = 0.0;
    CGC_TRY(hipMemcpyAsync(&sum, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "download tadpole");
    CGC_TRY(hipStreamSynchronize(ctx->stream), "tadpole sync");
    *value = sum / T;
    return OPX_OK;
}

int opx_cgc_wilson_dipole(opx_ctx* ctx, int capacity, int* num_bins, double* bin_width, double* corr, int64_t* counts)
{
    using namespace opx_host;
    if (!ctx) return OPX_E_ARG;
    if (!ctx->cgc || !ctx->cgc->VT) return fail(ctx, OPX_E_STATE, "dipole: no light-cone solve yet");
    const int n1 = ctx->dims[1], n2 = ctx->dims[2];
    const int T = n1 * n2;
    const int nbins = (int)std::sqrt((double)T) / 2;
    const double ds = std::min(n1 * ctx->as[1], n2 * ctx->as[2]) * 0.5 / (double)nbins;
    if (num_bins) *num_bins = nbins;
    if (bin_width) *bin_width = ds;
    if (!corr && !counts) return OPX_OK;   // size query
    if (capacity < nbins) return fail(ctx, OPX_E_ARG, "dipole: output holds fewer than num_bins entries");
    if (nbins == 0) return OPX_OK;
    Scratch s(ctx);
    double *psum = nullptr, *dcorr = nullptr;
    unsigned long long *pcnt = nullptr, *dcnt = nullptr;
    CGC_TRY(s.alloc(&psum, (size_t)nbins * T), "alloc dipole partials");
    CGC_TRY(s.alloc(&pcnt, (size_t)nbins * T), "alloc dipole counts");
    CGC_TRY(s.alloc(&dcorr, (size_t)nbins), "alloc dipole");
    CGC_TRY(s.alloc(&dcnt, (size_t)nbins), "alloc dipole");
    CGC_TRY(opx::cgc_dipole(ctx->nc, (uint32_t)n1, (uint32_t)n2, ctx->as[1], ctx->as[2], ds, nbins, ctx->cgc->VT, psum,
                            pcnt, dcorr, dcnt, ctx->stream),
            "dipole");
    std::vector<double> hc((size_t)nbins);
    std::vector<unsigned long long> hn((size_t)nbins);
    CGC_TRY(hipMemcpyAsync(hc.data(), dcorr, hc.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
            "download dipole");
    CGC_TRY(hipMemcpyAsync(hn.data(), dcnt, hn.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream),
            "download dipole");
    CGC_TRY(hipStreamSynchronize(ctx->stream), "dipole sync");
    for (int b = 0; b < nbins; b++) {
        if (corr) corr[b] = hc[b];
        if (counts) counts[b] = (int64_t)hn[b];
    }
    return OPX_OK;
}

int opx_cgc_clear(opx_ctx* ctx)
{
    if (!ctx) return OPX_E_ARG;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    opx_host::cgc_release(ctx);
    return OPX_OK;
}

}  // extern "C"
