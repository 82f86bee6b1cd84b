// opx_particles.hip -- per-particle CGC NGP coupling on gfx950: particles moving
// on grid lines at light speed whose colour charges are parallel transported by
// the link they cross (physics/grid/CGCParticleInterpolationNGP.java:15-104,
// physics/movement/solver/CGCParticleSolver.java:15-36).
//
// One thread per particle. Positions stay in double in every precision so that
// the nearest grid point, (int) rint(pos / a) rounded half to even
// (physics/util/GridFunctions.java:67-73), is bit-exact; charges are in the
// context precision, [coefficient][particle]. Deposits onto rho and J are
// atomic additions (several particles share a cell).
#include "opx_group.h"
#include "opx_kernels.h"

namespace opx {

namespace {

constexpr int kPBlock = 256;

struct Ngp {
    int c[3];
};

__device__ __forceinline__ Ngp nearest_grid_point(const double* __restrict__ pos, uint32_t count, uint32_t k,
                                                  const Lattice& L)
{
    Ngp g;
#pragma unroll
    for (int i = 0; i < 3; i++) g.c[i] = (int)rint(pos[(size_t)i * count + k] / L.as[i]);
    return g;
}

// Grid.getCellIndex with periodic coordinates
__device__ __forceinline__ uint32_t cell_of(const Lattice& L, const Ngp& g)
{
    const int n[3] = {(int)L.d0, (int)L.d1, (int)L.d2};
    uint32_t idx = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        int p = g.c[i] % n[i];
        if (p < 0) p += n[i];
        idx = i == 0 ? (uint32_t)p : idx * (uint32_t)n[i] + (uint32_t)p;
    }
    return idx;
}

template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, T>::A load_charge(const T* __restrict__ Q, uint32_t count, uint32_t k)
{
    using GR = Group<NC, T>;
    typename GR::A a;
#pragma unroll
    for (int m = 0; m < GR::AD; m++) aref(a, m) = Q[(size_t)m * count + k];
    return a;
}

template <int NC, typename T>
__device__ __forceinline__ void store_charge(T* __restrict__ Q, uint32_t count, uint32_t k,
                                             const typename Group<NC, T>::A& a)
{
    using GR = Group<NC, T>;
#pragma unroll
    for (int m = 0; m < GR::AD; m++) Q[(size_t)m * count + k] = aget(a, m);
}

// interpolateChargedensity: rho(ngp(pos0)) += Q0
template <int NC, typename T>
__global__ __launch_bounds__(kPBlock) void k_p_chargedensity(Lattice L, uint32_t count,
                                                             const double* __restrict__ pos0,
                                                             const T* __restrict__ Q0, T* __restrict__ rho)
{
    using GR = Group<NC, T>;
    const uint32_t k = blockIdx.x * kPBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t cell = cell_of(L, nearest_grid_point(pos0, count, k, L));
#pragma unroll
    for (int m = 0; m < GR::AD; m++) atomicAdd(rho + (size_t)m * L.n + cell, Q0[(size_t)m * count + k]);
}

// updatePosition, interpolateToParticle and updateCharge of one particle:
// pos1 = pos0 + v dt; when the nearest grid point along the beam changes, the
// crossed link U = Unext(old cell) (v > 0) or Unext(new cell)^dagger (v < 0) transports
// Q1 = U^dagger Q0 U; otherwise Q0 and Q1 trade places.
template <int NC, typename T>
__global__ __launch_bounds__(kPBlock) void k_p_push(Lattice L, uint32_t count, const int* __restrict__ dir,
                                                    const double* __restrict__ pos0, double* __restrict__ pos1,
                                                    const double* __restrict__ vel, double dt,
                                                    const T* __restrict__ Unext, T* __restrict__ Q0,
                                                    T* __restrict__ Q1)
{
#pragma clang fp contract(off)
    using GR = Group<NC, T>;
    const uint32_t k = blockIdx.x * kPBlock + threadIdx.x;
    if (k >= count) return;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double step = vel[(size_t)i * count + k] * dt;
        pos1[(size_t)i * count + k] = pos0[(size_t)i * count + k] + step;
    }
    const int d = dir[k];
    const Ngp no = nearest_grid_point(pos0, count, k, L), nn = nearest_grid_point(pos1, count, k, L);
    const typename GR::A q0 = load_charge<NC, T>(Q0, count, k);
    if (no.c[d] != nn.c[d]) {
        const bool forward = vel[(size_t)d * count + k] > 0;
        const uint32_t cell = cell_of(L, forward ? no : nn);
        typename GR::M U;
        const T* p = Unext + (size_t)d * GR::S * L.n + cell;
#pragma unroll
        for (int c = 0; c < GR::S; c++) GR::gref(U, GR::slot(c)) = p[(size_t)c * L.n];
        GR::complete(U);
        if (!forward) U = adj(U);
        store_charge<NC, T>(Q1, count, k, act(q0, adj(U)));
    } else {
        const typename GR::A q1 = load_charge<NC, T>(Q1, count, k);
        store_charge<NC, T>(Q0, count, k, q1);
        store_charge<NC, T>(Q1, count, k, q0);
    }
}

// interpolateToGrid: on a crossing, J_d(old cell) += Q0 a_d / a_t (v > 0) or
// J_d(new cell) += -Q1 a_d / a_t (v < 0)
template <int NC, typename T>
__global__ __launch_bounds__(kPBlock) void k_p_to_grid(Lattice L, uint32_t count, const int* __restrict__ dir,
                                                       const double* __restrict__ pos0,
                                                       const double* __restrict__ pos1,
                                                       const double* __restrict__ vel, const T* __restrict__ Q0,
                                                       const T* __restrict__ Q1, T* __restrict__ J)
{
    using GR = Group<NC, T>;
    const uint32_t k = blockIdx.x * kPBlock + threadIdx.x;
    if (k >= count) return;
    const int d = dir[k];
    const Ngp no = nearest_grid_point(pos0, count, k, L), nn = nearest_grid_point(pos1, count, k, L);
    if (no.c[d] == nn.c[d]) return;
    const bool forward = vel[(size_t)d * count + k] > 0;
    const uint32_t cell = cell_of(L, forward ? no : nn);
    const T f = forward ? (T)(L.as[d] / L.at) : (T)(-L.as[d] / L.at);
    const T* Q = forward ? Q0 : Q1;
    T* Jc = J + (size_t)d * GR::AD * L.n + cell;
#pragma unroll
    for (int m = 0; m < GR::AD; m++) atomicAdd(Jc + (size_t)m * L.n, Q[(size_t)m * count + k] * f);
}

#define OPX_PDISPATCH(nc, prec, ...)                                                           \
    do {                                                                                       \
        if ((nc) == 2 && (prec) == 0) { constexpr int NC = 2; using T = double; __VA_ARGS__; } \
        else if ((nc) == 2) { constexpr int NC = 2; using T = float; __VA_ARGS__; }           \
        else if ((prec) == 0) { constexpr int NC = 3; using T = double; __VA_ARGS__; }        \
        else { constexpr int NC = 3; using T = float; __VA_ARGS__; }                          \
    } while (0)

inline dim3 pgrid(uint32_t count) { return dim3((count + kPBlock - 1) / kPBlock); }

}  // namespace

hipError_t launch_p_chargedensity(int nc, int prec, const Lattice& L, const ParticleArrays& p, void* rho,
                                  hipStream_t st)
{
    if (p.count == 0) return hipSuccess;
    OPX_PDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_p_chargedensity<NC, T>), pgrid(p.count), dim3(kPBlock), 0, st, L, p.count, p.pos0,
                           (const T*)p.Q0, (T*)rho);
    });
    return hipGetLastError();
}

hipError_t launch_p_push(int nc, int prec, const Lattice& L, const ParticleArrays& p, double dt, const void* Unext,
                         hipStream_t st)
{
    if (p.count == 0) return hipSuccess;
    OPX_PDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_p_push<NC, T>), pgrid(p.count), dim3(kPBlock), 0, st, L, p.count, p.dir, p.pos0, p.pos1,
                           p.vel, dt, (const T*)Unext, (T*)p.Q0, (T*)p.Q1);
    });
    return hipGetLastError();
}

hipError_t launch_p_to_grid(int nc, int prec, const Lattice& L, const ParticleArrays& p, void* J, hipStream_t st)
{
    if (p.count == 0) return hipSuccess;
    OPX_PDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_p_to_grid<NC, T>), pgrid(p.count), dim3(kPBlock), 0, st, L, p.count, p.dir, p.pos0,
                           p.pos1, p.vel, (const T*)p.Q0, (const T*)p.Q1, (T*)J);
    });
    return hipGetLastError();
}

}  // namespace opx
