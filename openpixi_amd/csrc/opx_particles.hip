// opx_particles.hip -- per-particle CGC NGP coupling on gfx950: particles moving
// on grid lines at light speed whose colour charges are parallel transported by
// the link they cross (physics/grid/CGCParticleInterpolationNGP.java:15-104,
// physics/movement/solver/CGCParticleSolver.java:15-36).
//
// One thread per particle. Positions stay in double in every precision so that
// the nearest grid point, (int) rint(pos / a) rounded half to even
// (physics/util/GridFunctions.java:67-73), is bit-exact; charges are in the
// context precision, [coefficient][particle]. Deposits onto rho and J are ordered:
// the (target, particle) pairs are sorted by target with a stable radix sort, so the
// particles of one cell stay in list order, and the first entry of each run adds the
// run's charges to the cell in that order, as the reference's sequential loop does
// (bit for bit in FP64; no atomics, so identical runs give identical fields).
#include "opx_group.h"
#include "opx_kernels.h"

#include <hipcub/device/device_radix_sort.hpp>

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

// interpolateChargedensity, key pass: particle k deposits Q0 onto rho(ngp(pos0))
__global__ __launch_bounds__(kPBlock) void k_p_rho_keys(Lattice L, uint32_t count, const double* __restrict__ pos0,
                                                        uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t k = blockIdx.x * kPBlock + threadIdx.x;
    if (k >= count) return;
    keys[k] = cell_of(L, nearest_grid_point(pos0, count, k, L));
    vals[k] = k;
}

// rho(cell) += Q0 of the run of sorted entries that starts here, in particle order
template <int NC, typename T>
__global__ __launch_bounds__(kPBlock) void k_p_rho_gather(Lattice L, uint32_t count, const uint32_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ vals, const T* __restrict__ Q0,
                                                          T* __restrict__ rho)
{
#pragma clang fp contract(off)
    using GR = Group<NC, T>;
    const uint32_t i = blockIdx.x * kPBlock + threadIdx.x;
    if (i >= count) return;
    const uint32_t cell = keys[i];
    if (i > 0 && keys[i - 1] == cell) return;
#pragma unroll
    for (int m = 0; m < GR::AD; m++) {
        const uint64_t r = fposT<T>(m, cell, GR::AD);
        T v = rho[r];
        for (uint32_t j = i; j < count && keys[j] == cell; j++) v += Q0[(size_t)m * count + vals[j]];
        rho[r] = v;
    }
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
        const T* p = fcell(Unext, cell, 3 * GR::S);
#pragma unroll
        for (int c = 0; c < GR::S; c++) GR::gref(U, GR::slot(c)) = p[fcomp<T>(d * GR::S + c, 3 * GR::S)];
        GR::complete(U);
        if (!forward) U = adj(U);
        store_charge<NC, T>(Q1, count, k, act(q0, adj(U)));
    } else {
        const typename GR::A q1 = load_charge<NC, T>(Q1, count, k);
        store_charge<NC, T>(Q0, count, k, q1);
        store_charge<NC, T>(Q1, count, k, q0);
    }
}

// interpolateToGrid, key pass: on a crossing, particle k deposits into J_d(old cell)
// (v > 0) or J_d(new cell) (v < 0), key d n + cell; the others get the key 3 n (after
// every deposit in the sorted order)
__global__ __launch_bounds__(kPBlock) void k_p_j_keys(Lattice L, uint32_t count, const int* __restrict__ dir,
                                                      const double* __restrict__ pos0, const double* __restrict__ pos1,
                                                      const double* __restrict__ vel, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals)
{
    const uint32_t k = blockIdx.x * kPBlock + threadIdx.x;
    if (k >= count) return;
    const int d = dir[k];
    const Ngp no = nearest_grid_point(pos0, count, k, L), nn = nearest_grid_point(pos1, count, k, L);
    uint32_t key = 3 * L.n;
    if (no.c[d] != nn.c[d]) {
        const bool forward = vel[(size_t)d * count + k] > 0;
        key = (uint32_t)d * L.n + cell_of(L, forward ? no : nn);
    }
    keys[k] = key;
    vals[k] = k;
}

// J_d(cell) += Q0 a_d / a_t (v > 0) or -Q1 a_d / a_t (v < 0) for the run of sorted entries that
// starts here, in particle order
template <int NC, typename T>
__global__ __launch_bounds__(kPBlock) void k_p_j_gather(Lattice L, uint32_t count, const uint32_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ vals,
                                                        const double* __restrict__ vel, const T* __restrict__ Q0,
                                                        const T* __restrict__ Q1, T* __restrict__ J)
{
#pragma clang fp contract(off)
    using GR = Group<NC, T>;
    const uint32_t i = blockIdx.x * kPBlock + threadIdx.x;
    if (i >= count) return;
    const uint32_t key = keys[i];
    if (key >= 3 * L.n || (i > 0 && keys[i - 1] == key)) return;
    const int d = (int)(key / L.n);
    const uint32_t cell = key - (uint32_t)d * L.n;
    const T fp = (T)(L.as[d] / L.at), fm = (T)(-L.as[d] / L.at);
#pragma unroll
    for (int m = 0; m < GR::AD; m++) {
        const uint64_t r = fposT<T>(d * GR::AD + m, cell, 3 * GR::AD);
        T v = J[r];
        for (uint32_t j = i; j < count && keys[j] == key; j++) {
            const uint32_t k = vals[j];
            const bool forward = vel[(size_t)d * count + k] > 0;
            const T q = forward ? Q0[(size_t)m * count + k] : Q1[(size_t)m * count + k];
            const T c = q * (forward ? fp : fm);
            v += c;
        }
        J[r] = v;
    }
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

size_t particle_sort_temp_bytes(uint32_t count)
{
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)count, 0, 32) !=
        hipSuccess)
        return 0;
    return bytes;
}

static hipError_t sort_pairs(const ParticleSort& w, uint32_t count, hipStream_t st)
{
    size_t bytes = w.temp_bytes;
    return hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, w.keys, w.keys_sorted, w.vals, w.vals_sorted, (int)count,
                                              0, 32, st);
}

hipError_t launch_p_chargedensity(int nc, int prec, const Lattice& L, const ParticleArrays& p, const ParticleSort& w,
                                  void* rho, hipStream_t st)
{
    if (p.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_p_rho_keys, pgrid(p.count), dim3(kPBlock), 0, st, L, p.count, p.pos0, w.keys, w.vals);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = sort_pairs(w, p.count, st);
    if (e != hipSuccess) return e;
    OPX_PDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_p_rho_gather<NC, T>), pgrid(p.count), dim3(kPBlock), 0, st, L, p.count, w.keys_sorted,
                           w.vals_sorted, (const T*)p.Q0, (T*)rho);
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

hipError_t launch_p_to_grid(int nc, int prec, const Lattice& L, const ParticleArrays& p, const ParticleSort& w, void* J,
                            hipStream_t st)
{
    if (p.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_p_j_keys, pgrid(p.count), dim3(kPBlock), 0, st, L, p.count, p.dir, p.pos0, p.pos1, p.vel,
                       w.keys, w.vals);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = sort_pairs(w, p.count, st);
    if (e != hipSuccess) return e;
    OPX_PDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_p_j_gather<NC, T>), pgrid(p.count), dim3(kPBlock), 0, st, L, p.count, w.keys_sorted,
                           w.vals_sorted, p.vel, (const T*)p.Q0, (const T*)p.Q1, (T*)J);
    });
    return hipGetLastError();
}

}  // namespace opx
