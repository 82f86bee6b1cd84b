// opx_kernels.h -- launch interfaces of the gfx950 kernels (shared by the C-ABI layer).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace opx {

// Device layout of a cell field (links, E, J, rho and the scratch fields built like them): the
// cells are cut into blocks of kCellBlock = 64 consecutive cells (one wavefront), and a block
// holds every component of its cells. Within a block the components come in groups of V: group g
// holds components gV .. gV + V - 1 of the block's 64 cells, cell by cell (V values of a cell
// adjacent), so component c of cell s of a field with `comps` components per cell sits at element
//   ((s / 64) * comps + (c / V) * V) * 64 + (s % 64) * V + c % V.
// V = 2 for the double fields of SU(3) (links 54, algebra 24, charge density 8 components per
// cell): a lane reads a pair of its cell's components in one 16-byte access, a wavefront one
// contiguous 1 KiB run; per-direction counts (18 link reals, 8 coefficients) are even, so each
// direction starts a pair. Every other field keeps V = 1 (one component of 64 cells contiguous):
// four floats per 16-byte access made the SU(3) float cell form no faster and the SU(2) float
// cell form 7 % slower, and pairs made the SU(2) double cell form 6.5 % slower; the SU(2) algebra
// fields (3 per direction) could not be paired anyway. The components of a cell lie within
// comps * 64 elements of each other (one array per component, 2^k cells apart, sends every
// component of a cell to the same memory channels). A field of n cells holds
// padded_cells(n) * comps elements.
constexpr uint32_t kCellBlock = 64;
__host__ __device__ constexpr uint32_t fvec(uint32_t comps, uint32_t tsz)
{
    return (tsz == 8 && comps % 2 == 0 && comps != 12) ? 2u : 1u;
}
// the group width of the link and algebra fields of SU(NC) in T (as fvec: SU(3) double fields only)
template <int NC, typename T>
constexpr uint32_t kGroup = (NC == 3 && sizeof(T) == 8) ? 2u : 1u;
// element index of component c of cell s (tsz: bytes per real of the field)
__host__ __device__ __forceinline__ uint64_t fpos(uint32_t c, uint64_t s, uint32_t comps, uint32_t tsz)
{
    const uint32_t v = fvec(comps, tsz);
    return ((s >> 6) * comps + (c / v) * v) * 64 + (s & 63) * v + c % v;
}
template <typename T>
__host__ __device__ __forceinline__ uint64_t fposT(uint32_t c, uint64_t s, uint32_t comps)
{
    return fpos(c, s, comps, (uint32_t)sizeof(T));
}
__host__ __device__ __forceinline__ uint64_t padded_cells(uint64_t n) { return (n + 63) & ~(uint64_t)63; }
// byte offsets of element (c, s) for buffer instructions: the per-cell part and the per-component
// part (uniform over a wavefront, kept in a scalar register)
__host__ __device__ __forceinline__ uint32_t fvoff(uint32_t s, uint32_t comps, uint32_t sz)
{
    return ((s >> 6) * comps * 64 + (s & 63) * fvec(comps, sz)) * sz;
}
__host__ __device__ __forceinline__ uint32_t fsoff(uint32_t c, uint32_t comps, uint32_t sz)
{
    const uint32_t v = fvec(comps, sz);
    return ((c / v) * v * 64 + c % v) * sz;
}
// the first element of cell s (component 0); component c is fcell(f, s, comps)[fcomp<T>(c, comps)]
template <typename T>
__host__ __device__ __forceinline__ T* fcell(T* f, uint32_t s, uint32_t comps)
{
    return f + (uint64_t)(s >> 6) * (comps * 64u) + (s & 63u) * fvec(comps, (uint32_t)sizeof(T));
}
template <typename T>
__host__ __device__ constexpr uint32_t fcomp(uint32_t c, uint32_t comps)
{
    return (c / fvec(comps, (uint32_t)sizeof(T))) * fvec(comps, (uint32_t)sizeof(T)) * 64 +
           c % fvec(comps, (uint32_t)sizeof(T));
}

// Inclusive box of cells; all = 1 means every cell.
struct Box {
    int lo[3];
    int hi[3];
    int all;
};

// Lattice geometry and time-step constants for the field kernels.
struct Lattice {
    uint32_t n;        // cells
    uint32_t d0, d1, d2;
    uint32_t s0, s1;   // strides of x0 and x1 (x2 stride is 1)
    double at;         // time step argument of step/stepLinks
    double uf[3];      // -g a_i a_t (FastTYMSolver.step unitFactor)
    double area[3];    // 1/a_i^2 (Grid.getStapleSum areaFactor)
    // the cell form's staple-sum factors of direction d, whose planes are (d, j1), (d, j2), j1 < j2:
    // cell_ratio[d] = area[j2] / area[j1], cell_scale[d] = at area[j1]
    double cell_ratio[3];
    double cell_scale[3];
    double as[3];      // lattice spacings
    double g;          // coupling
    uint32_t force_global = 0;   // 1: the fused kernels read through plain pointers below 4 GiB too
    // x0 planes [j_lo[i], j_hi[i]), i < j_slabs, outside which the current J is zero: the cell and
    // per-link forms read J only there (every other reader reads all of it)
    uint32_t j_slabs = 1;
    uint32_t j_lo[4] = {0, 0, 0, 0};
    uint32_t j_hi[4] = {0xffffffffu, 0, 0, 0};
};
constexpr int kMaxCurrentSlabs = 4;

// One event of a batch: its lattice, the element offsets of its link block
// (3*G*n values) and E block (3*AD*n values) in the batch buffers, and the first
// of its 3*ceil(n/256) work items.
struct BatchEvent {
    Lattice L;
    uint64_t link_offset;
    uint64_t alg_offset;
    uint32_t item_start;
    uint32_t pad;
};

// Cell segments of the super-particle charge-density gather. The cells any particle
// sits on at time t are cut into segments of consecutive cells that one fixed set of
// particles covers: segment s holds cells [first[s], first[s] + cells_prefix[s+1] -
// cells_prefix[s]) and the particles list[list_prefix[s] .. list_prefix[s+1]) in
// particle-list order, each as qbase = (first charge index) - (first cell), so that the
// charge of that particle on cell c sits at q[qbase + c]. Built on the host per deposit.
struct CellSegments {
    uint32_t count;                  // segments
    uint32_t cells;                  // cells_prefix[count]
    const uint32_t* first;           // [count]
    const uint32_t* cells_prefix;    // [count + 1]
    const uint32_t* list_prefix;     // [count + 1]
    const int64_t* qbase;            // [list_prefix[count]]
};

// Super-particle descriptors (device arrays of `count` entries).
struct SuperParticles {
    int count;
    int max_num;
    const int* orientation;
    const int* num;
    const int* index_offset;
    const int* per_plane;
    const int* shift;
    const int* ppc;
    const int64_t* q_offset;
    int64_t total;     // charges in total
};

// Per-particle CGC NGP particles (opx_particles.hip): positions and velocities
// [3][count] in double, charges [coefficient][count] in the context precision.
struct ParticleArrays {
    uint32_t count;
    const int* dir;
    const double* pos0;
    double* pos1;
    const double* vel;
    void* Q0;
    void* Q1;
};

// Streaming (2.5D, LDS-exchanged plaquette) form of the fused update.
// Tile shapes (TY x TX threads over (x1, x2)): 0: 16x16, 1: 8x32, 2: 16x32, 3: 8x64, 4: 8x16.
constexpr int kStreamTiles = 5;
bool stream_supported(int nc, int prec, const Lattice& L, int tile);
void stream_tile_shape(int tile, int* ty, int* tx);
uint32_t stream_blocks_per_chunk(const Lattice& L, int tile);
hipError_t launch_stream_update(int nc, int prec, const Lattice& L, const Box& active, int tile, uint32_t nchunks,
                                const void* U, void* Unext, void* E, const void* J, bool has_j, hipStream_t st);

// Gauss-law projection building blocks (double precision work arrays).
hipError_t gauss_apply_D(int nc, int prec, const Lattice& L, const void* U, const double* phi, double* F, hipStream_t st);
hipError_t gauss_apply_Dt(int nc, int prec, const Lattice& L, const void* U, const double* F, const void* rho,
                          double* out, hipStream_t st);
// field (block layout, m / n components per cell) <-> [component][cell] double work array
hipError_t gauss_to_double(int prec, uint64_t m, uint32_t n, const void* a, double* b, hipStream_t st);
hipError_t gauss_sub_into(int prec, uint64_t m, uint32_t n, const double* Ed, const double* F, void* E, hipStream_t st);
size_t gauss_dot_scratch();
hipError_t gauss_dot(uint64_t m, const double* a, const double* b, double* scratch, double* out, hipStream_t st);
hipError_t gauss_axpy(uint64_t m, double alpha, const double* x, double* y, hipStream_t st);
hipError_t gauss_xpby(uint64_t m, const double* x, double beta, double* y, hipStream_t st);

// GaussConstraintRestoration building blocks: G = proj of the violation times factor, R = G - rho
// ([coefficient][cell] doubles; rho may be null), and E_j(x) += f2 [act(G(x+j), U_j(x)) - G(x)]
hipError_t restore_violation(int nc, int prec, const Lattice& L, const void* U, const void* Unext, const void* rho,
                             double factor, double* G, double* R, hipStream_t st);
hipError_t restore_correct(int nc, int prec, const Lattice& L, const void* U, const double* G, double f2, void* E,
                           hipStream_t st);
// UnitarityTester: flags[cell * 3 + dir] = 1 where |Re tr U| or |Re tr Unext| >= N + tol; *count
// (device) is incremented once per flagged link
hipError_t unitarity_flags(int nc, int prec, uint32_t n, const void* U, const void* Unext, double tol,
                           unsigned char* flags, unsigned long long* count, hipStream_t st);

// Batched events: persistent grid of `grid` workgroups over total_items work items.
// Work arrays of the ordered per-particle deposits: (target, particle) keys and values, their
// sorted copies (count entries each) and the radix sort's temporary storage
struct ParticleSort {
    uint32_t* keys;
    uint32_t* vals;
    uint32_t* keys_sorted;
    uint32_t* vals_sorted;
    void* temp;
    size_t temp_bytes;
};
size_t particle_sort_temp_bytes(uint32_t count);
hipError_t launch_p_chargedensity(int nc, int prec, const Lattice& L, const ParticleArrays& p, const ParticleSort& w,
                                  void* rho, hipStream_t st);
hipError_t launch_p_push(int nc, int prec, const Lattice& L, const ParticleArrays& p, double dt, const void* Unext,
                         hipStream_t st);
hipError_t launch_p_to_grid(int nc, int prec, const Lattice& L, const ParticleArrays& p, const ParticleSort& w, void* J,
                            hipStream_t st);
hipError_t launch_batch_update(int nc, int prec, const BatchEvent* ev, int nev, uint32_t total_items, uint32_t grid,
                               const void* U, void* Unext, void* E, hipStream_t st);
hipError_t launch_batch_reunitarize(int nc, int prec, const BatchEvent* ev, int nev, uint32_t total_items,
                                    uint32_t grid, void* U, hipStream_t st);
// the cell form over a batch (total_items: the per-link item count; a third of it are cell items);
// buf: every event's link field is below 4 GiB, so the fields are read by buffer instructions;
// reunit: the new links are reprojected onto the group in double before they are stored
hipError_t launch_batch_cell(int nc, int prec, const BatchEvent* ev, int nev, uint32_t total_items, uint32_t grid,
                             bool buf, const void* U, void* Unext, void* E, bool reunit, hipStream_t st);

// All launchers enqueue on `st` and return hipGetLastError().
hipError_t launch_combined_update(int nc, int prec, const Lattice& L, const Box& active, const void* U,
                                  void* Unext, void* E, const void* J, bool has_j, hipStream_t st);
// cell form: the three link updates of a cell in one thread, the upper staples of a plane
// sharing one product; blocks sweep row tiles of kCellRows[tile] x1 rows along x0.
// reunit: the new links are reprojected onto the group in double (as launch_reunitarize) before
// they are stored
constexpr int kCellTiles = 3;
constexpr uint32_t kCellRows[kCellTiles] = {32, 16, 8};
hipError_t launch_cell_update(int nc, int prec, const Lattice& L, const Box& active, int tile, const void* U,
                              void* Unext, void* E, const void* J, bool has_j, bool reunit, hipStream_t st);
// SU(2) LDS-tile form of the fused update (opx_tile.hip): blocks of TY x TX cells
// march along x0 over nchunks chunks of planes.
constexpr int kTileTY32 = 16, kTileTY64 = 8, kTileTX = 16;
struct TileGeom {
    uint32_t tiles1, tiles2, nchunks;
};
// One work item of the batched tile form: planes [pb, pe) of the tile at (row0, col0) of an event.
struct TileItem {
    uint32_t event, row0, col0, pb, pe;
};
bool tile_supported(int nc);
TileGeom tile_geom(int prec, const Lattice& L, uint32_t nchunks);
// nt: non-temporal stores of E and Unext
hipError_t launch_tile_update(int nc, int prec, const Lattice& L, const Box& active, uint32_t nchunks, bool nt,
                              const void* U, void* Unext, void* E, const void* J, bool has_j, hipStream_t st);
hipError_t launch_batch_tile(int prec, bool nt, const BatchEvent* ev, const TileItem* items, uint32_t nitems,
                             uint32_t grid, const void* U, void* Unext, void* E, hipStream_t st);

// Plane-tile form of the fused update (opx_plane.hip): a block updates the three
// links of a TY x TX tile of one x0 plane, with the plane's links (tile plus halo) in LDS.
constexpr int kPlaneShapes = 3;
struct PlaneGeom {
    uint32_t tiles1, tiles2;
};
bool plane_supported(int nc, int prec, const Lattice& L, int shape);
hipError_t launch_plane_update(int nc, int prec, const Lattice& L, const Box& active, int shape, const void* U,
                               void* Unext, void* E, const void* J, bool has_j, hipStream_t st);

// Plaquette form of the fused update: proj P of the three planes of every cell into P
// (3 * AD * n values), then the per-link update from them.
hipError_t launch_plaq_update(int nc, int prec, const Lattice& L, const Box& active, const void* U, void* P,
                              void* Unext, void* E, const void* J, bool has_j, hipStream_t st);
hipError_t launch_step_links(int nc, int prec, const Lattice& L, const Box& active, const void* U, void* Unext,
                             const void* E, hipStream_t st);
hipError_t launch_reunitarize(int nc, int prec, uint32_t n, void* U, hipStream_t st);

// SU(2) adjoint (opx_adjoint.hip); prec 0 = fp64, 1 = fp32; Ubar 4 and Ebar 3 values per link
hipError_t launch_adj_seed(int prec, uint32_t n, const void* E, void* Ebar, void* Ubar, const double coefE[3],
                           hipStream_t st);
hipError_t launch_adj_link_step(int prec, uint32_t n, const void* U, const void* E, void* Ubar, void* Ebar, double at,
                                hipStream_t st);
hipError_t launch_adj_force(int prec, const Lattice& L, const void* U, const void* Ebar, void* Ubar, double scale,
                            hipStream_t st);
hipError_t launch_adj_bsq(int prec, const Lattice& L, const void* U, void* Ubar, const double coef[3], hipStream_t st);
hipError_t launch_links_from_coeffs(int nc, int prec, uint32_t n, const double* coeffs_soa, void* U,
                                    hipStream_t st);
// partials: device buffer of at least measure_partials_size() doubles; out: 8 device doubles
size_t measure_partials_size(uint32_t n);
hipError_t launch_measure(int nc, int prec, const Lattice& L, const Box& eval, const void* U, const void* Unext,
                          const void* E, const void* rho, double* partials, double* out, hipStream_t st);
// ImplicitTYMSolver (opx_implicit.hip): one iteration (ImplicitStep + ImplicitDamping) from the
// links U (time 0), Uold (time -1) and the implicit estimate iU (time +1) into iE, iUn; and
// ImplicitEnd (E, Unext of the active links from iE, iUn).
hipError_t launch_implicit_iteration(int nc, int prec, const Lattice& L, const Box& active, const void* U,
                                     const void* Uold, const void* iU, const void* E, const void* J, bool has_j,
                                     void* iE, void* iUn, double damping, hipStream_t st);
hipError_t launch_implicit_end(int nc, int prec, const Lattice& L, const Box& active, const void* iE, const void* iUn,
                               void* E, void* Unext, hipStream_t st);
// GroupElement.pow(x) of `count` elements (reference storage, G reals each, precision prec)
hipError_t launch_group_pow(int nc, int prec, int count, const void* U, double x, void* out, hipStream_t st);

// SU(2) longitudinal Poynting vector per slice: out2[0, nd) averaged, out2[nd, 2 nd) time averaged
hipError_t launch_projected_poynting(int nc, int prec, const Lattice& L, int dir, const void* U, const void* Unext,
                                     const void* E, double* out2, hipStream_t st);
hipError_t launch_projected_energy(int nc, int prec, const Lattice& L, const Box& eval, int dir, const void* U,
                                   const void* Unext, const void* E, double* out4, hipStream_t st);
// Static per-slice buffer of the allocation-free projected energy: 4 x kNoAllocMaxSlices doubles
// (128 KiB) on every GPU the library runs on. noalloc_scratch returns its address on `device`,
// which must be the current device the first time it is asked for (looked up once per device).
constexpr int kNoAllocMaxSlices = 4096;
constexpr int kNoAllocMaxDevices = 64;
hipError_t noalloc_scratch(int device, double** out);
// zero every component (comps per cell) of the cells of up to kMaxCurrentSlabs ranges
// [first[r], first[r] + count[r]) of a field, in one launch
struct CellRanges {
    uint32_t first[kMaxCurrentSlabs];
    uint32_t start[kMaxCurrentSlabs + 1];   // prefix sums of the counts
    int count;
};
hipError_t launch_zero_cell_ranges(int prec, void* f, uint32_t comps, const CellRanges& r, hipStream_t st);
hipError_t launch_health(int nc, int prec, uint32_t n, const void* U, double tol,
                         unsigned long long* counters, hipStream_t st);
hipError_t launch_health_alg(int nc, int prec, uint32_t n_values, const void* X,
                             unsigned long long* counters, hipStream_t st);
hipError_t launch_sp_to_grid(int nc, int prec, const Lattice& L, const SuperParticles& sp, int t, void* q,
                             const void* Unext, void* J, hipStream_t st);
// gathers rho over the cells of `cs` only (the cells the particles occupy at t); q is
// [coef][q_total]
hipError_t launch_sp_chargedensity(int nc, int prec, const Lattice& L, int64_t q_total, const void* q, void* rho,
                                   const CellSegments& cs, hipStream_t st);

// CGC initial condition with the beam along x0 (opx_cgc.hip). Work arrays are double,
// [component][cell]; complex ones double2; group elements [real][cell].
constexpr uint32_t kCgcMaxFft = 4096;   // longest transform axis
// A block of bw planes from zs carrying super-particle charges, ppc charges per cell, split into
// subdivisions of ws + 1 planes (the first nl of them) and ws planes.
struct SpBlock {
    int zs, bw, ppc, ws, nl;
};
bool cgc_fft_supported(uint32_t n);
hipError_t cgc_twiddles(uint32_t n, double2* tw, hipStream_t st);
// in-place transform along one axis of `total` complex values: lines of n elements `stride` apart
hipError_t cgc_fft_axis(double2* x, uint64_t total, uint32_t n, uint32_t stride, const double2* tw, int sign,
                        hipStream_t st);
hipError_t cgc_mv_fill(const Lattice& L, uint64_t m, const double* noise, double rcw, double location, double norm,
                       double i2s2, double2* x, hipStream_t st);
hipError_t cgc_mv_regulate(const Lattice& L, uint64_t m, double uvt2, double long_width, double ir2, double inv_n,
                           double2* x, hipStream_t st);
hipError_t cgc_mv_store(const Lattice& L, const Box& active, uint64_t m, int lmin, int lmax, const double2* x,
                        double* rho, hipStream_t st);
hipError_t cgc_to_complex(uint64_t m, const double* a, double2* x, hipStream_t st);
hipError_t cgc_real_part(uint64_t m, const double2* x, double* a, hipStream_t st);
hipError_t cgc_poisson_scale(const Lattice& L, uint64_t m, double inv_n, double2* x, hipStream_t st);
// Wilson-line steps (Wz), the lines at -at/2 (Vm) and at +at/2 (Vp), each nc-colour [real][cell]
hipError_t cgc_wilson_lines(int nc, const Lattice& L, int orientation, int mfrac, const double* phi, double* Wz,
                            double* Vm, double* Vp, hipStream_t st);
hipError_t cgc_links(int nc, int prec, const Lattice& L, const double* Vm, const double* Vp, void* U, void* Unext,
                     void* E, hipStream_t st);
hipError_t cgc_gauss(int nc, int prec, const Lattice& L, const Box& active, const double* Vm, const double* Vp,
                     const void* U, const void* Unext, const void* rho, double* gauss, hipStream_t st);
hipError_t cgc_plane_max(int nc, const Lattice& L, const double* gauss, double* mx, hipStream_t st);
// Lists of at most kRefineMatrixMax charges are refined by the 100th powers of the sweep
// matrices; longer ones sweep by sweep. work: cgc_sp_work doubles; q: [coefficient][q_total]
// in the context precision.
constexpr int kRefineMatrixMax = 2048;
size_t cgc_sp_work(const SpBlock& b, uint32_t nT, int ad);
hipError_t cgc_sp_charges(int nc, int prec, const Lattice& L, const SpBlock& b, const double* gauss, double* work,
                          void* q, int64_t q_total, int64_t q_base, hipStream_t st);
// per-particle CGC particles of planes b.zs .. b.zs + b.bw - 1 (b.ppc per cell) written at
// [base, base + b.bw N1 N2 b.ppc) of count-long particle arrays; work: cgc_sp_work doubles
hipError_t cgc_particles(int nc, int prec, const Lattice& L, const SpBlock& b, int orientation, const double* gauss,
                         double* work, int* dir, double* pos0, double* pos1, double* vel, void* Q0, void* Q1,
                         uint64_t count, uint64_t base, hipStream_t st);
// java.util.Random(seed).nextGaussian() x count on the device (seed already scrambled with
// 0x5DEECE66D), from `trials` polar-method trials; *accepted (host) receives the number of
// accepted trials, which must reach count / 2 rounded up. work: java_gaussian_work(trials) bytes.
size_t java_gaussian_work(uint64_t trials);
hipError_t java_gaussians(uint64_t scrambled_seed, uint64_t count, uint64_t trials, void* work, double* out,
                          uint64_t* accepted, hipStream_t st);
// Wilson-line observables of the plane behind a nucleus, V [real][N1 N2]: the sum of Re tr V,
// and tr(V_x V_y^+) binned by distance (psum, pcnt: nbins * N1 N2 partials)
hipError_t cgc_tadpole(int nc, uint32_t T, const double* V, double* out, hipStream_t st);
hipError_t cgc_dipole(int nc, uint32_t n1, uint32_t n2, double a1, double a2, double ds, int nbins, const double* V,
                      double* psum, unsigned long long* pcnt, double* corr, unsigned long long* cnt, hipStream_t st);

}  // namespace opx
