// opx_ctx.h -- the context behind the opaque opx_ctx handle of include/opx.h
// and the host helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "opx.h"
#include "opx_kernels.h"

namespace opx_host {
struct AdjointWork;
struct CgcWork;
}

struct opx_ctx {
    int nc = 2;
    int prec = OPX_FP64;
    int dims[3] = {0, 0, 0};
    double as[3] = {1, 1, 1};
    double at = 0.0;
    double g = 1.0;
    int64_t n = 0;
    int G = 4;     // reals per link in reference storage
    int S = 4;     // reals stored per link on the device (FP32 SU(3): the first two rows, 12)
    int AD = 3;    // algebra coefficients on the device
    int AS = 3;    // algebra reals in reference storage
    size_t tsz = 8;
    void* U = nullptr;
    void* Unext = nullptr;
    void* E = nullptr;
    void* J = nullptr;
    void* rho = nullptr;
    void* P = nullptr;              // proj of the three plaquettes per cell (plaquette form scratch)
    // field solver: OPX_SOLVER_FAST or OPX_SOLVER_IMPLICIT with its iterations and damping;
    // the implicit solver's copies of U, Unext and E (allocated on first use)
    int solver = 0;
    int implicit_iterations = 0;
    double implicit_damping = 0.0;
    void* iU = nullptr;
    void* iUn = nullptr;
    void* iE = nullptr;
    bool j_nonzero = false;
    // disjoint x0 plane ranges, in order, outside which J is zero (at most opx::kMaxCurrentSlabs;
    // j_nonzero: the list is not empty)
    std::vector<std::pair<int, int>> j_slabs;
    opx::Box active{{0, 0, 0}, {0, 0, 0}, 1};
    opx::Box eval{{0, 0, 0}, {0, 0, 0}, 1};
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    double* partials = nullptr;
    double* dout = nullptr;
    double* diag = nullptr;         // per-slice sums of the projected diagnostics (4 x the longest axis)
    unsigned long long* counters = nullptr;
    // super particles
    int sp_count = 0;
    int sp_max_num = 0;
    int64_t sp_total = 0;
    int* sp_desc = nullptr;         // 6 arrays of sp_count ints, back to back
    std::vector<int> sp_host;       // host copy of the descriptors (same layout)
    int64_t* sp_qoff = nullptr;
    void* sp_q = nullptr;           // [coef][total]
    // per-particle CGC NGP particles: [3][count] doubles, charges [coef][count]
    uint32_t p_count = 0;
    int* p_dir = nullptr;
    double* p_pos0 = nullptr;
    double* p_pos1 = nullptr;
    double* p_vel = nullptr;
    void* p_Q0 = nullptr;
    void* p_Q1 = nullptr;
    // work arrays of the ordered deposits (opx::ParticleSort), sized for ps_capacity particles
    uint32_t* ps_buf = nullptr;     // keys, values and their sorted copies, 4 x ps_capacity
    void* ps_temp = nullptr;
    size_t ps_temp_bytes = 0;
    uint32_t ps_capacity = 0;
    // cell segments of the super-particle charge-density gather (opx::CellSegments), rebuilt
    // per deposit: a ring of pinned host staging slots, each with the event that marks its
    // copy done (the host runs up to kSegSlots deposits ahead), and one device copy (stream
    // order keeps a copy behind the previous deposit's kernel)
    static constexpr int kSegSlots = 4;
    unsigned char* rs_host[kSegSlots] = {};
    hipEvent_t rs_copied[kSegSlots] = {};
    int rs_slot = 0;
    unsigned char* rs_dev = nullptr;
    size_t rs_capacity = 0;         // bytes of the device copy and of each slot
    bool rho_pending = false;       // charge density deposit deferred until rho is read
    int rho_pending_t = 0;
    int steps = 0;                  // Simulation.totalSimulationSteps
    // fused-update kernel selection (opx_set_option / opx_autotune)
    int kernel = OPX_KERNEL_AUTO;
    int tile = 0;
    int nchunks = 0;                // 0: derived from the CU count
    int num_cus = 256;
    bool force_global = false;      // OPX_OPT_GLOBAL_POINTERS
    int reunit_every = 0;           // OPX_OPT_REUNIT_EVERY
    int device = 0;                 // the GPU current when the context was created
    bool batch_view = false;        // event of an opx_batch: memory borrowed, stepping through the batch only
    opx_host::AdjointWork* adj = nullptr;   // checkpoint and cotangent buffers of opx_adjoint_gradient
    opx_host::CgcWork* cgc = nullptr;       // charge density and Gauss violation of the CGC initial condition
    std::string err;
};

namespace opx_host {

// record `msg` as the context's last error and return `code`
int fail(opx_ctx* c, int code, const std::string& msg);
// OPX_OK for hipSuccess, else OPX_E_DEVICE with the HIP error text
int hip_check(opx_ctx* c, hipError_t e, const char* what);
// geometry and step constants of the context's lattice for time step `at`
opx::Lattice lattice(const opx_ctx* c, double at);
// the fused E/link update with the context's configured kernel form
hipError_t launch_update(opx_ctx* c, const opx::Lattice& L, const void* U, void* Unext, void* E);
// frees the adjoint workspace (opx_adjoint_capi.cpp)
void adjoint_release(opx_ctx* c);
// frees the CGC initial-condition buffers (opx_cgc_capi.cpp)
void cgc_release(opx_ctx* c);
// one algebra element: reference storage (3 or 9 reals) <-> device coefficients (3 or 8)
void storage_to_coeffs(int nc, const double* v, double* c);
void coeffs_to_storage(int nc, const double* c, double* v);

}  // namespace opx_host
