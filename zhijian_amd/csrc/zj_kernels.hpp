// zj_kernels.hpp -- device-side data layout and launch interface of the fused
// FR stage kernels (zj_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <new>

namespace zj {

// Kernel execution timing.  While `start` is set (zj_time_stages and the
// triangle equivalent), stage launches pass these events to
// hipExtLaunchKernelGGL, which records them when the kernel starts and ends:
// the kernel's own execution, without the gaps between launches.  Host state,
// per thread; never set while a graph is captured.
struct KernelTimer {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
};
KernelTimer& kernel_timer();

// Runs `steps` rounds of the three stage launches enq(1), enq(2), enq(3), each
// with kernel timing events, and returns in stage_ms[3] the mean execution
// time of each stage's kernel (ms).  enq returns a hipError_t.
template <typename Enq>
hipError_t time_stage_kernels(Enq enq, int steps, hipStream_t s, double* stage_ms) {
    const int n = 3 * steps;
    hipEvent_t* ev = new (std::nothrow) hipEvent_t[2 * (size_t)n]();
    if (!ev) return hipErrorOutOfMemory;
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 * n && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
    KernelTimer& kt = kernel_timer();
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        kt.start = ev[2 * i];
        kt.stop = ev[2 * i + 1];
        e = enq(1 + i % 3);
    }
    kt = KernelTimer{};
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
        acc[i % 3] += ms;
    }
    for (int i = 0; i < 2 * n; ++i)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    delete[] ev;
    if (e == hipSuccess)
        for (int k = 0; k < 3; ++k) stage_ms[k] = acc[k] / steps;
    return e;
}

template <typename... Formals, typename... Actuals>
inline void klaunch(void (*k)(Formals...), dim3 grid, dim3 block, unsigned shmem, hipStream_t s,
                    Actuals... args) {
    const KernelTimer& t = kernel_timer();
    if (t.start)
        hipExtLaunchKernelGGL(k, grid, block, shmem, s, t.start, t.stop, 0u, args...);
    else
        hipLaunchKernelGGL(k, grid, block, shmem, s, args...);
}

constexpr int NV = 4;           // conservative variables
constexpr int BLOCK = 256;      // threads per workgroup of the stage kernels

// Role of one (element, local edge) slot on its face.
enum EdgeRole : int { ROLE_LEFT = 0, ROLE_RIGHT = 1, ROLE_BOUNDARY = 2, ROLE_NONE = 3 };

// Per (element, local edge): neighbour element, neighbour edge (or BC type on a
// boundary), role, BC data index.
struct alignas(16) EdgeLink {
    int nbr;
    int nbr_edge;
    int role;
    int bc_idx;
};

// Sum-factorised operators of one order, extracted from the reference's dense
// tables: D[i][k] = l_k'(x_i), Lm[j] = l_j(-1), Lp[j] = l_j(+1), corr[4][S].
template <typename R>
struct OpTables {
    R D[36];
    R Lm[6];
    R Lp[6];
    R corr[4 * 36];
    R gL[6];   // textbook mode: VCJH g_L', g_R' at the 1D solution points
    R gR[6];
};

enum StageMode : int { MODE_RESIDUAL = 0, MODE_STAGE1 = 1, MODE_STAGE2 = 2, MODE_STAGE3 = 3 };

// Arguments of one stage launch.  State planes are SoA: plane v holds E*S
// doubles, element-major then solution point.
//
// FP64 form: the stage state lives in U_in/U_out (U1, U2 in the W buffer).
// FP32-bulk form: the intermediate stage state is stored as an FP32 increment
// over U^n (dW_in/dW_out): U1 = U^n + dW1, U2 = U^n + dW2; the increments are
// of order dt*R, so their FP32 rounding sits far below the FP32 residual's.
template <typename R>
struct StageArgs {
    int E;
    const double* U_in;      // stage input state (FP64 form; U^n in stage 1 of both forms)
    const double* U0;        // U^n (stages 2, 3)
    double* U_out;           // stage output state (may alias U_in or U0 element-locally)
    const float* dW_in;      // FP32-bulk: increment of the stage input over U^n (stages 2, 3)
    float* dW_out;           // FP32-bulk: increment of the stage output (stages 1, 2)
    double* R_out;           // residual planes (MODE_RESIDUAL)
    const R* Jinv;           // 4 planes of E*S, or [E][4] for affine meshes
    const R* T_in;           // traces of U_in [E][4][Fp][4]
    R* T_out;                // traces of U_out
    const EdgeLink* links;   // [E][4]
    const R* normals;        // [E][4][2]
    const double* bc_data;   // [n][8]
    const double* dt_ptr;    // device scalar
    double* l2_partial;      // [grid] sum of R^2 per workgroup, or nullptr
    double gamma;
    int positivity;
    int xcd_groups;          // workgroup -> tile remap groups (1: identity)
    // textbook mode
    const R* edge_f;         // [E][4] sign * edge half length (/ J on affine meshes)
    const R* Jpt;            // [E*S] Jacobian determinant (non-affine meshes)
    int riemann;             // 0 Rusanov, 1 Roe, 2 HLLC
};

template <typename R>
struct TraceArgs {
    int E;
    const double* U;         // SoA planes
    R* T_out;
    int xcd_groups;
};

// Viscous terms of the textbook scheme (zj_visc.hip): constant viscosity and
// conductivity, BR2 penalty, and the correction derivatives g_L'(-1), g_R'(+1).
struct ViscParams {
    double mu, kappa, eta, gLe, gRe;
};

// One mesh of a batched launch: its stage arguments for one mode, its
// operator tables and order, its scheme (reference arithmetic or textbook,
// with the viscous parameters of the Navier-Stokes terms).
template <typename R>
struct BatchInst {
    StageArgs<R> args;
    OpTables<R> ops;
    int P;
    int affine;
    int textbook;
    int viscous;
    ViscParams vp;
};

// Launch wrappers (host).  `epb` = elements per workgroup (0 = default).
template <typename R>
hipError_t launch_stage(int P, int mode, const StageArgs<R>& a, const OpTables<R>& ops,
                        int epb, bool affine, hipStream_t s, bool textbook = false);
template <typename R>
hipError_t launch_trace(int P, const TraceArgs<R>& a, const OpTables<R>& ops, int epb,
                        hipStream_t s);
// items[k] = (instance, tile) over instances of order P; insts holds the
// instances' arguments for `mode`.
template <typename R>
hipError_t launch_batched_stage(int P, bool affine, int mode, const BatchInst<R>* insts, const int2* items,
                                int n_items, int grid, hipStream_t s, bool textbook = false);
// The same stages for every member of a batch in ONE launch per stage:
// items[k] = (index into insts, tile), insts across all orders and metric
// forms; lds_bytes = batched_all_lds of the largest order present.
template <typename R>
hipError_t launch_batched_all(int mode, const BatchInst<R>* insts, const int2* items, int n_items, int grid,
                              int lds_bytes, hipStream_t s);
int batched_all_lds(int P, bool fp32);
int default_epb(int P);
int max_epb(int P);

// Stage launch with viscous terms (affine quadrilaterals, no boundary faces,
// traces of 12 words per flux point: U, dU/dx, dU/dy); tile = max_epb(P).
template <typename R>
hipError_t launch_visc_stage(int P, int mode, const StageArgs<R>& a, const OpTables<R>& ops, const ViscParams& vp,
                             hipStream_t s);
template <typename R>
hipError_t launch_visc_trace(int P, int E, const double* U, const R* Jinv_e, R* T, const OpTables<R>& ops,
                             int xcd_groups, hipStream_t s);
// The viscous stage for many meshes of order P (items = (instance, tile)).
template <typename R>
hipError_t launch_batched_visc(int P, int mode, const BatchInst<R>* insts, const int2* items, int n_items, int grid,
                               hipStream_t s);
// The viscous stage for many meshes of every order in one launch
// (lds_bytes = visc_all_lds of the largest order present).
template <typename R>
hipError_t launch_batched_visc_all(int mode, const BatchInst<R>* insts, const int2* items, int n_items, int grid,
                                   int lds_bytes, hipStream_t s);
int visc_all_lds(int P, bool fp32);
int num_blocks(int E, int epb);

// perm[e] = caller index of device element e (nullptr: same order), S points per element
hipError_t launch_aos_to_soa(const double* aos, double* soa, int64_t n_points, hipStream_t s,
                             const int* perm = nullptr, int S = 1);
hipError_t launch_soa_to_aos(const double* soa, double* aos, int64_t n_points, hipStream_t s,
                             const int* perm = nullptr, int S = 1);
hipError_t launch_set_scalar(double* p, double v, hipStream_t s);
// dst[0..n) = v[0..n) by value through kernel arguments (no host staging)
struct ScalarChunk {
    double v[64];
    int n;
};
hipError_t launch_set_many(double* dst, const double* v, int n, hipStream_t s);
hipError_t launch_rk3_combine(int stage, double* out, const double* U0, const double* Uk, const double* R,
                              int64_t n, double dt, hipStream_t s);
hipError_t launch_euler_update(double* U, const double* R, int64_t n, double dt, hipStream_t s);
// dt = min over points of cfl*h_e/(ws*cf), h_e = sqrt(area_e); with visc > 0
// (the viscous terms: (p+1)^4 max(4/3 mu, (gamma-1) kappa)) the minimum of
// cfl / (cf*ws/h_e + visc/(rho h_e^2)); partial per block.
hipError_t launch_dt_partial(const double* U, const double* hcell, int E, int S, double gamma,
                             double cfl, double cf, double visc, double* partial, int nblk, hipStream_t s);
// Ub += dt_bar * d(dt)/dU, dt the minimum above (*dtmin, already on the
// device): the derivative at the minimising point (lowest index among ties),
// found into *idx.
hipError_t launch_dt_cfl_vjp(const double* U, const double* hcell, int E, int S, double gamma, double cfl,
                             double cf, double visc, const double* dtmin, unsigned long long* idx, double dt_bar,
                             double* Ub, int nblk, hipStream_t s);
hipError_t launch_count_nonfinite(const double* U, int64_t n, unsigned long long* count,
                                  int nblk, hipStream_t s);
// Sum over the points of w (m_x^2 + m_y^2) / (2 rho) as nblk partials, and its
// gradient with respect to the state into the planes G (perm as above, w in
// the caller's element order [E][S]).
hipError_t launch_kinetic_energy(const double* U, const double* w, const int* perm, int E, int S, double* G,
                                 double* partial, int nblk, hipStream_t s);
// Deterministic reduction of block partials (fixed order): sqrt of the sum
// (an L2 norm), the minimum, or the plain sum.
enum FinalizeOp : int { FIN_SQRT_SUM = 0, FIN_MIN = 1, FIN_SUM = 2 };
hipError_t launch_finalize(const double* partial, int n, int op, double* out, hipStream_t s);

}  // namespace zj
