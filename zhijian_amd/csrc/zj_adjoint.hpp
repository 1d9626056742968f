// zj_adjoint.hpp -- argument blocks and launch wrappers of the stage VJP
// kernels (zj_adjoint.hip).
#pragma once

#include "zj_kernels.hpp"

namespace zj {

// One adjoint launch for a stage u_n = a U^n + b q + c dt R(q), u_out = limiter(u_n)
// (MODE_STAGE1..3), or for the residual alone (MODE_RESIDUAL: Ub_in = cotangent of R).
template <typename R>
struct VjpArgs {
    int E;
    // forward inputs of the stage, as in StageArgs
    const double* U_in;      // q (FP64 form)
    const double* U0;        // U^n (stages 2, 3)
    const float* dW_in;      // FP32-bulk: q = U^n + dW_in (stages 2, 3)
    const R* Jinv;
    const R* T_in;           // traces of q on edges read across tiles
    const EdgeLink* links;
    const R* normals;
    const double* bc_data;
    const double* dt_ptr;
    // cotangents
    const double* Ub_in;     // partial cotangent of u_out (planes), or of R
    const R* Tb_in;          // cross-tile edge cotangents completing Ub_in (nullptr: none)
    double* Ub_out;          // partial cotangent of q (may alias Ub_in)
    R* Tb_out;               // cross-tile edge cotangents of q, written at the neighbour's slots
    double* U0b;             // cotangent of U^n: stage 3 writes, stage 2 adds, stage 1 reads
    double* dtb_partial;     // per-tile sum of c * R . du_n (nullptr: skip)
    double* mub_partial;     // viscous terms: per-tile running sum of <cotangent of F_v, F_v> (nullptr: skip)
    double gamma;
    int positivity;
    int xcd_groups;
    // textbook mode (as StageArgs)
    const R* edge_f;         // [E][4] sign * edge half length (/ J on affine meshes)
    const R* Jpt;            // [E*S] Jacobian determinant (non-affine meshes)
    int riemann;             // 0 Rusanov, 1 Roe, 2 HLLC
};

// One mesh of a batched VJP launch: its stage-VJP arguments, operator
// tables and viscous parameters.
template <typename R>
struct BatchVjpInst {
    VjpArgs<R> a;
    OpTables<R> ops;
    ViscParams vp;
};

template <typename R>
struct CompleteArgs {
    int E;
    const double* Ub;
    const R* Tb;
    double* out;
    const EdgeLink* links;
    int xcd_groups;
};

int adjoint_tile(int P);
// textbook: the VJP of the textbook scheme's stage (VCJH correction with
// metrics, the configured Riemann solver), else of the reference arithmetic.
template <typename R>
hipError_t launch_vjp(int P, int mode, const VjpArgs<R>& a, const OpTables<R>& ops, bool affine, hipStream_t s,
                      bool textbook = false);
template <typename R>
hipError_t launch_complete(int P, const CompleteArgs<R>& a, const OpTables<R>& ops, hipStream_t s);
// The same stage VJP over several meshes of one order, metric form and scheme
// in one launch: work items (member, tile at adjoint_tile(P)), one per workgroup.
template <typename R>
hipError_t launch_batched_vjp(int P, bool affine, bool textbook, int mode, const BatchVjpInst<R>* insts,
                              const int2* items, int n_items, hipStream_t s);

// The BR2 (Navier-Stokes) stage's adjoint (zj_visc.hip): its own tile
// (visc_adjoint_tile), the forward stages 1 and 2 recomputed at that tile,
// the stage VJP, and the completion of 12-word edge cotangents (U, grad U).
int visc_adjoint_tile(int P, bool fp64);
template <typename R>
hipError_t launch_visc_stage_adj(int P, int mode, const StageArgs<R>& a, const OpTables<R>& ops,
                                 const ViscParams& vp, hipStream_t s);
template <typename R>
hipError_t launch_visc_vjp(int P, int mode, const VjpArgs<R>& a, const OpTables<R>& ops, const ViscParams& vp,
                           hipStream_t s);
template <typename R>
hipError_t launch_visc_complete(int P, const CompleteArgs<R>& a, const R* Jinv_e, const OpTables<R>& ops,
                                hipStream_t s);
// The BR2 stage VJP over several meshes of one order (work items at
// visc_adjoint_tile), one per workgroup.
template <typename R>
hipError_t launch_batched_visc_vjp(int P, int mode, const BatchVjpInst<R>* insts, const int2* items, int n_items,
                                   hipStream_t s);

}  // namespace zj
