/*
 * zhijian_amd.h -- C ABI of the MI355X flux-reconstruction residual and SSP-RK3
 * time step.
 *
 * This is the boundary under the C++ FRSolver / SSPRK3 API (include/zhijian/):
 * plain pointers and sizes, int status codes (ZJ_OK = 0), never exit().  One
 * handle owns one mesh's device state on one GPU.  Host arrays are copied at
 * the call; the caller keeps ownership.  Streams are hipStream_t passed as
 * void* (NULL = the handle's own stream); work a call queues on a caller's
 * stream is ordered against the handle's other calls by that caller
 * (synchronize that stream, or pass it again).  Calls on one handle are not
 * thread-safe; different handles may be used from different threads.  The
 * error text of calls without a handle (a failed create, zj_gmsh_read, the
 * I/O calls) is kept per thread.
 *
 * Reference interfaces replaced (Zhijian, file:line):
 *   zj_create        FRSolver::initialize/allocateGPUMemory/initGeometry
 *                    (src/solver/fr_solver.cpp:30-266) -- the device upload set
 *   zj_set_state     FRSolver::copyToDevice  (fr_solver.cpp:854-871)
 *   zj_get_state     FRSolver::copyToHost    (fr_solver.cpp:835-852)
 *   zj_residual      FRSolver::computeResidual_GPU / computeInviscidFlux_GPU
 *                    (fr_solver.cpp:335-608): kernels.cu K4,K1,K9,K8,K11,K13
 *   zj_rk3_step      FRSolver::advance (fr_solver.cpp:305-333) with
 *                    gpu::rkStage1/2/3 + enforcePositivity (kernels.cu:1236-1349)
 *   zj_rk3_stage     SSPRK3::stage{1,2,3}_GPU (include/time/rk_integrator.hpp:33-58)
 *   zj_dt_cfl        FRSolver::computeTimeStep (fr_solver.cpp:784-816)
 *   zj_residual_l2   FRSolver::computeResidual (fr_solver.cpp:818-833)
 *   zj_check_finite  the checkNaN scans (fr_solver.cpp:11-22), as a status
 *   zj_quad_operators  FROperators::initQuad (src/basis/basis.cpp:464-613)
 *   zj_build_faces     Mesh::buildFaces (src/mesh/mesh.cpp:18-83)
 *   zj_face_normals    Mesh::faceNormal (mesh.cpp:159-183)
 *   zj_elem_areas      Mesh::elementArea (mesh.cpp:122-146)
 *   zj_geometry        FRSolver::initGeometry (fr_solver.cpp:85-129)
 *   zj_pair_periodic, zj_pair_periodic_nodes
 *                      no counterpart (the reference has no periodic faces)
 *
 * Array layouts at this boundary are the reference's (include/gpu/kernels.hpp):
 *   U, dUdt  [E][S][4] AoS (rho, rhou, rhov, rhoE), S = (p+1)^2
 *   Jinv     [E][S][4] (dxi/dx, dxi/dy, deta/dx, deta/dy);  J [E][S]
 *   diff_xi, diff_eta [S][S];  interp [4][p+1][S];  corr [4][S]
 *   normals  [F][2];  face_* [F];  bc_data [n_bc][8]
 */
#ifndef ZHIJIAN_AMD_H
#define ZHIJIAN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZJ_OK 0
#define ZJ_ERR_ARG 1          /* invalid argument or inconsistent tables */
#define ZJ_ERR_HIP 2          /* HIP runtime error (message in zj_last_error) */
#define ZJ_ERR_NOMEM 3
#define ZJ_ERR_UNSUPPORTED 4  /* valid input this library does not handle */
#define ZJ_ERR_NONFINITE 5    /* state or residual holds NaN/Inf */
#define ZJ_ERR_STATE 6        /* call out of order (e.g. no state set) */

#define ZJ_PREC_FP64 0        /* plain: FP64 end to end */
#define ZJ_PREC_FP32BULK 1    /* FP32 interpolation/flux/correction, FP64 RK state */

#define ZJ_COMPAT_REFERENCE 0 /* the reference's executed arithmetic (quirks kept) */
#define ZJ_COMPAT_TEXTBOOK 1  /* the scheme its settings name: VCJH FR with metrics (Huynh 2007,
                                 Vincent et al. 2011) and the configured Riemann solver */
#define ZJ_RIEMANN_RUSANOV 0  /* RiemannSolver (types.hpp:66-70), textbook mode only */
#define ZJ_RIEMANN_ROE 1
#define ZJ_RIEMANN_HLLC 2
#define ZJ_CORR_DG 0          /* FluxType (types.hpp:59-64), textbook mode only */
#define ZJ_CORR_SD 1
#define ZJ_CORR_HU 2
#define ZJ_CORR_GA 3

/* BCType values (include/common/types.hpp:47-56) */
#define ZJ_BC_INTERIOR 0
#define ZJ_BC_WALL 1
#define ZJ_BC_SLIPWALL 2
#define ZJ_BC_SYMMETRY 3
#define ZJ_BC_FARFIELD 4
#define ZJ_BC_INFLOW 5
#define ZJ_BC_OUTFLOW 6
#define ZJ_BC_PERIODIC 7

typedef struct zj_solver* zj_handle;

/* What FRSolver uploads: operators flattened as fr_solver.cpp:164-221, geometry
 * as fr_solver.cpp:85-129, connectivity and normals as fr_solver.cpp:232-265,
 * BC tables as fr_solver.cpp:432-494 (bc_type/bc_face_map per face, 8 doubles
 * of bc_data per boundary entry). */
typedef struct {
    int poly_order;              /* 1..5 */
    int n_elem;
    int n_faces;
    const double* diff_xi;       /* [S*S] */
    const double* diff_eta;      /* [S*S] */
    const double* interp;        /* [4*(p+1)*S] */
    const double* corr;          /* [4*S] */
    const double* J;             /* [E*S] or NULL (the reference residual never reads it) */
    const double* Jinv;          /* [E*S*4] */
    const double* normals;       /* [F*2] */
    const int* face_left_elem;   /* [F] */
    const int* face_left_local;  /* [F] */
    const int* face_right_elem;  /* [F], -1 on boundary faces */
    const int* face_right_local; /* [F] */
    const int* bc_type;          /* [F] or NULL: no boundary faces expected */
    const int* bc_face_map;      /* [F] or NULL */
    const double* bc_data;       /* [n_bc_data*8] or NULL */
    int n_bc_data;
    const double* elem_area;     /* [E] or NULL (then the CFL dt is unavailable) */
    const double* edge_half_len; /* [E][4] half lengths of the straight element edges
                                    (edge k = node k -> node k+1); textbook mode only */
    const double* elem_centroid; /* [E][2] or NULL.  When given, the library numbers
                                    the elements along a Hilbert curve internally
                                    (compact element tiles); every array the caller
                                    passes or receives stays in the caller's order. */
} zj_mesh_desc;

typedef struct {
    double gamma;                /* SimParams::gamma */
    double cfl;                  /* SimParams::CFL, used by the fused CFL dt */
    int precision;               /* ZJ_PREC_* */
    int compat;                  /* ZJ_COMPAT_REFERENCE */
    int positivity;              /* 1: the reference's limiter after each stage */
    int device;                  /* HIP device ordinal */
    int riemann;                 /* ZJ_RIEMANN_*, textbook mode */
    int corr;                    /* ZJ_CORR_*, textbook mode */
    int viscous;                 /* SimParams::viscous.  Reference mode: the reference computes the
                                    viscous terms and discards them (fr_solver.cpp:645-646), so they
                                    add nothing.  Textbook mode: compressible Navier-Stokes terms with
                                    BR2 (affine meshes without boundary faces) */
    double Re;                   /* SimParams::Re: mu = 1/Re */
    double Pr;                   /* SimParams::Pr: kappa = mu gamma / ((gamma - 1) Pr) */
    double br2_eta;              /* BR2 penalty (<= 0: 2) */
} zj_params;

/* ---- solver --------------------------------------------------------- */
int zj_create(zj_handle* out, const zj_mesh_desc* mesh, const zj_params* params);
int zj_destroy(zj_handle h);
/* Last error text for h, or the last zj_create failure when h is NULL. */
const char* zj_last_error(zj_handle h);
int zj_version(void);

int zj_set_state(zj_handle h, const double* U_aos);          /* host [E][S][4] */
int zj_get_state(zj_handle h, double* U_aos);                /* host, synchronises */
int zj_set_state_device(zj_handle h, const double* d_U_aos, void* stream);
int zj_get_state_device(zj_handle h, double* d_U_aos, void* stream);
/* Device-resident state slots (no reference counterpart): checkpoints of a
 * trajectory kept on the GPU in the library's own layout, for checkpointed
 * adjoints and restarts within a run.  zj_state_slots allocates n slots of
 * the handle's state (n = 0 frees them; existing slots are dropped);
 * zj_state_save copies the current state into a slot and zj_state_load makes
 * a slot the current state, both asynchronous on `stream`. */
int zj_state_slots(zj_handle h, int n);
int zj_state_save(zj_handle h, int slot, void* stream);
int zj_state_load(zj_handle h, int slot, void* stream);

/* dUdt of the current state into dUdt_aos (host pointer; NULL keeps it on the
 * device for zj_residual_l2).  Synchronises when dUdt_aos is not NULL. */
int zj_residual(zj_handle h, double* dUdt_aos, void* stream);
/* The same into a device buffer [E][S][4]; asynchronous on `stream`. */
int zj_residual_device(zj_handle h, double* d_dUdt_aos, void* stream);
/* One SSP-RK3 step of size dt: three fused stage kernels (hipGraph replay). */
int zj_rk3_step(zj_handle h, double dt, void* stream);
/* One stage (1, 2 or 3) of the step, for callers that time or interleave them.
 * Stage 1 must come first; the state reads as U^n again only after stage 3. */
int zj_rk3_stage(zj_handle h, int stage, double dt, void* stream);
/* Waits for the work queued on the handle's own stream. */
int zj_synchronize(zj_handle h);
/* Measurement (no reference counterpart): `steps` SSP-RK3 steps of size dt,
 * each stage launched on its own with events that the runtime records when
 * the kernel starts and ends; stage_ms[3] receives each stage kernel's mean
 * execution time in ms (the quantity a kernel profiler reports, without the
 * gaps between launches).  The state advances by `steps` steps. */
int zj_time_stages(zj_handle h, double dt, int steps, double* stage_ms, void* stream);
/* CFL time step of the current state (synchronises): the minimum over the
 * points of cfl h / ((2p+1) ws), h = sqrt(element area), ws = |u| + c; with
 * the viscous terms, of cfl / ((2p+1) ws / h + (p+1)^4 nu_k / (rho h^2)),
 * nu_k = max(4/3 mu, (gamma-1) kappa). */
int zj_dt_cfl(zj_handle h, double cfl, double* dt);
/* L2 norm of the last residual: that of zj_residual, or of stage 3 of the last
 * step (synchronises). */
int zj_residual_l2(zj_handle h, double* l2);
/* Number of non-finite values in the current state (synchronises). */
int zj_check_finite(zj_handle h, int64_t* n_bad);
/* Replaces the boundary tables given at zj_create (same layout: bc_type and
 * bc_face_map per face of the creation face list, 8 doubles of bc_data per
 * entry).  FRSolver::computeInviscidFlux_GPU rebuilds these tables at every
 * residual (fr_solver.cpp:432-494); here they change only when re-set. */
int zj_set_boundary(zj_handle h, const int* bc_type, const int* bc_face_map, const double* bc_data,
                    int n_bc_data);
/* Elementwise SSP-RK3 stage combination on device arrays of n doubles, as
 * gpu::rkStage1/2/3 (kernels.cu:1297-1349) without the limiter:
 *   stage 1: out = U0 + dt R;  stage 2: out = 3/4 U0 + 1/4 Uk + 1/4 dt R;
 *   stage 3: out = 1/3 U0 + 2/3 Uk + 2/3 dt R   (Uk unused for stage 1). */
int zj_rk3_combine(int stage, double* out, const double* U0, const double* Uk, const double* R,
                   int64_t n, double dt, void* stream);
/* U += dt R on device arrays (ForwardEuler::advance_GPU). */
int zj_euler_update(double* U, const double* R, int64_t n, double dt, void* stream);
/* Launch configuration of the stage kernel (elements per workgroup). */
int zj_set_tile(zj_handle h, int elems_per_block);
int zj_get_info(zj_handle h, int64_t* n_dof, int* elems_per_block, int* n_sp);
/* Element edges whose neighbour lies in another tile at the current tile size
 * (their traces go through HBM), and whether the elements were renumbered. */
int zj_get_tiling(zj_handle h, int64_t* cross_tile_edges, int* reordered);
/* Metric storage of the stage kernel.  The reference stores Jinv per solution
 * point (fr_solver.cpp:85-129); when every element is affine (its per-point
 * Jinv agree to 1e-13 relative) the library reads one Jinv per element and
 * differentiates the contravariant fluxes instead, which is the default for
 * such meshes.  mode: -1 query, 0 per-point, 1 per-element (ZJ_ERR_UNSUPPORTED
 * on a non-affine mesh).  *active (may be NULL) receives the mode in use. */
int zj_set_geometry_mode(zj_handle h, int mode, int* active);
/* Workgroup-to-tile order of the stage kernel: `groups` contiguous runs of
 * tiles, one per XCD (default: the device's XCD count; 1 = launch order).
 * Results do not depend on it. */
int zj_set_xcd_groups(zj_handle h, int groups);

/* ---- discrete adjoint (no reference counterpart) ----------------------
 * Vector-Jacobian products of the reference-arithmetic path: the derivative
 * of every executed operation (fluxes kernels.cu:157-194, Rusanov and ghost
 * states :453-701, flux difference :346-450, divergence and correction
 * :813-1034, SSP-RK3 and the limiter :1236-1349), taking the branch of each
 * clamp the forward took.  In textbook mode, of the textbook scheme (VCJH
 * correction with metrics, Roe / HLLC / Rusanov of euler.hpp:204-375) and,
 * with the Navier-Stokes terms, of BR2 (navierstokes.hpp:53-99, 170-205).
 *
 * The handle keeps the cotangent of its state on the device.  Reverse time
 * stepping: zj_adjoint_set(dJ/dU^N); then for n = N-1 .. 0:
 * zj_set_state(U^n), zj_rk3_step_vjp(dt_n) -> the cotangent of U^n (the step
 * is recomputed from U^n); finally zj_adjoint_get.  dt_bar (may be NULL)
 * receives dJ/d(dt_n) for that step. */
int zj_adjoint_set(zj_handle h, const double* ubar_aos);
/* Objective on the device (no reference counterpart): J = sum over the
 * points of w (m_x^2 + m_y^2) / (2 rho) of the current state, the domain
 * integral of the kinetic energy for w = quadrature weight x |J| (basis.cpp:484,
 * fr_solver.cpp:85-129); seeds the adjoint with dJ/dU as zj_adjoint_set would,
 * with no state or cotangent leaving the GPU.  d_weights: device [E][S] in the
 * caller's element order.  J (may be NULL) receives the value (synchronizes). */
int zj_adjoint_seed_kinetic_energy(zj_handle h, const double* d_weights, double* J, void* stream);
int zj_adjoint_set_device(zj_handle h, const double* d_ubar_aos, void* stream);
int zj_adjoint_get(zj_handle h, double* ubar_aos);
int zj_adjoint_get_device(zj_handle h, double* d_ubar_aos, void* stream);
int zj_rk3_step_vjp(zj_handle h, double dt, double* dt_bar, void* stream);
/* Vector-Jacobian product of the CFL step (FRSolver::computeTimeStep,
 * fr_solver.cpp:784-816, as zj_dt_cfl computes it): adds dt_bar times
 * d(dt_cfl)/dU at the handle's state to the held adjoint cotangent.  The
 * minimum's derivative is its minimising point's (the lowest point index
 * among ties); every max() takes the branch the forward took.  For a
 * trajectory stepped with dt_n = dt_cfl(U^n): after zj_rk3_step_vjp(dt_n,
 * &dt_bar) (state U^n set), call zj_dt_cfl_vjp(cfl, dt_bar). */
int zj_dt_cfl_vjp(zj_handle h, double cfl, double dt_bar, void* stream);
/* Checkpointed gradient, both passes native (no reference counterpart).
 * zj_forward_checkpointed: n_steps SSP-RK3 steps from the handle's state with
 * dt = dts[n], U^{c*every} saved in state slot c (the slots grow to
 * ceil(n_steps / every) if fewer).  zj_adjoint_sweep: with the cotangent of
 * U^N seeded (zj_adjoint_set), every step backwards: per segment, its
 * checkpoint, the segment's forward steps recorded on a device tape (stage
 * states and traces, every steps), then the stage VJPs from the tape; the
 * handle then holds U^0 and the cotangent of U^0 (zj_adjoint_get).  dt_bar
 * (n_steps values, or NULL) receives dJ/d(dt_n); with NULL the sweep does not
 * synchronize.
 * zj_adjoint_tape_segments: the last n segments (default 1; 0 = none) are
 * recorded on the tape by zj_forward_checkpointed itself, as it steps (at the
 * adjoint tile: same results), and zj_adjoint_sweep of the same trajectory
 * (n_steps, every, dts) then skips their recompute.  The tape holds
 * n * every steps of 3 states and 3 trace sets in device memory (TGV 128^2
 * P3 with the Navier-Stokes terms, FP64: 0.1 GB per step).  A sweep that
 * recomputes any segment, zj_state_save and zj_state_slots end the record. */
int zj_forward_checkpointed(zj_handle h, const double* dts, int n_steps, int every, void* stream);
int zj_adjoint_sweep(zj_handle h, const double* dts, int n_steps, int every, double* dt_bar, void* stream);
int zj_adjoint_tape_segments(zj_handle h, int n);
/* With the Navier-Stokes terms (textbook mode, BR2, navierstokes.hpp:53-99,
 * 170-205): dJ/dRe accumulated over the steps run backwards since
 * zj_adjoint_set (0 without viscous terms).  The viscous fluxes scale with
 * mu = 1/Re (kappa with it), so the kernels sum <cotangent of F_v, F_v> over
 * every point and face point; dJ/dRe = -mu * that sum.  Synchronizes the device. */
int zj_adjoint_get_re(zj_handle h, double* dJ_dRe);
/* (dR/dU)^T w at the handle's state, R = FRSolver::computeResidual_GPU's dU/dt
 * (fr_solver.cpp:335-608); w and ubar are [E][S][4]. */
int zj_residual_vjp(zj_handle h, const double* w_aos, double* ubar_aos);
int zj_residual_vjp_device(zj_handle h, const double* d_w_aos, double* d_ubar_aos, void* stream);

/* ---- batches and tuning (no reference counterpart) -------------------
 * A batch steps many independent meshes (each its own zj_handle, own order,
 * state, dt, scheme and Reynolds number) with one launch per stage and
 * (order, metric form, scheme) group -- reference arithmetic, textbook
 * Euler, textbook Navier-Stokes -- the groups running side by side in one
 * captured graph; each member's step equals its own zj_rk3_step bit for
 * bit.  Members keep working on their
 * own: a member's calls queue their work after the batch's last step (an
 * event the step records), so a state read right after zj_batch_rk3_step sees
 * the stepped state; a member must not be stepped on another stream while its
 * batch runs.
 * Members must use the full tile of their order (the default).  While a
 * batch holds a handle, zj_set_tile, zj_set_geometry_mode (a change),
 * zj_autotune and zj_destroy on it return ZJ_ERR_STATE: the batch's graph
 * reads the member's device buffers at that tiling.  zj_set_boundary on a
 * member is taken up by the batch's next step.
 * dts of zj_batch_rk3_step holds one time step per member. */
typedef struct zj_batch* zj_batch_handle;
int zj_batch_create(zj_batch_handle* out, const zj_handle* members, int n_members);
int zj_batch_destroy(zj_batch_handle b);
const char* zj_batch_last_error(zj_batch_handle b);
/* dts[i] is member i's step (FRSolver::advance, fr_solver.cpp:305-333, for all members). */
int zj_batch_rk3_step(zj_batch_handle b, const double* dts, void* stream);
/* One SSP-RK3 step backwards for every member (no reference counterpart; the
 * batched form of zj_rk3_step_vjp): each member must hold its state U^n and
 * an adjoint seed (zj_adjoint_set / zj_adjoint_set_device); each member's
 * forward stages are recomputed on its own, then every stage VJP runs as one
 * launch per (order, metric form, scheme) group over all members' tiles.
 * The cotangents equal the members' own zj_rk3_step_vjp bit for bit.
 * dt_bars (may be NULL) receives dJ/d(dts[i]) per member; with it the call
 * waits for the step. */
int zj_batch_rk3_step_vjp(zj_batch_handle b, const double* dts, double* dt_bars, void* stream);
int zj_batch_synchronize(zj_batch_handle b);
/* Workgroups of each launch: 0 = one per work item (element tile);
 * k > 0 = k per CU, each looping over its share of the items (persistent). */
int zj_batch_set_grid(zj_batch_handle b, int policy);
/* single = 1: ONE launch per stage over every member's tiles, whatever the
 * members' orders and metric forms (the stage body is picked per tile, the
 * LDS sized for the largest order present); 0 (default): one launch per
 * (order, metric form) group, the groups on parallel graph branches.
 * Results are the same either way.  single = 1 takes a batch whose members
 * all run the reference arithmetic, or all have the Navier-Stokes terms (one
 * BR2 launch per stage across the orders); any other batch returns
 * ZJ_ERR_UNSUPPORTED. */
int zj_batch_set_launch(zj_batch_handle b, int single);
/* Times `reps` batched steps of sizes dts for each launch form and grid
 * policy (0, 2, 4), keeps the fastest (*single, *policy, *ms per step; any may
 * be NULL), and restores every member's state as it was before the call;
 * the single launch is a candidate where it applies. */
int zj_batch_autotune(zj_batch_handle b, const double* dts, int reps, int* single, int* policy, double* ms);
/* Times the residual launch for every tile (full, half) and XCD order and
 * keeps the fastest; results do not depend on the choice. */
int zj_autotune(zj_handle h, int reps, int* elems_per_block, int* xcd_groups, double* ms);

/* ---- mesh ingestion: Gmsh MSH 2.2 / 4.1 ASCII -------------------------
 * GmshReader::read (src/mesh/gmsh_reader.cpp:161-760): nodes in file order,
 * kept 2D elements (3/6-node triangles, 4/9/25-node quads, 8-node
 * serendipity quads completed with a centre node), boundary line edges with
 * their physical tags, and boundary groups with names and inferred BC types.
 * Faces are built afterwards with zj_build_faces; zj_gmsh_edge_tag gives the
 * physical tag of a boundary face's vertex pair (0: none). */
typedef struct zj_gmsh* zj_gmsh_handle;
int zj_gmsh_read(const char* path, zj_gmsh_handle* out);
const char* zj_gmsh_error(void);
int zj_gmsh_free(zj_gmsh_handle g);
int zj_gmsh_counts(zj_gmsh_handle g, int64_t* n_nodes, int* n_elem, int* max_nodes_per_elem, int* n_groups);
int zj_gmsh_nodes(zj_gmsh_handle g, double* xy);
/* kind: ElementType (0 triangle, 1 quadrilateral); nodes [E][stride], padded with -1 */
int zj_gmsh_elements(zj_gmsh_handle g, int* kind, int* order, int* n_nodes, int64_t* nodes, int stride);
int zj_gmsh_edge_tag(zj_gmsh_handle g, int64_t a, int64_t b, int* tag);
int zj_gmsh_group(zj_gmsh_handle g, int i, int* tag, int* bc_type, char* name, int name_cap);
/* The file's $Periodic node correspondences as 0-based node indices (slave,
 * master), in file order; *n_pairs gets the count and the arrays, when not
 * null, are filled.  No counterpart: the reference reader skips $Periodic. */
int zj_gmsh_periodic(zj_gmsh_handle g, int64_t* n_pairs, int64_t* slave, int64_t* master);

/* ---- triangles (textbook mode only) ---------------------------------
 * The reference declares triangle elements (ElementType::Triangle,
 * include/common/types.hpp) and builds their operators as zero placeholders
 * (FROperators::initTriangle, src/basis/basis.cpp:615-669); FRSolver assumes
 * quadrilaterals (fr_solver.cpp:138-140).  These entry points run the scheme
 * the reference's settings name on straight-sided triangles: nodal FR with the
 * energy-stable correction field of the configured family (Castonguay,
 * Vincent and Jameson 2012: DG = the L2 lifting; SD = GA and HU with the
 * one-dimensional VCJH c), warp-and-blend solution points, p+1 Gauss flux
 * points per edge, the configured Riemann solver (euler.hpp:204-375), SSP-RK3
 * (include/time/rk_integrator.hpp:33-58).
 * Params: compat = ZJ_COMPAT_TEXTBOOK, corr = ZJ_CORR_*, positivity = 0.
 * State layout [E][Np][4], Np = (p+1)(p+2)/2, solution points in the order
 * zj_tri_operators returns.  Faces come from zj_build_faces (nv = 3) and
 * zj_pair_periodic; normals from zj_face_normals. */
typedef struct zj_tri* zj_tri_handle;
typedef struct {
    int poly_order;              /* 1..5 */
    int64_t n_nodes;
    const double* xy;            /* [n_nodes][2] */
    int n_elem;
    const int64_t* tri;          /* [E][3] corner nodes, counter-clockwise */
    int n_faces;
    const int* face_left_elem;   /* [F] */
    const int* face_left_local;  /* [F] edge k runs from corner k to corner k+1 */
    const int* face_right_elem;  /* [F], -1 on boundary faces */
    const int* face_right_local; /* [F] */
    const double* normals;       /* [F][2] unit normal, outward from the left element */
    const int* bc_type;          /* [F] BCType per face, or NULL when there are no boundary faces */
    const int* bc_face_map;      /* [F] bc_data row of far-field / outflow faces, or NULL */
    const double* bc_data;       /* [n_bc_data][8] */
    int n_bc_data;
    int reorder;                 /* 1: number elements along a Hilbert curve internally */
} zj_tri_desc;
int zj_tri_create(zj_tri_handle* out, const zj_tri_desc* mesh, const zj_params* params);
int zj_tri_destroy(zj_tri_handle h);
const char* zj_tri_last_error(zj_tri_handle h);
int zj_tri_set_state(zj_tri_handle h, const double* U_aos);
int zj_tri_get_state(zj_tri_handle h, double* U_aos);
int zj_tri_set_state_device(zj_tri_handle h, const double* d_U_aos, void* stream);
int zj_tri_get_state_device(zj_tri_handle h, double* d_U_aos, void* stream);
int zj_tri_residual(zj_tri_handle h, double* dUdt_aos, void* stream);  /* host array or NULL */
int zj_tri_residual_device(zj_tri_handle h, double* d_dUdt_aos, void* stream);
int zj_tri_rk3_step(zj_tri_handle h, double dt, void* stream);
/* As zj_time_stages, for a triangle mesh. */
int zj_tri_time_stages(zj_tri_handle h, double dt, int steps, double* stage_ms, void* stream);
/* dt = cfl / max((2p+1)(|u|+c) / d_e + (p+1)^4 nu_k / (rho d_e^2)), d_e the
 * inscribed-circle diameter; nu_k = mu max(4/3, gamma/Pr) with the viscous
 * terms, else 0 */
int zj_tri_dt_cfl(zj_tri_handle h, double cfl, double* dt);
int zj_tri_residual_l2(zj_tri_handle h, double* l2);
int zj_tri_synchronize(zj_tri_handle h);
int zj_tri_set_xcd_groups(zj_tri_handle h, int groups);
int zj_tri_get_info(zj_tri_handle h, int64_t* n_dof, int* elems_per_block, int* n_sp,
                    int64_t* cross_tile_edges);
/* Operators of order p: solution points [Np][2] (r, s), flux points
 * [3(p+1)][2], Gauss weights [p+1], Dr, Ds [Np][Np], interp [3(p+1)][Np],
 * lift [Np][3(p+1)] (before the |e|/(2J) edge factor).  Any pointer may be NULL. */
int zj_tri_operators(int p, double* sp, double* fp, double* wg, double* Dr, double* Ds,
                     double* interp, double* lift);
/* The lifting matrix [Np][3(p+1)] of correction family `corr` (ZJ_CORR_*):
 * the DG lifting filtered by (M + K)^-1 M, K the c-term of the family's
 * energy norm; equal to zj_tri_operators' lift for ZJ_CORR_DG. */
int zj_tri_lift(int p, int corr, double* lift);

/* Discrete adjoint on triangles (no reference counterpart; the quadrilateral
 * zj_adjoint_* protocol): vector-Jacobian products of the triangle stage --
 * the transposed lifting, the configured Riemann solver's and the ghost
 * states' VJPs (euler.hpp:204-375, kernels.cu:453-525), the transposed
 * divergence, interpolation and Euler flux Jacobian, SSP-RK3
 * (rk_integrator.hpp:33-58); on handles with the viscous terms also the
 * BR2 terms backwards (the viscous flux VJP, navierstokes.hpp:38-99, at the
 * points and face points, the corrected gradient's and face liftings' jumps,
 * the transposed gradient), at every order such a handle steps.  Reverse
 * stepping: zj_tri_adjoint_set(dJ/dU^N);
 * for n = N-1 .. 0: zj_tri_set_state(U^n), zj_tri_rk3_step_vjp(dt_n)
 * (stages recomputed from U^n; dt_bar, may be NULL, receives dJ/d(dt_n));
 * then zj_tri_adjoint_get.  Arrays are [E][Np][4]. */
int zj_tri_adjoint_set(zj_tri_handle h, const double* ubar_aos);
int zj_tri_adjoint_set_device(zj_tri_handle h, const double* d_ubar_aos, void* stream);
int zj_tri_adjoint_get(zj_tri_handle h, double* ubar_aos);
int zj_tri_adjoint_get_device(zj_tri_handle h, double* d_ubar_aos, void* stream);
int zj_tri_rk3_step_vjp(zj_tri_handle h, double dt, double* dt_bar, void* stream);
/* (dR/dU)^T w at the handle's state, R = zj_tri_residual's dU/dt. */
int zj_tri_residual_vjp(zj_tri_handle h, const double* w_aos, double* ubar_aos);
int zj_tri_residual_vjp_device(zj_tri_handle h, const double* d_w_aos, double* d_ubar_aos, void* stream);

/* Polynomial order and gamma a handle was created with. */
int zj_get_scheme(zj_handle h, int* poly_order, double* gamma);
int zj_tri_get_scheme(zj_tri_handle h, int* poly_order, double* gamma);

/* ---- restart files ----------------------------------------------------
 * RestartIO (src/io/restart.cpp:8-165): HDF5 file format 0, as libhdf5
 * writes it with default property lists, with root attributes time (f64),
 * iteration (i32), poly_order (i32), gamma (f64) and datasets "solution"
 * (f64, sum(elem_n_sp) * 4 values: element by element, point by point,
 * variable fastest) and "elem_n_sp" (i32, one per element).  The library's
 * own HDF5 codec writes and reads them (no HDF5 library is needed); files
 * libhdf5 wrote in that layout read back the same.  On failure these return
 * ZJ_ERR_ARG (or the handle call's code) with the message in
 * zj_io_last_error(), which is per thread. */
typedef struct {
    int64_t n_elem;              /* length of elem_n_sp */
    int64_t n_values;            /* length of solution */
    double time;
    int iteration;
    int poly_order;
    double gamma;
} zj_restart_info;
const char* zj_io_last_error(void);
/* RestartIO::write (restart.cpp:8-91). */
int zj_restart_write(const char* path, int64_t n_elem, const int* elem_n_sp, const double* solution,
                     double time, int iteration, int poly_order, double gamma);
/* RestartIO::read, attributes and sizes (restart.cpp:93-131). */
int zj_restart_query(const char* path, zj_restart_info* info);
/* RestartIO::read, the arrays (restart.cpp:133-165); n_elem and n_values
 * must match the file; solution may be NULL. */
int zj_restart_read(const char* path, int64_t n_elem, int* elem_n_sp, int64_t n_values, double* solution);
/* A handle's device state to a restart file (FRSolver::copyToHost followed by
 * RestartIO::write), and back: load checks the element count, points per
 * element, poly_order and gamma against the handle, then sets the state (the
 * copy into the solver that the reference's main() leaves out,
 * src/main.cpp:349-360).  time / iteration may be NULL. */
int zj_restart_save(zj_handle h, const char* path, double time, int iteration);
int zj_restart_load(zj_handle h, const char* path, double* time, int* iteration);
int zj_tri_restart_save(zj_tri_handle h, const char* path, double time, int iteration);
int zj_tri_restart_load(zj_tri_handle h, const char* path, double* time, int* iteration);

/* ---- visualisation output ---------------------------------------------
 * VTKWriter (src/io/visualization.cpp:8-386): each element is subdivided
 * into a lattice of `subdivisions` cells per edge (quadrilaterals: (n+1)^2
 * points, 2n^2 triangles; triangles: (n+1)(n+2)/2 points, n^2 triangles),
 * point coordinates from the element's shape functions, point data Density,
 * Pressure, Mach, Velocity.  Default: the reference's text byte for byte,
 * with the state of the element's first solution point at every lattice
 * point (visualization.cpp:327, 360). */
typedef struct {
    int64_t n_nodes;
    const double* xy;                /* [n_nodes][2] */
    int64_t n_elem;
    const int* elem_type;            /* ElementType: 0 triangle, 1 quadrilateral */
    const int* elem_order;           /* geometric order: 1, 2 (6 / 9 nodes), 4 (25-node quads) */
    const int64_t* elem_node_start;  /* [n_elem + 1] offsets into elem_nodes */
    const int64_t* elem_nodes;       /* node ids, Gmsh order */
} zj_vis_mesh;
#define ZJ_VTK_VTU 0                 /* XML .vtu, ascii (VTKWriter::writeVTU) */
#define ZJ_VTK_LEGACY 1              /* legacy .vtk, ascii (VTKWriter::writeLegacy) */
#define ZJ_VTK_BINARY 2              /* .vtu data arrays appended raw (header_type UInt64) */
#define ZJ_VTK_INTERPOLATE 4         /* the solution polynomial at the lattice points (textbook);
                                        elem_n_sp must be (p+1)^2 / (p+1)(p+2)/2 */
int zj_vtk_write(const char* path, const zj_vis_mesh* mesh, const int* elem_n_sp, const double* solution,
                 int poly_order, double gamma, int subdivisions, int flags);
/* VTKWriter::writePVTU (visualization.cpp:247-283). */
int zj_pvtu_write(const char* path, int n_pieces, const char* const* piece_files);

/* ---- measurement (no reference counterpart) -------------------------- */
/* Copies `bytes` bytes (rounded to 16) between two device buffers `reps` times
 * with a plain grid-stride copy kernel and reports the mean GB/s (read + write
 * bytes over kernel time, hipEvents).  Used to calibrate HBM counters and the
 * achievable bandwidth next to the stage kernels' roofline. */
int zj_copy_probe(int device, int64_t bytes, int reps, double* gbps);

/* ---- the gpu:: free functions (include/gpu/kernels.hpp:17-266) ---------
 * Raw device arrays in the reference's layouts: U-like arrays [E][S][4];
 * element-indexed flux-point arrays [E][n_edges][Fp][4]; face-indexed
 * [F][Fp][4]; Jinv [E][S][4]; interp [n_edges][Fp][S]; diff [S][S];
 * corr [n_edges][S]; normals [F][2] (one per face) unless stated.  Each
 * function is asynchronous on `stream` (NULL: the default stream) except
 * zj_gpu_compute_time_step and zj_gpu_residual_norm, which block and return
 * the value through their last pointer, as the reference's return it.  The
 * caller owns every array; nothing is allocated except those two functions'
 * partial-result buffers.  Errors: ZJ_ERR_ARG for sizes or null arrays,
 * ZJ_ERR_HIP for a failed launch (text in zj_last_error(NULL)).
 * The C++ declarations with the reference's names and hipStream_t are in
 * include/zhijian/gpu_kernels.hpp. */
#define ZJ_ENGINE_MFMA 0      /* dense operator products on the FP64 matrix cores (default) */
#define ZJ_ENGINE_VALU 1      /* the same products as FMA chains on the vector ALUs */
/* Engine of zj_gpu_interpolate_to_flux_points, zj_gpu_reference_gradients and
 * zj_gpu_flux_divergence (process-wide). */
int zj_gpu_set_operator_engine(int engine);
/* gpu::interpolateToFluxPoints (kernels.cu:30-67); n_edges * Fp <= 64, S <= 64. */
int zj_gpu_interpolate_to_flux_points(const double* U_sp, double* U_fp, const double* interp, int n_elem,
                                      int n_sp, int n_edges, int n_fp_per_edge, void* stream);
/* gpu::computeReferenceGradients (kernels.cu:73-108). */
int zj_gpu_reference_gradients(const double* U, double* dUdxi, double* dUdeta, const double* diff_xi,
                               const double* diff_eta, int n_elem, int n_sp, void* stream);
/* gpu::transformGradients (kernels.cu:110-151). */
int zj_gpu_transform_gradients(const double* dUdxi, const double* dUdeta, double* dUdx, double* dUdy,
                               const double* Jinv, int n_elem, int n_sp, void* stream);
/* gpu::computeInviscidFluxSP (kernels.cu:157-194). */
int zj_gpu_inviscid_flux_sp(const double* U, double* Fx, double* Fy, double gamma, int n_elem, int n_sp,
                            void* stream);
/* gpu::computeRiemannFlux (kernels.hpp:58-64; the reference's body is empty,
 * kernels.cu:704-714): RiemannSolvers::rusanov/roe/hllc (euler.hpp:184-375,
 * solver_type ZJ_RIEMANN_*) on face-indexed states, normals [F][Fp][2]. */
int zj_gpu_riemann_flux(const double* U_L, const double* U_R, double* F_common, const double* normals,
                        double gamma, int solver_type, int n_faces, int n_fp_per_face, void* stream);
/* gpu::computeRiemannFluxWithBC (kernels.cu:453-701): inline clamped Rusanov
 * (solver_type ignored, as there), U_R from the right slot Fp-1-f or the ghost
 * state of bc_type[face] with bc_data row bc_face_map[face]. */
int zj_gpu_riemann_flux_with_bc(const double* U_fp, double* F_common, const int* face_left_elem,
                                const int* face_left_local, const int* face_right_elem, const int* face_right_local,
                                const int* bc_type, const double* bc_data, const int* bc_face_map,
                                const double* normals, double gamma, int solver_type, int n_faces,
                                int n_fp_per_face, int n_faces_per_elem, void* stream);
/* gpu::computeViscousFluxSP (kernels.cu:720-807): mu = T^0.7 / Re. */
int zj_gpu_viscous_flux_sp(const double* U, const double* dUdx, const double* dUdy, double* Fv_x, double* Fv_y,
                           double gamma, double Re, double Pr, int n_elem, int n_sp, void* stream);
/* gpu::computeInteriorFluxAtFace (kernels.cu:198-257). */
int zj_gpu_interior_flux_at_face(const double* U_fp, double* F_int, const int* face_left_elem,
                                 const int* face_left_local, const double* face_normals, double gamma, int n_faces,
                                 int n_fp_per_face, int n_faces_per_elem, void* stream);
/* gpu::computeInviscidFluxAtFP (kernels.cu:260-319): reference-element edge
 * normals; face_normals and the face arrays are not read. */
int zj_gpu_inviscid_flux_at_fp(const double* U_fp, double* F_fp, const double* face_normals,
                               const int* face_left_elem, const int* face_left_local, double gamma, int n_elem,
                               int n_edges, int n_fp_per_edge, void* stream);
/* gpu::computeFluxDifference (kernels.cu:322-340): F_diff = F_common - F_int. */
int zj_gpu_flux_difference(const double* F_common, const double* F_int, double* F_diff, int n, void* stream);
/* gpu::computeFluxDiffForFR (kernels.cu:346-450); slots of faces not listed
 * are left untouched (the reference's caller zeroes the array first). */
int zj_gpu_flux_diff_for_fr(const double* U_fp, const double* F_common, double* F_diff_elem,
                            const int* face_left_elem, const int* face_left_local, const int* face_right_elem,
                            const int* face_right_local, const double* face_normals, double gamma, int n_faces,
                            int n_fp_per_face, int n_faces_per_elem, void* stream);
/* gpu::computeFluxDivergence (kernels.cu:813-872): div_F = -(physical
 * divergence); J is not read. */
int zj_gpu_flux_divergence(const double* Fx, const double* Fy, double* div_F, const double* diff_xi,
                           const double* diff_eta, const double* Jinv, const double* J, int n_elem, int n_sp,
                           void* stream);
/* gpu::scatterFluxToElements (kernels.cu:875-934). */
int zj_gpu_scatter_flux_to_elements(const double* F_face, double* F_elem, const int* face_left_elem,
                                    const int* face_left_local, const int* face_right_elem,
                                    const int* face_right_local, int n_faces, int n_fp_per_face,
                                    int n_faces_per_elem, void* stream);
/* gpu::applyFRCorrection (kernels.cu:948-1034): quadrilaterals (n_edges 4,
 * n_sp = Fp^2); J is not read. */
int zj_gpu_apply_fr_correction(double* div_F, const double* F_diff, const double* corr_deriv, const double* J,
                               int n_elem, int n_sp, int n_edges, int n_fp_per_edge, void* stream);
/* gpu::applyBoundaryConditions (kernels.cu:1040-1228): characteristic far
 * field, reflecting walls, outflow pressure, total-condition inflow; normals
 * [n_bc_faces][Fp][2] and bc_data [n_bc_faces][8] per boundary face. */
int zj_gpu_apply_boundary_conditions(double* U_fp, const int* elem_idx, const int* local_face, const int* bc_type,
                                     const double* bc_data, const double* normals, double gamma, int n_bc_faces,
                                     int n_fp_per_face, int n_faces_per_elem, void* stream);
/* gpu::enforcePositivity (kernels.cu:1236-1295) on AoS states. */
int zj_gpu_enforce_positivity(double* U, int n_states, void* stream);
/* gpu::computeTimeStep (kernels.cu:1355-1474): min over points of
 * CFL h_min[e] / (maxWaveSpeed(U) sqrt|J|), over the points its per-element
 * tree reduction reaches (all points when n_sp is a power of two). */
int zj_gpu_compute_time_step(const double* U, const double* J, const double* h_min, double CFL, double gamma,
                             int n_elem, int n_sp, double* dt, void* stream);
/* gpu::computeResidualNorm (kernels.cu:1476-1524): sqrt of the sum of squares
 * of the n_elem * n_sp * 4 values, summed in a fixed order. */
int zj_gpu_residual_norm(const double* R, int n_elem, int n_sp, double* norm, void* stream);
/* zj_gpu_residual_norm without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a static device array of 256 KiB per GPU, held until
 * the process ends; at most 67,108,864 values (n_elem * n_sp * 4), larger
 * inputs return ZJ_ERR_ARG.  The result equals zj_gpu_residual_norm's bit for
 * bit (same kernels, same summation order).  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (different streams or threads)
 * race.  Blocks like zj_gpu_residual_norm. */
int zj_gpu_residual_norm_noalloc(const double* R, int n_elem, int n_sp, double* norm, void* stream);
/* gpu::setInitialCondition, copySolution, scaleArray, axpy (kernels.cu:1526-1582). */
int zj_gpu_set_initial_condition(double* U, double rho, double u, double v, double p, double gamma, int n_elem,
                                 int n_sp, void* stream);
int zj_gpu_copy_solution(double* dst, const double* src, int n, void* stream);
int zj_gpu_scale_array(double* dst, const double* src, double alpha, int n, void* stream);
int zj_gpu_axpy(double* y, const double* x, double alpha, int n, void* stream);

/* ---- host mesh and operator tables (native, used by FRSolver) -------- */
int zj_quad_operators(int poly_order, double* sol_xi, double* sol_eta, double* quad_w,
                      double* diff_xi, double* diff_eta, double* interp, double* corr);
/* Faces of n_elem elements with nv corners each (row stride `stride` in
 * elem_nodes).  Arrays hold n_elem*nv entries; *n_faces receives the count. */
int zj_build_faces(int n_elem, int nv, int stride, const int64_t* elem_nodes, int* n_faces,
                   int* face_le, int* face_ll, int* face_re, int* face_rl);
/* Pairs boundary faces one period apart (Lx or Ly; 0 disables a direction): the
 * lower-indexed face is kept in place with the partner as its right side, the
 * partner is removed, the order stays stable.  Updates *n_faces. */
int zj_pair_periodic(int* n_faces, int* face_le, int* face_ll, int* face_re, int* face_rl,
                     int nv, int stride, const int64_t* elem_nodes, const double* xy,
                     double Lx, double Ly, double tol);
/* Pairs boundary faces by node correspondence instead of by fixed periods:
 * n_pairs links slave[i] -> master[i] (node indices, e.g. from
 * zj_gmsh_periodic) join nodes into classes (transitively).  Two boundary
 * faces become one interior face when their end nodes fall in the same two
 * classes and one face is the other moved by a link offset
 * xy[master] - xy[slave] (either sign, within tol).  Kept face, removal and
 * order as zj_pair_periodic.  ZJ_ERR_ARG if a face has two candidates or an
 * index is out of range.  Updates *n_faces. */
int zj_pair_periodic_nodes(int* n_faces, int* face_le, int* face_ll, int* face_re, int* face_rl,
                           int nv, int stride, int n_elem, const int64_t* elem_nodes, const double* xy, int64_t n_nodes,
                           int64_t n_pairs, const int64_t* slave, const int64_t* master, double tol);
int zj_face_normals(int n_faces, const int* face_le, const int* face_ll, int nv, int stride,
                    const int64_t* elem_nodes, const double* xy, double* normals);
int zj_elem_areas(int n_elem, int nv, int stride, const int64_t* elem_nodes,
                  const double* xy, double* area);
/* Quads with 4, 9 or 25 nodes (Gmsh order).  J [E*S], Jinv [E*S*4]; pos
 * [E*S*2] may be NULL. */
int zj_geometry(int n_elem, int node_count, int stride, const int64_t* elem_nodes,
                const double* xy, int S, const double* sol_xi, const double* sol_eta,
                double* J, double* Jinv, double* pos);

#ifdef __cplusplus
}
#endif
#endif
