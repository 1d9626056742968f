/*
 * zj_oracle.h -- CPU oracle for the flux-reconstruction residual and SSP-RK3 step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's executed path (Zhijian: src/gpu/kernels.cu, src/solver/fr_solver.cpp,
 * src/basis/basis.cpp, src/mesh/mesh.cpp, include/physics/euler.hpp).  It is the
 * checker for the HIP product path and the CPU baseline timed next to it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links, loads or calls anything here.
 *
 * Layout conventions are the reference's (kernels.hpp:14-16):
 *   U      [E][S][4]        AoS, variable fastest (rho, rhou, rhov, rhoE)
 *   U_fp   [E][4][Fp][4]    element-indexed flux-point states
 *   F_face [F][Fp][4]       face-indexed common flux
 *   Jinv   [E][S][4]        (dxi/dx, dxi/dy, deta/dx, deta/dy)
 *   diff_xi, diff_eta [S][S]; interp [4][Fp][S]; corr [4][S]
 *
 * Pinning status of every function is stated next to its declaration and
 * checked by tests/test_oracle_*.py.
 */
#ifndef ZJ_ORACLE_H
#define ZJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BCType values (include/common/types.hpp:47-56). */
enum { OZ_BC_INTERIOR = 0, OZ_BC_WALL = 1, OZ_BC_SLIPWALL = 2, OZ_BC_SYMMETRY = 3,
       OZ_BC_FARFIELD = 4, OZ_BC_INFLOW = 5, OZ_BC_OUTFLOW = 6, OZ_BC_PERIODIC = 7 };

/* ---------------------------------------------------------------- operators */
/* GaussLegendre::getQuadrature, basis.cpp:12-104 (n = 1..6 closed forms/table,
 * Newton otherwise).  Pinned: golden vectors from the compiled reference
 * (bit-exact) and sum(w)=2, exactness on x^(2n-1). */
int oz_gauss(int n, double* pts, double* wts);
/* Lagrange1D::eval / evalDeriv, basis.cpp:169-196.  Pinned: delta property,
 * exact derivative of degree<=n-1 polynomials, golden operator tables. */
double oz_lagrange(const double* nodes, int n, int j, double x);
double oz_lagrange_deriv(const double* nodes, int n, int j, double x);
/* CorrectionFunction::evalLeftDeriv/evalRightDeriv, basis.cpp:314-379
 * (FluxType is ignored by the reference: DG/Radau for all types).  Pinned:
 * closed form of the Radau derivative and golden tables. */
double oz_corr_left_deriv(int p, double xi);
double oz_corr_right_deriv(int p, double xi);
/* FROperators::initQuad, basis.cpp:464-613.  Output sizes: sol_xi/sol_eta/qw
 * [S], diff_xi/diff_eta [S*S], interp [4*Fp*S], corr [4*S].  Pinned: golden
 * tables (bit-exact), polynomial exactness and edge orientation. */
void oz_quad_operators(int p, double* sol_xi, double* sol_eta, double* qw,
                       double* diff_xi, double* diff_eta, double* interp,
                       double* corr);

/* ------------------------------------------------------------------- mesh */
/* Mesh::buildFaces, mesh.cpp:18-83: faces in ascending (min vertex, max
 * vertex) order; left = first element (element order), local edge k = (node k,
 * node k+1).  elem_nodes holds the nv corner ids of each element (row stride
 * `stride`).  Returns the face count, or -1 on allocation failure.  Arrays must
 * hold up to E*nv entries.  Pinned: golden face tables of the compiled
 * reference on its test meshes and generated meshes (bit-exact). */
int oz_build_faces(int n_elem, int nv, int stride, const int64_t* elem_nodes,
                   int* face_le, int* face_ll, int* face_re, int* face_rl);
/* Periodic pairing (the reference has none; the library uses the same convention): the
 * boundary face with the lower index is kept in place and gets the partner's
 * left element/edge as its right side; the partner is removed; order stays
 * stable.  Partners are boundary faces whose midpoints differ by (+-Lx, 0) or
 * (0, +-Ly) within tol.  A period of 0 disables that direction.  Returns the new
 * face count or -1 if some boundary face has several candidate partners.
 * Pinned: closed algorithm; tests check face counts, involution and geometry. */
int oz_pair_periodic(int n_faces, int* face_le, int* face_ll, int* face_re,
                     int* face_rl, int nv, int stride, const int64_t* elem_nodes,
                     const double* xy, double Lx, double Ly, double tol);
/* Mesh::faceNormal (mesh.cpp:159-183) per face, normals [F][2].  Pinned:
 * golden normals of the compiled reference (bit-exact). */
void oz_face_normals(int n_faces, const int* face_le, const int* face_ll, int nv,
                     int stride, const int64_t* elem_nodes, const double* xy,
                     double* normals);
/* Mesh::elementArea (mesh.cpp:122-146).  Pinned: golden areas (bit-exact). */
void oz_elem_areas(int n_elem, int nv, int stride, const int64_t* elem_nodes,
                   const double* xy, double* area);
/* FRSolver::initGeometry (fr_solver.cpp:85-129) with ElementGeometry::jacobian
 * (basis.cpp:775-863) for quads of geometric order 1, 2 or 4 (node_count 4, 9,
 * 25 in Gmsh order).  J [E][S], Jinv [E][S][4].  Pinned: bilinear closed form
 * and golden Jacobians of the compiled reference. */
int oz_geometry(int n_elem, int node_count, int stride, const int64_t* elem_nodes,
                const double* xy, int S, const double* sol_xi,
                const double* sol_eta, double* J, double* Jinv);
/* ElementGeometry::refToPhys (basis.cpp:882-893) for the same element types;
 * pos [E][S][2]. */
int oz_sp_coords(int n_elem, int node_count, int stride, const int64_t* elem_nodes,
                 const double* xy, int S, const double* sol_xi,
                 const double* sol_eta, double* pos);

/* --------------------------------------------------------------- residual */
typedef struct oz_solver oz_solver;

/* Everything FRSolver::allocateGPUMemory/initGeometry upload, plus the per
 * residual BC tables built at fr_solver.cpp:432-494 (bc_type/bc_face_map per
 * face, 8 doubles of bc_data per boundary entry).  Arrays are copied. */
typedef struct {
    int p;              /* polynomial order (1..5) */
    int n_elem;
    int n_faces;
    double gamma;
    const double* diff_xi; const double* diff_eta;
    const double* interp;  const double* corr;
    const double* J;       const double* Jinv;
    const double* normals;                     /* [F][2] */
    const int* face_le; const int* face_ll;
    const int* face_re; const int* face_rl;
    const int* bc_type;                        /* [F], may be NULL (all interior) */
    const int* bc_face_map;                    /* [F], may be NULL */
    const double* bc_data; int n_bc_data;      /* [n_bc_data][8] */
    const double* elem_area;                   /* [E] for the CFL dt (A19) */
} oz_desc;

oz_solver* oz_create(const oz_desc* d);
void oz_destroy(oz_solver* s);
void oz_set_threads(int n);   /* OpenMP threads for the loops below (1 = serial) */
int oz_threads(void);

/* Individual kernels, each a literal restatement of one reference kernel.
 * They read/write the solver's internal buffers; accessors below expose them. */
/* computeInviscidFluxSPKernel kernels.cu:157-194.  Pinned: closed-form Euler
 * flux (golden IdealGas::fluxX/fluxY vectors, bit-exact). */
void oz_k_flux_sp(oz_solver* s, const double* U);
/* interpolateToFluxPointsKernel kernels.cu:30-67.  Pinned: polynomial
 * exactness at the edge flux points, including the reversed edges 2 and 3. */
void oz_k_interp(oz_solver* s, const double* U);
/* computeRiemannFluxWithBCKernel + computeGhostState kernels.cu:453-701
 * (inline clamped Rusanov; solver_type ignored).  Pinned: golden
 * RiemannSolvers::rusanov vectors of the compiled reference (<=1e-13 rel),
 * consistency F(U,U,n)=F(U).n, conservation under (L,R,n)->(R,L,-n), and
 * hand-derived ghost-state vectors. */
void oz_k_riemann(oz_solver* s);
/* computeFluxDiffForFRKernel kernels.cu:346-450 after the memset at
 * fr_solver.cpp:501.  Pinned: impulse test of slot placement (same fp on the
 * right element, not reversed) and freestream zero. */
void oz_k_flux_diff(oz_solver* s);
/* computeFluxDivergenceKernel kernels.cu:813-872.  Pinned: exact divergence
 * of polynomial fluxes on affine elements. */
void oz_k_divergence(oz_solver* s, double* dUdt);
/* applyFRCorrectionKernel kernels.cu:948-1034.  Pinned: impulse test written
 * from kernels.cu:971-1007 (edge 0/2 use the xi index sp%Fp, edges 1/3 the
 * eta index sp/Fp, no mirroring, no metric, subtracted). */
void oz_k_correction(oz_solver* s, double* dUdt);

/* The residual of FRSolver::computeInviscidFlux_GPU (fr_solver.cpp:351-608)
 * without the NaN scans: K4, K1, K9, K8, K11, K13 in that order.  Pinned as a
 * whole by freestream preservation and by the kernels above. */
void oz_residual(oz_solver* s, const double* U, double* dUdt);

/* SSP-RK3 stages (kernels.cu:1297-1349) and the positivity limiter
 * (kernels.cu:1236-1295).  n = number of doubles.  Pinned: closed form of the
 * SSP-RK3 amplification polynomial and hand-derived limiter vectors. */
void oz_rk_stage1(double* U1, const double* U0, const double* R, int n, double dt);
void oz_rk_stage2(double* U2, const double* U0, const double* U1, const double* R,
                  int n, double dt);
void oz_rk_stage3(double* U, const double* U0, const double* U2, const double* R,
                  int n, double dt);
void oz_positivity(double* U, int n_states);

/* FRSolver::advance (fr_solver.cpp:305-333) on U in place; positivity 0/1. */
void oz_advance(oz_solver* s, double* U, double dt, int positivity);

/* FRSolver::computeTimeStep (fr_solver.cpp:784-816) evaluated on U.  Pinned:
 * brute-force min (order independent, bit-exact). */
double oz_dt_cfl(oz_solver* s, const double* U, double cfl);
/* FRSolver::computeResidual (fr_solver.cpp:818-833): sqrt of the sequential sum
 * of squares of the last residual.  Pinned: math.fsum within 1e-14. */
double oz_residual_l2(oz_solver* s);

/* Access to the intermediate buffers (owned by the solver). */
const double* oz_buf_fx(oz_solver* s);
const double* oz_buf_fy(oz_solver* s);
const double* oz_buf_ufp(oz_solver* s);
const double* oz_buf_fface(oz_solver* s);
const double* oz_buf_fdiff(oz_solver* s);
double* oz_buf_ufp_mut(oz_solver* s);
double* oz_buf_fdiff_mut(oz_solver* s);
const double* oz_buf_dudt(oz_solver* s);

/* Single-point physics, exposed for known-answer tests. */
void oz_flux_xy(const double* U, double gamma, double* Fx, double* Fy);
void oz_rusanov_inline(const double* UL, const double* UR, double nx, double ny,
                       double gamma, double* F);
void oz_ghost_state(const double* Uint, double nx, double ny, int bc_type,
                    const double* bc_data, int bc_idx, double gamma, double* ghost);

/* ---- textbook scheme (zj_textbook.c): what the reference's FluxType and
 * RiemannSolver settings name but its executed path does not run. */
enum { OZ_RIEMANN_RUSANOV = 0, OZ_RIEMANN_ROE = 1, OZ_RIEMANN_HLLC = 2 };
enum { OZ_CORR_DG = 0, OZ_CORR_SD = 1, OZ_CORR_HU = 2, OZ_CORR_GA = 3 };
/* IdealGas::fluxNormal and maxWaveSpeed (euler.hpp:133-147, 166-171). */
void oz_flux_normal(const double* U, double nx, double ny, double gamma, double* Fn);
double oz_max_wave_speed(const double* U, double gamma);
/* RiemannSolvers::rusanov / roe / hllc (euler.hpp:184-375).  Pinned: the
 * compiled reference's golden vectors (tests/golden/physics.npz). */
void oz_riemann(int kind, const double* UL, const double* UR, double nx, double ny, double gamma,
                double* F);
/* VCJH correction-function derivatives gL'(x), gR'(x) at n points.  Pinned:
 * DG = the reference's Radau tables; HU = Huynh's g2; SD/GA by their zeros. */
int oz_vcjh_deriv(int p, int kind, const double* x, int n, double* gLd, double* gRd);
/* oz_riemann over n point pairs ([n][4] states, [n][2] normals). */
void oz_riemann_batch(int kind, int n, const double* UL, const double* UR, const double* nrm, double gamma,
                      double* F);
typedef struct oz_tb oz_tb;
/* edge_half_len: [E][4] half lengths of the straight element edges. */
oz_tb* oz_tb_create(const oz_desc* d, const double* edge_half_len, int riemann, int corr);
void oz_tb_destroy(oz_tb* t);
void oz_tb_residual(oz_tb* t, const double* U, double* dUdt);
void oz_tb_advance(oz_tb* t, double* U, double dt, int positivity);
/* Viscous terms (BR2) of the textbook residual: constant mu, kappa, BR2
 * penalty eta (affine quadrilaterals; boundary faces carry the interior
 * viscous flux).  oz_visc_flux: the NS viscous flux pair at (U, dU/dx, dU/dy).
 * oz_tb_gradient: the corrected gradient [E*S][8] (d/dx of the 4 variables,
 * then d/dy). */
void oz_visc_flux(const double* U, const double* Qx, const double* Qy, double gamma, double mu,
                  double kappa, double* Fv, double* Gv);
void oz_tb_set_viscous(oz_tb* t, double mu, double kappa, double eta);
void oz_tb_gradient(oz_tb* t, const double* U, double* Q);

#ifdef __cplusplus
}
#endif
#endif
