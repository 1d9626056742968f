// zhijian/gpu_kernels.hpp -- the gpu:: free functions of the reference
// (include/gpu/kernels.hpp:17-266), same names, arguments and layouts, with
// hipStream_t in place of cudaStream_t.  Each one calls the C ABI
// (zj_gpu_* in zhijian_amd.h); a failed call throws std::runtime_error with
// the library's message (the reference does not check its launches).
//
// Layouts: U-like arrays [n_elem][n_sp][N_VARS]; element-indexed flux-point
// arrays [n_elem][n_edges][n_fp][N_VARS]; face-indexed [n_faces][n_fp][N_VARS];
// Jinv [n_elem][n_sp][4]; normals [n_faces][2].
#pragma once

#include "zhijian/types.hpp"
#include "zhijian_amd.h"

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

namespace zhijian {
namespace gpu {

namespace detail {
inline void check(int rc, const char* fn) {
    if (rc != ZJ_OK) throw std::runtime_error(std::string(fn) + ": " + zj_last_error(nullptr));
}
}  // namespace detail

// ---- solution point operations
inline void interpolateToFluxPoints(const Real* U_sp, Real* U_fp, const Real* interp, int n_elem, int n_sp,
                                    int n_edges, int n_fp_per_edge, hipStream_t stream = 0) {
    detail::check(zj_gpu_interpolate_to_flux_points(U_sp, U_fp, interp, n_elem, n_sp, n_edges, n_fp_per_edge, stream),
                  "gpu::interpolateToFluxPoints");
}
inline void computeReferenceGradients(const Real* U, Real* dUdxi, Real* dUdeta, const Real* diff_xi,
                                      const Real* diff_eta, int n_elem, int n_sp, hipStream_t stream = 0) {
    detail::check(zj_gpu_reference_gradients(U, dUdxi, dUdeta, diff_xi, diff_eta, n_elem, n_sp, stream),
                  "gpu::computeReferenceGradients");
}
inline void transformGradients(const Real* dUdxi, const Real* dUdeta, Real* dUdx, Real* dUdy, const Real* Jinv,
                               int n_elem, int n_sp, hipStream_t stream = 0) {
    detail::check(zj_gpu_transform_gradients(dUdxi, dUdeta, dUdx, dUdy, Jinv, n_elem, n_sp, stream),
                  "gpu::transformGradients");
}

// ---- flux computation
inline void computeInviscidFluxSP(const Real* U, Real* Fx, Real* Fy, Real gamma, int n_elem, int n_sp,
                                  hipStream_t stream = 0) {
    detail::check(zj_gpu_inviscid_flux_sp(U, Fx, Fy, gamma, n_elem, n_sp, stream), "gpu::computeInviscidFluxSP");
}
inline void computeRiemannFlux(const Real* U_L, const Real* U_R, Real* F_common, const Real* normals, Real gamma,
                               RiemannSolver solver_type, int n_faces, int n_fp_per_face, hipStream_t stream = 0) {
    detail::check(zj_gpu_riemann_flux(U_L, U_R, F_common, normals, gamma, static_cast<int>(solver_type), n_faces,
                                      n_fp_per_face, stream),
                  "gpu::computeRiemannFlux");
}
inline void computeRiemannFluxWithBC(const Real* U_fp, Real* F_common, const int* face_left_elem,
                                     const int* face_left_local, const int* face_right_elem,
                                     const int* face_right_local, const int* bc_type, const Real* bc_data,
                                     const int* bc_face_map, const Real* normals, Real gamma,
                                     RiemannSolver solver_type, int n_faces, int n_fp_per_face,
                                     int n_faces_per_elem, hipStream_t stream = 0) {
    detail::check(zj_gpu_riemann_flux_with_bc(U_fp, F_common, face_left_elem, face_left_local, face_right_elem,
                                              face_right_local, bc_type, bc_data, bc_face_map, normals, gamma,
                                              static_cast<int>(solver_type), n_faces, n_fp_per_face,
                                              n_faces_per_elem, stream),
                  "gpu::computeRiemannFluxWithBC");
}
inline void computeViscousFluxSP(const Real* U, const Real* dUdx, const Real* dUdy, Real* Fv_x, Real* Fv_y,
                                 Real gamma, Real Re, Real Pr, int n_elem, int n_sp, hipStream_t stream = 0) {
    detail::check(zj_gpu_viscous_flux_sp(U, dUdx, dUdy, Fv_x, Fv_y, gamma, Re, Pr, n_elem, n_sp, stream),
                  "gpu::computeViscousFluxSP");
}
inline void computeInteriorFluxAtFace(const Real* U_fp, Real* F_int, const int* face_left_elem,
                                      const int* face_left_local, const Real* face_normals, Real gamma, int n_faces,
                                      int n_fp_per_face, int n_faces_per_elem, hipStream_t stream = 0) {
    detail::check(zj_gpu_interior_flux_at_face(U_fp, F_int, face_left_elem, face_left_local, face_normals, gamma,
                                               n_faces, n_fp_per_face, n_faces_per_elem, stream),
                  "gpu::computeInteriorFluxAtFace");
}
inline void computeInviscidFluxAtFP(const Real* U_fp, Real* F_fp, const Real* face_normals,
                                    const int* face_left_elem, const int* face_left_local, Real gamma, int n_elem,
                                    int n_edges, int n_fp_per_edge, hipStream_t stream = 0) {
    detail::check(zj_gpu_inviscid_flux_at_fp(U_fp, F_fp, face_normals, face_left_elem, face_left_local, gamma,
                                             n_elem, n_edges, n_fp_per_edge, stream),
                  "gpu::computeInviscidFluxAtFP");
}
inline void computeFluxDifference(const Real* F_common, const Real* F_int, Real* F_diff, int n,
                                  hipStream_t stream = 0) {
    detail::check(zj_gpu_flux_difference(F_common, F_int, F_diff, n, stream), "gpu::computeFluxDifference");
}
inline void computeFluxDiffForFR(const Real* U_fp, const Real* F_common, Real* F_diff_elem,
                                 const int* face_left_elem, const int* face_left_local, const int* face_right_elem,
                                 const int* face_right_local, const Real* face_normals, Real gamma, int n_faces,
                                 int n_fp_per_face, int n_faces_per_elem, hipStream_t stream = 0) {
    detail::check(zj_gpu_flux_diff_for_fr(U_fp, F_common, F_diff_elem, face_left_elem, face_left_local,
                                          face_right_elem, face_right_local, face_normals, gamma, n_faces,
                                          n_fp_per_face, n_faces_per_elem, stream),
                  "gpu::computeFluxDiffForFR");
}
// computeViscousFluxFP (kernels.hpp:140-147) is declared by the reference and
// defined nowhere; it has no counterpart here either.

// ---- divergence and correction
inline void computeFluxDivergence(const Real* Fx, const Real* Fy, Real* div_F, const Real* diff_xi,
                                  const Real* diff_eta, const Real* Jinv, const Real* J, int n_elem, int n_sp,
                                  hipStream_t stream = 0) {
    detail::check(zj_gpu_flux_divergence(Fx, Fy, div_F, diff_xi, diff_eta, Jinv, J, n_elem, n_sp, stream),
                  "gpu::computeFluxDivergence");
}
inline void scatterFluxToElements(const Real* F_face, Real* F_elem, const int* face_left_elem,
                                  const int* face_left_local, const int* face_right_elem,
                                  const int* face_right_local, int n_faces, int n_fp_per_face,
                                  int n_faces_per_elem, hipStream_t stream = 0) {
    detail::check(zj_gpu_scatter_flux_to_elements(F_face, F_elem, face_left_elem, face_left_local, face_right_elem,
                                                  face_right_local, n_faces, n_fp_per_face, n_faces_per_elem,
                                                  stream),
                  "gpu::scatterFluxToElements");
}
inline void applyFRCorrection(Real* div_F, const Real* F_diff, const Real* corr_deriv, const Real* J, int n_elem,
                              int n_sp, int n_edges, int n_fp_per_edge, hipStream_t stream = 0) {
    detail::check(zj_gpu_apply_fr_correction(div_F, F_diff, corr_deriv, J, n_elem, n_sp, n_edges, n_fp_per_edge,
                                             stream),
                  "gpu::applyFRCorrection");
}

// ---- boundary conditions
inline void applyBoundaryConditions(Real* U_fp, const int* elem_idx, const int* local_face, const int* bc_type,
                                    const Real* bc_data, const Real* normals, Real gamma, int n_bc_faces,
                                    int n_fp_per_face, int n_faces_per_elem, hipStream_t stream = 0) {
    detail::check(zj_gpu_apply_boundary_conditions(U_fp, elem_idx, local_face, bc_type, bc_data, normals, gamma,
                                                   n_bc_faces, n_fp_per_face, n_faces_per_elem, stream),
                  "gpu::applyBoundaryConditions");
}

// ---- time integration (SSP-RK3)
inline void rkStage1(Real* U1, const Real* U0, const Real* R, int n, Real dt, hipStream_t stream = 0) {
    detail::check(zj_rk3_combine(1, U1, U0, nullptr, R, n, dt, stream), "gpu::rkStage1");
}
inline void rkStage2(Real* U2, const Real* U0, const Real* U1, const Real* R, int n, Real dt,
                     hipStream_t stream = 0) {
    detail::check(zj_rk3_combine(2, U2, U0, U1, R, n, dt, stream), "gpu::rkStage2");
}
inline void rkStage3(Real* U, const Real* U0, const Real* U2, const Real* R, int n, Real dt,
                     hipStream_t stream = 0) {
    detail::check(zj_rk3_combine(3, U, U0, U2, R, n, dt, stream), "gpu::rkStage3");
}
inline void enforcePositivity(Real* U, int n_states, hipStream_t stream = 0) {
    detail::check(zj_gpu_enforce_positivity(U, n_states, stream), "gpu::enforcePositivity");
}

// ---- utility operations (the two reductions block)
inline Real computeTimeStep(const Real* U, const Real* J, const Real* h_min, Real CFL, Real gamma, int n_elem,
                            int n_sp, hipStream_t stream = 0) {
    Real dt = 0;
    detail::check(zj_gpu_compute_time_step(U, J, h_min, CFL, gamma, n_elem, n_sp, &dt, stream),
                  "gpu::computeTimeStep");
    return dt;
}
inline Real computeResidualNorm(const Real* R, int n_elem, int n_sp, hipStream_t stream = 0) {
    Real norm = 0;
    detail::check(zj_gpu_residual_norm(R, n_elem, n_sp, &norm, stream), "gpu::computeResidualNorm");
    return norm;
}
inline void setInitialCondition(Real* U, Real rho, Real u, Real v, Real p, Real gamma, int n_elem, int n_sp,
                                hipStream_t stream = 0) {
    detail::check(zj_gpu_set_initial_condition(U, rho, u, v, p, gamma, n_elem, n_sp, stream),
                  "gpu::setInitialCondition");
}
inline void copySolution(Real* dst, const Real* src, int n, hipStream_t stream = 0) {
    detail::check(zj_gpu_copy_solution(dst, src, n, stream), "gpu::copySolution");
}
inline void scaleArray(Real* dst, const Real* src, Real alpha, int n, hipStream_t stream = 0) {
    detail::check(zj_gpu_scale_array(dst, src, alpha, n, stream), "gpu::scaleArray");
}
inline void axpy(Real* y, const Real* x, Real alpha, int n, hipStream_t stream = 0) {
    detail::check(zj_gpu_axpy(y, x, alpha, n, stream), "gpu::axpy");
}

}  // namespace gpu
}  // namespace zhijian
