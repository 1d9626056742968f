// zj_gpu_api.hip -- the reference's gpu:: free functions (include/gpu/kernels.hpp:17-266)
// over raw device arrays in the reference's AoS layouts:
//
//   U, dUdt, Fx, Fy, dUdxi ...  [E][S][4]      variable fastest
//   U_fp, F_diff_elem           [E][n_edges][Fp][4]
//   F_common, F_int             [F][Fp][4]     face-indexed
//   Jinv                        [E][S][4]      (dxi/dx, dxi/dy, deta/dx, deta/dy)
//   interp [n_edges][Fp][S], diff_xi/diff_eta [S][S], corr [n_edges][S]
//
// They are the unfused building blocks the reference's FRSolver strings
// together (fr_solver.cpp:335-608); the library's own residual runs fused
// (zj_kernels.hip).  Every function is asynchronous on `stream` (NULL: the
// default stream) except the two reductions, which block and return a host
// value, as the reference's do.
//
// The dense operator products -- interpolation to flux points (K1,
// kernels.cu:30-67), reference gradients (K2, :73-108) and the flux divergence
// (K11, :813-872) -- are batched small GEMMs on the FP64 matrix cores:
// Y_e = A X_e for every element e, with A the (padded) operator staged in LDS
// and four elements' [S][4] blocks side by side as the 16 columns of one
// v_mfma_f64_16x16x4 tile.  A VALU form of the same products is kept for
// measurement (zj_gpu_set_operator_engine).
#include "zhijian_amd.h"
#include "zj_kernels.hpp"

#include "zj_device.hpp"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace zj {
void set_thread_error(const std::string& msg);
}

using namespace zj;

namespace {

std::atomic<int> g_engine{ZJ_ENGINE_MFMA};

using f64x4 = __attribute__((ext_vector_type(4))) double;

constexpr int OPB = 256;        // threads per workgroup of the operator kernels (4 waves)
constexpr int MAX_OP_S = 64;    // largest operator dimension handled (S <= 36 for p <= 5)

enum OpMode : int { OP_INTERP = 0, OP_GRAD = 1, OP_DIV = 2 };

struct OpArgs {
    const double* X0;   // input [E][K][4] (Fx for the divergence)
    const double* X1;   // second input (Fy, divergence only)
    const double* A0;   // operator [M][K]
    const double* A1;   // second operator (diff_eta)
    const double* Jinv; // [E][K][4] (divergence only)
    double* Y0;         // output [E][M][4]
    double* Y1;         // second output (gradients)
    int E, M, K;
};

__host__ __device__ constexpr int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Operators padded with zeros to [Mp][Kp] (Mp a multiple of 16, Kp of 4) in LDS.
__device__ void stage_operator(double* sA, const double* A, int M, int K, int Mp, int Kp) {
    for (int t = threadIdx.x; t < Mp * Kp; t += blockDim.x) {
        const int r = t / Kp, c = t - r * Kp;
        sA[t] = (r < M && c < K) ? A[(int64_t)r * K + c] : 0.0;
    }
}

// Y_e = A X_e for four elements per wave and 16-row tiles of A.  Lane l of the
// wave holds A operand A[16 mt + l%16][4 kk + l/16], B operand X_{e0 + c/4}[4 kk + l/16][c%4]
// with c = l%16, and after the K loop D[16 mt + 4 i + l/16][c], i = 0..3 (the
// FP64 16x16x4 result interleaves its rows over the four lane groups).
template <int MODE>
__global__ __launch_bounds__(OPB) void op_mfma_kernel(const OpArgs a) {
    extern __shared__ double sA[];
    const int Mp = round_up(a.M, 16), Kp = round_up(a.K, 4);
    constexpr int NA = MODE == OP_INTERP ? 1 : 2;
    stage_operator(sA, a.A0, a.M, a.K, Mp, Kp);
    if (NA == 2) stage_operator(sA + Mp * Kp, a.A1, a.M, a.K, Mp, Kp);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, kq = lane >> 4;
    const int ce = c >> 2, cv = c & 3;
    const int waves = gridDim.x * (OPB / 64);
    const int groups = (a.E + 3) / 4;
    for (int grp = blockIdx.x * (OPB / 64) + wave; grp < groups; grp += waves) {
        const int e = grp * 4 + ce;
        const bool live = e < a.E;
        const double* x0 = a.X0 + (int64_t)e * a.K * 4 + cv;
        const double* x1 = MODE == OP_DIV ? a.X1 + (int64_t)e * a.K * 4 + cv : nullptr;
        for (int mt = 0; mt < Mp / 16; ++mt) {
            f64x4 d0 = {0.0, 0.0, 0.0, 0.0}, d1 = d0, d2 = d0, d3 = d0;
            for (int kk = 0; kk < Kp / 4; ++kk) {
                const int k = 4 * kk + kq;
                const bool in = live && k < a.K;
                const double b0 = in ? x0[k * 4] : 0.0;
                const double a0 = sA[(16 * mt + c) * Kp + k];
                d0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, d0, 0, 0, 0);
                if (NA == 2) {
                    const double a1 = sA[Mp * Kp + (16 * mt + c) * Kp + k];
                    d1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, d1, 0, 0, 0);
                }
                if (MODE == OP_DIV) {
                    const double b1 = in ? x1[k * 4] : 0.0;
                    d2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, d2, 0, 0, 0);
                    const double a1 = sA[Mp * Kp + (16 * mt + c) * Kp + k];
                    d3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, d3, 0, 0, 0);
                }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (!live) break;
                const int row = 16 * mt + 4 * i + kq;
                if (row >= a.M) continue;
                const int64_t o = ((int64_t)e * a.M + row) * 4 + cv;
                if (MODE == OP_INTERP) {
                    a.Y0[o] = d0[i];
                } else if (MODE == OP_GRAD) {
                    a.Y0[o] = d0[i];
                    a.Y1[o] = d1[i];
                } else {
                    // kernels.cu:853-857: (dF/dxi xi_x + dF/deta eta_x) + (dG/dxi xi_y + dG/deta eta_y)
                    const double* J = a.Jinv + ((int64_t)e * a.K + row) * 4;
                    const double div = (d0[i] * J[0] + d1[i] * J[2]) + (d2[i] * J[1] + d3[i] * J[3]);
                    a.Y0[o] = -div;
                }
            }
        }
    }
}

// The same products on the VALU: one thread per (element, output row, variable).
template <int MODE>
__global__ __launch_bounds__(OPB) void op_valu_kernel(const OpArgs a) {
    extern __shared__ double sA[];
    const int MK = a.M * a.K;
    constexpr int NA = MODE == OP_INTERP ? 1 : 2;
    for (int t = threadIdx.x; t < MK; t += blockDim.x) {
        sA[t] = a.A0[t];
        if (NA == 2) sA[MK + t] = a.A1[t];
    }
    __syncthreads();
    const int64_t n = (int64_t)a.E * a.M * 4;
    for (int64_t t = (int64_t)blockIdx.x * OPB + threadIdx.x; t < n; t += (int64_t)gridDim.x * OPB) {
        const int v = (int)(t & 3);
        const int64_t er = t >> 2;
        const int e = (int)(er / a.M), row = (int)(er - (int64_t)e * a.M);
        const double* x0 = a.X0 + (int64_t)e * a.K * 4 + v;
        const double* A0 = sA + row * a.K;
        if (MODE == OP_INTERP) {
            double s = 0.0;
            for (int k = 0; k < a.K; ++k) s = fma(A0[k], x0[4 * k], s);
            a.Y0[t] = s;
        } else if (MODE == OP_GRAD) {
            const double* A1 = sA + MK + row * a.K;
            double s0 = 0.0, s1 = 0.0;
            for (int k = 0; k < a.K; ++k) {
                const double xk = x0[4 * k];
                s0 = fma(A0[k], xk, s0);
                s1 = fma(A1[k], xk, s1);
            }
            a.Y0[t] = s0;
            a.Y1[t] = s1;
        } else {
            const double* x1 = a.X1 + (int64_t)e * a.K * 4 + v;
            const double* A1 = sA + MK + row * a.K;
            double fxi = 0.0, feta = 0.0, gxi = 0.0, geta = 0.0;
            for (int k = 0; k < a.K; ++k) {
                const double f = x0[4 * k], g = x1[4 * k];
                fxi = fma(A0[k], f, fxi);
                feta = fma(A1[k], f, feta);
                gxi = fma(A0[k], g, gxi);
                geta = fma(A1[k], g, geta);
            }
            const double* J = a.Jinv + ((int64_t)e * a.K + row) * 4;
            a.Y0[t] = -((fxi * J[0] + feta * J[2]) + (gxi * J[1] + geta * J[3]));
        }
    }
}

int grid_for(int64_t work_items, int per_block) {
    int64_t g = (work_items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > 256 * 16) g = 256 * 16;  // 16 workgroups per CU, the rest by the grid-stride loops
    return (int)g;
}

template <int MODE>
void launch_op(const OpArgs& a, hipStream_t s) {
    constexpr int NA = MODE == OP_INTERP ? 1 : 2;
    if (a.E == 0) return;
    if (g_engine.load() == ZJ_ENGINE_MFMA) {
        const size_t lds = (size_t)NA * round_up(a.M, 16) * round_up(a.K, 4) * sizeof(double);
        const int grid = grid_for((a.E + 3) / 4, OPB / 64);
        klaunch(op_mfma_kernel<MODE>, dim3(grid), dim3(OPB), (unsigned)lds, s, a);
    } else {
        const size_t lds = (size_t)NA * a.M * a.K * sizeof(double);
        const int grid = grid_for((int64_t)a.E * a.M * 4, OPB);
        klaunch(op_valu_kernel<MODE>, dim3(grid), dim3(OPB), (unsigned)lds, s, a);
    }
}

// ------------------------------------------------------------ pointwise
__global__ void transform_gradients_kernel(const double* __restrict__ dxi, const double* __restrict__ deta,
                                           double* __restrict__ dx, double* __restrict__ dy,
                                           const double* __restrict__ Jinv, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double* J = Jinv + (i >> 2) * 4;
        const double a = dxi[i], b = deta[i];
        dx[i] = a * J[0] + b * J[2];
        dy[i] = a * J[1] + b * J[3];
    }
}

__global__ void inviscid_flux_sp_kernel(const double* __restrict__ U, double* __restrict__ Fx,
                                        double* __restrict__ Fy, double g, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double u[NV], fx[NV], fy[NV];
        load4(U + 4 * i, u);
        euler_flux(u, g, fx, fy);
        store4(Fx + 4 * i, fx);
        store4(Fy + 4 * i, fy);
    }
}

// computeViscousFluxSPKernel (kernels.cu:720-793): power-law viscosity
// mu = T^0.7 / Re, kappa = mu gamma / ((gamma - 1) Pr), T = p / rho.
__global__ void viscous_flux_sp_kernel(const double* __restrict__ U, const double* __restrict__ Ux,
                                       const double* __restrict__ Uy, double* __restrict__ Fvx,
                                       double* __restrict__ Fvy, double g, double Re, double Pr, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s[NV], sx[NV], sy[NV];
        load4(U + 4 * i, s);
        load4(Ux + 4 * i, sx);
        load4(Uy + 4 * i, sy);
        const double rho = s[0], u = s[1] / rho, v = s[2] / rho;
        const double dudx = (sx[1] - u * sx[0]) / rho, dudy = (sy[1] - u * sy[0]) / rho;
        const double dvdx = (sx[2] - v * sx[0]) / rho, dvdy = (sy[2] - v * sy[0]) / rho;
        const double p = (g - 1.0) * rho * (s[3] / rho - 0.5 * (u * u + v * v));
        const double T = p / rho;
        const double dpdx = (g - 1.0) * (sx[3] - 0.5 * sx[0] * (u * u + v * v) - rho * (u * dudx + v * dvdx));
        const double dpdy = (g - 1.0) * (sy[3] - 0.5 * sy[0] * (u * u + v * v) - rho * (u * dudy + v * dvdy));
        const double dTdx = (dpdx - T * sx[0]) / rho, dTdy = (dpdy - T * sy[0]) / rho;
        const double mu = pow(T, 0.7) / Re;
        const double k = mu * g / ((g - 1.0) * Pr);
        const double dv = dudx + dvdy, lam = -2.0 / 3.0 * mu;
        const double txx = 2.0 * mu * dudx + lam * dv, tyy = 2.0 * mu * dvdy + lam * dv, txy = mu * (dudy + dvdx);
        const double fx[NV] = {0.0, txx, txy, u * txx + v * txy + k * dTdx};
        const double fy[NV] = {0.0, txy, tyy, u * txy + v * tyy + k * dTdy};
        store4(Fvx + 4 * i, fx);
        store4(Fvy + 4 * i, fy);
    }
}

// ------------------------------------------------------------ face kernels
// One thread per (face, flux point).
struct FaceArgs {
    const double* U_fp;
    const double* F_common;
    double* out;
    const int *le, *ll, *re, *rl;
    const int* bc_type;
    const double* bc_data;
    const int* bc_face_map;
    const double* normals;
    double gamma;
    int n_faces, nfp, nfe;
};

__device__ __forceinline__ int64_t fp_slot(int e, int edge, int f, int nfe, int nfp) {
    return (((int64_t)e * nfe + edge) * nfp + f) * 4;
}

// computeRiemannFluxWithBCKernel (kernels.cu:528-676): U_R from the right
// element's slot nfp-1-f or the ghost state, inline clamped Rusanov whatever
// the configured solver, NaN/Inf -> 0.
__global__ void riemann_with_bc_kernel(const FaceArgs a) {
    const int64_t n = (int64_t)a.n_faces * a.nfp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int face = (int)(t / a.nfp), f = (int)(t - (int64_t)face * a.nfp);
        double UL[NV], UR[NV], F[NV];
        load4(a.U_fp + fp_slot(a.le[face], a.ll[face], f, a.nfe, a.nfp), UL);
        const double nx = a.normals[2 * face], ny = a.normals[2 * face + 1];
        const int re = a.re[face];
        if (re >= 0) {
            load4(a.U_fp + fp_slot(re, a.rl[face], a.nfp - 1 - f, a.nfe, a.nfp), UR);
        } else {
            int bct = a.bc_type[face], bci = a.bc_face_map[face];
            if (bci < 0) { bct = ZJ_BC_SLIPWALL; bci = 0; }
            ghost_state(UL, nx, ny, bct, a.bc_data + (int64_t)bci * 8, a.gamma, UR);
        }
        rusanov(UL, UR, nx, ny, a.gamma, F);
        store4(a.out + 4 * t, F);
    }
}

// computeRiemannFlux (declared kernels.hpp:58-64; the reference's body is
// empty, kernels.cu:704-714): the declared operation, RiemannSolvers::rusanov /
// roe / hllc (euler.hpp:184-375) on face-indexed states [F][Fp][4] with
// normals [F][Fp][2].
__global__ void riemann_face_kernel(const double* __restrict__ UL, const double* __restrict__ UR,
                                    double* __restrict__ F, const double* __restrict__ nrm, double g, int kind,
                                    int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        double l[NV], r[NV], f[NV];
        load4(UL + 4 * t, l);
        load4(UR + 4 * t, r);
        riemann_by_kind(kind, l, r, nrm[2 * t], nrm[2 * t + 1], g, f);
        store4(F + 4 * t, f);
    }
}

// computeInteriorFluxAtFaceKernel (kernels.cu:198-239): F(U_L).n per face point.
__global__ void interior_flux_at_face_kernel(const FaceArgs a) {
    const int64_t n = (int64_t)a.n_faces * a.nfp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int face = (int)(t / a.nfp), f = (int)(t - (int64_t)face * a.nfp);
        double U[NV], Fn[NV];
        load4(a.U_fp + fp_slot(a.le[face], a.ll[face], f, a.nfe, a.nfp), U);
        normal_flux(U, a.gamma, a.normals[2 * face], a.normals[2 * face + 1], Fn);
        store4(a.out + 4 * t, Fn);
    }
}

// computeInviscidFluxAtFPKernel (kernels.cu:260-304): element-indexed, with
// the reference element's edge normals (0,-1), (1,0), (0,1), (-1,0).
__global__ void inviscid_flux_at_fp_kernel(const double* __restrict__ U_fp, double* __restrict__ F_fp, double g,
                                           int n_elem, int n_edges, int nfp) {
    const int64_t n = (int64_t)n_elem * n_edges * nfp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int edge = (int)((t / nfp) % n_edges);
        const double nx = edge == 0 ? 0.0 : edge == 1 ? 1.0 : edge == 2 ? 0.0 : edge == 3 ? -1.0 : 1.0;
        const double ny = edge == 0 ? -1.0 : edge == 2 ? 1.0 : 0.0;
        double U[NV], Fn[NV];
        load4(U_fp + 4 * t, U);
        normal_flux(U, g, nx, ny, Fn);
        store4(F_fp + 4 * t, Fn);
    }
}

// computeFluxDiffForFRKernel (kernels.cu:346-426): left slot f gets
// F_common - F(U_L).n; the right element's slot f (not reversed) gets
// -F_common + F(U_R).n with U_R from that same slot.
__global__ void flux_diff_for_fr_kernel(const FaceArgs a) {
    const int64_t n = (int64_t)a.n_faces * a.nfp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int face = (int)(t / a.nfp), f = (int)(t - (int64_t)face * a.nfp);
        const double nx = a.normals[2 * face], ny = a.normals[2 * face + 1];
        double Fc[NV], U[NV], Fn[NV], D[NV];
        load4(a.F_common + 4 * t, Fc);
        const int64_t lo = fp_slot(a.le[face], a.ll[face], f, a.nfe, a.nfp);
        load4(a.U_fp + lo, U);
        normal_flux(U, a.gamma, nx, ny, Fn);
#pragma unroll
        for (int v = 0; v < NV; ++v) D[v] = Fc[v] - Fn[v];
        store4(a.out + lo, D);
        const int re = a.re[face];
        if (re >= 0) {
            const int64_t ro = fp_slot(re, a.rl[face], f, a.nfe, a.nfp);
            load4(a.U_fp + ro, U);
            normal_flux(U, a.gamma, nx, ny, Fn);
#pragma unroll
            for (int v = 0; v < NV; ++v) D[v] = -Fc[v] + Fn[v];
            store4(a.out + ro, D);
        }
    }
}

// scatterFluxToElementsKernel (kernels.cu:875-916): the face's value at slot f
// of both elements.
__global__ void scatter_flux_kernel(const FaceArgs a) {
    const int64_t n = (int64_t)a.n_faces * a.nfp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int face = (int)(t / a.nfp), f = (int)(t - (int64_t)face * a.nfp);
        double Fc[NV];
        load4(a.F_common + 4 * t, Fc);
        store4(a.out + fp_slot(a.le[face], a.ll[face], f, a.nfe, a.nfp), Fc);
        const int re = a.re[face];
        if (re >= 0) store4(a.out + fp_slot(re, a.rl[face], f, a.nfe, a.nfp), Fc);
    }
}

// applyBoundaryConditionsKernel (kernels.cu:1040-1211): ghost states written
// over the interior flux-point states of boundary faces; normals and bc_data
// are indexed by boundary face (normals per flux point).
__global__ void apply_bc_kernel(double* __restrict__ U_fp, const int* __restrict__ elem_idx,
                                const int* __restrict__ local_face, const int* __restrict__ bc_type,
                                const double* __restrict__ bc_data, const double* __restrict__ normals, double g,
                                int n_bc, int nfp, int nfe) {
    const int64_t n = (int64_t)n_bc * nfp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int bf = (int)(t / nfp), f = (int)(t - (int64_t)bf * nfp);
        double* slot = U_fp + fp_slot(elem_idx[bf], local_face[bf], f, nfe, nfp);
        double U[NV], G[NV];
        load4(slot, U);
        const double nx = normals[2 * t], ny = normals[2 * t + 1];
        const double* bc = bc_data + (int64_t)bf * 8;
        const int type = bc_type[bf];
        const double gm1 = g - 1.0;
        // IdealGas::pressure (euler.hpp:16-23)
        auto pressure = [&](const double* s) {
            const double u = s[1] / s[0], v = s[2] / s[0];
            return gm1 * s[0] * (s[3] / s[0] - 0.5 * (u * u + v * v));
        };
        if (type == ZJ_BC_WALL || type == ZJ_BC_SLIPWALL || type == ZJ_BC_SYMMETRY) {
            const double rho = U[0], u = U[1] / rho, v = U[2] / rho, p = pressure(U);
            const double vn = u * nx + v * ny;
            prim2cons(rho, u - 2.0 * vn * nx, v - 2.0 * vn * ny, p, g, G);
        } else if (type == ZJ_BC_FARFIELD) {
            // Riemann invariants R+ of the interior, R- of the free stream
            const double rinf = bc[0], uinf = bc[1], vinf = bc[2], pinf = bc[3];
            double Uinf[NV];
            prim2cons(rinf, uinf, vinf, pinf, g, Uinf);
            const double cinf = sqrt(g * fmax(pressure(Uinf), 1e-10) / fmax(Uinf[0], 1e-10));
            const double ri = U[0], ui = U[1] / ri, vi = U[2] / ri;
            const double ci = sqrt(g * fmax(pressure(U), 1e-10) / fmax(U[0], 1e-10));
            const double vni = ui * nx + vi * ny, vninf = uinf * nx + vinf * ny;
            const double Rp = vni + 2.0 * ci / gm1, Rm = vninf - 2.0 * cinf / gm1;
            const double vnb = 0.5 * (Rp + Rm), cb = 0.25 * gm1 * (Rp - Rm);
            double rb, pb, ub, vb;
            if (vnb >= 0) {
                const double vti = ui * (-ny) + vi * nx;
                const double si = pressure(U) / pow(ri, g);
                rb = pow(cb * cb / (g * si), 1.0 / gm1);
                pb = rb * cb * cb / g;
                ub = vnb * nx - vti * ny;
                vb = vnb * ny + vti * nx;
            } else {
                const double vtinf = uinf * (-ny) + vinf * nx;
                const double sinf = pinf / pow(rinf, g);
                rb = pow(cb * cb / (g * sinf), 1.0 / gm1);
                pb = rb * cb * cb / g;
                ub = vnb * nx - vtinf * ny;
                vb = vnb * ny + vtinf * nx;
            }
            prim2cons(rb, ub, vb, pb, g, G);
        } else if (type == ZJ_BC_OUTFLOW) {
            const double rho = U[0], u = U[1] / rho, v = U[2] / rho;
            prim2cons(rho, u, v, bc[0], g, G);
        } else if (type == ZJ_BC_INFLOW) {
            // total pressure and temperature, flow direction, static pressure from the interior
            const double pt = bc[0], Tt = bc[1];
            double dx = bc[2], dy = bc[3];
            const double dm = sqrt(dx * dx + dy * dy);
            if (dm > 1e-10) { dx /= dm; dy /= dm; }
            else { dx = -nx; dy = -ny; }
            const double ps = pressure(U);
            const double pr = pt / ps, ex = gm1 / g;
            const double M2 = fmax(2.0 / gm1 * (pow(pr, ex) - 1.0), 0.0);
            const double M = sqrt(M2);
            const double Ts = Tt / (1.0 + 0.5 * gm1 * M2);
            const double rb = ps / Ts;
            const double vel = M * sqrt(g * ps / rb);
            prim2cons(rb, vel * dx, vel * dy, ps, g, G);
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) G[v] = U[v];
        }
        store4(slot, G);
    }
}

// applyFRCorrectionKernel (kernels.cu:948-1021): edges 0 and 2 read slot sp%n1d,
// edges 1 and 3 slot sp/n1d; no metric; subtracted from div_F.
__global__ void fr_correction_kernel(double* __restrict__ div, const double* __restrict__ Fd,
                                     const double* __restrict__ corr, int n_elem, int n_sp, int n_edges, int nfp) {
    const int64_t n = (int64_t)n_elem * n_sp * 4;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int v = (int)(t & 3);
        const int64_t es = t >> 2;
        const int e = (int)(es / n_sp), sp = (int)(es - (int64_t)e * n_sp);
        const int si = sp / nfp, sj = sp % nfp;
        const int64_t b = (int64_t)e * n_edges;
        double c = 0.0;
        c += Fd[((b + 0) * nfp + sj) * 4 + v] * corr[0 * n_sp + sp];
        c += Fd[((b + 1) * nfp + si) * 4 + v] * corr[1 * n_sp + sp];
        c += Fd[((b + 2) * nfp + sj) * 4 + v] * corr[2 * n_sp + sp];
        c += Fd[((b + 3) * nfp + si) * 4 + v] * corr[3 * n_sp + sp];
        div[t] -= c;
    }
}

// ------------------------------------------------------------ state kernels
__global__ void positivity_kernel(double* __restrict__ U, int64_t n_states) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_states;
         i += (int64_t)gridDim.x * blockDim.x) {
        double q[NV];
        load4(U + 4 * i, q);
        positivity<false>(q);
        store4(U + 4 * i, q);
    }
}

__global__ void set_ic_kernel(double* __restrict__ U, double rho, double u, double v, double p, double g,
                              int64_t n) {
    const double rE = p / (g - 1.0) + 0.5 * rho * (u * u + v * v);
    const double q[NV] = {rho, rho * u, rho * v, rE};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        store4(U + 4 * i, q);
}

__global__ void scale_kernel(double* __restrict__ dst, const double* __restrict__ src, double alpha, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = alpha * src[i];
}

__global__ void axpy_kernel(double* __restrict__ y, const double* __restrict__ x, double alpha, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += alpha * x[i];
}

// ------------------------------------------------------------ reductions
// computeTimeStepKernel (kernels.cu:1355-1403) folds the points of an element
// with the tree `for (s = n_sp/2; s > 0; s >>= 1) if (sp < s && sp + s < n_sp)`,
// which leaves some points out whenever n_sp is not a power of two (P2: 9
// points, P4: 25, P5: 36).  `mask` has bit sp set for the points that reach
// the element's minimum; the host derives it by running that tree.
__global__ __launch_bounds__(BLOCK) void dt_gpu_partial_kernel(const double* __restrict__ U,
                                                               const double* __restrict__ J,
                                                               const double* __restrict__ hmin, double cfl,
                                                               double g, int n_elem, int n_sp, uint64_t mask,
                                                               int per_block, double* __restrict__ partial) {
    __shared__ double sRed[BLOCK / 64];
    const int64_t n = (int64_t)n_elem * n_sp;
    const int64_t i0 = (int64_t)blockIdx.x * per_block;
    const int64_t i1 = i0 + per_block < n ? i0 + per_block : n;
    double m = 1e20;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += BLOCK) {
        const int e = (int)(i / n_sp), sp = (int)(i - (int64_t)e * n_sp);
        if (!((mask >> sp) & 1ull)) continue;
        double s[NV];
        load4(U + 4 * i, s);
        // IdealGas::maxWaveSpeed (euler.hpp:166-171) with pressure/soundSpeed (:16-41)
        const double u = s[1] / s[0], v = s[2] / s[0];
        const double p = (g - 1.0) * s[0] * (s[3] / s[0] - 0.5 * (u * u + v * v));
        const double ws = sqrt(u * u + v * v) + sqrt(g * fmax(p, 1e-10) / fmax(s[0], 1e-10));
        m = fmin(m, cfl * hmin[e] / (ws * sqrt(fabs(J[i]))));
    }
    const double r = block_reduce<OpMin>(m, sRed);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// Sum of squares over fixed chunks of `per_block` values, one partial per
// chunk; the partials are then added in index order, so the result does not
// depend on the launch.
__global__ __launch_bounds__(BLOCK) void sumsq_partial_kernel(const double* __restrict__ R, int64_t n,
                                                              int per_block, double* __restrict__ partial) {
    __shared__ double sRed[BLOCK / 64];
    const int64_t i0 = (int64_t)blockIdx.x * per_block;
    const int64_t i1 = i0 + per_block < n ? i0 + per_block : n;
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += BLOCK) s = fma(R[i], R[i], s);
    const double r = block_reduce<OpSum>(s, sRed);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

constexpr int RED_CHUNK = BLOCK * 8;  // values per partial of the reductions

int ew_grid(int64_t n) { return grid_for(n, BLOCK); }

int err(const char* fn, const char* what) {
    set_thread_error(std::string(fn) + ": " + what);
    return ZJ_ERR_ARG;
}

int launched(const char* fn) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return ZJ_OK;
    set_thread_error(std::string(fn) + ": " + hipGetErrorString(e));
    return ZJ_ERR_HIP;
}

bool bad_dims(int n_elem, int n_sp) {
    return n_elem < 0 || n_sp < 1 || n_sp > MAX_OP_S;
}

// Points reached by computeTimeStepKernel's per-element tree (blockDim = n_sp).
uint64_t dt_tree_mask(int n_sp) {
    std::vector<uint64_t> in((size_t)n_sp);
    for (int i = 0; i < n_sp; ++i) in[(size_t)i] = 1ull << i;
    for (int s = n_sp / 2; s > 0; s >>= 1)
        for (int sp = 0; sp < s; ++sp)
            if (sp + s < n_sp) in[(size_t)sp] |= in[(size_t)(sp + s)];
    return in[0];
}

// Shared by the allocating and the allocation-free residual norm: `work` holds
// ceil(n / RED_CHUNK) partials followed by one result slot.
int residual_norm_on(const double* R, int64_t n, double* work, int nchunk, double* norm, hipStream_t s,
                     const char* fn) {
    if (n == 0) {
        *norm = 0.0;
        return ZJ_OK;
    }
    klaunch(sumsq_partial_kernel, dim3(nchunk), dim3(BLOCK), 0, s, R, n, RED_CHUNK, work);
    if (int rc = launched(fn)) return rc;
    if (launch_finalize(work, nchunk, FIN_SQRT_SUM, work + nchunk, s) != hipSuccess) return launched(fn);
    double out = 0.0;
    hipError_t e = hipMemcpyAsync(&out, work + nchunk, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_thread_error(std::string(fn) + ": " + hipGetErrorString(e));
        return ZJ_ERR_HIP;
    }
    *norm = out;
    return ZJ_OK;
}

// ---- allocation-free residual norm
// The partial sums live in a file-scope device array instead of a buffer the
// call allocates.  In host code `scratch` names a host-side stand-in, so the
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) and only that address is passed to the kernels.
constexpr int NOALLOC_MAX_CHUNKS = 32768;  // 256 KiB of partials: up to 32768 * RED_CHUNK values
constexpr int MAX_DEVICES = 64;
__device__ __align__(256) double scratch[NOALLOC_MAX_CHUNKS + 1];

std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

int scratch_of_current_device(double** out, const char* fn) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess && (dev < 0 || dev >= MAX_DEVICES)) e = hipErrorInvalidDevice;
    if (e != hipSuccess) {
        set_thread_error(std::string(fn) + ": " + hipGetErrorString(e));
        return ZJ_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_thread_error(std::string(fn) + ": hipGetSymbolAddress: " + hipGetErrorString(e));
            return ZJ_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ZJ_OK;
}

}  // namespace

// ===================================================================== C ABI
extern "C" int zj_gpu_set_operator_engine(int engine) {
    if (engine != ZJ_ENGINE_MFMA && engine != ZJ_ENGINE_VALU) return err("zj_gpu_set_operator_engine", "unknown engine");
    g_engine.store(engine);
    return ZJ_OK;
}

extern "C" int zj_gpu_interpolate_to_flux_points(const double* U_sp, double* U_fp, const double* interp, int n_elem,
                                                 int n_sp, int n_edges, int n_fp_per_edge, void* stream) {
    static const char* fn = "zj_gpu_interpolate_to_flux_points";
    if (bad_dims(n_elem, n_sp) || n_edges < 1 || n_fp_per_edge < 1 || n_edges * n_fp_per_edge > MAX_OP_S)
        return err(fn, "sizes out of range");
    if (n_elem > 0 && (!U_sp || !U_fp || !interp)) return err(fn, "null array");
    OpArgs a{U_sp, nullptr, interp, nullptr, nullptr, U_fp, nullptr, n_elem, n_edges * n_fp_per_edge, n_sp};
    launch_op<OP_INTERP>(a, (hipStream_t)stream);
    return launched(fn);
}

extern "C" int zj_gpu_reference_gradients(const double* U, double* dUdxi, double* dUdeta, const double* diff_xi,
                                          const double* diff_eta, int n_elem, int n_sp, void* stream) {
    static const char* fn = "zj_gpu_reference_gradients";
    if (bad_dims(n_elem, n_sp)) return err(fn, "sizes out of range");
    if (n_elem > 0 && (!U || !dUdxi || !dUdeta || !diff_xi || !diff_eta)) return err(fn, "null array");
    OpArgs a{U, nullptr, diff_xi, diff_eta, nullptr, dUdxi, dUdeta, n_elem, n_sp, n_sp};
    launch_op<OP_GRAD>(a, (hipStream_t)stream);
    return launched(fn);
}

extern "C" int zj_gpu_transform_gradients(const double* dUdxi, const double* dUdeta, double* dUdx, double* dUdy,
                                          const double* Jinv, int n_elem, int n_sp, void* stream) {
    static const char* fn = "zj_gpu_transform_gradients";
    if (n_elem < 0 || n_sp < 1) return err(fn, "sizes out of range");
    const int64_t n = (int64_t)n_elem * n_sp * 4;
    if (n == 0) return ZJ_OK;
    if (!dUdxi || !dUdeta || !dUdx || !dUdy || !Jinv) return err(fn, "null array");
    klaunch(transform_gradients_kernel, dim3(ew_grid(n)), dim3(BLOCK), 0, (hipStream_t)stream, dUdxi, dUdeta, dUdx,
            dUdy, Jinv, n);
    return launched(fn);
}

extern "C" int zj_gpu_inviscid_flux_sp(const double* U, double* Fx, double* Fy, double gamma, int n_elem, int n_sp,
                                       void* stream) {
    static const char* fn = "zj_gpu_inviscid_flux_sp";
    if (n_elem < 0 || n_sp < 1) return err(fn, "sizes out of range");
    const int64_t n = (int64_t)n_elem * n_sp;
    if (n == 0) return ZJ_OK;
    if (!U || !Fx || !Fy) return err(fn, "null array");
    klaunch(inviscid_flux_sp_kernel, dim3(ew_grid(n)), dim3(BLOCK), 0, (hipStream_t)stream, U, Fx, Fy, gamma, n);
    return launched(fn);
}

extern "C" int zj_gpu_riemann_flux(const double* U_L, const double* U_R, double* F_common, const double* normals,
                                   double gamma, int solver_type, int n_faces, int n_fp_per_face, void* stream) {
    static const char* fn = "zj_gpu_riemann_flux";
    if (n_faces < 0 || n_fp_per_face < 1 || solver_type < ZJ_RIEMANN_RUSANOV || solver_type > ZJ_RIEMANN_HLLC)
        return err(fn, "arguments out of range");
    const int64_t n = (int64_t)n_faces * n_fp_per_face;
    if (n == 0) return ZJ_OK;
    if (!U_L || !U_R || !F_common || !normals) return err(fn, "null array");
    klaunch(riemann_face_kernel, dim3(ew_grid(n)), dim3(BLOCK), 0, (hipStream_t)stream, U_L, U_R, F_common, normals,
            gamma, solver_type, n);
    return launched(fn);
}

extern "C" int zj_gpu_riemann_flux_with_bc(const double* U_fp, double* F_common, const int* face_left_elem,
                                           const int* face_left_local, const int* face_right_elem,
                                           const int* face_right_local, const int* bc_type, const double* bc_data,
                                           const int* bc_face_map, const double* normals, double gamma,
                                           int solver_type, int n_faces, int n_fp_per_face, int n_faces_per_elem,
                                           void* stream) {
    static const char* fn = "zj_gpu_riemann_flux_with_bc";
    (void)solver_type;  // ignored, as kernels.cu:540 ignores it: the flux is the inline Rusanov
    if (n_faces < 0 || n_fp_per_face < 1 || n_faces_per_elem < 1) return err(fn, "sizes out of range");
    if (n_faces == 0) return ZJ_OK;
    if (!U_fp || !F_common || !face_left_elem || !face_left_local || !face_right_elem || !face_right_local ||
        !bc_type || !bc_face_map || !normals)
        return err(fn, "null array");
    FaceArgs a{U_fp, nullptr, F_common, face_left_elem, face_left_local, face_right_elem, face_right_local,
               bc_type, bc_data, bc_face_map, normals, gamma, n_faces, n_fp_per_face, n_faces_per_elem};
    klaunch(riemann_with_bc_kernel, dim3(ew_grid((int64_t)n_faces * n_fp_per_face)), dim3(BLOCK), 0,
            (hipStream_t)stream, a);
    return launched(fn);
}

extern "C" int zj_gpu_viscous_flux_sp(const double* U, const double* dUdx, const double* dUdy, double* Fv_x,
                                      double* Fv_y, double gamma, double Re, double Pr, int n_elem, int n_sp,
                                      void* stream) {
    static const char* fn = "zj_gpu_viscous_flux_sp";
    if (n_elem < 0 || n_sp < 1) return err(fn, "sizes out of range");
    const int64_t n = (int64_t)n_elem * n_sp;
    if (n == 0) return ZJ_OK;
    if (!U || !dUdx || !dUdy || !Fv_x || !Fv_y) return err(fn, "null array");
    klaunch(viscous_flux_sp_kernel, dim3(ew_grid(n)), dim3(BLOCK), 0, (hipStream_t)stream, U, dUdx, dUdy, Fv_x,
            Fv_y, gamma, Re, Pr, n);
    return launched(fn);
}

extern "C" int zj_gpu_interior_flux_at_face(const double* U_fp, double* F_int, const int* face_left_elem,
                                            const int* face_left_local, const double* face_normals, double gamma,
                                            int n_faces, int n_fp_per_face, int n_faces_per_elem, void* stream) {
    static const char* fn = "zj_gpu_interior_flux_at_face";
    if (n_faces < 0 || n_fp_per_face < 1 || n_faces_per_elem < 1) return err(fn, "sizes out of range");
    if (n_faces == 0) return ZJ_OK;
    if (!U_fp || !F_int || !face_left_elem || !face_left_local || !face_normals) return err(fn, "null array");
    FaceArgs a{U_fp, nullptr, F_int, face_left_elem, face_left_local, nullptr, nullptr, nullptr, nullptr, nullptr,
               face_normals, gamma, n_faces, n_fp_per_face, n_faces_per_elem};
    klaunch(interior_flux_at_face_kernel, dim3(ew_grid((int64_t)n_faces * n_fp_per_face)), dim3(BLOCK), 0,
            (hipStream_t)stream, a);
    return launched(fn);
}

extern "C" int zj_gpu_inviscid_flux_at_fp(const double* U_fp, double* F_fp, const double* face_normals,
                                          const int* face_left_elem, const int* face_left_local, double gamma,
                                          int n_elem, int n_edges, int n_fp_per_edge, void* stream) {
    static const char* fn = "zj_gpu_inviscid_flux_at_fp";
    (void)face_normals; (void)face_left_elem; (void)face_left_local;  // unread, as in kernels.cu:260-304
    if (n_elem < 0 || n_edges < 1 || n_fp_per_edge < 1) return err(fn, "sizes out of range");
    const int64_t n = (int64_t)n_elem * n_edges * n_fp_per_edge;
    if (n == 0) return ZJ_OK;
    if (!U_fp || !F_fp) return err(fn, "null array");
    klaunch(inviscid_flux_at_fp_kernel, dim3(ew_grid(n)), dim3(BLOCK), 0, (hipStream_t)stream, U_fp, F_fp, gamma,
            n_elem, n_edges, n_fp_per_edge);
    return launched(fn);
}

extern "C" int zj_gpu_flux_difference(const double* F_common, const double* F_int, double* F_diff, int n,
                                      void* stream) {
    static const char* fn = "zj_gpu_flux_difference";
    if (n < 0) return err(fn, "n < 0");
    if (n == 0) return ZJ_OK;
    if (!F_common || !F_int || !F_diff) return err(fn, "null array");
    // F_diff = F_common + (-1) F_int, which rounds as F_common - F_int
    (void)launch_rk3_combine(1, F_diff, F_common, nullptr, F_int, n, -1.0, (hipStream_t)stream);
    return launched(fn);
}

extern "C" int zj_gpu_flux_diff_for_fr(const double* U_fp, const double* F_common, double* F_diff_elem,
                                       const int* face_left_elem, const int* face_left_local,
                                       const int* face_right_elem, const int* face_right_local,
                                       const double* face_normals, double gamma, int n_faces, int n_fp_per_face,
                                       int n_faces_per_elem, void* stream) {
    static const char* fn = "zj_gpu_flux_diff_for_fr";
    if (n_faces < 0 || n_fp_per_face < 1 || n_faces_per_elem < 1) return err(fn, "sizes out of range");
    if (n_faces == 0) return ZJ_OK;
    if (!U_fp || !F_common || !F_diff_elem || !face_left_elem || !face_left_local || !face_right_elem ||
        !face_right_local || !face_normals)
        return err(fn, "null array");
    FaceArgs a{U_fp, F_common, F_diff_elem, face_left_elem, face_left_local, face_right_elem, face_right_local,
               nullptr, nullptr, nullptr, face_normals, gamma, n_faces, n_fp_per_face, n_faces_per_elem};
    klaunch(flux_diff_for_fr_kernel, dim3(ew_grid((int64_t)n_faces * n_fp_per_face)), dim3(BLOCK), 0,
            (hipStream_t)stream, a);
    return launched(fn);
}

extern "C" int zj_gpu_flux_divergence(const double* Fx, const double* Fy, double* div_F, const double* diff_xi,
                                      const double* diff_eta, const double* Jinv, const double* J, int n_elem,
                                      int n_sp, void* stream) {
    static const char* fn = "zj_gpu_flux_divergence";
    (void)J;  // unread, as in kernels.cu:813-859 (physical divergence)
    if (bad_dims(n_elem, n_sp)) return err(fn, "sizes out of range");
    if (n_elem > 0 && (!Fx || !Fy || !div_F || !diff_xi || !diff_eta || !Jinv)) return err(fn, "null array");
    OpArgs a{Fx, Fy, diff_xi, diff_eta, Jinv, div_F, nullptr, n_elem, n_sp, n_sp};
    launch_op<OP_DIV>(a, (hipStream_t)stream);
    return launched(fn);
}

extern "C" int zj_gpu_scatter_flux_to_elements(const double* F_face, double* F_elem, const int* face_left_elem,
                                               const int* face_left_local, const int* face_right_elem,
                                               const int* face_right_local, int n_faces, int n_fp_per_face,
                                               int n_faces_per_elem, void* stream) {
    static const char* fn = "zj_gpu_scatter_flux_to_elements";
    if (n_faces < 0 || n_fp_per_face < 1 || n_faces_per_elem < 1) return err(fn, "sizes out of range");
    if (n_faces == 0) return ZJ_OK;
    if (!F_face || !F_elem || !face_left_elem || !face_left_local || !face_right_elem || !face_right_local)
        return err(fn, "null array");
    FaceArgs a{nullptr, F_face, F_elem, face_left_elem, face_left_local, face_right_elem, face_right_local,
               nullptr, nullptr, nullptr, nullptr, 0.0, n_faces, n_fp_per_face, n_faces_per_elem};
    klaunch(scatter_flux_kernel, dim3(ew_grid((int64_t)n_faces * n_fp_per_face)), dim3(BLOCK), 0,
            (hipStream_t)stream, a);
    return launched(fn);
}

extern "C" int zj_gpu_apply_fr_correction(double* div_F, const double* F_diff, const double* corr_deriv,
                                          const double* J, int n_elem, int n_sp, int n_edges, int n_fp_per_edge,
                                          void* stream) {
    static const char* fn = "zj_gpu_apply_fr_correction";
    (void)J;  // unread, as in kernels.cu:948-1021 (no metric)
    if (n_elem < 0 || n_sp < 1 || n_edges != 4 || n_fp_per_edge < 1 || n_fp_per_edge * n_fp_per_edge != n_sp)
        return err(fn, "sizes out of range (tensor-product quadrilaterals: n_edges 4, n_sp = n_fp^2)");
    const int64_t n = (int64_t)n_elem * n_sp * 4;
    if (n == 0) return ZJ_OK;
    if (!div_F || !F_diff || !corr_deriv) return err(fn, "null array");
    klaunch(fr_correction_kernel, dim3(ew_grid(n)), dim3(BLOCK), 0, (hipStream_t)stream, div_F, F_diff, corr_deriv,
            n_elem, n_sp, n_edges, n_fp_per_edge);
    return launched(fn);
}

extern "C" int zj_gpu_apply_boundary_conditions(double* U_fp, const int* elem_idx, const int* local_face,
                                                const int* bc_type, const double* bc_data, const double* normals,
                                                double gamma, int n_bc_faces, int n_fp_per_face,
                                                int n_faces_per_elem, void* stream) {
    static const char* fn = "zj_gpu_apply_boundary_conditions";
    if (n_bc_faces < 0 || n_fp_per_face < 1 || n_faces_per_elem < 1) return err(fn, "sizes out of range");
    if (n_bc_faces == 0) return ZJ_OK;
    if (!U_fp || !elem_idx || !local_face || !bc_type || !bc_data || !normals) return err(fn, "null array");
    klaunch(apply_bc_kernel, dim3(ew_grid((int64_t)n_bc_faces * n_fp_per_face)), dim3(BLOCK), 0,
            (hipStream_t)stream, U_fp, elem_idx, local_face, bc_type, bc_data, normals, gamma, n_bc_faces,
            n_fp_per_face, n_faces_per_elem);
    return launched(fn);
}

extern "C" int zj_gpu_enforce_positivity(double* U, int n_states, void* stream) {
    static const char* fn = "zj_gpu_enforce_positivity";
    if (n_states < 0) return err(fn, "n_states < 0");
    if (n_states == 0) return ZJ_OK;
    if (!U) return err(fn, "null array");
    klaunch(positivity_kernel, dim3(ew_grid(n_states)), dim3(BLOCK), 0, (hipStream_t)stream, U, (int64_t)n_states);
    return launched(fn);
}

extern "C" int zj_gpu_compute_time_step(const double* U, const double* J, const double* h_min, double CFL,
                                        double gamma, int n_elem, int n_sp, double* dt, void* stream) {
    static const char* fn = "zj_gpu_compute_time_step";
    if (n_elem < 1 || n_sp < 1 || n_sp > 64 || !dt) return err(fn, "sizes out of range");
    if (!U || !J || !h_min) return err(fn, "null array");
    const int64_t n = (int64_t)n_elem * n_sp;
    const int nchunk = (int)((n + RED_CHUNK - 1) / RED_CHUNK);
    hipStream_t s = (hipStream_t)stream;
    double* work = nullptr;
    hipError_t e = hipMalloc((void**)&work, ((size_t)nchunk + 1) * sizeof(double));
    if (e != hipSuccess) {
        set_thread_error(std::string(fn) + ": " + hipGetErrorString(e));
        return ZJ_ERR_NOMEM;
    }
    klaunch(dt_gpu_partial_kernel, dim3(nchunk), dim3(BLOCK), 0, s, U, J, h_min, CFL, gamma, n_elem, n_sp,
            dt_tree_mask(n_sp), RED_CHUNK, work);
    int rc = launched(fn);
    if (!rc && launch_finalize(work, nchunk, FIN_MIN, work + nchunk, s) != hipSuccess) rc = launched(fn);
    double out = 0.0;
    if (!rc) {
        e = hipMemcpyAsync(&out, work + nchunk, sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            set_thread_error(std::string(fn) + ": " + hipGetErrorString(e));
            rc = ZJ_ERR_HIP;
        }
    }
    (void)hipFree(work);
    if (!rc) *dt = out;
    return rc;
}

extern "C" int zj_gpu_residual_norm(const double* R, int n_elem, int n_sp, double* norm, void* stream) {
    static const char* fn = "zj_gpu_residual_norm";
    if (n_elem < 0 || n_sp < 1 || !norm) return err(fn, "sizes out of range");
    const int64_t n = (int64_t)n_elem * n_sp * 4;
    if (n > 0 && !R) return err(fn, "null array");
    if (n == 0) {
        *norm = 0.0;
        return ZJ_OK;
    }
    const int nchunk = (int)((n + RED_CHUNK - 1) / RED_CHUNK);
    double* work = nullptr;
    const hipError_t e = hipMalloc((void**)&work, ((size_t)nchunk + 1) * sizeof(double));
    if (e != hipSuccess) {
        set_thread_error(std::string(fn) + ": " + hipGetErrorString(e));
        return ZJ_ERR_NOMEM;
    }
    const int rc = residual_norm_on(R, n, work, nchunk, norm, (hipStream_t)stream, fn);
    (void)hipFree(work);
    return rc;
}

extern "C" int zj_gpu_residual_norm_noalloc(const double* R, int n_elem, int n_sp, double* norm, void* stream) {
    static const char* fn = "zj_gpu_residual_norm_noalloc";
    if (n_elem < 0 || n_sp < 1 || !norm) return err(fn, "sizes out of range");
    const int64_t n = (int64_t)n_elem * n_sp * 4;
    if (n > (int64_t)NOALLOC_MAX_CHUNKS * RED_CHUNK)
        return err(fn, "n_elem * n_sp * 4 exceeds the allocation-free maximum of 67108864 values");
    if (n > 0 && !R) return err(fn, "null array");
    if (n == 0) {
        *norm = 0.0;
        return ZJ_OK;
    }
    double* work = nullptr;
    if (int rc = scratch_of_current_device(&work, fn)) return rc;
    const int nchunk = (int)((n + RED_CHUNK - 1) / RED_CHUNK);
    return residual_norm_on(R, n, work, nchunk, norm, (hipStream_t)stream, fn);
}

extern "C" int zj_gpu_set_initial_condition(double* U, double rho, double u, double v, double p, double gamma,
                                            int n_elem, int n_sp, void* stream) {
    static const char* fn = "zj_gpu_set_initial_condition";
    if (n_elem < 0 || n_sp < 1) return err(fn, "sizes out of range");
    const int64_t n = (int64_t)n_elem * n_sp;
    if (n == 0) return ZJ_OK;
    if (!U) return err(fn, "null array");
    klaunch(set_ic_kernel, dim3(ew_grid(n)), dim3(BLOCK), 0, (hipStream_t)stream, U, rho, u, v, p, gamma, n);
    return launched(fn);
}

extern "C" int zj_gpu_copy_solution(double* dst, const double* src, int n, void* stream) {
    static const char* fn = "zj_gpu_copy_solution";
    if (n < 0) return err(fn, "n < 0");
    if (n == 0) return ZJ_OK;
    if (!dst || !src) return err(fn, "null array");
    const hipError_t e = hipMemcpyAsync(dst, src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice,
                                        (hipStream_t)stream);
    if (e != hipSuccess) {
        set_thread_error(std::string(fn) + ": " + hipGetErrorString(e));
        return ZJ_ERR_HIP;
    }
    return ZJ_OK;
}

extern "C" int zj_gpu_scale_array(double* dst, const double* src, double alpha, int n, void* stream) {
    static const char* fn = "zj_gpu_scale_array";
    if (n < 0) return err(fn, "n < 0");
    if (n == 0) return ZJ_OK;
    if (!dst || !src) return err(fn, "null array");
    klaunch(scale_kernel, dim3(ew_grid(n)), dim3(BLOCK), 0, (hipStream_t)stream, dst, src, alpha, (int64_t)n);
    return launched(fn);
}

extern "C" int zj_gpu_axpy(double* y, const double* x, double alpha, int n, void* stream) {
    static const char* fn = "zj_gpu_axpy";
    if (n < 0) return err(fn, "n < 0");
    if (n == 0) return ZJ_OK;
    if (!y || !x) return err(fn, "null array");
    klaunch(axpy_kernel, dim3(ew_grid(n)), dim3(BLOCK), 0, (hipStream_t)stream, y, x, alpha, (int64_t)n);
    return launched(fn);
}
