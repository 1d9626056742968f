// zj_visc.hip -- the textbook stage kernel with the compressible Navier-Stokes
// viscous terms (BR2), for affine quadrilaterals without boundary faces.
//
// The reference computes viscous gradients and fluxes and then discards them
// (fr_solver.cpp:645-646), so in its executed arithmetic the viscous terms
// contribute nothing (reference mode keeps that).  Textbook mode with
// viscous terms runs the scheme its settings name (SimParams::viscous, Re,
// Pr; BR2Scheme in include/physics/navierstokes.hpp:173-204):
//
//   U^C = (U_L + U_R)/2 at every face point; jump J = U^C - U_own
//   Q = grad U + corrections (VCJH g', per reference direction, mapped by the
//       affine metric) -- the FR auxiliary gradient
//   total flux F = F_inv(U) - F_v(U, Q)  (F_v as ViscousFlux::fluxX/fluxY,
//       navierstokes.hpp:53-99, with mu = 1/Re, kappa = mu gamma/((gamma-1) Pr))
//   common normal flux F^I = Riemann(U_L, U_R) - F_v(Ubar, Qbar) . n, where
//       Qbar averages each side's BR2 gradient grad U_own + eta r_f and r_f,
//       the lifting of this face's jump alone, is g'(+-1) J along the edge's
//       reference direction at the face's own points
//   then the textbook correction and SSP-RK3 update of the inviscid kernel.
//
// Traces carry U and grad U per flux point ([E][4][p+1][12]), written for the
// edges another tile reads; a neighbour in the same tile is read from LDS.
#include "zj_kernels.hpp"

#include "zj_device.hpp"

#include <algorithm>
#include <type_traits>

namespace zj {

namespace {

constexpr int TW = 3 * NV;  // trace words per flux point: U, dU/dx, dU/dy

// ViscousFlux::velocityGradients / temperatureGradient / stressTensor /
// fluxX / fluxY in one pass, mu and kappa given.
template <typename R>
__device__ __forceinline__ void visc_flux(const R* U, const R* Qx, const R* Qy, R g, R mu, R kappa, R* Fv, R* Gv) {
    const R rho = U[0], ir = rcpR(rho), u = U[1] * ir, v = U[2] * ir;
    const R dudx = (Qx[1] - u * Qx[0]) * ir, dudy = (Qy[1] - u * Qy[0]) * ir;
    const R dvdx = (Qx[2] - v * Qx[0]) * ir, dvdy = (Qy[2] - v * Qy[0]) * ir;
    const R q2 = u * u + v * v;
    const R p = (g - R(1)) * rho * (U[3] * ir - R(0.5) * q2);
    const R dpdx = (g - R(1)) * (Qx[3] - R(0.5) * Qx[0] * q2 - rho * (u * dudx + v * dvdx));
    const R dpdy = (g - R(1)) * (Qy[3] - R(0.5) * Qy[0] * q2 - rho * (u * dudy + v * dvdy));
    const R T = p * ir;
    const R dTdx = (dpdx - T * Qx[0]) * ir, dTdy = (dpdy - T * Qy[0]) * ir;
    const R div = dudx + dvdy, lam = R(-2.0 / 3.0) * mu;
    const R txx = R(2) * mu * dudx + lam * div, tyy = R(2) * mu * dvdy + lam * div, txy = mu * (dudy + dvdx);
    Fv[0] = R(0); Fv[1] = txx; Fv[2] = txy; Fv[3] = u * txx + v * txy + kappa * dTdx;
    Gv[0] = R(0); Gv[1] = txy; Gv[2] = tyy; Gv[3] = u * txy + v * tyy + kappa * dTdy;
}

// A tile's arrays.  G holds the gradient (x then y, each [var][point]), then
// the contravariant total fluxes in the same slots; J holds the jumps, then
// the flux differences.
template <typename R, int P, int EPB>
struct ViscSmem {
    using Y = Lay<R, P>;
    static constexpr int N1 = P + 1, S = N1 * N1;
    __device__ static R* U() { __shared__ R a[EPB * Y::UE]; return a; }
    __device__ static R* Gx() { __shared__ R a[EPB * Y::UE]; return a; }
    __device__ static R* Gy() { __shared__ R a[EPB * Y::UE]; return a; }
    __device__ static R* Tr() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* J() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* FI() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* D() { __shared__ R a[N1 * N1]; return a; }
    __device__ static R* Lm() { __shared__ R a[N1]; return a; }
    __device__ static R* Lp() { __shared__ R a[N1]; return a; }
    __device__ static R* C() { __shared__ R a[2 * N1]; return a; }
    __device__ static double* Red() { __shared__ double a[BLOCK / 64]; return a; }
    __device__ static unsigned char* Out() { __shared__ unsigned char a[EPB * 4]; return a; }
};

// gradient of one variable at point (i, j) of an element (u[S], xi index
// fastest), physical through the affine metric
template <typename R, int P>
__device__ __forceinline__ void point_gradient(const R* u, const R* D, int i, int j, R jx, R jy, R kx, R ky, R& gx,
                                               R& gy) {
    constexpr int N1 = P + 1;
    R dxi = R(0), deta = R(0);
#pragma unroll
    for (int m = 0; m < N1; ++m) {
        dxi = fmaR(D[i * N1 + m], u[j * N1 + m], dxi);
        deta = fmaR(D[j * N1 + m], u[m * N1 + i], deta);
    }
    gx = jx * dxi + kx * deta;
    gy = jy * dxi + ky * deta;
}

template <typename R, int P, int EPB, int MODE>
__global__ __launch_bounds__(BLOCK) void visc_stage_kernel(const StageArgs<R> a, const OpTables<R> ops,
                                                           const ViscParams vp) {
    using Y = Lay<R, P>;
    using SM = ViscSmem<R, P, EPB>;
    constexpr bool INC = std::is_same<R, float>::value;
    constexpr bool HAS_U0 = MODE == MODE_STAGE2 || MODE == MODE_STAGE3;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1;
    constexpr int NT = EPB * S, NF = EPB * 4 * FP;
    static_assert(NT <= BLOCK, "tile exceeds the workgroup");
    R* const sU = SM::U();
    R* const sGx = SM::Gx();
    R* const sGy = SM::Gy();
    R* const sTr = SM::Tr();
    R* const sJ = SM::J();
    R* const sFI = SM::FI();
    R* const sD = SM::D();
    R* const sLm = SM::Lm();
    R* const sLp = SM::Lp();
    R* const sC = SM::C();
    double* const sRed = SM::Red();
    unsigned char* const sOut = SM::Out();

    const int tid = threadIdx.x;
    const int tile = tile_of_block(a.xcd_groups);
    const int e0 = tile * EPB;
    const int64_t PS = (int64_t)a.E * S;
    const R g = R(a.gamma), mu = R(vp.mu), kappa = R(vp.kappa), eta = R(vp.eta);
    const R gLe = R(vp.gLe), gRe = R(vp.gRe);

    for (int t = tid; t < N1 * N1; t += BLOCK) sD[t] = ops.D[t];
    if (tid < N1) { sLm[tid] = ops.Lm[tid]; sLp[tid] = ops.Lp[tid]; sC[tid] = ops.gL[tid]; sC[N1 + tid] = ops.gR[tid]; }

    const int el = tid / S, sp = tid - (tid / S) * S;
    const int e = e0 + el;
    const bool active = tid < NT && e < a.E;
    const int64_t gi = (int64_t)e * S + sp;
    const int ub = el * Y::UE + sp;
    const int i = sp % N1, j = sp / N1;

    // ---- A: state (read again for the update in phase E, from cache, rather
    // than held in registers through the face phases)
    auto stage_input = [&](int v) -> double {
        return (INC && HAS_U0) ? a.U0[v * PS + gi] + double(a.dW_in[v * PS + gi]) : a.U_in[v * PS + gi];
    };
    R jx = R(0), jy = R(0), kx = R(0), ky = R(0);
    if (active) {
        jx = a.Jinv[(int64_t)e * 4]; jy = a.Jinv[(int64_t)e * 4 + 1];
        kx = a.Jinv[(int64_t)e * 4 + 2]; ky = a.Jinv[(int64_t)e * 4 + 3];
#pragma unroll
        for (int v = 0; v < NV; ++v) sU[ub + v * Y::UV] = R(stage_input(v));
    }
    __syncthreads();

    // ---- B1: gradient at the solution points; own traces
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
            point_gradient<R, P>(sU + el * Y::UE + v * Y::UV, sD, i, j, jx, jy, kx, ky, sGx[ub + v * Y::UV],
                                 sGy[ub + v * Y::UV]);
    }
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        if (e0 + el2 < a.E) {
            const TraceLine tl = trace_line<P>(k, f);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                sTr[Y::tr(el2, k, f, v)] = trace_point<R, P>(sU + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
        }
    }
    __syncthreads();

    // ---- B2a: jumps and the common viscous normal flux per (edge, flux point)
    // (kept apart from the Riemann solve so that the two never hold their
    // registers at once)
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        const R nx = a.normals[((int64_t)e2 * 4 + k) * 2], ny = a.normals[((int64_t)e2 * 4 + k) * 2 + 1];
        const int nl = L.nbr - e0;
        const bool inside = nl >= 0 && nl < EPB;
        if (f == 0) sOut[el2 * 4 + k] = !inside;
        const int oslot = P - f;
        const TraceLine tl = trace_line<P>(k, f);
        const TraceLine tn = trace_line<P>(L.nbr_edge, oslot);
        const R* T = a.T_in + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + oslot) * TW;
        // BR2 gradients: own and neighbour, each with its own face's lifting
        // (this item's element e2 is not the thread's own element)
        const R* Jo = a.Jinv + (int64_t)e2 * 4;
        const R* Jn = a.Jinv + (int64_t)L.nbr * 4;
        const bool oxi = (k & 1) != 0, nxi = (L.nbr_edge & 1) != 0;
        const R go = eta * ((k == 0 || k == 3) ? gLe : gRe);
        const R gn = eta * ((L.nbr_edge == 0 || L.nbr_edge == 3) ? gLe : gRe);
        const R ax = go * (oxi ? Jo[0] : Jo[2]), ay = go * (oxi ? Jo[1] : Jo[3]);
        const R bx = gn * (nxi ? Jn[0] : Jn[2]), by = gn * (nxi ? Jn[1] : Jn[3]);
        R qx[NV], qy[NV], Um[NV], Fv[NV], Gv[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R own = sTr[Y::tr(el2, k, f, v)];
            R oth, gxn, gyn;
            if (inside) {
                oth = sTr[Y::tr(nl, L.nbr_edge, oslot, v)];
                gxn = trace_point<R, P>(sGx + nl * Y::UE + v * Y::UV, tn, sLm, sLp);
                gyn = trace_point<R, P>(sGy + nl * Y::UE + v * Y::UV, tn, sLm, sLp);
            } else {
                oth = T[v];
                gxn = T[NV + v];
                gyn = T[2 * NV + v];
            }
            const R gxo = trace_point<R, P>(sGx + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
            const R gyo = trace_point<R, P>(sGy + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
            const R uc = R(0.5) * (own + oth);
            const R jo = uc - own, jn = uc - oth;
            sJ[Y::tr(el2, k, f, v)] = jo;
            qx[v] = R(0.5) * ((gxo + ax * jo) + (gxn + bx * jn));
            qy[v] = R(0.5) * ((gyo + ay * jo) + (gyn + by * jn));
            Um[v] = uc;
        }
        visc_flux(Um, qx, qy, g, mu, kappa, Fv, Gv);
#pragma unroll
        for (int v = 0; v < NV; ++v) sFI[Y::tr(el2, k, f, v)] = Fv[v] * nx + Gv[v] * ny;  // as the left side sees it
    }
    __syncthreads();

    // ---- B2b: Riemann flux; total common normal flux, outward and scaled
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        const R nx = a.normals[((int64_t)e2 * 4 + k) * 2], ny = a.normals[((int64_t)e2 * 4 + k) * 2 + 1];
        const int nl = L.nbr - e0;
        const int oslot = P - f;
        R own[NV], oth[NV], UL[NV], UR[NV], Fc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) own[v] = sTr[Y::tr(el2, k, f, v)];
        if (nl >= 0 && nl < EPB) {
#pragma unroll
            for (int v = 0; v < NV; ++v) oth[v] = sTr[Y::tr(nl, L.nbr_edge, oslot, v)];
        } else {
            load4(a.T_in + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + oslot) * TW, oth);
        }
        const bool right = L.role == ROLE_RIGHT;
#pragma unroll
        for (int v = 0; v < NV; ++v) { UL[v] = right ? oth[v] : own[v]; UR[v] = right ? own[v] : oth[v]; }
        riemann_by_kind(a.riemann, UL, UR, nx, ny, g, Fc);
        const R ef = a.edge_f[(int64_t)e2 * 4 + k];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R fc = Fc[v] - sFI[Y::tr(el2, k, f, v)];  // as the face's left side sees it
            sFI[Y::tr(el2, k, f, v)] = ef * (right ? -fc : fc);
        }
    }
    __syncthreads();

    // ---- C: corrected gradient, viscous and inviscid fluxes at the points
    if (active) {
        R Uq[NV], Qx[NV], Qy[NV], fx[NV], fy[NV], Fv[NV], Gv[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Uq[v] = sU[ub + v * Y::UV];
            const R cxi = sC[i] * sJ[Y::tr(el, 3, P - j, v)] + sC[N1 + i] * sJ[Y::tr(el, 1, j, v)];
            const R ceta = sC[j] * sJ[Y::tr(el, 0, i, v)] + sC[N1 + j] * sJ[Y::tr(el, 2, P - i, v)];
            Qx[v] = sGx[ub + v * Y::UV] + jx * cxi + kx * ceta;
            Qy[v] = sGy[ub + v * Y::UV] + jy * cxi + ky * ceta;
        }
        euler_flux(Uq, g, fx, fy);
        visc_flux(Uq, Qx, Qy, g, mu, kappa, Fv, Gv);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R Fx = fx[v] - Fv[v], Fy = fy[v] - Gv[v];
            sGx[ub + v * Y::UV] = jx * Fx + jy * Fy;  // contravariant total fluxes replace the gradient
            sGy[ub + v * Y::UV] = kx * Fx + ky * Fy;
        }
    }
    __syncthreads();

    // ---- D: flux differences (into the jump slots)
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        if (e0 + el2 >= a.E) continue;
        const TraceLine tl = trace_line<P>(k, f);
        const R* Fb = ((k & 1) ? sGx : sGy) + el2 * Y::UE;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            sJ[Y::tr(el2, k, f, v)] = sFI[Y::tr(el2, k, f, v)] - trace_point<R, P>(Fb + v * Y::UV, tl, sLm, sLp);
    }
    __syncthreads();

    // ---- E: divergence, correction, stage update
    double rsq = 0.0;
    if (active) {
        double res[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R* Fx = sGx + el * Y::UE + v * Y::UV;
            const R* Fy = sGy + el * Y::UE + v * Y::UV;
            R dxi = 0, deta = 0;
#pragma unroll
            for (int m = 0; m < N1; ++m) {
                dxi += sD[i * N1 + m] * Fx[j * N1 + m];
                deta += sD[j * N1 + m] * Fy[m * N1 + i];
            }
            const R c = sJ[Y::tr(el, 3, P - j, v)] * sC[i] + sJ[Y::tr(el, 1, j, v)] * sC[N1 + i] +
                        sJ[Y::tr(el, 0, i, v)] * sC[j] + sJ[Y::tr(el, 2, P - i, v)] * sC[N1 + j];
            res[v] = double(-(dxi + deta + c));
            rsq += res[v] * res[v];
        }
        if (MODE == MODE_RESIDUAL) {
#pragma unroll
            for (int v = 0; v < NV; ++v) a.R_out[v * PS + gi] = res[v];
        } else {
            const double dt = *a.dt_ptr;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double q = stage_input(v);
                const double u0 = HAS_U0 ? a.U0[v * PS + gi] : 0.0;
                double un;
                if (MODE == MODE_STAGE1) un = q + dt * res[v];
                else if (MODE == MODE_STAGE2) un = 0.75 * u0 + 0.25 * q + 0.25 * dt * res[v];
                else un = (1.0 / 3.0) * u0 + (2.0 / 3.0) * q + (2.0 / 3.0) * dt * res[v];
                double next = un;
                if (INC && MODE != MODE_STAGE3) {
                    const double base = MODE == MODE_STAGE1 ? q : u0;
                    const float dw = float(un - base);
                    a.dW_out[v * PS + gi] = dw;
                    next = base + double(dw);
                } else {
                    a.U_out[v * PS + gi] = un;
                }
                sU[ub + v * Y::UV] = R(next);
            }
        }
    }
    if (a.l2_partial) {
        const double s = block_reduce<OpSum>(rsq, sRed);
        if (tid == 0) a.l2_partial[tile] = s;
    }
    if (MODE == MODE_RESIDUAL) return;
    __syncthreads();

    // ---- F: gradient of the new state, then its traces for the other tiles
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
            point_gradient<R, P>(sU + el * Y::UE + v * Y::UV, sD, i, j, jx, jy, kx, ky, sGx[ub + v * Y::UV],
                                 sGy[ub + v * Y::UV]);
    }
    __syncthreads();
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E || !sOut[el2 * 4 + k]) continue;
        const TraceLine tl = trace_line<P>(k, f);
        R tu[NV], tx[NV], ty[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            tu[v] = trace_point<R, P>(sU + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
            tx[v] = trace_point<R, P>(sGx + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
            ty[v] = trace_point<R, P>(sGy + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
        }
        R* T = a.T_out + (((int64_t)e2 * 4 + k) * FP + f) * TW;
        store4(T, tu);
        store4(T + NV, tx);
        store4(T + 2 * NV, ty);
    }
}

// U and grad U on every edge (after a new state is set).
template <typename R, int P, int EPB>
__global__ __launch_bounds__(BLOCK) void visc_trace_kernel(int E, const double* U, const R* Jinv, R* T_out,
                                                           const OpTables<R> ops, int xcd_groups) {
    using Y = Lay<R, P>;
    using SM = ViscSmem<R, P, EPB>;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1, NT = EPB * S, NF = EPB * 4 * FP;
    R* const sU = SM::U();
    R* const sGx = SM::Gx();
    R* const sGy = SM::Gy();
    R* const sD = SM::D();
    R* const sLm = SM::Lm();
    R* const sLp = SM::Lp();
    const int tid = threadIdx.x;
    const int e0 = tile_of_block(xcd_groups) * EPB;
    const int64_t PS = (int64_t)E * S;
    for (int t = tid; t < N1 * N1; t += BLOCK) sD[t] = ops.D[t];
    if (tid < N1) { sLm[tid] = ops.Lm[tid]; sLp[tid] = ops.Lp[tid]; }
    const int el = tid / S, sp = tid - (tid / S) * S, e = e0 + el;
    const bool active = tid < NT && e < E;
    const int ub = el * Y::UE + sp;
    if (active)
#pragma unroll
        for (int v = 0; v < NV; ++v) sU[ub + v * Y::UV] = R(U[v * PS + (int64_t)e * S + sp]);
    __syncthreads();
    if (active) {
        const R jx = Jinv[(int64_t)e * 4], jy = Jinv[(int64_t)e * 4 + 1];
        const R kx = Jinv[(int64_t)e * 4 + 2], ky = Jinv[(int64_t)e * 4 + 3];
#pragma unroll
        for (int v = 0; v < NV; ++v)
            point_gradient<R, P>(sU + el * Y::UE + v * Y::UV, sD, sp % N1, sp / N1, jx, jy, kx, ky,
                                 sGx[ub + v * Y::UV], sGy[ub + v * Y::UV]);
    }
    __syncthreads();
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= E) continue;
        const TraceLine tl = trace_line<P>(k, f);
        R tu[NV], tx[NV], ty[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            tu[v] = trace_point<R, P>(sU + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
            tx[v] = trace_point<R, P>(sGx + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
            ty[v] = trace_point<R, P>(sGy + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
        }
        R* T = T_out + (((int64_t)e2 * 4 + k) * FP + f) * TW;
        store4(T, tu);
        store4(T + NV, tx);
        store4(T + 2 * NV, ty);
    }
}

template <typename R, int P>
hipError_t visc_launch_p(int mode, const StageArgs<R>& a, const OpTables<R>& ops, const ViscParams& vp,
                         hipStream_t s) {
    constexpr int EPB = BLOCK / ((P + 1) * (P + 1));
    const int grid = (std::max(a.E, 1) + EPB - 1) / EPB;
    switch (mode) {
    case MODE_RESIDUAL: klaunch(visc_stage_kernel<R, P, EPB, MODE_RESIDUAL>, grid, dim3(BLOCK), 0, s, a, ops, vp); break;
    case MODE_STAGE1: klaunch(visc_stage_kernel<R, P, EPB, MODE_STAGE1>, grid, dim3(BLOCK), 0, s, a, ops, vp); break;
    case MODE_STAGE2: klaunch(visc_stage_kernel<R, P, EPB, MODE_STAGE2>, grid, dim3(BLOCK), 0, s, a, ops, vp); break;
    default: klaunch(visc_stage_kernel<R, P, EPB, MODE_STAGE3>, grid, dim3(BLOCK), 0, s, a, ops, vp); break;
    }
    return hipGetLastError();
}

template <typename R, int P>
hipError_t visc_trace_p(int E, const double* U, const R* Jinv, R* T, const OpTables<R>& ops, int groups,
                        hipStream_t s) {
    constexpr int EPB = BLOCK / ((P + 1) * (P + 1));
    const int grid = (std::max(E, 1) + EPB - 1) / EPB;
    visc_trace_kernel<R, P, EPB><<<grid, BLOCK, 0, s>>>(E, U, Jinv, T, ops, groups);
    return hipGetLastError();
}

}  // namespace

int visc_epb(int P) { return BLOCK / ((P + 1) * (P + 1)); }

template <typename R>
hipError_t launch_visc_stage(int P, int mode, const StageArgs<R>& a, const OpTables<R>& ops, const ViscParams& vp,
                             hipStream_t s) {
    switch (P) {
    case 1: return visc_launch_p<R, 1>(mode, a, ops, vp, s);
    case 2: return visc_launch_p<R, 2>(mode, a, ops, vp, s);
    case 3: return visc_launch_p<R, 3>(mode, a, ops, vp, s);
    case 4: return visc_launch_p<R, 4>(mode, a, ops, vp, s);
    default: return visc_launch_p<R, 5>(mode, a, ops, vp, s);
    }
}

template <typename R>
hipError_t launch_visc_trace(int P, int E, const double* U, const R* Jinv, R* T, const OpTables<R>& ops, int groups,
                             hipStream_t s) {
    switch (P) {
    case 1: return visc_trace_p<R, 1>(E, U, Jinv, T, ops, groups, s);
    case 2: return visc_trace_p<R, 2>(E, U, Jinv, T, ops, groups, s);
    case 3: return visc_trace_p<R, 3>(E, U, Jinv, T, ops, groups, s);
    case 4: return visc_trace_p<R, 4>(E, U, Jinv, T, ops, groups, s);
    default: return visc_trace_p<R, 5>(E, U, Jinv, T, ops, groups, s);
    }
}

template hipError_t launch_visc_stage<double>(int, int, const StageArgs<double>&, const OpTables<double>&,
                                              const ViscParams&, hipStream_t);
template hipError_t launch_visc_stage<float>(int, int, const StageArgs<float>&, const OpTables<float>&,
                                             const ViscParams&, hipStream_t);
template hipError_t launch_visc_trace<double>(int, int, const double*, const double*, double*,
                                              const OpTables<double>&, int, hipStream_t);
template hipError_t launch_visc_trace<float>(int, int, const double*, const float*, float*, const OpTables<float>&,
                                             int, hipStream_t);

}  // namespace zj
