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

#include "zj_adjoint.hpp"
#include "zj_device.hpp"
#include "zj_riemann_vjp.hpp"
#include "zj_viscous.hpp"

#include <algorithm>
#include <type_traits>

namespace zj {

namespace {

constexpr int TW = 3 * NV;  // trace words per flux point: U, dU/dx, dU/dy

// A tile's arrays.  G holds the gradient (x then y, each [var][point]), then
// the contravariant total fluxes in the same slots; Tr, TGx, TGy the traces
// of U and (FP32 only) of the gradient at the tile's own face points; J holds
// the jumps, then the flux differences.  ViscSmem: static arrays of the
// single-mesh and per-order kernels; ViscSmemDyn: the same arrays in one
// dynamic LDS block (the one-launch batch, sized for its largest order).
template <typename R, int P, int EPB>
struct ViscSmem {
    using Y = Lay<R, P>;
    static constexpr int N1 = P + 1, S = N1 * N1;
    __device__ static R* U() { __shared__ R a[EPB * Y::UE]; return a; }
    __device__ static R* Gx() { __shared__ R a[EPB * Y::UE]; return a; }
    __device__ static R* Gy() { __shared__ R a[EPB * Y::UE]; return a; }
    __device__ static R* Tr() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* TGx() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* TGy() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* J() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* FI() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* D() { __shared__ R a[N1 * N1]; return a; }
    __device__ static R* Lm() { __shared__ R a[N1]; return a; }
    __device__ static R* Lp() { __shared__ R a[N1]; return a; }
    __device__ static R* C() { __shared__ R a[2 * N1]; return a; }
    __device__ static double* Red() { __shared__ double a[BLOCK / 64]; return a; }
    __device__ static unsigned char* Out() { __shared__ unsigned char a[EPB * 4]; return a; }
};

template <typename R, int P, int EPB>
struct ViscSmemDyn {
    using Y = Lay<R, P>;
    static constexpr int N1 = P + 1;
    static constexpr bool PRE = std::is_same<R, float>::value || EPB * 4 * (P + 1) <= BLOCK;  // as visc_body
    static constexpr int oU = 0, oGx = oU + EPB * Y::UE, oGy = oGx + EPB * Y::UE, oTr = oGy + EPB * Y::UE;
    static constexpr int oTGx = oTr + EPB * Y::TE, oTGy = oTGx + (PRE ? EPB * Y::TE : 0);
    static constexpr int oJ = oTGy + (PRE ? EPB * Y::TE : 0), oFI = oJ + EPB * Y::TE, oD = oFI + EPB * Y::TE;
    static constexpr int oLm = oD + N1 * N1, oLp = oLm + N1, oC = oLp + N1, nR = oC + 2 * N1;
    static constexpr int red_byte = ((nR * (int)sizeof(R) + 15) / 16) * 16;
    static constexpr int bytes = red_byte + (BLOCK / 64) * 8 + EPB * 4;
    unsigned char* base;
    __device__ R* r(int o) const { return reinterpret_cast<R*>(base) + o; }
    __device__ R* U() const { return r(oU); }
    __device__ R* Gx() const { return r(oGx); }
    __device__ R* Gy() const { return r(oGy); }
    __device__ R* Tr() const { return r(oTr); }
    __device__ R* TGx() const { return r(oTGx); }
    __device__ R* TGy() const { return r(oTGy); }
    __device__ R* J() const { return r(oJ); }
    __device__ R* FI() const { return r(oFI); }
    __device__ R* D() const { return r(oD); }
    __device__ R* Lm() const { return r(oLm); }
    __device__ R* Lp() const { return r(oLp); }
    __device__ R* C() const { return r(oC); }
    __device__ double* Red() const { return reinterpret_cast<double*>(base + red_byte); }
    __device__ unsigned char* Out() const { return base + red_byte + (BLOCK / 64) * 8; }
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

// Threads of a workgroup for a tile of EPB elements: its points or its face
// points, whichever are more, in whole waves, at most BLOCK (the full tile
// takes BLOCK; the adjoint's smaller tiles leave no wave idle).
template <int P, int EPB>
constexpr int visc_threads() {
    constexpr int S = (P + 1) * (P + 1), NF = EPB * 4 * (P + 1);
    constexpr int n = EPB * S > NF ? EPB * S : NF;
    constexpr int r = (n + 63) / 64 * 64;
    return r < BLOCK ? r : BLOCK;
}

// One tile's viscous stage (the workgroup's NTHR threads, the tile's elements).
template <typename R, int P, int EPB, int MODE, int NTHR, class SM>
__device__ __forceinline__ void visc_body(const StageArgs<R>& a, const OpTables<R>& ops, const ViscParams& vp,
                                          const int tile, const int tid, const SM& sm) {
    using Y = Lay<R, P>;
    constexpr bool INC = std::is_same<R, float>::value;
    constexpr bool HAS_U0 = MODE == MODE_STAGE2 || MODE == MODE_STAGE3;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1;
    constexpr int NT = EPB * S, NF = EPB * 4 * FP;
    static_assert(NT <= NTHR, "tile exceeds the workgroup");
    R* const sU = sm.U();
    R* const sGx = sm.Gx();
    R* const sGy = sm.Gy();
    // The gradient traces once per face point, in LDS (fewer VALU
    // instructions and registers: FP32 71 instead of 112 VGPRs at P4, FP64
    // about 115 instead of 184).  FP64 with more face points than threads
    // (P1, P2) recomputes them where they are read: there the two extra
    // arrays would take its LDS past what the registers allow.
    constexpr bool PRE = std::is_same<R, float>::value || NF <= NTHR;
    // With one face point per thread (NF <= NTHR) the face phase B2a reads
    // all its inputs before a barrier and then writes, so the jumps and the
    // common fluxes can take the gradient traces' place in LDS (FP32 P3-P5:
    // 23 instead of 30 KB at P4, 31 instead of 41 KB at P3, so more
    // workgroups fit on a CU).
    constexpr bool ALIAS = PRE && NF <= NTHR;
    R* const sTr = sm.Tr();
    R* sTGx = nullptr;
    R* sTGy = nullptr;
    if constexpr (PRE) { sTGx = sm.TGx(); sTGy = sm.TGy(); }
    R* sJ;
    R* sFI;
    if constexpr (ALIAS) { sJ = sTGx; sFI = sTGy; } else { sJ = sm.J(); sFI = sm.FI(); }
    R* const sD = sm.D();
    R* const sLm = sm.Lm();
    R* const sLp = sm.Lp();
    R* const sC = sm.C();
    double* const sRed = sm.Red();
    unsigned char* const sOut = sm.Out();

    const int e0 = tile * EPB;
    const int64_t PS = (int64_t)a.E * S;
    const R g = R(a.gamma), mu = R(vp.mu), kappa = R(vp.kappa), eta = R(vp.eta);
    const R gLe = R(vp.gLe), gRe = R(vp.gRe);

    for (int t = tid; t < N1 * N1; t += NTHR) sD[t] = ops.D[t];
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
    for (int t = tid; t < NF; t += NTHR) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        if (e0 + el2 < a.E) {
            const TraceLine tl = trace_line<P>(k, f);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                sTr[Y::tr(el2, k, f, v)] = trace_point<R, P>(sU + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
        }
    }
    __syncthreads();

    // ---- B1b: traces of the gradient at the own face points (each once;
    // both sides of an in-tile face read them)
    for (int t = tid; PRE && t < NF; t += NTHR) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        if (e0 + el2 < a.E) {
            const TraceLine tl = trace_line<P>(k, f);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                sTGx[Y::tr(el2, k, f, v)] = trace_point<R, P>(sGx + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
                sTGy[Y::tr(el2, k, f, v)] = trace_point<R, P>(sGy + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
            }
        }
    }
    if (PRE) __syncthreads();

    // ---- B2a: jumps and the common viscous normal flux per (edge, flux point)
    // (kept apart from the Riemann solve so that the two never hold their
    // registers at once)
    for (int t0 = 0; t0 < NF; t0 += NTHR) {
        const int t = t0 + tid;
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        const bool item = t < NF && e2 < a.E;
        R qx[NV], qy[NV], Um[NV], jw[NV];
        R nx = R(0), ny = R(0);
        if (item) {
            const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
            nx = a.normals[((int64_t)e2 * 4 + k) * 2];
            ny = a.normals[((int64_t)e2 * 4 + k) * 2 + 1];
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
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const R own = sTr[Y::tr(el2, k, f, v)];
                R oth, gxn, gyn;
                if (inside) {
                    oth = sTr[Y::tr(nl, L.nbr_edge, oslot, v)];
                    gxn = PRE ? sTGx[Y::tr(nl, L.nbr_edge, oslot, v)]
                              : trace_point<R, P>(sGx + nl * Y::UE + v * Y::UV, tn, sLm, sLp);
                    gyn = PRE ? sTGy[Y::tr(nl, L.nbr_edge, oslot, v)]
                              : trace_point<R, P>(sGy + nl * Y::UE + v * Y::UV, tn, sLm, sLp);
                } else {
                    oth = T[v];
                    gxn = T[NV + v];
                    gyn = T[2 * NV + v];
                }
                const R gxo = PRE ? sTGx[Y::tr(el2, k, f, v)]
                                  : trace_point<R, P>(sGx + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
                const R gyo = PRE ? sTGy[Y::tr(el2, k, f, v)]
                                  : trace_point<R, P>(sGy + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
                const R uc = R(0.5) * (own + oth);
                const R jo = uc - own, jn = uc - oth;
                jw[v] = jo;
                qx[v] = R(0.5) * ((gxo + ax * jo) + (gxn + bx * jn));
                qy[v] = R(0.5) * ((gyo + ay * jo) + (gyn + by * jn));
                Um[v] = uc;
            }
        }
        if constexpr (ALIAS) __syncthreads();  // every gradient trace read: the jumps and fluxes take their place
        if (item) {
            R Fv[NV], Gv[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) sJ[Y::tr(el2, k, f, v)] = jw[v];
            visc_flux(Um, qx, qy, g, mu, kappa, Fv, Gv);
#pragma unroll
            for (int v = 0; v < NV; ++v) sFI[Y::tr(el2, k, f, v)] = Fv[v] * nx + Gv[v] * ny;  // as the left side sees it
        }
    }
    __syncthreads();

    // ---- B2b: Riemann flux; total common normal flux, outward and scaled
    for (int t = tid; t < NF; t += NTHR) {
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
    for (int t = tid; t < NF; t += NTHR) {
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
        const double s = block_reduce_n<OpSum, NTHR / 64>(rsq, sRed);
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
    for (int t = tid; t < NF; t += NTHR) {
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

template <typename R, int P, int EPB, int MODE, int NTHR = visc_threads<P, EPB>()>
__global__ __launch_bounds__(NTHR) void visc_stage_kernel(const StageArgs<R> a, const OpTables<R> ops,
                                                          const ViscParams vp) {
    visc_body<R, P, EPB, MODE, NTHR>(a, ops, vp, tile_of_block(a.xcd_groups), threadIdx.x, ViscSmem<R, P, EPB>{});
}

// The stage for many meshes of order P: work item w = (instance, tile); a
// workgroup takes items blockIdx.x, blockIdx.x + gridDim.x, ...
template <typename R, int P, int EPB, int MODE>
__global__ __launch_bounds__(BLOCK) void batched_visc_kernel(const BatchInst<R>* __restrict__ insts,
                                                             const int2* __restrict__ items, int n_items) {
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        // the item is the workgroup's: in scalar registers, so that the
        // instance's arguments are scalar loads, not per-lane copies
        const int inst = __builtin_amdgcn_readfirstlane(items[w].x);
        const int tile = __builtin_amdgcn_readfirstlane(items[w].y);
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));  // keeps the per-thread index math in the loop body (registers)
        const BatchInst<R>& in = insts[inst];
        visc_body<R, P, EPB, MODE, BLOCK>(in.args, in.ops, in.vp, tile, tid, ViscSmem<R, P, EPB>{});
        __syncthreads();  // the next item reuses the LDS
    }
}

// One launch per stage for a batch of Navier-Stokes members of every order:
// the item's instance picks the stage body; the LDS buffer is dynamic, sized
// for the largest order present.
template <typename R, int MODE>
__global__ __launch_bounds__(BLOCK) void batched_visc_all_kernel(const BatchInst<R>* __restrict__ insts,
                                                                 const int2* __restrict__ items, int n_items) {
    extern __shared__ __align__(16) unsigned char raw_visc[];
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const int inst = __builtin_amdgcn_readfirstlane(items[w].x);
        const int tile = __builtin_amdgcn_readfirstlane(items[w].y);
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const BatchInst<R>& in = insts[inst];
        const int P = __builtin_amdgcn_readfirstlane(in.P);
#define ZJ_VALL(PP)                                                                                            \
    case PP: {                                                                                                 \
        constexpr int EPB = BLOCK / ((PP + 1) * (PP + 1));                                                     \
        visc_body<R, PP, EPB, MODE, BLOCK>(in.args, in.ops, in.vp, tile, tid, ViscSmemDyn<R, PP, EPB>{raw_visc}); \
        break;                                                                                                 \
    }
        switch (P) {
            ZJ_VALL(1) ZJ_VALL(2) ZJ_VALL(3) ZJ_VALL(4) ZJ_VALL(5)
        default: break;
        }
#undef ZJ_VALL
        __syncthreads();  // the next item reuses the LDS
    }
}

// U and grad U on every edge (after a new state is set).
template <typename R, int P, int EPB>
__global__ __launch_bounds__(BLOCK) void visc_trace_kernel(int E, const double* U, const R* Jinv, R* T_out,
                                                           const OpTables<R> ops, int xcd_groups) {
    using Y = Lay<R, P>;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1, NT = EPB * S, NF = EPB * 4 * FP;
    __shared__ R sU[EPB * Y::UE], sGx[EPB * Y::UE], sGy[EPB * Y::UE];
    __shared__ R sD[N1 * N1], sLm[N1], sLp[N1];
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

// ------------------------------------------------------------ adjoint
// Vector-Jacobian product of one BR2 stage (MODE_STAGE1..3) or of the
// residual (MODE_RESIDUAL).  The forward values are recomputed as
// visc_stage_kernel forms them; then, in reverse:
//   R1  cotangent of the stage output: its partial value, plus the transposed
//       traces of U and of grad U (through the transposed gradient) of the
//       edge cotangents other tiles left; the RK combination
//   R2  the correction transposed: the flux differences' cotangents
//   R3  per point: the transposed divergence and traces of the contravariant
//       fluxes, the metric, the Euler and viscous flux VJPs, which give the
//       cotangent of the corrected gradient Q; its correction terms
//       transposed onto the jumps
//   R4  per face point, as the forward solves it (both sides solve each
//       face): the Riemann and viscous common-flux VJPs, the BR2 gradients
//       and jumps backwards, to this side's trace and gradient-trace
//       cotangents and the partner's (LDS in the tile, HBM across tiles)
//   R5  the transposed traces; the gradient's cotangent through the
//       transposed gradient operator; the cotangent of q
// Traces and their cotangents across tiles carry 12 words per face point:
// U, dU/dx, dU/dy.
template <typename R, int P>
constexpr int visc_adj_epb() {
    using Y = Lay<R, P>;
    constexpr int full = BLOCK / ((P + 1) * (P + 1));
    // LDS within 120 KiB (gfx950 gives a workgroup up to 160 KiB): every
    // order's full tile at both precisions
    constexpr int cap = 122880 / ((6 * Y::UE + 9 * Y::TE) * (int)sizeof(R));
    return cap < 1 ? 1 : (cap < full ? cap : full);
}

// Both traces and both BR2 lifting coefficients of one face point of
// element el2 (edge k, point f), the partner read from LDS or HBM.
template <typename R, int P>
struct ViscFacePoint {
    R own[NV], oth[NV], gxo[NV], gyo[NV], gxn[NV], gyn[NV];
    R ax, ay, bx, by;
};

template <typename R, int P>
__device__ __forceinline__ void visc_face_point(ViscFacePoint<R, P>& F, const R* sU, const R* sGx, const R* sGy,
                                                const R* sTr, const R* sLm, const R* sLp, const R* T_in,
                                                const R* Jinv, int e0, int epb, int el2, int e2, int k, int f,
                                                const EdgeLink& L, R eta, R gLe, R gRe) {
    using Y = Lay<R, P>;
    constexpr int FP = P + 1;
    const int nl = L.nbr - e0;
    const bool inside = nl >= 0 && nl < epb;
    const int oslot = P - f;
    const TraceLine tl = trace_line<P>(k, f);
    const TraceLine tn = trace_line<P>(L.nbr_edge, oslot);
    const R* T = T_in + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + oslot) * TW;
    const R* Jo = Jinv + (int64_t)e2 * 4;
    const R* Jn = Jinv + (int64_t)L.nbr * 4;
    const bool oxi = (k & 1) != 0, nxi = (L.nbr_edge & 1) != 0;
    const R go = eta * ((k == 0 || k == 3) ? gLe : gRe);
    const R gn = eta * ((L.nbr_edge == 0 || L.nbr_edge == 3) ? gLe : gRe);
    F.ax = go * (oxi ? Jo[0] : Jo[2]); F.ay = go * (oxi ? Jo[1] : Jo[3]);
    F.bx = gn * (nxi ? Jn[0] : Jn[2]); F.by = gn * (nxi ? Jn[1] : Jn[3]);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        F.own[v] = sTr[Y::tr(el2, k, f, v)];
        if (inside) {
            F.oth[v] = sTr[Y::tr(nl, L.nbr_edge, oslot, v)];
            F.gxn[v] = trace_point<R, P>(sGx + nl * Y::UE + v * Y::UV, tn, sLm, sLp);
            F.gyn[v] = trace_point<R, P>(sGy + nl * Y::UE + v * Y::UV, tn, sLm, sLp);
        } else {
            F.oth[v] = T[v];
            F.gxn[v] = T[NV + v];
            F.gyn[v] = T[2 * NV + v];
        }
        F.gxo[v] = trace_point<R, P>(sGx + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
        F.gyo[v] = trace_point<R, P>(sGy + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
    }
    (void)sU;
}

template <typename R, int P, int EPB, int MODE, int NTHR>
__device__ __forceinline__ void visc_vjp_body(const VjpArgs<R>& a, const OpTables<R>& ops, const ViscParams& vp,
                                              const int tile) {
    using Y = Lay<R, P>;
    constexpr bool INC = std::is_same<R, float>::value;
    constexpr bool HAS_U0 = MODE == MODE_STAGE2 || MODE == MODE_STAGE3;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1;
    constexpr int NT = EPB * S, NF = EPB * 4 * FP;
    static_assert(NT <= NTHR, "tile exceeds the workgroup");
    constexpr double ca = MODE == MODE_STAGE2 ? 0.75 : (MODE == MODE_STAGE3 ? 1.0 / 3.0 : 0.0);
    constexpr double cb = MODE == MODE_STAGE2 ? 0.25 : (MODE == MODE_STAGE3 ? 2.0 / 3.0 : 1.0);
    constexpr double cc = MODE == MODE_STAGE2 ? 0.25 : (MODE == MODE_STAGE3 ? 2.0 / 3.0 : 1.0);
    __shared__ R sU[EPB * Y::UE], sGx[EPB * Y::UE], sGy[EPB * Y::UE];
    __shared__ R sW1[EPB * Y::UE], sW2[EPB * Y::UE], sW3[EPB * Y::UE];
    __shared__ R sTr[EPB * Y::TE], sJ[EPB * Y::TE], sFI[EPB * Y::TE];
    __shared__ R sAu[EPB * Y::TE], sAx[EPB * Y::TE], sAy[EPB * Y::TE];  // incoming, then own-side cotangents
    __shared__ R sNu[EPB * Y::TE], sNx[EPB * Y::TE], sNy[EPB * Y::TE];  // cotangents from in-tile partners
    __shared__ R sD[N1 * N1], sLm[N1], sLp[N1], sC[2 * N1];
    __shared__ unsigned char sIn[EPB * 4];
    __shared__ double sRed[BLOCK / 64];

    const int tid = threadIdx.x;
    const int e0 = tile * EPB;
    const int64_t PS = (int64_t)a.E * S;
    const R g = R(a.gamma), mu = R(vp.mu), kappa = R(vp.kappa), eta = R(vp.eta);
    const R gLe = R(vp.gLe), gRe = R(vp.gRe);
    for (int t = tid; t < N1 * N1; t += NTHR) sD[t] = ops.D[t];
    if (tid < N1) { sLm[tid] = ops.Lm[tid]; sLp[tid] = ops.Lp[tid]; sC[tid] = ops.gL[tid]; sC[N1 + tid] = ops.gR[tid]; }

    const int el = tid / S, sp = tid - (tid / S) * S;
    const int e = e0 + el;
    const bool active = tid < NT && e < a.E;
    const int64_t gi = (int64_t)e * S + sp;
    const int ub = el * Y::UE + sp;
    const int i = sp % N1, j = sp / N1;
    const bool incoming = MODE != MODE_RESIDUAL && a.Tb_in != nullptr;
    // the forward residual is needed only for the dt cotangent
    const bool need_res = MODE != MODE_RESIDUAL && a.dtb_partial != nullptr;

    // ---- forward A: stage input
    double q[NV], u0[NV];
    R jx = R(0), jy = R(0), kx = R(0), ky = R(0);
    if (active) {
        jx = a.Jinv[(int64_t)e * 4]; jy = a.Jinv[(int64_t)e * 4 + 1];
        kx = a.Jinv[(int64_t)e * 4 + 2]; ky = a.Jinv[(int64_t)e * 4 + 3];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            u0[v] = HAS_U0 ? a.U0[v * PS + gi] : 0.0;
            q[v] = (INC && HAS_U0) ? u0[v] + double(a.dW_in[v * PS + gi]) : a.U_in[v * PS + gi];
            sU[ub + v * Y::UV] = R(q[v]);
        }
    }
    __syncthreads();

    // ---- forward B1: gradient, own traces; incoming edge cotangents
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
            point_gradient<R, P>(sU + el * Y::UE + v * Y::UV, sD, i, j, jx, jy, kx, ky, sGx[ub + v * Y::UV],
                                 sGy[ub + v * Y::UV]);
    }
    for (int t = tid; t < NF; t += NTHR) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const TraceLine tl = trace_line<P>(k, f);
#pragma unroll
        for (int v = 0; v < NV; ++v)
            sTr[Y::tr(el2, k, f, v)] = trace_point<R, P>(sU + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        const int nl = L.nbr - e0;
        const bool inside = nl >= 0 && nl < EPB;
        if (f == 0) sIn[el2 * 4 + k] = inside;
        R tb[TW];
#pragma unroll
        for (int w = 0; w < TW; ++w) tb[w] = R(0);
        if (incoming && !inside) {
            const R* src = a.Tb_in + (((int64_t)e2 * 4 + k) * FP + f) * TW;
            load4(src, tb);
            load4(src + NV, tb + NV);
            load4(src + 2 * NV, tb + 2 * NV);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sAu[Y::tr(el2, k, f, v)] = tb[v];
            sAx[Y::tr(el2, k, f, v)] = tb[NV + v];
            sAy[Y::tr(el2, k, f, v)] = tb[2 * NV + v];
        }
    }
    __syncthreads();

    // ---- forward B2: jumps, common fluxes (viscous and Riemann), outward and scaled
    for (int t = tid; t < NF; t += NTHR) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        const R nx = a.normals[((int64_t)e2 * 4 + k) * 2], ny = a.normals[((int64_t)e2 * 4 + k) * 2 + 1];
        ViscFacePoint<R, P> F;
        visc_face_point<R, P>(F, sU, sGx, sGy, sTr, sLm, sLp, a.T_in, a.Jinv, e0, EPB, el2, e2, k, f, L, eta, gLe,
                              gRe);
        R qx[NV], qy[NV], Um[NV], Fv[NV], Gv[NV], UL[NV], UR[NV], Fc[NV];
        const bool right = L.role == ROLE_RIGHT;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R uc = R(0.5) * (F.own[v] + F.oth[v]);
            const R jo = uc - F.own[v], jn = uc - F.oth[v];
            sJ[Y::tr(el2, k, f, v)] = jo;
            qx[v] = R(0.5) * ((F.gxo[v] + F.ax * jo) + (F.gxn[v] + F.bx * jn));
            qy[v] = R(0.5) * ((F.gyo[v] + F.ay * jo) + (F.gyn[v] + F.by * jn));
            Um[v] = uc;
            UL[v] = right ? F.oth[v] : F.own[v];
            UR[v] = right ? F.own[v] : F.oth[v];
        }
        if (!need_res) continue;  // the jumps are all the reverse phases use
        visc_flux(Um, qx, qy, g, mu, kappa, Fv, Gv);
        riemann_by_kind(a.riemann, UL, UR, nx, ny, g, Fc);
        const R ef = a.edge_f[(int64_t)e2 * 4 + k];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R fc = Fc[v] - (Fv[v] * nx + Gv[v] * ny);
            sFI[Y::tr(el2, k, f, v)] = ef * (right ? -fc : fc);
        }
    }
    __syncthreads();

    // ---- forward C: corrected gradient; contravariant total fluxes
    auto corrected = [&](int v, R& Qx, R& Qy) {
        const R cxi = sC[i] * sJ[Y::tr(el, 3, P - j, v)] + sC[N1 + i] * sJ[Y::tr(el, 1, j, v)];
        const R ceta = sC[j] * sJ[Y::tr(el, 0, i, v)] + sC[N1 + j] * sJ[Y::tr(el, 2, P - i, v)];
        Qx = sGx[ub + v * Y::UV] + jx * cxi + kx * ceta;
        Qy = sGy[ub + v * Y::UV] + jy * cxi + ky * ceta;
    };
    if (active && need_res) {
        R Uq[NV], Qx[NV], Qy[NV], fx[NV], fy[NV], Fv[NV], Gv[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Uq[v] = sU[ub + v * Y::UV];
            corrected(v, Qx[v], Qy[v]);
        }
        euler_flux(Uq, g, fx, fy);
        visc_flux(Uq, Qx, Qy, g, mu, kappa, Fv, Gv);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R Fx = fx[v] - Fv[v], Fy = fy[v] - Gv[v];
            sW1[ub + v * Y::UV] = jx * Fx + jy * Fy;
            sW2[ub + v * Y::UV] = kx * Fx + ky * Fy;
        }
    }
    __syncthreads();

    // ---- forward D: flux differences (in place of the common fluxes)
    for (int t = tid; need_res && t < NF; t += NTHR) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        if (e0 + el2 >= a.E) continue;
        const TraceLine tl = trace_line<P>(k, f);
        const R* Fb = ((k & 1) ? sW1 : sW2) + el2 * Y::UE;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            sFI[Y::tr(el2, k, f, v)] = sFI[Y::tr(el2, k, f, v)] - trace_point<R, P>(Fb + v * Y::UV, tl, sLm, sLp);
    }
    __syncthreads();

    // ---- forward E: residual
    double res[NV] = {0.0, 0.0, 0.0, 0.0};
    if (active && need_res) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R* Fx = sW1 + el * Y::UE + v * Y::UV;
            const R* Fy = sW2 + el * Y::UE + v * Y::UV;
            R dxi = 0, deta = 0;
#pragma unroll
            for (int m = 0; m < N1; ++m) {
                dxi += sD[i * N1 + m] * Fx[j * N1 + m];
                deta += sD[j * N1 + m] * Fy[m * N1 + i];
            }
            const R c = sFI[Y::tr(el, 3, P - j, v)] * sC[i] + sFI[Y::tr(el, 1, j, v)] * sC[N1 + i] +
                        sFI[Y::tr(el, 0, i, v)] * sC[j] + sFI[Y::tr(el, 2, P - i, v)] * sC[N1 + j];
            res[v] = double(-(dxi + deta + c));
        }
    }
    __syncthreads();  // the fluxes in sW1, sW2 are read

    // transposed traces (edges 0..3 at this point's slots) of a trace array
    auto trace_t = [&](const R* A, int v) -> R {
        return sLm[j] * A[Y::tr(el, 0, i, v)] + sLp[j] * A[Y::tr(el, 2, P - i, v)] +
               sLp[i] * A[Y::tr(el, 1, j, v)] + sLm[i] * A[Y::tr(el, 3, P - j, v)];
    };
    // transposed gradient at this point of the (dxi, deta) cotangents in sW1, sW2
    auto grad_t = [&](int v) -> R {
        const R* X = sW1 + el * Y::UE + v * Y::UV;
        const R* H = sW2 + el * Y::UE + v * Y::UV;
        R acc = R(0);
#pragma unroll
        for (int m = 0; m < N1; ++m) acc += sD[m * N1 + i] * X[j * N1 + m] + sD[m * N1 + j] * H[m * N1 + i];
        return acc;
    };

    // ---- R1a: incoming gradient-trace cotangents onto the reference derivatives
    if (incoming && active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R gbx = trace_t(sAx, v), gby = trace_t(sAy, v);
            sW1[ub + v * Y::UV] = jx * gbx + jy * gby;
            sW2[ub + v * Y::UV] = kx * gbx + ky * gby;
        }
    }
    __syncthreads();

    // ---- R1b: cotangent of the stage output, RK combination
    double rsum = 0.0;
    double qb[NV] = {0.0, 0.0, 0.0, 0.0};
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double rb;
            if (MODE == MODE_RESIDUAL) {
                rb = a.Ub_in[v * PS + gi];
            } else {
                const double dt = *a.dt_ptr;
                double unb = a.Ub_in[v * PS + gi];
                if (incoming) unb += double(trace_t(sAu, v)) + double(grad_t(v));
                if (MODE == MODE_STAGE3) a.U0b[v * PS + gi] = ca * unb;
                if (MODE == MODE_STAGE2) a.U0b[v * PS + gi] += ca * unb;
                qb[v] = cb * unb;
                rb = cc * dt * unb;
                rsum += cc * res[v] * unb;
            }
            sW3[ub + v * Y::UV] = R(rb);
        }
    }
    if (a.dtb_partial) {
        const double s = block_reduce_n<OpSum, NTHR / 64>(rsum, sRed);
        if (tid == 0) a.dtb_partial[tile] = s;
    }
    __syncthreads();

    // ---- R2: cotangents of the flux differences (in place)
    for (int t = tid; t < NF; t += NTHR) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        if (e0 + el2 >= a.E) continue;
        const R* Rb = sW3 + el2 * Y::UE;
        const bool odd = (k & 1) != 0;
        const int line = (k == 0 || k == 1) ? f : P - f;
        const R* gw = sC + ((k == 1 || k == 2) ? N1 : 0);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R acc = R(0);
#pragma unroll
            for (int m = 0; m < N1; ++m) acc -= gw[m] * Rb[v * Y::UV + (odd ? line * N1 + m : m * N1 + line)];
            sFI[Y::tr(el2, k, f, v)] = acc;
        }
    }
    __syncthreads();

    // ---- R3: fluxes at the points backwards; the corrected gradient's cotangent
    R Qxb[NV] = {R(0), R(0), R(0), R(0)}, Qyb[NV] = {R(0), R(0), R(0), R(0)};
    double musum = 0.0;  // <cotangent of F_v, F_v> over this thread's points and face points
    if (active) {
        const R* rbp = sW3 + el * Y::UE;
        R Uq[NV], Qx[NV], Qy[NV], Fxb[NV], Fyb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Uq[v] = sU[ub + v * Y::UV];
            corrected(v, Qx[v], Qy[v]);
            const R* rv = rbp + v * Y::UV;
            // the flux traces enter the differences negatively
            R ftb = -(sLp[i] * sFI[Y::tr(el, 1, j, v)] + sLm[i] * sFI[Y::tr(el, 3, P - j, v)]);
            R gtb = -(sLm[j] * sFI[Y::tr(el, 0, i, v)] + sLp[j] * sFI[Y::tr(el, 2, P - i, v)]);
#pragma unroll
            for (int m = 0; m < N1; ++m) {
                ftb -= sD[m * N1 + i] * rv[j * N1 + m];
                gtb -= sD[m * N1 + j] * rv[m * N1 + i];
            }
            Fxb[v] = jx * ftb + kx * gtb;
            Fyb[v] = jy * ftb + ky * gtb;
        }
        R ubq[NV] = {R(0), R(0), R(0), R(0)}, nFx[NV], nFy[NV];
        euler_flux_vjp(Uq, g, Fxb, Fyb, ubq);
#pragma unroll
        for (int v = 0; v < NV; ++v) { nFx[v] = -Fxb[v]; nFy[v] = -Fyb[v]; }
        visc_flux_vjp(Uq, Qx, Qy, g, mu, kappa, nFx, nFy, ubq, Qxb, Qyb);
        if (a.mub_partial) {  // F_v is proportional to mu (kappa too): dF_v/dmu = F_v / mu
            R Fv[NV], Gv[NV];
            visc_flux(Uq, Qx, Qy, g, mu, kappa, Fv, Gv);
#pragma unroll
            for (int v = 1; v < NV; ++v) musum += double(nFx[v] * Fv[v] + nFy[v] * Gv[v]);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            qb[v] += double(ubq[v]);
            sW1[ub + v * Y::UV] = jx * Qxb[v] + jy * Qyb[v];  // cotangent of the xi correction term
            sW2[ub + v * Y::UV] = kx * Qxb[v] + ky * Qyb[v];  // and of the eta one
        }
    }
    __syncthreads();

    // ---- R4: face points backwards
    for (int t = tid; t < NF; t += NTHR) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        const R nx = a.normals[((int64_t)e2 * 4 + k) * 2], ny = a.normals[((int64_t)e2 * 4 + k) * 2 + 1];
        const int nl = L.nbr - e0;
        const bool inside = nl >= 0 && nl < EPB;
        const int oslot = P - f;
        // the jump's cotangent from the correction terms of the points on this slot's line
        const bool odd = (k & 1) != 0;
        const int line = (k == 0 || k == 1) ? f : P - f;
        const R* gw = sC + ((k == 1 || k == 2) ? N1 : 0);
        const R* Cb = (odd ? sW1 : sW2) + el2 * Y::UE;
        ViscFacePoint<R, P> F;
        visc_face_point<R, P>(F, sU, sGx, sGy, sTr, sLm, sLp, a.T_in, a.Jinv, e0, EPB, el2, e2, k, f, L, eta, gLe,
                              gRe);
        const bool right = L.role == ROLE_RIGHT;
        R Jb[NV], uc[NV], jo[NV], jn[NV], qx[NV], qy[NV], UL[NV], UR[NV], Fcb[NV], Fvb[NV], Gvb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R acc = R(0);
#pragma unroll
            for (int m = 0; m < N1; ++m) acc += gw[m] * Cb[v * Y::UV + (odd ? line * N1 + m : m * N1 + line)];
            Jb[v] = acc;
            uc[v] = R(0.5) * (F.own[v] + F.oth[v]);
            jo[v] = uc[v] - F.own[v];
            jn[v] = uc[v] - F.oth[v];
            qx[v] = R(0.5) * ((F.gxo[v] + F.ax * jo[v]) + (F.gxn[v] + F.bx * jn[v]));
            qy[v] = R(0.5) * ((F.gyo[v] + F.ay * jo[v]) + (F.gyn[v] + F.by * jn[v]));
            UL[v] = right ? F.oth[v] : F.own[v];
            UR[v] = right ? F.own[v] : F.oth[v];
        }
        const R ef = a.edge_f[(int64_t)e2 * 4 + k];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R d = sFI[Y::tr(el2, k, f, v)];
            Fcb[v] = ef * (right ? -d : d);
            Fvb[v] = -Fcb[v] * nx;
            Gvb[v] = -Fcb[v] * ny;
        }
        R ULb[NV] = {R(0), R(0), R(0), R(0)}, URb[NV] = {R(0), R(0), R(0), R(0)};
        riemann_by_kind_vjp(a.riemann, UL, UR, nx, ny, g, Fcb, ULb, URb);
        R ucb[NV] = {R(0), R(0), R(0), R(0)}, qxb[NV] = {R(0), R(0), R(0), R(0)}, qyb[NV] = {R(0), R(0), R(0), R(0)};
        visc_flux_vjp(uc, qx, qy, g, mu, kappa, Fvb, Gvb, ucb, qxb, qyb);
        if (a.mub_partial) {
            R Fv[NV], Gv[NV];
            visc_flux(uc, qx, qy, g, mu, kappa, Fv, Gv);
#pragma unroll
            for (int v = 1; v < NV; ++v) musum += double(Fvb[v] * Fv[v] + Gvb[v] * Gv[v]);
        }
        R nb[TW];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R job = Jb[v] + R(0.5) * (F.ax * qxb[v] + F.ay * qyb[v]);
            const R jnb = R(0.5) * (F.bx * qxb[v] + F.by * qyb[v]);
            const R uct = ucb[v] + job + jnb;
            const R ownb = (right ? URb[v] : ULb[v]) + R(0.5) * uct - job;
            nb[v] = (right ? ULb[v] : URb[v]) + R(0.5) * uct - jnb;
            nb[NV + v] = R(0.5) * qxb[v];
            nb[2 * NV + v] = R(0.5) * qyb[v];
            sAu[Y::tr(el2, k, f, v)] = ownb;
            sAx[Y::tr(el2, k, f, v)] = R(0.5) * qxb[v];
            sAy[Y::tr(el2, k, f, v)] = R(0.5) * qyb[v];
            if (inside) {
                sNu[Y::tr(nl, L.nbr_edge, oslot, v)] = nb[v];
                sNx[Y::tr(nl, L.nbr_edge, oslot, v)] = nb[NV + v];
                sNy[Y::tr(nl, L.nbr_edge, oslot, v)] = nb[2 * NV + v];
            }
        }
        if (!inside) {
            R* dst = a.Tb_out + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + oslot) * TW;
            store4(dst, nb);
            store4(dst + NV, nb + NV);
            store4(dst + 2 * NV, nb + 2 * NV);
        }
    }
    __syncthreads();

    // ---- R5: transposed traces of U and grad U; the gradient's cotangent onto
    // the reference derivatives
    if (active) {
        // a partner's slot is read only where the partner is in this tile
        // (elsewhere it holds no value: selected, never multiplied by 0)
        const bool in0 = sIn[el * 4 + 0], in1 = sIn[el * 4 + 1], in2 = sIn[el * 4 + 2], in3 = sIn[el * 4 + 3];
        auto both_t = [&](const R* A, const R* N, int v) -> R {
            const int s0 = Y::tr(el, 0, i, v), s1 = Y::tr(el, 1, j, v);
            const int s2 = Y::tr(el, 2, P - i, v), s3 = Y::tr(el, 3, P - j, v);
            return sLm[j] * (A[s0] + (in0 ? N[s0] : R(0))) + sLp[j] * (A[s2] + (in2 ? N[s2] : R(0))) +
                   sLp[i] * (A[s1] + (in1 ? N[s1] : R(0))) + sLm[i] * (A[s3] + (in3 ? N[s3] : R(0)));
        };
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            qb[v] += double(both_t(sAu, sNu, v));
            const R gbx = Qxb[v] + both_t(sAx, sNx, v), gby = Qyb[v] + both_t(sAy, sNy, v);
            sW3[ub + v * Y::UV] = jx * gbx + jy * gby;
            sGx[ub + v * Y::UV] = kx * gbx + ky * gby;
        }
    }
    __syncthreads();

    if (a.mub_partial) {
        const double sm = block_reduce_n<OpSum, NTHR / 64>(musum, sRed);
        if (tid == 0) a.mub_partial[tile] += sm;  // summed over the launches since the seed
    }

    // ---- R6: transposed gradient operator; cotangent of q
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R* X = sW3 + el * Y::UE + v * Y::UV;
            const R* H = sGx + el * Y::UE + v * Y::UV;
            R acc = R(0);
#pragma unroll
            for (int m = 0; m < N1; ++m) acc += sD[m * N1 + i] * X[j * N1 + m] + sD[m * N1 + j] * H[m * N1 + i];
            double out = qb[v] + double(acc);
            if (MODE == MODE_STAGE1) out += a.U0b[v * PS + gi];
            a.Ub_out[v * PS + gi] = out;
        }
    }
}

template <typename R, int P, int EPB, int MODE, int NTHR = visc_threads<P, EPB>()>
__global__ __launch_bounds__(NTHR) void visc_vjp_kernel(const VjpArgs<R> a, const OpTables<R> ops,
                                                         const ViscParams vp) {
    visc_vjp_body<R, P, EPB, MODE, NTHR>(a, ops, vp, tile_of_block(a.xcd_groups));
}

// The BR2 stage VJP over the tiles of several meshes of one order in one
// launch: one work item (member, tile) per workgroup, each member with its own
// arguments, operator tables and viscous parameters.
template <typename R, int P, int EPB, int MODE, int NTHR = visc_threads<P, EPB>()>
__global__ __launch_bounds__(NTHR) void batched_visc_vjp_kernel(const BatchVjpInst<R>* __restrict__ insts,
                                                                 const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    visc_vjp_body<R, P, EPB, MODE, NTHR>(insts[it.x].a, insts[it.x].ops, insts[it.x].vp, it.y);
}

// Completion of a BR2 cotangent: U_bar plus the transposed traces of U and
// grad U of the cross-tile edge cotangents the last adjoint launch left.
template <typename R, int P, int EPB>
__global__ __launch_bounds__(BLOCK) void visc_complete_kernel(const CompleteArgs<R> a, const R* __restrict__ Jinv,
                                                              const OpTables<R> ops) {
    using Y = Lay<R, P>;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1, NT = EPB * S, NF = EPB * 4 * FP;
    __shared__ R sAu[EPB * Y::TE], sAx[EPB * Y::TE], sAy[EPB * Y::TE];
    __shared__ R sX[EPB * Y::UE], sH[EPB * Y::UE];
    __shared__ R sD[N1 * N1], sLm[N1], sLp[N1];
    const int tid = threadIdx.x;
    const int e0 = tile_of_block(a.xcd_groups) * EPB;
    const int64_t PS = (int64_t)a.E * S;
    for (int t = tid; t < N1 * N1; t += BLOCK) sD[t] = ops.D[t];
    if (tid < N1) { sLm[tid] = ops.Lm[tid]; sLp[tid] = ops.Lp[tid]; }
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        const int nl = L.nbr - e0;
        R tb[TW];
#pragma unroll
        for (int w = 0; w < TW; ++w) tb[w] = R(0);
        if (a.Tb && !(nl >= 0 && nl < EPB)) {
            const R* src = a.Tb + (((int64_t)e2 * 4 + k) * FP + f) * TW;
            load4(src, tb);
            load4(src + NV, tb + NV);
            load4(src + 2 * NV, tb + 2 * NV);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sAu[Y::tr(el2, k, f, v)] = tb[v];
            sAx[Y::tr(el2, k, f, v)] = tb[NV + v];
            sAy[Y::tr(el2, k, f, v)] = tb[2 * NV + v];
        }
    }
    __syncthreads();
    const int el = tid / S, sp = tid - (tid / S) * S;
    const int e = e0 + el;
    const bool active = tid < NT && e < a.E;
    const int ub = el * Y::UE + sp;
    const int i = sp % N1, j = sp / N1;
    auto trace_t = [&](const R* A, int v) -> R {
        return sLm[j] * A[Y::tr(el, 0, i, v)] + sLp[j] * A[Y::tr(el, 2, P - i, v)] +
               sLp[i] * A[Y::tr(el, 1, j, v)] + sLm[i] * A[Y::tr(el, 3, P - j, v)];
    };
    if (active) {
        const R jx = Jinv[(int64_t)e * 4], jy = Jinv[(int64_t)e * 4 + 1];
        const R kx = Jinv[(int64_t)e * 4 + 2], ky = Jinv[(int64_t)e * 4 + 3];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R gbx = trace_t(sAx, v), gby = trace_t(sAy, v);
            sX[ub + v * Y::UV] = jx * gbx + jy * gby;
            sH[ub + v * Y::UV] = kx * gbx + ky * gby;
        }
    }
    __syncthreads();
    if (!active) return;
    const int64_t gi = (int64_t)e * S + sp;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const R* X = sX + el * Y::UE + v * Y::UV;
        const R* H = sH + el * Y::UE + v * Y::UV;
        R acc = trace_t(sAu, v);
#pragma unroll
        for (int m = 0; m < N1; ++m) acc += sD[m * N1 + i] * X[j * N1 + m] + sD[m * N1 + j] * H[m * N1 + i];
        a.out[v * PS + gi] = a.Ub[v * PS + gi] + double(acc);
    }
}

template <typename R, int P>
hipError_t visc_adj_stage_p(int mode, const StageArgs<R>& a, const OpTables<R>& ops, const ViscParams& vp,
                            hipStream_t s) {
    constexpr int EPB = visc_adj_epb<R, P>(), NTHR = visc_threads<P, EPB>();
    const int grid = (std::max(a.E, 1) + EPB - 1) / EPB;
    switch (mode) {
    case MODE_STAGE1: klaunch(visc_stage_kernel<R, P, EPB, MODE_STAGE1>, grid, dim3(NTHR), 0, s, a, ops, vp); break;
    case MODE_STAGE2: klaunch(visc_stage_kernel<R, P, EPB, MODE_STAGE2>, grid, dim3(NTHR), 0, s, a, ops, vp); break;
    case MODE_STAGE3: klaunch(visc_stage_kernel<R, P, EPB, MODE_STAGE3>, grid, dim3(NTHR), 0, s, a, ops, vp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, int P>
hipError_t visc_vjp_p(int mode, const VjpArgs<R>& a, const OpTables<R>& ops, const ViscParams& vp, hipStream_t s) {
    constexpr int EPB = visc_adj_epb<R, P>(), NTHR = visc_threads<P, EPB>();
    const int grid = (std::max(a.E, 1) + EPB - 1) / EPB;
    switch (mode) {
    case MODE_RESIDUAL: visc_vjp_kernel<R, P, EPB, MODE_RESIDUAL><<<grid, NTHR, 0, s>>>(a, ops, vp); break;
    case MODE_STAGE1: visc_vjp_kernel<R, P, EPB, MODE_STAGE1><<<grid, NTHR, 0, s>>>(a, ops, vp); break;
    case MODE_STAGE2: visc_vjp_kernel<R, P, EPB, MODE_STAGE2><<<grid, NTHR, 0, s>>>(a, ops, vp); break;
    default: visc_vjp_kernel<R, P, EPB, MODE_STAGE3><<<grid, NTHR, 0, s>>>(a, ops, vp); break;
    }
    return hipGetLastError();
}

template <typename R, int P>
hipError_t batched_visc_vjp_p(int mode, const BatchVjpInst<R>* insts, const int2* items, int n_items, hipStream_t s) {
    constexpr int EPB = visc_adj_epb<R, P>(), NTHR = visc_threads<P, EPB>();
    switch (mode) {
    case MODE_STAGE1: batched_visc_vjp_kernel<R, P, EPB, MODE_STAGE1><<<n_items, NTHR, 0, s>>>(insts, items); break;
    case MODE_STAGE2: batched_visc_vjp_kernel<R, P, EPB, MODE_STAGE2><<<n_items, NTHR, 0, s>>>(insts, items); break;
    case MODE_STAGE3: batched_visc_vjp_kernel<R, P, EPB, MODE_STAGE3><<<n_items, NTHR, 0, s>>>(insts, items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, int P>
hipError_t visc_complete_p(const CompleteArgs<R>& a, const R* Jinv, const OpTables<R>& ops, hipStream_t s) {
    constexpr int EPB = visc_adj_epb<R, P>();
    const int grid = (std::max(a.E, 1) + EPB - 1) / EPB;
    visc_complete_kernel<R, P, EPB><<<grid, BLOCK, 0, s>>>(a, Jinv, ops);
    return hipGetLastError();
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

int visc_adjoint_tile(int P, bool fp64) {
    switch (P) {
    case 1: return fp64 ? visc_adj_epb<double, 1>() : visc_adj_epb<float, 1>();
    case 2: return fp64 ? visc_adj_epb<double, 2>() : visc_adj_epb<float, 2>();
    case 3: return fp64 ? visc_adj_epb<double, 3>() : visc_adj_epb<float, 3>();
    case 4: return fp64 ? visc_adj_epb<double, 4>() : visc_adj_epb<float, 4>();
    case 5: return fp64 ? visc_adj_epb<double, 5>() : visc_adj_epb<float, 5>();
    default: return 0;
    }
}

#define ZJ_VISC_P(call)                 \
    switch (P) {                         \
    case 1: return call(1);              \
    case 2: return call(2);              \
    case 3: return call(3);              \
    case 4: return call(4);              \
    case 5: return call(5);              \
    default: return hipErrorInvalidValue; \
    }

template <typename R>
hipError_t launch_visc_stage_adj(int P, int mode, const StageArgs<R>& a, const OpTables<R>& ops,
                                 const ViscParams& vp, hipStream_t s) {
#define ZJ_C(PP) visc_adj_stage_p<R, PP>(mode, a, ops, vp, s)
    ZJ_VISC_P(ZJ_C)
#undef ZJ_C
}

template <typename R>
hipError_t launch_visc_vjp(int P, int mode, const VjpArgs<R>& a, const OpTables<R>& ops, const ViscParams& vp,
                           hipStream_t s) {
    if (a.E <= 0) return hipSuccess;
#define ZJ_C(PP) visc_vjp_p<R, PP>(mode, a, ops, vp, s)
    ZJ_VISC_P(ZJ_C)
#undef ZJ_C
}

template <typename R>
hipError_t launch_batched_visc_vjp(int P, int mode, const BatchVjpInst<R>* insts, const int2* items, int n_items,
                                   hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
#define ZJ_C(PP) batched_visc_vjp_p<R, PP>(mode, insts, items, n_items, s)
    ZJ_VISC_P(ZJ_C)
#undef ZJ_C
}

template <typename R>
hipError_t launch_visc_complete(int P, const CompleteArgs<R>& a, const R* Jinv_e, const OpTables<R>& ops,
                                hipStream_t s) {
    if (a.E <= 0) return hipSuccess;
#define ZJ_C(PP) visc_complete_p<R, PP>(a, Jinv_e, ops, s)
    ZJ_VISC_P(ZJ_C)
#undef ZJ_C
}
#undef ZJ_VISC_P

template hipError_t launch_visc_stage_adj<double>(int, int, const StageArgs<double>&, const OpTables<double>&,
                                                  const ViscParams&, hipStream_t);
template hipError_t launch_visc_stage_adj<float>(int, int, const StageArgs<float>&, const OpTables<float>&,
                                                 const ViscParams&, hipStream_t);
template hipError_t launch_visc_vjp<double>(int, int, const VjpArgs<double>&, const OpTables<double>&,
                                            const ViscParams&, hipStream_t);
template hipError_t launch_visc_vjp<float>(int, int, const VjpArgs<float>&, const OpTables<float>&,
                                           const ViscParams&, hipStream_t);
template hipError_t launch_batched_visc_vjp<double>(int, int, const BatchVjpInst<double>*, const int2*, int,
                                                    hipStream_t);
template hipError_t launch_batched_visc_vjp<float>(int, int, const BatchVjpInst<float>*, const int2*, int,
                                                   hipStream_t);
template hipError_t launch_visc_complete<double>(int, const CompleteArgs<double>&, const double*,
                                                 const OpTables<double>&, hipStream_t);
template hipError_t launch_visc_complete<float>(int, const CompleteArgs<float>&, const float*,
                                                const OpTables<float>&, hipStream_t);

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

template <typename R, int P>
hipError_t batched_visc_p(int mode, const BatchInst<R>* insts, const int2* items, int n_items, int grid,
                          hipStream_t s) {
    constexpr int EPB = BLOCK / ((P + 1) * (P + 1));
    const dim3 g(std::max(1, std::min(grid, n_items))), b(BLOCK);
    switch (mode) {
    case MODE_STAGE1: klaunch(batched_visc_kernel<R, P, EPB, MODE_STAGE1>, g, b, 0, s, insts, items, n_items); break;
    case MODE_STAGE2: klaunch(batched_visc_kernel<R, P, EPB, MODE_STAGE2>, g, b, 0, s, insts, items, n_items); break;
    case MODE_STAGE3: klaunch(batched_visc_kernel<R, P, EPB, MODE_STAGE3>, g, b, 0, s, insts, items, n_items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
hipError_t launch_batched_visc(int P, int mode, const BatchInst<R>* insts, const int2* items, int n_items, int grid,
                               hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    switch (P) {
    case 1: return batched_visc_p<R, 1>(mode, insts, items, n_items, grid, s);
    case 2: return batched_visc_p<R, 2>(mode, insts, items, n_items, grid, s);
    case 3: return batched_visc_p<R, 3>(mode, insts, items, n_items, grid, s);
    case 4: return batched_visc_p<R, 4>(mode, insts, items, n_items, grid, s);
    case 5: return batched_visc_p<R, 5>(mode, insts, items, n_items, grid, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R>
hipError_t launch_batched_visc_all(int mode, const BatchInst<R>* insts, const int2* items, int n_items, int grid,
                                   int lds_bytes, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    const dim3 g(std::max(1, std::min(grid, n_items))), b(BLOCK);
    const unsigned lds = (unsigned)lds_bytes;
    switch (mode) {
    case MODE_STAGE1: klaunch(batched_visc_all_kernel<R, MODE_STAGE1>, g, b, lds, s, insts, items, n_items); break;
    case MODE_STAGE2: klaunch(batched_visc_all_kernel<R, MODE_STAGE2>, g, b, lds, s, insts, items, n_items); break;
    case MODE_STAGE3: klaunch(batched_visc_all_kernel<R, MODE_STAGE3>, g, b, lds, s, insts, items, n_items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int visc_all_lds(int P, bool fp32) {
    switch (P) {
#define ZJ_VL(PP)                                                                       \
    case PP: {                                                                          \
        constexpr int EPB = BLOCK / ((PP + 1) * (PP + 1));                              \
        return fp32 ? ViscSmemDyn<float, PP, EPB>::bytes : ViscSmemDyn<double, PP, EPB>::bytes; \
    }
        ZJ_VL(1) ZJ_VL(2) ZJ_VL(3) ZJ_VL(4) ZJ_VL(5)
#undef ZJ_VL
    default: return 0;
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
template hipError_t launch_batched_visc_all<double>(int, const BatchInst<double>*, const int2*, int, int, int,
                                                    hipStream_t);
template hipError_t launch_batched_visc_all<float>(int, const BatchInst<float>*, const int2*, int, int, int,
                                                   hipStream_t);
template hipError_t launch_batched_visc<double>(int, int, const BatchInst<double>*, const int2*, int, int,
                                                hipStream_t);
template hipError_t launch_batched_visc<float>(int, int, const BatchInst<float>*, const int2*, int, int,
                                               hipStream_t);
template hipError_t launch_visc_trace<double>(int, int, const double*, const double*, double*,
                                              const OpTables<double>&, int, hipStream_t);
template hipError_t launch_visc_trace<float>(int, int, const double*, const float*, float*, const OpTables<float>&,
                                             int, hipStream_t);

}  // namespace zj
