// zj_device.hpp -- device-side building blocks shared by the forward stage
// kernels (zj_kernels.hip) and their adjoints (zj_adjoint.hip): precision
// helpers, the Euler physics of the reference path, wave/block reductions,
// the sum-factorised trace operator, the per-(precision, order) LDS layout
// and the XCD-aware tile order.
#pragma once

#include "zj_kernels.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

namespace zj {
namespace {

template <typename R> __device__ __forceinline__ R fmaR(R a, R b, R c);
template <> __device__ __forceinline__ double fmaR<double>(double a, double b, double c) { return fma(a, b, c); }
template <> __device__ __forceinline__ float fmaR<float>(float a, float b, float c) { return fmaf(a, b, c); }

// FP64 keeps correctly rounded square root and division (the plain form must
// track the reference to 1e-12); the FP32 bulk uses the 1-ulp hardware
// instructions, whose error sits far below the FP32 rounding of the state.
template <typename R> __device__ __forceinline__ R sqrtR(R x);
template <> __device__ __forceinline__ double sqrtR<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float sqrtR<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename R> __device__ __forceinline__ R rcpR(R x);
template <> __device__ __forceinline__ double rcpR<double>(double x) { return 1.0 / x; }
template <> __device__ __forceinline__ float rcpR<float>(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename R> __device__ __forceinline__ R fmaxR(R a, R b);
template <> __device__ __forceinline__ double fmaxR<double>(double a, double b) { return fmax(a, b); }
template <> __device__ __forceinline__ float fmaxR<float>(float a, float b) { return fmaxf(a, b); }
template <typename R> __device__ __forceinline__ R fminR(R a, R b);
template <> __device__ __forceinline__ double fminR<double>(double a, double b) { return fmin(a, b); }
template <> __device__ __forceinline__ float fminR<float>(float a, float b) { return fminf(a, b); }

// 4-wide loads and stores of one state (vector memory ops).
__device__ __forceinline__ void load4(const double* p, double* v) {
    const double2 a = reinterpret_cast<const double2*>(p)[0];
    const double2 b = reinterpret_cast<const double2*>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void load4(const float* p, float* v) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void store4(double* p, const double* v) {
    reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void store4(float* p, const float* v) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}

// ---------------------------------------------------------------- physics
// IdealGas::fluxX/fluxY (euler.hpp:101-129), no clamps.
template <typename R>
__device__ __forceinline__ void euler_flux(const R* U, R g, R* Fx, R* Fy) {
    const R inv = rcpR(U[0]);
    const R u = U[1] * inv, v = U[2] * inv, E = U[3] * inv;
    const R p = (g - R(1)) * U[0] * (E - R(0.5) * (u * u + v * v));
    Fx[0] = U[1]; Fx[1] = U[1] * u + p; Fx[2] = U[1] * v; Fx[3] = (U[3] + p) * u;
    Fy[0] = U[2]; Fy[1] = U[2] * u; Fy[2] = U[2] * v + p; Fy[3] = (U[3] + p) * v;
}

template <typename R>
__device__ __forceinline__ void normal_flux(const R* U, R g, R nx, R ny, R* Fn) {
    R fx[NV], fy[NV];
    euler_flux(U, g, fx, fy);
#pragma unroll
    for (int k = 0; k < NV; ++k) Fn[k] = fx[k] * nx + fy[k] * ny;
}

// Inline clamped Rusanov of computeRiemannFluxWithBCKernel (kernels.cu:601-668).
template <typename R>
__device__ __forceinline__ void rusanov(const R* UL, const R* UR, R nx, R ny, R g, R* F) {
    const R tiny = R(1e-10);
    const R rhoL = fmaxR(UL[0], tiny), rhoR = fmaxR(UR[0], tiny);
    const R iL = rcpR(rhoL), iR = rcpR(rhoR);
    const R uL = UL[1] * iL, vL = UL[2] * iL, uR = UR[1] * iR, vR = UR[2] * iR;
    const R EL = UL[3] * iL, ER = UR[3] * iR;
    const R qL = uL * uL + vL * vL, qR = uR * uR + vR * vR;
    const R pL = fmaxR((g - R(1)) * rhoL * (EL - R(0.5) * qL), tiny);
    const R pR = fmaxR((g - R(1)) * rhoR * (ER - R(0.5) * qR), tiny);
    const R cL = sqrtR(fmaxR(g * pL * iL, tiny)), cR = sqrtR(fmaxR(g * pR * iR, tiny));
    const R vnL = uL * nx + vL * ny, vnR = uR * nx + vR * ny;
    const R smax = fmaxR(sqrtR(qL) + cL, sqrtR(qR) + cR);
    const R HL = EL + pL * iL, HR = ER + pR * iR;
    const R FL[NV] = {rhoL * vnL, rhoL * uL * vnL + pL * nx, rhoL * vL * vnL + pL * ny, rhoL * HL * vnL};
    const R FR[NV] = {rhoR * vnR, rhoR * uR * vnR + pR * nx, rhoR * vR * vnR + pR * ny, rhoR * HR * vnR};
    const R hs = R(0.5) * smax;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        R f = R(0.5) * (FL[k] + FR[k]) - hs * (UR[k] - UL[k]);
        F[k] = isfinite(f) ? f : R(0);
    }
}

template <typename R>
__device__ __forceinline__ void prim2cons(R rho, R u, R v, R p, R g, R* U) {
    U[0] = rho; U[1] = rho * u; U[2] = rho * v;
    U[3] = p / (g - R(1)) + R(0.5) * rho * (u * u + v * v);
}

// computeGhostState (kernels.cu:453-525).
template <typename R>
__device__ void ghost_state(const R* Ui, R nx, R ny, int bct, const double* bc, R g, R* G) {
    if (bct == 1 || bct == 2 || bct == 3) {  // Wall, SlipWall, Symmetry: reflect
        const R rho = Ui[0], u = Ui[1] / rho, v = Ui[2] / rho;
        const R p = (g - R(1)) * rho * (Ui[3] / rho - R(0.5) * (u * u + v * v));
        const R vn = u * nx + v * ny;
        prim2cons(rho, u - R(2) * vn * nx, v - R(2) * vn * ny, p, g, G);
    } else if (bct == 4) {  // FarField: averaged pressure with clamps
        const R rinf = R(bc[0]), uinf = R(bc[1]), vinf = R(bc[2]), pinf = R(bc[3]);
        const R ri = fmaxR(Ui[0], R(1e-10));
        const R ui = Ui[1] / ri, vi = Ui[2] / ri;
        const R pi = fmaxR((g - R(1)) * (Ui[3] - R(0.5) * ri * (ui * ui + vi * vi)), R(1e-10));
        const R vninf = uinf * nx + vinf * ny;
        R rb, ub, vb;
        if (vninf > R(0)) { rb = rinf; ub = uinf; vb = vinf; }
        else { rb = ri; ub = ui; vb = vi; }
        R pb = R(0.5) * (pinf + pi);
        rb = fmaxR(fminR(rb, R(100)), R(0.01));
        ub = fmaxR(fminR(ub, R(1000)), R(-1000));
        vb = fmaxR(fminR(vb, R(1000)), R(-1000));
        pb = fmaxR(fminR(pb, R(1e7)), R(1e3));
        prim2cons(rb, ub, vb, pb, g, G);
    } else if (bct == 6) {  // Outflow: prescribed static pressure
        const R rho = Ui[0], u = Ui[1] / rho, v = Ui[2] / rho;
        prim2cons(rho, u, v, R(bc[0]), g, G);
    } else {
#pragma unroll
        for (int k = 0; k < NV; ++k) G[k] = Ui[k];
    }
}

// ------------------------------------------- textbook-mode Riemann solvers
// The solvers of the reference's euler.hpp with every quotient a / b taken as
// a * (1/b), each reciprocal computed once: FP64 a Newton-refined hardware
// reciprocal (within 2 ulp of 1/b), FP32 the hardware reciprocal.  An IEEE
// division costs about ten VALU instructions in either precision.
__device__ __forceinline__ double rcp_refined(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
template <typename R> __device__ __forceinline__ R recipR(R x);
template <> __device__ __forceinline__ double recipR<double>(double x) { return rcp_refined(x); }
template <> __device__ __forceinline__ float recipR<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// One side of a face: floored density rho = max(U0, 1e-10) and 1/rho, 1/U0
// (the reference divides by the unfloored density in pressure and fluxNormal),
// velocity from the floored density, pressure from the unfloored one
// (IdealGas::pressure, euler.hpp:16-23).
template <typename R>
struct Side {
    R rho, ir, i0, u, v, p;
};
template <typename R>
__device__ __forceinline__ Side<R> side_of(const R* U, R g) {
    Side<R> s;
    s.rho = fmaxR(U[0], R(1e-10));
    s.ir = recipR(s.rho);
    s.i0 = U[0] == s.rho ? s.ir : recipR(U[0]);
    s.u = U[1] * s.ir;
    s.v = U[2] * s.ir;
    const R u0 = U[1] * s.i0, v0 = U[2] * s.i0;
    s.p = (g - R(1)) * U[0] * (U[3] * s.i0 - R(0.5) * (u0 * u0 + v0 * v0));
    return s;
}
// IdealGas::fluxNormal with its floors (euler.hpp:133-147).
template <typename R>
__device__ __forceinline__ void flux_normal_floor(const R* U, const Side<R>& s, R nx, R ny, R* Fn) {
    const R p = fmaxR(s.p, R(1e-10));
    const R vn = s.u * nx + s.v * ny;
    Fn[0] = s.rho * vn;
    Fn[1] = s.rho * s.u * vn + p * nx;
    Fn[2] = s.rho * s.v * vn + p * ny;
    Fn[3] = (U[3] + s.p) * vn;
}
template <typename R>
__device__ __forceinline__ void flux_normal_floor(const R* U, R nx, R ny, R g, R* Fn) {
    flux_normal_floor(U, side_of(U, g), nx, ny, Fn);
}
// IdealGas::soundSpeed (euler.hpp:32-41)
template <typename R>
__device__ __forceinline__ R sound_speed_of(const Side<R>& s, R g) {
    return sqrtR(g * fmaxR(s.p, R(1e-10)) * s.ir);
}
// RiemannSolvers::rusanov (euler.hpp:184-200); its velocities use the unfloored
// density.  One side at a time: F holds F_L until F_R is known.
template <typename R>
__device__ __forceinline__ R rusanov_side(const R* U, R nx, R ny, R g, R* Fn) {
    const Side<R> s = side_of(U, g);
    flux_normal_floor(U, s, nx, ny, Fn);
    const R u = U[1] * s.i0, v = U[2] * s.i0;
    return sqrtR(u * u + v * v) + sound_speed_of(s, g);
}
template <typename R>
__device__ __forceinline__ void riemann_rusanov(const R* UL, const R* UR, R nx, R ny, R g, R* F) {
    const R sL = rusanov_side(UL, nx, ny, g, F);
    R FR[NV];
    const R sR = rusanov_side(UR, nx, ny, g, FR);
    const R smax = fmaxR(sL, sR);
#pragma unroll
    for (int i = 0; i < NV; ++i) F[i] = R(0.5) * (F[i] + FR[i]) - R(0.5) * smax * (UR[i] - UL[i]);
}
// RiemannSolvers::roe with Harten's entropy fix (euler.hpp:204-294).  The
// average of the two sides' normal fluxes goes into F first, so that neither
// side's state stays live through the wave decomposition.
template <typename R>
__device__ __forceinline__ void riemann_roe(const R* UL, const R* UR, R nx, R ny, R g, R* F) {
    R rhoL, rhoR, uL, vL, uR, vR, pL, pR, HL, HR;
    {
        const Side<R> L = side_of(UL, g);
        flux_normal_floor(UL, L, nx, ny, F);
        rhoL = L.rho; uL = L.u; vL = L.v; pL = L.p;
        HL = (UL[3] + pL) * L.ir;
    }
    {
        const Side<R> Rt = side_of(UR, g);
        R FR[NV];
        flux_normal_floor(UR, Rt, nx, ny, FR);
#pragma unroll
        for (int i = 0; i < NV; ++i) F[i] = R(0.5) * (F[i] + FR[i]);
        rhoR = Rt.rho; uR = Rt.u; vR = Rt.v; pR = Rt.p;
        HR = (UR[3] + pR) * Rt.ir;
    }
    const R sL = sqrtR(rhoL), sR = sqrtR(rhoR);
    const R iden = recipR(fmaxR(sL + sR, R(1e-10)));
    const R rr = sL * sR;
    const R ur = (sL * uL + sR * uR) * iden, vr = (sL * vL + sR * vR) * iden, Hr = (sL * HL + sR * HR) * iden;
    const R vnr = ur * nx + vr * ny, q2 = ur * ur + vr * vr;
    const R c = sqrtR(fmaxR((g - R(1)) * (Hr - R(0.5) * q2), R(1e-10)));
    R l1 = fabs(vnr - c), l4 = fabs(vnr + c);
    const R l2 = fabs(vnr);
    const R eps = R(0.1) * c;
    if (l1 < eps || l4 < eps) {
        const R i2e = recipR(R(2) * eps);
        if (l1 < eps) l1 = (l1 * l1 + eps * eps) * i2e;
        if (l4 < eps) l4 = (l4 * l4 + eps * eps) * i2e;
    }
    const R du = uR - uL, dv = vR - vL, dp = pR - pL, dvn = du * nx + dv * ny;
    const R ic2 = recipR(fmaxR(c * c, R(1e-10)));
    const R i2c2 = R(0.5) * ic2;
    const R a1 = (dp - rr * c * dvn) * i2c2;
    const R a2 = (rhoR - rhoL) - dp * ic2;
    const R a3 = rr * (dv * nx - du * ny);
    const R a4 = (dp + rr * c * dvn) * i2c2;
    const R tx = -ny, ty = nx;
    const R w1 = l1 * a1, w2 = l2 * a2, w3 = l2 * a3, w4 = l4 * a4;
    const R dF[NV] = {w1 + w2 + w4,
                      w1 * (ur - c * nx) + w2 * ur + w3 * tx + w4 * (ur + c * nx),
                      w1 * (vr - c * ny) + w2 * vr + w3 * ty + w4 * (vr + c * ny),
                      w1 * (Hr - c * vnr) + w2 * R(0.5) * q2 + w3 * (ur * tx + vr * ty) + w4 * (Hr + c * vnr)};
#pragma unroll
    for (int i = 0; i < NV; ++i) F[i] = F[i] - R(0.5) * dF[i];
}
template <typename R>
__device__ __forceinline__ R guard_den(R d) {
    return fabs(d) < R(1e-10) ? (d >= R(0) ? R(1e-10) : R(-1e-10)) : d;
}
// RiemannSolvers::hllc (euler.hpp:298-375).  The wave estimates need a few
// scalars of each side; the side whose flux is taken is rebuilt from its state
// afterwards instead of keeping both sides' states live.
template <typename R>
__device__ __forceinline__ void riemann_hllc(const R* UL, const R* UR, R nx, R ny, R g, R* F) {
    R rhoL, rhoR, vnL, vnR, pL, pR, cL, cR;
    {
        const Side<R> L = side_of(UL, g);
        rhoL = L.rho; pL = fmaxR(L.p, R(1e-10)); cL = sound_speed_of(L, g); vnL = L.u * nx + L.v * ny;
    }
    {
        const Side<R> Rt = side_of(UR, g);
        rhoR = Rt.rho; pR = fmaxR(Rt.p, R(1e-10)); cR = sound_speed_of(Rt, g); vnR = Rt.u * nx + Rt.v * ny;
    }
    R ps = R(0.5) * (pL + pR) - R(0.5) * (vnR - vnL) * R(0.5) * (rhoL + rhoR) * R(0.5) * (cL + cR);
    ps = fmaxR(ps, R(1e-10));
    const R gp = (g + R(1)) * recipR(R(2) * g);
    const R qL = ps <= pL ? R(1) : sqrtR(R(1) + gp * (ps * recipR(pL) - R(1)));
    const R qR = ps <= pR ? R(1) : sqrtR(R(1) + gp * (ps * recipR(pR) - R(1)));
    const R SL = vnL - cL * qL, SR = vnR + cR * qR;
    const R dSM = guard_den(rhoL * (SL - vnL) - rhoR * (SR - vnR));
    const R SM = (pR - pL + rhoL * vnL * (SL - vnL) - rhoR * vnR * (SR - vnR)) * recipR(dSM);
    // the upwind side: left if SL >= 0, right if SR <= 0, else the side of the contact
    const bool plain = SL >= R(0) || SR <= R(0);
    const bool left = plain ? SL >= R(0) : SM >= R(0);
    const R* U = left ? UL : UR;
    const Side<R> K = side_of(U, g);
    R Fk[NV];
    flux_normal_floor(U, K, nx, ny, Fk);
    if (plain) {
#pragma unroll
        for (int i = 0; i < NV; ++i) F[i] = Fk[i];
        return;
    }
    const R vn = left ? vnL : vnR, S = left ? SL : SR, p = left ? pL : pR;
    const R rho = K.rho, u = K.u, v = K.v;
    const R dK = guard_den(S - SM), dvn = guard_den(S - vn);
    const R fac = rho * dvn * recipR(dK);
    const R Us[NV] = {fac, fac * (u + (SM - vn) * nx), fac * (v + (SM - vn) * ny),
                      fac * (U[3] * K.ir + (SM - vn) * (SM + p * recipR(rho * dvn)))};
#pragma unroll
    for (int i = 0; i < NV; ++i) F[i] = Fk[i] + S * (Us[i] - U[i]);
}
template <typename R>
__device__ __forceinline__ void riemann_by_kind(int kind, const R* UL, const R* UR, R nx, R ny, R g, R* F) {
    if (kind == 1) riemann_roe(UL, UR, nx, ny, g, F);
    else if (kind == 2) riemann_hllc(UL, UR, nx, ny, g, F);
    else riemann_rusanov(UL, UR, nx, ny, g, F);
}

// enforcePositivityKernel (kernels.cu:1236-1288), FP64, gamma fixed at 1.4.
// FAST (the FP32-bulk form): the two divisions become a Newton-refined
// hardware reciprocal and a multiplication by 1/(gamma-1), within 2 ulp of
// the divisions, far below the form's FP32 rounding.
template <bool FAST = false>
__device__ __forceinline__ void positivity(double* q) {
    constexpr double GM1 = 1.4 - 1.0;
    const double rn = fmax(0.01, fmin(100.0, q[0]));
    const double ir = FAST ? rcp_refined(fmax(rn, 1e-10)) : 1.0 / fmax(rn, 1e-10);
    double u = q[1] * ir, v = q[2] * ir;
    // |v| > 2000 needs the square root only when |v|^2 > 4e6 (sqrt is
    // correctly rounded and monotone, so below that it cannot exceed 2000)
    const double s2 = u * u + v * v;
    if (s2 > 4.0e6) {
        const double vm = sqrt(s2);
        if (vm > 2000.0) {
            const double sc = 2000.0 / vm;
            u *= sc; v *= sc;
        }
    }
    const double ke = 0.5 * rn * (u * u + v * v);
    const double p = fmax(100.0, fmin(1e8, GM1 * (q[3] - ke)));
    q[0] = rn; q[1] = rn * u; q[2] = rn * v;
    q[3] = (FAST ? p * (1.0 / GM1) : p / GM1) + ke;
}

// ------------------------------------------------------- wave reductions
// Cross-lane moves through DPP on the two halves of a double.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double read_lane(double x, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

struct OpSum { __device__ static double f(double a, double b) { return a + b; } };
struct OpMin { __device__ static double f(double a, double b) { return fmin(a, b); } };

// Whole-wavefront reduction: butterfly within rows of 16 lanes by DPP
// (quad_perm 1032, quad_perm 2301, row_half_mirror, row_mirror), then the four
// row results by readlane.  Every lane of the wave must call it.
template <typename Op>
__device__ __forceinline__ double wave_reduce(double v) {
    v = Op::f(v, dpp_mov<0xB1>(v));
    v = Op::f(v, dpp_mov<0x4E>(v));
    v = Op::f(v, dpp_mov<0x141>(v));
    v = Op::f(v, dpp_mov<0x140>(v));
    return Op::f(Op::f(read_lane(v, 0), read_lane(v, 16)), Op::f(read_lane(v, 32), read_lane(v, 48)));
}

// Block reduction over NW waves; the result is valid in thread 0.
template <typename Op, int NW>
__device__ __forceinline__ double block_reduce_n(double v, double* sred) {
    v = wave_reduce<Op>(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sred[w] = v;
    __syncthreads();
    double r = sred[0];
    if (threadIdx.x == 0)
        for (int k = 1; k < NW; ++k) r = Op::f(r, sred[k]);
    return r;
}
// ... over BLOCK threads.
template <typename Op>
__device__ __forceinline__ double block_reduce(double v, double* sred) {
    return block_reduce_n<Op, BLOCK / 64>(v, sred);
}

// ------------------------------------------------------------ operators
// Value of one variable at flux point f of local edge k, from the solution
// points of one element (u[S], xi index fastest).  Edge 0 (eta=-1) and 2
// (eta=+1) run along xi, edges 1 (xi=+1) and 3 (xi=-1) along eta; edges 2 and
// 3 are stored in reversed order (basis.cpp:540-552).  The explicit FMA chain
// fixes the rounding, so both elements of a face see identical traces.
// Where flux point f of edge k reads its line of solution points: first point
// `base`, step `stride`, weights Lm (edges 0, 3) or Lp (edges 1, 2).  Computed
// once per thread without branches.
struct TraceLine {
    int base, stride, plus;
};
template <int P>
__device__ __forceinline__ TraceLine trace_line(int k, int f) {
    constexpr int N1 = P + 1;
    const int ff = k >= 2 ? P - f : f;
    const bool along_xi = (k & 1) != 0;
    return TraceLine{along_xi ? ff * N1 : ff, along_xi ? 1 : N1, (k == 1 || k == 2) ? 1 : 0};
}
template <typename R, int P>
__device__ __forceinline__ R trace_point(const R* u, const TraceLine& t, const R* Lm, const R* Lp) {
    constexpr int N1 = P + 1;
    const R* L = t.plus ? Lp : Lm;
    R acc = L[0] * u[t.base];
#pragma unroll
    for (int m = 1; m < N1; ++m) acc = fmaR(L[m], u[t.base + m * t.stride], acc);
    return acc;
}

// ---------------------------------------------------------------- kernels
// LDS layout of one element tile, chosen per (precision, order) to minimise
// bank conflicts over the kernel's access patterns (32 banks x 4 B):
// solution-point arrays are [element][variable][point] with padded strides;
// trace and flux-difference arrays are either variable-major
// ([element][variable][edge*Fp+point]) or point-major with a padded point
// stride ([element][edge*Fp+point][variable]).
template <typename R, int P> struct Lay;
#define ZJ_LAYOUT(R_, P_, PAD_V, PAD_E, VAR_MAJOR, PAD_T)                                     \
    template <> struct Lay<R_, P_> {                                                           \
        static constexpr int N1 = P_ + 1, S = N1 * N1, FP = N1;                                \
        static constexpr int UV = S + PAD_V, UE = NV * UV + PAD_E;                             \
        static constexpr bool VM = VAR_MAJOR;                                                  \
        static constexpr int TE = (VAR_MAJOR ? 4 * FP * NV : 4 * FP * (NV + 1)) + PAD_T;       \
        __device__ static constexpr int tr(int el, int k, int f, int v) {                      \
            return VAR_MAJOR ? el * TE + v * (4 * FP) + k * FP + f : el * TE + (k * FP + f) * (NV + 1) + v; \
        }                                                                                      \
    };
ZJ_LAYOUT(double, 1, 0, 4, false, 0)
ZJ_LAYOUT(double, 2, 0, 5, false, 0)
ZJ_LAYOUT(double, 3, 0, 0, true, 0)
ZJ_LAYOUT(double, 4, 0, 5, true, 5)
ZJ_LAYOUT(double, 5, 0, 7, false, 0)
ZJ_LAYOUT(float, 1, 0, 4, false, 0)
ZJ_LAYOUT(float, 2, 0, 5, false, 0)
ZJ_LAYOUT(float, 3, 3, 4, false, 0)
ZJ_LAYOUT(float, 4, 1, 5, true, 4)
ZJ_LAYOUT(float, 5, 0, 0, false, 0)
#undef ZJ_LAYOUT

// Workgroups are handed to the XCDs round-robin (workgroup b runs on XCD
// b mod 8), and each XCD has its own L2.  Giving each XCD a contiguous run of
// tiles keeps a tile's mesh neighbours, whose traces and state lines it
// shares, in the same L2.  `groups` = 1 leaves the order unchanged.
__device__ __forceinline__ int tile_of_block(int groups) {
    const int nb = gridDim.x, b = blockIdx.x;
    if (groups <= 1) return b;
    const int q = nb / groups, r = nb - q * groups, x = b % groups, loc = b / groups;
    return x * q + min(x, r) + loc;
}

}  // namespace
}  // namespace zj
