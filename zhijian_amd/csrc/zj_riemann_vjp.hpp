// zj_riemann_vjp.hpp -- vector-Jacobian products of the textbook-mode Riemann
// solvers of zj_device.hpp (RiemannSolvers::rusanov / roe / hllc,
// euler.hpp:184-375): given the cotangent Fb of the common flux F(U_L, U_R, n),
// add (dF/dU_L)^T Fb to ULb and (dF/dU_R)^T Fb to URb.
//
// Each VJP recomputes the forward intermediates the way the forward solver
// forms them (floors, guarded denominators, Harten's entropy fix) and runs
// them backwards; every max, floor, guard and branch takes the side the
// forward took, so the result is the derivative of the executed arithmetic.
// The Euler flux and ghost-state VJPs shared by the quadrilateral and
// triangle adjoints sit here too.
#pragma once

#include "zj_device.hpp"

namespace zj {
namespace {

// Cotangents of the quantities a side contributes: rho (floored density), u,
// v (velocity from the floored density), p (unfloored pressure), u0, v0
// (velocity from the unfloored density), and U3 directly.
template <typename R>
struct SideBar {
    R rho = R(0), u = R(0), v = R(0), p = R(0), u0 = R(0), v0 = R(0), U3 = R(0);
};

// side_of backwards: p = (g-1) U0 (U3 i0 - (u0^2 + v0^2)/2), u0 = U1 i0,
// u = U1 / rho, rho = max(U0, 1e-10).
template <typename R>
__device__ __forceinline__ void side_vjp(const R* U, const Side<R>& s, R g, const SideBar<R>& b, R* Ub) {
    const R gm = g - R(1);
    const R u0 = U[1] * s.i0, v0 = U[2] * s.i0;
    const R u0b = b.u0 - b.p * gm * U[0] * u0;
    const R v0b = b.v0 - b.p * gm * U[0] * v0;
    Ub[3] += b.p * gm + b.U3;
    Ub[0] += -b.p * gm * R(0.5) * (u0 * u0 + v0 * v0);
    Ub[1] += u0b * s.i0 + b.u * s.ir;
    Ub[2] += v0b * s.i0 + b.v * s.ir;
    Ub[0] += -(u0b * u0 + v0b * v0) * s.i0;
    const R rb = b.rho - (b.u * s.u + b.v * s.v) * s.ir;
    if (U[0] >= R(1e-10)) Ub[0] += rb;
}

// flux_normal_floor backwards: Fn = [rho vn, rho u vn + pf nx, rho v vn + pf ny,
// (U3 + p) vn], pf = max(p, 1e-10).
template <typename R>
__device__ __forceinline__ void flux_normal_floor_vjp(const R* U, const Side<R>& s, R nx, R ny, const R* Fb,
                                                      SideBar<R>& b) {
    const R vn = s.u * nx + s.v * ny;
    const R vnb = s.rho * (Fb[0] + Fb[1] * s.u + Fb[2] * s.v) + Fb[3] * (U[3] + s.p);
    b.rho += vn * (Fb[0] + Fb[1] * s.u + Fb[2] * s.v);
    b.u += Fb[1] * s.rho * vn + vnb * nx;
    b.v += Fb[2] * s.rho * vn + vnb * ny;
    if (s.p >= R(1e-10)) b.p += Fb[1] * nx + Fb[2] * ny;
    b.p += Fb[3] * vn;
    b.U3 += Fb[3] * vn;
}

// sound_speed_of backwards: c = sqrt(g max(p, 1e-10) / rho).
template <typename R>
__device__ __forceinline__ void sound_speed_vjp(const Side<R>& s, R g, R c, R cb, SideBar<R>& b) {
    const R c2b = cb * R(0.5) * recipR(c);
    if (s.p >= R(1e-10)) b.p += c2b * g * s.ir;
    b.rho -= c2b * c * c * s.ir;  // d(g pf / rho)/d rho = -c^2 / rho
}

template <typename R>
__device__ __forceinline__ void riemann_rusanov_vjp(const R* UL, const R* UR, R nx, R ny, R g, const R* Fb, R* ULb,
                                                    R* URb) {
    const Side<R> L = side_of(UL, g), Q = side_of(UR, g);
    const R cL = sound_speed_of(L, g), cR = sound_speed_of(Q, g);
    const R uL0 = UL[1] * L.i0, vL0 = UL[2] * L.i0, uR0 = UR[1] * Q.i0, vR0 = UR[2] * Q.i0;
    const R mL = sqrtR(uL0 * uL0 + vL0 * vL0), mR = sqrtR(uR0 * uR0 + vR0 * vR0);
    const R sL = mL + cL, sR = mR + cR;
    const R smax = fmaxR(sL, sR);
    R smaxb = R(0);
    R Fh[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        smaxb -= R(0.5) * Fb[k] * (UR[k] - UL[k]);
        ULb[k] += R(0.5) * smax * Fb[k];
        URb[k] -= R(0.5) * smax * Fb[k];
        Fh[k] = R(0.5) * Fb[k];
    }
    const bool lmax = sL >= sR;
    SideBar<R> bl, br;
    flux_normal_floor_vjp(UL, L, nx, ny, Fh, bl);
    flux_normal_floor_vjp(UR, Q, nx, ny, Fh, br);
    const R sLb = lmax ? smaxb : R(0), sRb = lmax ? R(0) : smaxb;
    if (mL > R(0)) { const R im = recipR(mL); bl.u0 += sLb * uL0 * im; bl.v0 += sLb * vL0 * im; }
    if (mR > R(0)) { const R im = recipR(mR); br.u0 += sRb * uR0 * im; br.v0 += sRb * vR0 * im; }
    sound_speed_vjp(L, g, cL, sLb, bl);
    sound_speed_vjp(Q, g, cR, sRb, br);
    side_vjp(UL, L, g, bl, ULb);
    side_vjp(UR, Q, g, br, URb);
}

// Sign with sign(0) = 0, the derivative of |x| taken at a kink.
template <typename R>
__device__ __forceinline__ R sgnR(R x) {
    return x > R(0) ? R(1) : (x < R(0) ? R(-1) : R(0));
}

template <typename R>
__device__ __forceinline__ void riemann_roe_vjp(const R* UL, const R* UR, R nx, R ny, R g, const R* Fb, R* ULb,
                                                R* URb) {
    const R tiny = R(1e-10);
    const Side<R> L = side_of(UL, g), Q = side_of(UR, g);
    const R rhoL = L.rho, uL = L.u, vL = L.v, pL = L.p, HL = (UL[3] + pL) * L.ir;
    const R rhoR = Q.rho, uR = Q.u, vR = Q.v, pR = Q.p, HR = (UR[3] + pR) * Q.ir;
    const R sL = sqrtR(rhoL), sR = sqrtR(rhoR);
    const R ssum = sL + sR;
    const R iden = recipR(fmaxR(ssum, tiny));
    const R rr = sL * sR;
    const R ur = (sL * uL + sR * uR) * iden, vr = (sL * vL + sR * vR) * iden, Hr = (sL * HL + sR * HR) * iden;
    const R vnr = ur * nx + vr * ny, q2 = ur * ur + vr * vr;
    const R c2r = (g - R(1)) * (Hr - R(0.5) * q2);
    const R c = sqrtR(fmaxR(c2r, tiny));
    const R x1 = vnr - c, x4 = vnr + c;
    const R r1 = fabs(x1), r4 = fabs(x4), l2 = fabs(vnr);
    const R eps = R(0.1) * c;
    const bool f1 = r1 < eps, f4 = r4 < eps;
    const R i2e = recipR(R(2) * eps);
    const R l1 = f1 ? (r1 * r1 + eps * eps) * i2e : r1;
    const R l4 = f4 ? (r4 * r4 + eps * eps) * i2e : r4;
    const R du = uR - uL, dv = vR - vL, dp = pR - pL, dvn = du * nx + dv * ny;
    const R cc = c * c;
    const R ic2 = recipR(fmaxR(cc, tiny)), i2c2 = R(0.5) * ic2;
    const R a1 = (dp - rr * c * dvn) * i2c2;
    const R a2 = (rhoR - rhoL) - dp * ic2;
    const R a3 = rr * (dv * nx - du * ny);
    const R a4 = (dp + rr * c * dvn) * i2c2;
    const R tx = -ny, ty = nx;
    const R w1 = l1 * a1, w2 = l2 * a2, w3 = l2 * a3, w4 = l4 * a4;
    // F = (F_L + F_R)/2 - dF/2
    R dFb[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) dFb[k] = R(-0.5) * Fb[k];
    const R w1b = dFb[0] + dFb[1] * (ur - c * nx) + dFb[2] * (vr - c * ny) + dFb[3] * (Hr - c * vnr);
    const R w2b = dFb[0] + dFb[1] * ur + dFb[2] * vr + dFb[3] * R(0.5) * q2;
    const R w3b = dFb[1] * tx + dFb[2] * ty + dFb[3] * (ur * tx + vr * ty);
    const R w4b = dFb[0] + dFb[1] * (ur + c * nx) + dFb[2] * (vr + c * ny) + dFb[3] * (Hr + c * vnr);
    const R ws = w1 + w2 + w4;
    R urb = dFb[1] * ws + dFb[3] * w3 * tx;
    R vrb = dFb[2] * ws + dFb[3] * w3 * ty;
    R cb = (w4 - w1) * (dFb[1] * nx + dFb[2] * ny + dFb[3] * vnr);
    R Hrb = dFb[3] * (w1 + w4);
    R vnrb = dFb[3] * c * (w4 - w1);
    R q2b = dFb[3] * R(0.5) * w2;
    const R l1b = w1b * a1, a1b = w1b * l1;
    const R l2b = w2b * a2 + w3b * a3, a2b = w2b * l2, a3b = w3b * l2;
    const R l4b = w4b * a4, a4b = w4b * l4;
    const R dpb = (a1b + a4b) * i2c2 - a2b * ic2;
    const R rrb = (a4b - a1b) * c * dvn * i2c2 + a3b * (dv * nx - du * ny);
    cb += (a4b - a1b) * rr * dvn * i2c2;
    const R dvnb = (a4b - a1b) * rr * c * i2c2;
    const R i2c2b = a1b * (dp - rr * c * dvn) + a4b * (dp + rr * c * dvn);
    const R ic2b = -a2b * dp + R(0.5) * i2c2b;
    if (cc >= tiny) cb -= ic2b * ic2 * ic2 * R(2) * c;
    const R dub = -a3b * rr * ny + dvnb * nx, dvb = a3b * rr * nx + dvnb * ny;
    SideBar<R> bl, br;
    bl.rho -= a2b; br.rho += a2b;
    bl.u -= dub; br.u += dub;
    bl.v -= dvb; br.v += dvb;
    bl.p -= dpb; br.p += dpb;
    // Harten's entropy fix and the absolute values
    R epsb = R(0), r1b = l1b, r4b = l4b;
    if (f1) { epsb += l1b * (R(0.5) - r1 * r1 * R(0.5) * i2e * i2e * R(4)); r1b = l1b * r1 * R(2) * i2e; }
    if (f4) { epsb += l4b * (R(0.5) - r4 * r4 * R(0.5) * i2e * i2e * R(4)); r4b = l4b * r4 * R(2) * i2e; }
    const R x1b = r1b * sgnR(x1), x4b = r4b * sgnR(x4);
    vnrb += x1b + x4b + l2b * sgnR(vnr);
    cb += x4b - x1b + R(0.1) * epsb;
    if (c2r >= tiny) {
        const R c2rb = cb * R(0.5) * recipR(c);
        Hrb += c2rb * (g - R(1));
        q2b -= c2rb * (g - R(1)) * R(0.5);
    }
    urb += R(2) * ur * q2b + vnrb * nx;
    vrb += R(2) * vr * q2b + vnrb * ny;
    // Roe averages
    R sLb = (urb * uL + vrb * vL + Hrb * HL) * iden + rrb * sR;
    R sRb = (urb * uR + vrb * vR + Hrb * HR) * iden + rrb * sL;
    bl.u += urb * sL * iden; br.u += urb * sR * iden;
    bl.v += vrb * sL * iden; br.v += vrb * sR * iden;
    const R HLb = Hrb * sL * iden, HRb = Hrb * sR * iden;
    const R idenb = urb * (sL * uL + sR * uR) + vrb * (sL * vL + sR * vR) + Hrb * (sL * HL + sR * HR);
    if (ssum >= tiny) { const R sb = -idenb * iden * iden; sLb += sb; sRb += sb; }
    bl.rho += sLb * R(0.5) * recipR(sL);
    br.rho += sRb * R(0.5) * recipR(sR);
    // H = (U3 + p) / rho
    bl.U3 += HLb * L.ir; bl.p += HLb * L.ir; bl.rho -= HLb * HL * L.ir;
    br.U3 += HRb * Q.ir; br.p += HRb * Q.ir; br.rho -= HRb * HR * Q.ir;
    R Fh[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) Fh[k] = R(0.5) * Fb[k];
    flux_normal_floor_vjp(UL, L, nx, ny, Fh, bl);
    flux_normal_floor_vjp(UR, Q, nx, ny, Fh, br);
    side_vjp(UL, L, g, bl, ULb);
    side_vjp(UR, Q, g, br, URb);
}

template <typename R>
__device__ __forceinline__ void riemann_hllc_vjp(const R* UL, const R* UR, R nx, R ny, R g, const R* Fb, R* ULb,
                                                 R* URb) {
    const R tiny = R(1e-10);
    const Side<R> L = side_of(UL, g), Q = side_of(UR, g);
    const R rhoL = L.rho, pL = fmaxR(L.p, tiny), cL = sound_speed_of(L, g), vnL = L.u * nx + L.v * ny;
    const R rhoR = Q.rho, pR = fmaxR(Q.p, tiny), cR = sound_speed_of(Q, g), vnR = Q.u * nx + Q.v * ny;
    const R psr = R(0.5) * (pL + pR) - R(0.5) * (vnR - vnL) * R(0.5) * (rhoL + rhoR) * R(0.5) * (cL + cR);
    const R ps = fmaxR(psr, tiny);
    const R gp = (g + R(1)) * recipR(R(2) * g);
    const R ipL = recipR(pL), ipR = recipR(pR);
    const bool shL = !(ps <= pL), shR = !(ps <= pR);
    const R qL = shL ? sqrtR(R(1) + gp * (ps * ipL - R(1))) : R(1);
    const R qR = shR ? sqrtR(R(1) + gp * (ps * ipR - R(1))) : R(1);
    const R SL = vnL - cL * qL, SR = vnR + cR * qR;
    const R aL = SL - vnL, aR = SR - vnR;
    const R dSMr = rhoL * aL - rhoR * aR;
    const R dSM = guard_den(dSMr);
    const R idSM = recipR(dSM);
    const R num = pR - pL + rhoL * vnL * aL - rhoR * vnR * aR;
    const R SM = num * idSM;
    const bool plain = SL >= R(0) || SR <= R(0);
    const bool left = plain ? SL >= R(0) : SM >= R(0);
    const R* U = left ? UL : UR;
    const Side<R>& K = left ? L : Q;
    SideBar<R> bl, br;
    SideBar<R>& bk = left ? bl : br;
    flux_normal_floor_vjp(U, K, nx, ny, Fb, bk);
    // cotangents of the wave-speed chain
    R SLb = R(0), SRb = R(0), SMb = R(0);
    R vnLb = R(0), vnRb = R(0), pLb = R(0), pRb = R(0), cLb = R(0), cRb = R(0);
    R* Ubk = left ? ULb : URb;
    if (!plain) {
        const R vn = left ? vnL : vnR, S = left ? SL : SR, p = left ? pL : pR;
        const R rho = K.rho, u = K.u, v = K.v;
        const R dKr = S - SM, dvnr = S - vn;
        const R dK = guard_den(dKr), dvn = guard_den(dvnr);
        const R idK = recipR(dK), irdvn = recipR(rho * dvn);
        const R fac = rho * dvn * idK;
        const R m = SM - vn;
        const R t = p * irdvn;
        const R E = U[3] * K.ir;
        const R Us[NV] = {fac, fac * (u + m * nx), fac * (v + m * ny), fac * (E + m * (SM + t))};
        R Sb = R(0);
        R Usb[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            Sb += Fb[k] * (Us[k] - U[k]);
            Usb[k] = S * Fb[k];
            Ubk[k] -= S * Fb[k];  // F = Fk + S (Us - U): the direct U term
        }
        const R facb = Usb[0] + Usb[1] * (u + m * nx) + Usb[2] * (v + m * ny) + Usb[3] * (E + m * (SM + t));
        bk.u += Usb[1] * fac;
        bk.v += Usb[2] * fac;
        const R mb = fac * (Usb[1] * nx + Usb[2] * ny + Usb[3] * (SM + t));
        const R Eb = Usb[3] * fac;
        R SMbl = Usb[3] * fac * m;
        const R tb = Usb[3] * fac * m;
        const R pbk = tb * irdvn;
        R rhob = -tb * t * recipR(rho);
        R dvnb = -tb * t * recipR(dvn);
        SMbl += mb;
        R vnb = -mb;
        rhob += facb * dvn * idK;
        dvnb += facb * rho * idK;
        const R dKb = -facb * fac * idK;
        if (fabs(dKr) >= tiny) { Sb += dKb; SMbl -= dKb; }
        if (fabs(dvnr) >= tiny) { Sb += dvnb; vnb -= dvnb; }
        SMb += SMbl;
        bk.U3 += Eb * K.ir;
        bk.rho += rhob - Eb * E * K.ir;
        if (left) { pLb += pbk; vnLb += vnb; SLb += Sb; }
        else { pRb += pbk; vnRb += vnb; SRb += Sb; }
    }
    // SM = num / dSM
    const R numb = SMb * idSM;
    const R dSMb = fabs(dSMr) >= tiny ? -SMb * SM * idSM : R(0);
    pRb += numb; pLb -= numb;
    bl.rho += numb * vnL * aL + dSMb * aL;
    vnLb += numb * rhoL * aL;
    const R aLb = numb * rhoL * vnL + dSMb * rhoL;
    br.rho -= numb * vnR * aR + dSMb * aR;
    vnRb -= numb * rhoR * aR;
    const R aRb = -numb * rhoR * vnR - dSMb * rhoR;
    SLb += aLb; vnLb -= aLb;
    SRb += aRb; vnRb -= aRb;
    // SL = vnL - cL qL, SR = vnR + cR qR
    vnLb += SLb; cLb -= SLb * qL;
    const R qLb = -SLb * cL;
    vnRb += SRb; cRb += SRb * qR;
    const R qRb = SRb * cR;
    R psb = R(0);
    if (shL) { const R ib = qLb * R(0.5) * recipR(qL); psb += ib * gp * ipL; pLb -= ib * gp * ps * ipL * ipL; }
    if (shR) { const R ib = qRb * R(0.5) * recipR(qR); psb += ib * gp * ipR; pRb -= ib * gp * ps * ipR * ipR; }
    if (psr >= tiny) {
        pLb += R(0.5) * psb; pRb += R(0.5) * psb;
        const R dl = vnR - vnL, sr = rhoL + rhoR, sc = cL + cR;
        const R db = R(-0.125) * psb * sr * sc;
        vnRb += db; vnLb -= db;
        const R srb = R(-0.125) * psb * dl * sc;
        bl.rho += srb; br.rho += srb;
        const R scb = R(-0.125) * psb * dl * sr;
        cLb += scb; cRb += scb;
    }
    // per side: floored pressure, sound speed, normal velocity
    if (L.p >= tiny) bl.p += pLb;
    if (Q.p >= tiny) br.p += pRb;
    sound_speed_vjp(L, g, cL, cLb, bl);
    sound_speed_vjp(Q, g, cR, cRb, br);
    bl.u += vnLb * nx; bl.v += vnLb * ny;
    br.u += vnRb * nx; br.v += vnRb * ny;
    side_vjp(UL, L, g, bl, ULb);
    side_vjp(UR, Q, g, br, URb);
}

// ------------------------------------------------------------ physics VJPs
// Ub += (dFx/dU)^T Fxb + (dFy/dU)^T Fyb for IdealGas::fluxX/fluxY (euler.hpp:101-129).
template <typename R>
__device__ __forceinline__ void euler_flux_vjp(const R* U, R g, const R* a, const R* b, R* Ub) {
    const R inv = rcpR(U[0]), gm = g - R(1);
    const R u = U[1] * inv, v = U[2] * inv;
    const R p = gm * (U[3] - R(0.5) * (U[1] * u + U[2] * v));
    const R H = U[3] + p;
    R ub = a[1] * U[1] + a[3] * H + b[1] * U[2];
    R vb = a[2] * U[1] + b[2] * U[2] + b[3] * H;
    const R pb = a[1] + a[3] * u + b[2] + b[3] * v;
    R U1b = a[0] + a[1] * u + a[2] * v - pb * gm * u;
    R U2b = b[0] + b[1] * u + b[2] * v - pb * gm * v;
    const R U3b = a[3] * u + b[3] * v + pb * gm;
    U1b += ub * inv;
    U2b += vb * inv;
    const R U0b = -(ub * u + vb * v) * inv + pb * gm * R(0.5) * (u * u + v * v);
    Ub[0] += U0b; Ub[1] += U1b; Ub[2] += U2b; Ub[3] += U3b;
}

// VJP of computeGhostState (kernels.cu:453-525): Ub += (dG/dUi)^T Gb.
template <typename R>
__device__ void ghost_state_vjp(const R* Ui, R nx, R ny, int bct, const double* bc, R g, const R* Gb, R* Ub) {
    const R gm = g - R(1);
    if (bct == 1 || bct == 2 || bct == 3) {
        const R rho = Ui[0], u = Ui[1] / rho, v = Ui[2] / rho;
        const R vn = u * nx + v * ny;
        const R ur = u - R(2) * vn * nx, vr = v - R(2) * vn * ny;
        R rb = Gb[0] + Gb[1] * ur + Gb[2] * vr + Gb[3] * R(0.5) * (ur * ur + vr * vr);
        const R urb = Gb[1] * rho + Gb[3] * rho * ur, vrb = Gb[2] * rho + Gb[3] * rho * vr;
        const R pb = Gb[3] / gm;
        const R vnb = -R(2) * (urb * nx + vrb * ny);
        R ub = urb + vnb * nx, vb = vrb + vnb * ny;
        // p = (g-1) (U3 - rho (u^2+v^2)/2)
        Ub[3] += pb * gm;
        rb -= pb * gm * R(0.5) * (u * u + v * v);
        ub -= pb * gm * rho * u;
        vb -= pb * gm * rho * v;
        Ub[1] += ub / rho; Ub[2] += vb / rho;
        rb -= (ub * u + vb * v) / rho;
        Ub[0] += rb;
    } else if (bct == 4) {
        const R rinf = R(bc[0]), uinf = R(bc[1]), vinf = R(bc[2]), pinf = R(bc[3]);
        const bool rok = Ui[0] >= R(1e-10);
        const R ri = fmaxR(Ui[0], R(1e-10));
        const R ui = Ui[1] / ri, vi = Ui[2] / ri;
        const R praw = gm * (Ui[3] - R(0.5) * ri * (ui * ui + vi * vi));
        const bool pok = praw >= R(1e-10);
        const R pi = fmaxR(praw, R(1e-10));
        const bool inflow = uinf * nx + vinf * ny > R(0);
        const R rb0 = inflow ? rinf : ri, ub0 = inflow ? uinf : ui, vb0 = inflow ? vinf : vi;
        const R pb0 = R(0.5) * (pinf + pi);
        const R rb = fmaxR(fminR(rb0, R(100)), R(0.01));
        const R ub = fmaxR(fminR(ub0, R(1000)), R(-1000));
        const R vb = fmaxR(fminR(vb0, R(1000)), R(-1000));
        // G = prim2cons(rb, ub, vb, pb)
        R rbb = Gb[0] + Gb[1] * ub + Gb[2] * vb + Gb[3] * R(0.5) * (ub * ub + vb * vb);
        R ubb = Gb[1] * rb + Gb[3] * rb * ub, vbb = Gb[2] * rb + Gb[3] * rb * vb;
        R pbb = Gb[3] / gm;
        if (!(rb0 >= R(0.01) && rb0 <= R(100))) rbb = R(0);
        if (!(ub0 >= R(-1000) && ub0 <= R(1000))) ubb = R(0);
        if (!(vb0 >= R(-1000) && vb0 <= R(1000))) vbb = R(0);
        if (!(pb0 >= R(1e3) && pb0 <= R(1e7))) pbb = R(0);
        // p_i = (g-1)(U3 - ri (ui^2 + vi^2)/2) in the variables (ri, ui, vi, U3)
        const R pib = pok ? R(0.5) * pbb : R(0);
        R rib = -pib * gm * R(0.5) * (ui * ui + vi * vi);
        Ub[3] += pib * gm;
        R uib = -pib * gm * ri * ui, vib = -pib * gm * ri * vi;
        if (!inflow) { rib += rbb; uib += ubb; vib += vbb; }
        Ub[1] += uib / ri; Ub[2] += vib / ri;
        rib -= (uib * ui + vib * vi) / ri;
        if (rok) Ub[0] += rib;
    } else if (bct == 6) {
        const R rho = Ui[0], u = Ui[1] / rho, v = Ui[2] / rho;
        Ub[0] += Gb[0] - Gb[3] * R(0.5) * (u * u + v * v);
        Ub[1] += Gb[1] + Gb[3] * u;
        Ub[2] += Gb[2] + Gb[3] * v;
    } else {
#pragma unroll
        for (int k = 0; k < NV; ++k) Ub[k] += Gb[k];
    }
}

template <typename R>
__device__ __forceinline__ void riemann_by_kind_vjp(int kind, const R* UL, const R* UR, R nx, R ny, R g,
                                                    const R* Fb, R* ULb, R* URb) {
    if (kind == 1) riemann_roe_vjp(UL, UR, nx, ny, g, Fb, ULb, URb);
    else if (kind == 2) riemann_hllc_vjp(UL, UR, nx, ny, g, Fb, ULb, URb);
    else riemann_rusanov_vjp(UL, UR, nx, ny, g, Fb, ULb, URb);
}

}  // namespace
}  // namespace zj
