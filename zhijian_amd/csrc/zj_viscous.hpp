// zj_viscous.hpp -- the compressible Navier-Stokes viscous flux and its
// vector-Jacobian product, shared by the quadrilateral (zj_visc.hip) and
// triangle (zj_tri.hip) kernels.
#pragma once

#include "zj_device.hpp"

namespace zj {
namespace {

// ViscousFlux::velocityGradients / temperatureGradient / stressTensor /
// fluxX / fluxY (navierstokes.hpp:38-99) in one pass, mu and kappa given.
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

// visc_flux backwards: Ub += (dF_v/dU)^T Fb + (dG_v/dU)^T Gb, and likewise
// the cotangents of the gradient Qx, Qy.
template <typename R>
__device__ __forceinline__ void visc_flux_vjp(const R* U, const R* Qx, const R* Qy, R g, R mu, R kappa, const R* Fb,
                                              const R* Gb, R* Ub, R* Qxb, R* Qyb) {
    const R gm = g - R(1);
    const R rho = U[0], ir = rcpR(rho), u = U[1] * ir, v = U[2] * ir;
    const R dudx = (Qx[1] - u * Qx[0]) * ir, dudy = (Qy[1] - u * Qy[0]) * ir;
    const R dvdx = (Qx[2] - v * Qx[0]) * ir, dvdy = (Qy[2] - v * Qy[0]) * ir;
    const R q2 = u * u + v * v;
    const R p = gm * rho * (U[3] * ir - R(0.5) * q2);
    const R dpdx = gm * (Qx[3] - R(0.5) * Qx[0] * q2 - rho * (u * dudx + v * dvdx));
    const R dpdy = gm * (Qy[3] - R(0.5) * Qy[0] * q2 - rho * (u * dudy + v * dvdy));
    const R T = p * ir;
    const R div = dudx + dvdy, lam = R(-2.0 / 3.0) * mu;
    const R txx = R(2) * mu * dudx + lam * div, tyy = R(2) * mu * dvdy + lam * div, txy = mu * (dudy + dvdx);
    // F_v = [0, txx, txy, u txx + v txy + kappa dT/dx], G_v = [0, txy, tyy, u txy + v tyy + kappa dT/dy]
    const R txxb = Fb[1] + u * Fb[3], txyb = Fb[2] + v * Fb[3] + Gb[1] + u * Gb[3], tyyb = Gb[2] + v * Gb[3];
    R ub = txx * Fb[3] + txy * Gb[3], vb = txy * Fb[3] + tyy * Gb[3];
    const R dTdxb = kappa * Fb[3], dTdyb = kappa * Gb[3];
    const R divb = lam * (txxb + tyyb);
    R dudxb = R(2) * mu * txxb + divb, dvdyb = R(2) * mu * tyyb + divb, dudyb = mu * txyb, dvdxb = mu * txyb;
    // dT/dx = (dp/dx - T Qx0) / rho, T = p / rho
    const R dpdxb = dTdxb * ir, dpdyb = dTdyb * ir;
    const R Tb = -(dTdxb * Qx[0] + dTdyb * Qy[0]) * ir;
    R Qx0b = -dTdxb * T * ir, Qy0b = -dTdyb * T * ir;
    R irb = dTdxb * (dpdx - T * Qx[0]) + dTdyb * (dpdy - T * Qy[0]);
    const R pb = Tb * ir;
    irb += Tb * p;
    // dp/dx = (g-1) (Qx3 - Qx0 q2 / 2 - rho (u du/dx + v dv/dx))
    Qxb[3] += gm * dpdxb;
    Qyb[3] += gm * dpdyb;
    Qx0b -= R(0.5) * gm * q2 * dpdxb;
    Qy0b -= R(0.5) * gm * q2 * dpdyb;
    R q2b = -R(0.5) * gm * (Qx[0] * dpdxb + Qy[0] * dpdyb);
    R rhob = -gm * (dpdxb * (u * dudx + v * dvdx) + dpdyb * (u * dudy + v * dvdy));
    ub -= gm * rho * (dpdxb * dudx + dpdyb * dudy);
    vb -= gm * rho * (dpdxb * dvdx + dpdyb * dvdy);
    dudxb -= gm * rho * u * dpdxb;
    dvdxb -= gm * rho * v * dpdxb;
    dudyb -= gm * rho * u * dpdyb;
    dvdyb -= gm * rho * v * dpdyb;
    // p = (g-1) rho (U3 / rho - q2 / 2)
    rhob += pb * gm * (U[3] * ir - R(0.5) * q2);
    Ub[3] += pb * gm * rho * ir;
    irb += pb * gm * rho * U[3];
    q2b -= R(0.5) * pb * gm * rho;
    ub += R(2) * u * q2b;
    vb += R(2) * v * q2b;
    // velocity gradients (Q - u Q0) / rho
    Qxb[1] += dudxb * ir; ub -= dudxb * Qx[0] * ir; Qx0b -= dudxb * u * ir; irb += dudxb * (Qx[1] - u * Qx[0]);
    Qyb[1] += dudyb * ir; ub -= dudyb * Qy[0] * ir; Qy0b -= dudyb * u * ir; irb += dudyb * (Qy[1] - u * Qy[0]);
    Qxb[2] += dvdxb * ir; vb -= dvdxb * Qx[0] * ir; Qx0b -= dvdxb * v * ir; irb += dvdxb * (Qx[2] - v * Qx[0]);
    Qyb[2] += dvdyb * ir; vb -= dvdyb * Qy[0] * ir; Qy0b -= dvdyb * v * ir; irb += dvdyb * (Qy[2] - v * Qy[0]);
    Qxb[0] += Qx0b;
    Qyb[0] += Qy0b;
    // u = U1 / rho, v = U2 / rho
    Ub[1] += ub * ir;
    Ub[2] += vb * ir;
    irb += ub * U[1] + vb * U[2];
    rhob -= irb * ir * ir;
    Ub[0] += rhob;
}

}  // namespace
}  // namespace zj
