// zhijian/physics.hpp -- ideal-gas relations (host).
//
// API of the reference's IdealGas (include/physics/euler.hpp:8-177) and its
// Rusanov flux (euler.hpp:184-200), with the same safety clamps.  The device
// path evaluates the same relations inside the stage kernels.
#pragma once

#include "zhijian/types.hpp"

#include <algorithm>
#include <cmath>

namespace zhijian {

class IdealGas {
public:
    explicit IdealGas(Real gamma = 1.4) : gamma_(gamma) {}

    Real pressure(const State& U) const {
        const Real rho = U.rho(), u = U.rhou() / rho, v = U.rhov() / rho, E = U.rhoE() / rho;
        return (gamma_ - 1.0) * rho * (E - 0.5 * (u * u + v * v));
    }
    Real temperature(const State& U, Real R_gas = 287.0) const { return pressure(U) / (U.rho() * R_gas); }
    Real soundSpeed(const State& U) const {
        const Real p = std::fmax(pressure(U), 1e-10), rho = std::fmax(U.rho(), 1e-10);
        return std::sqrt(gamma_ * p / rho);
    }
    Real machNumber(const State& U) const {
        const Real u = U.rhou() / U.rho(), v = U.rhov() / U.rho();
        return std::sqrt(u * u + v * v) / soundSpeed(U);
    }
    void velocity(const State& U, Real& u, Real& v) const {
        u = U.rhou() / U.rho();
        v = U.rhov() / U.rho();
    }
    Real enthalpy(const State& U) const { return (U.rhoE() + pressure(U)) / U.rho(); }
    Real totalEnthalpy(const State& U) const { return U.rhoE() + pressure(U); }
    Real entropy(const State& U) const { return pressure(U) / std::pow(U.rho(), gamma_); }
    State primToConserv(Real rho, Real u, Real v, Real p) const {
        State U;
        U.rho() = rho; U.rhou() = rho * u; U.rhov() = rho * v;
        U.rhoE() = p / (gamma_ - 1.0) + 0.5 * rho * (u * u + v * v);
        return U;
    }
    void conservToPrim(const State& U, Real& rho, Real& u, Real& v, Real& p) const {
        rho = U.rho(); u = U.rhou() / rho; v = U.rhov() / rho; p = pressure(U);
    }
    State fluxX(const State& U) const {
        const Real u = U.rhou() / U.rho(), v = U.rhov() / U.rho(), p = pressure(U);
        State F;
        F[0] = U.rhou(); F[1] = U.rhou() * u + p; F[2] = U.rhou() * v; F[3] = (U.rhoE() + p) * u;
        return F;
    }
    State fluxY(const State& U) const {
        const Real u = U.rhou() / U.rho(), v = U.rhov() / U.rho(), p = pressure(U);
        State G;
        G[0] = U.rhov(); G[1] = U.rhov() * u; G[2] = U.rhov() * v + p; G[3] = (U.rhoE() + p) * v;
        return G;
    }
    State fluxNormal(const State& U, Real nx, Real ny) const {
        const Real rho = std::fmax(U.rho(), 1e-10);
        const Real u = U.rhou() / rho, v = U.rhov() / rho;
        const Real p = std::fmax(pressure(U), 1e-10), H = totalEnthalpy(U), vn = u * nx + v * ny;
        State Fn;
        Fn[0] = rho * vn; Fn[1] = rho * u * vn + p * nx; Fn[2] = rho * v * vn + p * ny; Fn[3] = H * vn;
        return Fn;
    }
    void eigenvalues(const State& U, Real nx, Real ny, Real& l1, Real& l2, Real& l3, Real& l4) const {
        Real u, v;
        velocity(U, u, v);
        const Real vn = u * nx + v * ny, c = soundSpeed(U);
        l1 = vn - c; l2 = vn; l3 = vn; l4 = vn + c;
    }
    Real maxWaveSpeed(const State& U) const {
        Real u, v;
        velocity(U, u, v);
        return std::sqrt(u * u + v * v) + soundSpeed(U);
    }
    Real gamma() const { return gamma_; }

private:
    Real gamma_;
};

class RiemannSolvers {
public:
    static State rusanov(const State& UL, const State& UR, Real nx, Real ny, Real gamma) {
        IdealGas gas(gamma);
        const State FL = gas.fluxNormal(UL, nx, ny), FR = gas.fluxNormal(UR, nx, ny);
        const Real smax = std::fmax(gas.maxWaveSpeed(UL), gas.maxWaveSpeed(UR));
        State F;
        for (int i = 0; i < N_VARS; ++i) F[i] = 0.5 * (FL[i] + FR[i]) - 0.5 * smax * (UR[i] - UL[i]);
        return F;
    }
};

}  // namespace zhijian
