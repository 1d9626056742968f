// zhijian/rk_integrator.hpp -- time-integration API.
//
// The reference declares these classes (include/time/rk_integrator.hpp:11-79)
// and defines none of them; this library defines all of them.  The *_GPU
// functions take device pointers and run elementwise on the given HIP stream;
// they do not apply the positivity limiter (FRSolver::advance does, as the
// reference's advance does through gpu::rkStage*, kernels.cu:1305-1349).
#pragma once

#include "zhijian/types.hpp"

#include <functional>
#include <hip/hip_runtime_api.h>

namespace zhijian {

class TimeIntegrator {
public:
    virtual ~TimeIntegrator() = default;
    virtual void advance(Real* U, const Real* dUdt, int n, Real dt, hipStream_t stream = 0) = 0;
    virtual int numStages() const = 0;
    virtual void getStageCoeffs(int stage, Real& alpha, Real& beta) const = 0;
};

// u1 = u + dt L(u);  u2 = 3/4 u + 1/4 u1 + 1/4 dt L(u1);  u = 1/3 u + 2/3 u2 + 2/3 dt L(u2)
class SSPRK3 {
public:
    SSPRK3() = default;
    static constexpr int numStages() { return 3; }

    static void stage1_CPU(Real* U1, const Real* U0, const Real* R, int n, Real dt);
    static void stage2_CPU(Real* U2, const Real* U0, const Real* U1, const Real* R, int n, Real dt);
    static void stage3_CPU(Real* U, const Real* U0, const Real* U2, const Real* R, int n, Real dt);

    static void stage1_GPU(Real* U1, const Real* U0, const Real* R, int n, Real dt, hipStream_t stream = 0);
    static void stage2_GPU(Real* U2, const Real* U0, const Real* U1, const Real* R, int n, Real dt,
                           hipStream_t stream = 0);
    static void stage3_GPU(Real* U, const Real* U0, const Real* U2, const Real* R, int n, Real dt,
                           hipStream_t stream = 0);
};

class ForwardEuler {
public:
    static constexpr int numStages() { return 1; }
    static void advance_CPU(Real* U, const Real* R, int n, Real dt);
    static void advance_GPU(Real* U, const Real* R, int n, Real dt, hipStream_t stream = 0);
};

// Classical RK4 on the host: compute_rhs(U_in, R_out) evaluates the residual.
class RK4 {
public:
    static constexpr int numStages() { return 4; }
    static void advance_CPU(Real* U, Real* k1, Real* k2, Real* k3, Real* k4, const Real* R, int n, Real dt,
                            const std::function<void(Real*, Real*)>& compute_rhs);
};

}  // namespace zhijian
