// zhijian/fr_solver.hpp -- the flux-reconstruction solver (drop-in API).
//
// Public interface of the reference's FRSolver (include/solver/fr_solver.hpp:
// 77-123): initialize, setInitialCondition, setBoundaryCondition, advance,
// computeTimeStep, computeResidual, currentTime, iteration, solution,
// copyToHost, copyToDevice, params, mesh, computeDerivedQuantities.  Behind it,
// one zj_handle (include/zhijian_amd.h) owns the mesh's device state on one
// MI355X and runs each SSP-RK3 stage as one fused kernel.
//
// Differences from the reference, all deliberate:
//   * errors throw std::runtime_error (the reference calls exit());
//   * computeTimeStep evaluates the CFL rule on the current device state (the
//     reference reads a host mirror refreshed only at output time);
//   * the solver keeps a non-owning pointer to the Mesh, like the reference;
//     boundary tables are re-read from the mesh's faces at each advance() and
//     uploaded only when they changed.
#pragma once

#include "zhijian/basis.hpp"
#include "zhijian/mesh.hpp"
#include "zhijian/physics.hpp"
#include "zhijian/types.hpp"

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

struct zj_solver;

namespace zhijian {

class BoundaryCondition;

struct ElementSolution {
    std::vector<State> sol_pts;
    std::vector<State> flux_pts;
    std::vector<State> dUdx;
    std::vector<State> dUdy;
};

class FRSolver {
public:
    FRSolver();
    ~FRSolver();
    FRSolver(const FRSolver&) = delete;
    FRSolver& operator=(const FRSolver&) = delete;

    void initialize(const Mesh& mesh, const SimParams& params);
    void setInitialCondition(const std::function<State(Vec2)>& ic_func);
    void setBoundaryCondition(int bc_tag, std::shared_ptr<BoundaryCondition> bc);
    void advance(Real dt);
    Real computeTimeStep() const;
    Real computeResidual() const;
    Real currentTime() const { return time_; }
    int iteration() const { return iter_; }
    const std::vector<ElementSolution>& solution() const { return solution_; }
    void copyToHost();
    void copyToDevice();
    const SimParams& params() const { return params_; }
    const Mesh& mesh() const { return *mesh_; }
    void computeDerivedQuantities(std::vector<std::vector<Real>>& pressure,
                                  std::vector<std::vector<Real>>& mach) const;

    // ---- additions of this library
    // Residual dU/dt of the current state, AoS [E][S][4] on the host.
    std::vector<Real> residualField();
    // Resume from a restart file's solution (RestartIO::read): sets the state,
    // the time and the iteration count.  The reference's main() reads the file
    // but never hands it to the solver (src/main.cpp:349-360).
    void setSolution(const std::vector<ElementSolution>& solution, Real time, int iteration);
    // The underlying C ABI handle (for callers that drive stages directly).
    zj_solver* handle() const { return handle_; }

private:
    void uploadBoundaryTables(bool force);
    void check(int rc, const char* what) const;

    const Mesh* mesh_ = nullptr;
    SimParams params_;
    std::vector<ElementSolution> solution_;
    std::map<ElementType, FROperators> operators_;
    std::map<int, std::shared_ptr<BoundaryCondition>> bc_map_;
    zj_solver* handle_ = nullptr;
    Real time_ = 0.0;
    int iter_ = 0;
    int n_sp_tri_ = 0;
    int n_sp_quad_ = 0;
    std::vector<int> bc_type_cache_;
    std::vector<int> bc_face_map_cache_;
    std::vector<double> bc_data_cache_;
};

}  // namespace zhijian
