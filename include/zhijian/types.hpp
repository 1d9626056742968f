// zhijian/types.hpp -- basic types of the FR solver API.
//
// Same names, fields and defaults as the reference's include/common/types.hpp
// (Real, Index, LocalIndex, N_VARS, the enums, Vec2, State, Flux2D, SimParams,
// PartitionInfo) so that code written against the reference compiles
// unchanged.  SimParams gains fields of its own in this library: precision, compat
// and positivity.
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace zhijian {

using Real = double;
using Index = int64_t;
using LocalIndex = int32_t;

constexpr int MAX_POLY_ORDER = 5;
constexpr int N_VARS = 4;  // rho, rho*u, rho*v, rho*E

enum class Var : int { Density = 0, MomentumX = 1, MomentumY = 2, Energy = 3 };
enum class ElementType : int { Triangle = 0, Quadrilateral = 1 };
enum class BCType : int {
    Interior = 0, Wall = 1, SlipWall = 2, Symmetry = 3,
    FarField = 4, Inflow = 5, Outflow = 6, Periodic = 7
};
enum class FluxType : int { DG = 0, SD = 1, HU = 2, GA = 3 };
enum class RiemannSolver : int { Rusanov = 0, Roe = 1, HLLC = 2 };

// Arithmetic of the device path.
enum class Precision : int { FP64 = 0, FP32Bulk = 1 };
enum class Compat : int { Reference = 0, Textbook = 1 };

struct Vec2 {
    Real x, y;
    Vec2() : x(0), y(0) {}
    Vec2(Real x_, Real y_) : x(x_), y(y_) {}
    Vec2 operator+(const Vec2& v) const { return Vec2(x + v.x, y + v.y); }
    Vec2 operator-(const Vec2& v) const { return Vec2(x - v.x, y - v.y); }
    Vec2 operator*(Real s) const { return Vec2(x * s, y * s); }
    Vec2 operator/(Real s) const { return Vec2(x / s, y / s); }
    Real dot(const Vec2& v) const { return x * v.x + y * v.y; }
    Real norm() const { return std::sqrt(x * x + y * y); }
    Real norm2() const { return x * x + y * y; }
};

struct State {
    Real data[N_VARS];
    State() { for (int i = 0; i < N_VARS; ++i) data[i] = 0; }
    Real& operator[](int i) { return data[i]; }
    const Real& operator[](int i) const { return data[i]; }
    Real& rho() { return data[0]; }
    Real& rhou() { return data[1]; }
    Real& rhov() { return data[2]; }
    Real& rhoE() { return data[3]; }
    const Real& rho() const { return data[0]; }
    const Real& rhou() const { return data[1]; }
    const Real& rhov() const { return data[2]; }
    const Real& rhoE() const { return data[3]; }
    State operator+(const State& s) const {
        State r;
        for (int i = 0; i < N_VARS; ++i) r.data[i] = data[i] + s.data[i];
        return r;
    }
    State operator*(Real a) const {
        State r;
        for (int i = 0; i < N_VARS; ++i) r.data[i] = data[i] * a;
        return r;
    }
};

struct Flux2D {
    State Fx;
    State Fy;
};

struct SimParams {
    Real gamma = 1.4;
    Real Pr = 0.72;
    Real Re = 1000.0;
    Real Mach_inf = 0.5;
    Real T_inf = 300.0;
    Real p_inf = 101325.0;
    Real rho_inf = 1.225;
    Real AoA = 0.0;

    int poly_order = 3;
    FluxType flux_type = FluxType::HU;
    RiemannSolver riemann = RiemannSolver::Roe;

    Real CFL = 0.5;
    Real dt = 0.0;
    Real t_final = 1.0;
    int max_iter = 100000;

    // Reference arithmetic: computed and discarded, as the reference does.
    // Textbook scheme: compressible Navier-Stokes terms with BR2 (mu = 1/Re,
    // affine elements, periodic meshes; initialize() throws otherwise).
    bool viscous = true;

    int output_freq = 100;
    int restart_freq = 1000;
    std::string output_dir = "./output";
    std::string case_name = "simulation";

    struct BCSpec {
        std::string type_str;
        Real p_static = 0.0;
        Real p_total = 0.0;
        Real T_total = 0.0;
        Real T_wall = 0.0;
        Real dir_x = 1.0, dir_y = 0.0;
    };
    std::map<int, BCSpec> bc_specs;

    // ---- additions of this library
    Precision precision = Precision::FP64;  // FP32Bulk: FP32 flux/interp/correction, FP64 RK state
    Compat compat = Compat::Reference;      // the reference's executed arithmetic
    bool positivity = true;                 // the reference's limiter after every RK stage
    int device = 0;                         // HIP device owning the mesh
};

struct PartitionInfo {
    int rank = 0;
    int num_procs = 1;
    int num_local_elems = 0;
    int num_ghost_elems = 0;
    std::vector<int> neighbor_ranks;
};

}  // namespace zhijian
