// frsolver_driver.cpp -- exercises the C++ drop-in API the way the reference's
// main() drives its solver (src/main.cpp:291-416): build a mesh, initialize,
// set an initial condition, then per iteration dt = computeTimeStep(),
// dt = min(dt, params.dt), advance(dt), residual = computeResidual().
//
//   frsolver_driver mesh   <n> <out.bin>
//       (no GPU) faces and normals of the periodic n x n unit mesh
//   frsolver_driver gmsh   <file.msh> <out.bin>
//       (no GPU) faces and boundary tags of a Gmsh mesh read by GmshReader
//   frsolver_driver run    <vortex|tgv> <n> <p> <steps> <dt_max> <fp64|fp32bulk> <prefix>
//       writes <prefix>_U0.bin, <prefix>_U.bin and one JSON line on stdout
//   frsolver_driver io     <dir>
//       (no GPU) RestartIO, VTKWriter and ResidualWriter on a 3 x 3 mesh
//       with a synthetic P2 solution; files under <dir>, one JSON line
//   frsolver_driver resume <dir> <n> <p> <steps_a> <steps_b>
//       vortex steps_a steps, restart file, a new solver resumed from it
//       (RestartIO::read + setSolution), steps_b more; and one solver run
//       straight through; writes <dir>/resumed.bin, <dir>/straight.bin
#include "zhijian/fr_solver.hpp"
#include "zhijian/io.hpp"
#include "zhijian/rk_integrator.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

using namespace zhijian;

static Mesh uniform_mesh(int n, Real x0, Real x1) {
    Mesh m;
    const Real h = (x1 - x0) / n;
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) m.nodes().push_back(Vec2(x0 + i * h, x0 + j * h));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            Element e;
            e.type = ElementType::Quadrilateral;
            e.order = 1;
            const Index n0 = (Index)j * (n + 1) + i;
            e.node_ids = {n0, n0 + 1, n0 + n + 2, n0 + n + 1};
            m.elements().push_back(e);
        }
    m.buildFaces();
    m.pairPeriodic(x1 - x0, x1 - x0);
    m.computeGeometry();
    return m;
}

static void dump(const std::string& path, const std::vector<double>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

static std::vector<double> state(const FRSolver& s) {
    std::vector<double> U;
    for (const ElementSolution& e : s.solution())
        for (const State& q : e.sol_pts)
            for (int v = 0; v < N_VARS; ++v) U.push_back(q[v]);
    return U;
}

static std::vector<ElementSolution> synthetic_solution(Index E, int nsp) {
    std::vector<ElementSolution> sol(E);
    for (Index e = 0; e < E; ++e) {
        sol[e].sol_pts.resize(nsp);
        for (int k = 0; k < nsp; ++k) {
            const Real rho = 1.0 + 0.01 * ((e * 7 + k * 3) % 11), u = 0.1 * ((e + 2 * k) % 5) - 0.2,
                       v = 0.05 * ((3 * e + k) % 7) - 0.15, pr = 1.0 + 0.02 * ((e + k) % 13);
            State& q = sol[e].sol_pts[k];
            q[0] = rho;
            q[1] = rho * u;
            q[2] = rho * v;
            q[3] = pr / 0.4 + 0.5 * rho * (u * u + v * v);
        }
    }
    return sol;
}

static bool same_solution(const std::vector<ElementSolution>& a, const std::vector<ElementSolution>& b) {
    if (a.size() != b.size()) return false;
    for (size_t e = 0; e < a.size(); ++e) {
        if (a[e].sol_pts.size() != b[e].sol_pts.size()) return false;
        for (size_t k = 0; k < a[e].sol_pts.size(); ++k)
            for (int v = 0; v < N_VARS; ++v)
                if (std::memcmp(&a[e].sol_pts[k][v], &b[e].sol_pts[k][v], sizeof(Real))) return false;
    }
    return true;
}

static int io_case(const std::string& dir) {
    Mesh mesh = uniform_mesh(3, 0.0, 1.0);
    SimParams params;
    params.poly_order = 2;
    params.gamma = 1.4;
    const auto sol = synthetic_solution(mesh.numElements(), 9);
    std::vector<double> U;
    for (const auto& e : sol)
        for (const auto& q : e.sol_pts)
            for (int v = 0; v < N_VARS; ++v) U.push_back(q[v]);
    dump(dir + "/U.bin", U);

    RestartIO rio;
    bool ok = rio.write(dir + "/r.h5", mesh, sol, params, 0.75, 120);
    std::vector<ElementSolution> back;
    SimParams rp;
    Real t = 0;
    int it = 0;
    ok = ok && rio.read(dir + "/r.h5", mesh, back, rp, t, it);
    const bool restart_ok = ok && same_solution(sol, back) && t == 0.75 && it == 120 && rp.poly_order == 2 &&
                            rp.gamma == 1.4;
    ok = rio.writeParallel(dir + "/p.h5", 0, 1, mesh, sol, params, 0.5, 60);
    std::vector<ElementSolution> pback;
    ok = ok && rio.readParallel(dir + "/p.h5", 0, 1, mesh, pback, rp, t, it);
    const bool parallel_ok = ok && same_solution(sol, pback) && t == 0.5 && it == 60;
    std::vector<ElementSolution> none;
    const bool missing_fails = !rio.read(dir + "/missing.h5", mesh, none, rp, t, it) && !rio.errorMessage().empty();

    VTKWriter vtk;
    bool vtk_ok = vtk.writeVTU(dir + "/a.vtu", mesh, sol, params) && vtk.writeLegacy(dir + "/a.vtk", mesh, sol, params);
    vtk.setSubdivisions(2);
    vtk_ok = vtk_ok && vtk.writeParallelVTU(dir + "/par", 0, 2, mesh, sol, params);
    vtk.setSubdivisions(4);
    vtk.setBinary(true);
    vtk_ok = vtk_ok && vtk.writeVTU(dir + "/b.vtu", mesh, sol, params);
    params.compat = Compat::Textbook;
    vtk_ok = vtk_ok && vtk.writeVTU(dir + "/h.vtu", mesh, sol, params);

    ResidualWriter res;
    const bool res_ok = res.open(dir + "/res.csv");
    res.write(10, 0.5, 1.25e-3, {1.0, 2.5e-7, -3.0, 4.125});
    res.write(20, 1.0, 6.5e-4, {0.0, 0.0, 0.0, 0.0});
    res.close();
    std::printf("{\"restart\": %s, \"parallel\": %s, \"missing_fails\": %s, \"vtk\": %s, \"residual\": %s}\n",
                restart_ok ? "true" : "false", parallel_ok ? "true" : "false", missing_fails ? "true" : "false",
                vtk_ok ? "true" : "false", res_ok ? "true" : "false");
    return 0;
}

static int resume_case(const std::string& dir, int n, int p, int steps_a, int steps_b) {
    Mesh mesh = uniform_mesh(n, -5.0, 5.0);
    SimParams params;
    params.poly_order = p;
    params.gamma = 1.4;
    params.CFL = 0.5;
    params.dt = 1e-3;
    params.viscous = false;
    const Real g = params.gamma, beta = 5.0;
    IdealGas gas(g);
    auto ic = [&](Vec2 x) {
        const Real r2 = x.x * x.x + x.y * x.y, f = std::exp(0.5 * (1.0 - r2));
        const Real T = 1.0 - (g - 1.0) * beta * beta / (8.0 * g * M_PI * M_PI) * std::exp(1.0 - r2);
        const Real rho = std::pow(T, 1.0 / (g - 1.0));
        return gas.primToConserv(rho, 1.0 - beta / (2 * M_PI) * x.y * f, beta / (2 * M_PI) * x.x * f,
                                 std::pow(rho, g));
    };
    auto run = [&](FRSolver& s, int steps) {
        for (int k = 0; k < steps; ++k) {
            Real dt = std::min(s.computeTimeStep(), params.dt);
            s.advance(dt);
        }
    };
    FRSolver straight;
    straight.initialize(mesh, params);
    straight.setInitialCondition(ic);
    run(straight, steps_a + steps_b);
    straight.copyToHost();
    dump(dir + "/straight.bin", state(straight));

    RestartIO rio;
    {
        FRSolver first;
        first.initialize(mesh, params);
        first.setInitialCondition(ic);
        run(first, steps_a);
        first.copyToHost();
        if (!rio.write(dir + "/resume.h5", mesh, first.solution(), params, first.currentTime(), first.iteration())) {
            std::fprintf(stderr, "%s\n", rio.errorMessage().c_str());
            return 1;
        }
    }
    std::vector<ElementSolution> sol;
    SimParams rp = params;
    Real t = 0;
    int it = 0;
    if (!rio.read(dir + "/resume.h5", mesh, sol, rp, t, it)) {
        std::fprintf(stderr, "%s\n", rio.errorMessage().c_str());
        return 1;
    }
    FRSolver second;
    second.initialize(mesh, rp);
    second.setSolution(sol, t, it);
    run(second, steps_b);
    second.copyToHost();
    dump(dir + "/resumed.bin", state(second));
    std::printf("{\"time\": %.17g, \"time_straight\": %.17g, \"iteration\": %d, \"iteration_straight\": %d}\n",
                second.currentTime(), straight.currentTime(), second.iteration(), straight.iteration());
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 4 && std::strcmp(argv[1], "mesh") == 0) {
        Mesh m = uniform_mesh(std::atoi(argv[2]), 0.0, 1.0);
        std::vector<double> out;
        for (Index f = 0; f < m.numFaces(); ++f) {
            const Face& fc = m.face(f);
            const Vec2 n = m.faceNormal(f);
            out.insert(out.end(), {(double)fc.left_elem, (double)fc.left_face, (double)fc.right_elem,
                                   (double)fc.right_face, n.x, n.y});
        }
        dump(argv[3], out);
        std::printf("{\"faces\": %lld}\n", (long long)m.numFaces());
        return 0;
    }
    if (argc >= 4 && std::strcmp(argv[1], "gmsh") == 0) {
        // faces of a Gmsh mesh read through GmshReader: le ll re rl bc_type bc_tag per face;
        // with a 5th argument "periodic", after pairing by the file's $Periodic links
        Mesh m;
        GmshReader reader;
        if (!reader.read(argv[2], m)) {
            std::fprintf(stderr, "%s\n", reader.errorMessage().c_str());
            return 3;
        }
        if (argc >= 5 && std::strcmp(argv[4], "periodic") == 0) {
            try {
                m.pairPeriodicNodes();
            } catch (const std::exception& e) {
                std::fprintf(stderr, "%s\n", e.what());
                return 4;
            }
        }
        std::vector<double> out;
        for (Index f = 0; f < m.numFaces(); ++f) {
            const Face& fc = m.face(f);
            out.insert(out.end(), {(double)fc.left_elem, (double)fc.left_face, (double)fc.right_elem,
                                   (double)fc.right_face, (double)static_cast<int>(fc.bc_type), (double)fc.bc_tag});
        }
        dump(argv[3], out);
        std::printf("{\"nodes\": %lld, \"elements\": %lld, \"faces\": %lld, \"periodic_links\": %lld}\n",
                    (long long)m.numNodes(), (long long)m.numElements(), (long long)m.numFaces(),
                    (long long)m.periodicNodes().size());
        return 0;
    }
    if (argc >= 7 && std::strcmp(argv[1], "gmshrun") == 0) {
        // a periodic Gmsh mesh ($Periodic links) through GmshReader, Mesh::pairPeriodicNodes
        // and FRSolver in textbook mode (Roe, HU, no limiter, inviscid): TGV state, fixed dt
        Mesh mesh;
        GmshReader reader;
        if (!reader.read(argv[2], mesh)) {
            std::fprintf(stderr, "%s\n", reader.errorMessage().c_str());
            return 3;
        }
        mesh.pairPeriodicNodes();
        SimParams params;
        params.poly_order = std::atoi(argv[3]);
        params.compat = Compat::Textbook;
        params.flux_type = FluxType::HU;
        params.riemann = RiemannSolver::Roe;
        params.positivity = false;
        params.viscous = false;
        FRSolver solver;
        solver.initialize(mesh, params);
        const Real g = params.gamma;
        IdealGas gas(g);
        solver.setInitialCondition([&](Vec2 x) {
            return gas.primToConserv(1.0, std::sin(x.x) * std::cos(x.y), -std::cos(x.x) * std::sin(x.y),
                                     1.0 / (g * 0.3 * 0.3) + 0.25 * (std::cos(2 * x.x) + std::cos(2 * x.y)));
        });
        const std::string prefix = argv[6];
        solver.copyToHost();
        dump(prefix + "_U0.bin", state(solver));
        for (int it = 0; it < std::atoi(argv[4]); ++it) solver.advance(std::atof(argv[5]));
        solver.copyToHost();
        dump(prefix + "_U.bin", state(solver));
        std::printf("{\"elements\": %lld, \"faces\": %lld, \"boundary_faces\": %lld}\n",
                    (long long)mesh.numElements(), (long long)mesh.numFaces(), (long long)mesh.numBoundaryFaces());
        return 0;
    }
    if (argc >= 3 && std::strcmp(argv[1], "io") == 0) return io_case(argv[2]);
    if (argc >= 7 && std::strcmp(argv[1], "resume") == 0)
        return resume_case(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6]));
    if (argc < 9 || std::strcmp(argv[1], "run") != 0) {
        std::fprintf(stderr, "usage: %s mesh <n> <out> | run <vortex|tgv> <n> <p> <steps> <dt_max> <prec> <prefix>\n",
                     argv[0]);
        return 2;
    }
    const std::string kind = argv[2];
    const int n = std::atoi(argv[3]), p = std::atoi(argv[4]), steps = std::atoi(argv[5]);
    const double dt_max = std::atof(argv[6]);
    const std::string prec = argv[7], prefix = argv[8];
    const bool vortex = kind == "vortex";
    Mesh mesh = vortex ? uniform_mesh(n, -5.0, 5.0) : uniform_mesh(n, 0.0, 2.0 * M_PI);

    SimParams params;
    params.poly_order = p;
    params.gamma = 1.4;
    params.CFL = 0.5;
    params.dt = dt_max;
    params.viscous = false;
    params.flux_type = FluxType::DG;
    params.riemann = RiemannSolver::Rusanov;
    params.precision = prec == "fp32bulk" ? Precision::FP32Bulk : Precision::FP64;

    FRSolver solver;
    solver.initialize(mesh, params);
    const Real g = params.gamma;
    IdealGas gas(g);
    if (vortex) {
        const Real beta = 5.0;
        solver.setInitialCondition([&](Vec2 x) {
            const Real r2 = x.x * x.x + x.y * x.y, f = std::exp(0.5 * (1.0 - r2));
            const Real T = 1.0 - (g - 1.0) * beta * beta / (8.0 * g * M_PI * M_PI) * std::exp(1.0 - r2);
            const Real rho = std::pow(T, 1.0 / (g - 1.0));
            return gas.primToConserv(rho, 1.0 - beta / (2 * M_PI) * x.y * f, beta / (2 * M_PI) * x.x * f,
                                     std::pow(rho, g));
        });
    } else {
        const Real mach = 0.1;
        solver.setInitialCondition([&](Vec2 x) {
            return gas.primToConserv(1.0, std::sin(x.x) * std::cos(x.y), -std::cos(x.x) * std::sin(x.y),
                                     1.0 / (g * mach * mach) + 0.25 * (std::cos(2 * x.x) + std::cos(2 * x.y)));
        });
    }
    solver.copyToHost();
    dump(prefix + "_U0.bin", state(solver));

    std::vector<double> dts;
    Real residual = 0.0;
    for (int it = 0; it < steps; ++it) {
        Real dt = solver.computeTimeStep();
        if (params.dt > 0) dt = std::min(dt, params.dt);
        solver.advance(dt);
        residual = solver.computeResidual();
        dts.push_back(dt);
    }
    solver.copyToHost();
    dump(prefix + "_U.bin", state(solver));
    std::vector<std::vector<Real>> pr, ma;
    solver.computeDerivedQuantities(pr, ma);

    std::printf("{\"steps\": %d, \"iteration\": %d, \"time\": %.17g, \"residual\": %.17g, \"E\": %lld, \"S\": %d, "
                "\"p00\": %.17g, \"dt\": [",
                steps, solver.iteration(), solver.currentTime(), residual, (long long)mesh.numElements(),
                (int)solver.solution()[0].sol_pts.size(), pr[0][0]);
    for (size_t k = 0; k < dts.size(); ++k) std::printf("%s%.17g", k ? ", " : "", dts[k]);
    std::printf("]}\n");
    return 0;
}
