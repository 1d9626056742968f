// ref_shim.cpp -- C entry points over the reference's own host sources.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's
// src/basis/basis.cpp, src/mesh/mesh.cpp and src/mesh/gmsh_reader.cpp (read in
// place under /root/reference by oracle/Makefile) into oracle/_ref/libzjref.so.
// tests/golden/make_golden.py calls it to write the golden fixtures; the
// oracle tests compare against it directly when it is present.  Nothing here
// restates the reference: every value comes from the reference's own code.
#include "basis/basis.hpp"
#include "mesh/mesh.hpp"
#include "physics/euler.hpp"
#include "physics/navierstokes.hpp"

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace zhijian;

extern "C" {

int ref_gauss(int n, double* pts, double* wts) {
    std::vector<Real> p, w;
    GaussLegendre::getQuadrature(n, p, w);
    for (int i = 0; i < n; ++i) { pts[i] = p[i]; wts[i] = w[i]; }
    return 0;
}

double ref_corr_deriv(int flux_type, int p, double xi, int right) {
    FluxType ft = static_cast<FluxType>(flux_type);
    return right ? CorrectionFunction::evalRightDeriv(ft, p, xi)
                 : CorrectionFunction::evalLeftDeriv(ft, p, xi);
}

// Flattening as in FRSolver::allocateGPUMemory (fr_solver.cpp:164-221).
int ref_quad_operators(int p, int flux_type, double* sol_xi, double* sol_eta, double* qw,
                       double* diff_xi, double* diff_eta, double* interp, double* corr) {
    FROperators ops;
    ops.init(ElementType::Quadrilateral, p, static_cast<FluxType>(flux_type));
    int S = ops.numSolutionPoints(), Fp = ops.numFluxPointsPerEdge();
    for (int i = 0; i < S; ++i) {
        sol_xi[i] = ops.solutionXi()[i];
        sol_eta[i] = ops.solutionEta()[i];
        qw[i] = ops.quadWeights()[i];
        for (int j = 0; j < S; ++j) {
            diff_xi[i * S + j] = ops.diffXi()[i][j];
            diff_eta[i * S + j] = ops.diffEta()[i][j];
        }
    }
    for (int e = 0; e < 4; ++e)
        for (int f = 0; f < Fp; ++f)
            for (int s = 0; s < S; ++s) interp[(e * Fp + f) * S + s] = ops.interpToFlux(e)[f][s];
    for (int e = 0; e < 4; ++e)
        for (int s = 0; s < S; ++s) corr[e * S + s] = ops.correctionDeriv()[e][s];
    return S;
}

// A mesh built by the reference's Mesh class.
struct RefMesh {
    Mesh mesh;
};

static void* finish(std::unique_ptr<RefMesh> m) { return m.release(); }

// Corner-only quads/triangles from arrays (order 1 elements).
void* ref_mesh_from_arrays(int n_nodes, const double* xy, int n_elem, int nv,
                           const int64_t* elem_nodes) {
    auto m = std::make_unique<RefMesh>();
    for (int i = 0; i < n_nodes; ++i) m->mesh.nodes().push_back(Vec2(xy[2 * i], xy[2 * i + 1]));
    for (int e = 0; e < n_elem; ++e) {
        Element el;
        el.type = (nv == 3) ? ElementType::Triangle : ElementType::Quadrilateral;
        el.order = 1;
        el.partition = 0;
        for (int k = 0; k < nv; ++k) el.node_ids.push_back(elem_nodes[(int64_t)e * nv + k]);
        m->mesh.elements().push_back(el);
    }
    m->mesh.buildFaces();
    m->mesh.computeGeometry();
    return finish(std::move(m));
}

void* ref_mesh_read_gmsh(const char* path) {
    auto m = std::make_unique<RefMesh>();
    GmshReader reader;
    if (!reader.read(path, m->mesh)) return nullptr;
    return finish(std::move(m));
}

void ref_mesh_free(void* h) { delete static_cast<RefMesh*>(h); }

int ref_mesh_counts(void* h, int* n_nodes, int* n_elem, int* n_faces, int* max_nodes_per_elem) {
    const Mesh& m = static_cast<RefMesh*>(h)->mesh;
    *n_nodes = (int)m.numNodes();
    *n_elem = (int)m.numElements();
    *n_faces = (int)m.numFaces();
    int mx = 0;
    for (const auto& e : m.elements()) mx = std::max(mx, (int)e.node_ids.size());
    *max_nodes_per_elem = mx;
    return 0;
}

// elem_nodes [E][stride] padded with -1; elem_info [E][3] = (type, order, n_nodes)
int ref_mesh_dump(void* h, double* xy, int stride, int64_t* elem_nodes, int* elem_info,
                  int* face_le, int* face_ll, int* face_re, int* face_rl, int* face_bc_type,
                  int* face_bc_tag, double* normals, double* area, double* face_len) {
    const Mesh& m = static_cast<RefMesh*>(h)->mesh;
    for (Index i = 0; i < m.numNodes(); ++i) { xy[2 * i] = m.node(i).x; xy[2 * i + 1] = m.node(i).y; }
    for (Index e = 0; e < m.numElements(); ++e) {
        const Element& el = m.element(e);
        for (int k = 0; k < stride; ++k)
            elem_nodes[e * stride + k] = k < (int)el.node_ids.size() ? el.node_ids[k] : -1;
        elem_info[3 * e] = static_cast<int>(el.type);
        elem_info[3 * e + 1] = el.order;
        elem_info[3 * e + 2] = (int)el.node_ids.size();
        area[e] = m.elementArea(e);
    }
    for (Index f = 0; f < m.numFaces(); ++f) {
        const Face& fc = m.face(f);
        face_le[f] = fc.left_elem; face_ll[f] = fc.left_face;
        face_re[f] = fc.right_elem; face_rl[f] = fc.right_face;
        face_bc_type[f] = static_cast<int>(fc.bc_type);
        face_bc_tag[f] = fc.bc_tag;
        Vec2 n = m.faceNormal(f);
        normals[2 * f] = n.x; normals[2 * f + 1] = n.y;
        face_len[f] = m.faceLength(f);
    }
    return 0;
}

// FRSolver::initGeometry (fr_solver.cpp:85-129) for every element of the mesh.
int ref_mesh_geometry(void* h, int p, double* J, double* Jinv) {
    const Mesh& m = static_cast<RefMesh*>(h)->mesh;
    FROperators ops;
    ops.init(ElementType::Quadrilateral, p, FluxType::DG);
    int S = ops.numSolutionPoints();
    size_t o = 0;
    for (Index e = 0; e < m.numElements(); ++e) {
        const Element& el = m.element(e);
        std::vector<Vec2> nodes(el.node_ids.size());
        for (size_t j = 0; j < el.node_ids.size(); ++j) nodes[j] = m.node(el.node_ids[j]);
        for (int sp = 0; sp < S; ++sp, ++o) {
            Real a, b, c, d;
            ElementGeometry::jacobian(nodes, el.type, el.order, ops.solutionXi()[sp],
                                      ops.solutionEta()[sp], a, b, c, d);
            Real Jd = a * d - b * c;
            Real inv = 1.0 / Jd;
            J[o] = Jd;
            Jinv[4 * o] = d * inv; Jinv[4 * o + 1] = -b * inv;
            Jinv[4 * o + 2] = -c * inv; Jinv[4 * o + 3] = a * inv;
        }
    }
    return 0;
}

// Physics header vectors.
void ref_flux_xy(const double* U, double gamma, double* Fx, double* Fy) {
    IdealGas gas(gamma);
    State s;
    for (int v = 0; v < N_VARS; ++v) s[v] = U[v];
    State fx = gas.fluxX(s), fy = gas.fluxY(s);
    for (int v = 0; v < N_VARS; ++v) { Fx[v] = fx[v]; Fy[v] = fy[v]; }
}

double ref_max_wave_speed(const double* U, double gamma) {
    IdealGas gas(gamma);
    State s;
    for (int v = 0; v < N_VARS; ++v) s[v] = U[v];
    return gas.maxWaveSpeed(s);
}

void ref_riemann(int which, const double* UL, const double* UR, double nx, double ny,
                 double gamma, double* F) {
    State l, r, out;
    for (int v = 0; v < N_VARS; ++v) { l[v] = UL[v]; r[v] = UR[v]; }
    if (which == 0) out = RiemannSolvers::rusanov(l, r, nx, ny, gamma);
    else if (which == 1) out = RiemannSolvers::roe(l, r, nx, ny, gamma);
    else out = RiemannSolvers::hllc(l, r, nx, ny, gamma);
    for (int v = 0; v < N_VARS; ++v) F[v] = out[v];
}

// ViscousFlux at (U, dU/dx, dU/dy): velocityGradients, temperatureGradient,
// fluxX, fluxY, and the viscosity and conductivity the two fluxes used.
void ref_visc_flux(const double* U, const double* dUdx, const double* dUdy, double gamma, double Re,
                   double Pr, double* Fx, double* Fy, double* mu_k) {
    ViscousFlux vf(gamma, Re, Pr);
    IdealGas gas(gamma);
    State s, gx, gy;
    for (int v = 0; v < N_VARS; ++v) { s[v] = U[v]; gx[v] = dUdx[v]; gy[v] = dUdy[v]; }
    Real dudx, dudy, dvdx, dvdy, dTdx, dTdy;
    ViscousFlux::velocityGradients(s, gx, gy, dudx, dudy, dvdx, dvdy);
    vf.temperatureGradient(s, gx, gy, dTdx, dTdy);
    State fx = vf.fluxX(s, dudx, dudy, dvdx, dvdy, dTdx, dTdy);
    State fy = vf.fluxY(s, dudx, dudy, dvdx, dvdy, dTdx, dTdy);
    for (int v = 0; v < N_VARS; ++v) { Fx[v] = fx[v]; Fy[v] = fy[v]; }
    const Real T = gas.pressure(s) / s.rho();
    const Real mu = vf.dynamicViscosity(T) / Re;
    mu_k[0] = mu;
    mu_k[1] = vf.thermalConductivity(mu) / Re;
}

}  // extern "C"
