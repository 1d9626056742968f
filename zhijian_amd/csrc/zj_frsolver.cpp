// zj_frsolver.cpp -- the C++ drop-in API (include/zhijian/*.hpp) over the C
// ABI (include/zhijian_amd.h).
//
// Host-side behaviour follows the reference where the API promises it:
//   Mesh            src/mesh/mesh.cpp:10-194
//   basis classes   src/basis/basis.cpp:12-934
//   FRSolver        src/solver/fr_solver.cpp:24-892 (upload set, BC tables,
//                   advance, computeTimeStep, computeResidual, host mirror)
//   SSPRK3 etc.     include/time/rk_integrator.hpp:11-79 (declared there only)
#include "zhijian/basis.hpp"
#include "zhijian/fr_solver.hpp"
#include "zhijian/mesh.hpp"
#include "zhijian/rk_integrator.hpp"
#include "zhijian_amd.h"
#include "zj_host_math.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>

namespace zhijian {

namespace {
std::vector<int64_t> corners(const std::vector<Element>& els, int& nv) {
    nv = els.empty() ? 4 : els[0].numVertices();
    std::vector<int64_t> out;
    out.reserve(els.size() * nv);
    for (const Element& e : els) {
        if (e.numVertices() != nv) throw std::runtime_error("mixed triangle/quad meshes are not supported");
        for (int k = 0; k < nv; ++k) out.push_back(e.node_ids[k]);
    }
    return out;
}
std::vector<double> flat_xy(const std::vector<Vec2>& nodes) {
    std::vector<double> xy(nodes.size() * 2);
    for (size_t i = 0; i < nodes.size(); ++i) { xy[2 * i] = nodes[i].x; xy[2 * i + 1] = nodes[i].y; }
    return xy;
}
void need(int rc, const char* what) {
    if (rc != ZJ_OK) throw std::runtime_error(std::string(what) + ": " + zj_last_error(nullptr));
}
}  // namespace

// =================================================================== Mesh
Index Mesh::numBoundaryFaces() const {
    Index c = 0;
    for (const Face& f : faces_) c += f.is_boundary ? 1 : 0;
    return c;
}

void Mesh::rebuildFaceIds() {
    for (Element& e : elements_) e.face_ids.clear();
    for (size_t f = 0; f < faces_.size(); ++f) {
        elements_[faces_[f].left_elem].face_ids.push_back((LocalIndex)f);
        if (!faces_[f].is_boundary) elements_[faces_[f].right_elem].face_ids.push_back((LocalIndex)f);
    }
}

void Mesh::buildFaces() {
    faces_.clear();
    int nv;
    std::vector<int64_t> en = corners(elements_, nv);
    const int E = (int)elements_.size();
    std::vector<int> le(std::max(E * nv, 1)), ll(le.size()), re(le.size()), rl(le.size());
    int nf = 0;
    need(zj_build_faces(E, nv, nv, en.data(), &nf, le.data(), ll.data(), re.data(), rl.data()), "buildFaces");
    faces_.reserve(nf);
    for (int f = 0; f < nf; ++f) {
        Face fc;
        fc.left_elem = le[f]; fc.left_face = ll[f];
        fc.right_elem = re[f]; fc.right_face = rl[f];
        fc.is_boundary = re[f] < 0;
        fc.bc_type = fc.is_boundary ? BCType::FarField : BCType::Interior;  // mesh.cpp:55,65
        fc.bc_tag = fc.is_boundary ? 0 : -1;
        fc.mpi_rank = -1;
        faces_.push_back(fc);
    }
    rebuildFaceIds();
}

void Mesh::pairPeriodic(Real Lx, Real Ly, Real tol) {
    if (tol < 0) tol = 1e-9 * std::max(Lx, Ly);
    pairFaces([&](int* nf, int* le, int* ll, int* re, int* rl, int nv, const int64_t* en) {
        const std::vector<double> xy = flat_xy(nodes_);
        return zj_pair_periodic(nf, le, ll, re, rl, nv, nv, en, xy.data(), Lx, Ly, tol);
    }, "pairPeriodic");
}

void Mesh::pairPeriodicNodes() { pairPeriodicNodes(periodic_nodes_); }

void Mesh::pairPeriodicNodes(const std::vector<std::pair<Index, Index>>& links) {
    std::vector<int64_t> slave(links.size()), master(links.size());
    for (size_t i = 0; i < links.size(); ++i) {
        slave[i] = links[i].first;
        master[i] = links[i].second;
    }
    const std::vector<double> xy = flat_xy(nodes_);
    double lo[2] = {0, 0}, hi[2] = {0, 0};
    for (size_t i = 0; i < nodes_.size(); ++i)
        for (int d = 0; d < 2; ++d) {
            lo[d] = i ? std::min(lo[d], xy[2 * i + d]) : xy[2 * i + d];
            hi[d] = i ? std::max(hi[d], xy[2 * i + d]) : xy[2 * i + d];
        }
    const double tol = 1e-9 * std::max(hi[0] - lo[0], hi[1] - lo[1]);
    pairFaces([&](int* nf, int* le, int* ll, int* re, int* rl, int nv, const int64_t* en) {
        return zj_pair_periodic_nodes(nf, le, ll, re, rl, nv, nv, (int)elements_.size(), en, xy.data(), numNodes(),
                                      (int64_t)links.size(), slave.data(), master.data(), tol);
    }, "pairPeriodicNodes");
}

// Runs a pairing routine on the face table and keeps each paired face's BC
// fields as an interior face's.
void Mesh::pairFaces(const std::function<int(int*, int*, int*, int*, int*, int, const int64_t*)>& pair,
                     const char* what) {
    int nv;
    std::vector<int64_t> en = corners(elements_, nv);
    const int F = (int)faces_.size();
    std::vector<int> le(std::max(F, 1)), ll(le.size()), re(le.size()), rl(le.size());
    for (int f = 0; f < F; ++f) {
        le[f] = faces_[f].left_elem; ll[f] = faces_[f].left_face;
        re[f] = faces_[f].right_elem; rl[f] = faces_[f].right_face;
    }
    int nf = F;
    need(pair(&nf, le.data(), ll.data(), re.data(), rl.data(), nv, en.data()), what);
    std::vector<Face> out;
    out.reserve(nf);
    size_t src = 0;
    for (int f = 0; f < nf; ++f) {
        while (faces_[src].left_elem != le[f] || faces_[src].left_face != ll[f]) ++src;
        Face fc = faces_[src++];
        if (fc.is_boundary && re[f] >= 0) {
            fc.right_elem = re[f]; fc.right_face = rl[f];
            fc.is_boundary = false; fc.bc_type = BCType::Interior; fc.bc_tag = -1;
        }
        out.push_back(fc);
    }
    faces_ = std::move(out);
    rebuildFaceIds();
}

void Mesh::computeGeometry() {
    const Index E = numElements(), F = numFaces();
    elem_centroids_.resize(E); elem_areas_.resize(E);
    face_lengths_.resize(F); face_normals_.resize(F); face_centroids_.resize(F);
    for (Index i = 0; i < E; ++i) { elem_centroids_[i] = elementCentroid(i); elem_areas_[i] = elementArea(i); }
    for (Index i = 0; i < F; ++i) {
        face_lengths_[i] = faceLength(i); face_normals_[i] = faceNormal(i); face_centroids_[i] = faceCentroid(i);
    }
}

Vec2 Mesh::elementCentroid(Index e) const {
    const Element& el = elements_[e];
    Vec2 c(0, 0);
    const int nv = el.numVertices();
    for (int i = 0; i < nv; ++i) c = c + nodes_[el.node_ids[i]];
    return c / static_cast<Real>(nv);
}

Real Mesh::elementArea(Index e) const {
    const Element& el = elements_[e];
    auto tri = [&](Index a, Index b, Index c) {
        const Vec2 &p0 = nodes_[a], &p1 = nodes_[b], &p2 = nodes_[c];
        return 0.5 * std::abs((p1.x - p0.x) * (p2.y - p0.y) - (p2.x - p0.x) * (p1.y - p0.y));
    };
    if (el.type == ElementType::Triangle) return tri(el.node_ids[0], el.node_ids[1], el.node_ids[2]);
    return tri(el.node_ids[0], el.node_ids[1], el.node_ids[2]) + tri(el.node_ids[0], el.node_ids[2], el.node_ids[3]);
}

Real Mesh::faceLength(Index f) const {
    const Face& fc = faces_[f];
    const Element& el = elements_[fc.left_elem];
    const int nv = el.numVertices();
    return (nodes_[el.node_ids[(fc.left_face + 1) % nv]] - nodes_[el.node_ids[fc.left_face]]).norm();
}

Vec2 Mesh::faceNormal(Index f) const {
    const Face& fc = faces_[f];
    const Element& el = elements_[fc.left_elem];
    const int nv = el.numVertices();
    const Vec2 t = nodes_[el.node_ids[(fc.left_face + 1) % nv]] - nodes_[el.node_ids[fc.left_face]];
    const Real len = t.norm();
    Vec2 n(t.y / len, -t.x / len);
    if (n.dot(faceCentroid(f) - elementCentroid(fc.left_elem)) < 0) n = n * (-1.0);
    return n;
}

Vec2 Mesh::faceCentroid(Index f) const {
    const Face& fc = faces_[f];
    const Element& el = elements_[fc.left_elem];
    const int nv = el.numVertices();
    return (nodes_[el.node_ids[fc.left_face]] + nodes_[el.node_ids[(fc.left_face + 1) % nv]]) * 0.5;
}

void Mesh::printStats() const {
    std::cout << "=== Mesh Statistics ===\nNodes: " << numNodes() << "\nElements: " << numElements()
              << "\nFaces: " << numFaces() << "\nBoundary faces: " << numBoundaryFaces() << "\n";
}

// ================================================================== basis
void GaussLegendre::getQuadrature(int n, std::vector<Real>& points, std::vector<Real>& weights) {
    points.resize(n); weights.resize(n);
    if (n > 0) zj::host::gauss_rule(n, points.data(), weights.data());
}

void GaussLegendre::getGLL(int n, std::vector<Real>& points, std::vector<Real>& weights) {
    if (n < 2) throw std::invalid_argument("GLL requires at least 2 points");
    points.assign(n, 0.0); weights.assign(n, 0.0);
    points[0] = -1.0; points[n - 1] = 1.0;
    // interior nodes: roots of P'_{n-1}, Newton from Chebyshev-Gauss-Lobatto guesses
    for (int i = 1; i < n - 1; ++i) {
        Real x = -std::cos(M_PI * i / (n - 1));
        for (int it = 0; it < 50; ++it) {
            const Real dP = zj::host::legendre_dp(n - 1, x);
            const Real d2P = (2.0 * x * dP - (n - 1) * n * zj::host::legendre_p(n - 1, x)) / (1.0 - x * x);
            const Real dx = dP / d2P;
            x -= dx;
            if (std::abs(dx) < 1e-15) break;
        }
        points[i] = x;
    }
    std::sort(points.begin(), points.end());
    for (int i = 0; i < n; ++i) {
        const Real P = zj::host::legendre_p(n - 1, points[i]);
        weights[i] = 2.0 / (n * (n - 1) * P * P);
    }
}

Real Lagrange1D::eval(const std::vector<Real>& nodes, int j, Real x) {
    return zj::host::lag(nodes.data(), (int)nodes.size(), j, x);
}
Real Lagrange1D::evalDeriv(const std::vector<Real>& nodes, int j, Real x) {
    return zj::host::lag_d(nodes.data(), (int)nodes.size(), j, x);
}
void Lagrange1D::evalAll(const std::vector<Real>& nodes, Real x, std::vector<Real>& values) {
    values.resize(nodes.size());
    for (size_t j = 0; j < nodes.size(); ++j) values[j] = eval(nodes, (int)j, x);
}
void Lagrange1D::evalAllDeriv(const std::vector<Real>& nodes, Real x, std::vector<Real>& derivs) {
    derivs.resize(nodes.size());
    for (size_t j = 0; j < nodes.size(); ++j) derivs[j] = evalDeriv(nodes, (int)j, x);
}
void Lagrange1D::buildVandermonde(const std::vector<Real>& nodes, std::vector<std::vector<Real>>& V) {
    const size_t n = nodes.size();
    V.assign(n, std::vector<Real>(n, 1.0));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 1; j < n; ++j) V[i][j] = V[i][j - 1] * nodes[i];
}
void Lagrange1D::buildDiffMatrix(const std::vector<Real>& nodes, std::vector<std::vector<Real>>& D) {
    const int n = (int)nodes.size();
    D.assign(n, std::vector<Real>(n));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) D[i][j] = evalDeriv(nodes, j, nodes[i]);
}

// The reference evaluates the Radau (DG) functions for every FluxType
// (basis.cpp:263-379); getC returns its c values, which its operators never use.
Real CorrectionFunction::evalLeft(FluxType, int p, Real xi) {
    const Real s = (p % 2 == 0) ? 1.0 : -1.0;
    return s * 0.5 * (zj::host::legendre_p(p, xi) - zj::host::legendre_p(p + 1, xi));
}
Real CorrectionFunction::evalRight(FluxType, int p, Real xi) {
    return 0.5 * (zj::host::legendre_p(p, xi) + zj::host::legendre_p(p + 1, xi));
}
Real CorrectionFunction::evalLeftDeriv(FluxType, int p, Real xi) { return zj::host::radau_left_d(p, xi); }
Real CorrectionFunction::evalRightDeriv(FluxType, int p, Real xi) { return zj::host::radau_right_d(p, xi); }
Real CorrectionFunction::getC(FluxType type, int p) {
    const Real fp = std::tgamma(p + 1.0);
    switch (type) {
    case FluxType::SD:
    case FluxType::HU: return 2.0 * (p + 1) / ((2 * p + 1) * p * fp * fp);
    case FluxType::GA: return 2.0 * (p + 1) * fp * fp / (std::tgamma(2.0 * p + 1.0) * (2 * p + 1));
    default: return 0.0;
    }
}

void SolutionPoints::get1D(int order, std::vector<Real>& points) {
    std::vector<Real> w;
    GaussLegendre::getQuadrature(order + 1, points, w);
}
void SolutionPoints::getTriangle(int order, std::vector<Real>& xi, std::vector<Real>& eta) {
    std::vector<Real> g, w;
    GaussLegendre::getQuadrature(order + 1, g, w);
    xi.clear(); eta.clear();
    for (int j = 0; j <= order; ++j)
        for (int i = 0; i <= order - j; ++i) {
            const Real s = 0.5 * (1.0 + g[i]), t = 0.5 * (1.0 + g[j]);
            xi.push_back(s * (1.0 - t)); eta.push_back(t);
        }
}
void SolutionPoints::getQuad(int order, std::vector<Real>& xi, std::vector<Real>& eta) {
    std::vector<Real> g, w;
    GaussLegendre::getQuadrature(order + 1, g, w);
    const int n = order + 1;
    xi.resize(n * n); eta.resize(n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) { xi[j * n + i] = g[i]; eta[j * n + i] = g[j]; }
}
void FluxPoints::get1D(int order, std::vector<Real>& points) { SolutionPoints::get1D(order, points); }

void FROperators::init(ElementType type, int p, FluxType ft) {
    elem_type_ = type; poly_order_ = p; flux_type_ = ft;
    num_flux_pts_per_edge_ = p + 1;
    std::vector<Real> g;
    FluxPoints::get1D(p, g);
    if (type == ElementType::Quadrilateral) {
        if (p < 1 || p > MAX_POLY_ORDER) throw std::invalid_argument("poly_order must be 1..5");
        const int n = p + 1, S = n * n;
        num_sol_pts_ = S; num_edges_ = 4;
        sol_xi_.resize(S); sol_eta_.resize(S); quad_weights_.resize(S);
        std::vector<double> dx(S * S), de(S * S), in(4 * n * S), co(4 * S);
        need(zj_quad_operators(p, sol_xi_.data(), sol_eta_.data(), quad_weights_.data(), dx.data(), de.data(),
                               in.data(), co.data()), "FROperators");
        diff_xi_.assign(S, std::vector<Real>(S)); diff_eta_.assign(S, std::vector<Real>(S));
        for (int r = 0; r < S; ++r)
            for (int c = 0; c < S; ++c) { diff_xi_[r][c] = dx[r * S + c]; diff_eta_[r][c] = de[r * S + c]; }
        interp_to_flux_.assign(4, std::vector<std::vector<Real>>(n, std::vector<Real>(S)));
        for (int e = 0; e < 4; ++e)
            for (int f = 0; f < n; ++f)
                for (int s = 0; s < S; ++s) interp_to_flux_[e][f][s] = in[(e * n + f) * S + s];
        correction_deriv_.assign(4, std::vector<Real>(S));
        for (int e = 0; e < 4; ++e)
            for (int s = 0; s < S; ++s) correction_deriv_[e][s] = co[e * S + s];
        flux_pt_loc_.assign(4, g);
    } else {
        // the reference's triangle placeholders (basis.cpp:615-669)
        num_sol_pts_ = (p + 1) * (p + 2) / 2; num_edges_ = 3;
        SolutionPoints::getTriangle(p, sol_xi_, sol_eta_);
        quad_weights_.assign(num_sol_pts_, 0.5 / num_sol_pts_);
        flux_pt_loc_.assign(3, std::vector<Real>(p + 1));
        for (int e = 0; e < 3; ++e)
            for (int f = 0; f <= p; ++f) flux_pt_loc_[e][f] = 0.5 * (1.0 + g[f]);
        interp_to_flux_.assign(3, std::vector<std::vector<Real>>(p + 1, std::vector<Real>(num_sol_pts_, 0.0)));
        correction_deriv_.assign(3, std::vector<Real>(num_sol_pts_, 0.0));
        diff_xi_.assign(num_sol_pts_, std::vector<Real>(num_sol_pts_, 0.0));
        diff_eta_ = diff_xi_;
    }
}

void ElementGeometry::shapeFunc(ElementType type, int order, Real xi, Real eta, std::vector<Real>& N) {
    std::vector<Real> a, b;
    if (type == ElementType::Triangle) {
        const Real L0 = 1.0 - xi - eta, L1 = xi, L2 = eta;
        if (order == 1) N = {L0, L1, L2};
        else N = {L0 * (2 * L0 - 1), L1 * (2 * L1 - 1), L2 * (2 * L2 - 1), 4 * L0 * L1, 4 * L1 * L2, 4 * L2 * L0};
        return;
    }
    const int nc = order == 1 ? 4 : order == 2 ? 9 : order == 4 ? 25 : 0;
    N.assign(nc, 0.0); a.assign(nc, 0.0); b.assign(nc, 0.0);
    if (!nc || !zj::host::shape(nc, xi, eta, N.data(), a.data(), b.data()))
        throw std::invalid_argument("unsupported quad geometric order");
}

void ElementGeometry::shapeFuncDeriv(ElementType type, int order, Real xi, Real eta, std::vector<Real>& dNdxi,
                                     std::vector<Real>& dNdeta) {
    if (type == ElementType::Triangle) {
        if (order == 1) {
            dNdxi = {-1.0, 1.0, 0.0};
            dNdeta = {-1.0, 0.0, 1.0};
        } else {
            const Real L0 = 1.0 - xi - eta, L1 = xi, L2 = eta;
            dNdxi = {-(4 * L0 - 1), 4 * L1 - 1, 0, 4 * (L0 - L1), 4 * L2, -4 * L2};
            dNdeta = {-(4 * L0 - 1), 0, 4 * L2 - 1, -4 * L1, 4 * L1, 4 * (L0 - L2)};
        }
        return;
    }
    const int nc = order == 1 ? 4 : order == 2 ? 9 : order == 4 ? 25 : 0;
    std::vector<Real> N(nc);
    dNdxi.assign(nc, 0.0); dNdeta.assign(nc, 0.0);
    if (!nc || !zj::host::shape(nc, xi, eta, N.data(), dNdxi.data(), dNdeta.data()))
        throw std::invalid_argument("unsupported quad geometric order");
}

void ElementGeometry::jacobian(const std::vector<Vec2>& nodes, ElementType type, int order, Real xi, Real eta,
                               Real& dxdxi, Real& dxdeta, Real& dydxi, Real& dydeta) {
    std::vector<Real> a, b;
    shapeFuncDeriv(type, order, xi, eta, a, b);
    dxdxi = dxdeta = dydxi = dydeta = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        dxdxi += a[i] * nodes[i].x; dxdeta += b[i] * nodes[i].x;
        dydxi += a[i] * nodes[i].y; dydeta += b[i] * nodes[i].y;
    }
}

Real ElementGeometry::jacobianDet(const std::vector<Vec2>& nodes, ElementType type, int order, Real xi, Real eta) {
    Real a, b, c, d;
    jacobian(nodes, type, order, xi, eta, a, b, c, d);
    return a * d - b * c;
}

void ElementGeometry::jacobianInverse(Real dxdxi, Real dxdeta, Real dydxi, Real dydeta, Real& dxidx, Real& dxidy,
                                      Real& detadx, Real& detady) {
    const Real inv = 1.0 / (dxdxi * dydeta - dxdeta * dydxi);
    dxidx = dydeta * inv; dxidy = -dxdeta * inv; detadx = -dydxi * inv; detady = dxdxi * inv;
}

Vec2 ElementGeometry::refToPhys(const std::vector<Vec2>& nodes, ElementType type, int order, Real xi, Real eta) {
    std::vector<Real> N;
    shapeFunc(type, order, xi, eta, N);
    Vec2 p(0, 0);
    for (size_t i = 0; i < N.size(); ++i) { p.x += N[i] * nodes[i].x; p.y += N[i] * nodes[i].y; }
    return p;
}

Vec2 ElementGeometry::refFaceNormal(ElementType type, int face_id) {
    if (type == ElementType::Triangle) {
        switch (face_id) {
        case 0: return Vec2(0, -1);
        case 1: return Vec2(1, 1) / std::sqrt(2.0);
        case 2: return Vec2(-1, 0);
        default: return Vec2(0, 0);
        }
    }
    switch (face_id) {
    case 0: return Vec2(0, -1);
    case 1: return Vec2(1, 0);
    case 2: return Vec2(0, 1);
    case 3: return Vec2(-1, 0);
    default: return Vec2(0, 0);
    }
}

void ElementGeometry::faceParamToRef(ElementType type, int face_id, Real s, Real& xi, Real& eta) {
    if (type == ElementType::Triangle) {
        switch (face_id) {
        case 0: xi = s; eta = 0; break;
        case 1: xi = 1 - s; eta = s; break;
        default: xi = 0; eta = 1 - s; break;
        }
        return;
    }
    switch (face_id) {
    case 0: xi = s; eta = -1; break;
    case 1: xi = 1; eta = s; break;
    case 2: xi = -s; eta = 1; break;
    default: xi = -1; eta = -s; break;
    }
}

// ============================================================ integrators
void SSPRK3::stage1_CPU(Real* U1, const Real* U0, const Real* R, int n, Real dt) {
    for (int i = 0; i < n; ++i) U1[i] = U0[i] + dt * R[i];
}
void SSPRK3::stage2_CPU(Real* U2, const Real* U0, const Real* U1, const Real* R, int n, Real dt) {
    for (int i = 0; i < n; ++i) U2[i] = 0.75 * U0[i] + 0.25 * U1[i] + 0.25 * dt * R[i];
}
void SSPRK3::stage3_CPU(Real* U, const Real* U0, const Real* U2, const Real* R, int n, Real dt) {
    for (int i = 0; i < n; ++i) U[i] = (1.0 / 3.0) * U0[i] + (2.0 / 3.0) * U2[i] + (2.0 / 3.0) * dt * R[i];
}
void SSPRK3::stage1_GPU(Real* U1, const Real* U0, const Real* R, int n, Real dt, hipStream_t s) {
    need(zj_rk3_combine(1, U1, U0, nullptr, R, n, dt, s), "SSPRK3::stage1_GPU");
}
void SSPRK3::stage2_GPU(Real* U2, const Real* U0, const Real* U1, const Real* R, int n, Real dt, hipStream_t s) {
    need(zj_rk3_combine(2, U2, U0, U1, R, n, dt, s), "SSPRK3::stage2_GPU");
}
void SSPRK3::stage3_GPU(Real* U, const Real* U0, const Real* U2, const Real* R, int n, Real dt, hipStream_t s) {
    need(zj_rk3_combine(3, U, U0, U2, R, n, dt, s), "SSPRK3::stage3_GPU");
}
void ForwardEuler::advance_CPU(Real* U, const Real* R, int n, Real dt) {
    for (int i = 0; i < n; ++i) U[i] += dt * R[i];
}
void ForwardEuler::advance_GPU(Real* U, const Real* R, int n, Real dt, hipStream_t s) {
    need(zj_euler_update(U, R, n, dt, s), "ForwardEuler::advance_GPU");
}
void RK4::advance_CPU(Real* U, Real* k1, Real* k2, Real* k3, Real* k4, const Real*, int n, Real dt,
                      const std::function<void(Real*, Real*)>& rhs) {
    std::vector<Real> tmp(n);
    rhs(U, k1);
    for (int i = 0; i < n; ++i) tmp[i] = U[i] + 0.5 * dt * k1[i];
    rhs(tmp.data(), k2);
    for (int i = 0; i < n; ++i) tmp[i] = U[i] + 0.5 * dt * k2[i];
    rhs(tmp.data(), k3);
    for (int i = 0; i < n; ++i) tmp[i] = U[i] + dt * k3[i];
    rhs(tmp.data(), k4);
    for (int i = 0; i < n; ++i) U[i] += dt / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
}

// ================================================================ FRSolver
FRSolver::FRSolver() = default;

FRSolver::~FRSolver() {
    if (handle_) zj_destroy(handle_);
}

// ------------------------------------------------------------- GmshReader
bool GmshReader::read(const std::string& filename, Mesh& mesh) {
    zj_gmsh_handle g = nullptr;
    if (zj_gmsh_read(filename.c_str(), &g) != ZJ_OK) {
        error_msg_ = zj_gmsh_error();
        return false;
    }
    int64_t nn = 0;
    int ne = 0, mx = 0, ng = 0;
    zj_gmsh_counts(g, &nn, &ne, &mx, &ng);
    std::vector<double> xy((size_t)nn * 2);
    zj_gmsh_nodes(g, xy.data());
    std::vector<int> kind(ne), order(ne), cnt(ne);
    std::vector<int64_t> nodes((size_t)ne * mx);
    zj_gmsh_elements(g, kind.data(), order.data(), cnt.data(), nodes.data(), mx);
    mesh = Mesh();
    for (int64_t i = 0; i < nn; ++i) mesh.nodes().push_back(Vec2(xy[2 * i], xy[2 * i + 1]));
    for (int e = 0; e < ne; ++e) {
        Element el;
        el.type = static_cast<ElementType>(kind[e]);
        el.order = order[e];
        el.node_ids.assign(nodes.begin() + (size_t)e * mx, nodes.begin() + (size_t)e * mx + cnt[e]);
        el.partition = 0;
        mesh.elements().push_back(el);
    }
    int64_t npl = 0;
    zj_gmsh_periodic(g, &npl, nullptr, nullptr);
    std::vector<int64_t> ps((size_t)npl), pm((size_t)npl);
    zj_gmsh_periodic(g, &npl, ps.data(), pm.data());
    for (int64_t i = 0; i < npl; ++i) mesh.periodicNodes().emplace_back(ps[(size_t)i], pm[(size_t)i]);
    for (int i = 0; i < ng; ++i) {
        int tag = 0, type = 0;
        char name[256];
        zj_gmsh_group(g, i, &tag, &type, name, (int)sizeof(name));
        BCInfo info{};
        info.tag = tag;
        info.name = name;
        info.type = static_cast<BCType>(type);
        info.flow_direction = Vec2(1, 0);
        mesh.addBCInfo(tag, info);
    }
    mesh.buildFaces();
    // GmshReader::applyBoundaryTags (gmsh_reader.cpp:736-758)
    for (Index f = 0; f < mesh.numFaces(); ++f) {
        Face& fc = mesh.face(f);
        if (!fc.is_boundary) continue;
        const Element& el = mesh.element(fc.left_elem);
        const int nv = el.numVertices();
        int tag = 0;
        zj_gmsh_edge_tag(g, el.node_ids[fc.left_face % nv], el.node_ids[(fc.left_face + 1) % nv], &tag);
        if (tag) {
            fc.bc_tag = tag;
            if (mesh.hasBCInfo(tag)) fc.bc_type = mesh.getBCInfo(tag).type;
        }
    }
    zj_gmsh_free(g);
    mesh.computeGeometry();
    return true;
}

void FRSolver::check(int rc, const char* what) const {
    if (rc != ZJ_OK) throw std::runtime_error(std::string("FRSolver::") + what + ": " + zj_last_error(handle_));
}

void FRSolver::initialize(const Mesh& mesh, const SimParams& params) {
    if (handle_) { zj_destroy(handle_); handle_ = nullptr; }
    mesh_ = &mesh;
    params_ = params;
    operators_[ElementType::Quadrilateral].init(ElementType::Quadrilateral, params.poly_order, params.flux_type);
    operators_[ElementType::Triangle].init(ElementType::Triangle, params.poly_order, params.flux_type);
    n_sp_quad_ = SolutionPoints::numPointsQuad(params.poly_order);
    n_sp_tri_ = SolutionPoints::numPointsTri(params.poly_order);
    const Index E = mesh.numElements(), F = mesh.numFaces();
    for (Index e = 0; e < E; ++e)
        if (mesh.element(e).type != ElementType::Quadrilateral)
            throw std::runtime_error("FRSolver::initialize: the device path supports quadrilaterals only "
                                     "(the reference's triangle operators are placeholders)");
    const int n_fp = FluxPoints::numPointsPerEdge(params.poly_order);
    solution_.assign(E, ElementSolution());
    for (Index e = 0; e < E; ++e) {
        solution_[e].sol_pts.resize(n_sp_quad_);
        solution_[e].flux_pts.resize(4 * n_fp);
        solution_[e].dUdx.resize(n_sp_quad_);
        solution_[e].dUdy.resize(n_sp_quad_);
    }
    const FROperators& ops = operators_.at(ElementType::Quadrilateral);
    const int S = ops.numSolutionPoints();
    std::vector<double> dx, de, in, co;
    for (int r = 0; r < S; ++r)
        for (int c = 0; c < S; ++c) { dx.push_back(ops.diffXi()[r][c]); de.push_back(ops.diffEta()[r][c]); }
    for (int k = 0; k < 4; ++k)
        for (int f = 0; f < n_fp; ++f)
            for (int s = 0; s < S; ++s) in.push_back(ops.interpToFlux(k)[f][s]);
    for (int k = 0; k < 4; ++k)
        for (int s = 0; s < S; ++s) co.push_back(ops.correctionDeriv()[k][s]);
    // geometry at solution points (fr_solver.cpp:85-129)
    std::vector<double> J, Jinv, area(E);
    J.reserve(E * S); Jinv.reserve(E * S * 4);
    for (Index e = 0; e < E; ++e) {
        const Element& el = mesh.element(e);
        std::vector<Vec2> nodes(el.node_ids.size());
        for (size_t j = 0; j < nodes.size(); ++j) nodes[j] = mesh.node(el.node_ids[j]);
        for (int sp = 0; sp < S; ++sp) {
            Real a, b, c, d;
            ElementGeometry::jacobian(nodes, el.type, el.order, ops.solutionXi()[sp], ops.solutionEta()[sp], a, b, c, d);
            const Real det = a * d - b * c, inv = 1.0 / det;
            J.push_back(det);
            Jinv.push_back(d * inv); Jinv.push_back(-b * inv); Jinv.push_back(-c * inv); Jinv.push_back(a * inv);
        }
        area[e] = mesh.elementArea(e);
    }
    std::vector<double> nrm(std::max<Index>(2 * F, 1));
    std::vector<int> le(std::max<Index>(F, 1)), ll(le.size()), re(le.size()), rl(le.size());
    for (Index f = 0; f < F; ++f) {
        const Vec2 n = mesh.faceNormal(f);
        nrm[2 * f] = n.x; nrm[2 * f + 1] = n.y;
        const Face& fc = mesh.face(f);
        le[f] = fc.left_elem; ll[f] = fc.left_face; re[f] = fc.right_elem; rl[f] = fc.right_face;
    }
    uploadBoundaryTables(true);  // fills the caches from the mesh (no handle yet)
    zj_mesh_desc md{};
    md.poly_order = params.poly_order; md.n_elem = (int)E; md.n_faces = (int)F;
    md.diff_xi = dx.data(); md.diff_eta = de.data(); md.interp = in.data(); md.corr = co.data();
    md.J = J.data(); md.Jinv = Jinv.data(); md.normals = nrm.data();
    md.face_left_elem = le.data(); md.face_left_local = ll.data();
    md.face_right_elem = re.data(); md.face_right_local = rl.data();
    md.bc_type = bc_type_cache_.data(); md.bc_face_map = bc_face_map_cache_.data();
    md.bc_data = bc_data_cache_.empty() ? nullptr : bc_data_cache_.data();
    md.n_bc_data = (int)(bc_data_cache_.size() / 8);
    md.elem_area = area.data();
    // half lengths of the straight element edges (textbook correction metrics)
    std::vector<double> ehl((size_t)std::max<Index>(E, 1) * 4);
    for (Index e = 0; e < E; ++e) {
        const Element& el = mesh.element(e);
        for (int k = 0; k < 4; ++k) {
            const Vec2 a = mesh.node(el.node_ids[k]), b = mesh.node(el.node_ids[(k + 1) % 4]);
            ehl[(size_t)e * 4 + k] = 0.5 * std::sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y));
        }
    }
    md.edge_half_len = ehl.data();
    // element centroids let the library tile elements along a Hilbert curve
    std::vector<double> cen((size_t)std::max<Index>(E, 1) * 2);
    for (Index e = 0; e < E; ++e) {
        const Vec2 c = mesh.elementCentroid(e);
        cen[(size_t)e * 2] = c.x; cen[(size_t)e * 2 + 1] = c.y;
    }
    md.elem_centroid = cen.data();
    zj_params zp{};
    zp.gamma = params.gamma; zp.cfl = params.CFL;
    zp.precision = params.precision == Precision::FP32Bulk ? ZJ_PREC_FP32BULK : ZJ_PREC_FP64;
    zp.compat = params.compat == Compat::Textbook ? ZJ_COMPAT_TEXTBOOK : ZJ_COMPAT_REFERENCE;
    zp.positivity = params.positivity ? 1 : 0; zp.device = params.device;
    zp.riemann = params.riemann == RiemannSolver::Roe ? ZJ_RIEMANN_ROE
                 : params.riemann == RiemannSolver::HLLC ? ZJ_RIEMANN_HLLC : ZJ_RIEMANN_RUSANOV;
    zp.corr = static_cast<int>(params.flux_type);  // DG, SD, HU, GA share ZJ_CORR_* values
    // SimParams::viscous: the reference's executed arithmetic discards the
    // viscous terms (fr_solver.cpp:645-646); the textbook scheme runs them (BR2)
    zp.viscous = (params.compat == Compat::Textbook && params.viscous) ? 1 : 0;
    zp.Re = params.Re; zp.Pr = params.Pr;
    zj_handle h = nullptr;
    const int rc = zj_create(&h, &md, &zp);
    if (rc != ZJ_OK) throw std::runtime_error(std::string("FRSolver::initialize: ") + zj_last_error(nullptr));
    handle_ = h;
    time_ = 0.0;
    iter_ = 0;
}

// BC tables as fr_solver.cpp:432-482 builds them from the mesh's faces.
void FRSolver::uploadBoundaryTables(bool force) {
    const Index F = mesh_->numFaces();
    std::vector<int> bct(F, static_cast<int>(BCType::Interior)), bmap(F, -1);
    std::vector<double> data;
    const Real rho_inf = params_.rho_inf;
    const Real c_inf = std::sqrt(params_.gamma * params_.p_inf / rho_inf);
    const Real vel = params_.Mach_inf * c_inf;
    const Real u_inf = vel * std::cos(params_.AoA * M_PI / 180.0), v_inf = vel * std::sin(params_.AoA * M_PI / 180.0);
    int count = 0;
    for (Index f = 0; f < F; ++f) {
        const Face& fc = mesh_->face(f);
        if (!fc.is_boundary) continue;
        bct[f] = static_cast<int>(fc.bc_type);
        bmap[f] = count++;
        Real p_static = params_.p_inf;
        if (fc.bc_tag > 0 && mesh_->hasBCInfo(fc.bc_tag)) {
            const BCInfo& info = mesh_->getBCInfo(fc.bc_tag);
            if (info.p_static > 0) p_static = info.p_static;
        }
        double e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (fc.bc_type == BCType::FarField) { e[0] = rho_inf; e[1] = u_inf; e[2] = v_inf; e[3] = params_.p_inf; }
        else if (fc.bc_type == BCType::Outflow) e[0] = p_static;
        data.insert(data.end(), e, e + 8);
    }
    if (!force && bct == bc_type_cache_ && bmap == bc_face_map_cache_ && data == bc_data_cache_) return;
    bc_type_cache_ = std::move(bct);
    bc_face_map_cache_ = std::move(bmap);
    bc_data_cache_ = std::move(data);
    if (handle_)
        check(zj_set_boundary(handle_, bc_type_cache_.data(), bc_face_map_cache_.data(),
                              bc_data_cache_.empty() ? nullptr : bc_data_cache_.data(),
                              (int)(bc_data_cache_.size() / 8)), "advance (boundary tables)");
}

void FRSolver::setInitialCondition(const std::function<State(Vec2)>& ic) {
    for (Index e = 0; e < mesh_->numElements(); ++e) {
        const Element& el = mesh_->element(e);
        const FROperators& ops = operators_.at(el.type);
        std::vector<Vec2> nodes(el.node_ids.size());
        for (size_t j = 0; j < nodes.size(); ++j) nodes[j] = mesh_->node(el.node_ids[j]);
        for (int sp = 0; sp < ops.numSolutionPoints(); ++sp)
            solution_[e].sol_pts[sp] =
                ic(ElementGeometry::refToPhys(nodes, el.type, el.order, ops.solutionXi()[sp], ops.solutionEta()[sp]));
    }
    copyToDevice();
    time_ = 0.0;
    iter_ = 0;
}

void FRSolver::setBoundaryCondition(int bc_tag, std::shared_ptr<BoundaryCondition> bc) { bc_map_[bc_tag] = bc; }

void FRSolver::advance(Real dt) {
    uploadBoundaryTables(false);
    check(zj_rk3_step(handle_, dt, nullptr), "advance");
    check(zj_synchronize(handle_), "advance");
    time_ += dt;
    iter_++;
}

Real FRSolver::computeTimeStep() const {
    double dt = 0.0;
    check(zj_dt_cfl(handle_, params_.CFL, &dt), "computeTimeStep");
    return dt;
}

Real FRSolver::computeResidual() const {
    double l2 = 0.0;
    check(zj_residual_l2(handle_, &l2), "computeResidual");
    return l2;
}

void FRSolver::copyToHost() {
    const Index E = mesh_->numElements();
    std::vector<double> U(E * n_sp_quad_ * N_VARS);
    check(zj_get_state(handle_, U.data()), "copyToHost");
    size_t k = 0;
    for (Index e = 0; e < E; ++e)
        for (int sp = 0; sp < n_sp_quad_; ++sp)
            for (int v = 0; v < N_VARS; ++v) solution_[e].sol_pts[sp][v] = U[k++];
}

void FRSolver::copyToDevice() {
    const Index E = mesh_->numElements();
    std::vector<double> U;
    U.reserve(E * n_sp_quad_ * N_VARS);
    for (Index e = 0; e < E; ++e)
        for (int sp = 0; sp < n_sp_quad_; ++sp)
            for (int v = 0; v < N_VARS; ++v) U.push_back(solution_[e].sol_pts[sp][v]);
    check(zj_set_state(handle_, U.data()), "copyToDevice");
}

void FRSolver::setSolution(const std::vector<ElementSolution>& solution, Real time, int iteration) {
    const Index E = mesh_->numElements();
    if ((Index)solution.size() != E) throw std::runtime_error("setSolution: element count differs from the mesh");
    for (Index e = 0; e < E; ++e)
        if ((int)solution[e].sol_pts.size() != n_sp_quad_)
            throw std::runtime_error("setSolution: points per element differ from the solver's");
    for (Index e = 0; e < E; ++e) solution_[e].sol_pts = solution[e].sol_pts;
    copyToDevice();
    time_ = time;
    iter_ = iteration;
}

std::vector<Real> FRSolver::residualField() {
    std::vector<Real> R(mesh_->numElements() * n_sp_quad_ * N_VARS);
    check(zj_residual(handle_, R.data(), nullptr), "residualField");
    return R;
}

void FRSolver::computeDerivedQuantities(std::vector<std::vector<Real>>& pressure,
                                        std::vector<std::vector<Real>>& mach) const {
    const Index E = mesh_->numElements();
    pressure.resize(E); mach.resize(E);
    IdealGas gas(params_.gamma);
    for (Index e = 0; e < E; ++e) {
        const size_t n = solution_[e].sol_pts.size();
        pressure[e].resize(n); mach[e].resize(n);
        for (size_t sp = 0; sp < n; ++sp) {
            pressure[e][sp] = gas.pressure(solution_[e].sol_pts[sp]);
            mach[e][sp] = gas.machNumber(solution_[e].sol_pts[sp]);
        }
    }
}

}  // namespace zhijian
