// zhijian/basis.hpp -- FR basis, operators and element geometry.
//
// API of the reference's include/basis/basis.hpp.  The quad tables are the
// reference's (basis.cpp:464-613): Gauss solution and flux points, tensor
// differentiation matrices, per-edge interpolation with edges 2 and 3
// reversed, and the Radau (DG) correction derivative for every FluxType, as
// the reference computes it.  Triangle tables are zero, as in the reference
// (basis.cpp:615-669); the device path rejects triangles.
#pragma once

#include "zhijian/types.hpp"

#include <vector>

namespace zhijian {

class GaussLegendre {
public:
    static void getQuadrature(int n, std::vector<Real>& points, std::vector<Real>& weights);
    static void getGLL(int n, std::vector<Real>& points, std::vector<Real>& weights);
};

class Lagrange1D {
public:
    static Real eval(const std::vector<Real>& nodes, int j, Real x);
    static Real evalDeriv(const std::vector<Real>& nodes, int j, Real x);
    static void evalAll(const std::vector<Real>& nodes, Real x, std::vector<Real>& values);
    static void evalAllDeriv(const std::vector<Real>& nodes, Real x, std::vector<Real>& derivs);
    static void buildVandermonde(const std::vector<Real>& nodes, std::vector<std::vector<Real>>& V);
    static void buildDiffMatrix(const std::vector<Real>& nodes, std::vector<std::vector<Real>>& D);
};

class CorrectionFunction {
public:
    static Real evalLeft(FluxType type, int order, Real xi);
    static Real evalRight(FluxType type, int order, Real xi);
    static Real evalLeftDeriv(FluxType type, int order, Real xi);
    static Real evalRightDeriv(FluxType type, int order, Real xi);
    static Real getC(FluxType type, int order);
};

class SolutionPoints {
public:
    static void get1D(int order, std::vector<Real>& points);
    static void getTriangle(int order, std::vector<Real>& xi, std::vector<Real>& eta);
    static void getQuad(int order, std::vector<Real>& xi, std::vector<Real>& eta);
    static int numPoints1D(int order) { return order + 1; }
    static int numPointsTri(int order) { return (order + 1) * (order + 2) / 2; }
    static int numPointsQuad(int order) { return (order + 1) * (order + 1); }
};

class FluxPoints {
public:
    static void get1D(int order, std::vector<Real>& points);
    static int numPointsPerEdge(int order) { return order + 1; }
};

class FROperators {
public:
    FROperators() = default;
    void init(ElementType elem_type, int poly_order, FluxType flux_type);

    const std::vector<std::vector<Real>>& interpToFlux(int edge) const { return interp_to_flux_[edge]; }
    const std::vector<Real>& extrapToFluxPoint(int edge, int fp) const { return interp_to_flux_[edge][fp]; }
    const std::vector<std::vector<Real>>& diffXi() const { return diff_xi_; }
    const std::vector<std::vector<Real>>& diffEta() const { return diff_eta_; }
    const std::vector<std::vector<Real>>& correctionDeriv() const { return correction_deriv_; }
    const std::vector<Real>& quadWeights() const { return quad_weights_; }
    const std::vector<Real>& solutionXi() const { return sol_xi_; }
    const std::vector<Real>& solutionEta() const { return sol_eta_; }
    const std::vector<std::vector<Real>>& fluxPointLoc() const { return flux_pt_loc_; }
    int numSolutionPoints() const { return num_sol_pts_; }
    int numFluxPointsPerEdge() const { return num_flux_pts_per_edge_; }
    int numEdges() const { return num_edges_; }
    ElementType elementType() const { return elem_type_; }
    int polyOrder() const { return poly_order_; }

private:
    ElementType elem_type_ = ElementType::Quadrilateral;
    int poly_order_ = 0;
    FluxType flux_type_ = FluxType::DG;
    int num_sol_pts_ = 0;
    int num_flux_pts_per_edge_ = 0;
    int num_edges_ = 0;
    std::vector<Real> sol_xi_, sol_eta_, quad_weights_;
    std::vector<std::vector<Real>> flux_pt_loc_;
    std::vector<std::vector<std::vector<Real>>> interp_to_flux_;
    std::vector<std::vector<Real>> diff_xi_, diff_eta_;
    std::vector<std::vector<Real>> correction_deriv_;
};

class ElementGeometry {
public:
    static void jacobian(const std::vector<Vec2>& nodes, ElementType type, int order, Real xi, Real eta,
                         Real& dxdxi, Real& dxdeta, Real& dydxi, Real& dydeta);
    static Real jacobianDet(const std::vector<Vec2>& nodes, ElementType type, int order, Real xi, Real eta);
    static void jacobianInverse(Real dxdxi, Real dxdeta, Real dydxi, Real dydeta,
                                Real& dxidx, Real& dxidy, Real& detadx, Real& detady);
    static Vec2 refToPhys(const std::vector<Vec2>& nodes, ElementType type, int order, Real xi, Real eta);
    static void shapeFunc(ElementType type, int order, Real xi, Real eta, std::vector<Real>& N);
    static void shapeFuncDeriv(ElementType type, int order, Real xi, Real eta,
                               std::vector<Real>& dNdxi, std::vector<Real>& dNdeta);
    static Vec2 refFaceNormal(ElementType type, int face_id);
    static void faceParamToRef(ElementType type, int face_id, Real s, Real& xi, Real& eta);
};

}  // namespace zhijian
