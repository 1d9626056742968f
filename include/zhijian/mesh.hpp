// zhijian/mesh.hpp -- unstructured 2D mesh with the face list the solver uses.
//
// API of the reference's include/mesh/mesh.hpp (Face, Element, BCInfo, Mesh).
// buildFaces() reproduces the reference's face order and left/right roles
// (mesh.cpp:18-83): faces in ascending (min vertex, max vertex) order, left =
// first element in element order.  pairPeriodic() and pairPeriodicNodes()
// have no counterpart in the reference: they join boundary faces one period
// apart, or whose end nodes a mesh file's $Periodic links make correspond.
#pragma once

#include "zhijian/types.hpp"

#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace zhijian {

struct Face {
    LocalIndex left_elem;
    LocalIndex right_elem;   // -1 on a boundary
    LocalIndex left_face;
    LocalIndex right_face;
    BCType bc_type;
    int bc_tag;
    bool is_boundary;
    int mpi_rank;
};

struct Element {
    ElementType type = ElementType::Quadrilateral;
    int order = 1;                     // geometric order (1: Q1, 2: Q2, 4: Q4)
    std::vector<Index> node_ids;
    std::vector<LocalIndex> face_ids;
    int partition = 0;
    int numVertices() const { return (type == ElementType::Triangle) ? 3 : 4; }
    int numNodes() const {
        if (type == ElementType::Triangle) return (order == 1) ? 3 : 6;
        return (order == 1) ? 4 : 9;
    }
};

struct BCInfo {
    BCType type;
    std::string name;
    int tag;
    Real p_total;
    Real T_total;
    Vec2 flow_direction;
    Real p_static;
    Real T_wall;
    State far_field_state;
};

class Mesh {
public:
    Mesh() = default;

    Index numNodes() const { return (Index)nodes_.size(); }
    Index numElements() const { return (Index)elements_.size(); }
    Index numFaces() const { return (Index)faces_.size(); }
    Index numBoundaryFaces() const;

    const Vec2& node(Index i) const { return nodes_[i]; }
    Vec2& node(Index i) { return nodes_[i]; }
    const Element& element(Index i) const { return elements_[i]; }
    Element& element(Index i) { return elements_[i]; }
    const Face& face(Index i) const { return faces_[i]; }
    Face& face(Index i) { return faces_[i]; }

    const std::vector<Vec2>& nodes() const { return nodes_; }
    const std::vector<Element>& elements() const { return elements_; }
    const std::vector<Face>& faces() const { return faces_; }
    std::vector<Vec2>& nodes() { return nodes_; }
    std::vector<Element>& elements() { return elements_; }
    std::vector<Face>& faces() { return faces_; }

    void addBCInfo(int tag, const BCInfo& info) { bc_info_[tag] = info; }
    const BCInfo& getBCInfo(int tag) const { return bc_info_.at(tag); }
    bool hasBCInfo(int tag) const { return bc_info_.count(tag) > 0; }
    const std::map<int, BCInfo>& allBCInfo() const { return bc_info_; }

    void buildFaces();
    // Periodic pairing: boundary faces whose midpoints differ by (+-Lx, 0) or
    // (0, +-Ly) within tol become interior faces.  The lower-indexed face stays
    // in place with its partner's left element/edge as right side; the partner
    // is removed and the order is otherwise kept.  A period <= 0 disables that
    // direction.  Throws std::runtime_error if a face has several partners.
    void pairPeriodic(Real Lx, Real Ly, Real tol = -1.0);
    // Periodic pairing by node correspondence (slave node, master node), e.g.
    // a Gmsh file's $Periodic section: two boundary faces become one interior
    // face when their end nodes correspond (directly or through a chain of
    // links) and one face is the other moved by a link's offset; kept,
    // removed and ordered as pairPeriodic() does.  The no-argument form uses
    // periodicNodes().  Throws std::runtime_error when a face has two
    // candidates or a link names a node out of range.
    void pairPeriodicNodes();
    void pairPeriodicNodes(const std::vector<std::pair<Index, Index>>& links);
    // Node correspondences read from the mesh file (GmshReader fills them;
    // the reference's reader skips $Periodic and has no such field).
    const std::vector<std::pair<Index, Index>>& periodicNodes() const { return periodic_nodes_; }
    std::vector<std::pair<Index, Index>>& periodicNodes() { return periodic_nodes_; }
    void computeGeometry();

    Vec2 elementCentroid(Index elem_id) const;
    Real elementArea(Index elem_id) const;
    Real faceLength(Index face_id) const;
    Vec2 faceNormal(Index face_id) const;
    Vec2 faceCentroid(Index face_id) const;
    void printStats() const;

private:
    void rebuildFaceIds();
    void pairFaces(const std::function<int(int*, int*, int*, int*, int*, int, const int64_t*)>& pair,
                   const char* what);

    std::vector<Vec2> nodes_;
    std::vector<Element> elements_;
    std::vector<Face> faces_;
    std::map<int, BCInfo> bc_info_;
    std::vector<std::pair<Index, Index>> periodic_nodes_;
    std::vector<Vec2> elem_centroids_;
    std::vector<Real> elem_areas_;
    std::vector<Real> face_lengths_;
    std::vector<Vec2> face_normals_;
    std::vector<Vec2> face_centroids_;
};

// Gmsh MSH 2.2 / 4.1 ASCII reader (src/mesh/gmsh_reader.cpp:161-760): fills
// the nodes, elements and boundary groups, then buildFaces(), the boundary
// tags of the faces, and computeGeometry().  Returns false with
// errorMessage() set on failure.
class GmshReader {
public:
    GmshReader() = default;
    bool read(const std::string& filename, Mesh& mesh);
    const std::string& errorMessage() const { return error_msg_; }

private:
    std::string error_msg_;
};

}  // namespace zhijian
