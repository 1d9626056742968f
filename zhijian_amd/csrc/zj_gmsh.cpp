// zj_gmsh.cpp -- native Gmsh MSH reader (ASCII 2.2 and 4.1) for the mesh
// ingestion in front of the solver, with the reference's conventions
// (src/mesh/gmsh_reader.cpp:161-760):
//   - nodes numbered 0.. in file order; element node lists remapped;
//   - kept 2D elements: 3-node triangle, 6-node triangle, 4-node quad,
//     9-node quad, 25-node quad, and the 8-node serendipity quad, which gets a
//     centre node 0.5*(mid-sides) - 0.25*(corners) appended to the node list
//     (exact for bilinear geometry) and becomes a 9-node quad;
//   - line elements of any order are boundary edges, keyed by their two end
//     vertices, carrying the physical tag (v2: the first element tag; v4: the
//     first physical tag of the curve entity);
//   - boundary groups get the $PhysicalNames name (else "BC_<tag>") and a BC
//     type inferred from it (gmsh_reader.cpp:117-158);
//   - $Periodic node correspondences (slave node -> master node) are kept as
//     node-index pairs.  The reference skips this section; the pairs feed
//     zj_pair_periodic_nodes, which joins the faces they make periodic.
// Exposed through a small C ABI (include/zhijian_amd.h) and, in C++, through
// zhijian::GmshReader (include/zhijian/mesh.hpp).
#include "zhijian_amd.h"

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstring>
#include <exception>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace zj {
namespace gmsh {

struct Elem {
    int kind;   // ElementType: 0 triangle, 1 quadrilateral (types.hpp:41-44)
    int order;  // geometric order
    std::vector<int64_t> nodes;
};

struct Group {
    std::string name;
    int bc_type;
};

struct Parsed {
    std::vector<double> xy;  // [n][2]
    std::vector<Elem> elems;
    std::map<std::pair<int64_t, int64_t>, int> edge_tag;  // sorted vertex pair -> physical tag
    std::map<int, Group> groups;                           // physical tag -> group
    std::vector<std::pair<int64_t, int64_t>> periodic;     // (slave, master) node indices
    std::string err;
};

// Nodes per Gmsh element type (types listed by the Gmsh format; -1 unknown).
static int nodes_of(int t) {
    switch (t) {
    case 1: return 2;   case 2: return 3;   case 3: return 4;   case 4: return 4;
    case 5: return 8;   case 6: return 6;   case 7: return 5;   case 8: return 3;
    case 9: return 6;   case 10: return 9;  case 11: return 10; case 12: return 27;
    case 13: return 18; case 14: return 14; case 15: return 1;  case 16: return 8;
    case 20: return 9;  case 21: return 10; case 26: return 4;  case 27: return 5;
    case 28: return 6;  case 36: return 16; case 37: return 25;
    default: return -1;
    }
}

static bool is_line(int t) { return t == 1 || t == 8 || t == 26 || t == 27 || t == 28; }

static bool has(const std::string& s, const char* w) { return s.find(w) != std::string::npos; }

// BC type from a physical-group name (gmsh_reader.cpp:117-158), as ZJ_BC_*.
int infer_bc(const std::string& name) {
    std::string l = name;
    for (char& c : l) c = (char)std::tolower((unsigned char)c);
    if (has(l, "wall")) return (has(l, "slip") || has(l, "inviscid")) ? ZJ_BC_SLIPWALL : ZJ_BC_WALL;
    if (has(l, "airfoil") || has(l, "wing") || has(l, "body") || has(l, "surface")) return ZJ_BC_WALL;
    if (has(l, "farfield") || has(l, "far_field") || has(l, "far-field") || has(l, "freestream"))
        return ZJ_BC_FARFIELD;
    if (has(l, "inflow") || has(l, "inlet")) return ZJ_BC_INFLOW;
    if (has(l, "outflow") || has(l, "outlet") || has(l, "exit") || l == "out") return ZJ_BC_OUTFLOW;
    if (has(l, "symmetry") || has(l, "sym")) return ZJ_BC_SYMMETRY;
    if (has(l, "periodic")) return ZJ_BC_PERIODIC;
    return ZJ_BC_WALL;
}

namespace {

class Reader {
public:
    explicit Reader(std::istream& in) : in_(in) {}
    bool line(std::string& s) { return (bool)std::getline(in_, s); }
    // The next line of `section`; at end of file, false with the error set.
    bool need(std::string& s, const char* section, std::string& err) {
        if (line(s)) return true;
        err = std::string("unexpected end of file in ") + section;
        return false;
    }
    // skip to the line containing `marker`
    void skip_to(const std::string& marker) {
        std::string s;
        while (line(s))
            if (has(s, marker.c_str())) return;
    }

private:
    std::istream& in_;
};

struct State {
    Parsed& out;
    std::map<int64_t, int64_t> node_index;     // Gmsh node tag -> 0-based index
    std::map<int, int> curve_phys;             // v4 curve entity -> physical tag
    std::map<std::pair<int, int>, std::string> names;
    std::vector<std::pair<std::vector<int64_t>, int>> bedges;  // (vertex nodes, physical tag)
    std::vector<std::pair<int64_t, int64_t>> periodic_tags;    // $Periodic (slave, master) node tags
};

bool add_element(State& st, int type, int phys, std::istringstream& ls) {
    const int nn = nodes_of(type);
    std::vector<int64_t> ids((size_t)nn);
    for (int j = 0; j < nn; ++j) {
        int64_t tag = 0;
        if (!(ls >> tag)) {
            st.out.err = "element line with fewer than " + std::to_string(nn) + " node tags";
            return false;
        }
        auto it = st.node_index.find(tag);
        if (it == st.node_index.end()) {
            st.out.err = "element references unknown node tag " + std::to_string(tag);
            return false;
        }
        ids[(size_t)j] = it->second;
    }
    if (is_line(type)) {
        st.bedges.emplace_back(std::move(ids), phys);
        return true;
    }
    Elem e;
    switch (type) {
    case 2: e = Elem{0, 1, ids}; break;
    case 9: e = Elem{0, 2, ids}; break;
    case 3: e = Elem{1, 1, ids}; break;
    case 10: e = Elem{1, 2, ids}; break;
    case 37: e = Elem{1, 4, ids}; break;
    case 16: {  // serendipity: synthesise the centre node, keep as a 9-node quad
        const std::vector<double>& xy = st.out.xy;
        double c[2];
        for (int d = 0; d < 2; ++d) {
            const double mid = xy[2 * ids[4] + d] + xy[2 * ids[5] + d] + xy[2 * ids[6] + d] + xy[2 * ids[7] + d];
            const double cor = xy[2 * ids[0] + d] + xy[2 * ids[1] + d] + xy[2 * ids[2] + d] + xy[2 * ids[3] + d];
            c[d] = 0.5 * mid - 0.25 * cor;
        }
        const int64_t idx = (int64_t)(st.out.xy.size() / 2);
        st.out.xy.push_back(c[0]);
        st.out.xy.push_back(c[1]);
        ids.push_back(idx);
        e = Elem{1, 2, ids};
        break;
    }
    default: return true;  // points and 3D elements are not part of a 2D mesh
    }
    st.out.elems.push_back(std::move(e));
    return true;
}

// Node and element sections.  Counts in section and block headers are
// checked before use, storage grows with the lines actually read (a corrupt
// count cannot size an allocation), and a section that ends early is an error.
bool read_nodes(Reader& r, State& st, bool v4) {
    std::string s, &err = st.out.err;
    if (!r.need(s, "$Nodes", err)) return false;
    std::istringstream h(s);
    std::vector<double>& xy = st.out.xy;
    xy.clear();
    if (!v4) {
        int64_t n = -1;
        if (!(h >> n) || n < 0) { err = "malformed $Nodes header"; return false; }
        for (int64_t i = 0; i < n; ++i) {
            if (!r.need(s, "$Nodes", err)) return false;
            std::istringstream ls(s);
            int64_t tag;
            double x, y, z;
            if (!(ls >> tag >> x >> y >> z)) { err = "malformed node line: " + s; return false; }
            st.node_index[tag] = i;
            xy.push_back(x);
            xy.push_back(y);
        }
    } else {
        int64_t blocks = -1, n = -1, tmin, tmax;
        if (!(h >> blocks >> n >> tmin >> tmax) || blocks < 0 || n < 0) {
            err = "malformed $Nodes header";
            return false;
        }
        int64_t base = 0;
        for (int64_t b = 0; b < blocks; ++b) {
            if (!r.need(s, "$Nodes", err)) return false;
            std::istringstream bl(s);
            int dim, ent, param;
            int64_t cnt = -1;
            if (!(bl >> dim >> ent >> param >> cnt) || cnt < 0 || cnt > n - base) {
                err = "malformed $Nodes entity block: " + s;
                return false;
            }
            for (int64_t i = 0; i < cnt; ++i) {
                if (!r.need(s, "$Nodes", err)) return false;
                std::istringstream ls(s);
                int64_t tag;
                if (!(ls >> tag)) { err = "malformed node tag: " + s; return false; }
                st.node_index[tag] = base + i;
            }
            for (int64_t i = 0; i < cnt; ++i) {
                if (!r.need(s, "$Nodes", err)) return false;
                std::istringstream ls(s);
                double x, y, z;
                if (!(ls >> x >> y >> z)) { err = "malformed node coordinates: " + s; return false; }
                xy.push_back(x);
                xy.push_back(y);
            }
            base += cnt;
        }
    }
    r.skip_to("$EndNodes");
    return true;
}

bool read_elements(Reader& r, State& st, bool v4) {
    std::string s, &err = st.out.err;
    if (!r.need(s, "$Elements", err)) return false;
    std::istringstream h(s);
    if (!v4) {
        int64_t n = -1;
        if (!(h >> n) || n < 0) { err = "malformed $Elements header"; return false; }
        for (int64_t i = 0; i < n; ++i) {
            if (!r.need(s, "$Elements", err)) return false;
            std::istringstream ls(s);
            int id, type, ntags;
            if (!(ls >> id >> type >> ntags) || ntags < 0) { err = "malformed element line: " + s; return false; }
            int phys = 0;
            for (int t = 0; t < ntags; ++t) {
                int tag;
                if (!(ls >> tag)) { err = "malformed element line: " + s; return false; }
                if (t == 0) phys = tag;
            }
            if (nodes_of(type) < 0) continue;
            if (!add_element(st, type, phys, ls)) return false;
        }
    } else {
        int64_t blocks = -1, n, tmin, tmax;
        if (!(h >> blocks >> n >> tmin >> tmax) || blocks < 0) { err = "malformed $Elements header"; return false; }
        for (int64_t b = 0; b < blocks; ++b) {
            if (!r.need(s, "$Elements", err)) return false;
            std::istringstream bl(s);
            int dim, ent, type;
            int64_t cnt = -1;
            if (!(bl >> dim >> ent >> type >> cnt) || cnt < 0) {
                err = "malformed $Elements entity block: " + s;
                return false;
            }
            int phys = 0;
            if (dim == 1) {
                auto it = st.curve_phys.find(ent);
                if (it != st.curve_phys.end()) phys = it->second;
            }
            for (int64_t e = 0; e < cnt; ++e) {
                if (!r.need(s, "$Elements", err)) return false;
                if (nodes_of(type) < 0 || type == 15) continue;
                std::istringstream ls(s);
                int64_t id;
                if (!(ls >> id)) { err = "malformed element line: " + s; return false; }
                if (!add_element(st, type, phys, ls)) return false;
            }
        }
    }
    r.skip_to("$EndElements");
    return true;
}

// $Periodic.  MSH 2.2: per link "dim slave master", an optional "Affine ..."
// line, the node count, then "slave master" node-tag lines.  MSH 4.1: per link
// "dim tag master_tag", a line "n_affine values...", the node count, then the
// node-tag lines.  Only the node correspondences are kept.
bool read_periodic(Reader& r, State& st, bool v4) {
    std::string s, &err = st.out.err;
    if (!r.need(s, "$Periodic", err)) return false;
    std::istringstream h(s);
    int64_t links = -1;
    if (!(h >> links) || links < 0) { err = "malformed $Periodic header"; return false; }
    for (int64_t k = 0; k < links; ++k) {
        if (!r.need(s, "$Periodic", err)) return false;
        std::istringstream ls(s);
        int dim, tag, master;
        if (!(ls >> dim >> tag >> master)) { err = "malformed periodic link: " + s; return false; }
        if (!r.need(s, "$Periodic", err)) return false;
        if (v4) {
            std::istringstream as(s);
            int64_t n_affine = -1;
            if (!(as >> n_affine) || n_affine < 0) { err = "malformed periodic affine line: " + s; return false; }
            if (!r.need(s, "$Periodic", err)) return false;
        } else if (s.compare(0, 6, "Affine") == 0) {
            if (!r.need(s, "$Periodic", err)) return false;
        }
        std::istringstream cs(s);
        int64_t n = -1;
        if (!(cs >> n) || n < 0) { err = "malformed periodic node count: " + s; return false; }
        for (int64_t i = 0; i < n; ++i) {
            if (!r.need(s, "$Periodic", err)) return false;
            std::istringstream ns(s);
            int64_t a, b;
            if (!(ns >> a >> b)) { err = "malformed periodic node pair: " + s; return false; }
            st.periodic_tags.emplace_back(a, b);
        }
    }
    r.skip_to("$EndPeriodic");
    return true;
}

}  // namespace

bool parse(const std::string& path, Parsed& out) {
    std::ifstream f(path);
    if (!f) {
        out.err = "cannot open " + path;
        return false;
    }
    Reader r(f);
    State st{out, {}, {}, {}, {}};
    std::string s;
    double version = 0.0;
    int ftype = -1;
    bool have_nodes = false, have_elems = false;
    while (r.line(s)) {
        if (s.empty() || s[0] != '$') continue;
        if (has(s, "$MeshFormat")) {
            if (!r.need(s, "$MeshFormat", out.err)) return false;
            std::istringstream ls(s);
            int dsize;
            if (!(ls >> version >> ftype >> dsize) || version <= 0.0) {
                out.err = "malformed $MeshFormat line: " + s;
                return false;
            }
            if (ftype != 0) {
                out.err = "only ASCII Gmsh files are supported";
                return false;
            }
            r.skip_to("$EndMeshFormat");
        } else if (has(s, "$PhysicalNames")) {
            if (!r.need(s, "$PhysicalNames", out.err)) return false;
            std::istringstream hs(s);
            int n = -1;
            if (!(hs >> n) || n < 0) { out.err = "malformed $PhysicalNames header"; return false; }
            for (int i = 0; i < n; ++i) {
                if (!r.need(s, "$PhysicalNames", out.err)) return false;
                std::istringstream ls(s);
                int dim, tag;
                if (!(ls >> dim >> tag)) { out.err = "malformed physical name: " + s; return false; }
                std::string name;
                std::getline(ls, name);
                const size_t a = name.find_first_not_of(" \t");
                name = a == std::string::npos ? std::string() : name.substr(a);
                if (!name.empty() && name.front() == '"') name.erase(0, 1);
                if (!name.empty() && name.back() == '"') name.pop_back();
                st.names[{dim, tag}] = name;
            }
            r.skip_to("$EndPhysicalNames");
        } else if (has(s, "$Entities")) {
            if (!r.need(s, "$Entities", out.err)) return false;
            std::istringstream hs(s);
            int np = -1, nc = -1, ns, nv;
            if (!(hs >> np >> nc >> ns >> nv) || np < 0 || nc < 0) {
                out.err = "malformed $Entities header";
                return false;
            }
            for (int i = 0; i < np; ++i)
                if (!r.need(s, "$Entities", out.err)) return false;
            for (int i = 0; i < nc; ++i) {
                if (!r.need(s, "$Entities", out.err)) return false;
                std::istringstream ls(s);
                int tag, nphys = 0;
                double bb[6];
                if (!(ls >> tag >> bb[0] >> bb[1] >> bb[2] >> bb[3] >> bb[4] >> bb[5] >> nphys)) {
                    out.err = "malformed curve entity: " + s;
                    return false;
                }
                int ph;
                if (nphys > 0 && (ls >> ph)) st.curve_phys[tag] = ph;
            }
            r.skip_to("$EndEntities");
        } else if (has(s, "$Nodes")) {
            if (version == 0.0) { out.err = "$Nodes before $MeshFormat"; return false; }
            if (have_nodes) { out.err = "more than one $Nodes section"; return false; }
            if (!read_nodes(r, st, version >= 4.0)) return false;
            have_nodes = true;
        } else if (has(s, "$Elements")) {
            if (!have_nodes) { out.err = "$Elements before $Nodes"; return false; }
            if (have_elems) { out.err = "more than one $Elements section"; return false; }
            if (!read_elements(r, st, version >= 4.0)) return false;
            have_elems = true;
        } else if (has(s, "$Periodic")) {
            if (version == 0.0) { out.err = "$Periodic before $MeshFormat"; return false; }
            if (!read_periodic(r, st, version >= 4.0)) return false;
        } else if (!has(s, "$End")) {
            r.skip_to("$End" + s.substr(1));
        }
    }
    if (version == 0.0) { out.err = "no $MeshFormat section"; return false; }
    if (!have_nodes) { out.err = "no $Nodes section"; return false; }
    if (!have_elems) { out.err = "no $Elements section"; return false; }
    if (out.elems.empty()) { out.err = "no 2D elements (triangles or quadrilaterals) in the file"; return false; }
    for (const auto& pt : st.periodic_tags) {
        auto a = st.node_index.find(pt.first), b = st.node_index.find(pt.second);
        if (a == st.node_index.end() || b == st.node_index.end()) {
            out.err = "$Periodic names unknown node tag " +
                      std::to_string(a == st.node_index.end() ? pt.first : pt.second);
            return false;
        }
        out.periodic.emplace_back(a->second, b->second);
    }
    std::set<int> tags;
    for (const auto& be : st.bedges)
        if (be.second > 0) tags.insert(be.second);
    for (int t : tags) {
        auto it = st.names.find({1, t});
        const std::string name = it != st.names.end() ? it->second : "BC_" + std::to_string(t);
        out.groups[t] = Group{name, infer_bc(name)};
    }
    for (const auto& be : st.bedges) {
        if (be.first.size() < 2 || be.second <= 0) continue;
        const int64_t a = std::min(be.first[0], be.first[1]), b = std::max(be.first[0], be.first[1]);
        out.edge_tag[{a, b}] = be.second;
    }
    return true;
}

}  // namespace gmsh
}  // namespace zj

// ------------------------------------------------------------------ C ABI
struct zj_gmsh {
    zj::gmsh::Parsed p;
    std::vector<int> group_tags;
};

static thread_local std::string g_gmsh_error;

extern "C" int zj_gmsh_read(const char* path, zj_gmsh** out) {
    if (!path || !out) return ZJ_ERR_ARG;
    *out = nullptr;
    zj_gmsh* g = new (std::nothrow) zj_gmsh();
    if (!g) return ZJ_ERR_NOMEM;
    bool ok;
    try {
        ok = zj::gmsh::parse(path, g->p);
    } catch (const std::exception& e) {   // std::bad_alloc on a mesh too large for host memory
        g->p.err = std::string("reading ") + path + ": " + e.what();
        ok = false;
    }
    if (!ok) {
        g_gmsh_error = g->p.err;
        delete g;
        return ZJ_ERR_ARG;
    }
    for (const auto& kv : g->p.groups) g->group_tags.push_back(kv.first);
    *out = g;
    return ZJ_OK;
}

extern "C" const char* zj_gmsh_error(void) { return g_gmsh_error.c_str(); }

extern "C" int zj_gmsh_free(zj_gmsh* g) {
    delete g;
    return ZJ_OK;
}

extern "C" int zj_gmsh_counts(zj_gmsh* g, int64_t* n_nodes, int* n_elem, int* max_nodes, int* n_groups) {
    if (!g) return ZJ_ERR_ARG;
    size_t mx = 0;
    for (const auto& e : g->p.elems) mx = std::max(mx, e.nodes.size());
    if (n_nodes) *n_nodes = (int64_t)(g->p.xy.size() / 2);
    if (n_elem) *n_elem = (int)g->p.elems.size();
    if (max_nodes) *max_nodes = (int)mx;
    if (n_groups) *n_groups = (int)g->group_tags.size();
    return ZJ_OK;
}

extern "C" int zj_gmsh_nodes(zj_gmsh* g, double* xy) {
    if (!g || !xy) return ZJ_ERR_ARG;
    std::memcpy(xy, g->p.xy.data(), g->p.xy.size() * sizeof(double));
    return ZJ_OK;
}

extern "C" int zj_gmsh_elements(zj_gmsh* g, int* kind, int* order, int* n_nodes, int64_t* nodes, int stride) {
    if (!g || !kind || !order || !n_nodes || !nodes) return ZJ_ERR_ARG;
    for (size_t e = 0; e < g->p.elems.size(); ++e) {
        const auto& el = g->p.elems[e];
        if ((int)el.nodes.size() > stride) return ZJ_ERR_ARG;
        kind[e] = el.kind;
        order[e] = el.order;
        n_nodes[e] = (int)el.nodes.size();
        for (int k = 0; k < stride; ++k) nodes[e * stride + k] = k < (int)el.nodes.size() ? el.nodes[k] : -1;
    }
    return ZJ_OK;
}

extern "C" int zj_gmsh_edge_tag(zj_gmsh* g, int64_t a, int64_t b, int* tag) {
    if (!g || !tag) return ZJ_ERR_ARG;
    auto it = g->p.edge_tag.find({std::min(a, b), std::max(a, b)});
    *tag = it == g->p.edge_tag.end() ? 0 : it->second;
    return ZJ_OK;
}

extern "C" int zj_gmsh_periodic(zj_gmsh* g, int64_t* n_pairs, int64_t* slave, int64_t* master) {
    if (!g || !n_pairs) return ZJ_ERR_ARG;
    *n_pairs = (int64_t)g->p.periodic.size();
    for (size_t i = 0; i < g->p.periodic.size(); ++i) {
        if (slave) slave[i] = g->p.periodic[i].first;
        if (master) master[i] = g->p.periodic[i].second;
    }
    return ZJ_OK;
}

extern "C" int zj_gmsh_group(zj_gmsh* g, int i, int* tag, int* bc_type, char* name, int name_cap) {
    if (!g || i < 0 || i >= (int)g->group_tags.size()) return ZJ_ERR_ARG;
    const int t = g->group_tags[(size_t)i];
    const auto& grp = g->p.groups.at(t);
    if (tag) *tag = t;
    if (bc_type) *bc_type = grp.bc_type;
    if (name && name_cap > 0) {
        std::strncpy(name, grp.name.c_str(), (size_t)name_cap - 1);
        name[name_cap - 1] = '\0';
    }
    return ZJ_OK;
}
