// zj_host.cpp -- host-side tables the device path consumes: FR operators,
// faces (with periodic pairing), normals, areas and Jacobians.
//
// These run once per mesh.  They follow the reference's definitions so that
// the tables agree with what the reference uploads:
//   operators   src/basis/basis.cpp:12-104 (Gauss), 169-224 (Lagrange),
//               314-379 (Radau correction derivative), 464-613 (quad tables)
//   faces       src/mesh/mesh.cpp:18-83 (std::map order of sorted vertex pairs);
//               periodic pairing (by period or by node links) has no
//               counterpart in the reference
//   normals     mesh.cpp:159-183;  areas mesh.cpp:122-146
//   geometry    src/solver/fr_solver.cpp:85-129, basis.cpp:717-863
#include "zhijian_amd.h"
#include "zj_host_math.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

using namespace zj::host;

extern "C" int zj_quad_operators(int p, double* sol_xi, double* sol_eta, double* qw,
                                 double* diff_xi, double* diff_eta, double* interp, double* corr) {
    if (p < 1 || p > 5) return ZJ_ERR_ARG;
    const int n = p + 1, S = n * n;
    double g[8], w[8];
    gauss_rule(n, g, w);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            if (sol_xi) sol_xi[j * n + i] = g[i];
            if (sol_eta) sol_eta[j * n + i] = g[j];
            if (qw) qw[j * n + i] = w[i] * w[j];
        }
    double D[6][6];
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) D[i][k] = lag_d(g, n, k, g[i]);
    if (diff_xi && diff_eta) {
        std::fill(diff_xi, diff_xi + S * S, 0.0);
        std::fill(diff_eta, diff_eta + S * S, 0.0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < n; ++k) {
                    diff_xi[(j * n + i) * S + j * n + k] = D[i][k];
                    diff_eta[(j * n + i) * S + k * n + i] = D[j][k];
                }
    }
    if (interp) {
        // edge 0: (s,-1), 1: (1,s), 2: (-s,1), 3: (-1,-s)
        for (int e = 0; e < 4; ++e)
            for (int f = 0; f < n; ++f) {
                double s = g[f];
                double xf = e == 0 ? s : e == 1 ? 1.0 : e == 2 ? -s : -1.0;
                double ef = e == 0 ? -1.0 : e == 1 ? s : e == 2 ? 1.0 : -s;
                double Lx[6], Le[6];
                for (int i = 0; i < n; ++i) { Lx[i] = lag(g, n, i, xf); Le[i] = lag(g, n, i, ef); }
                for (int j = 0; j < n; ++j)
                    for (int i = 0; i < n; ++i) interp[(e * n + f) * S + j * n + i] = Lx[i] * Le[j];
            }
    }
    if (corr) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int sp = j * n + i;
                double li = lag(g, n, i, g[i]), lj = lag(g, n, j, g[j]);
                corr[0 * S + sp] = radau_left_d(p, g[j]) * li;
                corr[2 * S + sp] = radau_right_d(p, g[j]) * li;
                corr[3 * S + sp] = radau_left_d(p, g[i]) * lj;
                corr[1 * S + sp] = radau_right_d(p, g[i]) * lj;
            }
    }
    return ZJ_OK;
}

extern "C" int zj_build_faces(int n_elem, int nv, int stride, const int64_t* en, int* n_faces,
                              int* le, int* ll, int* re, int* rl) {
    if (n_elem < 0 || nv < 3 || stride < nv || !n_faces || (n_elem > 0 && !en)) return ZJ_ERR_ARG;
    const int64_t ne = (int64_t)n_elem * nv;
    struct Rec { int64_t lo, hi; int32_t slot; };  // slot = element*nv + local edge
    std::vector<Rec> r((size_t)ne);
    for (int e = 0; e < n_elem; ++e)
        for (int k = 0; k < nv; ++k) {
            int64_t a = node_of(en, stride, e, k), b = node_of(en, stride, e, (k + 1) % nv);
            r[(size_t)e * nv + k] = {std::min(a, b), std::max(a, b), (int32_t)(e * nv + k)};
        }
    // std::map iteration order = ascending (lo, hi); ties keep insertion order.
    std::stable_sort(r.begin(), r.end(), [](const Rec& x, const Rec& y) {
        return x.lo != y.lo ? x.lo < y.lo : x.hi < y.hi;
    });
    int nf = 0;
    for (size_t i = 0; i < r.size();) {
        size_t j = i + 1;
        while (j < r.size() && r[j].lo == r[i].lo && r[j].hi == r[i].hi) ++j;
        if (j - i == 1) {
            le[nf] = r[i].slot / nv; ll[nf] = r[i].slot % nv; re[nf] = -1; rl[nf] = -1; ++nf;
        } else if (j - i == 2) {
            le[nf] = r[i].slot / nv; ll[nf] = r[i].slot % nv;
            re[nf] = r[i + 1].slot / nv; rl[nf] = r[i + 1].slot % nv; ++nf;
        }  // edges shared by more than two elements are dropped, as in the reference
        i = j;
    }
    *n_faces = nf;
    return ZJ_OK;
}

extern "C" int zj_pair_periodic(int* n_faces, int* le, int* ll, int* re, int* rl, int nv,
                                int stride, const int64_t* en, const double* xy, double Lx,
                                double Ly, double tol) {
    if (!n_faces || *n_faces < 0) return ZJ_ERR_ARG;
    const int nf = *n_faces;
    // boundary faces sorted by midpoint so each partner search is a binary search
    struct B { double x, y; int f; };
    std::vector<B> b;
    for (int f = 0; f < nf; ++f) {
        if (re[f] >= 0) continue;
        int64_t v0 = node_of(en, stride, le[f], ll[f]), v1 = node_of(en, stride, le[f], (ll[f] + 1) % nv);
        b.push_back({(xy[2 * v0] + xy[2 * v1]) * 0.5, (xy[2 * v0 + 1] + xy[2 * v1 + 1]) * 0.5, f});
    }
    std::vector<B> byx = b;
    std::sort(byx.begin(), byx.end(), [](const B& p, const B& q) { return p.x < q.x; });
    std::vector<char> dead(nf, 0), taken(nf, 0);
    std::vector<int> partner(nf, -1);
    auto candidates = [&](double x, double y, int self, std::vector<int>& out) {
        out.clear();
        auto lo = std::lower_bound(byx.begin(), byx.end(), x - tol, [](const B& p, double v) { return p.x < v; });
        for (auto it = lo; it != byx.end() && it->x <= x + tol; ++it)
            if (it->f != self && std::fabs(it->y - y) <= tol) out.push_back(it->f);
    };
    std::vector<int> cand, all;
    for (const B& q : b) {  // ascending face order
        int f = q.f;
        if (taken[f]) continue;
        all.clear();
        const double shifts[4][2] = {{Lx, 0}, {-Lx, 0}, {0, Ly}, {0, -Ly}};
        for (int s = 0; s < 4; ++s) {
            if ((s < 2 && Lx <= 0) || (s >= 2 && Ly <= 0)) continue;
            candidates(q.x + shifts[s][0], q.y + shifts[s][1], f, cand);
            for (int g : cand)
                if (g > f && !taken[g]) all.push_back(g);
        }
        std::sort(all.begin(), all.end());
        all.erase(std::unique(all.begin(), all.end()), all.end());
        if (all.size() > 1) return ZJ_ERR_ARG;  // ambiguous: not a periodic-conforming mesh
        if (all.size() == 1) {
            int g = all[0];
            partner[f] = g; taken[f] = taken[g] = 1; dead[g] = 1;
        }
    }
    int out = 0;
    for (int f = 0; f < nf; ++f) {
        if (dead[f]) continue;
        int g = partner[f];
        int nre = g >= 0 ? le[g] : re[f], nrl = g >= 0 ? ll[g] : rl[f];
        le[out] = le[f]; ll[out] = ll[f]; re[out] = nre; rl[out] = nrl; ++out;
    }
    *n_faces = out;
    return ZJ_OK;
}

// Pairing by node correspondence (a mesh file's $Periodic links).  Two
// boundary faces are partners when
//   - their end nodes correspond: the links join nodes into classes
//     (transitively, so a corner of a doubly periodic mesh is one class with
//     its three images), and both faces have the same two end-node classes;
//   - and one is the other translated: both end nodes move by the same
//     vector, which is (up to sign and tol) the offset master - slave of some
//     link.
// The first rule alone would pair the two wall edges of a mesh two elements
// across a periodic direction; the second alone would pair faces the file
// does not link.  More than one candidate is an error.  Kept face, removal
// and order follow zj_pair_periodic, so on a translation-periodic mesh both
// give the same table.
extern "C" int zj_pair_periodic_nodes(int* n_faces, int* le, int* ll, int* re, int* rl, int nv, int stride,
                                      int n_elem, const int64_t* en, const double* xy, int64_t n_nodes,
                                      int64_t n_pairs, const int64_t* slave, const int64_t* master, double tol) {
    if (!n_faces || *n_faces < 0 || nv < 3 || stride < nv || n_elem < 0 || n_nodes < 0 || n_pairs < 0 ||
        !(tol >= 0.0) || (n_pairs > 0 && (!slave || !master || !xy)) ||
        (*n_faces > 0 && (!le || !ll || !re || !rl || !en || !xy)))
        return ZJ_ERR_ARG;
    const int nf = *n_faces;
    for (int f = 0; f < nf; ++f)  // every face's left side indexes en
        if (le[f] < 0 || le[f] >= n_elem || ll[f] < 0 || ll[f] >= nv || re[f] >= n_elem ||
            (re[f] >= 0 && (rl[f] < 0 || rl[f] >= nv)))
            return ZJ_ERR_ARG;
    std::vector<int64_t> parent((size_t)n_nodes);
    std::iota(parent.begin(), parent.end(), int64_t(0));
    auto root = [&](int64_t a) {
        while (parent[(size_t)a] != a) {
            parent[(size_t)a] = parent[(size_t)parent[(size_t)a]];
            a = parent[(size_t)a];
        }
        return a;
    };
    std::vector<std::pair<double, double>> shifts;   // distinct link offsets, one sign
    auto close = [&](double x0, double y0, double x1, double y1) {
        return std::fabs(x0 - x1) <= tol && std::fabs(y0 - y1) <= tol;
    };
    for (int64_t i = 0; i < n_pairs; ++i) {
        const int64_t a = slave[i], b = master[i];
        if (a < 0 || a >= n_nodes || b < 0 || b >= n_nodes) return ZJ_ERR_ARG;
        const int64_t ra = root(a), rb = root(b);
        if (ra != rb) parent[(size_t)std::max(ra, rb)] = std::min(ra, rb);
        const double dx = xy[2 * b] - xy[2 * a], dy = xy[2 * b + 1] - xy[2 * a + 1];
        bool known = false;
        for (const auto& v : shifts)
            if (close(dx, dy, v.first, v.second) || close(-dx, -dy, v.first, v.second)) { known = true; break; }
        if (!known) shifts.emplace_back(dx, dy);
    }
    struct K { int64_t lo, hi; int f; };
    std::vector<K> keys;
    for (int f = 0; f < nf; ++f) {
        if (re[f] >= 0) continue;
        const int64_t v0 = node_of(en, stride, le[f], ll[f]), v1 = node_of(en, stride, le[f], (ll[f] + 1) % nv);
        if (v0 < 0 || v0 >= n_nodes || v1 < 0 || v1 >= n_nodes) return ZJ_ERR_ARG;
        const int64_t a = root(v0), b = root(v1);
        keys.push_back({std::min(a, b), std::max(a, b), f});
    }
    std::stable_sort(keys.begin(), keys.end(), [](const K& x, const K& y) {
        return x.lo != y.lo ? x.lo < y.lo : x.hi < y.hi;
    });
    // g is f translated by a link offset: both end nodes move by +-v
    auto translate_of = [&](int f, int g) {
        const int64_t a0 = node_of(en, stride, le[f], ll[f]), a1 = node_of(en, stride, le[f], (ll[f] + 1) % nv);
        const int64_t b0 = node_of(en, stride, le[g], ll[g]), b1 = node_of(en, stride, le[g], (ll[g] + 1) % nv);
        for (int swap = 0; swap < 2; ++swap) {
            const int64_t c0 = swap ? b1 : b0, c1 = swap ? b0 : b1;
            const double dx = xy[2 * c0] - xy[2 * a0], dy = xy[2 * c0 + 1] - xy[2 * a0 + 1];
            if (!close(xy[2 * c1] - xy[2 * a1], xy[2 * c1 + 1] - xy[2 * a1 + 1], dx, dy)) continue;
            for (const auto& v : shifts)
                if (close(dx, dy, v.first, v.second) || close(-dx, -dy, v.first, v.second)) return true;
        }
        return false;
    };
    std::vector<char> dead((size_t)std::max(nf, 1), 0), taken(dead.size(), 0);
    std::vector<int> partner((size_t)std::max(nf, 1), -1);
    std::vector<int> group;
    for (size_t i = 0; i < keys.size();) {
        size_t j = i + 1;
        while (j < keys.size() && keys[j].lo == keys[i].lo && keys[j].hi == keys[i].hi) ++j;
        group.clear();
        for (size_t k = i; k < j; ++k) group.push_back(keys[k].f);   // ascending face order
        for (size_t x = 0; x < group.size(); ++x) {
            const int f = group[x];
            if (taken[(size_t)f]) continue;
            int found = -1, count = 0;
            for (size_t y = x + 1; y < group.size(); ++y) {
                const int g = group[y];
                if (!taken[(size_t)g] && translate_of(f, g)) { found = g; ++count; }
            }
            if (count > 1) return ZJ_ERR_ARG;  // ambiguous partners
            if (count == 1) {
                partner[(size_t)f] = found;
                taken[(size_t)f] = taken[(size_t)found] = 1;
                dead[(size_t)found] = 1;
            }
        }
        i = j;
    }
    int out = 0;
    for (int f = 0; f < nf; ++f) {
        if (dead[(size_t)f]) continue;
        const int g = partner[(size_t)f];
        const int nre = g >= 0 ? le[g] : re[f], nrl = g >= 0 ? ll[g] : rl[f];
        le[out] = le[f]; ll[out] = ll[f]; re[out] = nre; rl[out] = nrl; ++out;
    }
    *n_faces = out;
    return ZJ_OK;
}

extern "C" int zj_face_normals(int nf, const int* le, const int* ll, int nv, int stride,
                               const int64_t* en, const double* xy, double* normals) {
    for (int f = 0; f < nf; ++f) {
        const int e = le[f];
        int64_t v0 = node_of(en, stride, e, ll[f]), v1 = node_of(en, stride, e, (ll[f] + 1) % nv);
        double tx = xy[2 * v1] - xy[2 * v0], ty = xy[2 * v1 + 1] - xy[2 * v0 + 1];
        double len = std::sqrt(tx * tx + ty * ty);
        double nx = ty / len, ny = -tx / len;
        double cx = 0.0, cy = 0.0;
        for (int k = 0; k < nv; ++k) {
            int64_t v = node_of(en, stride, e, k);
            cx = cx + xy[2 * v]; cy = cy + xy[2 * v + 1];
        }
        cx = cx / (double)nv; cy = cy / (double)nv;
        double mx = (xy[2 * v0] + xy[2 * v1]) * 0.5, my = (xy[2 * v0 + 1] + xy[2 * v1 + 1]) * 0.5;
        if (nx * (mx - cx) + ny * (my - cy) < 0) { nx = nx * (-1.0); ny = ny * (-1.0); }
        normals[2 * f] = nx; normals[2 * f + 1] = ny;
    }
    return ZJ_OK;
}

extern "C" int zj_elem_areas(int n_elem, int nv, int stride, const int64_t* en, const double* xy,
                             double* area) {
    auto tri = [&](int64_t a, int64_t b, int64_t c) {
        return 0.5 * std::fabs((xy[2 * b] - xy[2 * a]) * (xy[2 * c + 1] - xy[2 * a + 1]) -
                               (xy[2 * c] - xy[2 * a]) * (xy[2 * b + 1] - xy[2 * a + 1]));
    };
    for (int e = 0; e < n_elem; ++e) {
        int64_t v0 = node_of(en, stride, e, 0), v1 = node_of(en, stride, e, 1), v2 = node_of(en, stride, e, 2);
        if (nv == 3) area[e] = tri(v0, v1, v2);
        else area[e] = tri(v0, v1, v2) + tri(v0, v2, node_of(en, stride, e, 3));
    }
    return ZJ_OK;
}

extern "C" int zj_geometry(int n_elem, int nc, int stride, const int64_t* en, const double* xy, int S,
                           const double* sx, const double* se, double* J, double* Jinv, double* pos) {
    double N[25], Nx[25], Ne[25];
    for (int sp = 0; sp < S; ++sp)
        if (!shape(nc, sx[sp], se[sp], N, Nx, Ne)) return ZJ_ERR_UNSUPPORTED;
    for (int e = 0; e < n_elem; ++e)
        for (int sp = 0; sp < S; ++sp) {
            shape(nc, sx[sp], se[sp], N, Nx, Ne);
            double xx = 0, xe = 0, yx = 0, ye = 0, px = 0, py = 0;
            for (int i = 0; i < nc; ++i) {
                int64_t v = node_of(en, stride, e, i);
                const double X = xy[2 * v], Y = xy[2 * v + 1];
                xx += Nx[i] * X; xe += Ne[i] * X; yx += Nx[i] * Y; ye += Ne[i] * Y;
                px += N[i] * X; py += N[i] * Y;
            }
            const double det = xx * ye - xe * yx, inv = 1.0 / det;
            const int64_t o = (int64_t)e * S + sp;
            if (J) J[o] = det;
            Jinv[4 * o + 0] = ye * inv; Jinv[4 * o + 1] = -xe * inv;
            Jinv[4 * o + 2] = -yx * inv; Jinv[4 * o + 3] = xx * inv;
            if (pos) { pos[2 * o] = px; pos[2 * o + 1] = py; }
        }
    return ZJ_OK;
}
