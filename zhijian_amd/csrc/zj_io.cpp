// zj_io.cpp -- restart files and visualisation output: the zj_restart_*,
// zj_vtk_write and zj_pvtu_write C ABI and the RestartIO / VTKWriter /
// ResidualWriter classes of include/zhijian/io.hpp.
//
// Restart files: src/io/restart.cpp of the reference, over the HDF5 codec in
// zj_h5.cpp.  Visualisation: src/io/visualization.cpp.  The lattice of every
// element is the same in reference coordinates, so the shape-function values
// (and, for interpolated output, the nodal-basis values) at the lattice points
// are tabulated once per element kind; each element then costs one small
// matrix product.  Elements are processed in parallel and text is formatted
// per thread with std::to_chars, which prints exactly what the reference's
// ostream does ("%.6g" by default, "%.12e" under std::scientific with
// setprecision(12)).
#include "zhijian_amd.h"

#include "zhijian/basis.hpp"
#include "zhijian/io.hpp"
#include "zhijian/physics.hpp"
#include "zj_h5.hpp"
#include "zj_host_math.hpp"
#include "zj_tri_ops.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int N_VARS_C = 4;   // conservative variables per point (rho, rho u, rho v, rho E)

thread_local std::string g_io_err;

int io_fail(int code, const std::string& msg) {
    g_io_err = msg;
    return code;
}

// Runs the body of a C entry point; a C++ exception (std::bad_alloc for a
// mesh or file too large for host memory) becomes an error code instead of
// crossing into the caller.
template <class F>
int guarded(F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return io_fail(ZJ_ERR_NOMEM, "out of host memory");
    } catch (const std::exception& e) {
        return io_fail(ZJ_ERR_ARG, e.what());
    }
}

// ============================================================================
// restart files
// ============================================================================

int restart_write(const char* path, int64_t n_elem, const int* nsp, const double* sol, double time, int iteration,
                  int poly_order, double gamma) {
    if (!path) return io_fail(ZJ_ERR_ARG, "null path");
    if (n_elem < 0 || (n_elem && !nsp)) return io_fail(ZJ_ERR_ARG, "elem_n_sp missing");
    uint64_t total = 0;
    for (int64_t e = 0; e < n_elem; ++e) {
        if (nsp[e] < 0) return io_fail(ZJ_ERR_ARG, "negative elem_n_sp");
        total += uint64_t(nsp[e]) * N_VARS_C;
    }
    if (total && !sol) return io_fail(ZJ_ERR_ARG, "null solution");
    using namespace zj::h5;
    std::vector<Attribute> at(4);
    at[0].name = "time";       at[0].type = Type::F64; at[0].f = time;
    at[1].name = "iteration";  at[1].type = Type::I32; at[1].i = iteration;
    at[2].name = "poly_order"; at[2].type = Type::I32; at[2].i = poly_order;
    at[3].name = "gamma";      at[3].type = Type::F64; at[3].f = gamma;
    std::vector<DatasetOut> ds(2);
    ds[0].name = "solution";  ds[0].type = Type::F64; ds[0].data = sol; ds[0].count = total;
    ds[1].name = "elem_n_sp"; ds[1].type = Type::I32; ds[1].data = nsp; ds[1].count = uint64_t(n_elem);
    std::string err;
    if (!write_file(path, at, ds, err)) return io_fail(ZJ_ERR_ARG, err);
    return ZJ_OK;
}

bool scalar_attr(const zj::h5::Reader& r, const char* name, double& v, std::string& err) {
    return r.attribute(name, v, err);
}

bool int_attr(const zj::h5::Reader& r, const char* name, int& v, std::string& err) {
    int64_t i;
    if (!r.attribute(name, i, err)) return false;
    if (i < INT32_MIN || i > INT32_MAX) {
        err = std::string("attribute '") + name + "' out of int range";
        return false;
    }
    v = int(i);
    return true;
}

}  // namespace

extern "C" const char* zj_io_last_error(void) { return g_io_err.c_str(); }

extern "C" int zj_restart_write(const char* path, int64_t n_elem, const int* elem_n_sp, const double* solution,
                                double time, int iteration, int poly_order, double gamma) {
    return guarded([&]() -> int {
        return restart_write(path, n_elem, elem_n_sp, solution, time, iteration, poly_order, gamma);
    });
}

extern "C" int zj_restart_query(const char* path, zj_restart_info* info) {
    return guarded([&]() -> int {
        if (!path || !info) return io_fail(ZJ_ERR_ARG, "null argument");
        zj::h5::Reader r;
        std::string err;
        if (!r.open(path, err)) return io_fail(ZJ_ERR_ARG, err);
        zj_restart_info o{};
        zj::h5::DatasetInfo dn, ds;
        if (!scalar_attr(r, "time", o.time, err) || !int_attr(r, "iteration", o.iteration, err) ||
            !int_attr(r, "poly_order", o.poly_order, err) || !scalar_attr(r, "gamma", o.gamma, err) ||
            !r.dataset("elem_n_sp", dn, err) || !r.dataset("solution", ds, err))
            return io_fail(ZJ_ERR_ARG, std::string(path) + ": " + err);
        o.n_elem = int64_t(dn.count);
        o.n_values = int64_t(ds.count);
        *info = o;
        return ZJ_OK;
    });
}

extern "C" int zj_restart_read(const char* path, int64_t n_elem, int* elem_n_sp, int64_t n_values, double* solution) {
    return guarded([&]() -> int {
        if (!path || n_elem < 0 || n_values < 0) return io_fail(ZJ_ERR_ARG, "bad argument");
        zj::h5::Reader r;
        std::string err;
        if (!r.open(path, err)) return io_fail(ZJ_ERR_ARG, err);
        zj::h5::DatasetInfo dn;
        if (!r.dataset("elem_n_sp", dn, err)) return io_fail(ZJ_ERR_ARG, err);
        if (dn.count != uint64_t(n_elem))
            return io_fail(ZJ_ERR_ARG, std::string(path) + ": elem_n_sp holds " + std::to_string(dn.count) +
                                           " entries, not " + std::to_string(n_elem));
        std::vector<int> nsp(static_cast<size_t>(n_elem));
        if (!r.read("elem_n_sp", nsp.data(), uint64_t(n_elem), err)) return io_fail(ZJ_ERR_ARG, err);
        uint64_t total = 0;
        for (int v : nsp) {
            if (v < 0) return io_fail(ZJ_ERR_ARG, "negative elem_n_sp in file");
            total += uint64_t(v) * N_VARS_C;
        }
        if (total != uint64_t(n_values))
            return io_fail(ZJ_ERR_ARG, "solution holds " + std::to_string(n_values) +
                                           " values but elem_n_sp adds up to " + std::to_string(total));
        if (solution && !r.read("solution", solution, uint64_t(n_values), err)) return io_fail(ZJ_ERR_ARG, err);
        if (elem_n_sp) std::copy(nsp.begin(), nsp.end(), elem_n_sp);
        return ZJ_OK;
    });
}

namespace {

// Device state of a handle -> restart file, and back.  get/set/info are the
// handle kind's own C entry points.
struct HandleOps {
    std::function<int(int64_t&, int&)> size;         // elements, points per element
    std::function<int(int&, double&)> scheme;        // poly order, gamma
    std::function<int(double*)> get;
    std::function<int(const double*)> set;
    std::function<const char*()> error;
};

// Pinned host staging for handle saves and loads: the device <-> host copy
// runs at full link speed and the file is written from / read into the same
// pages.  One buffer per process, grown to the largest state seen and kept
// until the process ends; saves and loads take turns on it.
class Staging {
public:
    double* get(size_t n, std::string& err) {
        if (n <= cap_) return buf_;
        if (buf_) (void)hipHostFree(buf_);
        buf_ = nullptr;
        cap_ = 0;
        void* p = nullptr;
        const hipError_t e = hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(double), hipHostMallocPortable);
        if (e != hipSuccess) {
            err = std::string("hipHostMalloc: ") + hipGetErrorString(e);
            return nullptr;
        }
        buf_ = static_cast<double*>(p);
        cap_ = n;
        return buf_;
    }
    std::mutex mu;

private:
    double* buf_ = nullptr;
    size_t cap_ = 0;
};

Staging& staging() {
    static Staging* s = new Staging();   // never destroyed: the HIP runtime may be gone at exit
    return *s;
}

int handle_save(const HandleOps& o, const char* path, double time, int iteration) {
    int64_t E;
    int S, p;
    double gamma;
    int rc = o.size(E, S);
    if (!rc) rc = o.scheme(p, gamma);
    if (rc) return io_fail(rc, o.error());
    std::lock_guard<std::mutex> lock(staging().mu);
    std::string err;
    double* U = staging().get(size_t(E) * S * N_VARS_C, err);
    if (!U) return io_fail(ZJ_ERR_NOMEM, err);
    if ((rc = o.get(U))) return io_fail(rc, o.error());
    std::vector<int> nsp(size_t(E), S);
    return restart_write(path, E, nsp.data(), U, time, iteration, p, gamma);
}

int handle_load(const HandleOps& o, const char* path, double* time, int* iteration) {
    int64_t E;
    int S, p;
    double gamma;
    int rc = o.size(E, S);
    if (!rc) rc = o.scheme(p, gamma);
    if (rc) return io_fail(rc, o.error());
    zj_restart_info info;
    if ((rc = zj_restart_query(path, &info))) return rc;
    if (info.n_elem != E || info.n_values != E * S * N_VARS_C)
        return io_fail(ZJ_ERR_ARG, std::string(path) + ": " + std::to_string(info.n_elem) + " elements / " +
                                       std::to_string(info.n_values) + " values do not match the mesh (" +
                                       std::to_string(E) + " elements of " + std::to_string(S) + " points)");
    if (info.poly_order != p)
        return io_fail(ZJ_ERR_ARG, std::string(path) + ": poly_order " + std::to_string(info.poly_order) +
                                       " differs from the solver's " + std::to_string(p));
    if (info.gamma != gamma) return io_fail(ZJ_ERR_ARG, std::string(path) + ": gamma differs from the solver's");
    std::vector<int> nsp(static_cast<size_t>(E));
    std::lock_guard<std::mutex> lock(staging().mu);
    std::string err;
    double* U = staging().get(size_t(info.n_values), err);
    if (!U) return io_fail(ZJ_ERR_NOMEM, err);
    if ((rc = zj_restart_read(path, E, nsp.data(), info.n_values, U))) return rc;
    for (int v : nsp)
        if (v != S) return io_fail(ZJ_ERR_ARG, std::string(path) + ": elem_n_sp differs from the mesh");
    if ((rc = o.set(U))) return io_fail(rc, o.error());
    if (time) *time = info.time;
    if (iteration) *iteration = info.iteration;
    return ZJ_OK;
}

HandleOps quad_ops(zj_handle h) {
    HandleOps o;
    o.size = [h](int64_t& E, int& S) {
        int64_t dof;
        const int rc = zj_get_info(h, &dof, nullptr, &S);
        E = S ? dof / (int64_t(S) * N_VARS_C) : 0;
        return rc;
    };
    o.scheme = [h](int& p, double& g) { return zj_get_scheme(h, &p, &g); };
    o.get = [h](double* U) { return zj_get_state(h, U); };
    o.set = [h](const double* U) { return zj_set_state(h, U); };
    o.error = [h]() { return zj_last_error(h); };
    return o;
}

HandleOps tri_ops(zj_tri_handle h) {
    HandleOps o;
    o.size = [h](int64_t& E, int& S) {
        int64_t dof;
        const int rc = zj_tri_get_info(h, &dof, nullptr, &S, nullptr);
        E = S ? dof / (int64_t(S) * N_VARS_C) : 0;
        return rc;
    };
    o.scheme = [h](int& p, double& g) { return zj_tri_get_scheme(h, &p, &g); };
    o.get = [h](double* U) { return zj_tri_get_state(h, U); };
    o.set = [h](const double* U) { return zj_tri_set_state(h, U); };
    o.error = [h]() { return zj_tri_last_error(h); };
    return o;
}

}  // namespace

extern "C" int zj_restart_save(zj_handle h, const char* path, double time, int iteration) {
    return guarded([&]() -> int {
        if (!h) return io_fail(ZJ_ERR_ARG, "null handle");
        return handle_save(quad_ops(h), path, time, iteration);
    });
}

extern "C" int zj_restart_load(zj_handle h, const char* path, double* time, int* iteration) {
    return guarded([&]() -> int {
        if (!h) return io_fail(ZJ_ERR_ARG, "null handle");
        return handle_load(quad_ops(h), path, time, iteration);
    });
}

extern "C" int zj_tri_restart_save(zj_tri_handle h, const char* path, double time, int iteration) {
    return guarded([&]() -> int {
        if (!h) return io_fail(ZJ_ERR_ARG, "null handle");
        return handle_save(tri_ops(h), path, time, iteration);
    });
}

extern "C" int zj_tri_restart_load(zj_tri_handle h, const char* path, double* time, int* iteration) {
    return guarded([&]() -> int {
        if (!h) return io_fail(ZJ_ERR_ARG, "null handle");
        return handle_load(tri_ops(h), path, time, iteration);
    });
}

// ============================================================================
// visualisation
// ============================================================================

namespace {

using zhijian::ElementType;
using zhijian::Real;
using zhijian::State;

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& fn) {
    const int64_t grain = 4096;
    int nt = int(std::min<int64_t>(std::max<unsigned>(1u, std::thread::hardware_concurrency()), 16));
    nt = int(std::min<int64_t>(nt, (n + grain - 1) / grain));
    if (nt <= 1) {
        if (n > 0) fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back([&, t] { fn(n * t / nt, n * (t + 1) / nt); });
    for (auto& x : th) x.join();
}

// One element kind (type, geometric order) on an n-cell lattice: shape
// functions and, for interpolated output, nodal-basis values at the points.
struct Kind {
    int npts = 0, ncells = 0, nnodes = 0, nsp = 0;
    std::vector<Real> N;        // [npts][nnodes]
    std::vector<Real> L;        // [npts][nsp] (interpolated output only)
    std::vector<int> cells;     // [ncells][3] local point indices
};

bool make_kind(ElementType type, int order, int n, bool interp, int p, Kind& k, std::string& err) {
    std::vector<std::pair<Real, Real>> ref;
    if (type == ElementType::Quadrilateral) {
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) ref.push_back({-1.0 + 2.0 * i / n, -1.0 + 2.0 * j / n});
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const int v00 = j * (n + 1) + i, v10 = v00 + 1, v01 = v00 + n + 1, v11 = v01 + 1;
                k.cells.insert(k.cells.end(), {v00, v10, v11, v00, v11, v01});
            }
    } else {
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n - j; ++i) ref.push_back({static_cast<Real>(i) / n, static_cast<Real>(j) / n});
        for (int j = 0; j < n; ++j) {
            const int row = j * (n + 1) - j * (j - 1) / 2, next = (j + 1) * (n + 1) - (j + 1) * j / 2;
            for (int i = 0; i < n - j; ++i) {
                k.cells.insert(k.cells.end(), {row + i, row + i + 1, next + i});
                if (i < n - j - 1) k.cells.insert(k.cells.end(), {row + i + 1, next + i + 1, next + i});
            }
        }
    }
    k.npts = int(ref.size());
    k.ncells = int(k.cells.size() / 3);
    std::vector<Real> Nq;
    try {
        zhijian::ElementGeometry::shapeFunc(type, order, 0.0, 0.0, Nq);
    } catch (const std::exception& e) {
        err = e.what();
        return false;
    }
    k.nnodes = int(Nq.size());
    k.N.resize(size_t(k.npts) * k.nnodes);
    for (int q = 0; q < k.npts; ++q) {
        zhijian::ElementGeometry::shapeFunc(type, order, ref[q].first, ref[q].second, Nq);
        std::copy(Nq.begin(), Nq.end(), k.N.begin() + size_t(q) * k.nnodes);
    }
    if (!interp) return true;
    if (p < 1 || p > 5) {
        err = "interpolated output needs poly_order 1..5";
        return false;
    }
    if (type == ElementType::Quadrilateral) {
        const int n1 = p + 1;
        k.nsp = n1 * n1;
        double x[8], w[8];
        zj::host::gauss_rule(n1, x, w);
        k.L.resize(size_t(k.npts) * k.nsp);
        for (int q = 0; q < k.npts; ++q)
            for (int b = 0; b < n1; ++b)
                for (int a = 0; a < n1; ++a)
                    k.L[size_t(q) * k.nsp + b * n1 + a] =
                        zj::host::lag(x, n1, a, ref[q].first) * zj::host::lag(x, n1, b, ref[q].second);
    } else {
        zj::tri::Operators ops;
        zj::tri::operators(p, ops);
        const int np = ops.np;
        k.nsp = np;
        std::vector<double> V(size_t(np) * np), Vr(V.size()), Vs(V.size());
        zj::tri::dubiner(p, np, ops.sp.data(), V.data(), Vr.data(), Vs.data());
        if (!zj::tri::invert(np, V)) {
            err = "singular triangle Vandermonde matrix";
            return false;
        }
        std::vector<double> rs(size_t(k.npts) * 2);
        for (int q = 0; q < k.npts; ++q) {
            rs[2 * q] = 2.0 * ref[q].first - 1.0;
            rs[2 * q + 1] = 2.0 * ref[q].second - 1.0;
        }
        std::vector<double> P(size_t(k.npts) * np), Pr(P.size()), Ps(P.size());
        zj::tri::dubiner(p, k.npts, rs.data(), P.data(), Pr.data(), Ps.data());
        k.L.assign(size_t(k.npts) * np, 0.0);
        for (int q = 0; q < k.npts; ++q)
            for (int j = 0; j < np; ++j) {
                double s = 0.0;
                for (int m = 0; m < np; ++m) s += P[size_t(q) * np + m] * V[size_t(m) * np + j];
                k.L[size_t(q) * np + j] = s;
            }
    }
    return true;
}

struct VisIn {
    const zj_vis_mesh* m;
    const int* nsp;
    const double* sol;
    int p;
    double gamma;
    int n;
    bool interp;
};

struct Lattice {
    int64_t npts = 0, ncells = 0;
    std::vector<double> xyz;      // [npts][3]
    std::vector<int64_t> conn;    // [ncells][3]
    std::vector<double> rho, pres, mach, vel;   // vel [npts][3]
};

bool build_lattice(const VisIn& in, Lattice& out, std::string& err) {
    const zj_vis_mesh& m = *in.m;
    const int64_t E = m.n_elem;
    if (in.n < 1) { err = "subdivisions must be >= 1"; return false; }
    if (E > 0 && (!m.elem_type || !m.elem_order || !m.elem_node_start || !m.elem_nodes || !m.xy || !in.nsp)) {
        err = "null mesh or solution array";
        return false;
    }
    std::map<std::pair<int, int>, std::unique_ptr<Kind>> kinds;
    std::vector<const Kind*> ek(static_cast<size_t>(E));
    std::vector<int64_t> soff(size_t(E) + 1, 0), poff(size_t(E) + 1, 0), coff(size_t(E) + 1, 0);
    for (int64_t e = 0; e < E; ++e) {
        const int t = m.elem_type[e], o = m.elem_order[e];
        if (t != 0 && t != 1) { err = "element " + std::to_string(e) + ": unknown type"; return false; }
        auto key = std::make_pair(t, o);
        auto it = kinds.find(key);
        if (it == kinds.end()) {
            auto k = std::make_unique<Kind>();
            if (!make_kind(ElementType(t), o, in.n, in.interp, in.p, *k, err)) return false;
            it = kinds.emplace(key, std::move(k)).first;
        }
        const Kind& k = *it->second;
        ek[e] = &k;
        const int64_t nn = m.elem_node_start[e + 1] - m.elem_node_start[e];
        if (m.elem_node_start[e] < 0 || nn < k.nnodes) {
            err = "element " + std::to_string(e) + ": too few nodes for its order";
            return false;
        }
        if (in.nsp[e] < 0 || (in.interp && in.nsp[e] != k.nsp)) {
            err = "element " + std::to_string(e) + ": " + std::to_string(in.nsp[e]) +
                  " solution points do not match poly_order " + std::to_string(in.p);
            return false;
        }
        soff[e + 1] = soff[e] + int64_t(in.nsp[e]) * N_VARS_C;
        poff[e + 1] = poff[e] + k.npts;
        coff[e + 1] = coff[e] + k.ncells;
    }
    if (soff[E] && !in.sol) { err = "null solution"; return false; }
    for (int64_t i = E ? m.elem_node_start[0] : 0; i < (E ? m.elem_node_start[E] : 0); ++i)
        if (m.elem_nodes[i] < 0 || m.elem_nodes[i] >= m.n_nodes) { err = "node index out of range"; return false; }
    out.npts = poff[E];
    out.ncells = coff[E];
    out.xyz.assign(size_t(out.npts) * 3, 0.0);
    out.conn.resize(size_t(out.ncells) * 3);
    out.rho.resize(size_t(out.npts));
    out.pres.resize(size_t(out.npts));
    out.mach.resize(size_t(out.npts));
    out.vel.assign(size_t(out.npts) * 3, 0.0);
    const zhijian::IdealGas gas(in.gamma);
    parallel_for(E, [&](int64_t e0, int64_t e1) {
        for (int64_t e = e0; e < e1; ++e) {
            const Kind& k = *ek[e];
            const int64_t* nodes = m.elem_nodes + m.elem_node_start[e];
            const double* U = in.sol + soff[e];
            for (int q = 0; q < k.npts; ++q) {
                const Real* N = k.N.data() + size_t(q) * k.nnodes;
                Real x = 0.0, y = 0.0;
                for (int i = 0; i < k.nnodes; ++i) {
                    x += N[i] * m.xy[2 * nodes[i]];
                    y += N[i] * m.xy[2 * nodes[i] + 1];
                }
                State s;
                if (in.interp) {
                    const Real* L = k.L.data() + size_t(q) * k.nsp;
                    for (int v = 0; v < N_VARS_C; ++v) {
                        Real a = 0.0;
                        for (int j = 0; j < k.nsp; ++j) a += L[j] * U[j * N_VARS_C + v];
                        s[v] = a;
                    }
                } else if (in.nsp[e] > 0) {
                    for (int v = 0; v < N_VARS_C; ++v) s[v] = U[v];
                }
                const int64_t g = poff[e] + q;
                out.xyz[3 * g] = x;
                out.xyz[3 * g + 1] = y;
                out.rho[g] = s.rho();
                out.pres[g] = gas.pressure(s);
                out.mach[g] = gas.machNumber(s);
                out.vel[3 * g] = s.rhou() / s.rho();
                out.vel[3 * g + 1] = s.rhov() / s.rho();
            }
            for (int c = 0; c < k.ncells; ++c)
                for (int v = 0; v < 3; ++v) out.conn[size_t(coff[e] + c) * 3 + v] = poff[e] + k.cells[3 * c + v];
        }
    });
    return true;
}

// ---- text output
struct Fmt {
    enum Mode { G6, E12 } mode;
    void num(std::string& s, double v) const {
        char b[40];
        const auto r = (mode == G6) ? std::to_chars(b, b + sizeof b, v, std::chars_format::general, 6)
                                    : std::to_chars(b, b + sizeof b, v, std::chars_format::scientific, 12);
        s.append(b, r.ptr);
    }
    static void integer(std::string& s, int64_t v) {
        char b[24];
        const auto r = std::to_chars(b, b + sizeof b, v);
        s.append(b, r.ptr);
    }
};

class Sink {
public:
    explicit Sink(const std::string& path) { f_ = std::fopen(path.c_str(), "wb"); }
    ~Sink() { if (f_) std::fclose(f_); }
    bool ok() const { return f_ && !bad_; }
    void put(const std::string& s) { put(s.data(), s.size()); }
    void put(const void* p, size_t n) {
        if (f_ && !bad_ && n && std::fwrite(p, 1, n, f_) != n) bad_ = true;
    }
    bool close() {
        if (!f_) return false;
        const bool good = !bad_ && std::fclose(f_) == 0;
        f_ = nullptr;
        return good;
    }

private:
    std::FILE* f_ = nullptr;
    bool bad_ = false;
};

// Formats items [0, n) in parallel chunks, `line` appending item i, and writes
// the chunks in order.
void put_lines(Sink& out, int64_t n, const std::function<void(std::string&, int64_t)>& line) {
    const int64_t chunk = 1 << 16;
    for (int64_t s0 = 0; s0 < n; s0 += 16 * chunk) {
        const int64_t s1 = std::min(n, s0 + 16 * chunk);
        const int64_t nchunks = (s1 - s0 + chunk - 1) / chunk;
        std::vector<std::string> bufs(static_cast<size_t>(nchunks));
        parallel_for(nchunks * 4096, [&](int64_t a, int64_t b) {
            for (int64_t c = (a + 4095) / 4096; c < (b + 4095) / 4096; ++c) {
                std::string& s = bufs[size_t(c)];
                const int64_t i0 = s0 + c * chunk, i1 = std::min(s1, i0 + chunk);
                s.reserve(size_t(i1 - i0) * 48);
                for (int64_t i = i0; i < i1; ++i) line(s, i);
            }
        });
        for (auto& s : bufs) out.put(s);
    }
}

const char* const IND = "          ";

bool write_vtu_ascii(const std::string& path, const Lattice& L, std::string& err) {
    Sink f(path);
    if (!f.ok()) { err = "Failed to open file: " + path; return false; }
    const Fmt g{Fmt::G6};
    std::string h;
    h += "<?xml version=\"1.0\"?>\n";
    h += "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    h += "  <UnstructuredGrid>\n";
    h += "    <Piece NumberOfPoints=\"" + std::to_string(L.npts) + "\" NumberOfCells=\"" +
         std::to_string(L.ncells) + "\">\n";
    h += "      <Points>\n";
    h += "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    f.put(h);
    put_lines(f, L.npts, [&](std::string& s, int64_t i) {
        s += IND; g.num(s, L.xyz[3 * i]); s += ' '; g.num(s, L.xyz[3 * i + 1]); s += " 0.0\n";
    });
    f.put("        </DataArray>\n      </Points>\n      <Cells>\n"
          "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
    put_lines(f, L.ncells, [&](std::string& s, int64_t i) {
        s += IND; Fmt::integer(s, L.conn[3 * i]); s += ' '; Fmt::integer(s, L.conn[3 * i + 1]); s += ' ';
        Fmt::integer(s, L.conn[3 * i + 2]); s += '\n';
    });
    f.put("        </DataArray>\n        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
    put_lines(f, L.ncells, [&](std::string& s, int64_t i) {
        s += IND; Fmt::integer(s, (i + 1) * 3); s += '\n';
    });
    f.put("        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
    put_lines(f, L.ncells, [&](std::string& s, int64_t) { s += IND; s += "5\n"; });
    f.put("        </DataArray>\n      </Cells>\n      <PointData Scalars=\"Density\" Vectors=\"Velocity\">\n");
    const std::pair<const char*, const std::vector<double>*> scal[3] = {
        {"Density", &L.rho}, {"Pressure", &L.pres}, {"Mach", &L.mach}};
    for (const auto& sc : scal) {
        f.put(std::string("        <DataArray type=\"Float64\" Name=\"") + sc.first + "\" format=\"ascii\">\n");
        const std::vector<double>& a = *sc.second;
        put_lines(f, L.npts, [&](std::string& s, int64_t i) { s += IND; g.num(s, a[i]); s += '\n'; });
        f.put("        </DataArray>\n");
    }
    f.put("        <DataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    put_lines(f, L.npts, [&](std::string& s, int64_t i) {
        s += IND; g.num(s, L.vel[3 * i]); s += ' '; g.num(s, L.vel[3 * i + 1]); s += " 0.0\n";
    });
    f.put("        </DataArray>\n      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
    if (!f.close()) { err = "write failed: " + path; return false; }
    return true;
}

bool write_vtu_binary(const std::string& path, const Lattice& L, std::string& err) {
    Sink f(path);
    if (!f.ok()) { err = "Failed to open file: " + path; return false; }
    const bool wide = L.npts > INT32_MAX || L.ncells * 3 > INT32_MAX;
    const char* itype = wide ? "Int64" : "Int32";
    const uint64_t isz = wide ? 8 : 4;
    struct Arr { std::string xml; uint64_t bytes; };
    std::vector<Arr> arrs;
    uint64_t off = 0;
    auto add = [&](const std::string& head, uint64_t bytes) {
        arrs.push_back({head + " format=\"appended\" offset=\"" + std::to_string(off) + "\"/>\n", bytes});
        off += 8 + bytes;
    };
    const uint64_t np = uint64_t(L.npts), nc = uint64_t(L.ncells);
    add("        <DataArray type=\"Float64\" NumberOfComponents=\"3\"", np * 24);
    add(std::string("        <DataArray type=\"") + itype + "\" Name=\"connectivity\"", nc * 3 * isz);
    add(std::string("        <DataArray type=\"") + itype + "\" Name=\"offsets\"", nc * isz);
    add("        <DataArray type=\"UInt8\" Name=\"types\"", nc);
    add("        <DataArray type=\"Float64\" Name=\"Density\"", np * 8);
    add("        <DataArray type=\"Float64\" Name=\"Pressure\"", np * 8);
    add("        <DataArray type=\"Float64\" Name=\"Mach\"", np * 8);
    add("        <DataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"3\"", np * 24);
    std::string h;
    h += "<?xml version=\"1.0\"?>\n";
    h += "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    h += "  <UnstructuredGrid>\n";
    h += "    <Piece NumberOfPoints=\"" + std::to_string(L.npts) + "\" NumberOfCells=\"" +
         std::to_string(L.ncells) + "\">\n";
    h += "      <Points>\n" + arrs[0].xml + "      </Points>\n";
    h += "      <Cells>\n" + arrs[1].xml + arrs[2].xml + arrs[3].xml + "      </Cells>\n";
    h += "      <PointData Scalars=\"Density\" Vectors=\"Velocity\">\n";
    for (int a = 4; a < 8; ++a) h += arrs[a].xml;
    h += "      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n   _";
    f.put(h);
    auto block = [&](uint64_t bytes, const void* p) {
        f.put(&bytes, 8);
        f.put(p, size_t(bytes));
    };
    block(arrs[0].bytes, L.xyz.data());
    if (wide) {
        block(arrs[1].bytes, L.conn.data());
        std::vector<int64_t> o(nc);
        for (uint64_t i = 0; i < nc; ++i) o[i] = int64_t(i + 1) * 3;
        block(arrs[2].bytes, o.data());
    } else {
        std::vector<int32_t> c(nc * 3), o(nc);
        parallel_for(int64_t(nc * 3), [&](int64_t a, int64_t b) {
            for (int64_t i = a; i < b; ++i) c[size_t(i)] = int32_t(L.conn[size_t(i)]);
        });
        for (uint64_t i = 0; i < nc; ++i) o[i] = int32_t((i + 1) * 3);
        block(arrs[1].bytes, c.data());
        block(arrs[2].bytes, o.data());
    }
    std::vector<uint8_t> types(nc, 5);
    block(arrs[3].bytes, types.data());
    block(arrs[4].bytes, L.rho.data());
    block(arrs[5].bytes, L.pres.data());
    block(arrs[6].bytes, L.mach.data());
    block(arrs[7].bytes, L.vel.data());
    f.put("\n  </AppendedData>\n</VTKFile>\n");
    if (!f.close()) { err = "write failed: " + path; return false; }
    return true;
}

bool write_legacy(const std::string& path, const Lattice& L, std::string& err) {
    Sink f(path);
    if (!f.ok()) { err = "Failed to open file: " + path; return false; }
    const Fmt e{Fmt::E12};
    f.put("# vtk DataFile Version 3.0\nZhijian FR Solver Output\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    f.put("POINTS " + std::to_string(L.npts) + " double\n");
    put_lines(f, L.npts, [&](std::string& s, int64_t i) {
        e.num(s, L.xyz[3 * i]); s += ' '; e.num(s, L.xyz[3 * i + 1]); s += " 0.0\n";
    });
    const int64_t nc = L.ncells;
    f.put("CELLS " + std::to_string(nc) + " " + std::to_string(nc * 4) + "\n");
    put_lines(f, L.ncells, [&](std::string& s, int64_t i) {
        s += "3 "; Fmt::integer(s, L.conn[3 * i]); s += ' '; Fmt::integer(s, L.conn[3 * i + 1]); s += ' ';
        Fmt::integer(s, L.conn[3 * i + 2]); s += '\n';
    });
    f.put("CELL_TYPES " + std::to_string(nc) + "\n");
    put_lines(f, L.ncells, [&](std::string& s, int64_t) { s += "5\n"; });
    f.put("POINT_DATA " + std::to_string(L.npts) + "\n");
    const std::pair<const char*, const std::vector<double>*> scal[3] = {
        {"Density", &L.rho}, {"Pressure", &L.pres}, {"Mach", &L.mach}};
    for (const auto& sc : scal) {
        f.put(std::string("SCALARS ") + sc.first + " double 1\nLOOKUP_TABLE default\n");
        const std::vector<double>& a = *sc.second;
        put_lines(f, L.npts, [&](std::string& s, int64_t i) { e.num(s, a[i]); s += '\n'; });
    }
    f.put("VECTORS Velocity double\n");
    put_lines(f, L.npts, [&](std::string& s, int64_t i) {
        e.num(s, L.vel[3 * i]); s += ' '; e.num(s, L.vel[3 * i + 1]); s += " 0.0\n";
    });
    if (!f.close()) { err = "write failed: " + path; return false; }
    return true;
}

bool write_pvtu(const std::string& path, const std::vector<std::string>& pieces, std::string& err) {
    Sink f(path);
    if (!f.ok()) { err = "Failed to open file: " + path; return false; }
    std::string s;
    s += "<?xml version=\"1.0\"?>\n";
    s += "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    s += "  <PUnstructuredGrid GhostLevel=\"0\">\n";
    s += "    <PPointData Scalars=\"Density\" Vectors=\"Velocity\">\n";
    s += "      <PDataArray type=\"Float64\" Name=\"Density\"/>\n";
    s += "      <PDataArray type=\"Float64\" Name=\"Pressure\"/>\n";
    s += "      <PDataArray type=\"Float64\" Name=\"Mach\"/>\n";
    s += "      <PDataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"3\"/>\n";
    s += "    </PPointData>\n";
    s += "    <PPoints>\n      <PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n    </PPoints>\n";
    for (const auto& p : pieces) s += "    <Piece Source=\"" + p + "\"/>\n";
    s += "  </PUnstructuredGrid>\n</VTKFile>\n";
    f.put(s);
    if (!f.close()) { err = "write failed: " + path; return false; }
    return true;
}

int vtk_write(const std::string& path, const VisIn& in, int flags, std::string& err) {
    Lattice L;
    if (!build_lattice(in, L, err)) return ZJ_ERR_ARG;
    bool ok;
    if (flags & ZJ_VTK_LEGACY) ok = write_legacy(path, L, err);
    else if (flags & ZJ_VTK_BINARY) ok = write_vtu_binary(path, L, err);
    else ok = write_vtu_ascii(path, L, err);
    return ok ? ZJ_OK : ZJ_ERR_ARG;
}

}  // namespace

extern "C" int zj_vtk_write(const char* path, const zj_vis_mesh* mesh, const int* elem_n_sp, const double* solution,
                            int poly_order, double gamma, int subdivisions, int flags) {
    return guarded([&]() -> int {
        if (!path || !mesh) return io_fail(ZJ_ERR_ARG, "null argument");
        if (flags & ~(ZJ_VTK_LEGACY | ZJ_VTK_BINARY | ZJ_VTK_INTERPOLATE)) return io_fail(ZJ_ERR_ARG, "unknown flags");
        VisIn in{mesh, elem_n_sp, solution, poly_order, gamma, subdivisions, (flags & ZJ_VTK_INTERPOLATE) != 0};
        std::string err;
        const int rc = vtk_write(path, in, flags, err);
        return rc ? io_fail(rc, err) : ZJ_OK;
    });
}

extern "C" int zj_pvtu_write(const char* path, int n_pieces, const char* const* piece_files) {
    return guarded([&]() -> int {
        if (!path || n_pieces < 0 || (n_pieces && !piece_files)) return io_fail(ZJ_ERR_ARG, "null argument");
        std::vector<std::string> pieces;
        for (int k = 0; k < n_pieces; ++k) {
            if (!piece_files[k]) return io_fail(ZJ_ERR_ARG, "null piece name");
            pieces.push_back(piece_files[k]);
        }
        std::string err;
        return write_pvtu(path, pieces, err) ? ZJ_OK : io_fail(ZJ_ERR_ARG, err);
    });
}

// ============================================================================
// C++ drop-in classes
// ============================================================================

namespace zhijian {

namespace {

struct FlatMesh {
    std::vector<double> xy;
    std::vector<int> type, order;
    std::vector<int64_t> start, nodes;
    zj_vis_mesh desc{};
    explicit FlatMesh(const Mesh& m) {
        xy.reserve(size_t(m.numNodes()) * 2);
        for (const auto& p : m.nodes()) { xy.push_back(p.x); xy.push_back(p.y); }
        start.push_back(0);
        for (const auto& e : m.elements()) {
            type.push_back(int(e.type));
            order.push_back(e.order);
            for (Index n : e.node_ids) nodes.push_back(n);
            start.push_back(int64_t(nodes.size()));
        }
        desc = zj_vis_mesh{m.numNodes(), xy.data(), m.numElements(), type.data(), order.data(), start.data(),
                           nodes.data()};
    }
};

void flatten(const std::vector<ElementSolution>& sol, size_t n_elem, std::vector<int>& nsp, std::vector<double>& U) {
    nsp.resize(n_elem);
    U.clear();
    for (size_t e = 0; e < n_elem; ++e) {
        nsp[e] = int(sol[e].sol_pts.size());
        for (const State& s : sol[e].sol_pts)
            for (int v = 0; v < N_VARS; ++v) U.push_back(s[v]);
    }
}

}  // namespace

bool RestartIO::write(const std::string& filename, const Mesh&, const std::vector<ElementSolution>& solution,
                      const SimParams& params, Real time, int iteration) {
    std::vector<int> nsp;
    std::vector<double> U;
    flatten(solution, solution.size(), nsp, U);
    if (zj_restart_write(filename.c_str(), int64_t(nsp.size()), nsp.data(), U.data(), time, iteration,
                         params.poly_order, params.gamma)) {
        error_msg_ = zj_io_last_error();
        return false;
    }
    std::cout << "Restart file written: " << filename << std::endl;
    return true;
}

bool RestartIO::read(const std::string& filename, Mesh&, std::vector<ElementSolution>& solution, SimParams& params,
                     Real& time, int& iteration) {
    zj_restart_info info;
    if (zj_restart_query(filename.c_str(), &info)) {
        error_msg_ = zj_io_last_error();
        return false;
    }
    std::vector<int> nsp(static_cast<size_t>(info.n_elem));
    std::vector<double> U(static_cast<size_t>(info.n_values));
    if (zj_restart_read(filename.c_str(), info.n_elem, nsp.data(), info.n_values, U.data())) {
        error_msg_ = zj_io_last_error();
        return false;
    }
    time = info.time;
    iteration = info.iteration;
    params.poly_order = info.poly_order;
    params.gamma = info.gamma;
    solution.resize(nsp.size());
    size_t k = 0;
    for (size_t e = 0; e < nsp.size(); ++e) {
        solution[e].sol_pts.resize(size_t(nsp[e]));
        for (auto& s : solution[e].sol_pts)
            for (int v = 0; v < N_VARS; ++v) s[v] = U[k++];
    }
    std::cout << "Restart file read: " << filename << " (time=" << time << ", iter=" << iteration << ")"
              << std::endl;
    return true;
}

bool RestartIO::writeParallel(const std::string& filename, int rank, int num_procs, const Mesh&,
                              const std::vector<ElementSolution>& solution, const SimParams&, Real time,
                              int iteration) {
    std::vector<int> nsp;
    std::vector<double> U;
    flatten(solution, solution.size(), nsp, U);
    using namespace zj::h5;
    std::vector<Attribute> at(4);
    at[0].name = "time";      at[0].type = Type::F64; at[0].f = time;
    at[1].name = "iteration"; at[1].type = Type::I32; at[1].i = iteration;
    at[2].name = "num_procs"; at[2].type = Type::I32; at[2].i = num_procs;
    at[3].name = "rank";      at[3].type = Type::I32; at[3].i = rank;
    std::vector<DatasetOut> ds(1);
    ds[0].name = "solution"; ds[0].type = Type::F64; ds[0].data = U.data(); ds[0].count = U.size();
    std::string err;
    if (!write_file(filename + ".rank" + std::to_string(rank), at, ds, err)) {
        error_msg_ = "Failed to create HDF5 file: " + filename + " (" + err + ")";
        return false;
    }
    return true;
}

bool RestartIO::readParallel(const std::string& filename, int rank, int, Mesh& local_mesh,
                             std::vector<ElementSolution>& solution, SimParams& params, Real& time, int& iteration) {
    zj::h5::Reader r;
    std::string err;
    const std::string path = filename + ".rank" + std::to_string(rank);
    int it = 0;
    double t = 0.0;
    zj::h5::DatasetInfo info;
    if (!r.open(path, err) || !scalar_attr(r, "time", t, err) || !int_attr(r, "iteration", it, err) ||
        !r.dataset("solution", info, err)) {
        error_msg_ = "Failed to open HDF5 file: " + filename + " (" + err + ")";
        return false;
    }
    const int p = params.poly_order;
    uint64_t total = 0;
    for (const auto& e : local_mesh.elements())
        total += uint64_t(e.type == ElementType::Quadrilateral ? (p + 1) * (p + 1) : (p + 1) * (p + 2) / 2) * N_VARS;
    if (total != info.count) {
        error_msg_ = path + ": " + std::to_string(info.count) + " values do not match the local mesh at poly_order " +
                     std::to_string(p);
        return false;
    }
    std::vector<double> U(static_cast<size_t>(total));
    if (!r.read("solution", U.data(), total, err)) {
        error_msg_ = path + ": " + err;
        return false;
    }
    time = t;
    iteration = it;
    solution.resize(size_t(local_mesh.numElements()));
    size_t k = 0;
    for (Index e = 0; e < local_mesh.numElements(); ++e) {
        const auto& el = local_mesh.element(e);
        solution[e].sol_pts.resize(size_t(el.type == ElementType::Quadrilateral ? (p + 1) * (p + 1)
                                                                                 : (p + 1) * (p + 2) / 2));
        for (auto& s : solution[e].sol_pts)
            for (int v = 0; v < N_VARS; ++v) s[v] = U[k++];
    }
    return true;
}

bool VTKWriter::write(const std::string& filename, const Mesh& mesh, const std::vector<ElementSolution>& solution,
                      const SimParams& params, int format) {
    if (solution.size() < size_t(mesh.numElements())) {
        error_msg_ = "solution has fewer entries than the mesh has elements";
        return false;
    }
    FlatMesh fm(mesh);
    std::vector<int> nsp;
    std::vector<double> U;
    flatten(solution, size_t(mesh.numElements()), nsp, U);
    int flags = format;
    if (format == ZJ_VTK_VTU && binary_) flags |= ZJ_VTK_BINARY;
    if (high_order_ && params.compat == Compat::Textbook) flags |= ZJ_VTK_INTERPOLATE;
    if (zj_vtk_write(filename.c_str(), &fm.desc, nsp.data(), U.data(), params.poly_order, params.gamma,
                     subdivisions_, flags)) {
        error_msg_ = zj_io_last_error();
        return false;
    }
    return true;
}

bool VTKWriter::writeLegacy(const std::string& filename, const Mesh& mesh, const std::vector<ElementSolution>& solution,
                            const SimParams& params) {
    return write(filename, mesh, solution, params, ZJ_VTK_LEGACY);
}

bool VTKWriter::writeVTU(const std::string& filename, const Mesh& mesh, const std::vector<ElementSolution>& solution,
                         const SimParams& params) {
    return write(filename, mesh, solution, params, ZJ_VTK_VTU);
}

bool VTKWriter::writeParallelVTU(const std::string& base_filename, int rank, int num_procs, const Mesh& local_mesh,
                                 const std::vector<ElementSolution>& solution, const SimParams& params) {
    if (!writeVTU(base_filename + "_" + std::to_string(rank) + ".vtu", local_mesh, solution, params)) return false;
    if (rank != 0) return true;
    std::vector<std::string> pieces;
    for (int p = 0; p < num_procs; ++p) pieces.push_back(base_filename + "_" + std::to_string(p) + ".vtu");
    return writePVTU(base_filename + ".pvtu", num_procs, pieces);
}

bool VTKWriter::writePVTU(const std::string& filename, int, const std::vector<std::string>& piece_filenames) {
    std::string err;
    if (!write_pvtu(filename, piece_filenames, err)) {
        error_msg_ = err;
        return false;
    }
    return true;
}

bool ResidualWriter::open(const std::string& filename) {
    file_.open(filename);
    if (!file_.is_open()) return false;
    file_ << "# Iteration, Time, Residual, Rho_res, RhoU_res, RhoV_res, RhoE_res\n";
    return true;
}

void ResidualWriter::write(int iteration, Real time, Real residual,
                           const std::array<Real, N_VARS>& component_residuals) {
    if (!file_.is_open()) return;
    char b[64];
    std::snprintf(b, sizeof b, "%d, %.8e, %.8e", iteration, time, residual);
    file_ << b;
    for (int v = 0; v < N_VARS; ++v) {
        std::snprintf(b, sizeof b, ", %.8e", component_residuals[v]);
        file_ << b;
    }
    file_ << "\n";
    file_.flush();
}

void ResidualWriter::close() {
    if (file_.is_open()) file_.close();
}

}  // namespace zhijian
