// zj_tri_ops.hpp -- host construction of the straight-sided triangle FR
// operators (orders 1..5).
//
// The reference declares triangle operators but fills them with zeros
// (src/basis/basis.cpp:615-669), so these follow the textbook nodal simplex
// construction (Hesthaven & Warburton, "Nodal Discontinuous Galerkin Methods",
// 2008, sections 3.1 and 6.1):
//   reference triangle (-1,-1), (1,-1), (-1,1); edges e0 = v0->v1, e1 = v1->v2,
//   e2 = v2->v0, counter-clockwise;
//   solution points: warp-and-blend nodes with the optimised blend parameter,
//   numbered row by row from v0;
//   flux points: p+1 Gauss-Legendre points per edge, running along the edge
//   (the neighbour across a face sees point f at its slot p-f);
//   orthonormal Dubiner basis psi_m; V[i][m] = psi_m(x_i);
//   Dr = Vr V^-1, Ds = Vs V^-1, interp = Vf V^-1;
//   lift[i][j] = w_j sum_m psi_m(x_i) psi_m(y_j), the L2 lifting of a unit
//   value at flux point j (the DG-equivalent correction field, c = 0).
// Other correction fields: the energy-stable (VCJH) family on triangles of
// Castonguay, Vincent and Jameson (J. Sci. Comput. 51, 2012), whose norm is
//   |u|^2 = int_T u^2 + (c/2) sum_v C(p,v) (d^p u / dr^v ds^(p-v))^2;
// its correction field is the DG lifting filtered by (M + K)^-1 M, with K the
// c-term's matrix.  In the orthonormal modal basis M = I and, the p-th
// derivatives of degree-p modes being constants (zero below degree p) and the
// reference triangle's area 2, K = c sum_v C(p,v) d_v d_v^T with
// d_v[m] = d^p psi_m / dr^v ds^(p-v).  c takes the one-dimensional VCJH
// values of the family (Vincent, Castonguay and Jameson, J. Sci. Comput. 47,
// 2011): c_SD = 2p / ((2p+1)(p+1)(a_p p!)^2), c_HU = 2(p+1) / ((2p+1) p
// (a_p p!)^2), a_p = (2p)! / (2^p (p!)^2); GA is Huynh's g_Ga, the SD member,
// as on quadrilaterals.
#pragma once

#include "zj_host_math.hpp"

#include <cmath>
#include <vector>

namespace zj {
namespace tri {

inline int n_sp(int p) { return (p + 1) * (p + 2) / 2; }
inline int n_fp(int p) { return 3 * (p + 1); }

// Orthonormal Jacobi polynomial P_n^(a,b)(x) by its three-term recurrence.
inline double jacobi_p(double x, double a, double b, int n) {
    const double g0 = std::pow(2.0, a + b + 1) / (a + b + 1) * std::tgamma(a + 1) * std::tgamma(b + 1) /
                      std::tgamma(a + b + 1);
    double p0 = 1.0 / std::sqrt(g0);
    if (n == 0) return p0;
    const double g1 = (a + 1) * (b + 1) / (a + b + 3) * g0;
    double p1 = ((a + b + 2) * x / 2 + (a - b) / 2) / std::sqrt(g1);
    if (n == 1) return p1;
    double aold = 2 / (2 + a + b) * std::sqrt((a + 1) * (b + 1) / (a + b + 3));
    for (int i = 1; i < n; ++i) {
        const double h1 = 2 * i + a + b;
        const double anew = 2 / (h1 + 2) *
                            std::sqrt((i + 1) * (i + 1 + a + b) * (i + 1 + a) * (i + 1 + b) / (h1 + 1) / (h1 + 3));
        const double bnew = -(a * a - b * b) / h1 / (h1 + 2);
        const double p2 = (-aold * p0 + (x - bnew) * p1) / anew;
        p0 = p1; p1 = p2;
        aold = anew;
    }
    return p1;
}

inline double grad_jacobi_p(double x, double a, double b, int n) {
    return n == 0 ? 0.0 : std::sqrt(n * (n + a + b + 1)) * jacobi_p(x, a + 1, b + 1, n - 1);
}

// Gauss-Lobatto nodes (n+1 of them, ascending): the endpoints and the roots of
// P_n', found by Newton from the Chebyshev-Lobatto points.
inline void gauss_lobatto(int n, double* x) {
    x[0] = -1.0; x[n] = 1.0;
    for (int k = 1; k < n; ++k) {
        double t = -std::cos(M_PI * k / n);
        for (int it = 0; it < 100; ++it) {
            const double d1 = host::legendre_dp(n, t);
            const double d2 = (2 * t * d1 - n * (n + 1) * host::legendre_p(n, t)) / (1 - t * t);
            const double dx = d1 / d2;
            t -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        x[k] = t;
    }
}

// Warp of the equidistant nodes onto the Gauss-Lobatto nodes along an edge,
// divided by the edge blend 1 - r^2 (zero at the vertices).
inline double warp_factor(int n, double r) {
    std::vector<double> lgl(n + 1), req(n + 1);
    gauss_lobatto(n, lgl.data());
    for (int k = 0; k <= n; ++k) req[k] = -1.0 + 2.0 * k / n;
    double w = 0.0;
    for (int k = 0; k <= n; ++k) w += host::lag(req.data(), n + 1, k, r) * (lgl[k] - req[k]);
    if (std::fabs(r) >= 1.0 - 1e-10) return 0.0;
    return w / (1.0 - r * r);
}

// Warp-and-blend nodes [n_sp][2] in (r, s).
inline void solution_points(int p, double* rs) {
    static const double alpopt[10] = {0.0, 0.0, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832, 1.3648, 1.4773};
    const double alpha = p < 10 ? alpopt[p] : 5.0 / 3.0;
    const double c2 = std::cos(2 * M_PI / 3), c4 = std::cos(4 * M_PI / 3);
    const double s2 = std::sin(2 * M_PI / 3), s4 = std::sin(4 * M_PI / 3), r3 = std::sqrt(3.0);
    int k = 0;
    for (int n = 0; n <= p; ++n)
        for (int m = 0; m <= p - n; ++m, ++k) {
            const double L1 = double(n) / p, L3 = double(m) / p, L2 = 1.0 - L1 - L3;
            double x = -L2 + L3, y = (-L2 - L3 + 2 * L1) / r3;
            const double w1 = 4 * L2 * L3 * warp_factor(p, L3 - L2) * (1 + (alpha * L1) * (alpha * L1));
            const double w2 = 4 * L1 * L3 * warp_factor(p, L1 - L3) * (1 + (alpha * L2) * (alpha * L2));
            const double w3 = 4 * L1 * L2 * warp_factor(p, L2 - L1) * (1 + (alpha * L3) * (alpha * L3));
            x += w1 + c2 * w2 + c4 * w3;
            y += s2 * w2 + s4 * w3;
            // equilateral triangle -> reference triangle through barycentrics
            const double l1 = (r3 * y + 1.0) / 3.0;
            const double l2 = (-3.0 * x - r3 * y + 2.0) / 6.0;
            const double l3 = (3.0 * x - r3 * y + 2.0) / 6.0;
            rs[2 * k] = -l2 + l3 - l1;
            rs[2 * k + 1] = -l2 - l3 + l1;
        }
}

// Flux points [3][p+1][2] and the 1D Gauss weights [p+1].
inline void flux_points(int p, double* rs, double* w) {
    static const double V[3][2] = {{-1.0, -1.0}, {1.0, -1.0}, {-1.0, 1.0}};
    std::vector<double> xg(p + 1);
    host::gauss_rule(p + 1, xg.data(), w);
    for (int e = 0; e < 3; ++e) {
        const double* a = V[e];
        const double* b = V[(e + 1) % 3];
        for (int f = 0; f <= p; ++f) {
            const double t = 0.5 * (1 + xg[f]);
            rs[(e * (p + 1) + f) * 2] = a[0] + t * (b[0] - a[0]);
            rs[(e * (p + 1) + f) * 2 + 1] = a[1] + t * (b[1] - a[1]);
        }
    }
}

// Orthonormal Dubiner basis at m points: V, dV/dr, dV/ds, each [m][n_sp].
inline void dubiner(int p, int m, const double* rs, double* V, double* Vr, double* Vs) {
    const int np = n_sp(p);
    for (int q = 0; q < m; ++q) {
        const double r = rs[2 * q], s = rs[2 * q + 1];
        const double a = std::fabs(1 - s) > 1e-14 ? 2 * (1 + r) / (1 - s) - 1 : -1.0, b = s;
        int k = 0;
        for (int i = 0; i <= p; ++i)
            for (int j = 0; j <= p - i; ++j, ++k) {
                const double h1 = jacobi_p(a, 0, 0, i), h2 = jacobi_p(b, 2 * i + 1, 0, j);
                const double dh1 = grad_jacobi_p(a, 0, 0, i), dh2 = grad_jacobi_p(b, 2 * i + 1, 0, j);
                const double hb = 0.5 * (1 - b);
                V[q * np + k] = std::sqrt(2.0) * h1 * h2 * std::pow(1 - b, i);
                double dr = dh1 * h2, ds = dh1 * (h2 * (0.5 * (1 + a)));
                if (i > 0) { dr *= std::pow(hb, i - 1); ds *= std::pow(hb, i - 1); }
                double t = dh2 * std::pow(hb, i);
                if (i > 0) t -= 0.5 * i * h2 * std::pow(hb, i - 1);
                ds += h1 * t;
                Vr[q * np + k] = dr * std::pow(2.0, i + 0.5);
                Vs[q * np + k] = ds * std::pow(2.0, i + 0.5);
            }
    }
}

// In-place inverse of an n x n matrix (Gauss-Jordan, partial pivoting).
inline bool invert(int n, std::vector<double>& A) {
    std::vector<double> B((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) B[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)piv * n + c])) piv = r;
        if (std::fabs(A[(size_t)piv * n + c]) < 1e-300) return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(A[(size_t)c * n + k], A[(size_t)piv * n + k]);
                std::swap(B[(size_t)c * n + k], B[(size_t)piv * n + k]);
            }
        const double d = 1.0 / A[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) { A[(size_t)c * n + k] *= d; B[(size_t)c * n + k] *= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = A[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
                B[(size_t)r * n + k] -= f * B[(size_t)c * n + k];
            }
        }
    }
    A.swap(B);
    return true;
}

// One-dimensional VCJH parameter of a correction family (0 DG, 1 SD, 2 HU,
// 3 GA = SD); negative for an unknown family.
inline double vcjh_c(int p, int family) {
    double fact = 1.0, fact2 = 1.0;
    for (int k = 2; k <= p; ++k) fact *= k;
    for (int k = 2; k <= 2 * p; ++k) fact2 *= k;
    const double ap = fact2 / (std::pow(2.0, p) * fact * fact), q = (ap * fact) * (ap * fact);
    switch (family) {
    case 0: return 0.0;
    case 1: case 3: return 2.0 * p / ((2.0 * p + 1) * (p + 1) * q);
    case 2: return 2.0 * (p + 1) / ((2.0 * p + 1) * p * q);
    default: return -1.0;
    }
}

// d^p psi_m / dr^v ds^(p-v) of every Dubiner mode (constant for degree p,
// zero below): each mode is fitted exactly by the monomials r^a s^b, a+b <= p,
// on the np solution points, and the coefficient of r^v s^(p-v) times
// v! (p-v)! is the derivative.  Out: d[v][m], v = 0..p.
inline bool top_derivatives(int p, const std::vector<double>& sp, const std::vector<double>& V,
                            std::vector<double>& d) {
    const int np = n_sp(p);
    std::vector<double> Mono((size_t)np * np);   // [point][monomial], monomials by degree, then a descending
    std::vector<std::pair<int, int>> mono;
    for (int deg = 0; deg <= p; ++deg)
        for (int a = deg; a >= 0; --a) mono.push_back({a, deg - a});
    for (int q = 0; q < np; ++q)
        for (int k = 0; k < np; ++k)
            Mono[(size_t)q * np + k] =
                std::pow(sp[2 * q], mono[k].first) * std::pow(sp[2 * q + 1], mono[k].second);
    if (!invert(np, Mono)) return false;
    d.assign((size_t)(p + 1) * np, 0.0);
    for (int v = 0; v <= p; ++v) {
        int k = 0;
        while (!(mono[k].first == v && mono[k].second == p - v)) ++k;
        double f = 1.0;
        for (int t = 2; t <= v; ++t) f *= t;
        for (int t = 2; t <= p - v; ++t) f *= t;
        for (int m = 0; m < np; ++m) {
            double coef = 0.0;   // monomial coefficient k of mode m = (Mono^-1 V)[k][m]
            for (int q = 0; q < np; ++q) coef += Mono[(size_t)k * np + q] * V[(size_t)q * np + m];
            d[(size_t)v * np + m] = coef * f;
        }
    }
    return true;
}

struct Operators {
    int p = 0, np = 0, nf = 0;
    std::vector<double> sp, fp, wg;   // [np][2], [nf][2], [p+1]
    std::vector<double> Dr, Ds;       // [np][np]
    std::vector<double> interp;       // [nf][np]
    std::vector<double> lift;         // [np][nf]
};

inline bool operators(int p, Operators& o, int family = 0) {
    if (p < 1 || p > 5) return false;
    const double c = vcjh_c(p, family);
    if (c < 0.0) return false;
    const int np = n_sp(p), nf = n_fp(p);
    o.p = p; o.np = np; o.nf = nf;
    o.sp.assign((size_t)np * 2, 0.0);
    o.fp.assign((size_t)nf * 2, 0.0);
    o.wg.assign((size_t)p + 1, 0.0);
    solution_points(p, o.sp.data());
    flux_points(p, o.fp.data(), o.wg.data());
    std::vector<double> V((size_t)np * np), Vr((size_t)np * np), Vs((size_t)np * np);
    dubiner(p, np, o.sp.data(), V.data(), Vr.data(), Vs.data());
    std::vector<double> Vi = V;
    if (!invert(np, Vi)) return false;
    std::vector<double> Vf((size_t)nf * np), Vfr((size_t)nf * np), Vfs((size_t)nf * np);
    dubiner(p, nf, o.fp.data(), Vf.data(), Vfr.data(), Vfs.data());
    auto mul = [&](const std::vector<double>& A, int rows, std::vector<double>& out) {
        out.assign((size_t)rows * np, 0.0);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < np; ++j) {
                double acc = 0.0;
                for (int k = 0; k < np; ++k) acc += A[(size_t)i * np + k] * Vi[(size_t)k * np + j];
                out[(size_t)i * np + j] = acc;
            }
    };
    mul(Vr, np, o.Dr);
    mul(Vs, np, o.Ds);
    mul(Vf, nf, o.interp);
    // modal lifting coefficients of a unit value at flux point j, filtered by
    // (I + K)^-1 (the identity for DG)
    std::vector<double> F((size_t)np * np, 0.0);   // (I + K)^-1, modal
    for (int m = 0; m < np; ++m) F[(size_t)m * np + m] = 1.0;
    if (c > 0.0) {
        std::vector<double> d;
        if (!top_derivatives(p, o.sp, V, d)) return false;
        double binom = 1.0;
        for (int v = 0; v <= p; ++v) {
            for (int m = 0; m < np; ++m)
                for (int n = 0; n < np; ++n)
                    F[(size_t)m * np + n] += c * binom * d[(size_t)v * np + m] * d[(size_t)v * np + n];
            binom = binom * (p - v) / (v + 1);
        }
        if (!invert(np, F)) return false;
    }
    o.lift.assign((size_t)np * nf, 0.0);
    std::vector<double> hat(np);
    for (int j = 0; j < nf; ++j) {
        for (int m = 0; m < np; ++m) {
            double acc = 0.0;
            for (int n = 0; n < np; ++n) acc += F[(size_t)m * np + n] * Vf[(size_t)j * np + n];
            hat[m] = acc;
        }
        for (int i = 0; i < np; ++i) {
            double acc = 0.0;
            for (int m = 0; m < np; ++m) acc += V[(size_t)i * np + m] * hat[m];
            o.lift[(size_t)i * nf + j] = acc * o.wg[(size_t)(j % (p + 1))];
        }
    }
    return true;
}

}  // namespace tri
}  // namespace zj
