// zj_host_math.hpp -- host-side basis mathematics shared by the C ABI tables
// (zj_host.cpp) and the C++ drop-in classes (zj_frsolver.cpp).
//
// Definitions follow the reference: Gauss rule basis.cpp:12-104, Lagrange
// basis.cpp:169-224, Radau correction derivatives basis.cpp:314-379, quad
// shape functions basis.cpp:717-847.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

namespace zj {
namespace host {
// 1D Gauss-Legendre rule with the reference's closed forms (n <= 5), its
// 15-digit table (n = 6) and Newton iteration beyond.
inline void gauss_rule(int n, double* x, double* w) {
    switch (n) {
    case 1: x[0] = 0.0; w[0] = 2.0; return;
    case 2: {
        double a = 1.0 / std::sqrt(3.0);
        x[0] = -a; x[1] = a; w[0] = w[1] = 1.0;
        return;
    }
    case 3: {
        double a = std::sqrt(3.0 / 5.0);
        x[0] = -a; x[1] = 0.0; x[2] = a;
        w[0] = w[2] = 5.0 / 9.0; w[1] = 8.0 / 9.0;
        return;
    }
    case 4: {
        double r = std::sqrt(6.0 / 5.0);
        double a = std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * r), b = std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * r);
        x[0] = -a; x[1] = -b; x[2] = b; x[3] = a;
        double s = std::sqrt(30.0);
        w[0] = w[3] = (18.0 - s) / 36.0; w[1] = w[2] = (18.0 + s) / 36.0;
        return;
    }
    case 5: {
        double r = std::sqrt(10.0 / 7.0);
        double a = std::sqrt(5.0 + 2.0 * r) / 3.0, b = std::sqrt(5.0 - 2.0 * r) / 3.0;
        x[0] = -a; x[1] = -b; x[2] = 0.0; x[3] = b; x[4] = a;
        double s = std::sqrt(70.0);
        w[0] = w[4] = (322.0 - 13.0 * s) / 900.0;
        w[1] = w[3] = (322.0 + 13.0 * s) / 900.0;
        w[2] = 128.0 / 225.0;
        return;
    }
    case 6: {
        static const double X[3] = {0.932469514203152, 0.661209386466265, 0.238619186083197};
        static const double W[3] = {0.171324492379170, 0.360761573048139, 0.467913934572691};
        for (int i = 0; i < 3; ++i) {
            x[i] = -X[i]; x[5 - i] = X[i]; w[i] = w[5 - i] = W[i];
        }
        return;
    }
    default:
        for (int i = 0; i < n; ++i) {
            double t = std::cos(M_PI * (i + 0.75) / (n + 0.5));
            auto leg = [n](double s, double& Pn, double& Pm) {
                double a = 1.0, b = s;
                for (int j = 2; j <= n; ++j) {
                    double c = ((2 * j - 1) * s * b - (j - 1) * a) / j;
                    a = b; b = c;
                }
                Pn = b; Pm = a;
            };
            for (int it = 0; it < 20; ++it) {
                double Pn, Pm;
                leg(t, Pn, Pm);
                double d = n * (t * Pn - Pm) / (t * t - 1.0);
                double dx = Pn / d;
                t -= dx;
                if (std::fabs(dx) < 1e-15) break;
            }
            double Pn, Pm;
            leg(t, Pn, Pm);
            double d = n * (t * Pn - Pm) / (t * t - 1.0);
            x[i] = t;
            w[i] = 2.0 / ((1.0 - t * t) * d * d);
        }
    }
}

// Lagrange basis on `nodes` (product and sum-of-products forms).
inline double lag(const double* nodes, int n, int j, double x) {
    double r = 1.0;
    for (int i = 0; i < n; ++i)
        if (i != j) r *= (x - nodes[i]) / (nodes[j] - nodes[i]);
    return r;
}

inline double lag_d(const double* nodes, int n, int j, double x) {
    double r = 0.0;
    for (int k = 0; k < n; ++k) {
        if (k == j) continue;
        double t = 1.0 / (nodes[j] - nodes[k]);
        for (int i = 0; i < n; ++i)
            if (i != j && i != k) t *= (x - nodes[i]) / (nodes[j] - nodes[i]);
        r += t;
    }
    return r;
}

inline double legendre_p(int n, double x) {
    if (n == 0) return 1.0;
    if (n == 1) return x;
    double a = 1.0, b = x;
    for (int i = 2; i <= n; ++i) {
        double c = ((2 * i - 1) * x * b - (i - 1) * a) / i;
        a = b; b = c;
    }
    return b;
}

inline double legendre_dp(int n, double x) {
    if (n == 0) return 0.0;
    double Pn = legendre_p(n, x), Pm = legendre_p(n - 1, x);
    if (std::fabs(x * x - 1.0) < 1e-14)  // the reference's endpoint branch (basis.cpp:336-339)
        return n * (n + 1) / 2.0 * ((x > 0) ? 1.0 : ((n % 2 == 0) ? 1.0 : -1.0));
    return n / (x * x - 1.0) * (x * Pn - Pm);
}

// Radau (DG) correction derivatives g_L', g_R' (Huynh 2007); the reference uses
// these for every FluxType.
inline double radau_left_d(int p, double x) {
    double s = (p % 2 == 0) ? 1.0 : -1.0;
    return s * 0.5 * (legendre_dp(p, x) - legendre_dp(p + 1, x));
}
inline double radau_right_d(int p, double x) { return 0.5 * (legendre_dp(p, x) + legendre_dp(p + 1, x)); }

// VCJH correction-function derivatives (Vincent, Castonguay, Jameson 2011):
// g_L = (-1)^p/2 [P_p - (eta P_{p-1} + P_{p+1})/(1+eta)], g_R(x) = g_L(-x), with
// eta = 0 (DG), p/(p+1) (SD, equal to Huynh's g_Ga), (p+1)/p (HU, Huynh's g2).
// kind: 0 DG, 1 SD, 2 HU, 3 GA.
inline double vcjh_eta(int p, int kind) {
    if (kind == 1 || kind == 3) return (double)p / (p + 1);
    if (kind == 2) return (double)(p + 1) / p;
    return 0.0;
}
// P_k'(x) for k = 0..n by P_k' = P_{k-2}' + (2k-1) P_{k-1}, exact at x = +-1
// as well (the reference's legendre_dp endpoint branch is not: basis.cpp:336-339).
inline void legendre_dp_all(int n, double x, double* dP) {
    double P[12];
    P[0] = 1.0; dP[0] = 0.0;
    if (n >= 1) { P[1] = x; dP[1] = 1.0; }
    for (int k = 2; k <= n; ++k) {
        P[k] = ((2 * k - 1) * x * P[k - 1] - (k - 1) * P[k - 2]) / k;
        dP[k] = dP[k - 2] + (2 * k - 1) * P[k - 1];
    }
}
inline double vcjh_left_d(int p, int kind, double x) {
    const double eta = vcjh_eta(p, kind), sgn = (p % 2) ? -1.0 : 1.0;
    double dP[12];
    legendre_dp_all(p + 1, x, dP);
    return sgn * 0.5 * (dP[p] - (eta * dP[p - 1] + dP[p + 1]) / (1.0 + eta));
}
inline double vcjh_right_d(int p, int kind, double x) { return -vcjh_left_d(p, kind, -x); }

inline int64_t node_of(const int64_t* en, int stride, int e, int k) { return en[(int64_t)e * stride + k]; }

// Quad shape functions in Gmsh node order: bilinear (4), biquadratic (9) and
// biquartic (25) tensor-product Lagrange.
inline const int Q9X[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1}, Q9Y[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
inline const int Q25X[25] = {0, 4, 4, 0, 1, 2, 3, 4, 4, 4, 3, 2, 1, 0, 0, 0, 1, 2, 3, 1, 2, 3, 1, 2, 3};
inline const int Q25Y[25] = {0, 0, 4, 4, 0, 0, 0, 1, 2, 3, 4, 4, 4, 3, 2, 1, 1, 1, 1, 2, 2, 2, 3, 3, 3};

inline bool shape(int nc, double xi, double eta, double* N, double* Nx, double* Ne) {
    if (nc == 4) {
        const double a = 1 - xi, b = 1 + xi, c = 1 - eta, d = 1 + eta;
        N[0] = 0.25 * a * c; N[1] = 0.25 * b * c; N[2] = 0.25 * b * d; N[3] = 0.25 * a * d;
        Nx[0] = -0.25 * c; Nx[1] = 0.25 * c; Nx[2] = 0.25 * d; Nx[3] = -0.25 * d;
        Ne[0] = -0.25 * a; Ne[1] = -0.25 * b; Ne[2] = 0.25 * b; Ne[3] = 0.25 * a;
        return true;
    }
    int m = nc == 9 ? 2 : nc == 25 ? 4 : 0;
    if (!m) return false;
    double nodes[5], Lx[5], dLx[5], Le[5], dLe[5];
    for (int i = 0; i <= m; ++i) nodes[i] = -1.0 + 2.0 * i / m;
    for (int i = 0; i <= m; ++i) {
        Lx[i] = lag(nodes, m + 1, i, xi); dLx[i] = lag_d(nodes, m + 1, i, xi);
        Le[i] = lag(nodes, m + 1, i, eta); dLe[i] = lag_d(nodes, m + 1, i, eta);
    }
    const int* IX = m == 2 ? Q9X : Q25X;
    const int* IY = m == 2 ? Q9Y : Q25Y;
    for (int k = 0; k < nc; ++k) {
        N[k] = Lx[IX[k]] * Le[IY[k]];
        Nx[k] = dLx[IX[k]] * Le[IY[k]];
        Ne[k] = Lx[IX[k]] * dLe[IY[k]];
    }
    return true;
}

// Index of (x, y) along the Hilbert curve of an n x n grid (n a power of two).
inline uint64_t hilbert_index(uint32_t n, uint32_t x, uint32_t y) {
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) { x = n - 1 - x; y = n - 1 - y; }
            std::swap(x, y);
        }
    }
    return d;
}

// Device element order: the caller's elements sorted along a Hilbert curve
// through their centroids, so that a tile of consecutive elements is a
// compact patch and fewer of its edges face another tile (55% -> 37% on a
// 256^2 grid at 10 elements per tile).  perm[new] = old.
inline std::vector<int> hilbert_order(int E, const double* c) {
    double x0 = c[0], x1 = c[0], y0 = c[1], y1 = c[1];
    for (int e = 0; e < E; ++e) {
        x0 = std::min(x0, c[2 * e]); x1 = std::max(x1, c[2 * e]);
        y0 = std::min(y0, c[2 * e + 1]); y1 = std::max(y1, c[2 * e + 1]);
    }
    const double span = std::max(std::max(x1 - x0, y1 - y0), 1e-300);
    const uint32_t n = 1u << 16;
    std::vector<std::pair<uint64_t, int>> key((size_t)E);
    for (int e = 0; e < E; ++e) {
        const uint32_t qx = (uint32_t)std::min<double>(n - 1, std::floor((c[2 * e] - x0) / span * n));
        const uint32_t qy = (uint32_t)std::min<double>(n - 1, std::floor((c[2 * e + 1] - y0) / span * n));
        key[(size_t)e] = {hilbert_index(n, qx, qy), e};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> perm((size_t)E);
    for (int e = 0; e < E; ++e) perm[(size_t)e] = key[(size_t)e].second;
    return perm;
}

}  // namespace host
}  // namespace zj
