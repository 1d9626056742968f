/*
 * zj_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see zj_oracle.h).
 *
 * Each function restates one piece of the reference's executed path and cites
 * the reference file:line it follows.  Loop order and arithmetic order follow
 * the reference so that results agree to rounding; no FMA contraction is used
 * (the build passes -ffp-contract=off).
 */
#include "zj_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NV 4 /* conservative variables */

static int g_threads = 1;
void oz_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int oz_threads(void) { return g_threads; }

/* ======================================================================= */
/* Operators                                                               */
/* ======================================================================= */

/* basis.cpp:12-104 */
int oz_gauss(int n, double* pts, double* wts) {
    switch (n) {
    case 1: pts[0] = 0.0; wts[0] = 2.0; return 0;
    case 2:
        pts[0] = -1.0 / sqrt(3.0); pts[1] = 1.0 / sqrt(3.0);
        wts[0] = wts[1] = 1.0; return 0;
    case 3:
        pts[0] = -sqrt(3.0 / 5.0); pts[1] = 0.0; pts[2] = sqrt(3.0 / 5.0);
        wts[0] = wts[2] = 5.0 / 9.0; wts[1] = 8.0 / 9.0; return 0;
    case 4: {
        double a = sqrt(3.0 / 7.0 + 2.0 / 7.0 * sqrt(6.0 / 5.0));
        double b = sqrt(3.0 / 7.0 - 2.0 / 7.0 * sqrt(6.0 / 5.0));
        pts[0] = -a; pts[1] = -b; pts[2] = b; pts[3] = a;
        wts[0] = wts[3] = (18.0 - sqrt(30.0)) / 36.0;
        wts[1] = wts[2] = (18.0 + sqrt(30.0)) / 36.0;
        return 0;
    }
    case 5: {
        double a = sqrt(5.0 + 2.0 * sqrt(10.0 / 7.0)) / 3.0;
        double b = sqrt(5.0 - 2.0 * sqrt(10.0 / 7.0)) / 3.0;
        pts[0] = -a; pts[1] = -b; pts[2] = 0.0; pts[3] = b; pts[4] = a;
        wts[0] = wts[4] = (322.0 - 13.0 * sqrt(70.0)) / 900.0;
        wts[1] = wts[3] = (322.0 + 13.0 * sqrt(70.0)) / 900.0;
        wts[2] = 128.0 / 225.0;
        return 0;
    }
    case 6: /* the reference's 15-digit table, basis.cpp:57-67 */
        pts[0] = -0.932469514203152; pts[1] = -0.661209386466265;
        pts[2] = -0.238619186083197; pts[3] = 0.238619186083197;
        pts[4] = 0.661209386466265;  pts[5] = 0.932469514203152;
        wts[0] = wts[5] = 0.171324492379170;
        wts[1] = wts[4] = 0.360761573048139;
        wts[2] = wts[3] = 0.467913934572691;
        return 0;
    default:
        if (n < 1) return -1;
        for (int i = 0; i < n; ++i) {
            double x = cos(M_PI * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 20; ++it) {
                double P0 = 1.0, P1 = x;
                for (int j = 2; j <= n; ++j) {
                    double Pn = ((2 * j - 1) * x * P1 - (j - 1) * P0) / j;
                    P0 = P1; P1 = Pn;
                }
                double dP = n * (x * P1 - P0) / (x * x - 1.0);
                double dx = P1 / dP;
                x -= dx;
                if (fabs(dx) < 1e-15) break;
            }
            pts[i] = x;
            double P0 = 1.0, P1 = x;
            for (int j = 2; j <= n; ++j) {
                double Pn = ((2 * j - 1) * x * P1 - (j - 1) * P0) / j;
                P0 = P1; P1 = Pn;
            }
            double dP = n * (x * P1 - P0) / (x * x - 1.0);
            wts[i] = 2.0 / ((1.0 - x * x) * dP * dP);
        }
        return 0;
    }
}

/* basis.cpp:169-178 */
double oz_lagrange(const double* nodes, int n, int j, double x) {
    double r = 1.0;
    for (int i = 0; i < n; ++i)
        if (i != j) r *= (x - nodes[i]) / (nodes[j] - nodes[i]);
    return r;
}

/* basis.cpp:180-196 */
double oz_lagrange_deriv(const double* nodes, int n, int j, double x) {
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

static double legendre(int n, double x) {
    if (n == 0) return 1.0;
    if (n == 1) return x;
    double P0 = 1.0, P1 = x;
    for (int i = 2; i <= n; ++i) {
        double Pn = ((2 * i - 1) * x * P1 - (i - 1) * P0) / i;
        P0 = P1; P1 = Pn;
    }
    return P1;
}

/* basis.cpp:332-341, including the endpoint branch */
static double legendre_deriv(int n, double x) {
    if (n == 0) return 0.0;
    double Pn = legendre(n, x);
    double Pn1 = legendre(n - 1, x);
    if (fabs(x * x - 1.0) < 1e-14)
        return n * (n + 1) / 2.0 * ((x > 0) ? 1.0 : ((n % 2 == 0) ? 1.0 : -1.0));
    return n / (x * x - 1.0) * (x * Pn - Pn1);
}

/* basis.cpp:314-348 */
double oz_corr_left_deriv(int p, double xi) {
    double dPp = legendre_deriv(p, xi), dPp1 = legendre_deriv(p + 1, xi);
    double sign = (p % 2 == 0) ? 1.0 : -1.0;
    return sign * 0.5 * (dPp - dPp1);
}

/* basis.cpp:350-379 */
double oz_corr_right_deriv(int p, double xi) {
    double dPp = legendre_deriv(p, xi), dPp1 = legendre_deriv(p + 1, xi);
    return 0.5 * (dPp + dPp1);
}

/* FROperators::initQuad, basis.cpp:464-613 */
void oz_quad_operators(int p, double* sol_xi, double* sol_eta, double* qw,
                       double* diff_xi, double* diff_eta, double* interp,
                       double* corr) {
    const int n = p + 1, S = n * n, Fp = n;
    double pts[8], w[8], D[8][8];
    oz_gauss(n, pts, w);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            int idx = j * n + i;
            sol_xi[idx] = pts[i]; sol_eta[idx] = pts[j];
            qw[idx] = w[i] * w[j];
        }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) D[i][k] = oz_lagrange_deriv(pts, n, k, pts[i]);
    memset(diff_xi, 0, sizeof(double) * S * S);
    memset(diff_eta, 0, sizeof(double) * S * S);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            int row = j * n + i;
            for (int k = 0; k < n; ++k) {
                diff_xi[row * S + j * n + k] = D[i][k];
                diff_eta[row * S + k * n + i] = D[j][k];
            }
        }
    /* flux points are the 1D Gauss points (FluxPoints::get1D, basis.cpp:442-446);
     * edges 2 and 3 run in the reversed direction (basis.cpp:540-552) */
    for (int e = 0; e < 4; ++e)
        for (int fp = 0; fp < Fp; ++fp) {
            double s = pts[fp], xf, ef;
            if (e == 0) { xf = s; ef = -1.0; }
            else if (e == 1) { xf = 1.0; ef = s; }
            else if (e == 2) { xf = -s; ef = 1.0; }
            else { xf = -1.0; ef = -s; }
            double Lx[8], Le[8];
            for (int i = 0; i < n; ++i) { Lx[i] = oz_lagrange(pts, n, i, xf); Le[i] = oz_lagrange(pts, n, i, ef); }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    interp[(e * Fp + fp) * S + j * n + i] = Lx[i] * Le[j];
        }
    /* basis.cpp:574-612 */
    for (int e = 0; e < 4; ++e)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int sp = j * n + i;
                double xi = sol_xi[sp], eta = sol_eta[sp], v;
                if (e == 0) v = oz_corr_left_deriv(p, eta) * oz_lagrange(pts, n, i, xi);
                else if (e == 2) v = oz_corr_right_deriv(p, eta) * oz_lagrange(pts, n, i, xi);
                else if (e == 3) v = oz_corr_left_deriv(p, xi) * oz_lagrange(pts, n, j, eta);
                else v = oz_corr_right_deriv(p, xi) * oz_lagrange(pts, n, j, eta);
                corr[e * S + sp] = v;
            }
}

/* ======================================================================= */
/* Mesh                                                                    */
/* ======================================================================= */

typedef struct { int64_t a, b; int elem, local, seq; } edge_rec;

static int edge_cmp(const void* pa, const void* pb) {
    const edge_rec* x = (const edge_rec*)pa;
    const edge_rec* y = (const edge_rec*)pb;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    if (x->b != y->b) return x->b < y->b ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/* mesh.cpp:18-83: std::map over sorted vertex pairs = lexicographic order;
 * within a key, elements in push order (element order, then local edge). */
int oz_build_faces(int n_elem, int nv, int stride, const int64_t* en,
                   int* face_le, int* face_ll, int* face_re, int* face_rl) {
    int ne = n_elem * nv;
    edge_rec* r = (edge_rec*)malloc(sizeof(edge_rec) * (ne > 0 ? ne : 1));
    if (!r) return -1;
    int c = 0;
    for (int e = 0; e < n_elem; ++e)
        for (int k = 0; k < nv; ++k) {
            int64_t v0 = en[(int64_t)e * stride + k];
            int64_t v1 = en[(int64_t)e * stride + (k + 1) % nv];
            r[c].a = v0 < v1 ? v0 : v1; r[c].b = v0 < v1 ? v1 : v0;
            r[c].elem = e; r[c].local = k; r[c].seq = c; ++c;
        }
    qsort(r, ne, sizeof(edge_rec), edge_cmp);
    int nf = 0;
    for (int i = 0; i < ne;) {
        int j = i;
        while (j < ne && r[j].a == r[i].a && r[j].b == r[i].b) ++j;
        int cnt = j - i;
        if (cnt == 1) {
            face_le[nf] = r[i].elem; face_ll[nf] = r[i].local;
            face_re[nf] = -1; face_rl[nf] = -1; ++nf;
        } else if (cnt == 2) {
            face_le[nf] = r[i].elem; face_ll[nf] = r[i].local;
            face_re[nf] = r[i + 1].elem; face_rl[nf] = r[i + 1].local; ++nf;
        } /* non-manifold edges are skipped, mesh.cpp:68-72 */
        i = j;
    }
    free(r);
    return nf;
}

static void face_mid(int le, int ll, int nv, int stride, const int64_t* en,
                     const double* xy, double* mx, double* my) {
    int64_t v0 = en[(int64_t)le * stride + ll], v1 = en[(int64_t)le * stride + (ll + 1) % nv];
    /* mesh.cpp:185-194 */
    *mx = (xy[2 * v0] + xy[2 * v1]) * 0.5;
    *my = (xy[2 * v0 + 1] + xy[2 * v1 + 1]) * 0.5;
}

int oz_pair_periodic(int nf, int* le, int* ll, int* re, int* rl, int nv, int stride,
                     const int64_t* en, const double* xy, double Lx, double Ly,
                     double tol) {
    int* dead = (int*)calloc(nf > 0 ? nf : 1, sizeof(int));
    double* mx = (double*)malloc(sizeof(double) * (nf > 0 ? nf : 1));
    double* my = (double*)malloc(sizeof(double) * (nf > 0 ? nf : 1));
    if (!dead || !mx || !my) { free(dead); free(mx); free(my); return -1; }
    for (int f = 0; f < nf; ++f)
        if (re[f] < 0) face_mid(le[f], ll[f], nv, stride, en, xy, &mx[f], &my[f]);
    int bad = 0;
    for (int f = 0; f < nf && !bad; ++f) {
        if (re[f] >= 0 || dead[f]) continue;
        int found = -1;
        for (int g = f + 1; g < nf; ++g) {
            if (re[g] >= 0 || dead[g]) continue;
            double dx = mx[g] - mx[f], dy = my[g] - my[f];
            int hit = 0;
            if (Lx > 0 && fabs(fabs(dx) - Lx) <= tol && fabs(dy) <= tol) hit = 1;
            if (Ly > 0 && fabs(fabs(dy) - Ly) <= tol && fabs(dx) <= tol) hit = 1;
            if (hit) {
                if (found >= 0) { bad = 1; break; }
                found = g;
            }
        }
        if (found >= 0) {
            re[f] = le[found]; rl[f] = ll[found];
            dead[found] = 1;
        }
    }
    int out = -1;
    if (!bad) {
        out = 0;
        for (int f = 0; f < nf; ++f) {
            if (dead[f]) continue;
            le[out] = le[f]; ll[out] = ll[f]; re[out] = re[f]; rl[out] = rl[f]; ++out;
        }
    }
    free(dead); free(mx); free(my);
    return out;
}

/* mesh.cpp:109-120 */
static void elem_centroid(int e, int nv, int stride, const int64_t* en, const double* xy,
                          double* cx, double* cy) {
    double x = 0, y = 0;
    for (int i = 0; i < nv; ++i) {
        int64_t v = en[(int64_t)e * stride + i];
        x = x + xy[2 * v]; y = y + xy[2 * v + 1];
    }
    *cx = x / (double)nv; *cy = y / (double)nv;
}

/* mesh.cpp:159-183 */
void oz_face_normals(int nf, const int* le, const int* ll, int nv, int stride,
                     const int64_t* en, const double* xy, double* normals) {
    for (int f = 0; f < nf; ++f) {
        int64_t v0 = en[(int64_t)le[f] * stride + ll[f]];
        int64_t v1 = en[(int64_t)le[f] * stride + (ll[f] + 1) % nv];
        double tx = xy[2 * v1] - xy[2 * v0], ty = xy[2 * v1 + 1] - xy[2 * v0 + 1];
        double len = sqrt(tx * tx + ty * ty);
        double nx = ty / len, ny = -tx / len;
        double cx, cy, fx, fy;
        elem_centroid(le[f], nv, stride, en, xy, &cx, &cy);
        face_mid(le[f], ll[f], nv, stride, en, xy, &fx, &fy);
        double tfx = fx - cx, tfy = fy - cy;
        if (nx * tfx + ny * tfy < 0) { nx = nx * (-1.0); ny = ny * (-1.0); }
        normals[2 * f] = nx; normals[2 * f + 1] = ny;
    }
}

/* mesh.cpp:122-146 */
void oz_elem_areas(int n_elem, int nv, int stride, const int64_t* en, const double* xy,
                   double* area) {
    for (int e = 0; e < n_elem; ++e) {
        const int64_t* v = en + (int64_t)e * stride;
        double x0 = xy[2 * v[0]], y0 = xy[2 * v[0] + 1];
        double x1 = xy[2 * v[1]], y1 = xy[2 * v[1] + 1];
        double x2 = xy[2 * v[2]], y2 = xy[2 * v[2] + 1];
        if (nv == 3) {
            area[e] = 0.5 * fabs((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0));
        } else {
            double x3 = xy[2 * v[3]], y3 = xy[2 * v[3] + 1];
            double a1 = 0.5 * fabs((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0));
            double a2 = 0.5 * fabs((x2 - x0) * (y3 - y0) - (x3 - x0) * (y2 - y0));
            area[e] = a1 + a2;
        }
    }
}

/* Quad shape functions (basis.cpp:717-847): bilinear Q1, biquadratic Q2 and
 * biquartic Q4, all tensor-product Lagrange in Gmsh node order. */
static const int Q2_XI[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1};
static const int Q2_ETA[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
static const int Q4_XI[25] = {0, 4, 4, 0, 1, 2, 3, 4, 4, 4, 3, 2, 1, 0, 0, 0, 1, 2, 3, 1, 2, 3, 1, 2, 3};
static const int Q4_ETA[25] = {0, 0, 4, 4, 0, 0, 0, 1, 2, 3, 4, 4, 4, 3, 2, 1, 1, 1, 1, 2, 2, 2, 3, 3, 3};

static void lag_equi(int m, double t, double* L, double* dL) {
    /* m+1 equispaced nodes on [-1,1] */
    double nodes[5];
    for (int i = 0; i <= m; ++i) nodes[i] = -1.0 + 2.0 * i / m;
    for (int i = 0; i <= m; ++i) {
        L[i] = oz_lagrange(nodes, m + 1, i, t);
        dL[i] = oz_lagrange_deriv(nodes, m + 1, i, t);
    }
}

static int shape(int nc, double xi, double eta, double* N, double* dNx, double* dNe) {
    if (nc == 4) {
        N[0] = 0.25 * (1 - xi) * (1 - eta); N[1] = 0.25 * (1 + xi) * (1 - eta);
        N[2] = 0.25 * (1 + xi) * (1 + eta); N[3] = 0.25 * (1 - xi) * (1 + eta);
        dNx[0] = -0.25 * (1 - eta); dNe[0] = -0.25 * (1 - xi);
        dNx[1] = 0.25 * (1 - eta);  dNe[1] = -0.25 * (1 + xi);
        dNx[2] = 0.25 * (1 + eta);  dNe[2] = 0.25 * (1 + xi);
        dNx[3] = -0.25 * (1 + eta); dNe[3] = 0.25 * (1 - xi);
        return 0;
    }
    int m = (nc == 9) ? 2 : (nc == 25) ? 4 : -1;
    if (m < 0) return -1;
    const int* IX = (m == 2) ? Q2_XI : Q4_XI;
    const int* IE = (m == 2) ? Q2_ETA : Q4_ETA;
    double Lx[5], dLx[5], Le[5], dLe[5];
    lag_equi(m, xi, Lx, dLx);
    lag_equi(m, eta, Le, dLe);
    for (int k = 0; k < nc; ++k) {
        N[k] = Lx[IX[k]] * Le[IE[k]];
        dNx[k] = dLx[IX[k]] * Le[IE[k]];
        dNe[k] = Lx[IX[k]] * dLe[IE[k]];
    }
    return 0;
}

/* fr_solver.cpp:85-129 + basis.cpp:849-863 */
int oz_geometry(int n_elem, int nc, int stride, const int64_t* en, const double* xy,
                int S, const double* sx, const double* se, double* J, double* Jinv) {
    double N[25], dNx[25], dNe[25];
    for (int e = 0; e < n_elem; ++e)
        for (int sp = 0; sp < S; ++sp) {
            if (shape(nc, sx[sp], se[sp], N, dNx, dNe)) return -1;
            double xx = 0, xe = 0, yx = 0, ye = 0;
            for (int i = 0; i < nc; ++i) {
                int64_t v = en[(int64_t)e * stride + i];
                xx += dNx[i] * xy[2 * v]; xe += dNe[i] * xy[2 * v];
                yx += dNx[i] * xy[2 * v + 1]; ye += dNe[i] * xy[2 * v + 1];
            }
            double Jd = xx * ye - xe * yx;
            double inv = 1.0 / Jd;
            int64_t o = (int64_t)e * S + sp;
            J[o] = Jd;
            Jinv[4 * o + 0] = ye * inv; Jinv[4 * o + 1] = -xe * inv;
            Jinv[4 * o + 2] = -yx * inv; Jinv[4 * o + 3] = xx * inv;
        }
    return 0;
}

/* basis.cpp:882-893 */
int oz_sp_coords(int n_elem, int nc, int stride, const int64_t* en, const double* xy,
                 int S, const double* sx, const double* se, double* pos) {
    double N[25], dNx[25], dNe[25];
    for (int e = 0; e < n_elem; ++e)
        for (int sp = 0; sp < S; ++sp) {
            if (shape(nc, sx[sp], se[sp], N, dNx, dNe)) return -1;
            double x = 0, y = 0;
            for (int i = 0; i < nc; ++i) {
                int64_t v = en[(int64_t)e * stride + i];
                x += N[i] * xy[2 * v]; y += N[i] * xy[2 * v + 1];
            }
            pos[2 * ((int64_t)e * S + sp)] = x; pos[2 * ((int64_t)e * S + sp) + 1] = y;
        }
    return 0;
}

/* ======================================================================= */
/* Physics                                                                 */
/* ======================================================================= */

/* IdealGas::pressure euler.hpp:16-23 */
static double pressure(const double* U, double g) {
    double rho = U[0], u = U[1] / rho, v = U[2] / rho, E = U[3] / rho;
    double ke = 0.5 * (u * u + v * v);
    return (g - 1.0) * rho * (E - ke);
}

/* IdealGas::fluxX / fluxY euler.hpp:101-129 */
void oz_flux_xy(const double* U, double g, double* Fx, double* Fy) {
    double rho = U[0], u = U[1] / rho, v = U[2] / rho, p = pressure(U, g);
    Fx[0] = U[1]; Fx[1] = U[1] * u + p; Fx[2] = U[1] * v; Fx[3] = (U[3] + p) * u;
    Fy[0] = U[2]; Fy[1] = U[2] * u; Fy[2] = U[2] * v + p; Fy[3] = (U[3] + p) * v;
}

/* IdealGas::primToConserv euler.hpp:81-88 */
static void prim2cons(double rho, double u, double v, double p, double g, double* U) {
    U[0] = rho; U[1] = rho * u; U[2] = rho * v;
    U[3] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v);
}

/* computeGhostState kernels.cu:453-525 */
void oz_ghost_state(const double* Ui, double nx, double ny, int bct, const double* bc_data,
                    int bc_idx, double g, double* G) {
    switch (bct) {
    case OZ_BC_WALL: case OZ_BC_SLIPWALL: case OZ_BC_SYMMETRY: {
        double rho = Ui[0], u = Ui[1] / rho, v = Ui[2] / rho, p = pressure(Ui, g);
        double vn = u * nx + v * ny;
        prim2cons(rho, u - 2.0 * vn * nx, v - 2.0 * vn * ny, p, g, G);
        return;
    }
    case OZ_BC_FARFIELD: {
        int o = bc_idx * 8;
        double rinf = bc_data[o], uinf = bc_data[o + 1], vinf = bc_data[o + 2], pinf = bc_data[o + 3];
        double ri = fmax(Ui[0], 1e-10);
        double ui = Ui[1] / ri, vi = Ui[2] / ri;
        double pi = fmax((g - 1.0) * (Ui[3] - 0.5 * ri * (ui * ui + vi * vi)), 1e-10);
        double vninf = uinf * nx + vinf * ny;
        double rb, ub, vb, pb;
        if (vninf > 0) { rb = rinf; ub = uinf; vb = vinf; pb = 0.5 * (pinf + pi); }
        else { rb = ri; ub = ui; vb = vi; pb = 0.5 * (pinf + pi); }
        rb = fmax(fmin(rb, 100.0), 0.01);
        ub = fmax(fmin(ub, 1000.0), -1000.0);
        vb = fmax(fmin(vb, 1000.0), -1000.0);
        pb = fmax(fmin(pb, 1e7), 1e3);
        prim2cons(rb, ub, vb, pb, g, G);
        return;
    }
    case OZ_BC_OUTFLOW: {
        double rho = Ui[0], u = Ui[1] / rho, v = Ui[2] / rho;
        prim2cons(rho, u, v, bc_data[bc_idx * 8], g, G);
        return;
    }
    default:
        for (int k = 0; k < NV; ++k) G[k] = Ui[k];
        return;
    }
}

/* inline Rusanov of computeRiemannFluxWithBCKernel, kernels.cu:597-668 */
void oz_rusanov_inline(const double* UL, const double* UR, double nx, double ny, double g,
                       double* F) {
    double rhoL = fmax(UL[0], 1e-10), rhoR = fmax(UR[0], 1e-10);
    double uL = UL[1] / rhoL, vL = UL[2] / rhoL, uR = UR[1] / rhoR, vR = UR[2] / rhoR;
    double EL = UL[3] / rhoL, ER = UR[3] / rhoR;
    double pL = fmax((g - 1.0) * rhoL * (EL - 0.5 * (uL * uL + vL * vL)), 1e-10);
    double pR = fmax((g - 1.0) * rhoR * (ER - 0.5 * (uR * uR + vR * vR)), 1e-10);
    double cL = sqrt(fmax(g * pL / rhoL, 1e-10));
    double cR = sqrt(fmax(g * pR / rhoR, 1e-10));
    double vnL = uL * nx + vL * ny, vnR = uR * nx + vR * ny;
    double velL = sqrt(uL * uL + vL * vL), velR = sqrt(uR * uR + vR * vR);
    double smax = fmax(velL + cL, velR + cR);
    double HL = EL + pL / rhoL, HR = ER + pR / rhoR;
    double FL0 = rhoL * vnL, FL1 = rhoL * uL * vnL + pL * nx;
    double FL2 = rhoL * vL * vnL + pL * ny, FL3 = rhoL * HL * vnL;
    double FR0 = rhoR * vnR, FR1 = rhoR * uR * vnR + pR * nx;
    double FR2 = rhoR * vR * vnR + pR * ny, FR3 = rhoR * HR * vnR;
    F[0] = 0.5 * (FL0 + FR0) - 0.5 * smax * (UR[0] - UL[0]);
    F[1] = 0.5 * (FL1 + FR1) - 0.5 * smax * (UR[1] - UL[1]);
    F[2] = 0.5 * (FL2 + FR2) - 0.5 * smax * (UR[2] - UL[2]);
    F[3] = 0.5 * (FL3 + FR3) - 0.5 * smax * (UR[3] - UL[3]);
    for (int k = 0; k < NV; ++k)
        if (isnan(F[k]) || isinf(F[k])) F[k] = 0.0;
}

/* ======================================================================= */
/* Residual                                                                */
/* ======================================================================= */

struct oz_solver {
    int p, n1, S, Fp, E, F;
    double gamma;
    double *diff_xi, *diff_eta, *interp, *corr, *J, *Jinv, *normals, *area;
    int *le, *ll, *re, *rl, *bct, *bcmap;
    double* bcd; int nbcd;
    /* buffers mirroring GPUSolverData and the per-call temporaries */
    double *Fx, *Fy, *Ufp, *Fface, *Fdiff, *dUdt, *U0, *U1;
};

static double* dup_d(const double* src, size_t n) {
    double* p = (double*)malloc(sizeof(double) * (n ? n : 1));
    if (p && src) memcpy(p, src, sizeof(double) * n);
    else if (p) memset(p, 0, sizeof(double) * (n ? n : 1));
    return p;
}
static int* dup_i(const int* src, size_t n, int fill) {
    int* p = (int*)malloc(sizeof(int) * (n ? n : 1));
    if (!p) return p;
    if (src) memcpy(p, src, sizeof(int) * n);
    else for (size_t i = 0; i < n; ++i) p[i] = fill;
    return p;
}

oz_solver* oz_create(const oz_desc* d) {
    if (d->p < 1 || d->p > 5 || d->n_elem < 0 || d->n_faces < 0) return NULL;
    oz_solver* s = (oz_solver*)calloc(1, sizeof(oz_solver));
    if (!s) return NULL;
    s->p = d->p; s->n1 = d->p + 1; s->S = s->n1 * s->n1; s->Fp = s->n1;
    s->E = d->n_elem; s->F = d->n_faces; s->gamma = d->gamma;
    size_t S = s->S, E = s->E, F = s->F, Fp = s->Fp;
    s->diff_xi = dup_d(d->diff_xi, S * S); s->diff_eta = dup_d(d->diff_eta, S * S);
    s->interp = dup_d(d->interp, 4 * Fp * S); s->corr = dup_d(d->corr, 4 * S);
    s->J = dup_d(d->J, E * S); s->Jinv = dup_d(d->Jinv, E * S * 4);
    s->normals = dup_d(d->normals, F * 2);
    s->area = dup_d(d->elem_area, E);
    s->le = dup_i(d->face_le, F, 0); s->ll = dup_i(d->face_ll, F, 0);
    s->re = dup_i(d->face_re, F, -1); s->rl = dup_i(d->face_rl, F, -1);
    s->bct = dup_i(d->bc_type, F, OZ_BC_INTERIOR); s->bcmap = dup_i(d->bc_face_map, F, -1);
    s->nbcd = d->n_bc_data; s->bcd = dup_d(d->bc_data, (size_t)(d->n_bc_data > 0 ? d->n_bc_data : 1) * 8);
    size_t sol = E * S * NV, fp = E * 4 * Fp * NV;
    s->Fx = dup_d(NULL, sol); s->Fy = dup_d(NULL, sol); s->dUdt = dup_d(NULL, sol);
    s->U0 = dup_d(NULL, sol); s->U1 = dup_d(NULL, sol);
    s->Ufp = dup_d(NULL, fp); s->Fdiff = dup_d(NULL, fp);
    s->Fface = dup_d(NULL, F * Fp * NV);
    if (!s->diff_xi || !s->diff_eta || !s->interp || !s->corr || !s->J || !s->Jinv ||
        !s->normals || !s->area || !s->le || !s->ll || !s->re || !s->rl || !s->bct ||
        !s->bcmap || !s->bcd || !s->Fx || !s->Fy || !s->dUdt || !s->U0 || !s->U1 ||
        !s->Ufp || !s->Fdiff || !s->Fface) {
        oz_destroy(s);
        return NULL;
    }
    return s;
}

void oz_destroy(oz_solver* s) {
    if (!s) return;
    void* ptrs[] = {s->diff_xi, s->diff_eta, s->interp, s->corr, s->J, s->Jinv, s->normals,
                    s->area, s->le, s->ll, s->re, s->rl, s->bct, s->bcmap, s->bcd, s->Fx,
                    s->Fy, s->Ufp, s->Fface, s->Fdiff, s->dUdt, s->U0, s->U1};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(s);
}

/* kernels.cu:157-194 */
void oz_k_flux_sp(oz_solver* s, const double* U) {
    const int total = s->E * s->S;
    const double g = s->gamma;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int idx = 0; idx < total; ++idx)
        oz_flux_xy(U + (size_t)idx * NV, g, s->Fx + (size_t)idx * NV, s->Fy + (size_t)idx * NV);
}

/* kernels.cu:30-67 */
void oz_k_interp(oz_solver* s, const double* U) {
    const int S = s->S, Fp = s->Fp, E = s->E;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int e = 0; e < E; ++e)
        for (int t = 0; t < 4 * Fp; ++t) {
            int edge = t / Fp, f = t % Fp;
            for (int v = 0; v < NV; ++v) {
                double val = 0.0;
                for (int sp = 0; sp < S; ++sp)
                    val += s->interp[edge * Fp * S + f * S + sp] * U[(size_t)e * S * NV + sp * NV + v];
                s->Ufp[((size_t)e * 4 + edge) * Fp * NV + f * NV + v] = val;
            }
        }
}

/* kernels.cu:528-676 */
void oz_k_riemann(oz_solver* s) {
    const int Fp = s->Fp, F = s->F;
    const double g = s->gamma;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int f = 0; f < F; ++f)
        for (int fp = 0; fp < Fp; ++fp) {
            size_t lo = (size_t)s->le[f] * 4 * Fp * NV + (size_t)s->ll[f] * Fp * NV + fp * NV;
            const double* UL = s->Ufp + lo;
            double nx = s->normals[2 * f], ny = s->normals[2 * f + 1];
            double UR[NV];
            if (s->re[f] >= 0) {
                int rfp = Fp - 1 - fp;
                size_t ro = (size_t)s->re[f] * 4 * Fp * NV + (size_t)s->rl[f] * Fp * NV + rfp * NV;
                for (int v = 0; v < NV; ++v) UR[v] = s->Ufp[ro + v];
            } else {
                int bct = s->bct[f], bci = s->bcmap[f];
                if (bci < 0) { bct = OZ_BC_SLIPWALL; bci = 0; }
                oz_ghost_state(UL, nx, ny, bct, s->bcd, bci, g, UR);
            }
            oz_rusanov_inline(UL, UR, nx, ny, g, s->Fface + ((size_t)f * Fp + fp) * NV);
        }
}

/* fr_solver.cpp:501 memset + kernels.cu:346-426 */
void oz_k_flux_diff(oz_solver* s) {
    const int Fp = s->Fp, F = s->F;
    const double g = s->gamma;
    memset(s->Fdiff, 0, sizeof(double) * (size_t)s->E * 4 * Fp * NV);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int f = 0; f < F; ++f)
        for (int fp = 0; fp < Fp; ++fp) {
            double nx = s->normals[2 * f], ny = s->normals[2 * f + 1];
            const double* Fc = s->Fface + ((size_t)f * Fp + fp) * NV;
            double Fx[NV], Fy[NV];
            size_t lo = (size_t)s->le[f] * 4 * Fp * NV + (size_t)s->ll[f] * Fp * NV + fp * NV;
            oz_flux_xy(s->Ufp + lo, g, Fx, Fy);
            for (int v = 0; v < NV; ++v) {
                double Fint = Fx[v] * nx + Fy[v] * ny;
                s->Fdiff[lo + v] = Fc[v] - Fint;
            }
            if (s->re[f] >= 0) {
                size_t ro = (size_t)s->re[f] * 4 * Fp * NV + (size_t)s->rl[f] * Fp * NV + fp * NV;
                oz_flux_xy(s->Ufp + ro, g, Fx, Fy);
                for (int v = 0; v < NV; ++v) {
                    double Fn = Fx[v] * nx + Fy[v] * ny;
                    s->Fdiff[ro + v] = -Fc[v] + Fn;
                }
            }
        }
}

/* kernels.cu:813-859 */
void oz_k_divergence(oz_solver* s, double* dUdt) {
    const int S = s->S, E = s->E;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int e = 0; e < E; ++e)
        for (int sp = 0; sp < S; ++sp)
            for (int v = 0; v < NV; ++v) {
                size_t ji = (size_t)e * S * 4 + sp * 4;
                double dxidx = s->Jinv[ji], dxidy = s->Jinv[ji + 1];
                double detadx = s->Jinv[ji + 2], detady = s->Jinv[ji + 3];
                double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
                for (int k = 0; k < S; ++k) {
                    size_t ki = (size_t)e * S * NV + k * NV + v;
                    double fx = s->Fx[ki], fy = s->Fy[ki];
                    a += s->diff_xi[sp * S + k] * fx;
                    b += s->diff_eta[sp * S + k] * fx;
                    c += s->diff_xi[sp * S + k] * fy;
                    d += s->diff_eta[sp * S + k] * fy;
                }
                double div = (a * dxidx + b * detadx) + (c * dxidy + d * detady);
                dUdt[(size_t)e * S * NV + sp * NV + v] = -div;
            }
}

/* kernels.cu:948-1021 */
void oz_k_correction(oz_solver* s, double* dUdt) {
    const int S = s->S, E = s->E, Fp = s->Fp, n1 = Fp;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int e = 0; e < E; ++e)
        for (int sp = 0; sp < S; ++sp) {
            int si = sp / n1, sj = sp % n1;
            for (int v = 0; v < NV; ++v) {
                double c = 0.0;
                c += s->Fdiff[(((size_t)e * 4 + 0) * Fp + sj) * NV + v] * s->corr[0 * S + sp];
                c += s->Fdiff[(((size_t)e * 4 + 1) * Fp + si) * NV + v] * s->corr[1 * S + sp];
                c += s->Fdiff[(((size_t)e * 4 + 2) * Fp + sj) * NV + v] * s->corr[2 * S + sp];
                c += s->Fdiff[(((size_t)e * 4 + 3) * Fp + si) * NV + v] * s->corr[3 * S + sp];
                dUdt[(size_t)e * S * NV + sp * NV + v] -= c;
            }
        }
}

/* fr_solver.cpp:351-608 (NaN scans and debug prints omitted) */
void oz_residual(oz_solver* s, const double* U, double* dUdt) {
    oz_k_flux_sp(s, U);
    oz_k_interp(s, U);
    oz_k_riemann(s);
    oz_k_flux_diff(s);
    oz_k_divergence(s, s->dUdt);
    oz_k_correction(s, s->dUdt);
    if (dUdt) memcpy(dUdt, s->dUdt, sizeof(double) * (size_t)s->E * s->S * NV);
}

/* kernels.cu:1297-1349 */
void oz_rk_stage1(double* U1, const double* U0, const double* R, int n, double dt) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 0; i < n; ++i) U1[i] = U0[i] + dt * R[i];
}
void oz_rk_stage2(double* U2, const double* U0, const double* U1, const double* R, int n,
                  double dt) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 0; i < n; ++i) U2[i] = 0.75 * U0[i] + 0.25 * U1[i] + 0.25 * dt * R[i];
}
void oz_rk_stage3(double* U, const double* U0, const double* U2, const double* R, int n,
                  double dt) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 0; i < n; ++i)
        U[i] = (1.0 / 3.0) * U0[i] + (2.0 / 3.0) * U2[i] + (2.0 / 3.0) * dt * R[i];
}

/* kernels.cu:1236-1288 */
void oz_positivity(double* U, int n_states) {
    const double RHO_MIN = 0.01, RHO_MAX = 100.0, P_MIN = 100.0, P_MAX = 1e8;
    const double VEL_MAX = 2000.0, GPP = 1.4;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int k = 0; k < n_states; ++k) {
        double* q = U + (size_t)k * NV;
        double rho = q[0], rhou = q[1], rhov = q[2], rhoE = q[3];
        double rn = fmax(RHO_MIN, fmin(RHO_MAX, rho));
        double u = rhou / fmax(rn, 1e-10), v = rhov / fmax(rn, 1e-10);
        double vm = sqrt(u * u + v * v);
        if (vm > VEL_MAX) { double sc = VEL_MAX / vm; u *= sc; v *= sc; }
        double ke = 0.5 * rn * (u * u + v * v);
        double p = (GPP - 1.0) * (rhoE - ke);
        p = fmax(P_MIN, fmin(P_MAX, p));
        double Et = p / (GPP - 1.0) + ke;
        q[0] = rn; q[1] = rn * u; q[2] = rn * v; q[3] = Et;
    }
}

/* fr_solver.cpp:305-333 with gpu::rkStage* (kernels.cu:1305-1349) */
void oz_advance(oz_solver* s, double* U, double dt, int positivity) {
    const int n = s->E * s->S * NV, ns = s->E * s->S;
    memcpy(s->U0, U, sizeof(double) * n);
    oz_residual(s, U, NULL);
    oz_rk_stage1(U, s->U0, s->dUdt, n, dt);
    if (positivity) oz_positivity(U, ns);
    memcpy(s->U1, U, sizeof(double) * n);
    oz_residual(s, U, NULL);
    oz_rk_stage2(U, s->U0, s->U1, s->dUdt, n, dt);
    if (positivity) oz_positivity(U, ns);
    oz_residual(s, U, NULL);
    oz_rk_stage3(U, s->U0, U, s->dUdt, n, dt);
    if (positivity) oz_positivity(U, ns);
}

/* fr_solver.cpp:784-816 with IdealGas::maxWaveSpeed euler.hpp:34-41,166-171 */
double oz_dt_cfl(oz_solver* s, const double* U, double cfl) {
    double dt_min = 1e20;
    double cf = 2.0 * s->p + 1.0;
    for (int e = 0; e < s->E; ++e) {
        double h = sqrt(s->area[e]);
        for (int sp = 0; sp < s->S; ++sp) {
            const double* q = U + ((size_t)e * s->S + sp) * NV;
            double u = q[1] / q[0], v = q[2] / q[0];
            double vel = sqrt(u * u + v * v);
            double p = fmax(pressure(q, s->gamma), 1e-10), rho = fmax(q[0], 1e-10);
            double ws = vel + sqrt(s->gamma * p / rho);
            double dtl = cfl * h / (ws * cf);
            dt_min = fmin(dt_min, dtl);
        }
    }
    return dt_min;
}

/* fr_solver.cpp:818-833 */
double oz_residual_l2(oz_solver* s) {
    size_t n = (size_t)s->E * s->S * NV;
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) acc += s->dUdt[i] * s->dUdt[i];
    return sqrt(acc);
}

const double* oz_buf_fx(oz_solver* s) { return s->Fx; }
const double* oz_buf_fy(oz_solver* s) { return s->Fy; }
const double* oz_buf_ufp(oz_solver* s) { return s->Ufp; }
const double* oz_buf_fface(oz_solver* s) { return s->Fface; }
const double* oz_buf_fdiff(oz_solver* s) { return s->Fdiff; }
double* oz_buf_ufp_mut(oz_solver* s) { return s->Ufp; }
double* oz_buf_fdiff_mut(oz_solver* s) { return s->Fdiff; }
const double* oz_buf_dudt(oz_solver* s) { return s->dUdt; }
