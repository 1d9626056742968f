/* sanitize_oracle.c -- the CPU oracle (oracle/zj_oracle.c, zj_textbook.c)
 * under AddressSanitizer and UBSan (built with gcc -fsanitize=address,undefined
 * by tests/test_sanitize.py): a jittered periodic n x n quadrilateral mesh at
 * every order 1..5 through the reference-arithmetic residual, three limited
 * SSP-RK3 steps, the CFL dt and the norm, then the textbook residual and a
 * step for every Riemann solver and correction family, and the BR2 viscous
 * terms.  Prints one line with a checksum. */
#include "zj_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

int main(int argc, char** argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 5;
    const int E = n * n, N = (n + 1) * (n + 1);
    double* xy = malloc(sizeof(double) * 2 * N);
    int64_t* en = malloc(sizeof(int64_t) * 4 * E);
    unsigned seed = 7;
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            const int inner = i > 0 && i < n && j > 0 && j < n;
            seed = seed * 1103515245u + 12345u;
            const double d = inner ? 0.15 * (((seed >> 8) % 1000) / 1000.0 - 0.5) / n : 0.0;
            xy[2 * (j * (n + 1) + i)] = (double)i / n + d;
            xy[2 * (j * (n + 1) + i) + 1] = (double)j / n - d;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int64_t a = j * (n + 1) + i;
            int64_t* e = en + 4 * (j * n + i);
            e[0] = a; e[1] = a + 1; e[2] = a + n + 2; e[3] = a + n + 1;
        }
    int* le = malloc(sizeof(int) * 4 * E); int* ll = malloc(sizeof(int) * 4 * E);
    int* re = malloc(sizeof(int) * 4 * E); int* rl = malloc(sizeof(int) * 4 * E);
    int F = oz_build_faces(E, 4, 4, en, le, ll, re, rl);
    F = oz_pair_periodic(F, le, ll, re, rl, 4, 4, en, xy, 1.0, 1.0, 1e-9);
    if (F != 2 * E) return 3;
    double* nrm = malloc(sizeof(double) * 2 * F);
    double* area = malloc(sizeof(double) * E);
    double* ehl = malloc(sizeof(double) * 4 * E);
    oz_face_normals(F, le, ll, 4, 4, en, xy, nrm);
    oz_elem_areas(E, 4, 4, en, xy, area);
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < 4; ++k) {
            const int64_t a = en[4 * e + k], b = en[4 * e + (k + 1) % 4];
            ehl[4 * e + k] = 0.5 * hypot(xy[2 * b] - xy[2 * a], xy[2 * b + 1] - xy[2 * a + 1]);
        }
    double sum = 0.0;
    for (int p = 1; p <= 5; ++p) {
        const int S = (p + 1) * (p + 1), Fp = p + 1, nd = E * S * 4;
        double *sx = malloc(sizeof(double) * S), *se = malloc(sizeof(double) * S), *qw = malloc(sizeof(double) * S);
        double *dx = malloc(sizeof(double) * S * S), *de = malloc(sizeof(double) * S * S);
        double *ip = malloc(sizeof(double) * 4 * Fp * S), *co = malloc(sizeof(double) * 4 * S);
        double *J = malloc(sizeof(double) * E * S), *Ji = malloc(sizeof(double) * 4 * E * S);
        double *pos = malloc(sizeof(double) * 2 * E * S);
        double *U = malloc(sizeof(double) * nd), *R = malloc(sizeof(double) * nd), *Q = malloc(sizeof(double) * 2 * nd);
        oz_quad_operators(p, sx, se, qw, dx, de, ip, co);
        if (oz_geometry(E, 4, 4, en, xy, S, sx, se, J, Ji) || oz_sp_coords(E, 4, 4, en, xy, S, sx, se, pos)) return 4;
        for (int k = 0; k < E * S; ++k) {
            const double x = 6.283185307179586 * pos[2 * k], y = 6.283185307179586 * pos[2 * k + 1];
            const double u = 0.3 * sin(x) * cos(y), v = -0.3 * cos(x) * sin(y), pr = 150.0 + cos(2 * x);
            U[4 * k] = 1.0 + 0.05 * sin(x + y); U[4 * k + 1] = U[4 * k] * u; U[4 * k + 2] = U[4 * k] * v;
            U[4 * k + 3] = pr / 0.4 + 0.5 * U[4 * k] * (u * u + v * v);
        }
        oz_desc d = {0};
        d.p = p; d.n_elem = E; d.n_faces = F; d.gamma = 1.4;
        d.diff_xi = dx; d.diff_eta = de; d.interp = ip; d.corr = co; d.J = J; d.Jinv = Ji; d.normals = nrm;
        d.face_le = le; d.face_ll = ll; d.face_re = re; d.face_rl = rl; d.elem_area = area;
        oz_solver* s = oz_create(&d);
        oz_residual(s, U, R);
        for (int k = 0; k < nd; ++k) sum += fabs(R[k]) * 1e-6;
        const double dt = 0.2 * oz_dt_cfl(s, U, 0.5);
        for (int it = 0; it < 3; ++it) oz_advance(s, U, dt, 1);
        sum += oz_residual_l2(s) * 1e-6;
        for (int rk = 0; rk < 3; ++rk)
            for (int corr = 0; corr < 4; ++corr) {
                oz_tb* t = oz_tb_create(&d, ehl, rk, corr);
                oz_tb_residual(t, U, R);
                for (int k = 0; k < nd; ++k) sum += fabs(R[k]) * 1e-9;
                if (rk == 1 && corr == 2) {
                    oz_tb_set_viscous(t, 1e-3, 1.4e-3 / 0.4 / 0.72, 2.0);
                    oz_tb_gradient(t, U, Q);
                    oz_tb_advance(t, U, dt, 0);
                }
                oz_tb_destroy(t);
            }
        oz_destroy(s);
        free(sx); free(se); free(qw); free(dx); free(de); free(ip); free(co); free(J); free(Ji); free(pos);
        free(U); free(R); free(Q);
    }
    free(xy); free(en); free(le); free(ll); free(re); free(rl); free(nrm); free(area); free(ehl);
    printf("{\"checksum\": %.6e, \"finite\": %d}\n", sum, isfinite(sum) ? 1 : 0);
    return 0;
}
