// zj_adjoint.hip -- discrete adjoint (vector-Jacobian products) of the fused
// flux-reconstruction stage for gfx950.
//
// For one SSP-RK3 stage  u_n = a*U^n + b*q + c*dt*R(q),  u_out = limiter(u_n),
// the kernel takes the cotangent of u_out and produces the cotangents of q,
// U^n and dt.  It recomputes the stage's forward values in LDS (the same
// phases as the forward stage kernel) and then runs them backwards:
//
//   R1  cotangent of u_out = its partial value + the transposed trace
//       operator applied to the trace cotangents that the neighbouring tiles
//       produced (the completion of the previous adjoint launch); limiter
//       VJP; U^n and dt contributions; r = c*dt*du_n
//   R2  per solution point: transposed divergence (D^T along both lines, the
//       metric) and the Euler flux Jacobian^T
//       per (edge, flux point): transposed correction, the flux difference,
//       the Rusanov flux and the normal flux VJPs (or the ghost state VJP);
//       the cotangent of the neighbour's trace goes to LDS when the neighbour
//       is in this tile, else to HBM for the neighbour's tile to collect
//   R3  transposed trace operator onto the solution points; cotangent of q
//
// Every derivative is that of the reference's executed arithmetic
// (kernels.cu:157-194, 346-450, 453-701, 813-1034, 1236-1349): the clamps and
// maxima take the branch the forward took, and a flux component scrubbed to
// zero for being non-finite has zero derivative.
#include "zj_adjoint.hpp"
#include "zj_device.hpp"
#include "zj_riemann_vjp.hpp"

#include <cmath>
#include <type_traits>

namespace zj {
namespace {

// One side of the inline Rusanov flux (kernels.cu:601-648): the forward
// intermediates and the branch flags of its floors.
template <typename R>
struct RSide {
    R r, i, u, v, E, q, p, c, vn, s, H;
    bool rok, pok, cok;
};
template <typename R>
__device__ __forceinline__ RSide<R> rside(const R* U, R nx, R ny, R g) {
    const R tiny = R(1e-10);
    RSide<R> s;
    s.rok = U[0] >= tiny;
    s.r = fmaxR(U[0], tiny);
    s.i = rcpR(s.r);
    s.u = U[1] * s.i; s.v = U[2] * s.i; s.E = U[3] * s.i;
    s.q = s.u * s.u + s.v * s.v;
    const R praw = (g - R(1)) * s.r * (s.E - R(0.5) * s.q);
    s.pok = praw >= tiny;
    s.p = fmaxR(praw, tiny);
    const R c2 = g * s.p * s.i;
    s.cok = c2 >= tiny;
    s.c = sqrtR(fmaxR(c2, tiny));
    s.vn = s.u * nx + s.v * ny;
    s.s = sqrtR(s.q) + s.c;
    s.H = s.E + s.p * s.i;
    return s;
}
// Ub += VJP of F_side = [r vn, r u vn + p nx, r v vn + p ny, r H vn] (cotangent
// Fb) and of the side's wave speed s = |v| + c (cotangent sb).
template <typename R>
__device__ __forceinline__ void rside_vjp(const RSide<R>& s, const R* U, R nx, R ny, R g, const R* Fb, R sb,
                                          R* Ub) {
    const R gm = g - R(1);
    const R w = Fb[0] + Fb[1] * s.u + Fb[2] * s.v + Fb[3] * s.H;
    R rb = w * s.vn;
    const R vnb = s.r * w;
    R ub = Fb[1] * s.r * s.vn + vnb * nx;
    R vb = Fb[2] * s.r * s.vn + vnb * ny;
    R pb = Fb[1] * nx + Fb[2] * ny;
    const R Hb = Fb[3] * s.r * s.vn;
    R Eb = Hb;
    pb += Hb * s.i;
    R ib = Hb * s.p;
    R qb = s.q > R(0) ? sb * R(0.5) * rcpR(sqrtR(s.q)) : R(0);
    const R c2b = s.cok ? sb * R(0.5) * rcpR(s.c) : R(0);
    pb += c2b * g * s.i;
    ib += c2b * g * s.p;
    const R prb = s.pok ? pb : R(0);
    rb += prb * gm * (s.E - R(0.5) * s.q);
    Eb += prb * gm * s.r;
    qb -= R(0.5) * prb * gm * s.r;
    ub += R(2) * s.u * qb;
    vb += R(2) * s.v * qb;
    Ub[1] += ub * s.i; Ub[2] += vb * s.i; Ub[3] += Eb * s.i;
    ib += ub * U[1] + vb * U[2] + Eb * U[3];
    rb -= ib * s.i * s.i;
    if (s.rok) Ub[0] += rb;
}
// VJP of F = 1/2 (F_L + F_R) - 1/2 max(s_L, s_R) (U_R - U_L), scrubbed.
template <typename R>
__device__ __forceinline__ void rusanov_vjp(const R* UL, const R* UR, R nx, R ny, R g, const R* Fbin, R* ULb,
                                            R* URb) {
    const RSide<R> L = rside(UL, nx, ny, g), Q = rside(UR, nx, ny, g);
    const R smax = fmaxR(L.s, Q.s), hs = R(0.5) * smax;
    const R FL[NV] = {L.r * L.vn, L.r * L.u * L.vn + L.p * nx, L.r * L.v * L.vn + L.p * ny, L.r * L.H * L.vn};
    const R FR[NV] = {Q.r * Q.vn, Q.r * Q.u * Q.vn + Q.p * nx, Q.r * Q.v * Q.vn + Q.p * ny, Q.r * Q.H * Q.vn};
    R Fb[NV], Fh[NV], sb = R(0);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const R f = R(0.5) * (FL[k] + FR[k]) - hs * (UR[k] - UL[k]);
        Fb[k] = isfinite(f) ? Fbin[k] : R(0);
        Fh[k] = R(0.5) * Fb[k];
        sb -= Fh[k] * (UR[k] - UL[k]);
        ULb[k] += hs * Fb[k];
        URb[k] -= hs * Fb[k];
    }
    const bool lmax = L.s >= Q.s;
    rside_vjp(L, UL, nx, ny, g, Fh, lmax ? sb : R(0), ULb);
    rside_vjp(Q, UR, nx, ny, g, Fh, lmax ? R(0) : sb, URb);
}

// VJP of enforcePositivityKernel (kernels.cu:1236-1288) at the pre-limiter
// state q: qb = (d limiter / dq)^T ob.
__device__ __forceinline__ void positivity_vjp(const double* q, const double* ob, double* qb) {
    constexpr double GM1 = 1.4 - 1.0;
    const double rn = fmax(0.01, fmin(100.0, q[0]));
    const double ir = 1.0 / fmax(rn, 1e-10);
    const double u0 = q[1] * ir, v0 = q[2] * ir;
    const double vm = sqrt(u0 * u0 + v0 * v0);
    const bool capped = vm > 2000.0;
    const double sc = capped ? 2000.0 / vm : 1.0;
    const double u = u0 * sc, v = v0 * sc;
    const double ke = 0.5 * rn * (u * u + v * v);
    const double praw = GM1 * (q[3] - ke);
    const bool pact = praw >= 100.0 && praw <= 1e8;
    double rnb = ob[0] + ob[1] * u + ob[2] * v;
    double ub = ob[1] * rn, vb = ob[2] * rn;
    const double pb = ob[3] / GM1;
    double keb = ob[3];
    if (pact) {
        qb[3] = pb * GM1;
        keb -= pb * GM1;
    } else {
        qb[3] = 0.0;
    }
    rnb += keb * 0.5 * (u * u + v * v);
    ub += keb * rn * u;
    vb += keb * rn * v;
    if (capped) {  // (u, v) = 2000 (u0, v0)/|(u0, v0)|: project out the radial part
        const double hu = u0 / vm, hv = v0 / vm, d = ub * hu + vb * hv;
        ub = sc * (ub - d * hu);
        vb = sc * (vb - d * hv);
    }
    qb[1] = ub * ir;
    qb[2] = vb * ir;
    rnb -= (ub * q[1] + vb * q[2]) * ir * ir;
    qb[0] = (q[0] >= 0.01 && q[0] <= 100.0) ? rnb : 0.0;
}

// ------------------------------------------------------------------ kernel
template <typename R, int P, int EPB, int MODE, bool AFF, bool TB>
__device__ __forceinline__ void vjp_body(const VjpArgs<R>& a, const OpTables<R>& ops, const int tile) {
    using Y = Lay<R, P>;
    constexpr bool INC = std::is_same<R, float>::value;
    constexpr bool HAS_U0 = MODE == MODE_STAGE2 || MODE == MODE_STAGE3;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1;
    constexpr int NT = EPB * S, NF = EPB * 4 * FP;
    static_assert(NT <= BLOCK, "tile exceeds the workgroup");
    // stage coefficients u_n = ca U^n + cb q + cc dt R(q)  (kernels.cu:1297-1349)
    constexpr double ca = MODE == MODE_STAGE2 ? 0.75 : (MODE == MODE_STAGE3 ? 1.0 / 3.0 : 0.0);
    constexpr double cb = MODE == MODE_STAGE2 ? 0.25 : (MODE == MODE_STAGE3 ? 2.0 / 3.0 : 1.0);
    constexpr double cc = MODE == MODE_STAGE2 ? 0.25 : (MODE == MODE_STAGE3 ? 2.0 / 3.0 : 1.0);

    __shared__ R sU[EPB * Y::UE];
    __shared__ R sFx[EPB * Y::UE];   // contravariant/physical x flux, then the residual cotangent
    __shared__ R sFy[EPB * Y::UE];   // y flux, then (non-affine) the metric
    __shared__ R sTr[EPB * Y::TE];   // own traces
    __shared__ R sFd[EPB * Y::TE];   // flux differences, then own-slot trace cotangents
    __shared__ R sTm[EPB * Y::TE];   // trace cotangents at the mirrored slot (right role)
    __shared__ R sTb[EPB * Y::TE];   // incoming completion cotangents, then in-tile neighbour cotangents
    __shared__ R sD[N1 * N1], sLm[N1], sLp[N1], sC[4 * S];
    __shared__ unsigned char sIn[EPB * 4];   // edge partner in this tile
    __shared__ double sRed[BLOCK / 64];

    const int tid = threadIdx.x;
    const int e0 = tile * EPB;
    const int64_t PS = (int64_t)a.E * S;
    const R g = R(a.gamma);
    for (int t = tid; t < N1 * N1; t += BLOCK) sD[t] = ops.D[t];
    if (TB) {  // textbook: g_L', g_R' at the 1D points
        if (tid < N1) { sC[tid] = ops.gL[tid]; sC[N1 + tid] = ops.gR[tid]; }
    } else {
        for (int t = tid; t < 4 * S; t += BLOCK) sC[t] = ops.corr[t];
    }
    if (tid < N1) { sLm[tid] = ops.Lm[tid]; sLp[tid] = ops.Lp[tid]; }

    const int el = tid / S, sp = tid - (tid / S) * S;
    const int e = e0 + el;
    const bool active = tid < NT && e < a.E;
    const int64_t gi = (int64_t)e * S + sp;
    const int ub = el * Y::UE + sp;
    const int i = sp % N1, j = sp / N1;

    // The forward residual is needed only for the dt cotangent and the
    // limiter's VJP; without either, the kernel keeps just the traces.
    const bool need_res = MODE != MODE_RESIDUAL && (a.dtb_partial != nullptr || a.positivity);

    // ---- forward A: state, flux
    double q[NV], u0[NV];
    R jx = R(0), jy = R(0), kx = R(0), ky = R(0);
    R jd = R(1);  // textbook, non-affine: the transformed fluxes carry J
    R uq[NV];
    if (active) {
        if (AFF) {
            jx = a.Jinv[(int64_t)e * 4]; jy = a.Jinv[(int64_t)e * 4 + 1];
            kx = a.Jinv[(int64_t)e * 4 + 2]; ky = a.Jinv[(int64_t)e * 4 + 3];
        } else {
            jx = a.Jinv[gi]; jy = a.Jinv[PS + gi]; kx = a.Jinv[2 * PS + gi]; ky = a.Jinv[3 * PS + gi];
            if (TB) jd = a.Jpt[gi];
        }
        R fx[NV], fy[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (HAS_U0) u0[v] = a.U0[v * PS + gi];
            if (INC && HAS_U0) q[v] = u0[v] + double(a.dW_in[v * PS + gi]);
            else q[v] = a.U_in[v * PS + gi];
            uq[v] = R(q[v]);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sU[ub + v * Y::UV] = uq[v];
        if (need_res) {
        euler_flux(uq, g, fx, fy);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (AFF) {
                sFx[ub + v * Y::UV] = jx * fx[v] + jy * fy[v];
                sFy[ub + v * Y::UV] = kx * fx[v] + ky * fy[v];
            } else if (TB) {
                sFx[ub + v * Y::UV] = jd * (jx * fx[v] + jy * fy[v]);
                sFy[ub + v * Y::UV] = jd * (kx * fx[v] + ky * fy[v]);
            } else {
                sFx[ub + v * Y::UV] = fx[v]; sFy[ub + v * Y::UV] = fy[v];
            }
        }
        }
    }
    __syncthreads();

    // ---- forward B1: own traces
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        if (e0 + el2 < a.E) {
            const TraceLine tl = trace_line<P>(k, f);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                sTr[Y::tr(el2, k, f, v)] = trace_point<R, P>(sU + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
        }
    }
    __syncthreads();

    // ---- forward B2: flux differences; incoming completion cotangents
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        const int nl = L.nbr - e0;
        const bool inside = (L.role == ROLE_LEFT || L.role == ROLE_RIGHT) && nl >= 0 && nl < EPB;
        if (f == 0) sIn[el2 * 4 + k] = inside;
        const bool cross = (L.role == ROLE_LEFT || L.role == ROLE_RIGHT) && !inside;
        R tb[NV] = {R(0), R(0), R(0), R(0)};
        if (cross && a.Tb_in) load4(a.Tb_in + (((int64_t)e2 * 4 + k) * FP + f) * NV, tb);
#pragma unroll
        for (int v = 0; v < NV; ++v) sTb[Y::tr(el2, k, f, v)] = tb[v];
        R Fd[NV] = {R(0), R(0), R(0), R(0)};
        if (need_res && L.role != ROLE_NONE) {
            const R nx = a.normals[((int64_t)e2 * 4 + k) * 2], ny = a.normals[((int64_t)e2 * 4 + k) * 2 + 1];
            const bool right = L.role == ROLE_RIGHT;
            R own[NV], mirr[NV], other[NV], UL[NV], UR[NV], Fc[NV], Fn[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) { own[v] = sTr[Y::tr(el2, k, f, v)]; mirr[v] = sTr[Y::tr(el2, k, P - f, v)]; }
            // reference: the right role reads the neighbour's slot f (kernels.cu:577);
            // textbook: the geometrically matching slot p-f on both sides
            const int oslot = (TB || !right) ? P - f : f;
            if (L.role == ROLE_BOUNDARY) {
                ghost_state(own, nx, ny, L.nbr_edge, a.bc_data + (int64_t)L.bc_idx * 8, g, other);
            } else if (inside) {
#pragma unroll
                for (int v = 0; v < NV; ++v) other[v] = sTr[Y::tr(nl, L.nbr_edge, oslot, v)];
            } else {
                load4(a.T_in + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + oslot) * NV, other);
            }
            if (TB) {
#pragma unroll
                for (int v = 0; v < NV; ++v) { UL[v] = right ? other[v] : own[v]; UR[v] = right ? own[v] : other[v]; }
                riemann_by_kind(a.riemann, UL, UR, nx, ny, g, Fc);
                const R ef = a.edge_f[(int64_t)e2 * 4 + k];
                const TraceLine tl = trace_line<P>(k, f);
                const R* Fbase = ((k & 1) ? sFx : sFy) + el2 * Y::UE;
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    Fd[v] = ef * (right ? -Fc[v] : Fc[v]) - trace_point<R, P>(Fbase + v * Y::UV, tl, sLm, sLp);
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) { UL[v] = right ? other[v] : own[v]; UR[v] = right ? mirr[v] : other[v]; }
                rusanov(UL, UR, nx, ny, g, Fc);
                normal_flux(own, g, nx, ny, Fn);
#pragma unroll
                for (int v = 0; v < NV; ++v) Fd[v] = right ? Fn[v] - Fc[v] : Fc[v] - Fn[v];
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sFd[Y::tr(el2, k, f, v)] = Fd[v];
    }
    __syncthreads();

    // ---- forward C: residual
    double rsum = 0.0;  // dt cotangent partial
    double qb[NV] = {0.0, 0.0, 0.0, 0.0};
    double res[NV] = {0.0, 0.0, 0.0, 0.0};
    if (active && need_res) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R* Fx = sFx + el * Y::UE + v * Y::UV;
            const R* Fy = sFy + el * Y::UE + v * Y::UV;
            R div;
            if (AFF || TB) {
                R dxi = 0, deta = 0;
#pragma unroll
                for (int m = 0; m < N1; ++m) {
                    dxi += sD[i * N1 + m] * Fx[j * N1 + m];
                    deta += sD[j * N1 + m] * Fy[m * N1 + i];
                }
                div = dxi + deta;
            } else {
                R dFdxi = 0, dFdeta = 0, dGdxi = 0, dGdeta = 0;
#pragma unroll
                for (int m = 0; m < N1; ++m) {
                    const R di = sD[i * N1 + m], dj = sD[j * N1 + m];
                    dFdxi += di * Fx[j * N1 + m];
                    dGdxi += di * Fy[j * N1 + m];
                    dFdeta += dj * Fx[m * N1 + i];
                    dGdeta += dj * Fy[m * N1 + i];
                }
                div = (dFdxi * jx + dFdeta * kx) + (dGdxi * jy + dGdeta * ky);
            }
            if (TB) {
                const R c = sFd[Y::tr(el, 3, P - j, v)] * sC[i] + sFd[Y::tr(el, 1, j, v)] * sC[N1 + i] +
                            sFd[Y::tr(el, 0, i, v)] * sC[j] + sFd[Y::tr(el, 2, P - i, v)] * sC[N1 + j];
                res[v] = double(-(div + c) * (AFF ? R(1) : rcpR(jd)));
            } else {
                const R c = sFd[Y::tr(el, 0, i, v)] * sC[sp] + sFd[Y::tr(el, 1, j, v)] * sC[S + sp] +
                            sFd[Y::tr(el, 2, i, v)] * sC[2 * S + sp] + sFd[Y::tr(el, 3, j, v)] * sC[3 * S + sp];
                res[v] = double(-div - c);
            }
        }
    }
    __syncthreads();  // sFx, sFy, sFd are reused below

    // ---- R1: cotangent of the stage output, limiter, RK combination
    if (active) {
        double rb[NV];
        if (MODE == MODE_RESIDUAL) {
#pragma unroll
            for (int v = 0; v < NV; ++v) rb[v] = a.Ub_in[v * PS + gi];
        } else {
            const double dt = *a.dt_ptr;
            // cotangent of u_out: partial + transposed traces of the incoming edge cotangents
            double uo[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const R tsum = sLm[j] * sTb[Y::tr(el, 0, i, v)] + sLp[j] * sTb[Y::tr(el, 2, P - i, v)] +
                               sLp[i] * sTb[Y::tr(el, 1, j, v)] + sLm[i] * sTb[Y::tr(el, 3, P - j, v)];
                uo[v] = a.Ub_in[v * PS + gi] + double(tsum);
            }
            double un[NV], unb[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) un[v] = ca * (HAS_U0 ? u0[v] : 0.0) + cb * q[v] + cc * dt * res[v];
            if (a.positivity) {
                positivity_vjp(un, uo, unb);
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) unb[v] = uo[v];
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if (MODE == MODE_STAGE3) a.U0b[v * PS + gi] = ca * unb[v];
                if (MODE == MODE_STAGE2) a.U0b[v * PS + gi] += ca * unb[v];
                qb[v] = cb * unb[v];
                rb[v] = cc * dt * unb[v];
                rsum += cc * res[v] * unb[v];
            }
        }
        // textbook, non-affine: R = -(div + c) / J, so div and c see rb / J
        const R rs = (TB && !AFF) ? rcpR(jd) : R(1);
#pragma unroll
        for (int v = 0; v < NV; ++v) sFx[ub + v * Y::UV] = R(rb[v]) * rs;
        if (!AFF && !TB) {  // the per-point metric for the transposed divergence
            sFy[ub] = jx; sFy[ub + Y::UV] = jy; sFy[ub + 2 * Y::UV] = kx; sFy[ub + 3 * Y::UV] = ky;
        }
    }
    if (a.dtb_partial) {
        const double s = block_reduce<OpSum>(rsum, sRed);
        if (tid == 0) a.dtb_partial[e0 / EPB] = s;
    }
    __syncthreads();

    if (TB) {
    // ---- R2b (textbook): per (edge, flux point) the cotangent of
    // Delta = ef (+-F^I) - (interpolated transformed flux): the correction's
    // transpose gathers it from the points of the slot's grid line; its flux
    // part waits in sTm for the points, its common-flux part runs through the
    // Riemann solver's VJP to this side (sFd) and the other (sTb or HBM)
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        // edges 1, 3 correct along the xi line j = f (1) / p-f (3); edges 0, 2 the eta line i = f / p-f
        const R* Rb = sFx + el2 * Y::UE;
        const bool odd = (k & 1) != 0;
        const int line = (k == 0 || k == 1) ? f : P - f;
        const R* gw = sC + ((k == 1 || k == 2) ? N1 : 0);
        R Fdb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R acc = 0;
#pragma unroll
            for (int m = 0; m < N1; ++m) acc -= gw[m] * Rb[v * Y::UV + (odd ? line * N1 + m : m * N1 + line)];
            Fdb[v] = acc;
            sTm[Y::tr(el2, k, f, v)] = -acc;
        }
        R ownb[NV] = {R(0), R(0), R(0), R(0)};
        if (L.role != ROLE_NONE) {
            const R nx = a.normals[((int64_t)e2 * 4 + k) * 2], ny = a.normals[((int64_t)e2 * 4 + k) * 2 + 1];
            const bool right = L.role == ROLE_RIGHT;
            const int nl = L.nbr - e0;
            const bool inside = L.role != ROLE_BOUNDARY && nl >= 0 && nl < EPB;
            R own[NV], other[NV], UL[NV], UR[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) own[v] = sTr[Y::tr(el2, k, f, v)];
            const double* bcd = a.bc_data + (int64_t)(L.role == ROLE_BOUNDARY ? L.bc_idx : 0) * 8;
            if (L.role == ROLE_BOUNDARY) {
                ghost_state(own, nx, ny, L.nbr_edge, bcd, g, other);
            } else if (inside) {
#pragma unroll
                for (int v = 0; v < NV; ++v) other[v] = sTr[Y::tr(nl, L.nbr_edge, P - f, v)];
            } else {
                load4(a.T_in + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + (P - f)) * NV, other);
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) { UL[v] = right ? other[v] : own[v]; UR[v] = right ? own[v] : other[v]; }
            const R ef = a.edge_f[(int64_t)e2 * 4 + k];
            R Fcb[NV], ULb[NV] = {R(0), R(0), R(0), R(0)}, URb[NV] = {R(0), R(0), R(0), R(0)};
#pragma unroll
            for (int v = 0; v < NV; ++v) Fcb[v] = ef * (right ? -Fdb[v] : Fdb[v]);
            riemann_by_kind_vjp(a.riemann, UL, UR, nx, ny, g, Fcb, ULb, URb);
            R othb[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                ownb[v] = right ? URb[v] : ULb[v];
                othb[v] = right ? ULb[v] : URb[v];
            }
            if (L.role == ROLE_BOUNDARY) {
                ghost_state_vjp(own, nx, ny, L.nbr_edge, bcd, g, othb, ownb);
            } else if (inside) {
#pragma unroll
                for (int v = 0; v < NV; ++v) sTb[Y::tr(nl, L.nbr_edge, P - f, v)] = othb[v];
            } else {
                store4(a.Tb_out + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + (P - f)) * NV, othb);
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sFd[Y::tr(el2, k, f, v)] = ownb[v];  // sFd was last read in forward C
    }
    __syncthreads();

    // ---- R2a (textbook): transposed divergence plus the interpolated-flux
    // cotangents of the edge points on this point's lines, then the metric and
    // the Euler flux Jacobian
    if (active) {
        const R* Rb = sFx + el * Y::UE;
        R Fxb[NV], Fyb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R* rv = Rb + v * Y::UV;
            R ftb = sLp[i] * sTm[Y::tr(el, 1, j, v)] + sLm[i] * sTm[Y::tr(el, 3, P - j, v)];
            R gtb = sLm[j] * sTm[Y::tr(el, 0, i, v)] + sLp[j] * sTm[Y::tr(el, 2, P - i, v)];
#pragma unroll
            for (int m = 0; m < N1; ++m) {
                ftb -= sD[m * N1 + i] * rv[j * N1 + m];
                gtb -= sD[m * N1 + j] * rv[m * N1 + i];
            }
            Fxb[v] = jd * (jx * ftb + kx * gtb);
            Fyb[v] = jd * (jy * ftb + ky * gtb);
        }
        R ubq[NV] = {R(0), R(0), R(0), R(0)};
        euler_flux_vjp(uq, g, Fxb, Fyb, ubq);
#pragma unroll
        for (int v = 0; v < NV; ++v) qb[v] += double(ubq[v]);
    }
    } else {
    // ---- R2a: transposed divergence and Euler flux Jacobian at the solution points
    if (active) {
        const R* Rb = sFx + el * Y::UE;
        const R* Jm = sFy + el * Y::UE;
        R Fxb[NV], Fyb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R* rv = Rb + v * Y::UV;
            if (AFF) {
                // res = -(sum_m D[i][m] Ft[j][m] + sum_m D[j][m] Gt[m][i])
                R ftb = 0, gtb = 0;
#pragma unroll
                for (int m = 0; m < N1; ++m) {
                    ftb -= sD[m * N1 + i] * rv[j * N1 + m];   // xi line j, output point m
                    gtb -= sD[m * N1 + j] * rv[m * N1 + i];   // eta line i, output point m
                }
                Fxb[v] = jx * ftb + kx * gtb;
                Fyb[v] = jy * ftb + ky * gtb;
            } else {
                R fb = 0, gb = 0;
#pragma unroll
                for (int m = 0; m < N1; ++m) {
                    const int pxi = j * N1 + m, peta = m * N1 + i;
                    const R dxi = sD[m * N1 + i], deta = sD[m * N1 + j];
                    fb -= dxi * Jm[pxi] * rv[pxi] + deta * Jm[2 * Y::UV + peta] * rv[peta];
                    gb -= dxi * Jm[Y::UV + pxi] * rv[pxi] + deta * Jm[3 * Y::UV + peta] * rv[peta];
                }
                Fxb[v] = fb;
                Fyb[v] = gb;
            }
        }
        R ubq[NV] = {R(0), R(0), R(0), R(0)};
        euler_flux_vjp(uq, g, Fxb, Fyb, ubq);
#pragma unroll
        for (int v = 0; v < NV; ++v) qb[v] += double(ubq[v]);
    }

    // ---- R2b: transposed correction, flux difference, Riemann and normal flux
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        R ownb[NV] = {R(0), R(0), R(0), R(0)}, mirb[NV] = {R(0), R(0), R(0), R(0)};
        if (L.role != ROLE_NONE) {
            // Fd[k][f] enters R at the points with xi index f (edges 0, 2) or eta index f (1, 3)
            const R* Rb = sFx + el2 * Y::UE;
            R Fdb[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                R acc = 0;
#pragma unroll
                for (int m = 0; m < N1; ++m) {
                    const int p = (k & 1) ? f * N1 + m : m * N1 + f;
                    acc -= sC[k * S + p] * Rb[v * Y::UV + p];
                }
                Fdb[v] = acc;
            }
            const R nx = a.normals[((int64_t)e2 * 4 + k) * 2], ny = a.normals[((int64_t)e2 * 4 + k) * 2 + 1];
            const bool right = L.role == ROLE_RIGHT;
            const int nl = L.nbr - e0;
            const bool inside = L.role != ROLE_BOUNDARY && nl >= 0 && nl < EPB;
            R own[NV], mirr[NV], other[NV], UL[NV], UR[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) { own[v] = sTr[Y::tr(el2, k, f, v)]; mirr[v] = sTr[Y::tr(el2, k, P - f, v)]; }
            const double* bcd = a.bc_data + (int64_t)(L.role == ROLE_BOUNDARY ? L.bc_idx : 0) * 8;
            if (L.role == ROLE_BOUNDARY) {
                ghost_state(own, nx, ny, L.nbr_edge, bcd, g, other);
            } else if (inside) {
#pragma unroll
                for (int v = 0; v < NV; ++v) other[v] = sTr[Y::tr(nl, L.nbr_edge, right ? f : P - f, v)];
            } else {
                load4(a.T_in + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + (right ? f : P - f)) * NV, other);
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) { UL[v] = right ? other[v] : own[v]; UR[v] = right ? mirr[v] : other[v]; }
            // Fd = right ? Fn - Fc : Fc - Fn
            R Fcb[NV], Fnb_x[NV], Fnb_y[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const R fnb = right ? Fdb[v] : -Fdb[v];
                Fcb[v] = -fnb;
                Fnb_x[v] = fnb * nx;
                Fnb_y[v] = fnb * ny;
            }
            euler_flux_vjp(own, g, Fnb_x, Fnb_y, ownb);
            R ULb[NV] = {R(0), R(0), R(0), R(0)}, URb[NV] = {R(0), R(0), R(0), R(0)};
            rusanov_vjp(UL, UR, nx, ny, g, Fcb, ULb, URb);
            R othb[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if (right) { othb[v] = ULb[v]; mirb[v] = URb[v]; }
                else { ownb[v] += ULb[v]; othb[v] = URb[v]; }
            }
            if (L.role == ROLE_BOUNDARY) {
                ghost_state_vjp(own, nx, ny, L.nbr_edge, bcd, g, othb, ownb);
            } else if (inside) {
#pragma unroll
                for (int v = 0; v < NV; ++v) sTb[Y::tr(nl, L.nbr_edge, right ? f : P - f, v)] = othb[v];
            } else {
                store4(a.Tb_out + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + (right ? f : P - f)) * NV, othb);
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sFd[Y::tr(el2, k, f, v)] = ownb[v];      // sFd was last read in forward C
            sTm[Y::tr(el2, k, P - f, v)] = mirb[v];  // one writer per slot
        }
    }
    }  // textbook / reference reverse phases
    __syncthreads();

    // ---- R3: transposed trace operator; cotangent of q
    if (active) {
        const bool in0 = sIn[el * 4 + 0], in1 = sIn[el * 4 + 1], in2 = sIn[el * 4 + 2], in3 = sIn[el * 4 + 3];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int s0 = Y::tr(el, 0, i, v), s1 = Y::tr(el, 1, j, v);
            const int s2 = Y::tr(el, 2, P - i, v), s3 = Y::tr(el, 3, P - j, v);
            // (textbook: sTm holds the flux part, already taken up in R2a)
            const R t0 = sFd[s0] + (TB ? R(0) : sTm[s0]) + (in0 ? sTb[s0] : R(0));
            const R t1 = sFd[s1] + (TB ? R(0) : sTm[s1]) + (in1 ? sTb[s1] : R(0));
            const R t2 = sFd[s2] + (TB ? R(0) : sTm[s2]) + (in2 ? sTb[s2] : R(0));
            const R t3 = sFd[s3] + (TB ? R(0) : sTm[s3]) + (in3 ? sTb[s3] : R(0));
            const R tsum = sLm[j] * t0 + sLp[j] * t2 + sLp[i] * t1 + sLm[i] * t3;
            double out = qb[v] + double(tsum);
            if (MODE == MODE_STAGE1) out += a.U0b[v * PS + gi];
            a.Ub_out[v * PS + gi] = out;
        }
    }
}

template <typename R, int P, int EPB, int MODE, bool AFF, bool TB>
__global__ __launch_bounds__(BLOCK) void vjp_kernel(const VjpArgs<R> a, const OpTables<R> ops) {
    vjp_body<R, P, EPB, MODE, AFF, TB>(a, ops, tile_of_block(a.xcd_groups));
}

// The stage VJP over the tiles of several meshes of one order, metric form
// and scheme in one launch: one work item (member, tile) per workgroup, each
// member with its own arguments and operator tables.
template <typename R, int P, int EPB, int MODE, bool AFF, bool TB>
__global__ __launch_bounds__(BLOCK) void batched_vjp_kernel(const BatchVjpInst<R>* __restrict__ insts,
                                                            const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    vjp_body<R, P, EPB, MODE, AFF, TB>(insts[it.x].a, insts[it.x].ops, it.y);
}

// Completion of a cotangent: U_bar + transposed traces of the cross-tile edge
// cotangents that the last adjoint launch left for this tile.
template <typename R, int P, int EPB>
__global__ __launch_bounds__(BLOCK) void complete_kernel(const CompleteArgs<R> a, const OpTables<R> ops) {
    using Y = Lay<R, P>;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1, NT = EPB * S, NF = EPB * 4 * FP;
    __shared__ R sTb[EPB * Y::TE];
    __shared__ R sLm[N1], sLp[N1];
    const int tid = threadIdx.x;
    const int e0 = tile_of_block(a.xcd_groups) * EPB;
    const int64_t PS = (int64_t)a.E * S;
    if (tid < N1) { sLm[tid] = ops.Lm[tid]; sLp[tid] = ops.Lp[tid]; }
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        const int nl = L.nbr - e0;
        const bool cross = (L.role == ROLE_LEFT || L.role == ROLE_RIGHT) && !(nl >= 0 && nl < EPB);
        R tb[NV] = {R(0), R(0), R(0), R(0)};
        if (cross && a.Tb) load4(a.Tb + (((int64_t)e2 * 4 + k) * FP + f) * NV, tb);
#pragma unroll
        for (int v = 0; v < NV; ++v) sTb[Y::tr(el2, k, f, v)] = tb[v];
    }
    __syncthreads();
    const int el = tid / S, sp = tid - (tid / S) * S;
    const int e = e0 + el;
    if (tid >= NT || e >= a.E) return;
    const int64_t gi = (int64_t)e * S + sp;
    const int i = sp % N1, j = sp / N1;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const R tsum = sLm[j] * sTb[Y::tr(el, 0, i, v)] + sLp[j] * sTb[Y::tr(el, 2, P - i, v)] +
                       sLp[i] * sTb[Y::tr(el, 1, j, v)] + sLm[i] * sTb[Y::tr(el, 3, P - j, v)];
        a.out[v * PS + gi] = a.Ub[v * PS + gi] + double(tsum);
    }
}

template <int P>
constexpr int adj_tile_max() {
    // the adjoint tile keeps its LDS within 64 KiB at every order
    return P <= 2 ? (BLOCK / ((P + 1) * (P + 1))) / 2 : BLOCK / ((P + 1) * (P + 1));
}

template <typename R, int P, bool AFF, bool TB>
static hipError_t vjp_dispatch(int mode, const VjpArgs<R>& a, const OpTables<R>& ops, hipStream_t s) {
    constexpr int T = adj_tile_max<P>();
    const dim3 grid(num_blocks(a.E, T)), block(BLOCK);
    switch (mode) {
    case MODE_RESIDUAL: vjp_kernel<R, P, T, MODE_RESIDUAL, AFF, TB><<<grid, block, 0, s>>>(a, ops); break;
    case MODE_STAGE1: vjp_kernel<R, P, T, MODE_STAGE1, AFF, TB><<<grid, block, 0, s>>>(a, ops); break;
    case MODE_STAGE2: vjp_kernel<R, P, T, MODE_STAGE2, AFF, TB><<<grid, block, 0, s>>>(a, ops); break;
    case MODE_STAGE3: vjp_kernel<R, P, T, MODE_STAGE3, AFF, TB><<<grid, block, 0, s>>>(a, ops); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, int P>
static hipError_t complete_dispatch(const CompleteArgs<R>& a, const OpTables<R>& ops, hipStream_t s) {
    constexpr int T = adj_tile_max<P>();
    complete_kernel<R, P, T><<<dim3(num_blocks(a.E, T)), dim3(BLOCK), 0, s>>>(a, ops);
    return hipGetLastError();
}

}  // namespace

int adjoint_tile(int P) {
    switch (P) {
    case 1: return adj_tile_max<1>();
    case 2: return adj_tile_max<2>();
    case 3: return adj_tile_max<3>();
    case 4: return adj_tile_max<4>();
    case 5: return adj_tile_max<5>();
    default: return 0;
    }
}

template <typename R>
hipError_t launch_vjp(int P, int mode, const VjpArgs<R>& a, const OpTables<R>& ops, bool affine, hipStream_t s,
                      bool textbook) {
    if (a.E <= 0) return hipSuccess;
#define ZJ_VJP_P(PP)                                                                                      \
    case PP:                                                                                              \
        if (textbook)                                                                                     \
            return affine ? vjp_dispatch<R, PP, true, true>(mode, a, ops, s)                              \
                          : vjp_dispatch<R, PP, false, true>(mode, a, ops, s);                            \
        return affine ? vjp_dispatch<R, PP, true, false>(mode, a, ops, s)                                 \
                      : vjp_dispatch<R, PP, false, false>(mode, a, ops, s);
    switch (P) {
        ZJ_VJP_P(1) ZJ_VJP_P(2) ZJ_VJP_P(3) ZJ_VJP_P(4) ZJ_VJP_P(5)
    default: return hipErrorInvalidValue;
    }
#undef ZJ_VJP_P
}

template <typename R>
hipError_t launch_batched_vjp(int P, bool affine, bool textbook, int mode, const BatchVjpInst<R>* insts,
                              const int2* items, int n_items, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
#define ZJ_BV_M(PP, AF, TBK)                                                                                    \
    {                                                                                                           \
        constexpr int T = adj_tile_max<PP>();                                                                   \
        switch (mode) {                                                                                         \
        case MODE_STAGE1: batched_vjp_kernel<R, PP, T, MODE_STAGE1, AF, TBK><<<n_items, BLOCK, 0, s>>>(insts, items); break; \
        case MODE_STAGE2: batched_vjp_kernel<R, PP, T, MODE_STAGE2, AF, TBK><<<n_items, BLOCK, 0, s>>>(insts, items); break; \
        case MODE_STAGE3: batched_vjp_kernel<R, PP, T, MODE_STAGE3, AF, TBK><<<n_items, BLOCK, 0, s>>>(insts, items); break; \
        default: return hipErrorInvalidValue;                                                                   \
        }                                                                                                       \
        return hipGetLastError();                                                                               \
    }
#define ZJ_BV_P(PP)                                                                                             \
    case PP:                                                                                                    \
        if (textbook) {                                                                                         \
            if (affine) ZJ_BV_M(PP, true, true) else ZJ_BV_M(PP, false, true)                                   \
        }                                                                                                       \
        if (affine) ZJ_BV_M(PP, true, false) else ZJ_BV_M(PP, false, false)
    switch (P) {
        ZJ_BV_P(1) ZJ_BV_P(2) ZJ_BV_P(3) ZJ_BV_P(4) ZJ_BV_P(5)
    default: return hipErrorInvalidValue;
    }
#undef ZJ_BV_P
#undef ZJ_BV_M
}

template <typename R>
hipError_t launch_complete(int P, const CompleteArgs<R>& a, const OpTables<R>& ops, hipStream_t s) {
    if (a.E <= 0) return hipSuccess;
    switch (P) {
    case 1: return complete_dispatch<R, 1>(a, ops, s);
    case 2: return complete_dispatch<R, 2>(a, ops, s);
    case 3: return complete_dispatch<R, 3>(a, ops, s);
    case 4: return complete_dispatch<R, 4>(a, ops, s);
    case 5: return complete_dispatch<R, 5>(a, ops, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_vjp<double>(int, int, const VjpArgs<double>&, const OpTables<double>&, bool, hipStream_t,
                                      bool);
template hipError_t launch_vjp<float>(int, int, const VjpArgs<float>&, const OpTables<float>&, bool, hipStream_t,
                                     bool);
template hipError_t launch_batched_vjp<double>(int, bool, bool, int, const BatchVjpInst<double>*, const int2*, int,
                                              hipStream_t);
template hipError_t launch_batched_vjp<float>(int, bool, bool, int, const BatchVjpInst<float>*, const int2*, int,
                                             hipStream_t);
template hipError_t launch_complete<double>(int, const CompleteArgs<double>&, const OpTables<double>&, hipStream_t);
template hipError_t launch_complete<float>(int, const CompleteArgs<float>&, const OpTables<float>&, hipStream_t);

}  // namespace zj
