// zj_kernels.hip -- fused flux-reconstruction stage kernels for gfx950.
//
// One launch per SSP-RK3 stage does the whole residual and the stage update
// of a tile of elements (the reference needs 7 kernels, a memset and two D2D
// copies per stage, kernels.cu:30-1349):
//
//   A  load U_in (SoA, coalesced), Euler flux at every solution point
//      [K4 computeInviscidFluxSPKernel, kernels.cu:157-194]       -> LDS
//   B1 the element's own flux-point states, sum-factorised 1D
//      interpolation from LDS [K1 interpolateToFluxPointsKernel, :30-67]
//   B2 per (edge, flux point): common flux against the neighbour's trace
//      written by the previous stage, or a ghost state on a boundary, then the
//      flux difference of this element's side
//      [K9 computeRiemannFluxWithBCKernel :453-676, K8 computeFluxDiffForFRKernel
//       :346-426]; each face's common flux is evaluated by both of its elements
//      with identical inputs, so no face buffer goes through HBM
//   C  divergence with the 1D derivative matrix from LDS [K11 :813-859],
//      FR correction [K13 :948-1021], SSP-RK3 combination [K16-18 :1297-1349]
//      and the positivity limiter [K15 :1236-1288] in FP64; U_out written
//   D  trace of U_out for the next stage (written once, read once by neighbours)
//
// Every reference quirk on this path is kept: the clamped inline Rusanov that
// ignores the configured solver, the NaN/Inf scrub, the right element pairing
// F_common[fp] with its own slot fp, the correction indexed by the unmirrored
// xi index on edges 0/2 and eta index on edges 1/3 without metric, and the
// limiter's SI constants with gamma = 1.4.
#include "zj_kernels.hpp"

#include "zj_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <type_traits>

namespace zj {

// Elements per workgroup of the full tile at order P.
#define ZJ_EPB(P) (BLOCK / (((P) + 1) * ((P) + 1)))

namespace {

// LDS of one element tile of the stage kernel, as a policy of accessors.
// SepSmem gives every array its own LDS object (distinct objects never alias,
// which the scheduler exploits); the batched kernel carves the same arrays
// out of one buffer (BatchSmem, below).
template <typename R, int P, int EPB>
struct SepSmem {
    using Y = Lay<R, P>;
    static constexpr int N1 = P + 1, S = N1 * N1;
    __device__ static R* U() { __shared__ R a[EPB * Y::UE]; return a; }
    __device__ static R* Fx() { __shared__ R a[EPB * Y::UE]; return a; }
    __device__ static R* Fy() { __shared__ R a[EPB * Y::UE]; return a; }
    __device__ static R* Tr() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* Fd() { __shared__ R a[EPB * Y::TE]; return a; }
    __device__ static R* D() { __shared__ R a[N1 * N1]; return a; }
    __device__ static R* Lm() { __shared__ R a[N1]; return a; }
    __device__ static R* Lp() { __shared__ R a[N1]; return a; }
    __device__ static R* C() { __shared__ R a[4 * S]; return a; }
    __device__ static double* Red() { __shared__ double a[BLOCK / 64]; return a; }
    __device__ static unsigned char* Out() { __shared__ unsigned char a[EPB * 4]; return a; }
};

// One stage over tile `tile` (elements tile*EPB ...) as thread `tid`, shared
// by the single-mesh kernel and the batched persistent kernel.
template <typename R, int P, int EPB, int MODE, bool AFF, bool TB = false, class SM>
__device__ __forceinline__ void stage_body(const StageArgs<R>& a, const OpTables<R>& ops, int tile, const SM& sm,
                                           int tid) {
    using Y = Lay<R, P>;
    // FP32-bulk keeps the intermediate stage state as FP32 increments over U^n
    constexpr bool INC = std::is_same<R, float>::value;
    constexpr bool HAS_U0 = MODE == MODE_STAGE2 || MODE == MODE_STAGE3;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1;
    constexpr int NT = EPB * S, NF = EPB * 4 * FP;
    static_assert(NT <= BLOCK, "tile exceeds the workgroup");
    R* const sU = sm.U();
    R* const sFx = sm.Fx();
    R* const sFy = sm.Fy();
    R* const sTr = sm.Tr();
    R* const sFd = sm.Fd();
    R* const sD = sm.D();
    R* const sLm = sm.Lm();
    R* const sLp = sm.Lp();
    R* const sC = sm.C();
    double* const sRed = sm.Red();
    unsigned char* const sOut = sm.Out();  // edge trace needed outside this tile

    const int e0 = tile * EPB;
    const int64_t PS = (int64_t)a.E * S;
    const R g = R(a.gamma);

    for (int t = tid; t < N1 * N1; t += BLOCK) sD[t] = ops.D[t];
    if (TB) {  // textbook: g_L', g_R' at the 1D points in place of the reference's corr table
        if (tid < N1) { sC[tid] = ops.gL[tid]; sC[N1 + tid] = ops.gR[tid]; }
    } else {
        for (int t = tid; t < 4 * S; t += BLOCK) sC[t] = ops.corr[t];
    }
    if (tid < N1) { sLm[tid] = ops.Lm[tid]; sLp[tid] = ops.Lp[tid]; }

    const int el = tid / S, sp = tid - (tid / S) * S;
    const int e = e0 + el;
    const bool active = tid < NT && e < a.E;
    const int64_t gi = (int64_t)e * S + sp;
    const int ub = el * Y::UE + sp;  // this thread's point in the solution-point arrays

    // ---- A: state and volume flux
    // Affine elements (constant metric): the divergence regroups into the
    // derivatives of the contravariant fluxes xi_x F + xi_y G and eta_x F + eta_y G,
    // with the metric read once per element.
    double q[NV], u0[NV];
    R jx = R(0), jy = R(0), kx = R(0), ky = R(0);
    if (active) {
        R jd = R(1);  // textbook, non-affine: transformed fluxes carry J
        if (AFF) {
            jx = a.Jinv[(int64_t)e * 4]; jy = a.Jinv[(int64_t)e * 4 + 1];
            kx = a.Jinv[(int64_t)e * 4 + 2]; ky = a.Jinv[(int64_t)e * 4 + 3];
        } else if (TB) {
            jx = a.Jinv[gi]; jy = a.Jinv[PS + gi]; kx = a.Jinv[2 * PS + gi]; ky = a.Jinv[3 * PS + gi];
            jd = a.Jpt[gi];
        }
        R u[NV], fx[NV], fy[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (INC && HAS_U0) u0[v] = a.U0[v * PS + gi];
            if (INC && HAS_U0) q[v] = u0[v] + double(a.dW_in[v * PS + gi]);
            else q[v] = a.U_in[v * PS + gi];
            u[v] = R(q[v]);
        }
        euler_flux(u, g, fx, fy);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sU[ub + v * Y::UV] = u[v];
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
    __syncthreads();

    // ---- B1: own traces
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

    // ---- B2: common flux and flux difference per (edge, flux point)
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const EdgeLink L = a.links[(int64_t)e2 * 4 + k];
        R Fd[NV] = {R(0), R(0), R(0), R(0)};
        if (f == 0) sOut[el2 * 4 + k] = 0;
        if (L.role != ROLE_NONE) {
            const R nx = a.normals[((int64_t)e2 * 4 + k) * 2], ny = a.normals[((int64_t)e2 * 4 + k) * 2 + 1];
            const bool right = L.role == ROLE_RIGHT;
            R own[NV], mirr[NV], other[NV], UL[NV], UR[NV], Fc[NV], Fn[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) { own[v] = sTr[Y::tr(el2, k, f, v)]; mirr[v] = sTr[Y::tr(el2, k, P - f, v)]; }
            // The face's common flux as its left element computes it: U_L = left
            // slot f, U_R = right slot Fp-1-f (kernels.cu:577).  One call site
            // serves both roles so a wave never runs the flux twice.
            // A neighbour inside this tile has its traces in sTr already (the
            // same arithmetic on the same state as the copy in T_in), so only
            // edges that cross tiles travel through T_in/T_out.
            const int nl = L.nbr - e0;
            const bool inside = L.role != ROLE_BOUNDARY && nl >= 0 && nl < EPB;
            if (f == 0) sOut[el2 * 4 + k] = L.role != ROLE_BOUNDARY && !inside;
            // reference: the neighbour's slot f on the right role (kernels.cu:577 quirk);
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
                // one common flux per face point with the left element's normal,
                // outward on each side; Delta = sign |e|/2 (F.n) [/J] - interpolated transformed flux
#pragma unroll
                for (int v = 0; v < NV; ++v) { UL[v] = right ? other[v] : own[v]; UR[v] = right ? own[v] : other[v]; }
                riemann_by_kind(a.riemann, UL, UR, nx, ny, g, Fc);
                const R ef = a.edge_f[(int64_t)e2 * 4 + k];
                const TraceLine tl = trace_line<P>(k, f);
                const R* Fbase = ((k & 1) ? sFx : sFy) + el2 * Y::UE;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const R fdisc = trace_point<R, P>(Fbase + v * Y::UV, tl, sLm, sLp);
                    Fd[v] = ef * (right ? -Fc[v] : Fc[v]) - fdisc;
                }
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) { UL[v] = right ? other[v] : own[v]; UR[v] = right ? mirr[v] : other[v]; }
                rusanov(UL, UR, nx, ny, g, Fc);
                normal_flux(own, g, nx, ny, Fn);
                // left: F_com - F(U).n; right: -F_com + F(U).n at its own slot f (kernels.cu:397,423)
#pragma unroll
                for (int v = 0; v < NV; ++v) Fd[v] = right ? Fn[v] - Fc[v] : Fc[v] - Fn[v];
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sFd[Y::tr(el2, k, f, v)] = Fd[v];
    }
    __syncthreads();

    // ---- C: divergence, correction, stage update
    double rsq = 0.0;
    if (active) {
        const int i = sp % N1, j = sp / N1;  // xi index, eta index
        if (!AFF && !TB) { jx = a.Jinv[gi]; jy = a.Jinv[PS + gi]; kx = a.Jinv[2 * PS + gi]; ky = a.Jinv[3 * PS + gi]; }
        const R inv_j = (TB && !AFF) ? rcpR(a.Jpt[gi]) : R(1);
        double res[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R* Fx = sFx + el * Y::UE + v * Y::UV;
            const R* Fy = sFy + el * Y::UE + v * Y::UV;
            R div;
            if (AFF || TB) {  // Fx, Fy hold the contravariant (transformed) fluxes
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
            R c;
            if (TB) {  // xi line j: edges 3 (slot p-j) and 1 (slot j); eta line i: edges 0 (i), 2 (p-i)
                c = sFd[Y::tr(el, 3, P - j, v)] * sC[i] + sFd[Y::tr(el, 1, j, v)] * sC[N1 + i] +
                    sFd[Y::tr(el, 0, i, v)] * sC[j] + sFd[Y::tr(el, 2, P - i, v)] * sC[N1 + j];
            } else {
                c = sFd[Y::tr(el, 0, i, v)] * sC[sp] + sFd[Y::tr(el, 1, j, v)] * sC[S + sp] +
                    sFd[Y::tr(el, 2, i, v)] * sC[2 * S + sp] + sFd[Y::tr(el, 3, j, v)] * sC[3 * S + sp];
            }
            res[v] = TB ? double(-(div + c) * inv_j) : double(-div - c);
            rsq += res[v] * res[v];
        }
        if (MODE == MODE_RESIDUAL) {
#pragma unroll
            for (int v = 0; v < NV; ++v) a.R_out[v * PS + gi] = res[v];
        } else {
            const double dt = *a.dt_ptr;
            double un[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if (!INC && HAS_U0) u0[v] = a.U0[v * PS + gi];
                if (MODE == MODE_STAGE1) un[v] = q[v] + dt * res[v];
                else if (MODE == MODE_STAGE2) un[v] = 0.75 * u0[v] + 0.25 * q[v] + 0.25 * dt * res[v];
                else un[v] = (1.0 / 3.0) * u0[v] + (2.0 / 3.0) * q[v] + (2.0 / 3.0) * dt * res[v];
            }
            if (a.positivity) positivity<INC>(un);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double next = un[v];  // the state the next stage will read back
                if (INC && MODE != MODE_STAGE3) {
                    const double base = MODE == MODE_STAGE1 ? q[v] : u0[v];
                    const float dw = float(un[v] - base);
                    a.dW_out[v * PS + gi] = dw;
                    next = base + double(dw);
                } else {
                    a.U_out[v * PS + gi] = un[v];
                }
                sU[ub + v * Y::UV] = R(next);  // safe: sU was last read before the B1->B2 barrier
            }
        }
    }
    if (a.l2_partial) {  // one partial sum of R^2 per workgroup; reduced on request
        const double s = block_reduce<OpSum>(rsq, sRed);
        if (tid == 0) a.l2_partial[e0 / EPB] = s;  // tile order: the sum does not depend on the remap
    }
    if (MODE == MODE_RESIDUAL) return;
    __syncthreads();

    // ---- D: traces of the new state, for edges read by another tile
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E || !sOut[el2 * 4 + k]) continue;
        const TraceLine tl = trace_line<P>(k, f);
        R tr[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) tr[v] = trace_point<R, P>(sU + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
        store4(a.T_out + (((int64_t)e2 * 4 + k) * FP + f) * NV, tr);
    }
}

template <typename R, int P, int EPB, int MODE, bool AFF, bool TB = false>
__global__ __launch_bounds__(BLOCK) void stage_kernel(const StageArgs<R> a, const OpTables<R> ops) {
    stage_body<R, P, EPB, MODE, AFF, TB>(a, ops, tile_of_block(a.xcd_groups), SepSmem<R, P, EPB>{}, threadIdx.x);
}

// ------------------------------------------------------------ batched stage
// The same arrays carved out of one LDS buffer.
template <typename R, int P, int EPB>
struct BatchSmem {
    using Y = Lay<R, P>;
    static constexpr int N1 = P + 1, S = N1 * N1;
    static constexpr int oU = 0, oFx = oU + EPB * Y::UE, oFy = oFx + EPB * Y::UE, oTr = oFy + EPB * Y::UE;
    static constexpr int oFd = oTr + EPB * Y::TE, oD = oFd + EPB * Y::TE, oLm = oD + N1 * N1, oLp = oLm + N1;
    static constexpr int oC = oLp + N1, nR = oC + 4 * S;
    static constexpr int red_byte = ((nR * (int)sizeof(R) + 15) / 16) * 16;
    static constexpr int bytes = red_byte + (BLOCK / 64) * 8 + EPB * 4;
    unsigned char* base;
    __device__ R* U() const { return reinterpret_cast<R*>(base) + oU; }
    __device__ R* Fx() const { return reinterpret_cast<R*>(base) + oFx; }
    __device__ R* Fy() const { return reinterpret_cast<R*>(base) + oFy; }
    __device__ R* Tr() const { return reinterpret_cast<R*>(base) + oTr; }
    __device__ R* Fd() const { return reinterpret_cast<R*>(base) + oFd; }
    __device__ R* D() const { return reinterpret_cast<R*>(base) + oD; }
    __device__ R* Lm() const { return reinterpret_cast<R*>(base) + oLm; }
    __device__ R* Lp() const { return reinterpret_cast<R*>(base) + oLp; }
    __device__ R* C() const { return reinterpret_cast<R*>(base) + oC; }
    __device__ double* Red() const { return reinterpret_cast<double*>(base + red_byte); }
    __device__ unsigned char* Out() const { return base + red_byte + (BLOCK / 64) * 8; }
};

// One launch per (stage, order, metric form) for many meshes: workgroup b
// takes work items b, b + grid, ... (instance, tile), items ordered by
// instance size so that the static round-robin stays balanced.  Every
// instance uses its order's full tile, the tiling its trace buffers were
// written with.  The launches of one stage are independent branches of the
// step graph, so they share the GPU.
template <typename R, int P, int MODE, bool AFF, bool TB>
__global__ __launch_bounds__(BLOCK) void batched_stage_kernel(const BatchInst<R>* __restrict__ insts,
                                                              const int2* __restrict__ items, int n_items) {
    constexpr int T = BLOCK / ((P + 1) * (P + 1));
    using SM = BatchSmem<R, P, T>;
    __shared__ __align__(16) unsigned char raw[SM::bytes];
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const int2 it = items[w];
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));  // keeps the per-thread index math in the loop body (registers)
        stage_body<R, P, T, MODE, AFF, TB>(insts[it.x].args, insts[it.x].ops, it.y, SM{raw}, tid);
        __syncthreads();  // the next item reuses the LDS
    }
}

// One launch per stage for a whole batch, every order and metric form: work
// item w = (instance, tile) in a single list; the instance's order and metric
// form select the stage body, and the LDS buffer is sized for the largest tile
// layout among the instances (dynamic shared memory).
template <typename R, int MODE>
__global__ __launch_bounds__(BLOCK) void batched_all_kernel(const BatchInst<R>* __restrict__ insts,
                                                            const int2* __restrict__ items, int n_items) {
    extern __shared__ __align__(16) unsigned char raw_all[];
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const int2 it = items[w];
        const BatchInst<R>& in = insts[it.x];
        const int kind = __builtin_amdgcn_readfirstlane(in.P * 2 + in.affine);
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
#define ZJ_BALL(PP)                                                                                          \
    case 2 * PP:                                                                                             \
        stage_body<R, PP, ZJ_EPB(PP), MODE, false>(in.args, in.ops, it.y, BatchSmem<R, PP, ZJ_EPB(PP)>{raw_all}, tid); \
        break;                                                                                               \
    case 2 * PP + 1:                                                                                         \
        stage_body<R, PP, ZJ_EPB(PP), MODE, true>(in.args, in.ops, it.y, BatchSmem<R, PP, ZJ_EPB(PP)>{raw_all}, tid); \
        break;
        switch (kind) {
            ZJ_BALL(1) ZJ_BALL(2) ZJ_BALL(3) ZJ_BALL(4) ZJ_BALL(5)
        default: break;
        }
#undef ZJ_BALL
        __syncthreads();  // the next item reuses the LDS
    }
}

template <typename R, int P, int EPB>
__global__ __launch_bounds__(BLOCK) void trace_kernel(const TraceArgs<R> a, const OpTables<R> ops) {
    using Y = Lay<R, P>;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1, NT = EPB * S, NF = EPB * 4 * FP;
    __shared__ R sU[EPB * Y::UE];
    __shared__ R sLm[N1], sLp[N1];
    const int tid = threadIdx.x, e0 = tile_of_block(a.xcd_groups) * EPB;
    const int64_t PS = (int64_t)a.E * S;
    if (tid < N1) { sLm[tid] = ops.Lm[tid]; sLp[tid] = ops.Lp[tid]; }
    const int el = tid / S, sp = tid - (tid / S) * S;
    if (tid < NT && e0 + el < a.E) {
        const int64_t gi = (int64_t)(e0 + el) * S + sp;
#pragma unroll
        for (int v = 0; v < NV; ++v) sU[el * Y::UE + v * Y::UV + sp] = R(a.U[v * PS + gi]);
    }
    __syncthreads();
    for (int t = tid; t < NF; t += BLOCK) {
        const int el2 = t / (4 * FP), r = t - el2 * (4 * FP), k = r / FP, f = r - k * FP;
        const int e2 = e0 + el2;
        if (e2 >= a.E) continue;
        const TraceLine tl = trace_line<P>(k, f);
        R tr[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) tr[v] = trace_point<R, P>(sU + el2 * Y::UE + v * Y::UV, tl, sLm, sLp);
        store4(a.T_out + (((int64_t)e2 * 4 + k) * FP + f) * NV, tr);
    }
}

// AoS [E][S][4] in the caller's element order <-> SoA planes in the device
// order; perm[e] = caller index of device element e (nullptr: identity).
__global__ void aos_to_soa_kernel(const double* __restrict__ aos, double* __restrict__ soa, int64_t n,
                                  const int* __restrict__ perm, int S) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t src = perm ? (int64_t)perm[i / S] * S + i % S : i;
    double v[NV];
    load4(aos + 4 * src, v);
#pragma unroll
    for (int k = 0; k < NV; ++k) soa[k * n + i] = v[k];
}

__global__ void soa_to_aos_kernel(const double* __restrict__ soa, double* __restrict__ aos, int64_t n,
                                  const int* __restrict__ perm, int S) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t dst = perm ? (int64_t)perm[i / S] * S + i % S : i;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = soa[k * n + i];
    store4(aos + 4 * dst, v);
}

// gpu::rkStage1/2/3 arithmetic (kernels.cu:1297-1349) on arbitrary arrays.
__global__ void rk3_combine_kernel(int stage, double* out, const double* U0, const double* Uk,
                                   const double* R, int64_t n, double dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (stage == 1) out[i] = U0[i] + dt * R[i];
    else if (stage == 2) out[i] = 0.75 * U0[i] + 0.25 * Uk[i] + 0.25 * dt * R[i];
    else out[i] = (1.0 / 3.0) * U0[i] + (2.0 / 3.0) * Uk[i] + (2.0 / 3.0) * dt * R[i];
}

__global__ void euler_update_kernel(double* U, const double* R, int64_t n, double dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) U[i] = U[i] + dt * R[i];
}

__global__ void set_scalar_kernel(double* p, double v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void set_many_kernel(double* dst, ScalarChunk c) {
    if (threadIdx.x < c.n) dst[threadIdx.x] = c.v[threadIdx.x];
}

// FRSolver::computeTimeStep at one point (fr_solver.cpp:799-811,
// euler.hpp:34-41,166-171); with the viscous terms their rate
// (p+1)^4 nu_k / h^2 adds to the convective one (visc > 0).
__device__ __forceinline__ double cfl_dt_point(double rho, double ru, double rv, double rE, double h, double gamma,
                                               double cfl, double cf, double visc) {
#pragma clang fp contract(off)
    const double u = ru / rho, v = rv / rho;
    const double p = (gamma - 1.0) * rho * (rE / rho - 0.5 * (u * u + v * v));
    const double ws = sqrt(u * u + v * v) + sqrt(gamma * fmax(p, 1e-10) / fmax(rho, 1e-10));
    return visc == 0.0 ? cfl * h / (ws * cf) : cfl / (cf * ws / h + visc / (fmax(rho, 1e-10) * h * h));
}

__global__ __launch_bounds__(BLOCK) void dt_partial_kernel(const double* __restrict__ U,
                                                           const double* __restrict__ hcell,
                                                           int E, int S, double gamma, double cfl,
                                                           double cf, double visc, double* __restrict__ partial) {
    __shared__ double sRed[BLOCK / 64];
    const int64_t n = (int64_t)E * S;
    double m = 1e20;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const double d = cfl_dt_point(U[i], U[n + i], U[2 * n + i], U[3 * n + i], hcell[i / S], gamma, cfl, cf, visc);
        m = d < m ? d : m;
    }
    const double r = block_reduce<OpMin>(m, sRed);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ void fill_u64_kernel(unsigned long long* p, unsigned long long v) {
    if (threadIdx.x == 0) *p = v;
}

// The point whose CFL step is the minimum *dtmin (the lowest flat index among
// ties): the minimum's derivative is that point's.
__global__ __launch_bounds__(BLOCK) void dt_argmin_kernel(const double* __restrict__ U,
                                                          const double* __restrict__ hcell, int E, int S,
                                                          double gamma, double cfl, double cf, double visc,
                                                          const double* __restrict__ dtmin,
                                                          unsigned long long* __restrict__ idx) {
    const int64_t n = (int64_t)E * S;
    const double m = *dtmin;
    unsigned long long best = ~0ull;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const double d = cfl_dt_point(U[i], U[n + i], U[2 * n + i], U[3 * n + i], hcell[i / S], gamma, cfl, cf, visc);
        if (d == m && (unsigned long long)i < best) best = (unsigned long long)i;
    }
    if (best != ~0ull) atomicMin(idx, best);
}

// Ub += dt_bar * d(dt)/dU at the minimising point *idx: the derivative of
// cfl_dt_point through u, v, p, the sound speed and (viscous) the density
// floor, each max() taking the branch the forward took.
__global__ void dt_cfl_vjp_kernel(const double* __restrict__ U, const double* __restrict__ hcell, int E, int S,
                                  double gamma, double cfl, double cf, double visc,
                                  const unsigned long long* __restrict__ idx, double dt_bar, double* __restrict__ Ub) {
    if (threadIdx.x != 0) return;
    const unsigned long long i = *idx;
    const int64_t n = (int64_t)E * S;
    if (i >= (unsigned long long)n) return;
    const double rho = U[i], ru = U[n + i], rv = U[2 * n + i], rE = U[3 * n + i], h = hcell[i / S];
    const double gm = gamma - 1.0;
    const double u = ru / rho, v = rv / rho, q2 = u * u + v * v, sp = sqrt(q2);
    const double p = gm * rho * (rE / rho - 0.5 * q2);
    const double pp = fmax(p, 1e-10), rr = fmax(rho, 1e-10);
    const double c = sqrt(gamma * pp / rr), ws = sp + c;
    // speed |u|: d/du = u / |u| (0 at rest)
    const double su = sp > 0.0 ? u / sp : 0.0, sv = sp > 0.0 ? v / sp : 0.0;
    const double cp = p > 1e-10 ? gamma / (2.0 * c * rr) : 0.0;     // dc/dp
    const double cr = rho > 1e-10 ? -c / (2.0 * rr) : 0.0;           // dc/drho through the floor
    // p = (g-1)(rE - (ru^2 + rv^2) / (2 rho)); u = ru / rho, v = rv / rho
    const double ws_r = -(su * u + sv * v) / rho + cp * gm * 0.5 * q2 + cr;
    const double ws_ru = su / rho - cp * gm * u;
    const double ws_rv = sv / rho - cp * gm * v;
    const double ws_rE = cp * gm;
    double dws, drr = 0.0;
    if (visc == 0.0) {
        dws = -cfl * h / (ws * ws * cf);
    } else {
        const double D = cf * ws / h + visc / (rr * h * h);
        dws = -cfl * cf / (D * D * h);
        drr = rho > 1e-10 ? cfl * visc / (D * D * rr * rr * h * h) : 0.0;
    }
    Ub[i] += dt_bar * (dws * ws_r + drr);
    Ub[n + i] += dt_bar * dws * ws_ru;
    Ub[2 * n + i] += dt_bar * dws * ws_rv;
    Ub[3 * n + i] += dt_bar * dws * ws_rE;
}

__global__ __launch_bounds__(BLOCK) void nonfinite_kernel(const double* __restrict__ U, int64_t n,
                                                          unsigned long long* __restrict__ count) {
    __shared__ double sRed[BLOCK / 64];
    double c = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        c += isfinite(U[i]) ? 0.0 : 1.0;
    const double r = block_reduce<OpSum>(c, sRed);
    if (threadIdx.x == 0 && r > 0) atomicAdd(count, (unsigned long long)r);
}

// Kinetic energy sum_i w_i (m_x^2 + m_y^2) / (2 rho) of a state in SoA planes
// (device element order; the weights in the caller's order, through perm),
// one partial per workgroup, and its gradient into the planes G.
__global__ __launch_bounds__(BLOCK) void kinetic_energy_kernel(const double* __restrict__ U,
                                                               const double* __restrict__ w,
                                                               const int* __restrict__ perm, int E, int S,
                                                               double* __restrict__ G, double* __restrict__ partial) {
    __shared__ double sRed[BLOCK / 64];
    const int64_t n = (int64_t)E * S;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const int64_t e = i / S, sp = i - e * S;
        const double wi = w[(perm ? (int64_t)perm[e] : e) * S + sp];
        const double rho = U[i], mx = U[n + i], my = U[2 * n + i];
        const double u = mx / rho, v = my / rho;
        acc += wi * (0.5 * (mx * u + my * v));
        G[i] = -0.5 * wi * (u * u + v * v);
        G[n + i] = wi * u;
        G[2 * n + i] = wi * v;
        G[3 * n + i] = 0.0;
    }
    const double r = block_reduce<OpSum>(acc, sRed);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(BLOCK) void finalize_kernel(const double* __restrict__ partial, int n,
                                                         int op, double* __restrict__ out) {
    __shared__ double sRed[BLOCK / 64];
    const bool is_min = op == FIN_MIN;
    double acc = is_min ? 1e300 : 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK)
        acc = is_min ? fmin(acc, partial[i]) : acc + partial[i];
    const double r = is_min ? block_reduce<OpMin>(acc, sRed) : block_reduce<OpSum>(acc, sRed);
    if (threadIdx.x == 0) *out = op == FIN_SQRT_SUM ? sqrt(r) : r;
}

}  // namespace

// ------------------------------------------------------------------ launch
KernelTimer& kernel_timer() {
    static thread_local KernelTimer t;
    return t;
}

int max_epb(int P) { return BLOCK / ((P + 1) * (P + 1)); }
int default_epb(int P) { return max_epb(P); }
int num_blocks(int E, int epb) { return (E + epb - 1) / epb; }


template <typename R, int P, int EPB, bool AFF, bool TB>
static hipError_t stage_dispatch_mode(int mode, const StageArgs<R>& a, const OpTables<R>& ops, hipStream_t s) {
    const dim3 grid(num_blocks(a.E, EPB)), block(BLOCK);
    switch (mode) {
    case MODE_RESIDUAL: klaunch(stage_kernel<R, P, EPB, MODE_RESIDUAL, AFF, TB>, grid, block, 0, s, a, ops); break;
    case MODE_STAGE1: klaunch(stage_kernel<R, P, EPB, MODE_STAGE1, AFF, TB>, grid, block, 0, s, a, ops); break;
    case MODE_STAGE2: klaunch(stage_kernel<R, P, EPB, MODE_STAGE2, AFF, TB>, grid, block, 0, s, a, ops); break;
    case MODE_STAGE3: klaunch(stage_kernel<R, P, EPB, MODE_STAGE3, AFF, TB>, grid, block, 0, s, a, ops); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, int P, bool AFF>
static hipError_t stage_dispatch_epb(int mode, const StageArgs<R>& a, const OpTables<R>& ops, int epb, bool tb,
                                     hipStream_t s) {
    constexpr int M = ZJ_EPB(P);
    if (tb) {  // the textbook scheme runs at the full tile; P1 and P2 also at half (the adjoint's tile)
        if (epb == 0 || epb == M) return stage_dispatch_mode<R, P, M, AFF, true>(mode, a, ops, s);
        if constexpr (P <= 2) {
            if (epb == M / 2) return stage_dispatch_mode<R, P, M / 2, AFF, true>(mode, a, ops, s);
        }
        return hipErrorInvalidValue;
    }
    if (epb == 0 || epb == M) return stage_dispatch_mode<R, P, M, AFF, false>(mode, a, ops, s);
    if (M / 2 >= 1 && epb == M / 2)
        return stage_dispatch_mode<R, P, (M / 2 >= 1 ? M / 2 : 1), AFF, false>(mode, a, ops, s);
    return hipErrorInvalidValue;
}

template <typename R, int P>
static hipError_t stage_dispatch_aff(int mode, const StageArgs<R>& a, const OpTables<R>& ops, int epb, bool aff,
                                     bool tb, hipStream_t s) {
    return aff ? stage_dispatch_epb<R, P, true>(mode, a, ops, epb, tb, s)
               : stage_dispatch_epb<R, P, false>(mode, a, ops, epb, tb, s);
}

template <typename R>
hipError_t launch_stage(int P, int mode, const StageArgs<R>& a, const OpTables<R>& ops, int epb, bool affine,
                        hipStream_t s, bool textbook) {
    if (a.E <= 0) return hipSuccess;
    switch (P) {
    case 1: return stage_dispatch_aff<R, 1>(mode, a, ops, epb, affine, textbook, s);
    case 2: return stage_dispatch_aff<R, 2>(mode, a, ops, epb, affine, textbook, s);
    case 3: return stage_dispatch_aff<R, 3>(mode, a, ops, epb, affine, textbook, s);
    case 4: return stage_dispatch_aff<R, 4>(mode, a, ops, epb, affine, textbook, s);
    case 5: return stage_dispatch_aff<R, 5>(mode, a, ops, epb, affine, textbook, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R, int P>
static hipError_t trace_dispatch(const TraceArgs<R>& a, const OpTables<R>& ops, hipStream_t s) {
    constexpr int M = ZJ_EPB(P);
    trace_kernel<R, P, M><<<dim3(num_blocks(a.E, M)), dim3(BLOCK), 0, s>>>(a, ops);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_trace(int P, const TraceArgs<R>& a, const OpTables<R>& ops, int, hipStream_t s) {
    if (a.E <= 0) return hipSuccess;
    switch (P) {
    case 1: return trace_dispatch<R, 1>(a, ops, s);
    case 2: return trace_dispatch<R, 2>(a, ops, s);
    case 3: return trace_dispatch<R, 3>(a, ops, s);
    case 4: return trace_dispatch<R, 4>(a, ops, s);
    case 5: return trace_dispatch<R, 5>(a, ops, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R, int P, bool AFF, bool TB>
static hipError_t batched_mode(int mode, const BatchInst<R>* insts, const int2* items, int n_items, int grid,
                               hipStream_t s) {
    const dim3 g(std::max(1, std::min(grid, n_items))), b(BLOCK);
    switch (mode) {
    case MODE_STAGE1: klaunch(batched_stage_kernel<R, P, MODE_STAGE1, AFF, TB>, g, b, 0, s, insts, items, n_items); break;
    case MODE_STAGE2: klaunch(batched_stage_kernel<R, P, MODE_STAGE2, AFF, TB>, g, b, 0, s, insts, items, n_items); break;
    case MODE_STAGE3: klaunch(batched_stage_kernel<R, P, MODE_STAGE3, AFF, TB>, g, b, 0, s, insts, items, n_items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
hipError_t launch_batched_stage(int P, bool affine, int mode, const BatchInst<R>* insts, const int2* items,
                                int n_items, int grid, hipStream_t s, bool textbook) {
    if (n_items <= 0) return hipSuccess;
#define ZJ_BP(PP)                                                                                    \
    case PP:                                                                                         \
        if (textbook)                                                                                \
            return affine ? batched_mode<R, PP, true, true>(mode, insts, items, n_items, grid, s)    \
                          : batched_mode<R, PP, false, true>(mode, insts, items, n_items, grid, s);  \
        return affine ? batched_mode<R, PP, true, false>(mode, insts, items, n_items, grid, s)       \
                      : batched_mode<R, PP, false, false>(mode, insts, items, n_items, grid, s);
    switch (P) {
        ZJ_BP(1) ZJ_BP(2) ZJ_BP(3) ZJ_BP(4) ZJ_BP(5)
    default: return hipErrorInvalidValue;
    }
#undef ZJ_BP
}

int batched_all_lds(int P, bool fp32) {
    switch (P) {
#define ZJ_BL(PP) \
    case PP: return fp32 ? BatchSmem<float, PP, ZJ_EPB(PP)>::bytes : BatchSmem<double, PP, ZJ_EPB(PP)>::bytes;
        ZJ_BL(1) ZJ_BL(2) ZJ_BL(3) ZJ_BL(4) ZJ_BL(5)
#undef ZJ_BL
    default: return 0;
    }
}

template <typename R>
hipError_t launch_batched_all(int mode, const BatchInst<R>* insts, const int2* items, int n_items, int grid,
                              int lds_bytes, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    const dim3 g(std::max(1, std::min(grid, n_items))), b(BLOCK);
    const unsigned lds = (unsigned)lds_bytes;
    switch (mode) {
    case MODE_STAGE1: klaunch(batched_all_kernel<R, MODE_STAGE1>, g, b, lds, s, insts, items, n_items); break;
    case MODE_STAGE2: klaunch(batched_all_kernel<R, MODE_STAGE2>, g, b, lds, s, insts, items, n_items); break;
    case MODE_STAGE3: klaunch(batched_all_kernel<R, MODE_STAGE3>, g, b, lds, s, insts, items, n_items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_batched_all<double>(int, const BatchInst<double>*, const int2*, int, int, int, hipStream_t);
template hipError_t launch_batched_all<float>(int, const BatchInst<float>*, const int2*, int, int, int, hipStream_t);
template hipError_t launch_batched_stage<double>(int, bool, int, const BatchInst<double>*, const int2*, int, int,
                                                 hipStream_t, bool);
template hipError_t launch_batched_stage<float>(int, bool, int, const BatchInst<float>*, const int2*, int, int,
                                                hipStream_t, bool);

template hipError_t launch_stage<double>(int, int, const StageArgs<double>&, const OpTables<double>&, int, bool,
                                        hipStream_t, bool);
template hipError_t launch_stage<float>(int, int, const StageArgs<float>&, const OpTables<float>&, int, bool,
                                       hipStream_t, bool);
template hipError_t launch_trace<double>(int, const TraceArgs<double>&, const OpTables<double>&, int, hipStream_t);
template hipError_t launch_trace<float>(int, const TraceArgs<float>&, const OpTables<float>&, int, hipStream_t);

hipError_t launch_aos_to_soa(const double* aos, double* soa, int64_t n, hipStream_t s, const int* perm, int S) {
    if (n <= 0) return hipSuccess;
    aos_to_soa_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(aos, soa, n, perm, S);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(const double* soa, double* aos, int64_t n, hipStream_t s, const int* perm, int S) {
    if (n <= 0) return hipSuccess;
    soa_to_aos_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(soa, aos, n, perm, S);
    return hipGetLastError();
}

hipError_t launch_set_many(double* dst, const double* v, int n, hipStream_t s) {
    for (int k = 0; k < n; k += 64) {
        ScalarChunk c{};
        c.n = std::min(64, n - k);
        for (int t = 0; t < c.n; ++t) c.v[t] = v[k + t];
        set_many_kernel<<<1, 64, 0, s>>>(dst + k, c);
    }
    return hipGetLastError();
}

hipError_t launch_set_scalar(double* p, double v, hipStream_t s) {
    set_scalar_kernel<<<1, 64, 0, s>>>(p, v);
    return hipGetLastError();
}

hipError_t launch_rk3_combine(int stage, double* out, const double* U0, const double* Uk, const double* R,
                              int64_t n, double dt, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    rk3_combine_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(stage, out, U0, Uk, R, n, dt);
    return hipGetLastError();
}

hipError_t launch_euler_update(double* U, const double* R, int64_t n, double dt, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    euler_update_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(U, R, n, dt);
    return hipGetLastError();
}

hipError_t launch_dt_partial(const double* U, const double* hcell, int E, int S, double gamma, double cfl,
                             double cf, double visc, double* partial, int nblk, hipStream_t s) {
    dt_partial_kernel<<<dim3(nblk), dim3(BLOCK), 0, s>>>(U, hcell, E, S, gamma, cfl, cf, visc, partial);
    return hipGetLastError();
}

hipError_t launch_dt_cfl_vjp(const double* U, const double* hcell, int E, int S, double gamma, double cfl,
                             double cf, double visc, const double* dtmin, unsigned long long* idx, double dt_bar,
                             double* Ub, int nblk, hipStream_t s) {
    fill_u64_kernel<<<1, 64, 0, s>>>(idx, ~0ull);
    dt_argmin_kernel<<<dim3((unsigned)std::max(nblk, 1)), dim3(BLOCK), 0, s>>>(U, hcell, E, S, gamma, cfl, cf, visc,
                                                                               dtmin, idx);
    dt_cfl_vjp_kernel<<<1, 64, 0, s>>>(U, hcell, E, S, gamma, cfl, cf, visc, idx, dt_bar, Ub);
    return hipGetLastError();
}

hipError_t launch_count_nonfinite(const double* U, int64_t n, unsigned long long* count, int nblk, hipStream_t s) {
    nonfinite_kernel<<<dim3(nblk), dim3(BLOCK), 0, s>>>(U, n, count);
    return hipGetLastError();
}

hipError_t launch_kinetic_energy(const double* U, const double* w, const int* perm, int E, int S, double* G,
                                 double* partial, int nblk, hipStream_t s) {
    kinetic_energy_kernel<<<dim3(nblk), dim3(BLOCK), 0, s>>>(U, w, perm, E, S, G, partial);
    return hipGetLastError();
}

hipError_t launch_finalize(const double* partial, int n, int op, double* out, hipStream_t s) {
    finalize_kernel<<<1, BLOCK, 0, s>>>(partial, n, op, out);
    return hipGetLastError();
}

}  // namespace zj
