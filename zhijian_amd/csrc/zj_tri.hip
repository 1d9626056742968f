// zj_tri.hip -- flux reconstruction on straight-sided triangles (textbook
// scheme) and its SSP-RK3 step: one fused kernel per stage, C ABI zj_tri_*
// (include/zhijian_amd.h).
//
// The reference has no triangle semantics to follow: FROperators::initTriangle
// leaves every table zero (src/basis/basis.cpp:615-669) and FRSolver assumes
// quadrilaterals (src/solver/fr_solver.cpp:138-140).  The scheme is the one its
// settings name on simplices: nodal FR with an energy-stable correction field
// (Castonguay, Vincent & Jameson 2012; the L2 lifting for DG, c = 0, and its
// (M + K)^-1 M filtering for SD = GA and HU, zj_tri_ops.hpp), transformed
// fluxes through the constant affine metric, one common flux per face point
// from the configured Riemann solver (include/physics/euler.hpp:204-375).
//
//   dU/dt = -(Dr Fr + Ds Fs) - sum_j lift[:, j] (F^I_j s_j - Fn_j)
//   Fr = r_x F + r_y G,  Fs = s_x F + s_y G   (contravariant fluxes)
//   s_j = |e|/(2J) of flux point j's edge:  |grad s|, |grad r + grad s|, |grad r|
//   Fn_j = interpolated transformed normal flux: -Fs, Fr + Fs, -Fr on edges 0, 1, 2
//   F^I_j = Riemann flux along the face normal, outward on each side
//
// Work mapping: one lane per (element, conservative variable); the four lanes
// of an element form a DPP quad.  A lane keeps its variable at every solution
// point in registers and applies the dense operators as sums of outer
// products whose entries are wave-uniform, read from constant memory through
// scalar loads: no operator or state traffic through LDS.  The Euler flux of
// point m gathers the point's four variables across the quad by DPP.  LDS
// holds only the element tile's flux-point traces (for the Riemann solves of
// in-tile faces) and the common fluxes on their way back to their variables.
//
// Device layout (one GPU holds the whole mesh):
//   U, W, R  state in blocks of 16 elements: [E/16][Np][16][4] FP64, so the 64
//            lanes of a wave (16 elements x 4 variables) touch 512 contiguous
//            bytes per solution point (W = U1, then U2 in place)
//   T[3]     flux-point traces [E][3(p+1)][4] in the compute precision;
//            only edges whose neighbour lies in another tile of 64 elements are
//            written (a neighbour in the same tile is read from LDS), rotating
//            as the quadrilateral path's: stage 1 T0->T1, 2 T1->T2, 3 T2->T0
//   geo      [E][10]: r_x, r_y, s_x, s_y, then the face normal (as the face's
//            left element sees it) of edges 0, 1, 2
//   links    [E][3] int2: x = neighbour element (-1 on a boundary), y = packed
//            neighbour edge (or BC type) | role << 3 | BC data index << 6
//   ops      constant memory, per order: Dr, Ds [Np][Np], interp [3(p+1)][Np],
//            lift [Np][3(p+1)]
#include "zhijian_amd.h"
#include "zj_device.hpp"
#include "zj_host_math.hpp"
#include "zj_kernels.hpp"
#include "zj_riemann_vjp.hpp"
#include "zj_tri_ops.hpp"
#include "zj_viscous.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <new>
#include <vector>

using namespace zj;

namespace {

template <int P>
struct TriShape {
    static constexpr int NP = (P + 1) * (P + 2) / 2;        // solution points
    static constexpr int FP = P + 1;                        // flux points per edge
    static constexpr int NFP = 3 * FP;
};

constexpr int TILE = BLOCK / NV;  // elements per workgroup (one lane per element and variable)
constexpr int EB = 16;            // elements per block of the state and trace layouts
constexpr int GEO = 10;           // geometry words per element

// Operator tables of order P in constant memory, laid out for wide scalar
// loads in the kernels' loops:
//   A  for each solution point m: column m of Dr, of Ds (NP each) and of the
//      interpolation to the flux points (NFP), contiguous -- the point loop of
//      the volume phase reads one such block per iteration;
//   B  the lifting matrix by columns: column j (NP values) contiguous;
//   C  the interpolation matrix by rows (row j: NP values), for the traces;
//   D  the lifting matrices of the SD (= GA) and HU correction fields, laid
//      out as B, one after the other.
// The kernel reads the lifting matrix at a per-handle offset (B or D).
template <int P> struct TriOps {
    static constexpr int NP = (P + 1) * (P + 2) / 2, NFP = 3 * (P + 1);
    static constexpr int CA = 2 * NP + NFP;         // words per block of A
    static constexpr int B = NP * CA, C = B + NFP * NP, D = C + NFP * NP;
};
constexpr int ops_words(int P) {
    return 2 * ((P + 1) * (P + 2) / 2) * ((P + 1) * (P + 2) / 2) + 5 * 3 * (P + 1) * ((P + 1) * (P + 2) / 2);
}
// Offset of a correction family's lifting matrix within the order-P block.
inline int lift_offset(int P, int family) {
    const int np = (P + 1) * (P + 2) / 2, nfp = 3 * (P + 1), b = np * (2 * np + nfp);
    if (family == ZJ_CORR_DG) return b;
    const int d = b + 2 * nfp * np;
    return family == ZJ_CORR_HU ? d + nfp * np : d;   // SD and GA share one field
}
constexpr int ops_offset(int P) {
    int o = 0;
    for (int q = 1; q < P; ++q) o += ops_words(q);
    return o;
}
constexpr int OPS_TOTAL = ops_offset(6);

__constant__ double c_ops64[OPS_TOTAL];
__constant__ float c_ops32[OPS_TOTAL];

// BR2 face liftings, per order and correction family (DG, SD = GA, HU):
// LF_k = interp(edge k rows) . lift(edge k columns), [3][FP][FP] -- the lifting
// of one edge's jumps evaluated at that edge's own flux points.
constexpr int lf_words(int P) { return 9 * (P + 1) * (P + 1); }
constexpr int lf_offset(int P) {
    int o = 0;
    for (int q = 1; q < P; ++q) o += 3 * lf_words(q);
    return o;
}
constexpr int LF_TOTAL = lf_offset(6);
inline int lf_family(int corr) { return corr == ZJ_CORR_DG ? 0 : (corr == ZJ_CORR_HU ? 2 : 1); }
__constant__ double c_lf64[LF_TOTAL];
__constant__ float c_lf32[LF_TOTAL];
template <typename R> __device__ __forceinline__ R clf(int k);
template <> __device__ __forceinline__ double clf<double>(int k) { return c_lf64[k]; }
template <> __device__ __forceinline__ float clf<float>(int k) { return c_lf32[k]; }

template <typename R> __device__ __forceinline__ R cop(int k);
template <> __device__ __forceinline__ double cop<double>(int k) { return c_ops64[k]; }
template <> __device__ __forceinline__ float cop<float>(int k) { return c_ops32[k]; }

// value of lane W of this lane's quad
template <int W> __device__ __forceinline__ double quad_get(double x) { return dpp_mov<W * 0x55>(x); }
template <int W> __device__ __forceinline__ float quad_get(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), W * 0x55, 0xF, 0xF, false));
}

// offset of variable 0 of (element e, point j) in a [E/16][n][16][4] array
__device__ __forceinline__ int64_t blk(int e, int j, int n) {
    return ((int64_t)(e >> 4) * n + j) * (EB * NV) + (e & (EB - 1)) * NV;
}
// offset of variable 0 of (element e, flux point j) in the [E][n][4] trace
// array: the p+1 points of an edge share one 128-byte line (FP64, p = 3)
__device__ __forceinline__ int64_t trix(int e, int j, int n) { return ((int64_t)e * n + j) * NV; }

template <typename R>
struct TriArgs {
    int E;
    const double* U_in;   // stage input
    const double* U0;     // U^n (stages 2, 3)
    double* U_out;        // stage output (may alias U_in or U0 lane-locally)
    double* R_out;        // residual (MODE_RESIDUAL)
    const R* geo;
    const int2* links;
    const R* T_in;
    R* T_out;
    const double* bc_data;
    const double* dt_ptr;
    double* l2_partial;   // per tile sum of R^2, or nullptr
    double gamma;
    int riemann;
    int xcd_groups;
    int lift_off;         // the correction family's lifting matrix in the order's operator block
    // Navier-Stokes terms (tri_visc_stage_kernel): mu = 1/Re, kappa, BR2 penalty,
    // and the family's face liftings in c_lf
    double mu, kappa, eta;
    int lf_off;
};

__device__ __forceinline__ int link_edge(int info) { return info & 7; }
__device__ __forceinline__ int link_role(int info) { return (info >> 3) & 3; }
__device__ __forceinline__ int link_bc(int info) { return info >> 6; }

// 1/x: a Newton-refined hardware reciprocal in FP64 (within 2 ulp), the
// hardware reciprocal in FP32.
template <typename R> __device__ __forceinline__ R fast_rcp(R x);
template <> __device__ __forceinline__ double fast_rcp<double>(double x) { return rcp_refined(x); }
template <> __device__ __forceinline__ float fast_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// Waves per SIMD the stage kernel is compiled for: the FP64 P3 kernel fits
// three (at most 168 VGPRs) without scratch; the rest is left to the compiler.
template <typename R, int P> struct TriWaves { static constexpr int value = 1; };
template <> struct TriWaves<double, 3> { static constexpr int value = 3; };

template <typename R, int P, int MODE>
__global__ __launch_bounds__(BLOCK, (TriWaves<R, P>::value)) void tri_stage_kernel(TriArgs<R> a) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, FP = S::FP, NFP = S::NFP;
    constexpr int NW = NP > NFP ? NP : NFP;
    using OP = TriOps<P>;
    constexpr int O0 = ops_offset(P), CA = OP::CA, OI = O0 + OP::C;
    const int OLT = O0 + a.lift_off;
    __shared__ R sT[TILE * NFP * NV];  // own traces
    __shared__ R sW[TILE * NW * NV];   // the stage input state, then the transformed normal fluxes,
                                       // then the flux differences
    __shared__ double sred[BLOCK / 64];
    __shared__ R sEdge[TILE * 3];                // |a| of every element edge of the tile
    __shared__ unsigned char sList[TILE * 3];    // edges whose face points this tile solves (el * 3 + edge)
    __shared__ int sCnt[BLOCK / 64];             // such edges per wave

    const int tid = threadIdx.x, el = tid >> 2, v = tid & 3;
    const int tile = tile_of_block(a.xcd_groups);
    const int e0 = tile * TILE, e = e0 + el;
    const bool live = e < a.E;
    const R g = R(a.gamma);
    const int64_t base = blk(e, 0, NP) + v;  // this lane's variable at point 0; point i at + i*EB*NV

    double u[NP];
    R div[NP];
    R rx = R(0), ry = R(0), sx = R(0), sy = R(0);
    int2 L[3] = {make_int2(0, 0), make_int2(0, 0), make_int2(0, 0)};

    // ---- 0: this lane's variable at every solution point (one coalesced
    // 512-byte row per point and wave), kept for the stage update
    if (live) {
#pragma unroll
        for (int m = 0; m < NP; ++m) u[m] = a.U_in[base + (int64_t)m * EB * NV];
        const R* G = a.geo + (int64_t)e * GEO;
        rx = G[0]; ry = G[1]; sx = G[2]; sy = G[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) L[k] = a.links[(int64_t)e * 3 + k];
#pragma unroll
        for (int m = 0; m < NP; ++m) sW[(el * NP + m) * NV + v] = R(u[m]);
        if (v < 3) {  // a: the metric gradient normal to edge v (-Fs on edge 0, Fr + Fs on 1, -Fr on 2)
            const R ax = v == 0 ? sx : (v == 1 ? rx + sx : rx), ay = v == 0 ? sy : (v == 1 ? ry + sy : ry);
            sEdge[el * 3 + v] = sqrtR(ax * ax + ay * ay);
        }
    }
    // Lane v < 3 of an element stands for its edge v.  An in-tile face is
    // solved by its left element only; an edge on a boundary or shared with
    // another tile by its own element.
    bool owned = false;
    if (live && v < 3) {
        const int2 Lv = v == 0 ? L[0] : (v == 1 ? L[1] : L[2]);
        const int nl = Lv.x - e0;
        owned = !(link_role(Lv.y) == ROLE_RIGHT && nl >= 0 && nl < TILE);
    }
    const unsigned long long own_mask = __ballot(owned);
    const int lane = tid & 63, wave = tid >> 6;
    if (lane == 0) sCnt[wave] = __popcll(own_mask);
    __syncthreads();

    // ---- 1: fluxes, divergence, traces and interpolated normal fluxes, as
    // sums over the solution points m of outer products with wave-uniform
    // operator columns
    R fn[NFP];
    if (live) {
#pragma unroll
        for (int i = 0; i < NP; ++i) div[i] = R(0);
#pragma unroll
        for (int j = 0; j < NFP; ++j) fn[j] = R(0);
#pragma unroll 1
        for (int m = 0; m < NP; ++m) {
            const R um = sW[(el * NP + m) * NV + v];
            const R q0 = quad_get<0>(um), q1 = quad_get<1>(um), q2 = quad_get<2>(um), q3 = quad_get<3>(um);
            const R inv = fast_rcp(q0);
            const R ux = q1 * inv, uy = q2 * inv;
            const R p = (g - R(1)) * (q3 - R(0.5) * (q1 * ux + q2 * uy));
            const R fx = v == 0 ? q1 : (v == 1 ? q1 * ux + p : (v == 2 ? q1 * uy : (q3 + p) * ux));
            const R fy = v == 0 ? q2 : (v == 1 ? q2 * ux : (v == 2 ? q2 * uy + p : (q3 + p) * uy));
            const R fr = rx * fx + ry * fy, fs = sx * fx + sy * fy, frs = fr + fs;
#pragma unroll
            for (int i = 0; i < NP; ++i)
                div[i] = fmaR(cop<R>(O0 + m * CA + NP + i), fs, fmaR(cop<R>(O0 + m * CA + i), fr, div[i]));
#pragma unroll
            for (int j = 0; j < NFP; ++j)
                fn[j] = fmaR(cop<R>(O0 + m * CA + 2 * NP + j), j < FP ? fs : (j < 2 * FP ? frs : fr), fn[j]);
        }
        // traces in a second pass, so that they and the fluxes' sums are not live together
        R ut[NFP];
#pragma unroll
        for (int j = 0; j < NFP; ++j) ut[j] = R(0);
#pragma unroll 1
        for (int m = 0; m < NP; ++m) {
            const R um = sW[(el * NP + m) * NV + v];
#pragma unroll
            for (int j = 0; j < NFP; ++j) ut[j] = fmaR(cop<R>(O0 + m * CA + 2 * NP + j), um, ut[j]);
        }
#pragma unroll
        for (int j = 0; j < NFP; ++j) sT[(el * NFP + j) * NV + v] = ut[j];
    }
    if (owned) {  // compact the solved edges, in element order
        int pos = __popcll(own_mask & ((1ull << lane) - 1));
        for (int w = 0; w < wave; ++w) pos += sCnt[w];
        sList[pos] = (unsigned char)(el * 3 + v);
    }
    __syncthreads();  // traces in place, state reads done
    // transformed normal fluxes to their flux points' slots in sW:
    // -Fs on edge 0, Fr + Fs on edge 1, -Fr on edge 2
    if (live) {
#pragma unroll
        for (int j = 0; j < NFP; ++j) sW[(el * NFP + j) * NV + v] = (j >= FP && j < 2 * FP) ? fn[j] : -fn[j];
    }
    __syncthreads();

    // ---- 2: common fluxes over the tile's solved edges, spread evenly over
    // the workgroup: the flux difference sc F^I - Fn of this side, and of the
    // in-tile neighbour's side with the same common flux (both elements store
    // the face normal of the left element, so the neighbour would have solved
    // the same problem).  sc = +-|a|, negative on the right.
    {
        int nsolved = 0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) nsolved += sCnt[w];
#pragma unroll 1
        for (int t = tid; t < nsolved * FP; t += BLOCK) {
            const int c = sList[t / FP], f = t - (t / FP) * FP;
            const int elt = c / 3, kk = c - elt * 3, j = kk * FP + f;
            const int2 Lk = a.links[(int64_t)(e0 + elt) * 3 + kk];
            const int role = link_role(Lk.y);
            const R nx = a.geo[(int64_t)(e0 + elt) * GEO + 4 + 2 * kk];
            const R ny = a.geo[(int64_t)(e0 + elt) * GEO + 5 + 2 * kk];
            R own[NV], oth[NV], UL[NV], UR[NV], Fc[NV];
#pragma unroll
            for (int w = 0; w < NV; ++w) own[w] = sT[(elt * NFP + j) * NV + w];
            int nl = -1, oj = 0;
            if (role == ROLE_BOUNDARY) {
                ghost_state(own, nx, ny, link_edge(Lk.y), a.bc_data + (int64_t)link_bc(Lk.y) * 8, g, oth);
            } else {
                nl = Lk.x - e0;
                oj = link_edge(Lk.y) * FP + (P - f);
                if (nl >= 0 && nl < TILE) {
#pragma unroll
                    for (int w = 0; w < NV; ++w) oth[w] = sT[(nl * NFP + oj) * NV + w];
                } else {
                    load4(a.T_in + trix(Lk.x, oj, NFP), oth);
                    nl = -1;
                }
            }
            const bool right = role == ROLE_RIGHT;
#pragma unroll
            for (int w = 0; w < NV; ++w) { UL[w] = right ? oth[w] : own[w]; UR[w] = right ? own[w] : oth[w]; }
            riemann_by_kind(a.riemann, UL, UR, nx, ny, g, Fc);
            {
                const R sc = right ? -sEdge[c] : sEdge[c];
                R* d = sW + (elt * NFP + j) * NV;
#pragma unroll
                for (int w = 0; w < NV; ++w) d[w] = sc * Fc[w] - d[w];
            }
            if (nl >= 0) {  // the in-tile neighbour, whose side is the right one
                const R sc = -sEdge[nl * 3 + link_edge(Lk.y)];
                R* d = sW + (nl * NFP + oj) * NV;
#pragma unroll
                for (int w = 0; w < NV; ++w) d[w] = sc * Fc[w] - d[w];
            }
        }
    }
    __syncthreads();

    // ---- 3: lifted flux differences, RK stage update
    double l2 = 0.0;
    double un[NP];
    if (live) {
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
            const R d = sW[(el * NFP + j) * NV + v];
#pragma unroll
            for (int i = 0; i < NP; ++i) div[i] = fmaR(cop<R>(OLT + j * NP + i), d, div[i]);
        }
        const double dt = MODE == MODE_RESIDUAL ? 0.0 : *a.dt_ptr;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const double res = -double(div[i]);
            const int64_t gi = base + (int64_t)i * EB * NV;
            if (MODE == MODE_RESIDUAL) {
                a.R_out[gi] = res;
                l2 += res * res;
            } else {
                if (MODE == MODE_STAGE1) {
                    un[i] = u[i] + dt * res;
                } else {
                    const double u0 = a.U0[gi];
                    un[i] = MODE == MODE_STAGE2 ? 0.75 * u0 + 0.25 * (u[i] + dt * res)
                                                : (1.0 / 3.0) * u0 + (2.0 / 3.0) * (u[i] + dt * res);
                    if (MODE == MODE_STAGE3) l2 += res * res;
                }
                a.U_out[gi] = un[i];
            }
        }
    }
    if ((MODE == MODE_RESIDUAL || MODE == MODE_STAGE3) && a.l2_partial) {
        const double s = block_reduce<OpSum>(l2, sred);
        if (tid == 0) a.l2_partial[tile] = s;
    }
    if (MODE == MODE_RESIDUAL || !live) return;

    // ---- 4: traces of the new stage state on edges another tile reads
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
        const int nl = L[kk].x - e0;
        if (link_role(L[kk].y) == ROLE_BOUNDARY || (nl >= 0 && nl < TILE)) continue;
#pragma unroll 1
        for (int j = kk * FP; j < kk * FP + FP; ++j) {
            R t = R(0);
#pragma unroll
            for (int m = 0; m < NP; ++m) t = fmaR(cop<R>(OI + j * NP + m), R(un[m]), t);
            a.T_out[trix(e, j, NFP) + v] = t;
        }
    }
}

// ------------------------------------------------------------ Navier-Stokes terms
// The stage with the compressible Navier-Stokes viscous terms, BR2 (textbook
// scheme; Bassi-Rebay 2 on the triangle's FR correction field), periodic
// meshes:
//   U^C = (U_own + U_nbr)/2 at every face point, jump U^C - U_own
//   Q = grad U + sum_j lift[:, j] s_j n_j (U^C - U_own)_j   (corrected gradient)
//   F = F_inv(U) - F_v(U, Q) at the points (F_v: ViscousFlux, mu = 1/Re)
//   F^I = Riemann(U_L, U_R) - F_v(U^C, Qbar) . n, Qbar the mean of both sides'
//       trace of grad U plus eta times the lifting of this face's jumps
//       alone at the face's points (LF_k = interp . lift of edge k)
//   dU/dt = -(Dr Fr + Ds Fs) - sum_j lift[:, j] (s_j F^I_j - Fn_j)
// Lanes as in tri_stage_kernel (one per element and variable), but the point
// values (U, then Q) live in LDS, and each lane reads the four variables of a
// point from there: the per-lane registers hold only the divergence and the
// normal-flux sums.  A tile holds TT elements, sized so that the trace arrays
// (U, dU/dx, dU/dy) and the point arrays (U, Qx, Qy; later the flux
// differences) fit the LDS.  Traces across tiles carry 12 words per flux
// point ([E][3(p+1)][12]: U, dU/dx, dU/dy).
constexpr int TWV = 12;

template <typename R, int P> struct TriViscTile {
    static constexpr int NP = (P + 1) * (P + 2) / 2, NFP = 3 * (P + 1);
    static constexpr int bytes(int tt) { return tt * (3 * NFP + 3 * NP) * NV * (int)sizeof(R); }
    // within 120 KiB of LDS (gfx950 gives a workgroup up to 160 KiB)
    static constexpr int value = bytes(64) <= 122880 ? 64 : (bytes(32) <= 122880 ? 32 : 16);
    static_assert(bytes(16) <= 122880, "the smallest tile exceeds the LDS");
};

template <typename R, int P, int MODE, int TT>
__global__ __launch_bounds__(TT * NV) void tri_visc_stage_kernel(TriArgs<R> a) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, FP = S::FP, NFP = S::NFP;
    constexpr int NB = TT * NV, NWV = NB / 64;
    static_assert(NB % 64 == 0 && 2 * NP >= NFP, "tile and point-array layout");
    using OP = TriOps<P>;
    constexpr int O0 = ops_offset(P), CA = OP::CA, OI = O0 + OP::C;
    const int OLT = O0 + a.lift_off;
    const int LFO = lf_offset(P) + a.lf_off;
    __shared__ R sTu[TT * NFP * NV], sTx[TT * NFP * NV], sTy[TT * NFP * NV];
    __shared__ R sU[TT * NP * NV];       // the stage input, later the stage output
    __shared__ R sQ[2 * TT * NP * NV];   // Qx, Qy; later the flux differences [TT][NFP][NV]
    __shared__ double sred[NWV];
    __shared__ R sEdge[TT * 3];
    __shared__ int2 sLink[TT * 3];  // the tile's edge links and face normals
    __shared__ R sNrm[TT * 3 * 2];
    __shared__ unsigned char sList[TT * 3];
    __shared__ int sCnt[NWV];
    R* const sQx = sQ;
    R* const sQy = sQ + TT * NP * NV;
    R* const sD = sQ;

    const int tid = threadIdx.x, el = tid >> 2, v = tid & 3;
    const int tile = tile_of_block(a.xcd_groups);
    const int e0 = tile * TT, e = e0 + el;
    const bool live = e < a.E;
    const R g = R(a.gamma), mu = R(a.mu), kappa = R(a.kappa), eta = R(a.eta);
    const int64_t base = blk(e, 0, NP) + v;
    auto sl = [](int elx, int j, int w) { return (elx * NFP + j) * NV + w; };   // trace slots
    auto sp = [](int elx, int m, int w) { return (elx * NP + m) * NV + w; };    // point slots

    // physical gradient of this lane's variable at the points, from sU into gx,
    // gy (every point's sums advance together: 2 NP independent FMA chains)
    auto gradient_of = [&](R rx_, R ry_, R sx_, R sy_, R* gx, R* gy) __attribute__((always_inline)) {
        R gr[NP], gs[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) { gr[i] = R(0); gs[i] = R(0); }
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const R um = sU[sp(el, m, v)];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                gr[i] = fmaR(cop<R>(O0 + m * CA + i), um, gr[i]);
                gs[i] = fmaR(cop<R>(O0 + m * CA + NP + i), um, gs[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            gx[sp(el, i, v)] = rx_ * gr[i] + sx_ * gs[i];
            gy[sp(el, i, v)] = ry_ * gr[i] + sy_ * gs[i];
        }
    };
    // traces of U, gx, gy (this lane's variable) at the FP points of edge kk
    // (3 FP independent FMA chains), for each through out(j, tu, tx, ty)
    auto edge_traces = [&](int kk, int op_row, const R* gx, const R* gy, auto&& out) __attribute__((always_inline)) {
        R tu[FP], tx[FP], ty[FP];
#pragma unroll
        for (int f = 0; f < FP; ++f) { tu[f] = R(0); tx[f] = R(0); ty[f] = R(0); }
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const R um = sU[sp(el, m, v)], xm = gx[sp(el, m, v)], ym = gy[sp(el, m, v)];
#pragma unroll
            for (int f = 0; f < FP; ++f) {
                const int j = kk * FP + f;
                const R w = op_row ? cop<R>(OI + j * NP + m) : cop<R>(O0 + m * CA + 2 * NP + j);
                tu[f] = fmaR(w, um, tu[f]);
                tx[f] = fmaR(w, xm, tx[f]);
                ty[f] = fmaR(w, ym, ty[f]);
            }
        }
#pragma unroll
        for (int f = 0; f < FP; ++f) out(kk * FP + f, tu[f], tx[f], ty[f]);
    };

    // ---- 0: state, gradient, traces of U and grad U
    R rx = R(0), ry = R(0), sx = R(0), sy = R(0);
    int2 L[3] = {make_int2(0, 0), make_int2(0, 0), make_int2(0, 0)};
    if (live) {
#pragma unroll
        for (int m = 0; m < NP; ++m) sU[sp(el, m, v)] = R(a.U_in[base + (int64_t)m * EB * NV]);
        const R* G = a.geo + (int64_t)e * GEO;
        rx = G[0]; ry = G[1]; sx = G[2]; sy = G[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) L[k] = a.links[(int64_t)e * 3 + k];
        if (v < 3) {
            const R ax = v == 0 ? sx : (v == 1 ? rx + sx : rx), ay = v == 0 ? sy : (v == 1 ? ry + sy : ry);
            sEdge[el * 3 + v] = sqrtR(ax * ax + ay * ay);
            // the edge's link and normal, for whichever lane solves its face points
            sLink[el * 3 + v] = v == 0 ? L[0] : (v == 1 ? L[1] : L[2]);
            sNrm[(el * 3 + v) * 2] = G[4 + 2 * v];
            sNrm[(el * 3 + v) * 2 + 1] = G[5 + 2 * v];
        }
        gradient_of(rx, ry, sx, sy, sQx, sQy);
#pragma unroll 1
        for (int kk = 0; kk < 3; ++kk)
            edge_traces(kk, 0, sQx, sQy, [&](int j, R tu, R tx, R ty) __attribute__((always_inline)) {
                sTu[sl(el, j, v)] = tu;
                sTx[sl(el, j, v)] = tx;
                sTy[sl(el, j, v)] = ty;
            });
    }
    bool owned = false;
    if (live && v < 3) {
        const int2 Lv = v == 0 ? L[0] : (v == 1 ? L[1] : L[2]);
        const int nl = Lv.x - e0;
        owned = !(link_role(Lv.y) == ROLE_RIGHT && nl >= 0 && nl < TT);
    }
    const unsigned long long own_mask = __ballot(owned);
    const int lane = tid & 63, wave = tid >> 6;
    if (lane == 0) sCnt[wave] = __popcll(own_mask);
    __syncthreads();

    // ---- 1: corrected gradient: the lifted jumps along the outward normals
    if (live) {
        R jwx[NFP], jwy[NFP];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
            const int2 Lk = L[kk];
            const R sg = link_role(Lk.y) == ROLE_RIGHT ? R(-1) : R(1);
            const R s = sEdge[el * 3 + kk];
            const R wx = sg * s * a.geo[(int64_t)e * GEO + 4 + 2 * kk];
            const R wy = sg * s * a.geo[(int64_t)e * GEO + 5 + 2 * kk];
            const int nl = Lk.x - e0;
            const bool in = nl >= 0 && nl < TT;
            const int ek = link_edge(Lk.y);
#pragma unroll
            for (int f = 0; f < FP; ++f) {
                const int j = kk * FP + f, oj = ek * FP + (P - f);
                const R own = sTu[sl(el, j, v)];
                R oth;
                if (in) oth = sTu[sl(nl, oj, v)];
                else oth = a.T_in[((int64_t)Lk.x * NFP + oj) * TWV + v];
                const R jmp = R(0.5) * (own + oth) - own;
                jwx[j] = wx * jmp;
                jwy[j] = wy * jmp;
            }
        }
#pragma unroll 2
        for (int i = 0; i < NP; ++i) {
            R qx = sQx[sp(el, i, v)], qy = sQy[sp(el, i, v)];
#pragma unroll
            for (int j = 0; j < NFP; ++j) {
                const R l = cop<R>(OLT + j * NP + i);
                qx = fmaR(l, jwx[j], qx);
                qy = fmaR(l, jwy[j], qy);
            }
            sQx[sp(el, i, v)] = qx;
            sQy[sp(el, i, v)] = qy;
        }
    }
    if (owned) {
        int pos = __popcll(own_mask & ((1ull << lane) - 1));
        for (int w = 0; w < wave; ++w) pos += sCnt[w];
        sList[pos] = (unsigned char)(el * 3 + v);
    }
    __syncthreads();  // Q of every variable in place

    // ---- 2: total fluxes at the points, each point's four components formed
    // once: this lane takes points v, v + 4, ... of its element and leaves
    // their contravariant fluxes where Q was; then every lane accumulates its
    // own component's divergence and interpolated normal fluxes
    constexpr int MPL = (NP + NV - 1) / NV;  // points per lane
    R FR[MPL][NV], FS[MPL][NV];
    if (live) {
#pragma unroll
        for (int q = 0; q < MPL; ++q) {
            const int m = v + NV * q;
            if (m >= NP) break;
            R U4[NV], X4[NV], Y4[NV];
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                U4[w] = sU[sp(el, m, w)];
                X4[w] = sQx[sp(el, m, w)];
                Y4[w] = sQy[sp(el, m, w)];
            }
            const R inv = fast_rcp(U4[0]);
            const R ux = U4[1] * inv, uy = U4[2] * inv;
            const R p = (g - R(1)) * (U4[3] - R(0.5) * (U4[1] * ux + U4[2] * uy));
            R Fv[NV], Gv[NV];
            visc_flux(U4, X4, Y4, g, mu, kappa, Fv, Gv);
            const R fx[NV] = {U4[1], U4[1] * ux + p, U4[1] * uy, (U4[3] + p) * ux};
            const R fy[NV] = {U4[2], U4[2] * ux, U4[2] * uy + p, (U4[3] + p) * uy};
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                const R Fx = fx[w] - Fv[w], Fy = fy[w] - Gv[w];
                FR[q][w] = rx * Fx + ry * Fy;
                FS[q][w] = sx * Fx + sy * Fy;
            }
        }
    }
    __syncthreads();  // every read of Q done: the contravariant fluxes take its place
    if (live) {
#pragma unroll
        for (int q = 0; q < MPL; ++q) {
            const int m = v + NV * q;
            if (m >= NP) break;
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                sQx[sp(el, m, w)] = FR[q][w];
                sQy[sp(el, m, w)] = FS[q][w];
            }
        }
    }
    __syncthreads();
    R div[NP], fn[NFP];
#pragma unroll
    for (int i = 0; i < NP; ++i) div[i] = R(0);
#pragma unroll
    for (int j = 0; j < NFP; ++j) fn[j] = R(0);
    if (live) {
#pragma unroll 1
        for (int m = 0; m < NP; ++m) {
            const R fr = sQx[sp(el, m, v)], fs = sQy[sp(el, m, v)], frs = fr + fs;
#pragma unroll
            for (int i = 0; i < NP; ++i)
                div[i] = fmaR(cop<R>(O0 + m * CA + NP + i), fs, fmaR(cop<R>(O0 + m * CA + i), fr, div[i]));
#pragma unroll
            for (int j = 0; j < NFP; ++j)
                fn[j] = fmaR(cop<R>(O0 + m * CA + 2 * NP + j), j < FP ? fs : (j < 2 * FP ? frs : fr), fn[j]);
        }
    }
    __syncthreads();  // every read of the point fluxes done: the flux differences take their place
    if (live) {
#pragma unroll
        for (int j = 0; j < NFP; ++j) sD[sl(el, j, v)] = (j >= FP && j < 2 * FP) ? fn[j] : -fn[j];
    }
    __syncthreads();
    int nsolved = 0;
#pragma unroll
    for (int w = 0; w < NWV; ++w) nsolved += sCnt[w];

    // ---- 3: common fluxes of the solved face points: Riemann minus the BR2
    // viscous flux; flux differences of both sides (in-tile partner included)
#pragma unroll 1
    for (int t = tid; t < nsolved * FP; t += NB) {
        const int c = sList[t / FP], f = t - (t / FP) * FP;
        const int elt = c / 3, kk = c - elt * 3, j = kk * FP + f;
        const int2 Lk = sLink[elt * 3 + kk];
        const bool right = link_role(Lk.y) == ROLE_RIGHT;
        const R nx = sNrm[(elt * 3 + kk) * 2];
        const R ny = sNrm[(elt * 3 + kk) * 2 + 1];
        const int ek = link_edge(Lk.y), oj = ek * FP + (P - f);
        const int nl = Lk.x - e0;
        const bool in = nl >= 0 && nl < TT;
        const R* Tn = a.T_in + (int64_t)Lk.x * NFP * TWV;  // the partner's traces (across tiles)
        R own[NV], oth[NV], gxo[NV], gyo[NV], gxn[NV], gyn[NV];
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            own[w] = sTu[sl(elt, j, w)];
            gxo[w] = sTx[sl(elt, j, w)];
            gyo[w] = sTy[sl(elt, j, w)];
        }
        if (in) {
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                oth[w] = sTu[sl(nl, oj, w)];
                gxn[w] = sTx[sl(nl, oj, w)];
                gyn[w] = sTy[sl(nl, oj, w)];
            }
        } else {
            load4(Tn + (int64_t)oj * TWV, oth);
            load4(Tn + (int64_t)oj * TWV + NV, gxn);
            load4(Tn + (int64_t)oj * TWV + 2 * NV, gyn);
        }
        // each side's edge factor |e|/(2J)
        const R so = sEdge[elt * 3 + kk];
        R sn;
        if (in) {
            sn = sEdge[nl * 3 + ek];
        } else {
            const R* Gn = a.geo + (int64_t)Lk.x * GEO;
            const R ax = ek == 0 ? Gn[2] : (ek == 1 ? Gn[0] + Gn[2] : Gn[0]);
            const R ay = ek == 0 ? Gn[3] : (ek == 1 ? Gn[1] + Gn[3] : Gn[1]);
            sn = sqrtR(ax * ax + ay * ay);
        }
        // liftings of this face's jumps at both sides' points; the partner's
        // jump at its slot P-f2 is minus this side's at f2
        R lo[NV] = {R(0), R(0), R(0), R(0)}, ln[NV] = {R(0), R(0), R(0), R(0)};
#pragma unroll 2
        for (int f2 = 0; f2 < FP; ++f2) {
            const int j2 = kk * FP + f2, o2 = ek * FP + (P - f2);
            R uo[NV], un[NV];
#pragma unroll
            for (int w = 0; w < NV; ++w) uo[w] = sTu[sl(elt, j2, w)];
            if (in) {
#pragma unroll
                for (int w = 0; w < NV; ++w) un[w] = sTu[sl(nl, o2, w)];
            } else {
                load4(Tn + (int64_t)o2 * TWV, un);
            }
            const R wo = clf<R>(LFO + (kk * FP + f) * FP + f2);
            const R wn = clf<R>(LFO + (ek * FP + (P - f)) * FP + (P - f2));
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                const R jo = R(0.5) * (uo[w] + un[w]) - uo[w];
                lo[w] = fmaR(wo, jo, lo[w]);
                ln[w] = fmaR(wn, jo, ln[w]);
            }
        }
        const R sg = right ? R(-1) : R(1);
        const R nox = sg * nx, noy = sg * ny;  // this side's outward normal
        R Uc[NV], Qbx[NV], Qby[NV], UL[NV], UR[NV], Fc[NV], Fv[NV], Gv[NV];
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            Uc[w] = R(0.5) * (own[w] + oth[w]);
            Qbx[w] = R(0.5) * ((gxo[w] + eta * so * nox * lo[w]) + (gxn[w] + eta * sn * nox * ln[w]));
            Qby[w] = R(0.5) * ((gyo[w] + eta * so * noy * lo[w]) + (gyn[w] + eta * sn * noy * ln[w]));
            UL[w] = right ? oth[w] : own[w];
            UR[w] = right ? own[w] : oth[w];
        }
        visc_flux(Uc, Qbx, Qby, g, mu, kappa, Fv, Gv);
        riemann_by_kind(a.riemann, UL, UR, nx, ny, g, Fc);
        {
            const R sc = right ? -so : so;
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                const R fi = Fc[w] - (Fv[w] * nx + Gv[w] * ny);
                sD[sl(elt, j, w)] = sc * fi - sD[sl(elt, j, w)];
            }
        }
        if (in) {  // the in-tile partner, whose side is the right one
            const R sc = -sEdge[nl * 3 + ek];
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                const R fi = Fc[w] - (Fv[w] * nx + Gv[w] * ny);
                sD[sl(nl, oj, w)] = sc * fi - sD[sl(nl, oj, w)];
            }
        }
    }
    __syncthreads();

    // ---- 4: lifted flux differences, RK stage update (the stage input read
    // again from memory); the stage output to sU for its traces
    double l2 = 0.0;
    if (live) {
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
            const R d = sD[sl(el, j, v)];
#pragma unroll
            for (int i = 0; i < NP; ++i) div[i] = fmaR(cop<R>(OLT + j * NP + i), d, div[i]);
        }
        const double dt = MODE == MODE_RESIDUAL ? 0.0 : *a.dt_ptr;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const double res = -double(div[i]);
            const int64_t gi = base + (int64_t)i * EB * NV;
            if (MODE == MODE_RESIDUAL) {
                a.R_out[gi] = res;
                l2 += res * res;
            } else {
                const double q = a.U_in[gi];
                double un;
                if (MODE == MODE_STAGE1) {
                    un = q + dt * res;
                } else {
                    const double u0 = a.U0[gi];
                    un = MODE == MODE_STAGE2 ? 0.75 * u0 + 0.25 * (q + dt * res)
                                             : (1.0 / 3.0) * u0 + (2.0 / 3.0) * (q + dt * res);
                    if (MODE == MODE_STAGE3) l2 += res * res;
                }
                a.U_out[gi] = un;
                sU[sp(el, i, v)] = R(un);  // only this lane reads its variable's points back
            }
        }
    }
    if ((MODE == MODE_RESIDUAL || MODE == MODE_STAGE3) && a.l2_partial) {
        const double s2 = block_reduce_n<OpSum, NWV>(l2, sred);
        if (tid == 0) a.l2_partial[tile] = s2;
    }
    if (MODE == MODE_RESIDUAL) return;
    __syncthreads();  // every read of the flux differences done: their space is reused below
    if (!live) return;

    // ---- 5: traces of the new stage state and of its gradient on edges
    // another tile reads (the gradient into the Q arrays, own variable only)
    bool any_out = false;
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
        const int nl = L[kk].x - e0;
        any_out = any_out || !(nl >= 0 && nl < TT);
    }
    if (!any_out) return;
    R* const gx = sQx;  // each lane reads only its own variable's slots from here on
    R* const gy = sQy;
    gradient_of(rx, ry, sx, sy, gx, gy);
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
        const int nl = L[kk].x - e0;
        if (nl >= 0 && nl < TT) continue;
        edge_traces(kk, 1, gx, gy, [&](int j, R tu, R tx, R ty) __attribute__((always_inline)) {
            R* T = a.T_out + ((int64_t)e * NFP + j) * TWV;
            T[v] = tu;
            T[NV + v] = tx;
            T[2 * NV + v] = ty;
        });
    }
}

// Traces of U and grad U on every edge (after a new state is set).
template <typename R, int P>
__global__ __launch_bounds__(BLOCK) void tri_visc_trace_kernel(int E, const double* U, const R* geo, R* T_out) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, NFP = S::NFP, O0 = ops_offset(P), CA = TriOps<P>::CA;
    const int e = blockIdx.x * TILE + (threadIdx.x >> 2), v = threadIdx.x & 3;
    if (e >= E) return;
    const R* G = geo + (int64_t)e * GEO;
    const R rx = G[0], ry = G[1], sx = G[2], sy = G[3];
    R u[NP], gr[NP], gs[NP];
#pragma unroll
    for (int m = 0; m < NP; ++m) { u[m] = R(U[blk(e, m, NP) + v]); gr[m] = R(0); gs[m] = R(0); }
#pragma unroll
    for (int m = 0; m < NP; ++m)
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            gr[i] = fmaR(cop<R>(O0 + m * CA + i), u[m], gr[i]);
            gs[i] = fmaR(cop<R>(O0 + m * CA + NP + i), u[m], gs[i]);
        }
#pragma unroll 1
    for (int j = 0; j < NFP; ++j) {
        R tu = R(0), tx = R(0), ty = R(0);
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const R w = cop<R>(O0 + m * CA + 2 * NP + j);
            tu = fmaR(w, u[m], tu);
            tx = fmaR(w, rx * gr[m] + sx * gs[m], tx);
            ty = fmaR(w, ry * gr[m] + sy * gs[m], ty);
        }
        R* T = T_out + ((int64_t)e * NFP + j) * TWV;
        T[v] = tu;
        T[NV + v] = tx;
        T[2 * NV + v] = ty;
    }
}

// ------------------------------------------------------------ adjoint
// Vector-Jacobian product of one triangle stage u_n = a U^n + b q + c dt R(q)
// (MODE_STAGE1..3), or of the residual alone (MODE_RESIDUAL).  The forward
// values are recomputed as tri_stage_kernel forms them; then, per lane
// (element, variable):
//   R1  cotangent of u_out: its partial value plus the transposed trace
//       operator on the edge cotangents other tiles left (cross-tile edges);
//       the RK combination's cotangents of U^n, q and dt
//   R2  the lifting transposed: the flux differences' cotangents (LDS)
//   R3  per solved face point (as the forward solves them): the common flux's
//       cotangent from both sides' differences, the Riemann solver's VJP, to
//       this side's trace cotangent (LDS), an in-tile neighbour's (LDS) or a
//       cross-tile neighbour's (HBM, at its slot) or through the ghost state
//   R4  per solution point: Dr, Ds and the interpolated normal flux
//       transposed, the metric, the Euler flux Jacobian (the four variables'
//       cotangents gathered across the element's DPP quad), and the
//       transposed trace operator
template <typename R>
struct TriVjpArgs {
    int E;
    const double* U_in;     // q
    const double* U0;       // U^n (stages 2, 3)
    const R* geo;
    const int2* links;
    const R* T_in;          // traces of q on edges read across tiles
    const double* bc_data;
    const double* dt_ptr;
    const double* Ub_in;    // partial cotangent of u_out, or of R
    const R* Tb_in;         // cross-tile edge cotangents completing Ub_in (nullptr: none)
    double* Ub_out;         // partial cotangent of q (may alias Ub_in)
    R* Tb_out;              // cross-tile edge cotangents of q, at the neighbour's slots
    double* U0b;            // cotangent of U^n: stage 3 writes, stage 2 adds, stage 1 reads
    double* dtb_partial;    // per-tile sum of c R . du_n (nullptr: skip)
    double gamma;
    int riemann;
    int xcd_groups;
    int lift_off;
    double mu, kappa, eta;  // Navier-Stokes terms (tri_visc_vjp_kernel)
    int lf_off;
};

template <typename R, int P, int MODE>
__global__ __launch_bounds__(BLOCK) void tri_vjp_kernel(TriVjpArgs<R> a) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, FP = S::FP, NFP = S::NFP;
    constexpr int NW = NP > NFP ? NP : NFP;
    using OP = TriOps<P>;
    constexpr int O0 = ops_offset(P), CA = OP::CA, OI = O0 + OP::C;
    constexpr double ca = MODE == MODE_STAGE2 ? 0.75 : (MODE == MODE_STAGE3 ? 1.0 / 3.0 : 0.0);
    constexpr double cb = MODE == MODE_STAGE2 ? 0.25 : (MODE == MODE_STAGE3 ? 2.0 / 3.0 : 1.0);
    constexpr double cc = MODE == MODE_STAGE2 ? 0.25 : (MODE == MODE_STAGE3 ? 2.0 / 3.0 : 1.0);
    const int OLT = O0 + a.lift_off;
    __shared__ R sT[TILE * NFP * NV];   // own traces
    __shared__ R sW[TILE * NW * NV];    // q, then Fn, then the flux differences d, then their cotangents
    __shared__ R sTb[TILE * NFP * NV];  // trace cotangents
    __shared__ double sred[BLOCK / 64];
    __shared__ R sEdge[TILE * 3];
    __shared__ unsigned char sList[TILE * 3];
    __shared__ int sCnt[BLOCK / 64];

    const int tid = threadIdx.x, el = tid >> 2, v = tid & 3;
    const int tile = tile_of_block(a.xcd_groups);
    const int e0 = tile * TILE, e = e0 + el;
    const bool live = e < a.E;
    const R g = R(a.gamma);
    const int64_t base = blk(e, 0, NP) + v;

    double q[NP];
    R rx = R(0), ry = R(0), sx = R(0), sy = R(0);
    int2 L[3] = {make_int2(0, 0), make_int2(0, 0), make_int2(0, 0)};
    if (live) {
#pragma unroll
        for (int m = 0; m < NP; ++m) q[m] = a.U_in[base + (int64_t)m * EB * NV];
        const R* G = a.geo + (int64_t)e * GEO;
        rx = G[0]; ry = G[1]; sx = G[2]; sy = G[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) L[k] = a.links[(int64_t)e * 3 + k];
#pragma unroll
        for (int m = 0; m < NP; ++m) sW[(el * NP + m) * NV + v] = R(q[m]);
        if (v < 3) {
            const R ax = v == 0 ? sx : (v == 1 ? rx + sx : rx), ay = v == 0 ? sy : (v == 1 ? ry + sy : ry);
            sEdge[el * 3 + v] = sqrtR(ax * ax + ay * ay);
        }
    } else {
#pragma unroll
        for (int m = 0; m < NP; ++m) q[m] = 0.0;
    }
    bool owned = false;
    if (live && v < 3) {
        const int2 Lv = v == 0 ? L[0] : (v == 1 ? L[1] : L[2]);
        const int nl = Lv.x - e0;
        owned = !(link_role(Lv.y) == ROLE_RIGHT && nl >= 0 && nl < TILE);
    }
    const unsigned long long own_mask = __ballot(owned);
    const int lane = tid & 63, wave = tid >> 6;
    if (lane == 0) sCnt[wave] = __popcll(own_mask);
    __syncthreads();

    // ---- forward 1: divergence, interpolated normal fluxes, traces
    R div[NP], fn[NFP];
    if (live) {
#pragma unroll
        for (int i = 0; i < NP; ++i) div[i] = R(0);
#pragma unroll
        for (int j = 0; j < NFP; ++j) fn[j] = R(0);
#pragma unroll 1
        for (int m = 0; m < NP; ++m) {
            const R um = sW[(el * NP + m) * NV + v];
            const R q0 = quad_get<0>(um), q1 = quad_get<1>(um), q2 = quad_get<2>(um), q3 = quad_get<3>(um);
            const R inv = fast_rcp(q0);
            const R ux = q1 * inv, uy = q2 * inv;
            const R p = (g - R(1)) * (q3 - R(0.5) * (q1 * ux + q2 * uy));
            const R fx = v == 0 ? q1 : (v == 1 ? q1 * ux + p : (v == 2 ? q1 * uy : (q3 + p) * ux));
            const R fy = v == 0 ? q2 : (v == 1 ? q2 * ux : (v == 2 ? q2 * uy + p : (q3 + p) * uy));
            const R fr = rx * fx + ry * fy, fs = sx * fx + sy * fy, frs = fr + fs;
#pragma unroll
            for (int i = 0; i < NP; ++i)
                div[i] = fmaR(cop<R>(O0 + m * CA + NP + i), fs, fmaR(cop<R>(O0 + m * CA + i), fr, div[i]));
#pragma unroll
            for (int j = 0; j < NFP; ++j)
                fn[j] = fmaR(cop<R>(O0 + m * CA + 2 * NP + j), j < FP ? fs : (j < 2 * FP ? frs : fr), fn[j]);
        }
        R ut[NFP];
#pragma unroll
        for (int j = 0; j < NFP; ++j) ut[j] = R(0);
#pragma unroll 1
        for (int m = 0; m < NP; ++m) {
            const R um = sW[(el * NP + m) * NV + v];
#pragma unroll
            for (int j = 0; j < NFP; ++j) ut[j] = fmaR(cop<R>(O0 + m * CA + 2 * NP + j), um, ut[j]);
        }
#pragma unroll
        for (int j = 0; j < NFP; ++j) sT[(el * NFP + j) * NV + v] = ut[j];
    }
    if (owned) {
        int pos = __popcll(own_mask & ((1ull << lane) - 1));
        for (int w = 0; w < wave; ++w) pos += sCnt[w];
        sList[pos] = (unsigned char)(el * 3 + v);
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int j = 0; j < NFP; ++j) sW[(el * NFP + j) * NV + v] = (j >= FP && j < 2 * FP) ? fn[j] : -fn[j];
    }
    __syncthreads();
    int nsolved = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) nsolved += sCnt[w];

    // ---- forward 2: flux differences d = sc F^I - Fn of both sides of every solved face point
#pragma unroll 1
    for (int t = tid; t < nsolved * FP; t += BLOCK) {
        const int c = sList[t / FP], f = t - (t / FP) * FP;
        const int elt = c / 3, kk = c - elt * 3, j = kk * FP + f;
        const int2 Lk = a.links[(int64_t)(e0 + elt) * 3 + kk];
        const int role = link_role(Lk.y);
        const R nx = a.geo[(int64_t)(e0 + elt) * GEO + 4 + 2 * kk];
        const R ny = a.geo[(int64_t)(e0 + elt) * GEO + 5 + 2 * kk];
        R own[NV], oth[NV], UL[NV], UR[NV], Fc[NV];
#pragma unroll
        for (int w = 0; w < NV; ++w) own[w] = sT[(elt * NFP + j) * NV + w];
        int nl = -1, oj = 0;
        if (role == ROLE_BOUNDARY) {
            ghost_state(own, nx, ny, link_edge(Lk.y), a.bc_data + (int64_t)link_bc(Lk.y) * 8, g, oth);
        } else {
            nl = Lk.x - e0;
            oj = link_edge(Lk.y) * FP + (P - f);
            if (nl >= 0 && nl < TILE) {
#pragma unroll
                for (int w = 0; w < NV; ++w) oth[w] = sT[(nl * NFP + oj) * NV + w];
            } else {
                load4(a.T_in + trix(Lk.x, oj, NFP), oth);
                nl = -1;
            }
        }
        const bool right = role == ROLE_RIGHT;
#pragma unroll
        for (int w = 0; w < NV; ++w) { UL[w] = right ? oth[w] : own[w]; UR[w] = right ? own[w] : oth[w]; }
        riemann_by_kind(a.riemann, UL, UR, nx, ny, g, Fc);
        {
            const R sc = right ? -sEdge[c] : sEdge[c];
            R* d = sW + (elt * NFP + j) * NV;
#pragma unroll
            for (int w = 0; w < NV; ++w) d[w] = sc * Fc[w] - d[w];
        }
        if (nl >= 0) {
            const R sc = -sEdge[nl * 3 + link_edge(Lk.y)];
            R* d = sW + (nl * NFP + oj) * NV;
#pragma unroll
            for (int w = 0; w < NV; ++w) d[w] = sc * Fc[w] - d[w];
        }
    }
    __syncthreads();

    // ---- forward 3: residual; R1: cotangent of the stage output, RK combination
    double rsum = 0.0;
    R divb[NP];
    double qb[NP];
    if (live) {
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
            const R d = sW[(el * NFP + j) * NV + v];
#pragma unroll
            for (int i = 0; i < NP; ++i) div[i] = fmaR(cop<R>(OLT + j * NP + i), d, div[i]);
        }
        // incoming edge cotangents of u_out on this element's cross-tile edges
        R tin[NFP];
#pragma unroll
        for (int j = 0; j < NFP; ++j) tin[j] = R(0);
        if (MODE != MODE_RESIDUAL && a.Tb_in) {
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                const int nl = L[kk].x - e0;
                const int role = link_role(L[kk].y);
                if (role == ROLE_BOUNDARY || (nl >= 0 && nl < TILE)) continue;
#pragma unroll
                for (int f = 0; f < FP; ++f) tin[kk * FP + f] = a.Tb_in[trix(e, kk * FP + f, NFP) + v];
            }
        }
        const double dt = MODE == MODE_RESIDUAL ? 0.0 : *a.dt_ptr;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int64_t gi = base + (int64_t)i * EB * NV;
            double rb;
            if (MODE == MODE_RESIDUAL) {
                rb = a.Ub_in[gi];
                qb[i] = 0.0;
            } else {
                R tsum = R(0);
#pragma unroll
                for (int j = 0; j < NFP; ++j) tsum = fmaR(cop<R>(OI + j * NP + i), tin[j], tsum);
                const double unb = a.Ub_in[gi] + double(tsum);
                if (MODE == MODE_STAGE3) a.U0b[gi] = ca * unb;
                if (MODE == MODE_STAGE2) a.U0b[gi] += ca * unb;
                qb[i] = cb * unb;
                rb = cc * dt * unb;
                rsum += cc * (-double(div[i])) * unb;
            }
            divb[i] = R(-rb);  // R = -div
        }
    } else {
#pragma unroll
        for (int i = 0; i < NP; ++i) { divb[i] = R(0); qb[i] = 0.0; }
    }
    if (a.dtb_partial) {
        const double s = block_reduce<OpSum>(rsum, sred);
        if (tid == 0) a.dtb_partial[tile] = s;
    }
    __syncthreads();  // every lane's reads of sW (forward 3) are done

    // ---- R2: the lifting transposed, d_j's cotangent into sW; trace cotangents cleared
    if (live) {
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
            R acc = R(0);
#pragma unroll
            for (int i = 0; i < NP; ++i) acc = fmaR(cop<R>(OLT + j * NP + i), divb[i], acc);
            sW[(el * NFP + j) * NV + v] = acc;
            sTb[(el * NFP + j) * NV + v] = R(0);
        }
    }
    __syncthreads();

    // ---- R3: the common flux of every solved face point, backwards
#pragma unroll 1
    for (int t = tid; t < nsolved * FP; t += BLOCK) {
        const int c = sList[t / FP], f = t - (t / FP) * FP;
        const int elt = c / 3, kk = c - elt * 3, j = kk * FP + f;
        const int2 Lk = a.links[(int64_t)(e0 + elt) * 3 + kk];
        const int role = link_role(Lk.y);
        const R nx = a.geo[(int64_t)(e0 + elt) * GEO + 4 + 2 * kk];
        const R ny = a.geo[(int64_t)(e0 + elt) * GEO + 5 + 2 * kk];
        R own[NV], oth[NV], UL[NV], UR[NV], Fcb[NV];
#pragma unroll
        for (int w = 0; w < NV; ++w) own[w] = sT[(elt * NFP + j) * NV + w];
        int nl = -1, oj = 0;
        bool cross = false;
        const double* bcd = a.bc_data + (int64_t)(role == ROLE_BOUNDARY ? link_bc(Lk.y) : 0) * 8;
        if (role == ROLE_BOUNDARY) {
            ghost_state(own, nx, ny, link_edge(Lk.y), bcd, g, oth);
        } else {
            nl = Lk.x - e0;
            oj = link_edge(Lk.y) * FP + (P - f);
            if (nl >= 0 && nl < TILE) {
#pragma unroll
                for (int w = 0; w < NV; ++w) oth[w] = sT[(nl * NFP + oj) * NV + w];
            } else {
                load4(a.T_in + trix(Lk.x, oj, NFP), oth);
                nl = -1;
                cross = true;
            }
        }
        const bool right = role == ROLE_RIGHT;
        const R sc = right ? -sEdge[c] : sEdge[c];
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            UL[w] = right ? oth[w] : own[w];
            UR[w] = right ? own[w] : oth[w];
            Fcb[w] = sc * sW[(elt * NFP + j) * NV + w];
        }
        if (nl >= 0) {
            const R scn = -sEdge[nl * 3 + link_edge(Lk.y)];
#pragma unroll
            for (int w = 0; w < NV; ++w) Fcb[w] += scn * sW[(nl * NFP + oj) * NV + w];
        }
        R ULb[NV] = {R(0), R(0), R(0), R(0)}, URb[NV] = {R(0), R(0), R(0), R(0)};
        riemann_by_kind_vjp(a.riemann, UL, UR, nx, ny, g, Fcb, ULb, URb);
        R ownb[NV], othb[NV];
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            ownb[w] = right ? URb[w] : ULb[w];
            othb[w] = right ? ULb[w] : URb[w];
        }
        if (role == ROLE_BOUNDARY) {
            ghost_state_vjp(own, nx, ny, link_edge(Lk.y), bcd, g, othb, ownb);
        } else if (nl >= 0) {
#pragma unroll
            for (int w = 0; w < NV; ++w) sTb[(nl * NFP + oj) * NV + w] = othb[w];
        } else if (cross) {
            store4(a.Tb_out + trix(Lk.x, oj, NFP), othb);
        }
#pragma unroll
        for (int w = 0; w < NV; ++w) sTb[(elt * NFP + j) * NV + w] = ownb[w];
    }
    __syncthreads();

    // ---- R4: volume terms at the solution points, transposed traces; cotangent of q
    if (!live) return;
#pragma unroll 1
    for (int m = 0; m < NP; ++m) {
        R frb = R(0), fsb = R(0);
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            frb = fmaR(cop<R>(O0 + m * CA + i), divb[i], frb);
            fsb = fmaR(cop<R>(O0 + m * CA + NP + i), divb[i], fsb);
        }
        R tb = R(0);
#pragma unroll
        for (int j = 0; j < NFP; ++j) {
            const R im = cop<R>(O0 + m * CA + 2 * NP + j);
            // Fn = -fn (edges 0, 2), +fn (edge 1); d = sc F^I - Fn; fn_j = sum_m interp[j][m] X_m
            // with X = fs, fr + fs, fr on edges 0, 1, 2
            const R db = sW[(el * NFP + j) * NV + v];
            const R fnb = (j >= FP && j < 2 * FP) ? -db : db;
            if (j >= FP) frb = fmaR(im, fnb, frb);
            if (j < 2 * FP) fsb = fmaR(im, fnb, fsb);
            tb = fmaR(im, sTb[(el * NFP + j) * NV + v], tb);
        }
        const R fxb = rx * frb + sx * fsb, fyb = ry * frb + sy * fsb;
        const R um = R(q[m]);
        const R Uq[NV] = {quad_get<0>(um), quad_get<1>(um), quad_get<2>(um), quad_get<3>(um)};
        const R A[NV] = {quad_get<0>(fxb), quad_get<1>(fxb), quad_get<2>(fxb), quad_get<3>(fxb)};
        const R B[NV] = {quad_get<0>(fyb), quad_get<1>(fyb), quad_get<2>(fyb), quad_get<3>(fyb)};
        R Ub[NV] = {R(0), R(0), R(0), R(0)};
        euler_flux_vjp(Uq, g, A, B, Ub);
        const R mine = v == 0 ? Ub[0] : (v == 1 ? Ub[1] : (v == 2 ? Ub[2] : Ub[3]));
        const int64_t gi = base + (int64_t)m * EB * NV;
        double out = qb[m] + double(mine) + double(tb);
        if (MODE == MODE_STAGE1) out += a.U0b[gi];
        a.Ub_out[gi] = out;
    }
}

// Completion of a cotangent: U_bar plus the transposed traces of the
// cross-tile edge cotangents that the last adjoint launch left for this tile.
template <typename R, int P>
__global__ __launch_bounds__(BLOCK) void tri_complete_kernel(int E, const int2* __restrict__ links,
                                                             const double* __restrict__ Ub, const R* __restrict__ Tb,
                                                             double* __restrict__ out, int xcd_groups) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, FP = S::FP, NFP = S::NFP;
    constexpr int OI = ops_offset(P) + TriOps<P>::C;
    const int tid = threadIdx.x, el = tid >> 2, v = tid & 3;
    const int e0 = tile_of_block(xcd_groups) * TILE, e = e0 + el;
    if (e >= E) return;
    const int64_t base = blk(e, 0, NP) + v;
    R tin[NFP];
#pragma unroll
    for (int j = 0; j < NFP; ++j) tin[j] = R(0);
    if (Tb) {
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
            const int2 Lk = links[(int64_t)e * 3 + kk];
            const int nl = Lk.x - e0;
            if (link_role(Lk.y) == ROLE_BOUNDARY || (nl >= 0 && nl < TILE)) continue;
#pragma unroll
            for (int f = 0; f < FP; ++f) tin[kk * FP + f] = Tb[trix(e, kk * FP + f, NFP) + v];
        }
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        R tsum = R(0);
#pragma unroll
        for (int j = 0; j < NFP; ++j) tsum = fmaR(cop<R>(OI + j * NP + i), tin[j], tsum);
        const int64_t gi = base + (int64_t)i * EB * NV;
        out[gi] = Ub[gi] + double(tsum);
    }
}

// ------------------------------------------------------------ adjoint, BR2
// Vector-Jacobian product of one triangle stage with the Navier-Stokes terms
// (tri_visc_stage_kernel), or of its residual alone.  The forward values are
// recomputed as that kernel forms them (traces of U and of its gradient, the
// corrected gradient, the point fluxes, the flux differences), with the point
// fluxes kept apart from the corrected gradient; then, per lane (element,
// variable) unless stated:
//   R1  cotangent of u_out: its partial value plus, on cross-tile edges, the
//       transposed traces of U and of its gradient; the RK combination
//   R2  the lifting transposed: the flux differences' cotangents
//   R4  the divergence and the interpolated normal fluxes transposed; then one
//       lane per point (all four variables): the Euler and viscous flux VJPs,
//       the cotangents of U and of the corrected gradient at the point
//   R5  the corrected gradient's lifted jumps transposed: each edge point's
//       jump cotangent
//   R3  per solved face point: the common flux's cotangent, the Riemann and
//       viscous flux VJPs at the face, the jump cotangents of both sides'
//       corrected gradients; the cotangents of both sides' traces (of U and
//       of the gradient) at their slots (LDS, or the cross-tile neighbour's
//       slot in HBM, 12 words); then the face liftings transposed, gathered
//       over the edge's points
//   R6  the traces and the gradient transposed: cotangent of q
template <typename R, int P> struct TriViscVjpTile {
    static constexpr int NP = (P + 1) * (P + 2) / 2, NFP = 3 * (P + 1);
    static constexpr int XW = 2 * (NP > NFP ? NP : NFP);  // the point fluxes, later the jump cotangents
    static constexpr int words = 4 * NFP + 3 * NP + XW;    // per element and variable
    static constexpr int bytes(int tt) { return tt * words * NV * (int)sizeof(R); }
    static constexpr int value = bytes(64) <= 61440 ? 64 : (bytes(32) <= 61440 ? 32 : 16);
    // gfx950 gives a workgroup up to 160 KiB of LDS: FP64 P4 and P5 take 69
    // and 91 KiB at 16 elements (one workgroup per CU)
    static constexpr bool fits = bytes(16) <= 122880;
};

template <typename R, int P, int MODE, int TT>
__global__ __launch_bounds__(TT * NV) void tri_visc_vjp_kernel(TriVjpArgs<R> a) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, FP = S::FP, NFP = S::NFP;
    constexpr int NB = TT * NV, NWV = NB / 64;
    static_assert(NB % 64 == 0, "tile layout");
    using OP = TriOps<P>;
    constexpr int O0 = ops_offset(P), CA = OP::CA, OI = O0 + OP::C;
    constexpr int XW = TriViscVjpTile<R, P>::XW;
    constexpr double ca = MODE == MODE_STAGE2 ? 0.75 : (MODE == MODE_STAGE3 ? 1.0 / 3.0 : 0.0);
    constexpr double cb = MODE == MODE_STAGE2 ? 0.25 : (MODE == MODE_STAGE3 ? 2.0 / 3.0 : 1.0);
    constexpr double cc = MODE == MODE_STAGE2 ? 0.25 : (MODE == MODE_STAGE3 ? 2.0 / 3.0 : 1.0);
    const int OLT = O0 + a.lift_off;
    const int LFO = lf_offset(P) + a.lf_off;
    __shared__ R sTu[TT * NFP * NV];   // traces of U
    __shared__ R sTx[TT * NFP * NV];   // traces of the gradient, then their cotangents
    __shared__ R sTy[TT * NFP * NV];
    __shared__ R sD[TT * NFP * NV];    // flux differences, their cotangents, then the U traces' cotangents
    __shared__ R sU[TT * NP * NV];     // q, then the flux VJPs' cotangents of U at the points
    __shared__ R sQx[TT * NP * NV];    // corrected gradient, then its cotangent
    __shared__ R sQy[TT * NP * NV];
    __shared__ R sX[TT * XW * NV];     // contravariant point fluxes, then their x / y cotangents;
                                       // later the jump cotangents and the face liftings' seeds
    __shared__ double sred[NWV];
    __shared__ R sEdge[TT * 3];
    __shared__ int2 sLink[TT * 3];
    __shared__ R sNrm[TT * 3 * 2];
    __shared__ unsigned char sList[TT * 3];
    __shared__ int sCnt[NWV];
    R* const sFR = sX;
    R* const sFS = sX + TT * NP * NV;
    R* const sJ = sX;
    R* const sC = sX + TT * NFP * NV;

    const int tid = threadIdx.x, el = tid >> 2, v = tid & 3;
    const int tile = tile_of_block(a.xcd_groups);
    const int e0 = tile * TT, e = e0 + el;
    const bool live = e < a.E;
    const R g = R(a.gamma), mu = R(a.mu), kappa = R(a.kappa), eta = R(a.eta);
    const int64_t base = blk(e, 0, NP) + v;
    auto sl = [](int elx, int j, int w) { return (elx * NFP + j) * NV + w; };
    auto sp = [](int elx, int m, int w) { return (elx * NP + m) * NV + w; };

    // ---- forward 0: state, gradient, traces of U and of the gradient
    R rx = R(0), ry = R(0), sx = R(0), sy = R(0);
    R jwx[3] = {R(0), R(0), R(0)}, jwy[3] = {R(0), R(0), R(0)};  // each edge's jump weights
    int2 L[3] = {make_int2(0, 0), make_int2(0, 0), make_int2(0, 0)};
    if (live) {
#pragma unroll
        for (int m = 0; m < NP; ++m) sU[sp(el, m, v)] = R(a.U_in[base + (int64_t)m * EB * NV]);
        const R* G = a.geo + (int64_t)e * GEO;
        rx = G[0]; ry = G[1]; sx = G[2]; sy = G[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) L[k] = a.links[(int64_t)e * 3 + k];
        R se[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const R ax = k == 0 ? sx : (k == 1 ? rx + sx : rx), ay = k == 0 ? sy : (k == 1 ? ry + sy : ry);
            se[k] = sqrtR(ax * ax + ay * ay);
            const R sg = link_role(L[k].y) == ROLE_RIGHT ? R(-1) : R(1);
            jwx[k] = sg * se[k] * G[4 + 2 * k];
            jwy[k] = sg * se[k] * G[5 + 2 * k];
        }
        if (v < 3) {
            sEdge[el * 3 + v] = se[v];
            sLink[el * 3 + v] = L[v];
            sNrm[(el * 3 + v) * 2] = G[4 + 2 * v];
            sNrm[(el * 3 + v) * 2 + 1] = G[5 + 2 * v];
        }
        R gr[NP], gs[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) { gr[i] = R(0); gs[i] = R(0); }
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const R um = sU[sp(el, m, v)];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                gr[i] = fmaR(cop<R>(O0 + m * CA + i), um, gr[i]);
                gs[i] = fmaR(cop<R>(O0 + m * CA + NP + i), um, gs[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            sQx[sp(el, i, v)] = rx * gr[i] + sx * gs[i];
            sQy[sp(el, i, v)] = ry * gr[i] + sy * gs[i];
        }
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
            R tu = R(0), tx = R(0), ty = R(0);
#pragma unroll
            for (int m = 0; m < NP; ++m) {
                const R w = cop<R>(OI + j * NP + m);
                tu = fmaR(w, sU[sp(el, m, v)], tu);
                tx = fmaR(w, sQx[sp(el, m, v)], tx);
                ty = fmaR(w, sQy[sp(el, m, v)], ty);
            }
            sTu[sl(el, j, v)] = tu;
            sTx[sl(el, j, v)] = tx;
            sTy[sl(el, j, v)] = ty;
        }
    }
    bool owned = false;
    if (live && v < 3) {
        const int2 Lv = v == 0 ? L[0] : (v == 1 ? L[1] : L[2]);
        const int nl = Lv.x - e0;
        owned = !(link_role(Lv.y) == ROLE_RIGHT && nl >= 0 && nl < TT);
    }
    const unsigned long long own_mask = __ballot(owned);
    const int lane = tid & 63, wave = tid >> 6;
    if (lane == 0) sCnt[wave] = __popcll(own_mask);
    __syncthreads();

    // ---- forward 1: corrected gradient
    if (live) {
        R jx[NFP], jy[NFP];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
            const int2 Lk = L[kk];
            const int nl = Lk.x - e0;
            const bool in = nl >= 0 && nl < TT;
            const int ek = link_edge(Lk.y);
#pragma unroll
            for (int f = 0; f < FP; ++f) {
                const int j = kk * FP + f, oj = ek * FP + (P - f);
                const R own = sTu[sl(el, j, v)];
                R oth;
                if (in) oth = sTu[sl(nl, oj, v)];
                else oth = a.T_in[((int64_t)Lk.x * NFP + oj) * TWV + v];
                const R jmp = R(0.5) * (own + oth) - own;
                jx[j] = jwx[kk] * jmp;
                jy[j] = jwy[kk] * jmp;
            }
        }
#pragma unroll 1
        for (int i = 0; i < NP; ++i) {
            R qx = sQx[sp(el, i, v)], qy = sQy[sp(el, i, v)];
#pragma unroll
            for (int j = 0; j < NFP; ++j) {
                const R l = cop<R>(OLT + j * NP + i);
                qx = fmaR(l, jx[j], qx);
                qy = fmaR(l, jy[j], qy);
            }
            sQx[sp(el, i, v)] = qx;
            sQy[sp(el, i, v)] = qy;
        }
    }
    if (owned) {
        int pos = __popcll(own_mask & ((1ull << lane) - 1));
        for (int w = 0; w < wave; ++w) pos += sCnt[w];
        sList[pos] = (unsigned char)(el * 3 + v);
    }
    __syncthreads();

    // ---- forward 2: contravariant total fluxes at the points (points v, v + 4,
    // ... of the element, four variables each), then this lane's divergence
    // and interpolated normal fluxes
    constexpr int MPL = (NP + NV - 1) / NV;
    if (live) {
#pragma unroll 1
        for (int q = 0; q < MPL; ++q) {
            const int m = v + NV * q;
            if (m >= NP) break;
            R U4[NV], X4[NV], Y4[NV];
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                U4[w] = sU[sp(el, m, w)];
                X4[w] = sQx[sp(el, m, w)];
                Y4[w] = sQy[sp(el, m, w)];
            }
            const R inv = fast_rcp(U4[0]);
            const R ux = U4[1] * inv, uy = U4[2] * inv;
            const R p = (g - R(1)) * (U4[3] - R(0.5) * (U4[1] * ux + U4[2] * uy));
            R Fv[NV], Gv[NV];
            visc_flux(U4, X4, Y4, g, mu, kappa, Fv, Gv);
            const R fx[NV] = {U4[1], U4[1] * ux + p, U4[1] * uy, (U4[3] + p) * ux};
            const R fy[NV] = {U4[2], U4[2] * ux, U4[2] * uy + p, (U4[3] + p) * uy};
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                const R Fx = fx[w] - Fv[w], Fy = fy[w] - Gv[w];
                sFR[sp(el, m, w)] = rx * Fx + ry * Fy;
                sFS[sp(el, m, w)] = sx * Fx + sy * Fy;
            }
        }
    }
    __syncthreads();
    R div[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) div[i] = R(0);
    if (live) {
        R fn[NFP];
#pragma unroll
        for (int j = 0; j < NFP; ++j) fn[j] = R(0);
#pragma unroll 1
        for (int m = 0; m < NP; ++m) {
            const R fr = sFR[sp(el, m, v)], fs = sFS[sp(el, m, v)], frs = fr + fs;
#pragma unroll
            for (int i = 0; i < NP; ++i)
                div[i] = fmaR(cop<R>(O0 + m * CA + NP + i), fs, fmaR(cop<R>(O0 + m * CA + i), fr, div[i]));
#pragma unroll
            for (int j = 0; j < NFP; ++j)
                fn[j] = fmaR(cop<R>(O0 + m * CA + 2 * NP + j), j < FP ? fs : (j < 2 * FP ? frs : fr), fn[j]);
        }
#pragma unroll
        for (int j = 0; j < NFP; ++j) sD[sl(el, j, v)] = (j >= FP && j < 2 * FP) ? fn[j] : -fn[j];
    }
    __syncthreads();
    int nsolved = 0;
#pragma unroll
    for (int w = 0; w < NWV; ++w) nsolved += sCnt[w];

    // the values of one solved face point (slot j of element elt, edge kk):
    // both sides' traces, edge factors, the face liftings of the jumps and
    // the averaged state and BR2 gradient
    struct Face {
        int elt, kk, j, ek, oj, nl;
        bool in, right;
        int2 Lk;
        R nx, ny, so, sn;
        R own[NV], oth[NV], Uc[NV], Qbx[NV], Qby[NV];
    };
    auto face_at = [&](int t, Face& F) __attribute__((always_inline)) {
        const int c = sList[t / FP], f = t - (t / FP) * FP;
        F.elt = c / 3;
        F.kk = c - F.elt * 3;
        F.j = F.kk * FP + f;
        F.Lk = sLink[F.elt * 3 + F.kk];
        F.right = link_role(F.Lk.y) == ROLE_RIGHT;
        F.nx = sNrm[(F.elt * 3 + F.kk) * 2];
        F.ny = sNrm[(F.elt * 3 + F.kk) * 2 + 1];
        F.ek = link_edge(F.Lk.y);
        F.oj = F.ek * FP + (P - f);
        F.nl = F.Lk.x - e0;
        F.in = F.nl >= 0 && F.nl < TT;
        const R* Tn = a.T_in + (int64_t)F.Lk.x * NFP * TWV;
        R gxo[NV], gyo[NV], gxn[NV], gyn[NV];
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            F.own[w] = sTu[sl(F.elt, F.j, w)];
            gxo[w] = sTx[sl(F.elt, F.j, w)];
            gyo[w] = sTy[sl(F.elt, F.j, w)];
        }
        if (F.in) {
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                F.oth[w] = sTu[sl(F.nl, F.oj, w)];
                gxn[w] = sTx[sl(F.nl, F.oj, w)];
                gyn[w] = sTy[sl(F.nl, F.oj, w)];
            }
            F.sn = sEdge[F.nl * 3 + F.ek];
        } else {
            load4(Tn + (int64_t)F.oj * TWV, F.oth);
            load4(Tn + (int64_t)F.oj * TWV + NV, gxn);
            load4(Tn + (int64_t)F.oj * TWV + 2 * NV, gyn);
            const R* Gn = a.geo + (int64_t)F.Lk.x * GEO;
            const R ax = F.ek == 0 ? Gn[2] : (F.ek == 1 ? Gn[0] + Gn[2] : Gn[0]);
            const R ay = F.ek == 0 ? Gn[3] : (F.ek == 1 ? Gn[1] + Gn[3] : Gn[1]);
            F.sn = sqrtR(ax * ax + ay * ay);
        }
        F.so = sEdge[F.elt * 3 + F.kk];
        R lo[NV] = {R(0), R(0), R(0), R(0)}, ln[NV] = {R(0), R(0), R(0), R(0)};
#pragma unroll 1
        for (int f2 = 0; f2 < FP; ++f2) {
            const int j2 = F.kk * FP + f2, o2 = F.ek * FP + (P - f2);
            R uo[NV], un[NV];
#pragma unroll
            for (int w = 0; w < NV; ++w) uo[w] = sTu[sl(F.elt, j2, w)];
            if (F.in) {
#pragma unroll
                for (int w = 0; w < NV; ++w) un[w] = sTu[sl(F.nl, o2, w)];
            } else {
                load4(Tn + (int64_t)o2 * TWV, un);
            }
            const R wo = clf<R>(LFO + (F.kk * FP + f) * FP + f2);
            const R wn = clf<R>(LFO + (F.ek * FP + (P - f)) * FP + (P - f2));
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                const R jo = R(0.5) * (uo[w] + un[w]) - uo[w];
                lo[w] = fmaR(wo, jo, lo[w]);
                ln[w] = fmaR(wn, jo, ln[w]);
            }
        }
        const R sg = F.right ? R(-1) : R(1);
        const R nox = sg * F.nx, noy = sg * F.ny;
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            F.Uc[w] = R(0.5) * (F.own[w] + F.oth[w]);
            F.Qbx[w] = R(0.5) * ((gxo[w] + eta * F.so * nox * lo[w]) + (gxn[w] + eta * F.sn * nox * ln[w]));
            F.Qby[w] = R(0.5) * ((gyo[w] + eta * F.so * noy * lo[w]) + (gyn[w] + eta * F.sn * noy * ln[w]));
        }
    };

    // ---- forward 3: flux differences of both sides of every solved face point
#pragma unroll 1
    for (int t = tid; t < nsolved * FP; t += NB) {
        Face F;
        face_at(t, F);
        R UL[NV], UR[NV], Fc[NV], Fv[NV], Gv[NV];
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            UL[w] = F.right ? F.oth[w] : F.own[w];
            UR[w] = F.right ? F.own[w] : F.oth[w];
        }
        visc_flux(F.Uc, F.Qbx, F.Qby, g, mu, kappa, Fv, Gv);
        riemann_by_kind(a.riemann, UL, UR, F.nx, F.ny, g, Fc);
        const R sc = F.right ? -F.so : F.so;
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            const R fi = Fc[w] - (Fv[w] * F.nx + Gv[w] * F.ny);
            sD[sl(F.elt, F.j, w)] = sc * fi - sD[sl(F.elt, F.j, w)];
            if (F.in) sD[sl(F.nl, F.oj, w)] = -F.sn * fi - sD[sl(F.nl, F.oj, w)];
        }
    }
    __syncthreads();

    // ---- forward 4 and R1: residual; cotangent of u_out, RK combination
    double rsum = 0.0;
    R divb[NP];
    double qb[NP];
    if (live) {
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
            const R d = sD[sl(el, j, v)];
#pragma unroll
            for (int i = 0; i < NP; ++i) div[i] = fmaR(cop<R>(OLT + j * NP + i), d, div[i]);
        }
        double unb[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) unb[i] = a.Ub_in[base + (int64_t)i * EB * NV];
        if (MODE != MODE_RESIDUAL && a.Tb_in) {
            // edge cotangents of u_out's traces (U, gradient) left on this element's cross-tile edges
            R tu[NFP], tx[NFP], ty[NFP];
            bool any = false;
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                const int nl = L[kk].x - e0;
                const bool cross = !(nl >= 0 && nl < TT);
                any = any || cross;
#pragma unroll
                for (int f = 0; f < FP; ++f) {
                    const int j = kk * FP + f;
                    const R* Tb = a.Tb_in + ((int64_t)e * NFP + j) * TWV;
                    tu[j] = cross ? Tb[v] : R(0);
                    tx[j] = cross ? Tb[NV + v] : R(0);
                    ty[j] = cross ? Tb[2 * NV + v] : R(0);
                }
            }
            if (any) {
                R grb[NP], gsb[NP];
#pragma unroll
                for (int m = 0; m < NP; ++m) {
                    R su = R(0), qx = R(0), qy = R(0);
#pragma unroll
                    for (int j = 0; j < NFP; ++j) {
                        const R w = cop<R>(OI + j * NP + m);
                        su = fmaR(w, tu[j], su);
                        qx = fmaR(w, tx[j], qx);
                        qy = fmaR(w, ty[j], qy);
                    }
                    unb[m] += double(su);
                    grb[m] = rx * qx + ry * qy;
                    gsb[m] = sx * qx + sy * qy;
                }
#pragma unroll
                for (int m = 0; m < NP; ++m) {
                    R acc = R(0);
#pragma unroll
                    for (int i = 0; i < NP; ++i)
                        acc = fmaR(cop<R>(O0 + m * CA + NP + i), gsb[i], fmaR(cop<R>(O0 + m * CA + i), grb[i], acc));
                    unb[m] += double(acc);
                }
            }
        }
        const double dt = MODE == MODE_RESIDUAL ? 0.0 : *a.dt_ptr;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int64_t gi = base + (int64_t)i * EB * NV;
            double rb;
            if (MODE == MODE_RESIDUAL) {
                rb = unb[i];
                qb[i] = 0.0;
            } else {
                if (MODE == MODE_STAGE3) a.U0b[gi] = ca * unb[i];
                if (MODE == MODE_STAGE2) a.U0b[gi] += ca * unb[i];
                qb[i] = cb * unb[i];
                rb = cc * dt * unb[i];
                rsum += cc * (-double(div[i])) * unb[i];
            }
            divb[i] = R(-rb);  // R = -div
        }
    } else {
#pragma unroll
        for (int i = 0; i < NP; ++i) { divb[i] = R(0); qb[i] = 0.0; }
    }
    if (a.dtb_partial) {
        const double s2 = block_reduce_n<OpSum, NWV>(rsum, sred);
        if (tid == 0) a.dtb_partial[tile] = s2;
    }
    __syncthreads();

    // ---- R2, R4: the lifting transposed; the divergence and normal fluxes
    // transposed to the x / y flux cotangents at the points
    if (live) {
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
            R acc = R(0);
#pragma unroll
            for (int i = 0; i < NP; ++i) acc = fmaR(cop<R>(OLT + j * NP + i), divb[i], acc);
            sD[sl(el, j, v)] = acc;
        }
#pragma unroll 1
        for (int m = 0; m < NP; ++m) {
            R frb = R(0), fsb = R(0);
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                frb = fmaR(cop<R>(O0 + m * CA + i), divb[i], frb);
                fsb = fmaR(cop<R>(O0 + m * CA + NP + i), divb[i], fsb);
            }
#pragma unroll
            for (int j = 0; j < NFP; ++j) {
                // d = sc F^I - Fn with Fn = -fn (edges 0, 2), +fn (edge 1); fn_j
                // interpolates fs, fr + fs, fr on edges 0, 1, 2
                const R im = cop<R>(O0 + m * CA + 2 * NP + j);
                const R db = sD[sl(el, j, v)];
                const R fnb = (j >= FP && j < 2 * FP) ? -db : db;
                if (j >= FP) frb = fmaR(im, fnb, frb);
                if (j < 2 * FP) fsb = fmaR(im, fnb, fsb);
            }
            sFR[sp(el, m, v)] = rx * frb + sx * fsb;  // cotangent of F_x = f_x - F_v
            sFS[sp(el, m, v)] = ry * frb + sy * fsb;  // and of F_y
        }
    }
    __syncthreads();
    if (live) {
#pragma unroll 1
        for (int q = 0; q < MPL; ++q) {
            const int m = v + NV * q;
            if (m >= NP) break;
            R U4[NV], X4[NV], Y4[NV], A4[NV], B4[NV], Fb[NV], Gb[NV];
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                U4[w] = sU[sp(el, m, w)];
                X4[w] = sQx[sp(el, m, w)];
                Y4[w] = sQy[sp(el, m, w)];
                A4[w] = sFR[sp(el, m, w)];
                B4[w] = sFS[sp(el, m, w)];
                Fb[w] = -A4[w];
                Gb[w] = -B4[w];
            }
            R Ub4[NV] = {R(0), R(0), R(0), R(0)}, Xb4[NV] = {R(0), R(0), R(0), R(0)},
              Yb4[NV] = {R(0), R(0), R(0), R(0)};
            euler_flux_vjp(U4, g, A4, B4, Ub4);
            visc_flux_vjp(U4, X4, Y4, g, mu, kappa, Fb, Gb, Ub4, Xb4, Yb4);
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                sU[sp(el, m, w)] = Ub4[w];
                sQx[sp(el, m, w)] = Xb4[w];
                sQy[sp(el, m, w)] = Yb4[w];
            }
        }
    }
    __syncthreads();

    // ---- R5: the corrected gradient's lifted jumps transposed
    if (live) {
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
            R bx = R(0), by = R(0);
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const R l = cop<R>(OLT + j * NP + i);
                bx = fmaR(l, sQx[sp(el, i, v)], bx);
                by = fmaR(l, sQy[sp(el, i, v)], by);
            }
            const int kk = j / FP;
            sJ[sl(el, j, v)] = (kk == 0 ? jwx[0] : (kk == 1 ? jwx[1] : jwx[2])) * bx +
                               (kk == 0 ? jwy[0] : (kk == 1 ? jwy[1] : jwy[2])) * by;
        }
    }
    __syncthreads();

    // ---- R3: every solved face point backwards
#pragma unroll 1
    for (int t = tid; t < nsolved * FP; t += NB) {
        Face F;
        face_at(t, F);
        R UL[NV], UR[NV], Fcb[NV], Fvb[NV], Gvb[NV];
        const R sc = F.right ? -F.so : F.so;
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            UL[w] = F.right ? F.oth[w] : F.own[w];
            UR[w] = F.right ? F.own[w] : F.oth[w];
            R fb = sc * sD[sl(F.elt, F.j, w)];
            if (F.in) fb += -F.sn * sD[sl(F.nl, F.oj, w)];
            Fcb[w] = fb;
            Fvb[w] = -fb * F.nx;
            Gvb[w] = -fb * F.ny;
        }
        R ULb[NV] = {R(0), R(0), R(0), R(0)}, URb[NV] = {R(0), R(0), R(0), R(0)};
        riemann_by_kind_vjp(a.riemann, UL, UR, F.nx, F.ny, g, Fcb, ULb, URb);
        R Ucb[NV] = {R(0), R(0), R(0), R(0)}, Qxb[NV] = {R(0), R(0), R(0), R(0)}, Qyb[NV] = {R(0), R(0), R(0), R(0)};
        visc_flux_vjp(F.Uc, F.Qbx, F.Qby, g, mu, kappa, Fvb, Gvb, Ucb, Qxb, Qyb);
        const R sg = F.right ? R(-1) : R(1);
        const R nox = sg * F.nx, noy = sg * F.ny;
        R ownb[NV], othb[NV];
#pragma unroll
        for (int w = 0; w < NV; ++w) {
            const R jo = sJ[sl(F.elt, F.j, w)], jn = F.in ? sJ[sl(F.nl, F.oj, w)] : R(0);
            ownb[w] = (F.right ? URb[w] : ULb[w]) + R(0.5) * Ucb[w] - R(0.5) * jo + R(0.5) * jn;
            othb[w] = (F.right ? ULb[w] : URb[w]) + R(0.5) * Ucb[w] + R(0.5) * jo - R(0.5) * jn;
            Qxb[w] *= R(0.5);
            Qyb[w] *= R(0.5);
            sC[sl(F.elt, F.j, w)] = eta * (nox * Qxb[w] + noy * Qyb[w]);
            sD[sl(F.elt, F.j, w)] = ownb[w];
            sTx[sl(F.elt, F.j, w)] = Qxb[w];
            sTy[sl(F.elt, F.j, w)] = Qyb[w];
        }
        if (F.in) {
#pragma unroll
            for (int w = 0; w < NV; ++w) {
                sD[sl(F.nl, F.oj, w)] = othb[w];
                sTx[sl(F.nl, F.oj, w)] = Qxb[w];
                sTy[sl(F.nl, F.oj, w)] = Qyb[w];
            }
        } else {
            R* Tb = a.Tb_out + ((int64_t)F.Lk.x * NFP + F.oj) * TWV;
            store4(Tb, othb);
            store4(Tb + NV, Qxb);
            store4(Tb + 2 * NV, Qyb);
        }
    }
    __syncthreads();
    // the face liftings transposed: the jump at this point from every point's
    // lifting cotangents (the seeds c, times each side's edge factor)
#pragma unroll 1
    for (int t = tid; t < nsolved * FP; t += NB) {
        const int c = sList[t / FP], f = t - (t / FP) * FP;
        const int elt = c / 3, kk = c - elt * 3, j = kk * FP + f;
        const int2 Lk = sLink[elt * 3 + kk];
        const int ek = link_edge(Lk.y), oj = ek * FP + (P - f), nl = Lk.x - e0;
        const bool in = nl >= 0 && nl < TT;
        const R so = sEdge[elt * 3 + kk];
        R sn;
        if (in) {
            sn = sEdge[nl * 3 + ek];
        } else {
            const R* Gn = a.geo + (int64_t)Lk.x * GEO;
            const R ax = ek == 0 ? Gn[2] : (ek == 1 ? Gn[0] + Gn[2] : Gn[0]);
            const R ay = ek == 0 ? Gn[3] : (ek == 1 ? Gn[1] + Gn[3] : Gn[1]);
            sn = sqrtR(ax * ax + ay * ay);
        }
        R jb[NV] = {R(0), R(0), R(0), R(0)};
#pragma unroll 1
        for (int f1 = 0; f1 < FP; ++f1) {
            const R w1 = so * clf<R>(LFO + (kk * FP + f1) * FP + f) +
                         sn * clf<R>(LFO + (ek * FP + (P - f1)) * FP + (P - f));
#pragma unroll
            for (int w = 0; w < NV; ++w) jb[w] = fmaR(w1, sC[sl(elt, kk * FP + f1, w)], jb[w]);
        }
        // jump = (u_n - u_o) / 2 at this point
#pragma unroll
        for (int w = 0; w < NV; ++w) sD[sl(elt, j, w)] -= R(0.5) * jb[w];
        if (in) {
#pragma unroll
            for (int w = 0; w < NV; ++w) sD[sl(nl, oj, w)] += R(0.5) * jb[w];
        } else {
            R* Tb = a.Tb_out + ((int64_t)Lk.x * NFP + oj) * TWV;
#pragma unroll
            for (int w = 0; w < NV; ++w) Tb[w] += R(0.5) * jb[w];
        }
    }
    __syncthreads();

    // ---- R6: traces and gradient transposed; cotangent of q
    if (!live) return;
    R grb[NP], gsb[NP];
    double out[NP];
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        R su = R(0), qx = sQx[sp(el, m, v)], qy = sQy[sp(el, m, v)];
#pragma unroll
        for (int j = 0; j < NFP; ++j) {
            const R w = cop<R>(OI + j * NP + m);
            su = fmaR(w, sD[sl(el, j, v)], su);
            qx = fmaR(w, sTx[sl(el, j, v)], qx);
            qy = fmaR(w, sTy[sl(el, j, v)], qy);
        }
        out[m] = qb[m] + double(sU[sp(el, m, v)]) + double(su);
        grb[m] = rx * qx + ry * qy;
        gsb[m] = sx * qx + sy * qy;
    }
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        R acc = R(0);
#pragma unroll
        for (int i = 0; i < NP; ++i)
            acc = fmaR(cop<R>(O0 + m * CA + NP + i), gsb[i], fmaR(cop<R>(O0 + m * CA + i), grb[i], acc));
        const int64_t gi = base + (int64_t)m * EB * NV;
        double o = out[m] + double(acc);
        if (MODE == MODE_STAGE1) o += a.U0b[gi];
        a.Ub_out[gi] = o;
    }
}

// Completion of a cotangent with the Navier-Stokes terms: U_bar plus the
// transposed traces of U and of its gradient for the cross-tile edge
// cotangents (12 words per point) that the last adjoint launch left.
template <typename R, int P, int TT>
__global__ __launch_bounds__(TT * NV) void tri_visc_complete_kernel(int E, const int2* __restrict__ links,
                                                                    const R* __restrict__ geo,
                                                                    const double* __restrict__ Ub,
                                                                    const R* __restrict__ Tb,
                                                                    double* __restrict__ out, int xcd_groups) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, FP = S::FP, NFP = S::NFP;
    constexpr int O0 = ops_offset(P), CA = TriOps<P>::CA, OI = O0 + TriOps<P>::C;
    const int tid = threadIdx.x, el = tid >> 2, v = tid & 3;
    const int e0 = tile_of_block(xcd_groups) * TT, e = e0 + el;
    if (e >= E) return;
    const int64_t base = blk(e, 0, NP) + v;
    double acc[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) acc[i] = Ub[base + (int64_t)i * EB * NV];
    if (Tb) {
        R tu[NFP], tx[NFP], ty[NFP];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
            const int2 Lk = links[(int64_t)e * 3 + kk];
            const int nl = Lk.x - e0;
            const bool cross = !(nl >= 0 && nl < TT);
#pragma unroll
            for (int f = 0; f < FP; ++f) {
                const int j = kk * FP + f;
                const R* T = Tb + ((int64_t)e * NFP + j) * TWV;
                tu[j] = cross ? T[v] : R(0);
                tx[j] = cross ? T[NV + v] : R(0);
                ty[j] = cross ? T[2 * NV + v] : R(0);
            }
        }
        const R* G = geo + (int64_t)e * GEO;
        const R rx = G[0], ry = G[1], sx = G[2], sy = G[3];
        R grb[NP], gsb[NP];
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            R su = R(0), qx = R(0), qy = R(0);
#pragma unroll
            for (int j = 0; j < NFP; ++j) {
                const R w = cop<R>(OI + j * NP + m);
                su = fmaR(w, tu[j], su);
                qx = fmaR(w, tx[j], qx);
                qy = fmaR(w, ty[j], qy);
            }
            acc[m] += double(su);
            grb[m] = rx * qx + ry * qy;
            gsb[m] = sx * qx + sy * qy;
        }
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            R s = R(0);
#pragma unroll
            for (int i = 0; i < NP; ++i)
                s = fmaR(cop<R>(O0 + m * CA + NP + i), gsb[i], fmaR(cop<R>(O0 + m * CA + i), grb[i], s));
            acc[m] += double(s);
        }
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) out[base + (int64_t)i * EB * NV] = acc[i];
}

// Traces of U on every edge (after a new state is set).
template <typename R, int P>
__global__ __launch_bounds__(BLOCK) void tri_trace_kernel(int E, const double* U, R* T_out) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, NFP = S::NFP, OI = ops_offset(P) + TriOps<P>::C;
    const int e = blockIdx.x * TILE + (threadIdx.x >> 2), v = threadIdx.x & 3;
    if (e >= E) return;
    R u[NP];
#pragma unroll
    for (int m = 0; m < NP; ++m) u[m] = R(U[blk(e, m, NP) + v]);
#pragma unroll 1
    for (int j = 0; j < NFP; ++j) {
        R t = R(0);
#pragma unroll
        for (int m = 0; m < NP; ++m) t = fmaR(cop<R>(OI + j * NP + m), u[m], t);
        T_out[trix(e, j, NFP) + v] = t;
    }
}

// Caller AoS [E][Np][4] (caller numbering) <-> blocked device layout.
__global__ void tri_to_blk_kernel(const double* __restrict__ aos, double* __restrict__ out, int E, int NP,
                                  const int* __restrict__ perm) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)E * NP) return;
    const int e = (int)(t / NP), i = (int)(t - (int64_t)e * NP);
    const int src = perm ? perm[e] : e;
    double q[NV];
    load4(aos + ((int64_t)src * NP + i) * NV, q);
    store4(out + blk(e, i, NP), q);
}

__global__ void tri_from_blk_kernel(const double* __restrict__ in, double* __restrict__ aos, int E, int NP,
                                    const int* __restrict__ perm) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)E * NP) return;
    const int e = (int)(t / NP), i = (int)(t - (int64_t)e * NP);
    const int dst = perm ? perm[e] : e;
    double q[NV];
    load4(in + blk(e, i, NP), q);
    store4(aos + ((int64_t)dst * NP + i) * NV, q);
}

// CFL step per point of the blocked state, minimum per workgroup.  With the
// viscous terms (nu_k = mu max(4/3, gamma/Pr), else 0) the diffusive rate
// (p+1)^4 nu_k / (rho d^2) joins the convective (2p+1)(|u|+c)/d: the
// spectral radius of the BR2 operator on triangles is about 1.3 (p+1)^4
// nu / d^2 (P2-P4; 1.7 at P1), d the inscribed-circle diameter.
__global__ __launch_bounds__(BLOCK) void tri_dt_kernel(const double* __restrict__ U, const double* __restrict__ hcell,
                                                       int E, int NP, double gamma, double cfl, double cf,
                                                       double visc, double* __restrict__ partial) {
    __shared__ double sRed[BLOCK / 64];
    const int64_t n = (int64_t)E * NP;
    double mn = 1e20;
    for (int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x; t < n; t += (int64_t)gridDim.x * BLOCK) {
        const int e = (int)(t / NP), i = (int)(t - (int64_t)e * NP);
        double q[NV];
        load4(U + blk(e, i, NP), q);
        const double u = q[1] / q[0], w = q[2] / q[0];
        const double p = (gamma - 1.0) * q[0] * (q[3] / q[0] - 0.5 * (u * u + w * w));
        const double ws = sqrt(u * u + w * w) + sqrt(gamma * fmax(p, 1e-10) / fmax(q[0], 1e-10));
        const double h = hcell[e];
        const double d = cfl / (cf * ws / h + visc / (fmax(q[0], 1e-10) * h * h));
        mn = d < mn ? d : mn;
    }
    const double r = block_reduce<OpMin>(mn, sRed);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

template <typename R, int P>
hipError_t tri_launch_p(int mode, const TriArgs<R>& a, hipStream_t s) {
    const int grid = (std::max(a.E, 1) + TILE - 1) / TILE;
    switch (mode) {
    case MODE_RESIDUAL: klaunch(tri_stage_kernel<R, P, MODE_RESIDUAL>, grid, dim3(BLOCK), 0, s, a); break;
    case MODE_STAGE1: klaunch(tri_stage_kernel<R, P, MODE_STAGE1>, grid, dim3(BLOCK), 0, s, a); break;
    case MODE_STAGE2: klaunch(tri_stage_kernel<R, P, MODE_STAGE2>, grid, dim3(BLOCK), 0, s, a); break;
    default: klaunch(tri_stage_kernel<R, P, MODE_STAGE3>, grid, dim3(BLOCK), 0, s, a); break;
    }
    return hipGetLastError();
}

template <typename R>
hipError_t tri_launch(int P, int mode, const TriArgs<R>& a, hipStream_t s) {
    switch (P) {
    case 1: return tri_launch_p<R, 1>(mode, a, s);
    case 2: return tri_launch_p<R, 2>(mode, a, s);
    case 3: return tri_launch_p<R, 3>(mode, a, s);
    case 4: return tri_launch_p<R, 4>(mode, a, s);
    default: return tri_launch_p<R, 5>(mode, a, s);
    }
}

template <typename R, int P>
hipError_t tri_visc_launch_p(int mode, const TriArgs<R>& a, hipStream_t s) {
    constexpr int TT = TriViscTile<R, P>::value;
    const int grid = (std::max(a.E, 1) + TT - 1) / TT;
    const dim3 blk(TT * NV);
    switch (mode) {
    case MODE_RESIDUAL: klaunch(tri_visc_stage_kernel<R, P, MODE_RESIDUAL, TT>, grid, blk, 0, s, a); break;
    case MODE_STAGE1: klaunch(tri_visc_stage_kernel<R, P, MODE_STAGE1, TT>, grid, blk, 0, s, a); break;
    case MODE_STAGE2: klaunch(tri_visc_stage_kernel<R, P, MODE_STAGE2, TT>, grid, blk, 0, s, a); break;
    default: klaunch(tri_visc_stage_kernel<R, P, MODE_STAGE3, TT>, grid, blk, 0, s, a); break;
    }
    return hipGetLastError();
}

template <typename R>
hipError_t tri_visc_launch(int P, int mode, const TriArgs<R>& a, hipStream_t s) {
    switch (P) {
    case 1: return tri_visc_launch_p<R, 1>(mode, a, s);
    case 2: return tri_visc_launch_p<R, 2>(mode, a, s);
    case 3: return tri_visc_launch_p<R, 3>(mode, a, s);
    case 4: return tri_visc_launch_p<R, 4>(mode, a, s);
    default: return tri_visc_launch_p<R, 5>(mode, a, s);
    }
}

template <typename R>
hipError_t tri_visc_trace(int P, int E, const double* U, const R* geo, R* T, hipStream_t s) {
    const int grid = (std::max(E, 1) + TILE - 1) / TILE;
    switch (P) {
    case 1: tri_visc_trace_kernel<R, 1><<<grid, BLOCK, 0, s>>>(E, U, geo, T); break;
    case 2: tri_visc_trace_kernel<R, 2><<<grid, BLOCK, 0, s>>>(E, U, geo, T); break;
    case 3: tri_visc_trace_kernel<R, 3><<<grid, BLOCK, 0, s>>>(E, U, geo, T); break;
    case 4: tri_visc_trace_kernel<R, 4><<<grid, BLOCK, 0, s>>>(E, U, geo, T); break;
    default: tri_visc_trace_kernel<R, 5><<<grid, BLOCK, 0, s>>>(E, U, geo, T); break;
    }
    return hipGetLastError();
}

int tri_visc_tile(int P, bool fp64) {
    switch (P) {
    case 1: return fp64 ? TriViscTile<double, 1>::value : TriViscTile<float, 1>::value;
    case 2: return fp64 ? TriViscTile<double, 2>::value : TriViscTile<float, 2>::value;
    case 3: return fp64 ? TriViscTile<double, 3>::value : TriViscTile<float, 3>::value;
    case 4: return fp64 ? TriViscTile<double, 4>::value : TriViscTile<float, 4>::value;
    default: return fp64 ? TriViscTile<double, 5>::value : TriViscTile<float, 5>::value;
    }
}

template <typename R, int P>
hipError_t tri_vjp_p(int mode, const TriVjpArgs<R>& a, hipStream_t s) {
    const int grid = (std::max(a.E, 1) + TILE - 1) / TILE;
    switch (mode) {
    case MODE_RESIDUAL: tri_vjp_kernel<R, P, MODE_RESIDUAL><<<grid, BLOCK, 0, s>>>(a); break;
    case MODE_STAGE1: tri_vjp_kernel<R, P, MODE_STAGE1><<<grid, BLOCK, 0, s>>>(a); break;
    case MODE_STAGE2: tri_vjp_kernel<R, P, MODE_STAGE2><<<grid, BLOCK, 0, s>>>(a); break;
    default: tri_vjp_kernel<R, P, MODE_STAGE3><<<grid, BLOCK, 0, s>>>(a); break;
    }
    return hipGetLastError();
}

template <typename R>
hipError_t tri_vjp(int P, int mode, const TriVjpArgs<R>& a, hipStream_t s) {
    switch (P) {
    case 1: return tri_vjp_p<R, 1>(mode, a, s);
    case 2: return tri_vjp_p<R, 2>(mode, a, s);
    case 3: return tri_vjp_p<R, 3>(mode, a, s);
    case 4: return tri_vjp_p<R, 4>(mode, a, s);
    default: return tri_vjp_p<R, 5>(mode, a, s);
    }
}

template <typename R>
hipError_t tri_complete(int P, int E, const int2* links, const double* Ub, const R* Tb, double* out, int xcd,
                        hipStream_t s) {
    const int grid = (std::max(E, 1) + TILE - 1) / TILE;
    switch (P) {
    case 1: tri_complete_kernel<R, 1><<<grid, BLOCK, 0, s>>>(E, links, Ub, Tb, out, xcd); break;
    case 2: tri_complete_kernel<R, 2><<<grid, BLOCK, 0, s>>>(E, links, Ub, Tb, out, xcd); break;
    case 3: tri_complete_kernel<R, 3><<<grid, BLOCK, 0, s>>>(E, links, Ub, Tb, out, xcd); break;
    case 4: tri_complete_kernel<R, 4><<<grid, BLOCK, 0, s>>>(E, links, Ub, Tb, out, xcd); break;
    default: tri_complete_kernel<R, 5><<<grid, BLOCK, 0, s>>>(E, links, Ub, Tb, out, xcd); break;
    }
    return hipGetLastError();
}

// The BR2 adjoint's elements per workgroup for an order and precision (0
// where its LDS arrays would exceed 120 KiB at 16 elements: none at P1-P5).
int tri_visc_vjp_tile(int P, bool fp64) {
    switch (P) {
#define ZJ_TV(PP)                                                                                           \
    case PP:                                                                                                \
        return fp64 ? (TriViscVjpTile<double, PP>::fits ? TriViscVjpTile<double, PP>::value : 0)            \
                    : (TriViscVjpTile<float, PP>::fits ? TriViscVjpTile<float, PP>::value : 0);
        ZJ_TV(1) ZJ_TV(2) ZJ_TV(3) ZJ_TV(4) ZJ_TV(5)
#undef ZJ_TV
    default: return 0;
    }
}

template <typename R, int P>
hipError_t tri_visc_vjp_p(int mode, const TriVjpArgs<R>& a, hipStream_t s) {
    if constexpr (!TriViscVjpTile<R, P>::fits) {
        return hipErrorInvalidValue;
    } else {
        constexpr int TT = TriViscVjpTile<R, P>::value;
        const int grid = (std::max(a.E, 1) + TT - 1) / TT;
        const dim3 blk(TT * NV);
        switch (mode) {
        case MODE_RESIDUAL: klaunch(tri_visc_vjp_kernel<R, P, MODE_RESIDUAL, TT>, grid, blk, 0, s, a); break;
        case MODE_STAGE1: klaunch(tri_visc_vjp_kernel<R, P, MODE_STAGE1, TT>, grid, blk, 0, s, a); break;
        case MODE_STAGE2: klaunch(tri_visc_vjp_kernel<R, P, MODE_STAGE2, TT>, grid, blk, 0, s, a); break;
        default: klaunch(tri_visc_vjp_kernel<R, P, MODE_STAGE3, TT>, grid, blk, 0, s, a); break;
        }
        return hipGetLastError();
    }
}

template <typename R>
hipError_t tri_visc_vjp(int P, int mode, const TriVjpArgs<R>& a, hipStream_t s) {
    switch (P) {
    case 1: return tri_visc_vjp_p<R, 1>(mode, a, s);
    case 2: return tri_visc_vjp_p<R, 2>(mode, a, s);
    case 3: return tri_visc_vjp_p<R, 3>(mode, a, s);
    case 4: return tri_visc_vjp_p<R, 4>(mode, a, s);
    default: return tri_visc_vjp_p<R, 5>(mode, a, s);
    }
}

template <typename R, int P>
hipError_t tri_visc_complete_p(int E, const int2* links, const R* geo, const double* Ub, const R* Tb, double* out,
                               int xcd, hipStream_t s) {
    if constexpr (!TriViscVjpTile<R, P>::fits) {
        return hipErrorInvalidValue;
    } else {
        constexpr int TT = TriViscVjpTile<R, P>::value;
        const int grid = (std::max(E, 1) + TT - 1) / TT;
        tri_visc_complete_kernel<R, P, TT><<<grid, TT * NV, 0, s>>>(E, links, geo, Ub, Tb, out, xcd);
        return hipGetLastError();
    }
}

template <typename R>
hipError_t tri_visc_complete(int P, int E, const int2* links, const R* geo, const double* Ub, const R* Tb, double* out,
                             int xcd, hipStream_t s) {
    switch (P) {
    case 1: return tri_visc_complete_p<R, 1>(E, links, geo, Ub, Tb, out, xcd, s);
    case 2: return tri_visc_complete_p<R, 2>(E, links, geo, Ub, Tb, out, xcd, s);
    case 3: return tri_visc_complete_p<R, 3>(E, links, geo, Ub, Tb, out, xcd, s);
    case 4: return tri_visc_complete_p<R, 4>(E, links, geo, Ub, Tb, out, xcd, s);
    default: return tri_visc_complete_p<R, 5>(E, links, geo, Ub, Tb, out, xcd, s);
    }
}

template <typename R>
hipError_t tri_trace(int P, int E, const double* U, R* T, hipStream_t s) {
    const int grid = (std::max(E, 1) + TILE - 1) / TILE;
    switch (P) {
    case 1: tri_trace_kernel<R, 1><<<grid, BLOCK, 0, s>>>(E, U, T); break;
    case 2: tri_trace_kernel<R, 2><<<grid, BLOCK, 0, s>>>(E, U, T); break;
    case 3: tri_trace_kernel<R, 3><<<grid, BLOCK, 0, s>>>(E, U, T); break;
    case 4: tri_trace_kernel<R, 4><<<grid, BLOCK, 0, s>>>(E, U, T); break;
    default: tri_trace_kernel<R, 5><<<grid, BLOCK, 0, s>>>(E, U, T); break;
    }
    return hipGetLastError();
}

hipError_t tri_convert(bool to_blk, const double* src, double* dst, int E, int NP, const int* perm, hipStream_t s) {
    const int64_t n = (int64_t)E * NP;
    const int grid = (int)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK);
    if (to_blk)
        tri_to_blk_kernel<<<grid, BLOCK, 0, s>>>(src, dst, E, NP, perm);
    else
        tri_from_blk_kernel<<<grid, BLOCK, 0, s>>>(src, dst, E, NP, perm);
    return hipGetLastError();
}

// Operator tables of every order into this device's constant memory, once.
std::mutex g_ops_mutex;
uint64_t g_ops_loaded = 0;  // bit d: device d holds the tables

hipError_t load_tri_ops(int device) {
    std::lock_guard<std::mutex> lock(g_ops_mutex);
    if (device >= 0 && device < 64 && (g_ops_loaded >> device) & 1) return hipSuccess;
    std::vector<double> all((size_t)OPS_TOTAL);
    for (int P = 1; P <= 5; ++P) {
        tri::Operators O, Osd, Ohu;
        if (!tri::operators(P, O) || !tri::operators(P, Osd, ZJ_CORR_SD) || !tri::operators(P, Ohu, ZJ_CORR_HU))
            return hipErrorInvalidValue;
        double* d = all.data() + ops_offset(P);
        const int np = O.np, nf = O.nf, ca = 2 * np + nf;   // TriOps<P> layout
        for (int m = 0; m < np; ++m) {
            for (int i = 0; i < np; ++i) {
                d[m * ca + i] = O.Dr[(size_t)i * np + m];
                d[m * ca + np + i] = O.Ds[(size_t)i * np + m];
            }
            for (int jj = 0; jj < nf; ++jj) d[m * ca + 2 * np + jj] = O.interp[(size_t)jj * np + m];
        }
        double* b = d + (size_t)np * ca;
        for (int jj = 0; jj < nf; ++jj)
            for (int i = 0; i < np; ++i) b[(size_t)jj * np + i] = O.lift[(size_t)i * nf + jj];
        std::copy(O.interp.begin(), O.interp.end(), b + (size_t)nf * np);
        for (int k = 0; k < 2; ++k) {
            const tri::Operators& F = k ? Ohu : Osd;
            double* dl = d + lift_offset(P, k ? ZJ_CORR_HU : ZJ_CORR_SD);
            for (int jj = 0; jj < nf; ++jj)
                for (int i = 0; i < np; ++i) dl[(size_t)jj * np + i] = F.lift[(size_t)i * nf + jj];
        }
    }
    std::vector<double> lf((size_t)LF_TOTAL, 0.0);
    for (int P = 1; P <= 5; ++P) {
        for (int fam = 0; fam < 3; ++fam) {
            tri::Operators O;
            if (!tri::operators(P, O, fam == 0 ? ZJ_CORR_DG : (fam == 1 ? ZJ_CORR_SD : ZJ_CORR_HU)))
                return hipErrorInvalidValue;
            const int np = O.np, nf = O.nf, fp = P + 1;
            double* d = lf.data() + lf_offset(P) + (size_t)fam * lf_words(P);
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < fp; ++a)
                    for (int b = 0; b < fp; ++b) {
                        double acc = 0.0;
                        for (int i = 0; i < np; ++i)
                            acc += O.interp[(size_t)(k * fp + a) * np + i] * O.lift[(size_t)i * nf + k * fp + b];
                        d[(k * fp + a) * fp + b] = acc;
                    }
        }
    }
    std::vector<float> allf(all.begin(), all.end()), lff(lf.begin(), lf.end());
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_ops64), all.data(), all.size() * sizeof(double), 0,
                                     hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpyToSymbol(HIP_SYMBOL(c_ops32), allf.data(), allf.size() * sizeof(float), 0,
                              hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpyToSymbol(HIP_SYMBOL(c_lf64), lf.data(), lf.size() * sizeof(double), 0, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpyToSymbol(HIP_SYMBOL(c_lf32), lff.data(), lff.size() * sizeof(float), 0, hipMemcpyHostToDevice);
    if (e == hipSuccess && device >= 0 && device < 64) g_ops_loaded |= 1ull << device;
    return e;
}

}  // namespace

// ------------------------------------------------------------------ host side
struct zj_tri {
    int device = 0;
    hipStream_t stream = nullptr;
    int P = 0, NP = 0, NFP = 0, E = 0, F = 0;
    int prec = ZJ_PREC_FP64;
    double gamma = 1.4;
    int riemann = 0;
    int corr = ZJ_CORR_DG;
    int xcd_groups = 1;
    int epb = 0, nblk = 0, nred = 0;
    int64_t cross_edges = 0;
    double* U = nullptr;
    double* W = nullptr;
    double* Rres = nullptr;
    double* aos = nullptr;
    void* geo = nullptr;
    int2* links = nullptr;
    void* T[3] = {nullptr, nullptr, nullptr};
    double* bc_data = nullptr;
    double* hcell = nullptr;
    double* d_scalars = nullptr;  // [0] dt, [1] l2, [2] dt_cfl
    double* partial = nullptr;
    double* partial_dt = nullptr;
    double* h_pinned = nullptr;
    int* perm = nullptr;
    int l2_blocks = 0;
    double dt_on_device = NAN;
    bool has_state = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // discrete adjoint, allocated on first use: the cotangent of the state
    // (blocked layout), of U^n within a step, the recomputed stage states and
    // their traces, the cross-tile edge cotangents (ping-pong, A_cur holds the
    // pending ones), per-tile dt partials of the three stages
    double* A_Ub = nullptr;
    double* A_U0b = nullptr;
    double* A_W1 = nullptr;
    double* A_W2 = nullptr;
    double* A_dtb = nullptr;
    void* A_T[3] = {nullptr, nullptr, nullptr};
    void* A_Tb[2] = {nullptr, nullptr};
    int A_cur = 0;
    int A_blk = 0;   // the adjoint kernels' tiles (the BR2 adjoint's tile may be smaller than the stage kernel's)
    bool A_has_tb = false;
    bool A_seeded = false;
    // Navier-Stokes terms (BR2): mu = 1/Re, kappa, penalty; traces of 12 words
    bool viscous = false;
    double mu = 0.0, kappa = 0.0, eta = 2.0;
    std::string err;
};

static thread_local std::string g_tri_error;

#define ZJ_THIP(h, call)                                                                 \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return ZJ_ERR_HIP;                                                           \
        }                                                                                \
    } while (0)

static int tfail(zj_tri* h, int code, const std::string& msg) {
    h->err = msg;
    return code;
}

static hipStream_t tpick(zj_tri* h, void* s) { return s ? (hipStream_t)s : h->stream; }

template <typename T>
static int tupload(zj_tri* h, void** dst, const std::vector<T>& v) {
    ZJ_THIP(h, hipMalloc(dst, std::max<size_t>(v.size(), 1) * sizeof(T)));
    if (!v.empty()) ZJ_THIP(h, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return ZJ_OK;
}

template <typename R>
static int tupload_as(zj_tri* h, void** dst, const std::vector<double>& v) {
    return tupload(h, dst, std::vector<R>(v.begin(), v.end()));
}

static void trelease(zj_tri* h) {
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    void* bufs[] = {h->U, h->W, h->Rres, h->aos, h->geo, h->links, h->T[0], h->T[1], h->T[2],
                    h->bc_data, h->hcell, h->d_scalars, h->partial, h->partial_dt, h->perm,
                    h->A_Ub, h->A_U0b, h->A_W1, h->A_W2, h->A_dtb, h->A_T[0], h->A_T[1], h->A_T[2],
                    h->A_Tb[0], h->A_Tb[1]};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (h->h_pinned) (void)hipHostFree(h->h_pinned);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

static int tbuild(zj_tri* h, const zj_tri_desc* m, const zj_params* p) {
    const int P = m->poly_order;
    if (P < 1 || P > 5) return tfail(h, ZJ_ERR_ARG, "poly_order must be 1..5");
    if (m->n_elem < 1 || m->n_faces < 1 || m->n_nodes < 3) return tfail(h, ZJ_ERR_ARG, "empty mesh");
    if (!m->xy || !m->tri || !m->normals || !m->face_left_elem || !m->face_left_local || !m->face_right_elem ||
        !m->face_right_local)
        return tfail(h, ZJ_ERR_ARG, "mesh tables missing");
    if (p->precision != ZJ_PREC_FP64 && p->precision != ZJ_PREC_FP32BULK)
        return tfail(h, ZJ_ERR_ARG, "unknown precision");
    if (p->compat != ZJ_COMPAT_TEXTBOOK)
        return tfail(h, ZJ_ERR_UNSUPPORTED, "triangles exist in textbook mode only (the reference's triangle "
                                            "operators are zero)");
    if (p->riemann < 0 || p->riemann > 2) return tfail(h, ZJ_ERR_ARG, "unknown Riemann solver");
    if (p->corr < ZJ_CORR_DG || p->corr > ZJ_CORR_GA) return tfail(h, ZJ_ERR_ARG, "unknown correction function");
    if (p->positivity) return tfail(h, ZJ_ERR_UNSUPPORTED, "the positivity limiter is quadrilateral-only");
    if (!(p->gamma > 1.0)) return tfail(h, ZJ_ERR_ARG, "gamma must exceed 1");
    const int NP = tri::n_sp(P), NFP = tri::n_fp(P);
    if ((int64_t)m->n_elem * NP * 4 > (int64_t)INT32_MAX * 4 || (int64_t)m->n_elem * 3 >= (1ll << 31))
        return tfail(h, ZJ_ERR_UNSUPPORTED, "mesh too large for 32-bit element indices");

    h->P = P; h->NP = NP; h->NFP = NFP; h->E = m->n_elem; h->F = m->n_faces;
    h->prec = p->precision; h->gamma = p->gamma; h->riemann = p->riemann; h->corr = p->corr;
    h->device = p->device;
    h->viscous = p->viscous != 0;
    if (h->viscous) {
        if (!(p->Re > 0.0) || !(p->Pr > 0.0)) return tfail(h, ZJ_ERR_ARG, "viscous terms need Re > 0 and Pr > 0");
        if (P == 4 && p->precision == ZJ_PREC_FP64)
            return tfail(h, ZJ_ERR_UNSUPPORTED, "FP64 P4 triangles with the viscous terms are switched off: an earlier "
                                                "build of their kernel faulted on the GPU, with no cause found");
        h->mu = 1.0 / p->Re;
        h->kappa = h->mu * p->gamma / ((p->gamma - 1.0) * p->Pr);
        h->eta = p->br2_eta > 0.0 ? p->br2_eta : 2.0;
    }
    h->epb = h->viscous ? tri_visc_tile(P, h->prec == ZJ_PREC_FP64) : TILE;
    h->nblk = (h->E + h->epb - 1) / h->epb;
    const int E = h->E;

    // affine geometry in the caller's numbering
    std::vector<double> met((size_t)E * 4), hc((size_t)E), cen((size_t)E * 2);
    for (int e = 0; e < E; ++e) {
        int64_t n[3];
        for (int k = 0; k < 3; ++k) {
            n[k] = m->tri[(size_t)e * 3 + k];
            if (n[k] < 0 || n[k] >= m->n_nodes) return tfail(h, ZJ_ERR_ARG, "triangle node index out of range");
        }
        const double ax = m->xy[2 * n[0]], ay = m->xy[2 * n[0] + 1];
        const double bx = m->xy[2 * n[1]], by = m->xy[2 * n[1] + 1];
        const double cx = m->xy[2 * n[2]], cy = m->xy[2 * n[2] + 1];
        const double xr = 0.5 * (bx - ax), yr = 0.5 * (by - ay), xs = 0.5 * (cx - ax), ys = 0.5 * (cy - ay);
        const double J = xr * ys - xs * yr;
        if (!(J > 0.0)) return tfail(h, ZJ_ERR_ARG, "triangle " + std::to_string(e) + " is degenerate or clockwise");
        met[(size_t)e * 4 + 0] = ys / J;
        met[(size_t)e * 4 + 1] = -xs / J;
        met[(size_t)e * 4 + 2] = -yr / J;
        met[(size_t)e * 4 + 3] = xr / J;
        const double per = std::hypot(bx - ax, by - ay) + std::hypot(cx - bx, cy - by) + std::hypot(ax - cx, ay - cy);
        hc[(size_t)e] = 4.0 * (2.0 * J) / per;  // inscribed-circle diameter
        cen[(size_t)e * 2] = (ax + bx + cx) / 3.0;
        cen[(size_t)e * 2 + 1] = (ay + by + cy) / 3.0;
    }
    // device numbering along a Hilbert curve: perm[new] = old, inv[old] = new
    std::vector<int> perm, inv((size_t)E);
    if (m->reorder && E > 1) perm = host::hilbert_order(E, cen.data());
    for (int e = 0; e < E; ++e) inv[(size_t)(perm.empty() ? e : perm[(size_t)e])] = e;
    auto dev = [&](int old) { return inv[(size_t)old]; };

    std::vector<int2> links((size_t)E * 3, make_int2(-1, -1));
    std::vector<double> geo((size_t)E * GEO, 0.0);
    for (int e = 0; e < E; ++e) {
        const int o = perm.empty() ? e : perm[(size_t)e];
        for (int c = 0; c < 4; ++c) geo[(size_t)e * GEO + c] = met[(size_t)o * 4 + c];
    }
    std::vector<double> hcd((size_t)E);
    for (int e = 0; e < E; ++e) hcd[(size_t)e] = hc[(size_t)(perm.empty() ? e : perm[(size_t)e])];
    for (int f = 0; f < h->F; ++f) {
        const int le = m->face_left_elem[f], ll = m->face_left_local[f];
        const int re = m->face_right_elem[f], rl = m->face_right_local[f];
        if (le < 0 || le >= E || ll < 0 || ll > 2) return tfail(h, ZJ_ERR_ARG, "face left side out of range");
        const int dl = dev(le);
        int2& L = links[(size_t)dl * 3 + ll];
        if (L.y != -1) return tfail(h, ZJ_ERR_ARG, "an element edge belongs to two faces");
        const double nx = m->normals[2 * f], ny = m->normals[2 * f + 1];
        geo[(size_t)dl * GEO + 4 + 2 * ll] = nx;
        geo[(size_t)dl * GEO + 5 + 2 * ll] = ny;
        if (re >= 0) {
            if (re >= E || rl < 0 || rl > 2) return tfail(h, ZJ_ERR_ARG, "face right side out of range");
            const int dr = dev(re);
            int2& Rr = links[(size_t)dr * 3 + rl];
            if (Rr.y != -1 || (re == le && rl == ll)) return tfail(h, ZJ_ERR_ARG, "an element edge belongs to two faces");
            L = make_int2(dr, rl | (ROLE_LEFT << 3));
            Rr = make_int2(dl, ll | (ROLE_RIGHT << 3));
            geo[(size_t)dr * GEO + 4 + 2 * rl] = nx;
            geo[(size_t)dr * GEO + 5 + 2 * rl] = ny;
            if (dl / h->epb != dr / h->epb) h->cross_edges += 2;
        } else {
            if (h->viscous)
                return tfail(h, ZJ_ERR_UNSUPPORTED, "the triangle viscous terms are built for meshes without "
                                                    "boundary faces");
            if (!m->bc_type) return tfail(h, ZJ_ERR_ARG, "boundary face without bc_type");
            const int bct = m->bc_type[f];
            if (bct < 0 || bct > 7) return tfail(h, ZJ_ERR_ARG, "unknown BC type");
            int bci = 0;
            if (bct == ZJ_BC_FARFIELD || bct == ZJ_BC_OUTFLOW) {
                bci = m->bc_face_map ? m->bc_face_map[f] : -1;
                if (!m->bc_data || bci < 0 || bci >= m->n_bc_data)
                    return tfail(h, ZJ_ERR_ARG, "boundary face refers to missing bc_data");
            }
            L = make_int2(-1, bct | (ROLE_BOUNDARY << 3) | (bci << 6));
        }
    }
    for (size_t k = 0; k < links.size(); ++k)
        if (links[k].y == -1) return tfail(h, ZJ_ERR_ARG, "an element edge has no face");

    ZJ_THIP(h, hipSetDevice(h->device));
    ZJ_THIP(h, load_tri_ops(h->device));
    ZJ_THIP(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    int nxcc = 1;
    if (hipDeviceGetAttribute(&nxcc, hipDeviceAttributeNumberOfXccs, h->device) != hipSuccess || nxcc < 1) nxcc = 1;
    h->xcd_groups = nxcc;
    const size_t npts = (size_t)E * NP;
    const size_t eb = ((size_t)E + EB - 1) / EB * EB;  // elements rounded up to whole layout blocks
    const size_t rsz = h->prec == ZJ_PREC_FP64 ? sizeof(double) : sizeof(float);
    ZJ_THIP(h, hipMalloc((void**)&h->U, eb * NP * 4 * sizeof(double)));
    ZJ_THIP(h, hipMalloc((void**)&h->W, eb * NP * 4 * sizeof(double)));
    ZJ_THIP(h, hipMalloc((void**)&h->Rres, eb * NP * 4 * sizeof(double)));
    ZJ_THIP(h, hipMalloc((void**)&h->aos, npts * 4 * sizeof(double)));
    for (int k = 0; k < 3; ++k) ZJ_THIP(h, hipMalloc(&h->T[k], eb * NFP * (h->viscous ? TWV : 4) * rsz));
    int rc;
    if (h->prec == ZJ_PREC_FP64) {
        if ((rc = tupload_as<double>(h, &h->geo, geo))) return rc;
    } else {
        if ((rc = tupload_as<float>(h, &h->geo, geo))) return rc;
    }
    if ((rc = tupload(h, (void**)&h->links, links))) return rc;
    std::vector<double> bcd;
    if (m->bc_data && m->n_bc_data > 0) bcd.assign(m->bc_data, m->bc_data + (size_t)m->n_bc_data * 8);
    if ((rc = tupload(h, (void**)&h->bc_data, bcd))) return rc;
    if ((rc = tupload(h, (void**)&h->hcell, hcd))) return rc;
    if (!perm.empty() && (rc = tupload(h, (void**)&h->perm, perm))) return rc;
    h->nred = (int)std::min<int64_t>(1024, std::max<int64_t>(1, ((int64_t)npts + BLOCK - 1) / BLOCK));
    ZJ_THIP(h, hipMalloc((void**)&h->partial, (size_t)std::max(h->nblk, 1) * sizeof(double)));
    ZJ_THIP(h, hipMalloc((void**)&h->partial_dt, (size_t)h->nred * sizeof(double)));
    ZJ_THIP(h, hipMalloc((void**)&h->d_scalars, 8 * sizeof(double)));
    ZJ_THIP(h, hipHostMalloc((void**)&h->h_pinned, 8 * sizeof(double), hipHostMallocDefault));
    ZJ_THIP(h, hipMemset(h->d_scalars, 0, 8 * sizeof(double)));
    ZJ_THIP(h, hipDeviceSynchronize());
    return ZJ_OK;
}

template <typename R>
static TriArgs<R> tri_args(zj_tri* h, int mode) {
    TriArgs<R> a{};
    a.E = h->E;
    a.geo = (const R*)h->geo;
    a.links = h->links;
    a.bc_data = h->bc_data;
    a.dt_ptr = h->d_scalars;
    a.gamma = h->gamma;
    a.riemann = h->riemann;
    a.xcd_groups = h->xcd_groups;
    a.lift_off = lift_offset(h->P, h->corr);
    a.mu = h->mu;
    a.kappa = h->kappa;
    a.eta = h->eta;
    a.lf_off = lf_family(h->corr) * lf_words(h->P);
    R** T = (R**)h->T;
    switch (mode) {
    case MODE_RESIDUAL:
        a.U_in = h->U; a.R_out = h->Rres; a.T_in = T[0]; a.l2_partial = h->partial;
        break;
    case MODE_STAGE1:
        a.U_in = h->U; a.U_out = h->W; a.T_in = T[0]; a.T_out = T[1];
        break;
    case MODE_STAGE2:
        a.U0 = h->U; a.U_in = h->W; a.U_out = h->W; a.T_in = T[1]; a.T_out = T[2];
        break;
    default:
        a.U0 = h->U; a.U_in = h->W; a.U_out = h->U; a.T_in = T[2]; a.T_out = T[0]; a.l2_partial = h->partial;
        break;
    }
    return a;
}

static hipError_t tenqueue(zj_tri* h, int mode, hipStream_t s) {
    hipError_t e;
    if (h->viscous)
        e = h->prec == ZJ_PREC_FP64 ? tri_visc_launch<double>(h->P, mode, tri_args<double>(h, mode), s)
                                    : tri_visc_launch<float>(h->P, mode, tri_args<float>(h, mode), s);
    else
        e = h->prec == ZJ_PREC_FP64 ? tri_launch<double>(h->P, mode, tri_args<double>(h, mode), s)
                                    : tri_launch<float>(h->P, mode, tri_args<float>(h, mode), s);
    if (e == hipSuccess && (mode == MODE_RESIDUAL || mode == MODE_STAGE3)) h->l2_blocks = h->nblk;
    return e;
}

static int trefresh(zj_tri* h, hipStream_t s) {
    hipError_t e;
    if (h->viscous)
        e = h->prec == ZJ_PREC_FP64
                ? tri_visc_trace<double>(h->P, h->E, h->U, (const double*)h->geo, (double*)h->T[0], s)
                : tri_visc_trace<float>(h->P, h->E, h->U, (const float*)h->geo, (float*)h->T[0], s);
    else
        e = h->prec == ZJ_PREC_FP64 ? tri_trace<double>(h->P, h->E, h->U, (double*)h->T[0], s)
                                    : tri_trace<float>(h->P, h->E, h->U, (float*)h->T[0], s);
    ZJ_THIP(h, e);
    return ZJ_OK;
}

static void tdrop_graph(zj_tri* h) {
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
    if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
}

extern "C" int zj_tri_create(zj_tri_handle* out, const zj_tri_desc* m, const zj_params* p) {
    if (!out || !m || !p) {
        g_tri_error = "null argument";
        return ZJ_ERR_ARG;
    }
    *out = nullptr;
    zj_tri* h = new (std::nothrow) zj_tri();
    if (!h) {
        g_tri_error = "out of host memory";
        return ZJ_ERR_NOMEM;
    }
    int rc;
    try {
        rc = tbuild(h, m, p);
    } catch (const std::bad_alloc&) {   // host tables of a mesh too large for host memory
        h->err = "out of host memory";
        rc = ZJ_ERR_NOMEM;
    }
    if (rc != ZJ_OK) {
        g_tri_error = h->err;
        trelease(h);
        delete h;
        return rc;
    }
    *out = h;
    return ZJ_OK;
}

extern "C" int zj_tri_destroy(zj_tri_handle h) {
    if (!h) return ZJ_ERR_ARG;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    trelease(h);
    delete h;
    return ZJ_OK;
}

extern "C" const char* zj_tri_last_error(zj_tri_handle h) { return h ? h->err.c_str() : g_tri_error.c_str(); }

extern "C" int zj_tri_set_state_device(zj_tri_handle h, const double* d_aos, void* stream) {
    if (!h || !d_aos) return ZJ_ERR_ARG;
    ZJ_THIP(h, hipSetDevice(h->device));
    hipStream_t s = tpick(h, stream);
    ZJ_THIP(h, tri_convert(true, d_aos, h->U, h->E, h->NP, h->perm, s));
    const int rc = trefresh(h, s);
    if (rc) return rc;
    h->has_state = true;
    return ZJ_OK;
}

extern "C" int zj_tri_set_state(zj_tri_handle h, const double* U_aos) {
    if (!h || !U_aos) return ZJ_ERR_ARG;
    ZJ_THIP(h, hipSetDevice(h->device));
    const size_t n = (size_t)h->E * h->NP * 4;
    ZJ_THIP(h, hipMemcpyAsync(h->aos, U_aos, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    const int rc = zj_tri_set_state_device(h, h->aos, nullptr);
    if (rc) return rc;
    ZJ_THIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

extern "C" int zj_tri_get_state_device(zj_tri_handle h, double* d_aos, void* stream) {
    if (!h || !d_aos) return ZJ_ERR_ARG;
    if (!h->has_state) return tfail(h, ZJ_ERR_STATE, "no state set");
    ZJ_THIP(h, hipSetDevice(h->device));
    ZJ_THIP(h, tri_convert(false, h->U, d_aos, h->E, h->NP, h->perm, tpick(h, stream)));
    return ZJ_OK;
}

extern "C" int zj_tri_get_state(zj_tri_handle h, double* U_aos) {
    if (!h || !U_aos) return ZJ_ERR_ARG;
    const int rc = zj_tri_get_state_device(h, h->aos, nullptr);
    if (rc) return rc;
    ZJ_THIP(h, hipMemcpyAsync(U_aos, h->aos, (size_t)h->E * h->NP * 4 * sizeof(double), hipMemcpyDeviceToHost,
                              h->stream));
    ZJ_THIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

extern "C" int zj_tri_residual(zj_tri_handle h, double* dUdt_aos, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (!h->has_state) return tfail(h, ZJ_ERR_STATE, "no state set");
    ZJ_THIP(h, hipSetDevice(h->device));
    hipStream_t s = tpick(h, stream);
    ZJ_THIP(h, tenqueue(h, MODE_RESIDUAL, s));
    if (dUdt_aos) {
        ZJ_THIP(h, tri_convert(false, h->Rres, h->aos, h->E, h->NP, h->perm, s));
        ZJ_THIP(h, hipMemcpyAsync(dUdt_aos, h->aos, (size_t)h->E * h->NP * 4 * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
        ZJ_THIP(h, hipStreamSynchronize(s));
    }
    return ZJ_OK;
}

extern "C" int zj_tri_residual_device(zj_tri_handle h, double* d_dUdt_aos, void* stream) {
    if (!h || !d_dUdt_aos) return ZJ_ERR_ARG;
    if (!h->has_state) return tfail(h, ZJ_ERR_STATE, "no state set");
    ZJ_THIP(h, hipSetDevice(h->device));
    hipStream_t s = tpick(h, stream);
    ZJ_THIP(h, tenqueue(h, MODE_RESIDUAL, s));
    ZJ_THIP(h, tri_convert(false, h->Rres, d_dUdt_aos, h->E, h->NP, h->perm, s));
    return ZJ_OK;
}

extern "C" int zj_tri_time_stages(zj_tri_handle h, double dt, int steps, double* stage_ms, void* stream) {
    if (!h || steps < 1 || !stage_ms) return ZJ_ERR_ARG;
    if (!h->has_state) return tfail(h, ZJ_ERR_STATE, "no state set");
    ZJ_THIP(h, hipSetDevice(h->device));
    hipStream_t s = tpick(h, stream);
    if (dt != h->dt_on_device) {
        h->dt_on_device = dt;
        ZJ_THIP(h, launch_set_scalar(h->d_scalars, dt, s));
    }
    ZJ_THIP(h, time_stage_kernels([&](int st) { return tenqueue(h, st, s); }, steps, s, stage_ms));
    return ZJ_OK;
}

extern "C" int zj_tri_rk3_step(zj_tri_handle h, double dt, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (!h->has_state) return tfail(h, ZJ_ERR_STATE, "no state set");
    ZJ_THIP(h, hipSetDevice(h->device));
    hipStream_t s = tpick(h, stream);
    if (!h->gexec) {
        ZJ_THIP(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int st = 1; st <= 3 && e == hipSuccess; ++st) e = tenqueue(h, st, h->stream);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(h->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            ZJ_THIP(h, e != hipSuccess ? e : e2);
        }
        h->graph = g;
        ZJ_THIP(h, hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
    }
    if (dt != h->dt_on_device) {
        h->dt_on_device = dt;
        ZJ_THIP(h, launch_set_scalar(h->d_scalars, dt, s));
    }
    ZJ_THIP(h, hipGraphLaunch(h->gexec, s));
    return ZJ_OK;
}

extern "C" int zj_tri_dt_cfl(zj_tri_handle h, double cfl, double* dt) {
    if (!h || !dt) return ZJ_ERR_ARG;
    if (!h->has_state) return tfail(h, ZJ_ERR_STATE, "no state set");
    ZJ_THIP(h, hipSetDevice(h->device));
    const double p1 = h->P + 1.0;
    const double visc = h->viscous ? p1 * p1 * p1 * p1 * std::max(h->mu * 4.0 / 3.0, h->kappa * (h->gamma - 1.0))
                                   : 0.0;  // kappa (gamma-1) = mu gamma / Pr
    tri_dt_kernel<<<h->nred, BLOCK, 0, h->stream>>>(h->U, h->hcell, h->E, h->NP, h->gamma, cfl, 2.0 * h->P + 1.0,
                                                    visc, h->partial_dt);
    ZJ_THIP(h, hipGetLastError());
    ZJ_THIP(h, launch_finalize(h->partial_dt, h->nred, FIN_MIN, h->d_scalars + 2, h->stream));
    ZJ_THIP(h, hipMemcpyAsync(h->h_pinned, h->d_scalars + 2, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ZJ_THIP(h, hipStreamSynchronize(h->stream));
    *dt = h->h_pinned[0];
    return ZJ_OK;
}

extern "C" int zj_tri_residual_l2(zj_tri_handle h, double* l2) {
    if (!h || !l2) return ZJ_ERR_ARG;
    if (h->l2_blocks == 0) return tfail(h, ZJ_ERR_STATE, "no residual has been evaluated");
    ZJ_THIP(h, hipSetDevice(h->device));
    ZJ_THIP(h, hipDeviceSynchronize());
    ZJ_THIP(h, launch_finalize(h->partial, h->l2_blocks, FIN_SQRT_SUM, h->d_scalars + 1, h->stream));
    ZJ_THIP(h, hipMemcpyAsync(h->h_pinned, h->d_scalars + 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ZJ_THIP(h, hipStreamSynchronize(h->stream));
    *l2 = h->h_pinned[0];
    return ZJ_OK;
}

extern "C" int zj_tri_synchronize(zj_tri_handle h) {
    if (!h) return ZJ_ERR_ARG;
    ZJ_THIP(h, hipSetDevice(h->device));
    ZJ_THIP(h, hipDeviceSynchronize());
    return ZJ_OK;
}

extern "C" int zj_tri_set_xcd_groups(zj_tri_handle h, int groups) {
    if (!h || groups < 1) return ZJ_ERR_ARG;
    h->xcd_groups = groups;
    tdrop_graph(h);
    return ZJ_OK;
}

extern "C" int zj_tri_get_scheme(zj_tri_handle h, int* poly_order, double* gamma) {
    if (!h) return ZJ_ERR_ARG;
    if (poly_order) *poly_order = h->P;
    if (gamma) *gamma = h->gamma;
    return ZJ_OK;
}

extern "C" int zj_tri_get_info(zj_tri_handle h, int64_t* n_dof, int* elems_per_block, int* n_sp,
                               int64_t* cross_tile_edges) {
    if (!h) return ZJ_ERR_ARG;
    if (n_dof) *n_dof = (int64_t)h->E * h->NP * 4;
    if (elems_per_block) *elems_per_block = h->epb;
    if (n_sp) *n_sp = h->NP;
    if (cross_tile_edges) *cross_tile_edges = h->cross_edges;
    return ZJ_OK;
}

// ------------------------------------------------------------ adjoint (host)
static int tensure_adjoint(zj_tri* h) {
    if (h->A_Ub) return ZJ_OK;
    int tt = TILE;
    if (h->viscous) {
        tt = tri_visc_vjp_tile(h->P, h->prec == ZJ_PREC_FP64);
        if (tt == 0) return tfail(h, ZJ_ERR_UNSUPPORTED, "the triangle adjoint's LDS arrays exceed the workgroup's");
    }
    ZJ_THIP(h, hipSetDevice(h->device));
    h->A_blk = (h->E + tt - 1) / tt;
    const size_t eb = ((size_t)h->E + EB - 1) / EB * EB;
    const size_t nst = eb * h->NP * 4 * sizeof(double);
    const size_t ntr = eb * h->NFP * (h->viscous ? TWV : 4) * (h->prec == ZJ_PREC_FP64 ? sizeof(double) : sizeof(float));
    for (double** b : {&h->A_Ub, &h->A_U0b, &h->A_W1, &h->A_W2}) ZJ_THIP(h, hipMalloc((void**)b, nst));
    for (int k = 0; k < 3; ++k) ZJ_THIP(h, hipMalloc(&h->A_T[k], ntr));
    for (int k = 0; k < 2; ++k) ZJ_THIP(h, hipMalloc(&h->A_Tb[k], ntr));
    ZJ_THIP(h, hipMalloc((void**)&h->A_dtb, (size_t)3 * std::max(h->A_blk, 1) * sizeof(double)));
    return ZJ_OK;
}

template <typename R>
static TriVjpArgs<R> tvjp_args(zj_tri* h) {
    TriVjpArgs<R> a{};
    a.E = h->E;
    a.geo = (const R*)h->geo;
    a.links = h->links;
    a.bc_data = h->bc_data;
    a.dt_ptr = h->d_scalars;
    a.gamma = h->gamma;
    a.riemann = h->riemann;
    a.xcd_groups = h->xcd_groups;
    a.lift_off = lift_offset(h->P, h->corr);
    a.mu = h->mu;
    a.kappa = h->kappa;
    a.eta = h->eta;
    a.lf_off = lf_family(h->corr) * lf_words(h->P);
    return a;
}

// One step backwards: traces of U^n, stages 1 and 2 recomputed into A_W1 and
// A_W2 (with their cross-tile traces), then the stage VJPs 3, 2, 1, each
// consuming the edge cotangents the previous launch left.
template <typename R>
static hipError_t tri_adjoint_step(zj_tri* h, bool want_dt, hipStream_t s) {
    R** AT = (R**)h->A_T;
    R** Tb = (R**)h->A_Tb;
    const R* geo = (const R*)h->geo;
    // With the viscous terms every edge's traces are formed after each
    // recomputed stage: the adjoint's tiles can be smaller than the stage
    // kernel's, whose own trace writes cover only the edges across its tiles.
    hipError_t e = h->viscous ? tri_visc_trace<R>(h->P, h->E, h->U, geo, AT[0], s) : tri_trace<R>(h->P, h->E, h->U, AT[0], s);
    if (e != hipSuccess) return e;
    TriArgs<R> f = tri_args<R>(h, MODE_STAGE1);
    f.U_in = h->U; f.U_out = h->A_W1; f.T_in = AT[0]; f.T_out = AT[1]; f.l2_partial = nullptr;
    if ((e = h->viscous ? tri_visc_launch<R>(h->P, MODE_STAGE1, f, s) : tri_launch<R>(h->P, MODE_STAGE1, f, s)) != hipSuccess)
        return e;
    if (h->viscous && (e = tri_visc_trace<R>(h->P, h->E, h->A_W1, geo, AT[1], s)) != hipSuccess) return e;
    f = tri_args<R>(h, MODE_STAGE2);
    f.U0 = h->U; f.U_in = h->A_W1; f.U_out = h->A_W2; f.T_in = AT[1]; f.T_out = AT[2]; f.l2_partial = nullptr;
    if ((e = h->viscous ? tri_visc_launch<R>(h->P, MODE_STAGE2, f, s) : tri_launch<R>(h->P, MODE_STAGE2, f, s)) != hipSuccess)
        return e;
    if (h->viscous && (e = tri_visc_trace<R>(h->P, h->E, h->A_W2, geo, AT[2], s)) != hipSuccess) return e;
    for (int st = 3; st >= 1; --st) {
        TriVjpArgs<R> a = tvjp_args<R>(h);
        a.U_in = st == 3 ? h->A_W2 : (st == 2 ? h->A_W1 : h->U);
        a.U0 = h->U;
        a.T_in = AT[st - 1];
        a.Ub_in = h->A_Ub;
        a.Ub_out = h->A_Ub;
        a.Tb_in = h->A_has_tb ? Tb[h->A_cur] : nullptr;
        a.Tb_out = Tb[1 - h->A_cur];
        a.U0b = h->A_U0b;
        a.dtb_partial = want_dt ? h->A_dtb + (size_t)(3 - st) * h->A_blk : nullptr;
        if ((e = h->viscous ? tri_visc_vjp<R>(h->P, st, a, s) : tri_vjp<R>(h->P, st, a, s)) != hipSuccess) return e;
        h->A_cur = 1 - h->A_cur;
        h->A_has_tb = true;
    }
    return hipSuccess;
}

extern "C" int zj_tri_adjoint_set_device(zj_tri_handle h, const double* d_ubar_aos, void* stream) {
    if (!h || !d_ubar_aos) return ZJ_ERR_ARG;
    const int rc = tensure_adjoint(h);
    if (rc) return rc;
    ZJ_THIP(h, hipSetDevice(h->device));
    ZJ_THIP(h, tri_convert(true, d_ubar_aos, h->A_Ub, h->E, h->NP, h->perm, tpick(h, stream)));
    h->A_has_tb = false;
    h->A_seeded = true;
    return ZJ_OK;
}

extern "C" int zj_tri_adjoint_set(zj_tri_handle h, const double* ubar_aos) {
    if (!h || !ubar_aos) return ZJ_ERR_ARG;
    ZJ_THIP(h, hipSetDevice(h->device));
    ZJ_THIP(h, hipMemcpyAsync(h->aos, ubar_aos, (size_t)h->E * h->NP * 4 * sizeof(double), hipMemcpyHostToDevice,
                              h->stream));
    const int rc = zj_tri_adjoint_set_device(h, h->aos, nullptr);
    if (rc) return rc;
    ZJ_THIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

extern "C" int zj_tri_adjoint_get_device(zj_tri_handle h, double* d_ubar_aos, void* stream) {
    if (!h || !d_ubar_aos) return ZJ_ERR_ARG;
    if (!h->A_seeded) return tfail(h, ZJ_ERR_STATE, "no adjoint seed set");
    ZJ_THIP(h, hipSetDevice(h->device));
    hipStream_t s = tpick(h, stream);
    const void* tb = h->A_has_tb ? h->A_Tb[h->A_cur] : nullptr;
    hipError_t e;
    if (h->viscous)
        e = h->prec == ZJ_PREC_FP64
                ? tri_visc_complete<double>(h->P, h->E, h->links, (const double*)h->geo, h->A_Ub, (const double*)tb,
                                            h->Rres, h->xcd_groups, s)
                : tri_visc_complete<float>(h->P, h->E, h->links, (const float*)h->geo, h->A_Ub, (const float*)tb,
                                           h->Rres, h->xcd_groups, s);
    else
        e = h->prec == ZJ_PREC_FP64
                ? tri_complete<double>(h->P, h->E, h->links, h->A_Ub, (const double*)tb, h->Rres, h->xcd_groups, s)
                : tri_complete<float>(h->P, h->E, h->links, h->A_Ub, (const float*)tb, h->Rres, h->xcd_groups, s);
    ZJ_THIP(h, e);
    ZJ_THIP(h, tri_convert(false, h->Rres, d_ubar_aos, h->E, h->NP, h->perm, s));
    return ZJ_OK;
}

extern "C" int zj_tri_adjoint_get(zj_tri_handle h, double* ubar_aos) {
    if (!h || !ubar_aos) return ZJ_ERR_ARG;
    const int rc = zj_tri_adjoint_get_device(h, h->aos, nullptr);
    if (rc) return rc;
    ZJ_THIP(h, hipMemcpyAsync(ubar_aos, h->aos, (size_t)h->E * h->NP * 4 * sizeof(double), hipMemcpyDeviceToHost,
                              h->stream));
    ZJ_THIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

extern "C" int zj_tri_rk3_step_vjp(zj_tri_handle h, double dt, double* dt_bar, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (!h->has_state) return tfail(h, ZJ_ERR_STATE, "no state set");
    if (!h->A_seeded) return tfail(h, ZJ_ERR_STATE, "no adjoint seed set");
    ZJ_THIP(h, hipSetDevice(h->device));
    hipStream_t s = tpick(h, stream);
    if (dt != h->dt_on_device) {
        h->dt_on_device = dt;
        ZJ_THIP(h, launch_set_scalar(h->d_scalars, dt, s));
    }
    const bool want = dt_bar != nullptr;
    ZJ_THIP(h, h->prec == ZJ_PREC_FP64 ? tri_adjoint_step<double>(h, want, s) : tri_adjoint_step<float>(h, want, s));
    if (want) {
        ZJ_THIP(h, launch_finalize(h->A_dtb, 3 * h->A_blk, FIN_SUM, h->d_scalars + 3, s));
        ZJ_THIP(h, hipMemcpyAsync(h->h_pinned + 3, h->d_scalars + 3, sizeof(double), hipMemcpyDeviceToHost, s));
        ZJ_THIP(h, hipStreamSynchronize(s));
        *dt_bar = h->h_pinned[3];
    }
    return ZJ_OK;
}

template <typename R>
static hipError_t tri_residual_vjp(zj_tri* h, hipStream_t s) {
    R** AT = (R**)h->A_T;
    R* tb = (R*)h->A_Tb[1 - h->A_cur];  // the pending edge cotangents of the accumulator stay in A_Tb[A_cur]
    const R* geo = (const R*)h->geo;
    hipError_t e = h->viscous ? tri_visc_trace<R>(h->P, h->E, h->U, geo, AT[0], s) : tri_trace<R>(h->P, h->E, h->U, AT[0], s);
    if (e != hipSuccess) return e;
    TriVjpArgs<R> a = tvjp_args<R>(h);
    a.U_in = h->U;
    a.T_in = AT[0];
    a.Ub_in = h->Rres;
    a.Ub_out = h->Rres;
    a.Tb_out = tb;
    if (h->viscous) {
        if ((e = tri_visc_vjp<R>(h->P, MODE_RESIDUAL, a, s)) != hipSuccess) return e;
        return tri_visc_complete<R>(h->P, h->E, h->links, geo, h->Rres, tb, h->W, h->xcd_groups, s);
    }
    if ((e = tri_vjp<R>(h->P, MODE_RESIDUAL, a, s)) != hipSuccess) return e;
    return tri_complete<R>(h->P, h->E, h->links, h->Rres, tb, h->W, h->xcd_groups, s);
}

extern "C" int zj_tri_residual_vjp_device(zj_tri_handle h, const double* d_w_aos, double* d_ubar_aos, void* stream) {
    if (!h || !d_w_aos || !d_ubar_aos) return ZJ_ERR_ARG;
    if (!h->has_state) return tfail(h, ZJ_ERR_STATE, "no state set");
    const int rc = tensure_adjoint(h);
    if (rc) return rc;
    ZJ_THIP(h, hipSetDevice(h->device));
    hipStream_t s = tpick(h, stream);
    ZJ_THIP(h, tri_convert(true, d_w_aos, h->Rres, h->E, h->NP, h->perm, s));
    ZJ_THIP(h, h->prec == ZJ_PREC_FP64 ? tri_residual_vjp<double>(h, s) : tri_residual_vjp<float>(h, s));
    ZJ_THIP(h, tri_convert(false, h->W, d_ubar_aos, h->E, h->NP, h->perm, s));
    return ZJ_OK;
}

extern "C" int zj_tri_residual_vjp(zj_tri_handle h, const double* w_aos, double* ubar_aos) {
    if (!h || !w_aos || !ubar_aos) return ZJ_ERR_ARG;
    ZJ_THIP(h, hipSetDevice(h->device));
    const size_t n = (size_t)h->E * h->NP * 4;
    ZJ_THIP(h, hipMemcpyAsync(h->aos, w_aos, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    const int rc = zj_tri_residual_vjp_device(h, h->aos, h->aos, nullptr);  // w is consumed before aos is rewritten
    if (rc) return rc;
    ZJ_THIP(h, hipMemcpyAsync(ubar_aos, h->aos, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ZJ_THIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

extern "C" int zj_tri_operators(int p, double* sp, double* fp, double* wg, double* Dr, double* Ds, double* interp,
                                double* lift) {
    tri::Operators O;
    if (!tri::operators(p, O)) return ZJ_ERR_ARG;
    auto put = [](double* dst, const std::vector<double>& v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
    };
    put(sp, O.sp); put(fp, O.fp); put(wg, O.wg);
    put(Dr, O.Dr); put(Ds, O.Ds); put(interp, O.interp); put(lift, O.lift);
    return ZJ_OK;
}

extern "C" int zj_tri_lift(int p, int corr, double* lift) {
    tri::Operators O;
    if (!lift || corr < ZJ_CORR_DG || corr > ZJ_CORR_GA || !tri::operators(p, O, corr)) return ZJ_ERR_ARG;
    std::memcpy(lift, O.lift.data(), O.lift.size() * sizeof(double));
    return ZJ_OK;
}
