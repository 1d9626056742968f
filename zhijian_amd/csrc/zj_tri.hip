// zj_tri.hip -- flux reconstruction on straight-sided triangles (textbook
// scheme, DG-equivalent correction) and its SSP-RK3 step: one fused kernel per
// stage, operators in LDS, C ABI zj_tri_* (include/zhijian_amd.h).
//
// The reference has no triangle semantics to follow: FROperators::initTriangle
// leaves every table zero (src/basis/basis.cpp:615-669) and FRSolver assumes
// quadrilaterals (src/solver/fr_solver.cpp:138-140).  The scheme is the one its
// settings name on simplices: nodal FR with the L2 lifting as correction field
// (Huynh 2007 with c = 0; Castonguay, Vincent & Jameson 2012), transformed
// fluxes through the constant affine metric, one common flux per face point
// from the configured Riemann solver (include/physics/euler.hpp:204-375).
//
//   dU/dt = -(Dr Fr + Ds Fs) - sum_j lift[:, j] (F^I_j s_j - Fn_j)
//   Fr = r_x F + r_y G,  Fs = s_x F + s_y G   (contravariant fluxes)
//   s_j = |e|/(2J) of flux point j's edge:  |grad s|, |grad r + grad s|, |grad r|
//   Fn_j = interpolated transformed normal flux: -Fs, Fr + Fs, -Fr on edges 0, 1, 2
//   F^I_j = Riemann flux along the face normal, outward on each side
//
// Device layout (one GPU holds the whole mesh):
//   U, W     state planes [4][E*Np] FP64 (W = U1, then U2 in place)
//   T[3]     flux-point traces [E][3][p+1][4] in the compute precision; only
//            edges whose neighbour lies in another element tile are written
//            (a neighbour in the same tile reads them from LDS), rotating as
//            the quadrilateral path's: stage 1 T0->T1, 2 T1->T2, 3 T2->T0
//   geo      [E][10]: r_x, r_y, s_x, s_y, then the face normal (as the face's
//            left element sees it) of edges 0, 1, 2
//   links    [E][3] int2: x = neighbour element (-1 on a boundary), y = packed
//            neighbour edge (or BC type) | role << 3 | BC data index << 6
//   ops      transposed operators in the compute precision:
//            Dr^T [Np][Np], Ds^T [Np][Np], interp^T [Np][3(p+1)], lift^T [3(p+1)][Np]
#include "zhijian_amd.h"
#include "zj_device.hpp"
#include "zj_host_math.hpp"
#include "zj_kernels.hpp"
#include "zj_tri_ops.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace zj;

namespace {

template <int P>
struct TriShape {
    static constexpr int NP = (P + 1) * (P + 2) / 2;        // solution points
    static constexpr int FP = P + 1;                        // flux points per edge
    static constexpr int NFP = 3 * FP;
    static constexpr int EPB = BLOCK / NP;                  // elements per workgroup
    static constexpr int NT = EPB * NP;                     // threads holding a solution point
    static constexpr int NI = EPB * NFP;                    // flux-point items per tile
    static constexpr int NIT = (NI + BLOCK - 1) / BLOCK;
    static constexpr int OPS = 2 * NP * NP + 2 * NFP * NP;  // operator words
};

inline int tri_epb(int P) { return BLOCK / ((P + 1) * (P + 2) / 2); }

constexpr int GEO = 10;  // geometry words per element

template <typename R>
struct TriArgs {
    int E;
    const double* U_in;   // stage input planes
    const double* U0;     // U^n (stages 2, 3)
    double* U_out;        // stage output planes (may alias U_in or U0 element-locally)
    double* R_out;        // residual planes (MODE_RESIDUAL)
    const R* geo;
    const int2* links;
    const R* T_in;
    R* T_out;
    const R* ops;
    const double* bc_data;
    const double* dt_ptr;
    double* l2_partial;   // per tile sum of R^2, or nullptr
    double gamma;
    int riemann;
    int xcd_groups;
};

__device__ __forceinline__ int link_edge(int info) { return info & 7; }
__device__ __forceinline__ int link_role(int info) { return (info >> 3) & 3; }
__device__ __forceinline__ int link_bc(int info) { return info >> 6; }

// |e|/(2J) of local edge k from the affine metric.
template <typename R>
__device__ __forceinline__ R edge_scale(int k, R rx, R ry, R sx, R sy) {
    const R ax = k == 0 ? sx : (k == 1 ? rx + sx : rx);
    const R ay = k == 0 ? sy : (k == 1 ? ry + sy : ry);
    return sqrtR(ax * ax + ay * ay);
}

template <typename R, int P, int MODE>
__global__ __launch_bounds__(BLOCK) void tri_stage_kernel(TriArgs<R> a) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, FP = S::FP, NFP = S::NFP, EPB = S::EPB, NIT = S::NIT;
    __shared__ R sOp[S::OPS];
    __shared__ R sU[EPB * NP * NV];
    __shared__ R sF[EPB * NP * 2 * NV];
    __shared__ R sT[EPB * NFP * NV];  // own traces, then the flux differences
    __shared__ double sred[BLOCK / 64];
    R* const sD = sT;

    const int tid = threadIdx.x;
    const int tile = tile_of_block(a.xcd_groups);
    const int e0 = tile * EPB;
    for (int k = tid; k < S::OPS; k += BLOCK) sOp[k] = a.ops[k];
    const R* DrT = sOp;
    const R* DsT = DrT + NP * NP;
    const R* IT = DsT + NP * NP;
    const R* LT = IT + NP * NFP;
    const R g = R(a.gamma);
    const int64_t PS = (int64_t)a.E * NP;

    // ---- 1: state and contravariant fluxes at the solution points
    const int el = tid / NP, i = tid - el * NP;
    const int e = e0 + el;
    const bool live = tid < S::NT && e < a.E;
    const int64_t gi = (int64_t)e * NP + i;
    double uin[NV] = {0.0, 0.0, 0.0, 0.0};
    if (live) {
        R u[NV], fx[NV], fy[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            uin[v] = a.U_in[v * PS + gi];
            u[v] = R(uin[v]);
            sU[(el * NP + i) * NV + v] = u[v];
        }
        euler_flux(u, g, fx, fy);
        const R* G = a.geo + (int64_t)e * GEO;
        const R rx = G[0], ry = G[1], sx = G[2], sy = G[3];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sF[(el * NP + i) * 2 * NV + v] = rx * fx[v] + ry * fy[v];
            sF[(el * NP + i) * 2 * NV + NV + v] = sx * fx[v] + sy * fy[v];
        }
    }
    __syncthreads();

    // ---- 2: own traces and the interpolated transformed normal flux
    R fn[NIT][NV];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int t = tid + it * BLOCK;
        const int el2 = t / NFP, j = t - el2 * NFP;
#pragma unroll
        for (int v = 0; v < NV; ++v) fn[it][v] = R(0);
        if (t < S::NI && e0 + el2 < a.E) {
            R ut[NV], ar[NV], as[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) { ut[v] = R(0); ar[v] = R(0); as[v] = R(0); }
#pragma unroll 2
            for (int m = 0; m < NP; ++m) {
                const R w = IT[m * NFP + j];
                const R* u = sU + (el2 * NP + m) * NV;
                const R* f = sF + (el2 * NP + m) * 2 * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    ut[v] = fmaR(w, u[v], ut[v]);
                    ar[v] = fmaR(w, f[v], ar[v]);
                    as[v] = fmaR(w, f[NV + v], as[v]);
                }
            }
            const int k = j / FP;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                fn[it][v] = k == 0 ? -as[v] : (k == 1 ? ar[v] + as[v] : -ar[v]);
                sT[(el2 * NFP + j) * NV + v] = ut[v];
            }
        }
    }
    __syncthreads();

    // ---- 3: common flux and flux difference per flux point
    bool cross[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int t = tid + it * BLOCK;
        const int el2 = t / NFP, j = t - el2 * NFP;
        const int e2 = e0 + el2;
        cross[it] = false;
        if (t < S::NI && e2 < a.E) {
            const int k = j / FP, f = j - k * FP;
            const int2 L = a.links[(int64_t)e2 * 3 + k];
            const int role = link_role(L.y);
            const R* G = a.geo + (int64_t)e2 * GEO;
            const R nx = G[4 + 2 * k], ny = G[5 + 2 * k];
            R own[NV], oth[NV], UL[NV], UR[NV], Fc[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) own[v] = sT[(el2 * NFP + j) * NV + v];
            if (role == ROLE_BOUNDARY) {
                ghost_state(own, nx, ny, link_edge(L.y), a.bc_data + (int64_t)link_bc(L.y) * 8, g, oth);
            } else {
                const int nl = L.x - e0;
                const int oj = link_edge(L.y) * FP + (P - f);
                if (nl >= 0 && nl < EPB) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) oth[v] = sT[(nl * NFP + oj) * NV + v];
                } else {
                    cross[it] = true;
                    load4(a.T_in + ((int64_t)L.x * NFP + oj) * NV, oth);
                }
            }
            const bool right = role == ROLE_RIGHT;
#pragma unroll
            for (int v = 0; v < NV; ++v) { UL[v] = right ? oth[v] : own[v]; UR[v] = right ? own[v] : oth[v]; }
            riemann_by_kind(a.riemann, UL, UR, nx, ny, g, Fc);
            const R sc = edge_scale(k, G[0], G[1], G[2], G[3]);
#pragma unroll
            for (int v = 0; v < NV; ++v) fn[it][v] = sc * (right ? -Fc[v] : Fc[v]) - fn[it][v];
        }
    }
    __syncthreads();  // every trace read; the flux differences take their place
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int t = tid + it * BLOCK;
        if (t < S::NI && e0 + t / NFP < a.E)
#pragma unroll
            for (int v = 0; v < NV; ++v) sD[t * NV + v] = fn[it][v];
    }
    __syncthreads();

    // ---- 4: divergence + lifted correction, RK stage update
    double l2 = 0.0;
    if (live) {
        R acc[NV] = {R(0), R(0), R(0), R(0)};
#pragma unroll 2
        for (int m = 0; m < NP; ++m) {
            const R dr = DrT[m * NP + i], ds = DsT[m * NP + i];
            const R* f = sF + (el * NP + m) * 2 * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] = fmaR(ds, f[NV + v], fmaR(dr, f[v], acc[v]));
        }
#pragma unroll 2
        for (int j = 0; j < NFP; ++j) {
            const R w = LT[j * NP + i];
            const R* d = sD + (el * NFP + j) * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] = fmaR(w, d[v], acc[v]);
        }
        double res[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) res[v] = -double(acc[v]);
        if (MODE == MODE_RESIDUAL) {
#pragma unroll
            for (int v = 0; v < NV; ++v) { a.R_out[v * PS + gi] = res[v]; l2 += res[v] * res[v]; }
        } else {
            const double dt = *a.dt_ptr;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double un;
                if (MODE == MODE_STAGE1) {
                    un = uin[v] + dt * res[v];
                } else {
                    const double u0 = a.U0[v * PS + gi];
                    un = MODE == MODE_STAGE2 ? 0.75 * u0 + 0.25 * (uin[v] + dt * res[v])
                                             : (1.0 / 3.0) * u0 + (2.0 / 3.0) * (uin[v] + dt * res[v]);
                    if (MODE == MODE_STAGE3) l2 += res[v] * res[v];
                }
                a.U_out[v * PS + gi] = un;
                sU[(el * NP + i) * NV + v] = R(un);
            }
        }
    }
    if ((MODE == MODE_RESIDUAL || MODE == MODE_STAGE3) && a.l2_partial) {
        const double s = block_reduce<OpSum>(l2, sred);
        if (tid == 0) a.l2_partial[tile] = s;
    }
    if (MODE == MODE_RESIDUAL) return;
    __syncthreads();

    // ---- 5: traces of the new stage state on edges another tile reads
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        if (!cross[it]) continue;
        const int t = tid + it * BLOCK;
        const int el2 = t / NFP, j = t - el2 * NFP;
        R ut[NV] = {R(0), R(0), R(0), R(0)};
#pragma unroll 2
        for (int m = 0; m < NP; ++m) {
            const R w = IT[m * NFP + j];
            const R* u = sU + (el2 * NP + m) * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) ut[v] = fmaR(w, u[v], ut[v]);
        }
        store4(a.T_out + ((int64_t)(e0 + el2) * NFP + j) * NV, ut);
    }
}

// Traces of U on every edge (after a new state is set).
template <typename R, int P>
__global__ __launch_bounds__(BLOCK) void tri_trace_kernel(int E, const double* U, const R* ops, R* T_out,
                                                          int xcd_groups) {
    using S = TriShape<P>;
    constexpr int NP = S::NP, NFP = S::NFP, EPB = S::EPB;
    __shared__ R sI[NP * NFP];
    __shared__ R sU[EPB * NP * NV];
    const int tid = threadIdx.x;
    const int e0 = tile_of_block(xcd_groups) * EPB;
    for (int k = tid; k < NP * NFP; k += BLOCK) sI[k] = ops[2 * NP * NP + k];
    const int64_t PS = (int64_t)E * NP;
    const int el = tid / NP, i = tid - el * NP;
    if (tid < S::NT && e0 + el < E)
#pragma unroll
        for (int v = 0; v < NV; ++v) sU[(el * NP + i) * NV + v] = R(U[v * PS + (int64_t)(e0 + el) * NP + i]);
    __syncthreads();
    for (int t = tid; t < S::NI; t += BLOCK) {
        const int el2 = t / NFP, j = t - el2 * NFP;
        if (e0 + el2 >= E) continue;
        R ut[NV] = {R(0), R(0), R(0), R(0)};
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const R w = sI[m * NFP + j];
#pragma unroll
            for (int v = 0; v < NV; ++v) ut[v] = fmaR(w, sU[(el2 * NP + m) * NV + v], ut[v]);
        }
        store4(T_out + ((int64_t)(e0 + el2) * NFP + j) * NV, ut);
    }
}

template <typename R, int P>
hipError_t tri_launch_p(int mode, const TriArgs<R>& a, hipStream_t s) {
    const int grid = (std::max(a.E, 1) + TriShape<P>::EPB - 1) / TriShape<P>::EPB;
    switch (mode) {
    case MODE_RESIDUAL: tri_stage_kernel<R, P, MODE_RESIDUAL><<<grid, BLOCK, 0, s>>>(a); break;
    case MODE_STAGE1: tri_stage_kernel<R, P, MODE_STAGE1><<<grid, BLOCK, 0, s>>>(a); break;
    case MODE_STAGE2: tri_stage_kernel<R, P, MODE_STAGE2><<<grid, BLOCK, 0, s>>>(a); break;
    default: tri_stage_kernel<R, P, MODE_STAGE3><<<grid, BLOCK, 0, s>>>(a); break;
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
hipError_t tri_trace_p(int E, const double* U, const R* ops, R* T, int groups, hipStream_t s) {
    const int grid = (std::max(E, 1) + TriShape<P>::EPB - 1) / TriShape<P>::EPB;
    tri_trace_kernel<R, P><<<grid, BLOCK, 0, s>>>(E, U, ops, T, groups);
    return hipGetLastError();
}

template <typename R>
hipError_t tri_trace(int P, int E, const double* U, const R* ops, R* T, int groups, hipStream_t s) {
    switch (P) {
    case 1: return tri_trace_p<R, 1>(E, U, ops, T, groups, s);
    case 2: return tri_trace_p<R, 2>(E, U, ops, T, groups, s);
    case 3: return tri_trace_p<R, 3>(E, U, ops, T, groups, s);
    case 4: return tri_trace_p<R, 4>(E, U, ops, T, groups, s);
    default: return tri_trace_p<R, 5>(E, U, ops, T, groups, s);
    }
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
    void* ops = nullptr;
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
    std::string err;
};

static std::string g_tri_error;

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
    void* bufs[] = {h->U, h->W, h->Rres, h->aos, h->geo, h->links, h->T[0], h->T[1], h->T[2], h->ops,
                    h->bc_data, h->hcell, h->d_scalars, h->partial, h->partial_dt, h->perm};
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
    if (p->corr != ZJ_CORR_DG)
        return tfail(h, ZJ_ERR_UNSUPPORTED, "triangles use the DG-equivalent correction field only");
    if (p->positivity) return tfail(h, ZJ_ERR_UNSUPPORTED, "the positivity limiter is quadrilateral-only");
    if (!(p->gamma > 1.0)) return tfail(h, ZJ_ERR_ARG, "gamma must exceed 1");
    const int NP = tri::n_sp(P), NFP = tri::n_fp(P);
    if ((int64_t)m->n_elem * NP * 4 > (int64_t)INT32_MAX * 4 || (int64_t)m->n_elem * 3 >= (1ll << 31))
        return tfail(h, ZJ_ERR_UNSUPPORTED, "mesh too large for 32-bit element indices");

    h->P = P; h->NP = NP; h->NFP = NFP; h->E = m->n_elem; h->F = m->n_faces;
    h->prec = p->precision; h->gamma = p->gamma; h->riemann = p->riemann; h->device = p->device;
    h->epb = tri_epb(P);
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

    tri::Operators O;
    if (!tri::operators(P, O)) return tfail(h, ZJ_ERR_UNSUPPORTED, "triangle operators failed");
    std::vector<double> ops((size_t)(2 * NP * NP + 2 * NFP * NP));
    {
        double* DrT = ops.data();
        double* DsT = DrT + NP * NP;
        double* IT = DsT + NP * NP;
        double* LT = IT + NP * NFP;
        for (int i = 0; i < NP; ++i)
            for (int j = 0; j < NP; ++j) {
                DrT[j * NP + i] = O.Dr[(size_t)i * NP + j];
                DsT[j * NP + i] = O.Ds[(size_t)i * NP + j];
            }
        for (int j = 0; j < NFP; ++j)
            for (int i = 0; i < NP; ++i) {
                IT[i * NFP + j] = O.interp[(size_t)j * NP + i];
                LT[j * NP + i] = O.lift[(size_t)i * NFP + j];
            }
    }

    ZJ_THIP(h, hipSetDevice(h->device));
    ZJ_THIP(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    int nxcc = 1;
    if (hipDeviceGetAttribute(&nxcc, hipDeviceAttributeNumberOfXccs, h->device) != hipSuccess || nxcc < 1) nxcc = 1;
    h->xcd_groups = nxcc;
    const size_t npts = (size_t)E * NP;
    const size_t rsz = h->prec == ZJ_PREC_FP64 ? sizeof(double) : sizeof(float);
    ZJ_THIP(h, hipMalloc((void**)&h->U, npts * 4 * sizeof(double)));
    ZJ_THIP(h, hipMalloc((void**)&h->W, npts * 4 * sizeof(double)));
    ZJ_THIP(h, hipMalloc((void**)&h->Rres, npts * 4 * sizeof(double)));
    ZJ_THIP(h, hipMalloc((void**)&h->aos, npts * 4 * sizeof(double)));
    for (int k = 0; k < 3; ++k) ZJ_THIP(h, hipMalloc(&h->T[k], (size_t)E * NFP * 4 * rsz));
    int rc;
    if (h->prec == ZJ_PREC_FP64) {
        if ((rc = tupload_as<double>(h, &h->geo, geo))) return rc;
        if ((rc = tupload_as<double>(h, &h->ops, ops))) return rc;
    } else {
        if ((rc = tupload_as<float>(h, &h->geo, geo))) return rc;
        if ((rc = tupload_as<float>(h, &h->ops, ops))) return rc;
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
    a.ops = (const R*)h->ops;
    a.bc_data = h->bc_data;
    a.dt_ptr = h->d_scalars;
    a.gamma = h->gamma;
    a.riemann = h->riemann;
    a.xcd_groups = h->xcd_groups;
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
    hipError_t e = h->prec == ZJ_PREC_FP64 ? tri_launch<double>(h->P, mode, tri_args<double>(h, mode), s)
                                           : tri_launch<float>(h->P, mode, tri_args<float>(h, mode), s);
    if (e == hipSuccess && (mode == MODE_RESIDUAL || mode == MODE_STAGE3)) h->l2_blocks = h->nblk;
    return e;
}

static int trefresh(zj_tri* h, hipStream_t s) {
    hipError_t e = h->prec == ZJ_PREC_FP64
                       ? tri_trace<double>(h->P, h->E, h->U, (const double*)h->ops, (double*)h->T[0], h->xcd_groups, s)
                       : tri_trace<float>(h->P, h->E, h->U, (const float*)h->ops, (float*)h->T[0], h->xcd_groups, s);
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
    const int rc = tbuild(h, m, p);
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
    ZJ_THIP(h, launch_aos_to_soa(d_aos, h->U, (int64_t)h->E * h->NP, s, h->perm, h->NP));
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
    ZJ_THIP(h, launch_soa_to_aos(h->U, d_aos, (int64_t)h->E * h->NP, tpick(h, stream), h->perm, h->NP));
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
        ZJ_THIP(h, launch_soa_to_aos(h->Rres, h->aos, (int64_t)h->E * h->NP, s, h->perm, h->NP));
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
    ZJ_THIP(h, launch_soa_to_aos(h->Rres, d_dUdt_aos, (int64_t)h->E * h->NP, s, h->perm, h->NP));
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
    ZJ_THIP(h, launch_dt_partial(h->U, h->hcell, h->E, h->NP, h->gamma, cfl, 2.0 * h->P + 1.0, h->partial_dt,
                                 h->nred, h->stream));
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

extern "C" int zj_tri_get_info(zj_tri_handle h, int64_t* n_dof, int* elems_per_block, int* n_sp,
                               int64_t* cross_tile_edges) {
    if (!h) return ZJ_ERR_ARG;
    if (n_dof) *n_dof = (int64_t)h->E * h->NP * 4;
    if (elems_per_block) *elems_per_block = h->epb;
    if (n_sp) *n_sp = h->NP;
    if (cross_tile_edges) *cross_tile_edges = h->cross_edges;
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
