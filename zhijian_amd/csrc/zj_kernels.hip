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

#include <cmath>
#include <cstdio>
#include <type_traits>

namespace zj {

namespace {

template <typename R> __device__ __forceinline__ R fmaR(R a, R b, R c);
template <> __device__ __forceinline__ double fmaR<double>(double a, double b, double c) { return fma(a, b, c); }
template <> __device__ __forceinline__ float fmaR<float>(float a, float b, float c) { return fmaf(a, b, c); }

// FP64 keeps correctly rounded square root and division (the plain form must
// track the reference to 1e-12); the FP32 bulk uses the 1-ulp hardware
// instructions, whose error sits far below the FP32 rounding of the state.
template <typename R> __device__ __forceinline__ R sqrtR(R x);
template <> __device__ __forceinline__ double sqrtR<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float sqrtR<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename R> __device__ __forceinline__ R rcpR(R x);
template <> __device__ __forceinline__ double rcpR<double>(double x) { return 1.0 / x; }
template <> __device__ __forceinline__ float rcpR<float>(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename R> __device__ __forceinline__ R fmaxR(R a, R b);
template <> __device__ __forceinline__ double fmaxR<double>(double a, double b) { return fmax(a, b); }
template <> __device__ __forceinline__ float fmaxR<float>(float a, float b) { return fmaxf(a, b); }
template <typename R> __device__ __forceinline__ R fminR(R a, R b);
template <> __device__ __forceinline__ double fminR<double>(double a, double b) { return fmin(a, b); }
template <> __device__ __forceinline__ float fminR<float>(float a, float b) { return fminf(a, b); }

// 4-wide loads and stores of one state (vector memory ops).
__device__ __forceinline__ void load4(const double* p, double* v) {
    const double2 a = reinterpret_cast<const double2*>(p)[0];
    const double2 b = reinterpret_cast<const double2*>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void load4(const float* p, float* v) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void store4(double* p, const double* v) {
    reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void store4(float* p, const float* v) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}

// ---------------------------------------------------------------- physics
// IdealGas::fluxX/fluxY (euler.hpp:101-129), no clamps.
template <typename R>
__device__ __forceinline__ void euler_flux(const R* U, R g, R* Fx, R* Fy) {
    const R inv = rcpR(U[0]);
    const R u = U[1] * inv, v = U[2] * inv, E = U[3] * inv;
    const R p = (g - R(1)) * U[0] * (E - R(0.5) * (u * u + v * v));
    Fx[0] = U[1]; Fx[1] = U[1] * u + p; Fx[2] = U[1] * v; Fx[3] = (U[3] + p) * u;
    Fy[0] = U[2]; Fy[1] = U[2] * u; Fy[2] = U[2] * v + p; Fy[3] = (U[3] + p) * v;
}

template <typename R>
__device__ __forceinline__ void normal_flux(const R* U, R g, R nx, R ny, R* Fn) {
    R fx[NV], fy[NV];
    euler_flux(U, g, fx, fy);
#pragma unroll
    for (int k = 0; k < NV; ++k) Fn[k] = fx[k] * nx + fy[k] * ny;
}

// Inline clamped Rusanov of computeRiemannFluxWithBCKernel (kernels.cu:601-668).
template <typename R>
__device__ __forceinline__ void rusanov(const R* UL, const R* UR, R nx, R ny, R g, R* F) {
    const R tiny = R(1e-10);
    const R rhoL = fmaxR(UL[0], tiny), rhoR = fmaxR(UR[0], tiny);
    const R iL = rcpR(rhoL), iR = rcpR(rhoR);
    const R uL = UL[1] * iL, vL = UL[2] * iL, uR = UR[1] * iR, vR = UR[2] * iR;
    const R EL = UL[3] * iL, ER = UR[3] * iR;
    const R qL = uL * uL + vL * vL, qR = uR * uR + vR * vR;
    const R pL = fmaxR((g - R(1)) * rhoL * (EL - R(0.5) * qL), tiny);
    const R pR = fmaxR((g - R(1)) * rhoR * (ER - R(0.5) * qR), tiny);
    const R cL = sqrtR(fmaxR(g * pL * iL, tiny)), cR = sqrtR(fmaxR(g * pR * iR, tiny));
    const R vnL = uL * nx + vL * ny, vnR = uR * nx + vR * ny;
    const R smax = fmaxR(sqrtR(qL) + cL, sqrtR(qR) + cR);
    const R HL = EL + pL * iL, HR = ER + pR * iR;
    const R FL[NV] = {rhoL * vnL, rhoL * uL * vnL + pL * nx, rhoL * vL * vnL + pL * ny, rhoL * HL * vnL};
    const R FR[NV] = {rhoR * vnR, rhoR * uR * vnR + pR * nx, rhoR * vR * vnR + pR * ny, rhoR * HR * vnR};
    const R hs = R(0.5) * smax;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        R f = R(0.5) * (FL[k] + FR[k]) - hs * (UR[k] - UL[k]);
        F[k] = isfinite(f) ? f : R(0);
    }
}

template <typename R>
__device__ __forceinline__ void prim2cons(R rho, R u, R v, R p, R g, R* U) {
    U[0] = rho; U[1] = rho * u; U[2] = rho * v;
    U[3] = p / (g - R(1)) + R(0.5) * rho * (u * u + v * v);
}

// computeGhostState (kernels.cu:453-525).
template <typename R>
__device__ void ghost_state(const R* Ui, R nx, R ny, int bct, const double* bc, R g, R* G) {
    if (bct == 1 || bct == 2 || bct == 3) {  // Wall, SlipWall, Symmetry: reflect
        const R rho = Ui[0], u = Ui[1] / rho, v = Ui[2] / rho;
        const R p = (g - R(1)) * rho * (Ui[3] / rho - R(0.5) * (u * u + v * v));
        const R vn = u * nx + v * ny;
        prim2cons(rho, u - R(2) * vn * nx, v - R(2) * vn * ny, p, g, G);
    } else if (bct == 4) {  // FarField: averaged pressure with clamps
        const R rinf = R(bc[0]), uinf = R(bc[1]), vinf = R(bc[2]), pinf = R(bc[3]);
        const R ri = fmaxR(Ui[0], R(1e-10));
        const R ui = Ui[1] / ri, vi = Ui[2] / ri;
        const R pi = fmaxR((g - R(1)) * (Ui[3] - R(0.5) * ri * (ui * ui + vi * vi)), R(1e-10));
        const R vninf = uinf * nx + vinf * ny;
        R rb, ub, vb;
        if (vninf > R(0)) { rb = rinf; ub = uinf; vb = vinf; }
        else { rb = ri; ub = ui; vb = vi; }
        R pb = R(0.5) * (pinf + pi);
        rb = fmaxR(fminR(rb, R(100)), R(0.01));
        ub = fmaxR(fminR(ub, R(1000)), R(-1000));
        vb = fmaxR(fminR(vb, R(1000)), R(-1000));
        pb = fmaxR(fminR(pb, R(1e7)), R(1e3));
        prim2cons(rb, ub, vb, pb, g, G);
    } else if (bct == 6) {  // Outflow: prescribed static pressure
        const R rho = Ui[0], u = Ui[1] / rho, v = Ui[2] / rho;
        prim2cons(rho, u, v, R(bc[0]), g, G);
    } else {
#pragma unroll
        for (int k = 0; k < NV; ++k) G[k] = Ui[k];
    }
}

// enforcePositivityKernel (kernels.cu:1236-1288), FP64, gamma fixed at 1.4.
__device__ __forceinline__ void positivity(double* q) {
    constexpr double GM1 = 1.4 - 1.0;
    const double rn = fmax(0.01, fmin(100.0, q[0]));
    const double ir = 1.0 / fmax(rn, 1e-10);
    double u = q[1] * ir, v = q[2] * ir;
    const double vm = sqrt(u * u + v * v);
    if (vm > 2000.0) {
        const double sc = 2000.0 / vm;
        u *= sc; v *= sc;
    }
    const double ke = 0.5 * rn * (u * u + v * v);
    const double p = fmax(100.0, fmin(1e8, GM1 * (q[3] - ke)));
    q[0] = rn; q[1] = rn * u; q[2] = rn * v; q[3] = p / GM1 + ke;
}

// ------------------------------------------------------- wave reductions
// Cross-lane moves through DPP on the two halves of a double.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double read_lane(double x, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

struct OpSum { __device__ static double f(double a, double b) { return a + b; } };
struct OpMin { __device__ static double f(double a, double b) { return fmin(a, b); } };

// Whole-wavefront reduction: butterfly within rows of 16 lanes by DPP
// (quad_perm 1032, quad_perm 2301, row_half_mirror, row_mirror), then the four
// row results by readlane.  Every lane of the wave must call it.
template <typename Op>
__device__ __forceinline__ double wave_reduce(double v) {
    v = Op::f(v, dpp_mov<0xB1>(v));
    v = Op::f(v, dpp_mov<0x4E>(v));
    v = Op::f(v, dpp_mov<0x141>(v));
    v = Op::f(v, dpp_mov<0x140>(v));
    return Op::f(Op::f(read_lane(v, 0), read_lane(v, 16)), Op::f(read_lane(v, 32), read_lane(v, 48)));
}

// Block reduction over BLOCK threads; the result is valid in thread 0.
template <typename Op>
__device__ __forceinline__ double block_reduce(double v, double* sred) {
    v = wave_reduce<Op>(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sred[w] = v;
    __syncthreads();
    double r = sred[0];
    if (threadIdx.x == 0)
        for (int k = 1; k < BLOCK / 64; ++k) r = Op::f(r, sred[k]);
    return r;
}

// ------------------------------------------------------------ operators
// Value of one variable at flux point f of local edge k, from the solution
// points of one element (u[S], xi index fastest).  Edge 0 (eta=-1) and 2
// (eta=+1) run along xi, edges 1 (xi=+1) and 3 (xi=-1) along eta; edges 2 and
// 3 are stored in reversed order (basis.cpp:540-552).  The explicit FMA chain
// fixes the rounding, so both elements of a face see identical traces.
// Where flux point f of edge k reads its line of solution points: first point
// `base`, step `stride`, weights Lm (edges 0, 3) or Lp (edges 1, 2).  Computed
// once per thread without branches.
struct TraceLine {
    int base, stride, plus;
};
template <int P>
__device__ __forceinline__ TraceLine trace_line(int k, int f) {
    constexpr int N1 = P + 1;
    const int ff = k >= 2 ? P - f : f;
    const bool along_xi = (k & 1) != 0;
    return TraceLine{along_xi ? ff * N1 : ff, along_xi ? 1 : N1, (k == 1 || k == 2) ? 1 : 0};
}
template <typename R, int P>
__device__ __forceinline__ R trace_point(const R* u, const TraceLine& t, const R* Lm, const R* Lp) {
    constexpr int N1 = P + 1;
    const R* L = t.plus ? Lp : Lm;
    R acc = L[0] * u[t.base];
#pragma unroll
    for (int m = 1; m < N1; ++m) acc = fmaR(L[m], u[t.base + m * t.stride], acc);
    return acc;
}

// ---------------------------------------------------------------- kernels
// LDS layout of one element tile, chosen per (precision, order) to minimise
// bank conflicts over the kernel's access patterns (32 banks x 4 B):
// solution-point arrays are [element][variable][point] with padded strides;
// trace and flux-difference arrays are either variable-major
// ([element][variable][edge*Fp+point]) or point-major with a padded point
// stride ([element][edge*Fp+point][variable]).
template <typename R, int P> struct Lay;
#define ZJ_LAYOUT(R_, P_, PAD_V, PAD_E, VAR_MAJOR, PAD_T)                                     \
    template <> struct Lay<R_, P_> {                                                           \
        static constexpr int N1 = P_ + 1, S = N1 * N1, FP = N1;                                \
        static constexpr int UV = S + PAD_V, UE = NV * UV + PAD_E;                             \
        static constexpr bool VM = VAR_MAJOR;                                                  \
        static constexpr int TE = (VAR_MAJOR ? 4 * FP * NV : 4 * FP * (NV + 1)) + PAD_T;       \
        __device__ static constexpr int tr(int el, int k, int f, int v) {                      \
            return VAR_MAJOR ? el * TE + v * (4 * FP) + k * FP + f : el * TE + (k * FP + f) * (NV + 1) + v; \
        }                                                                                      \
    };
ZJ_LAYOUT(double, 1, 0, 4, false, 0)
ZJ_LAYOUT(double, 2, 0, 5, false, 0)
ZJ_LAYOUT(double, 3, 0, 0, true, 0)
ZJ_LAYOUT(double, 4, 0, 5, true, 5)
ZJ_LAYOUT(double, 5, 0, 7, false, 0)
ZJ_LAYOUT(float, 1, 0, 4, false, 0)
ZJ_LAYOUT(float, 2, 0, 5, false, 0)
ZJ_LAYOUT(float, 3, 3, 4, false, 0)
ZJ_LAYOUT(float, 4, 1, 5, true, 4)
ZJ_LAYOUT(float, 5, 0, 0, false, 0)
#undef ZJ_LAYOUT

// Workgroups are handed to the XCDs round-robin (workgroup b runs on XCD
// b mod 8), and each XCD has its own L2.  Giving each XCD a contiguous run of
// tiles keeps a tile's mesh neighbours, whose traces and state lines it
// shares, in the same L2.  `groups` = 1 leaves the order unchanged.
__device__ __forceinline__ int tile_of_block(int groups) {
    const int nb = gridDim.x, b = blockIdx.x;
    if (groups <= 1) return b;
    const int q = nb / groups, r = nb - q * groups, x = b % groups, loc = b / groups;
    return x * q + min(x, r) + loc;
}

template <typename R, int P, int EPB, int MODE, bool AFF>
__global__ __launch_bounds__(BLOCK) void stage_kernel(const StageArgs<R> a, const OpTables<R> ops) {
    using Y = Lay<R, P>;
    // FP32-bulk keeps the intermediate stage state as FP32 increments over U^n
    constexpr bool INC = std::is_same<R, float>::value;
    constexpr bool HAS_U0 = MODE == MODE_STAGE2 || MODE == MODE_STAGE3;
    constexpr int N1 = P + 1, S = N1 * N1, FP = N1;
    constexpr int NT = EPB * S, NF = EPB * 4 * FP;
    static_assert(NT <= BLOCK, "tile exceeds the workgroup");
    __shared__ R sU[EPB * Y::UE];
    __shared__ R sFx[EPB * Y::UE];
    __shared__ R sFy[EPB * Y::UE];
    __shared__ R sTr[EPB * Y::TE];
    __shared__ R sFd[EPB * Y::TE];
    __shared__ R sD[N1 * N1], sLm[N1], sLp[N1], sC[4 * S];
    __shared__ double sRed[BLOCK / 64];
    __shared__ unsigned char sOut[EPB * 4];  // edge trace needed outside this tile

    const int tid = threadIdx.x;
    const int e0 = tile_of_block(a.xcd_groups) * EPB;
    const int64_t PS = (int64_t)a.E * S;
    const R g = R(a.gamma);

    for (int t = tid; t < N1 * N1; t += BLOCK) sD[t] = ops.D[t];
    for (int t = tid; t < 4 * S; t += BLOCK) sC[t] = ops.corr[t];
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
        if (AFF) {
            jx = a.Jinv[(int64_t)e * 4]; jy = a.Jinv[(int64_t)e * 4 + 1];
            kx = a.Jinv[(int64_t)e * 4 + 2]; ky = a.Jinv[(int64_t)e * 4 + 3];
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
            if (L.role == ROLE_BOUNDARY) {
                ghost_state(own, nx, ny, L.nbr_edge, a.bc_data + (int64_t)L.bc_idx * 8, g, other);
            } else if (inside) {
#pragma unroll
                for (int v = 0; v < NV; ++v) other[v] = sTr[Y::tr(nl, L.nbr_edge, right ? f : P - f, v)];
            } else {
                load4(a.T_in + (((int64_t)L.nbr * 4 + L.nbr_edge) * FP + (right ? f : P - f)) * NV, other);
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) { UL[v] = right ? other[v] : own[v]; UR[v] = right ? mirr[v] : other[v]; }
            rusanov(UL, UR, nx, ny, g, Fc);
            normal_flux(own, g, nx, ny, Fn);
            // left: F_com - F(U).n; right: -F_com + F(U).n at its own slot f (kernels.cu:397,423)
#pragma unroll
            for (int v = 0; v < NV; ++v) Fd[v] = right ? Fn[v] - Fc[v] : Fc[v] - Fn[v];
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sFd[Y::tr(el2, k, f, v)] = Fd[v];
    }
    __syncthreads();

    // ---- C: divergence, correction, stage update
    double rsq = 0.0;
    if (active) {
        const int i = sp % N1, j = sp / N1;  // xi index, eta index
        if (!AFF) { jx = a.Jinv[gi]; jy = a.Jinv[PS + gi]; kx = a.Jinv[2 * PS + gi]; ky = a.Jinv[3 * PS + gi]; }
        double res[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R* Fx = sFx + el * Y::UE + v * Y::UV;
            const R* Fy = sFy + el * Y::UE + v * Y::UV;
            R div;
            if (AFF) {  // Fx, Fy hold the contravariant fluxes
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
            const R c = sFd[Y::tr(el, 0, i, v)] * sC[sp] + sFd[Y::tr(el, 1, j, v)] * sC[S + sp] +
                        sFd[Y::tr(el, 2, i, v)] * sC[2 * S + sp] + sFd[Y::tr(el, 3, j, v)] * sC[3 * S + sp];
            res[v] = double(-div - c);
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
            if (a.positivity) positivity(un);
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

__global__ void aos_to_soa_kernel(const double* __restrict__ aos, double* __restrict__ soa, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[NV];
    load4(aos + 4 * i, v);
#pragma unroll
    for (int k = 0; k < NV; ++k) soa[k * n + i] = v[k];
}

__global__ void soa_to_aos_kernel(const double* __restrict__ soa, double* __restrict__ aos, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = soa[k * n + i];
    store4(aos + 4 * i, v);
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

// FRSolver::computeTimeStep per point (fr_solver.cpp:799-811, euler.hpp:34-41,166-171).
__global__ __launch_bounds__(BLOCK) void dt_partial_kernel(const double* __restrict__ U,
                                                           const double* __restrict__ hcell,
                                                           int E, int S, double gamma, double cfl,
                                                           double cf, double* __restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double sRed[BLOCK / 64];
    const int64_t n = (int64_t)E * S;
    double m = 1e20;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const double rho = U[i], ru = U[n + i], rv = U[2 * n + i], rE = U[3 * n + i];
        const double u = ru / rho, v = rv / rho;
        const double p = (gamma - 1.0) * rho * (rE / rho - 0.5 * (u * u + v * v));
        const double ws = sqrt(u * u + v * v) + sqrt(gamma * fmax(p, 1e-10) / fmax(rho, 1e-10));
        const double h = hcell[i / S];
        const double d = cfl * h / (ws * cf);
        m = d < m ? d : m;
    }
    const double r = block_reduce<OpMin>(m, sRed);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
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

__global__ __launch_bounds__(BLOCK) void finalize_kernel(const double* __restrict__ partial, int n,
                                                         int op_min, double* __restrict__ out) {
    __shared__ double sRed[BLOCK / 64];
    double acc = op_min ? 1e300 : 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK)
        acc = op_min ? fmin(acc, partial[i]) : acc + partial[i];
    const double r = op_min ? block_reduce<OpMin>(acc, sRed) : block_reduce<OpSum>(acc, sRed);
    if (threadIdx.x == 0) *out = op_min ? r : sqrt(r);
}

}  // namespace

// ------------------------------------------------------------------ launch
int max_epb(int P) { return BLOCK / ((P + 1) * (P + 1)); }
int default_epb(int P) { return max_epb(P); }
int num_blocks(int E, int epb) { return (E + epb - 1) / epb; }

#define ZJ_EPB(P) (BLOCK / (((P) + 1) * ((P) + 1)))

template <typename R, int P, int EPB, bool AFF>
static hipError_t stage_dispatch_mode(int mode, const StageArgs<R>& a, const OpTables<R>& ops, hipStream_t s) {
    const dim3 grid(num_blocks(a.E, EPB)), block(BLOCK);
    switch (mode) {
    case MODE_RESIDUAL: stage_kernel<R, P, EPB, MODE_RESIDUAL, AFF><<<grid, block, 0, s>>>(a, ops); break;
    case MODE_STAGE1: stage_kernel<R, P, EPB, MODE_STAGE1, AFF><<<grid, block, 0, s>>>(a, ops); break;
    case MODE_STAGE2: stage_kernel<R, P, EPB, MODE_STAGE2, AFF><<<grid, block, 0, s>>>(a, ops); break;
    case MODE_STAGE3: stage_kernel<R, P, EPB, MODE_STAGE3, AFF><<<grid, block, 0, s>>>(a, ops); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, int P, bool AFF>
static hipError_t stage_dispatch_epb(int mode, const StageArgs<R>& a, const OpTables<R>& ops, int epb, hipStream_t s) {
    constexpr int M = ZJ_EPB(P);
    if (epb == 0 || epb == M) return stage_dispatch_mode<R, P, M, AFF>(mode, a, ops, s);
    if (M / 2 >= 1 && epb == M / 2) return stage_dispatch_mode<R, P, (M / 2 >= 1 ? M / 2 : 1), AFF>(mode, a, ops, s);
    return hipErrorInvalidValue;
}

template <typename R, int P>
static hipError_t stage_dispatch_aff(int mode, const StageArgs<R>& a, const OpTables<R>& ops, int epb, bool aff,
                                     hipStream_t s) {
    return aff ? stage_dispatch_epb<R, P, true>(mode, a, ops, epb, s)
               : stage_dispatch_epb<R, P, false>(mode, a, ops, epb, s);
}

template <typename R>
hipError_t launch_stage(int P, int mode, const StageArgs<R>& a, const OpTables<R>& ops, int epb, bool affine,
                        hipStream_t s) {
    if (a.E <= 0) return hipSuccess;
    switch (P) {
    case 1: return stage_dispatch_aff<R, 1>(mode, a, ops, epb, affine, s);
    case 2: return stage_dispatch_aff<R, 2>(mode, a, ops, epb, affine, s);
    case 3: return stage_dispatch_aff<R, 3>(mode, a, ops, epb, affine, s);
    case 4: return stage_dispatch_aff<R, 4>(mode, a, ops, epb, affine, s);
    case 5: return stage_dispatch_aff<R, 5>(mode, a, ops, epb, affine, s);
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

template hipError_t launch_stage<double>(int, int, const StageArgs<double>&, const OpTables<double>&, int, bool,
                                        hipStream_t);
template hipError_t launch_stage<float>(int, int, const StageArgs<float>&, const OpTables<float>&, int, bool,
                                       hipStream_t);
template hipError_t launch_trace<double>(int, const TraceArgs<double>&, const OpTables<double>&, int, hipStream_t);
template hipError_t launch_trace<float>(int, const TraceArgs<float>&, const OpTables<float>&, int, hipStream_t);

hipError_t launch_aos_to_soa(const double* aos, double* soa, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    aos_to_soa_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(aos, soa, n);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(const double* soa, double* aos, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    soa_to_aos_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(soa, aos, n);
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
                             double cf, double* partial, int nblk, hipStream_t s) {
    dt_partial_kernel<<<dim3(nblk), dim3(BLOCK), 0, s>>>(U, hcell, E, S, gamma, cfl, cf, partial);
    return hipGetLastError();
}

hipError_t launch_count_nonfinite(const double* U, int64_t n, unsigned long long* count, int nblk, hipStream_t s) {
    nonfinite_kernel<<<dim3(nblk), dim3(BLOCK), 0, s>>>(U, n, count);
    return hipGetLastError();
}

hipError_t launch_finalize(const double* partial, int n, int op_min, double* out, hipStream_t s) {
    finalize_kernel<<<1, BLOCK, 0, s>>>(partial, n, op_min, out);
    return hipGetLastError();
}

}  // namespace zj
