// zj_solver.hip -- the C ABI (include/zhijian_amd.h) over the fused stage
// kernels: device state ownership, operator extraction, launch sequencing and
// hipGraph replay of the SSP-RK3 step.
//
// Device layout (one GPU owns the whole mesh):
//   U, W        state planes [4][E*S] FP64 (U = U^n; W = U1, then U2 in place);
//               in FP32-bulk W holds FP32 increments U1 - U^n, then U2 - U^n
//   T[3]        flux-point traces [E][4][Fp][4] in the compute precision;
//               stage 1 reads T[c] and writes T[c+1], stage 2 T[c+1]->T[c+2],
//               stage 3 T[c+2]->T[c], so every step is the same launch sequence
//   Jinv        4 planes [E*S] in the compute precision; Jinv_e [E][4] as well
//               when every element is affine (constant metric)
//   links       [E][4] EdgeLink; normals [E][4][2] (face normal of the left side)
#include "zhijian_amd.h"
#include "zj_kernels.hpp"
#include "zj_adjoint.hpp"
#include "zj_host_math.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <new>
#include <vector>

using namespace zj;

struct zj_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    int P = 0, N1 = 0, S = 0, E = 0, F = 0;
    int prec = ZJ_PREC_FP64;
    int positivity = 1;
    double gamma = 1.4, cfl = 0.5;
    int epb = 0;
    int xcd_groups = 1;  // workgroup -> tile remap (number of XCDs)
    int nblk = 0;        // workgroups of the stage kernel
    int nred = 0;        // workgroups of the standalone reductions
    double* U = nullptr;
    double* W = nullptr;
    double* Rres = nullptr;
    double* aos = nullptr;
    void* Jinv = nullptr;
    void* Jinv_e = nullptr;  // per-element metric (affine meshes only)
    bool affine_ok = false;  // every element's metric is constant
    bool affine = false;     // kernels read Jinv_e
    bool textbook = false;   // ZJ_COMPAT_TEXTBOOK
    bool viscous = false;    // textbook Navier-Stokes terms (BR2); traces carry U and grad U
    ViscParams vp{};
    int riemann = 0;
    void* edge_f_aff = nullptr;  // textbook: [E][4] sign * |edge|/2 / J (affine elements)
    void* edge_f_pt = nullptr;   // textbook: [E][4] sign * |edge|/2 (per-point metric)
    void* Jpt = nullptr;         // textbook: [E*S] J in the compute precision
    int* perm = nullptr;         // device element e is the caller's element perm[e] (nullptr: same order)
    void* T[3] = {nullptr, nullptr, nullptr};
    EdgeLink* links = nullptr;
    void* normals = nullptr;
    double* bc_data = nullptr;
    double* hcell = nullptr;
    double* d_scalars = nullptr;  // [0] dt, [1] l2, [2] dt_cfl, [3] dt_bar, [4] sum of <F_v cotangent, F_v>,
                                  // [5] kinetic energy
    double* partial = nullptr;     // per-workgroup sums of R^2 of the last residual / stage 3
    double* partial_dt = nullptr;  // per-workgroup minima of the CFL dt
    int l2_blocks = 0;             // valid entries of `partial` (0: no residual yet)
    unsigned long long* d_count = nullptr;
    double* h_pinned = nullptr;
    double dt_on_device = NAN;   // last dt written to d_scalars[0]
    OpTables<double> ops64{};
    OpTables<float> ops32{};
    bool has_state = false;
    bool has_area = false;
    int stage_next = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::string err;
    // host copies kept for zj_set_boundary
    std::vector<EdgeLink> links_host;
    std::vector<int> bface;      // creation-order index of each boundary face
    std::vector<int> bface_slot; // element*4 + local edge of that face
    int n_bc_alloc = 0;          // entries of bc_data allocated on the device
    // adjoint workspace, allocated on first use
    bool adj_ready = false;
    int adj_epb = 0;               // tile of the adjoint launches and their forward recompute
    int adj_nt = 0;                // tiles at adj_epb
    double* A_W1 = nullptr;        // U1 (FP64) or FP32 increment U1 - U^n
    double* A_W2 = nullptr;        // U2, likewise
    double* A_Ub = nullptr;        // partial cotangent of the current state (planes)
    double* A_U0b = nullptr;       // cotangent of U^n within a step
    void* A_Tb[2] = {nullptr, nullptr};  // cross-tile edge cotangents, ping-pong
    double* A_dtb = nullptr;       // per-tile dt cotangents of the three stages
    double* A_mub = nullptr;       // viscous terms: per-tile <F_v cotangent, F_v>, summed since the seed
    unsigned long long* A_idx = nullptr;  // the point that sets the CFL step (zj_dt_cfl_vjp)
    double* slots = nullptr;       // device copies of states (SoA planes), zj_state_slots
    int n_slots = 0;
    // tape of the native reverse sweep: per step of a segment, U^n, U1, U2
    // and their traces at the adjoint tile (tape_steps entries)
    char* tape = nullptr;
    int tape_steps = 0;
    hipEvent_t after_batch = nullptr;  // recorded after the last batch step of a batch holding this handle
    // trailing segments that zj_forward_checkpointed records on the tape as it
    // steps, and what the tape holds from it: segments rec_first ..
    // rec_first + rec_n - 1 of a trajectory of rec_steps steps cut every
    // rec_every, step sizes rec_dts (rec_n = 0: nothing usable)
    int tape_segs = 1;
    int rec_first = 0, rec_n = 0, rec_every = 0, rec_steps = 0;
    std::vector<double> rec_dts;
    double* d_dtb_steps = nullptr;  // dJ/d(dt_n) of every step of a sweep
    int n_dtb_steps = 0;
    int A_cur = 0;
    bool A_has_tb = false;
    bool A_seeded = false;
    // batches that hold this handle read its device buffers through kernel
    // arguments uploaded at batch creation: `gen` moves whenever those
    // arguments change, and tile, metric form and lifetime stay fixed while
    // `in_batches` > 0
    unsigned gen = 0;
    int in_batches = 0;
};

// Boundary link of one face (kernels.cu:584-595): a missing entry falls back
// to a slip wall; far field and outflow read their 8-double entry.
static int boundary_link(zj_solver* h, int f, const int* bc_type, const int* bc_face_map, int n_bc_data,
                         bool have_data, EdgeLink& L) {
    int bct = bc_type ? bc_type[f] : ZJ_BC_SLIPWALL;
    int bci = bc_face_map ? bc_face_map[f] : -1;
    if (bci < 0) { bct = ZJ_BC_SLIPWALL; bci = 0; }
    const bool needs_data = (bct == ZJ_BC_FARFIELD || bct == ZJ_BC_OUTFLOW);
    if (needs_data && (!have_data || bci >= n_bc_data)) {
        h->err = "boundary face refers to missing bc_data";
        return ZJ_ERR_ARG;
    }
    L = EdgeLink{-1, bct, ROLE_BOUNDARY, needs_data ? bci : 0};
    return ZJ_OK;
}

static thread_local std::string g_create_error;   // error text of handle-less calls, per thread

namespace zj {
// Error text of a handle-less call in another source file (zj_last_error(NULL)).
void set_thread_error(const std::string& msg) { g_create_error = msg; }
}  // namespace zj

#define ZJ_HIP(h, call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return ZJ_ERR_HIP;                                                           \
        }                                                                                \
    } while (0)

static int fail(zj_solver* h, int code, const std::string& msg) {
    h->err = msg;
    return code;
}

// The stream a call runs on: the caller's, or the handle's own.  Work queued
// there waits for the last batch step that wrote the handle's buffers (a
// batch steps on its own stream, asynchronously).
static hipStream_t pick(zj_solver* h, void* s) {
    hipStream_t st = s ? (hipStream_t)s : h->stream;
    if (h->after_batch) (void)hipStreamWaitEvent(st, h->after_batch, 0);
    return st;
}

// ---------------------------------------------------------------- tables
// The reference's dense tables (basis.cpp:508-569) are tensor products of 1D
// Gauss-point operators; extract the 1D factors and check every entry.
static bool extract_ops(const zj_mesh_desc* m, int P, OpTables<double>& o, std::string& why) {
    const int n = P + 1, S = n * n;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) o.D[i * n + k] = m->diff_xi[(size_t)i * S + k];
    for (int j = 0; j < n; ++j) o.Lm[j] = m->interp[(size_t)(0 * n + 0) * S + j * n + 0];
    for (int i = 0; i < n; ++i) o.Lp[i] = m->interp[(size_t)(1 * n + 0) * S + 0 * n + i];
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int row = j * n + i;
            for (int c = 0; c < S; ++c) {
                const int cj = c / n, ci = c % n;
                const double ex = (cj == j) ? o.D[i * n + ci] : 0.0;
                const double ee = (ci == i) ? o.D[j * n + cj] : 0.0;
                if (m->diff_xi[(size_t)row * S + c] != ex || m->diff_eta[(size_t)row * S + c] != ee) {
                    why = "differentiation tables are not tensor products of one 1D matrix";
                    return false;
                }
            }
        }
    for (int k = 0; k < 4; ++k)
        for (int f = 0; f < n; ++f)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double ex;
                    if (k == 0) ex = (i == f) ? o.Lm[j] : 0.0;
                    else if (k == 1) ex = (j == f) ? o.Lp[i] : 0.0;
                    else if (k == 2) ex = (i == P - f) ? o.Lp[j] : 0.0;
                    else ex = (j == P - f) ? o.Lm[i] : 0.0;
                    if (m->interp[(size_t)(k * n + f) * S + j * n + i] != ex) {
                        why = "interpolation tables do not place flux points at the solution-point "
                              "abscissae (edges 0/1 forward, 2/3 reversed)";
                        return false;
                    }
                }
    for (int t = 0; t < 4 * S; ++t) o.corr[t] = m->corr[t];
    return true;
}

static void ops_to_f32(const OpTables<double>& a, OpTables<float>& b) {
    for (int t = 0; t < 36; ++t) b.D[t] = (float)a.D[t];
    for (int t = 0; t < 6; ++t) { b.Lm[t] = (float)a.Lm[t]; b.Lp[t] = (float)a.Lp[t]; }
    for (int t = 0; t < 4 * 36; ++t) b.corr[t] = (float)a.corr[t];
    for (int t = 0; t < 6; ++t) { b.gL[t] = (float)a.gL[t]; b.gR[t] = (float)a.gR[t]; }
}

template <typename T>
static int upload(zj_solver* h, void** dst, const std::vector<T>& v) {
    ZJ_HIP(h, hipMalloc(dst, std::max<size_t>(v.size(), 1) * sizeof(T)));
    if (!v.empty()) ZJ_HIP(h, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return ZJ_OK;
}

static void release(zj_solver* h) {
    if (h->gexec) hipGraphExecDestroy(h->gexec);
    if (h->graph) hipGraphDestroy(h->graph);
    void* bufs[] = {h->perm, h->edge_f_aff, h->edge_f_pt, h->Jpt, h->U, h->W, h->Rres, h->aos, h->Jinv, h->Jinv_e, h->T[0], h->T[1], h->T[2], h->links,
                    h->normals, h->bc_data, h->hcell, h->d_scalars, h->partial, h->partial_dt, h->d_count,
                    h->A_W1, h->A_W2, h->A_Ub, h->A_U0b, h->A_Tb[0], h->A_Tb[1], h->A_dtb, h->A_mub, h->A_idx, h->slots, h->tape, h->d_dtb_steps};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (h->h_pinned) hipHostFree(h->h_pinned);
    if (h->stream) hipStreamDestroy(h->stream);
}

static int build_ordered(zj_solver* h, const zj_mesh_desc* m, const zj_params* p);

static int build(zj_solver* h, const zj_mesh_desc* m, const zj_params* p) {
    if (!m->elem_centroid || m->n_elem <= 1 || m->poly_order < 1 || m->poly_order > 5)
        return build_ordered(h, m, p);
    // the same mesh with its elements renumbered; faces keep their order
    const int E = m->n_elem, F = m->n_faces, S = (m->poly_order + 1) * (m->poly_order + 1);
    const std::vector<int> perm = zj::host::hilbert_order(E, m->elem_centroid);
    std::vector<int> inv((size_t)E);
    for (int e = 0; e < E; ++e) inv[(size_t)perm[(size_t)e]] = e;
    auto per_elem = [&](const double* src, int w) {
        std::vector<double> out;
        if (!src) return out;
        out.resize((size_t)E * w);
        for (int e = 0; e < E; ++e)
            std::copy(src + (size_t)perm[(size_t)e] * w, src + (size_t)perm[(size_t)e] * w + w, out.begin() + (size_t)e * w);
        return out;
    };
    const std::vector<double> J = per_elem(m->J, S), Jinv = per_elem(m->Jinv, 4 * S);
    const std::vector<double> area = per_elem(m->elem_area, 1), ehl = per_elem(m->edge_half_len, 4);
    std::vector<int> le((size_t)F), re((size_t)F);
    for (int f = 0; f < F; ++f) {
        const int a = m->face_left_elem ? m->face_left_elem[f] : -1;
        const int b = m->face_right_elem ? m->face_right_elem[f] : -1;
        if (a < 0 || a >= E || b >= E) return fail(h, ZJ_ERR_ARG, "face element index out of range");
        le[(size_t)f] = inv[(size_t)a];
        re[(size_t)f] = b >= 0 ? inv[(size_t)b] : b;
    }
    zj_mesh_desc pm = *m;
    pm.J = m->J ? J.data() : nullptr;
    pm.Jinv = m->Jinv ? Jinv.data() : nullptr;
    pm.elem_area = m->elem_area ? area.data() : nullptr;
    pm.edge_half_len = m->edge_half_len ? ehl.data() : nullptr;
    pm.face_left_elem = F ? le.data() : m->face_left_elem;
    pm.face_right_elem = F ? re.data() : m->face_right_elem;
    pm.elem_centroid = nullptr;
    int rc = build_ordered(h, &pm, p);
    if (rc) return rc;
    std::vector<int> pv = perm;
    return upload(h, (void**)&h->perm, pv);
}

static int build_ordered(zj_solver* h, const zj_mesh_desc* m, const zj_params* p) {
    const int P = m->poly_order;
    if (P < 1 || P > 5) return fail(h, ZJ_ERR_ARG, "poly_order must be 1..5");
    if (m->n_elem < 0 || m->n_faces < 0) return fail(h, ZJ_ERR_ARG, "negative mesh sizes");
    if (!m->diff_xi || !m->diff_eta || !m->interp || !m->corr || (m->n_elem > 0 && !m->Jinv))
        return fail(h, ZJ_ERR_ARG, "operator or geometry tables missing");
    if (m->n_faces > 0 && (!m->normals || !m->face_left_elem || !m->face_left_local ||
                           !m->face_right_elem || !m->face_right_local))
        return fail(h, ZJ_ERR_ARG, "face tables missing");
    if (p->precision != ZJ_PREC_FP64 && p->precision != ZJ_PREC_FP32BULK)
        return fail(h, ZJ_ERR_ARG, "unknown precision");
    if (p->compat != ZJ_COMPAT_REFERENCE && p->compat != ZJ_COMPAT_TEXTBOOK)
        return fail(h, ZJ_ERR_ARG, "unknown compat mode");
    if (p->compat == ZJ_COMPAT_TEXTBOOK) {
        if (m->n_elem > 0 && (!m->J || !m->edge_half_len))
            return fail(h, ZJ_ERR_ARG, "textbook mode needs J and edge_half_len");
        if (p->riemann < 0 || p->riemann > 2 || p->corr < 0 || p->corr > 3)
            return fail(h, ZJ_ERR_ARG, "unknown Riemann solver or correction function");
    }
    if (!(p->gamma > 1.0)) return fail(h, ZJ_ERR_ARG, "gamma must exceed 1");
    if ((int64_t)m->n_elem * (P + 1) * (P + 1) * 4 > (int64_t)INT32_MAX * 4)
        return fail(h, ZJ_ERR_UNSUPPORTED, "mesh too large for 32-bit element indices");

    h->P = P; h->N1 = P + 1; h->S = h->N1 * h->N1; h->E = m->n_elem; h->F = m->n_faces;
    h->prec = p->precision; h->positivity = p->positivity ? 1 : 0;
    h->gamma = p->gamma; h->cfl = p->cfl; h->device = p->device;
    h->epb = default_epb(P);
    h->nblk = num_blocks(std::max(h->E, 1), h->epb);

    std::string why;
    if (!extract_ops(m, P, h->ops64, why)) return fail(h, ZJ_ERR_UNSUPPORTED, why);
    h->textbook = p->compat == ZJ_COMPAT_TEXTBOOK;
    h->riemann = p->riemann;
    if (h->textbook && p->viscous) {
        if (!(p->Re > 0.0) || !(p->Pr > 0.0)) return fail(h, ZJ_ERR_ARG, "viscous terms need Re > 0 and Pr > 0");
        const double mu = 1.0 / p->Re;
        h->viscous = true;
        h->vp = ViscParams{mu, mu * p->gamma / ((p->gamma - 1.0) * p->Pr), p->br2_eta > 0.0 ? p->br2_eta : 2.0,
                           zj::host::vcjh_left_d(P, p->corr, -1.0), zj::host::vcjh_right_d(P, p->corr, 1.0)};
    }
    if (h->textbook) {  // VCJH g_L', g_R' at the 1D Gauss points
        double x[6], w[6];
        zj::host::gauss_rule(P + 1, x, w);
        for (int i = 0; i <= P; ++i) {
            h->ops64.gL[i] = zj::host::vcjh_left_d(P, p->corr, x[i]);
            h->ops64.gR[i] = zj::host::vcjh_right_d(P, p->corr, x[i]);
        }
    }
    ops_to_f32(h->ops64, h->ops32);

    // element-edge links from the face list (fr_solver.cpp:240-250, kernels.cu:549-595)
    const int E = h->E, F = h->F, Fp = h->N1;
    std::vector<EdgeLink> links((size_t)E * 4, EdgeLink{-1, 0, ROLE_NONE, 0});
    std::vector<double> nrm((size_t)E * 8, 0.0);
    for (int f = 0; f < F; ++f) {
        const int le = m->face_left_elem[f], ll = m->face_left_local[f];
        const int re = m->face_right_elem[f], rl = m->face_right_local[f];
        if (le < 0 || le >= E || ll < 0 || ll > 3) return fail(h, ZJ_ERR_ARG, "face left side out of range");
        EdgeLink& L = links[(size_t)le * 4 + ll];
        if (L.role != ROLE_NONE) return fail(h, ZJ_ERR_ARG, "an element edge belongs to two faces");
        const double nx = m->normals[2 * f], ny = m->normals[2 * f + 1];
        nrm[((size_t)le * 4 + ll) * 2] = nx; nrm[((size_t)le * 4 + ll) * 2 + 1] = ny;
        if (re >= 0) {
            if (re >= E || rl < 0 || rl > 3) return fail(h, ZJ_ERR_ARG, "face right side out of range");
            EdgeLink& Rr = links[(size_t)re * 4 + rl];
            if (Rr.role != ROLE_NONE || (re == le && rl == ll))
                return fail(h, ZJ_ERR_ARG, "an element edge belongs to two faces");
            L = EdgeLink{re, rl, ROLE_LEFT, 0};
            Rr = EdgeLink{le, ll, ROLE_RIGHT, 0};
            nrm[((size_t)re * 4 + rl) * 2] = nx; nrm[((size_t)re * 4 + rl) * 2 + 1] = ny;
        } else {
            int rc = boundary_link(h, f, m->bc_type, m->bc_face_map, m->n_bc_data, m->bc_data != nullptr, L);
            if (rc) return rc;
            h->bface.push_back(f);
            h->bface_slot.push_back(le * 4 + ll);
        }
    }

    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    int nxcc = 1;
    if (hipDeviceGetAttribute(&nxcc, hipDeviceAttributeNumberOfXccs, h->device) != hipSuccess || nxcc < 1) nxcc = 1;
    h->xcd_groups = nxcc;

    const size_t npts = (size_t)E * h->S;
    const size_t rsz = h->prec == ZJ_PREC_FP64 ? sizeof(double) : sizeof(float);
    ZJ_HIP(h, hipMalloc((void**)&h->U, std::max<size_t>(npts * 4, 1) * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->W, std::max<size_t>(npts * 4, 1) * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->Rres, std::max<size_t>(npts * 4, 1) * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->aos, std::max<size_t>(npts * 4, 1) * sizeof(double)));
    if (h->viscous)
        for (const EdgeLink& L : links)
            if (L.role != ROLE_LEFT && L.role != ROLE_RIGHT)
                return fail(h, ZJ_ERR_UNSUPPORTED, "viscous terms are built for meshes without boundary faces");
    const size_t tw = h->viscous ? 12 : 4;  // trace words per flux point
    for (int k = 0; k < 3; ++k) ZJ_HIP(h, hipMalloc(&h->T[k], std::max<size_t>((size_t)E * 4 * Fp * tw, 1) * rsz));

    // Jinv: [E][S][4] -> 4 planes, converted to the compute precision
    std::vector<double> jp(npts * 4);
    for (size_t g = 0; g < npts; ++g)
        for (int c = 0; c < 4; ++c) jp[c * npts + g] = m->Jinv[g * 4 + c];
    // Affine detection: an element whose per-point metric agrees with its first
    // point's to 1e-13 of its largest entry is stored once (the bilinear map of a
    // parallelogram has a constant Jacobian; the tables differ only by rounding).
    std::vector<double> je((size_t)E * 4);
    h->affine_ok = E > 0;
    for (int e = 0; e < E && h->affine_ok; ++e) {
        const double* J0 = m->Jinv + (size_t)e * h->S * 4;
        double mx = 0.0;
        for (int c = 0; c < 4; ++c) mx = std::max(mx, std::fabs(J0[c]));
        for (int sp = 1; sp < h->S && h->affine_ok; ++sp)
            for (int c = 0; c < 4; ++c)
                if (!(std::fabs(J0[sp * 4 + c] - J0[c]) <= 1e-13 * mx)) { h->affine_ok = false; break; }
        for (int c = 0; c < 4; ++c) je[(size_t)e * 4 + c] = J0[c];
    }
    if (!h->affine_ok) je.clear();
    h->affine = h->affine_ok;
    if (h->viscous && !h->affine_ok)
        return fail(h, ZJ_ERR_UNSUPPORTED, "viscous terms are built for affine elements (constant metric)");
    // textbook correction factors: sign (-1 on edges 0, 3) * half edge length [/ J]
    std::vector<double> efa, efp, jpt;
    if (h->textbook) {
        const double sgn[4] = {-1.0, 1.0, 1.0, -1.0};
        efa.resize((size_t)E * 4); efp.resize((size_t)E * 4);
        for (int e = 0; e < E; ++e)
            for (int k = 0; k < 4; ++k) {
                efp[(size_t)e * 4 + k] = sgn[k] * m->edge_half_len[(size_t)e * 4 + k];
                efa[(size_t)e * 4 + k] = efp[(size_t)e * 4 + k] / m->J[(size_t)e * h->S];
            }
        jpt.assign(m->J, m->J + npts);
    }
    int rc;
    if (h->prec == ZJ_PREC_FP64) {
        if ((rc = upload(h, &h->Jinv, jp))) return rc;
        if ((rc = upload(h, &h->Jinv_e, je))) return rc;
        if ((rc = upload(h, &h->normals, nrm))) return rc;
        if ((rc = upload(h, &h->edge_f_aff, efa))) return rc;
        if ((rc = upload(h, &h->edge_f_pt, efp))) return rc;
        if ((rc = upload(h, &h->Jpt, jpt))) return rc;
    } else {
        std::vector<float> efaf(efa.begin(), efa.end()), efpf(efp.begin(), efp.end()), jptf(jpt.begin(), jpt.end());
        if ((rc = upload(h, &h->edge_f_aff, efaf))) return rc;
        if ((rc = upload(h, &h->edge_f_pt, efpf))) return rc;
        if ((rc = upload(h, &h->Jpt, jptf))) return rc;
        std::vector<float> jf(jp.begin(), jp.end()), jef(je.begin(), je.end()), nf(nrm.begin(), nrm.end());
        if ((rc = upload(h, &h->Jinv, jf))) return rc;
        if ((rc = upload(h, &h->Jinv_e, jef))) return rc;
        if ((rc = upload(h, &h->normals, nf))) return rc;
    }
    if ((rc = upload(h, (void**)&h->links, links))) return rc;
    h->links_host = std::move(links);
    std::vector<double> bcd;
    if (m->bc_data && m->n_bc_data > 0) bcd.assign(m->bc_data, m->bc_data + (size_t)m->n_bc_data * 8);
    if ((rc = upload(h, (void**)&h->bc_data, bcd))) return rc;
    h->n_bc_alloc = (int)(bcd.size() / 8);
    std::vector<double> hc;
    if (m->elem_area) {
        hc.resize(E);
        for (int e = 0; e < E; ++e) hc[e] = std::sqrt(m->elem_area[e]);  // fr_solver.cpp:799
        h->has_area = true;
    }
    if ((rc = upload(h, (void**)&h->hcell, hc))) return rc;

    h->nred = (int)std::min<int64_t>(1024, std::max<int64_t>(1, ((int64_t)npts + BLOCK - 1) / BLOCK));
    const int npart = std::max(std::max(h->nblk, num_blocks(std::max(E, 1), 1)), h->nred);
    ZJ_HIP(h, hipMalloc((void**)&h->partial, (size_t)npart * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->partial_dt, (size_t)h->nred * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->d_scalars, 8 * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->d_count, sizeof(unsigned long long)));
    ZJ_HIP(h, hipHostMalloc((void**)&h->h_pinned, 8 * sizeof(double), hipHostMallocDefault));
    ZJ_HIP(h, hipMemset(h->d_scalars, 0, 8 * sizeof(double)));
    ZJ_HIP(h, hipDeviceSynchronize());
    return ZJ_OK;
}

extern "C" int zj_version(void) { return 1; }

extern "C" int zj_create(zj_handle* out, const zj_mesh_desc* m, const zj_params* p) {
    if (!out || !m || !p) {
        g_create_error = "null argument";
        return ZJ_ERR_ARG;
    }
    *out = nullptr;
    zj_solver* h = new (std::nothrow) zj_solver();
    if (!h) {
        g_create_error = "out of host memory";
        return ZJ_ERR_NOMEM;
    }
    int rc;
    try {
        rc = build(h, m, p);
    } catch (const std::bad_alloc&) {   // host tables of a mesh too large for host memory
        h->err = "out of host memory";
        rc = ZJ_ERR_NOMEM;
    }
    if (rc != ZJ_OK) {
        g_create_error = h->err;
        release(h);
        delete h;
        return rc;
    }
    *out = h;
    return ZJ_OK;
}

extern "C" int zj_destroy(zj_handle h) {
    if (!h) return ZJ_ERR_ARG;
    if (h->in_batches > 0) return fail(h, ZJ_ERR_STATE, "the handle belongs to a batch: destroy the batch first");
    hipSetDevice(h->device);
    hipStreamSynchronize(h->stream);
    release(h);
    delete h;
    return ZJ_OK;
}

extern "C" const char* zj_last_error(zj_handle h) { return h ? h->err.c_str() : g_create_error.c_str(); }

// ---------------------------------------------------------------- launches
template <typename R>
static StageArgs<R> stage_args(zj_solver* h, int mode) {
    StageArgs<R> a{};
    a.E = h->E;
    a.Jinv = (const R*)(h->affine ? h->Jinv_e : h->Jinv);
    a.links = h->links;
    a.normals = (const R*)h->normals;
    a.bc_data = h->bc_data;
    a.dt_ptr = h->d_scalars;
    a.gamma = h->gamma;
    a.positivity = h->positivity;
    a.xcd_groups = h->xcd_groups;
    a.edge_f = (const R*)(h->affine ? h->edge_f_aff : h->edge_f_pt);
    a.Jpt = (const R*)h->Jpt;
    a.riemann = h->riemann;
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
        a.U0 = h->U; a.U_in = h->W; a.U_out = h->U; a.T_in = T[2]; a.T_out = T[0];
        a.l2_partial = h->partial;
        break;
    }
    if (std::is_same<R, float>::value) {  // FP32-bulk: W holds FP32 increments over U^n
        float* dw = (float*)h->W;
        if (mode == MODE_STAGE1) { a.dW_out = dw; a.U_out = nullptr; }
        if (mode == MODE_STAGE2) { a.dW_in = dw; a.dW_out = dw; a.U_in = nullptr; a.U_out = nullptr; }
        if (mode == MODE_STAGE3) { a.dW_in = dw; a.U_in = nullptr; }
    }
    return a;
}

static hipError_t enqueue_stage(zj_solver* h, int mode, hipStream_t s) {
    hipError_t e;
    if (h->viscous) {
        e = h->prec == ZJ_PREC_FP64 ? launch_visc_stage<double>(h->P, mode, stage_args<double>(h, mode), h->ops64, h->vp, s)
                                    : launch_visc_stage<float>(h->P, mode, stage_args<float>(h, mode), h->ops32, h->vp, s);
        if (e == hipSuccess && (mode == MODE_RESIDUAL || mode == MODE_STAGE3)) h->l2_blocks = h->nblk;
        return e;
    }
    if (h->prec == ZJ_PREC_FP64)
        e = launch_stage<double>(h->P, mode, stage_args<double>(h, mode), h->ops64, h->epb, h->affine, s, h->textbook);
    else
        e = launch_stage<float>(h->P, mode, stage_args<float>(h, mode), h->ops32, h->epb, h->affine, s, h->textbook);
    if (e == hipSuccess && (mode == MODE_RESIDUAL || mode == MODE_STAGE3)) h->l2_blocks = h->nblk;
    return e;
}

// The step size lives in device memory so that the captured step graph serves
// every dt; it is rewritten only when it changes.
static hipError_t set_dt(zj_solver* h, double dt, hipStream_t s) {
    if (dt == h->dt_on_device) return hipSuccess;
    h->dt_on_device = dt;
    return launch_set_scalar(h->d_scalars, dt, s);
}

static int refresh_trace(zj_solver* h, hipStream_t s) {
    hipError_t e;
    if (h->viscous)
        e = h->prec == ZJ_PREC_FP64
                ? launch_visc_trace<double>(h->P, h->E, h->U, (const double*)h->Jinv_e, (double*)h->T[0], h->ops64,
                                            h->xcd_groups, s)
                : launch_visc_trace<float>(h->P, h->E, h->U, (const float*)h->Jinv_e, (float*)h->T[0], h->ops32,
                                           h->xcd_groups, s);
    else if (h->prec == ZJ_PREC_FP64)
        e = launch_trace<double>(h->P, TraceArgs<double>{h->E, h->U, (double*)h->T[0], h->xcd_groups}, h->ops64, 0, s);
    else
        e = launch_trace<float>(h->P, TraceArgs<float>{h->E, h->U, (float*)h->T[0], h->xcd_groups}, h->ops32, 0, s);
    ZJ_HIP(h, e);
    return ZJ_OK;
}

static void drop_graph(zj_solver* h) {
    if (h->gexec) { hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
    if (h->graph) { hipGraphDestroy(h->graph); h->graph = nullptr; }
}

extern "C" int zj_set_state_device(zj_handle h, const double* d_aos, void* stream) {
    if (!h || !d_aos) return ZJ_ERR_ARG;
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    ZJ_HIP(h, launch_aos_to_soa(d_aos, h->U, (int64_t)h->E * h->S, s, h->perm, h->S));
    int rc = refresh_trace(h, s);
    if (rc) return rc;
    h->has_state = true;
    h->stage_next = 1;
    return ZJ_OK;
}

// ---- device state slots: checkpoints that never leave the GPU
extern "C" int zj_state_slots(zj_handle h, int n) {
    if (!h || n < 0) return ZJ_ERR_ARG;
    h->rec_n = 0;
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    if (h->slots) {
        ZJ_HIP(h, hipFree(h->slots));
        h->slots = nullptr;
        h->n_slots = 0;
    }
    if (n > 0) {
        const size_t bytes = (size_t)n * h->E * h->S * 4 * sizeof(double);
        if (hipMalloc((void**)&h->slots, bytes) != hipSuccess) {
            h->slots = nullptr;
            return fail(h, ZJ_ERR_NOMEM, "state slots: device allocation failed");
        }
        h->n_slots = n;
    }
    return ZJ_OK;
}

extern "C" int zj_state_save(zj_handle h, int slot, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (slot < 0 || slot >= h->n_slots) return fail(h, ZJ_ERR_ARG, "slot out of range");
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "state requested in the middle of a step");
    ZJ_HIP(h, hipSetDevice(h->device));
    const size_t n = (size_t)h->E * h->S * 4;
    ZJ_HIP(h, hipMemcpyAsync(h->slots + (size_t)slot * n, h->U, n * sizeof(double), hipMemcpyDeviceToDevice,
                             pick(h, stream)));
    h->rec_n = 0;  // a recorded segment may start from this slot
    return ZJ_OK;
}

extern "C" int zj_state_load(zj_handle h, int slot, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (slot < 0 || slot >= h->n_slots) return fail(h, ZJ_ERR_ARG, "slot out of range");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "state change in the middle of a step");
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    const size_t n = (size_t)h->E * h->S * 4;
    ZJ_HIP(h, hipMemcpyAsync(h->U, h->slots + (size_t)slot * n, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    int rc = refresh_trace(h, s);
    if (rc) return rc;
    h->has_state = true;
    return ZJ_OK;
}

extern "C" int zj_set_state(zj_handle h, const double* U_aos) {
    if (!h || !U_aos) return ZJ_ERR_ARG;
    ZJ_HIP(h, hipSetDevice(h->device));
    const size_t n = (size_t)h->E * h->S * 4;
    ZJ_HIP(h, hipMemcpyAsync(h->aos, U_aos, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    int rc = zj_set_state_device(h, h->aos, nullptr);
    if (rc) return rc;
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

extern "C" int zj_get_state_device(zj_handle h, double* d_aos, void* stream) {
    if (!h || !d_aos) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "state requested in the middle of a step");
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, launch_soa_to_aos(h->U, d_aos, (int64_t)h->E * h->S, pick(h, stream), h->perm, h->S));
    return ZJ_OK;
}

extern "C" int zj_get_state(zj_handle h, double* U_aos) {
    if (!h || !U_aos) return ZJ_ERR_ARG;
    int rc = zj_get_state_device(h, h->aos, nullptr);
    if (rc) return rc;
    const size_t n = (size_t)h->E * h->S * 4;
    ZJ_HIP(h, hipMemcpyAsync(U_aos, h->aos, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

extern "C" int zj_residual(zj_handle h, double* dUdt_aos, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "residual requested in the middle of a step");
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    ZJ_HIP(h, enqueue_stage(h, MODE_RESIDUAL, s));
    if (dUdt_aos) {
        ZJ_HIP(h, launch_soa_to_aos(h->Rres, h->aos, (int64_t)h->E * h->S, s, h->perm, h->S));
        ZJ_HIP(h, hipMemcpyAsync(dUdt_aos, h->aos, (size_t)h->E * h->S * 4 * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
        ZJ_HIP(h, hipStreamSynchronize(s));
    }
    return ZJ_OK;
}

extern "C" int zj_residual_device(zj_handle h, double* d_dUdt_aos, void* stream) {
    if (!h || !d_dUdt_aos) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "residual requested in the middle of a step");
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    ZJ_HIP(h, enqueue_stage(h, MODE_RESIDUAL, s));
    ZJ_HIP(h, launch_soa_to_aos(h->Rres, d_dUdt_aos, (int64_t)h->E * h->S, s, h->perm, h->S));
    return ZJ_OK;
}

extern "C" int zj_rk3_stage(zj_handle h, int stage, double dt, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (stage != h->stage_next) return fail(h, ZJ_ERR_STATE, "stages must run in the order 1, 2, 3");
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    ZJ_HIP(h, set_dt(h, dt, s));
    ZJ_HIP(h, enqueue_stage(h, stage, s));
    h->stage_next = stage == 3 ? 1 : stage + 1;
    return ZJ_OK;
}

extern "C" int zj_time_stages(zj_handle h, double dt, int steps, double* stage_ms, void* stream) {
    if (!h || steps < 1 || !stage_ms) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "a partial step is in progress");
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    ZJ_HIP(h, set_dt(h, dt, s));
    ZJ_HIP(h, time_stage_kernels([&](int st) { return enqueue_stage(h, st, s); }, steps, s, stage_ms));
    return ZJ_OK;
}

static int capture_step(zj_solver* h) {
    drop_graph(h);
    ZJ_HIP(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int st = 1; st <= 3 && e == hipSuccess; ++st) e = enqueue_stage(h, st, h->stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        ZJ_HIP(h, e != hipSuccess ? e : e2);
    }
    h->graph = g;
    ZJ_HIP(h, hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
    return ZJ_OK;
}

extern "C" int zj_rk3_step(zj_handle h, double dt, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "a partial step is in progress");
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    if (!h->gexec) {
        int rc = capture_step(h);
        if (rc) return rc;
    }
    ZJ_HIP(h, set_dt(h, dt, s));
    ZJ_HIP(h, hipGraphLaunch(h->gexec, s));
    return ZJ_OK;
}

extern "C" int zj_dt_cfl(zj_handle h, double cfl, double* dt) {
    if (!h || !dt) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (!h->has_area) return fail(h, ZJ_ERR_STATE, "element areas were not provided");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "dt requested in the middle of a step");
    ZJ_HIP(h, hipSetDevice(h->device));
    // BR2 on quadrilaterals: the viscous operator's spectral radius is below
    // 2.0 (p+1)^4 nu_k / h^2 (h the side; measured on the oracle's Jacobian, P1-P5)
    const double p1 = h->P + 1.0;
    const double visc =
        h->viscous ? p1 * p1 * p1 * p1 * std::max(h->vp.mu * 4.0 / 3.0, h->vp.kappa * (h->gamma - 1.0)) : 0.0;
    (void)pick(h, nullptr);
    ZJ_HIP(h, launch_dt_partial(h->U, h->hcell, h->E, h->S, h->gamma, cfl, 2.0 * h->P + 1.0, visc, h->partial_dt,
                                h->nred, h->stream));
    ZJ_HIP(h, launch_finalize(h->partial_dt, h->nred, FIN_MIN, h->d_scalars + 2, h->stream));
    ZJ_HIP(h, hipMemcpyAsync(h->h_pinned, h->d_scalars + 2, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    *dt = h->h_pinned[0];
    return ZJ_OK;
}

// The CFL step's vector-Jacobian product: the held adjoint cotangent gains
// dt_bar * d(dt_cfl)/dU at the held state (the derivative of the minimum is
// that of its minimising point, the lowest index among ties).
extern "C" int zj_dt_cfl_vjp(zj_handle h, double cfl, double dt_bar, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (!h->has_area) return fail(h, ZJ_ERR_STATE, "element areas were not provided");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "a partial step is in progress");
    if (!h->A_seeded) return fail(h, ZJ_ERR_STATE, "no adjoint seed set");
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    const double p1 = h->P + 1.0;
    const double visc =
        h->viscous ? p1 * p1 * p1 * p1 * std::max(h->vp.mu * 4.0 / 3.0, h->vp.kappa * (h->gamma - 1.0)) : 0.0;
    const double cf = 2.0 * h->P + 1.0;
    ZJ_HIP(h, launch_dt_partial(h->U, h->hcell, h->E, h->S, h->gamma, cfl, cf, visc, h->partial_dt, h->nred, s));
    ZJ_HIP(h, launch_finalize(h->partial_dt, h->nred, FIN_MIN, h->d_scalars + 2, s));
    ZJ_HIP(h, launch_dt_cfl_vjp(h->U, h->hcell, h->E, h->S, h->gamma, cfl, cf, visc, h->d_scalars + 2, h->A_idx,
                                dt_bar, h->A_Ub, h->nred, s));
    return ZJ_OK;
}

extern "C" int zj_residual_l2(zj_handle h, double* l2) {
    if (!h || !l2) return ZJ_ERR_ARG;
    if (h->l2_blocks == 0) return fail(h, ZJ_ERR_STATE, "no residual has been evaluated");
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, hipDeviceSynchronize());  // the last residual may have run on a caller's stream
    ZJ_HIP(h, launch_finalize(h->partial, h->l2_blocks, FIN_SQRT_SUM, h->d_scalars + 1, h->stream));
    ZJ_HIP(h, hipMemcpyAsync(h->h_pinned, h->d_scalars + 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    *l2 = h->h_pinned[0];
    return ZJ_OK;
}

extern "C" int zj_check_finite(zj_handle h, int64_t* n_bad) {
    if (!h || !n_bad) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    ZJ_HIP(h, hipSetDevice(h->device));
    (void)pick(h, nullptr);
    ZJ_HIP(h, hipMemsetAsync(h->d_count, 0, sizeof(unsigned long long), h->stream));
    ZJ_HIP(h, launch_count_nonfinite(h->U, (int64_t)h->E * h->S * 4, h->d_count, h->nred, h->stream));
    unsigned long long c = 0;
    ZJ_HIP(h, hipMemcpyAsync(&c, h->d_count, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    *n_bad = (int64_t)c;
    return ZJ_OK;
}

extern "C" int zj_set_tile(zj_handle h, int epb) {
    if (!h) return ZJ_ERR_ARG;
    const int M = max_epb(h->P);
    if (epb != M && epb != M / 2) return fail(h, ZJ_ERR_UNSUPPORTED, "tile must be the maximum or half of it");
    if (h->textbook && epb != M) return fail(h, ZJ_ERR_UNSUPPORTED, "the textbook scheme runs at the full tile");
    if (epb == h->epb) return ZJ_OK;
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "tile change in the middle of a step");
    if (h->in_batches > 0) return fail(h, ZJ_ERR_STATE, "the handle belongs to a batch, which runs the full tile");
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    // the partial-sum buffer holds one entry per element, enough for any tile
    h->epb = epb;
    h->nblk = num_blocks(std::max(h->E, 1), epb);
    drop_graph(h);
    // traces of edges inside a tile never reach T; a new tiling needs all of them
    if (h->has_state) {
        int rc = refresh_trace(h, h->stream);
        if (rc) return rc;
        ZJ_HIP(h, hipStreamSynchronize(h->stream));
    }
    return ZJ_OK;
}

extern "C" int zj_set_xcd_groups(zj_handle h, int groups) {
    if (!h || groups < 1 || groups > 64) return ZJ_ERR_ARG;
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "launch change in the middle of a step");
    if (groups == h->xcd_groups) return ZJ_OK;
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    h->xcd_groups = groups;
    ++h->gen;
    drop_graph(h);
    return ZJ_OK;
}

extern "C" int zj_set_geometry_mode(zj_handle h, int mode, int* active) {
    if (!h || mode < -1 || mode > 1) return ZJ_ERR_ARG;
    if (mode == 1 && !h->affine_ok) return fail(h, ZJ_ERR_UNSUPPORTED, "the mesh has non-affine elements");
    if (mode == 0 && h->viscous) return fail(h, ZJ_ERR_UNSUPPORTED, "the viscous terms run with the per-element metric");
    if (mode >= 0 && (mode == 1) != h->affine) {
        if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "geometry change in the middle of a step");
        if (h->in_batches > 0) return fail(h, ZJ_ERR_STATE, "the handle belongs to a batch, grouped by metric form");
        ZJ_HIP(h, hipSetDevice(h->device));
        ZJ_HIP(h, hipStreamSynchronize(h->stream));
        h->affine = mode == 1;
        drop_graph(h);
    }
    if (active) *active = h->affine ? 1 : 0;
    return ZJ_OK;
}

extern "C" int zj_get_tiling(zj_handle h, int64_t* cross_tile_edges, int* reordered) {
    if (!h) return ZJ_ERR_ARG;
    int64_t c = 0;
    for (size_t s = 0; s < h->links_host.size(); ++s) {
        const EdgeLink& L = h->links_host[s];
        if ((L.role == ROLE_LEFT || L.role == ROLE_RIGHT) && L.nbr / h->epb != (int)(s / 4) / h->epb) ++c;
    }
    if (cross_tile_edges) *cross_tile_edges = c;
    if (reordered) *reordered = h->perm ? 1 : 0;
    return ZJ_OK;
}

extern "C" int zj_get_scheme(zj_handle h, int* poly_order, double* gamma) {
    if (!h) return ZJ_ERR_ARG;
    if (poly_order) *poly_order = h->P;
    if (gamma) *gamma = h->gamma;
    return ZJ_OK;
}

extern "C" int zj_get_info(zj_handle h, int64_t* n_dof, int* epb, int* n_sp) {
    if (!h) return ZJ_ERR_ARG;
    if (n_dof) *n_dof = (int64_t)h->E * h->S * 4;
    if (epb) *epb = h->epb;
    if (n_sp) *n_sp = h->S;
    return ZJ_OK;
}

// ------------------------------------------------------------ measurement
namespace {
// Grid-stride copy of 32-byte elements, U loads in flight per thread before
// their stores.
template <int U>
__global__ __launch_bounds__(256) void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst,
                                                   int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < n; i += U * stride) {
        double4 v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) v[k] = src[i + k * stride];
#pragma unroll
        for (int k = 0; k < U; ++k) dst[i + k * stride] = v[k];
    }
    for (; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" int zj_copy_probe(int device, int64_t bytes, int reps, double* gbps) {
    if (bytes < 32 || reps < 1 || !gbps) return ZJ_ERR_ARG;
    zj_solver tmp;  // only for error text plumbing
    zj_solver* h = &tmp;
    const int64_t n = bytes / 32;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ZJ_OK;
    auto fin = [&](int code) {
        if (a) hipFree(a);
        if (b) hipFree(b);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (code) g_create_error = h->err;
        return code;
    };
    if (hipSetDevice(device) != hipSuccess || hipMalloc(&a, n * 32) != hipSuccess ||
        hipMalloc(&b, n * 32) != hipSuccess || hipMemset(a, 0, n * 32) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        h->err = "copy probe: allocation failed";
        return fin(ZJ_ERR_HIP);
    }
    // the best of a few launch shapes: 8192..32768 workgroups, 1 or 8 loads in flight
    double best = 0.0;
    for (int grid : {8192, 16384, 32768}) {
        for (int unroll : {1, 8}) {
            auto launch = [&] {
                if (unroll == 1) copy_kernel<1><<<grid, 256>>>((const double4*)a, (double4*)b, n);
                else copy_kernel<8><<<grid, 256>>>((const double4*)a, (double4*)b, n);
            };
            launch();  // warm-up
            hipEventRecord(e0, nullptr);
            for (int r = 0; r < reps; ++r) launch();
            hipEventRecord(e1, nullptr);
            if (hipEventSynchronize(e1) != hipSuccess) {
                h->err = "copy probe: kernel failed";
                return fin(ZJ_ERR_HIP);
            }
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            best = std::max(best, 2.0 * (double)n * 32.0 * reps / (ms * 1e-3) / 1e9);
        }
    }
    *gbps = best;
    return fin(rc);
}

extern "C" int zj_set_boundary(zj_handle h, const int* bc_type, const int* bc_face_map, const double* bc_data,
                               int n_bc_data) {
    if (!h || n_bc_data < 0) return ZJ_ERR_ARG;
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "boundary change in the middle of a step");
    std::vector<EdgeLink> links = h->links_host;
    for (size_t b = 0; b < h->bface.size(); ++b) {
        int rc = boundary_link(h, h->bface[b], bc_type, bc_face_map, n_bc_data, bc_data != nullptr,
                               links[h->bface_slot[b]]);
        if (rc) return rc;
    }
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    if (n_bc_data > h->n_bc_alloc) {
        drop_graph(h);  // the data buffer moves
        ZJ_HIP(h, hipFree(h->bc_data));
        h->bc_data = nullptr;
        ZJ_HIP(h, hipMalloc((void**)&h->bc_data, (size_t)n_bc_data * 8 * sizeof(double)));
        h->n_bc_alloc = n_bc_data;
    }
    if (n_bc_data > 0 && bc_data)
        ZJ_HIP(h, hipMemcpy(h->bc_data, bc_data, (size_t)n_bc_data * 8 * sizeof(double), hipMemcpyHostToDevice));
    ZJ_HIP(h, hipMemcpy(h->links, links.data(), links.size() * sizeof(EdgeLink), hipMemcpyHostToDevice));
    h->links_host = std::move(links);
    ++h->gen;  // batches holding the handle re-read its bc_data pointer
    return ZJ_OK;
}

extern "C" int zj_rk3_combine(int stage, double* out, const double* U0, const double* Uk, const double* R,
                              int64_t n, double dt, void* stream) {
    if (stage < 1 || stage > 3 || n < 0 || (n > 0 && (!out || !U0 || !R || (stage > 1 && !Uk))))
        return ZJ_ERR_ARG;
    if (launch_rk3_combine(stage, out, U0, Uk, R, n, dt, (hipStream_t)stream) != hipSuccess) {
        g_create_error = "zj_rk3_combine: launch failed";
        return ZJ_ERR_HIP;
    }
    return ZJ_OK;
}

extern "C" int zj_euler_update(double* U, const double* R, int64_t n, double dt, void* stream) {
    if (n < 0 || (n > 0 && (!U || !R))) return ZJ_ERR_ARG;
    if (launch_euler_update(U, R, n, dt, (hipStream_t)stream) != hipSuccess) {
        g_create_error = "zj_euler_update: launch failed";
        return ZJ_ERR_HIP;
    }
    return ZJ_OK;
}

extern "C" int zj_synchronize(zj_handle h) {
    if (!h) return ZJ_ERR_ARG;
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

// ---------------------------------------------------------------- adjoint
// Discrete adjoint of the SSP-RK3 step (no reference counterpart): the
// cotangent of the handle's state is kept on the device as a partial value
// plus the cross-tile edge cotangents of the last adjoint launch, which the
// next launch (or zj_adjoint_get) folds in.
static int ensure_adjoint(zj_solver* h) {
    if (h->adj_ready) return ZJ_OK;
    const size_t npts = (size_t)h->E * h->S;
    const size_t rsz = h->prec == ZJ_PREC_FP64 ? sizeof(double) : sizeof(float);
    // edge cotangents: U (4 words per face point), with the viscous terms also grad U (12)
    const size_t ntr = std::max<size_t>((size_t)h->E * 4 * h->N1 * (h->viscous ? 12 : 4), 1);
    h->adj_epb = h->viscous ? visc_adjoint_tile(h->P, h->prec == ZJ_PREC_FP64) : adjoint_tile(h->P);
    h->adj_nt = num_blocks(std::max(h->E, 1), h->adj_epb);
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, hipMalloc((void**)&h->A_W1, std::max<size_t>(npts * 4, 1) * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->A_W2, std::max<size_t>(npts * 4, 1) * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->A_Ub, std::max<size_t>(npts * 4, 1) * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->A_U0b, std::max<size_t>(npts * 4, 1) * sizeof(double)));
    for (int k = 0; k < 2; ++k) ZJ_HIP(h, hipMalloc(&h->A_Tb[k], ntr * rsz));
    ZJ_HIP(h, hipMalloc((void**)&h->A_dtb, (size_t)3 * h->adj_nt * sizeof(double)));
    if (h->viscous) ZJ_HIP(h, hipMalloc((void**)&h->A_mub, (size_t)h->adj_nt * sizeof(double)));
    ZJ_HIP(h, hipMalloc((void**)&h->A_idx, sizeof(unsigned long long)));
    h->adj_ready = true;
    return ZJ_OK;
}

template <typename R>
static VjpArgs<R> vjp_args(zj_solver* h) {
    VjpArgs<R> a{};
    a.E = h->E;
    a.Jinv = (const R*)(h->affine ? h->Jinv_e : h->Jinv);
    a.links = h->links;
    a.normals = (const R*)h->normals;
    a.bc_data = h->bc_data;
    a.dt_ptr = h->d_scalars;
    a.gamma = h->gamma;
    a.positivity = h->positivity;
    a.xcd_groups = h->xcd_groups;
    a.edge_f = (const R*)(h->affine ? h->edge_f_aff : h->edge_f_pt);
    a.Jpt = (const R*)h->Jpt;
    a.riemann = h->riemann;
    return a;
}

// One step's forward data at the adjoint tile: U^n, the stage states U1, U2
// (FP64, or FP32 increments over U^n) and the traces of U^n, U1, U2 on the
// edges another adjoint tile reads.
struct StepBufs {
    double* U;
    double* W1;
    double* W2;
    void* T[3];
};

static StepBufs handle_step(zj_solver* h) { return StepBufs{h->U, h->A_W1, h->A_W2, {h->T[0], h->T[1], h->T[2]}}; }

template <typename R>
static hipError_t launch_stage_adj(zj_solver* h, int mode, const StageArgs<R>& a, const OpTables<R>& ops,
                                   hipStream_t s) {
    return h->viscous ? launch_visc_stage_adj<R>(h->P, mode, a, ops, h->vp, s)
                      : launch_stage<R>(h->P, mode, a, ops, h->adj_epb, h->affine, s, h->textbook);
}

// The forward stages 1 and 2 at the adjoint tile, keeping U1 and U2 apart;
// with U_next, stage 3 too (into U_next and its traces T_next).
template <typename R>
static hipError_t recompute_stages(zj_solver* h, const OpTables<R>& ops, hipStream_t s, const StepBufs& b,
                                   double* U_next = nullptr, void* T_next = nullptr) {
    const bool inc = std::is_same<R, float>::value;
    StageArgs<R> a = stage_args<R>(h, MODE_STAGE1);
    a.U_in = b.U; a.T_in = (const R*)b.T[0]; a.T_out = (R*)b.T[1]; a.l2_partial = nullptr;
    if (inc) { a.dW_out = (float*)b.W1; a.U_out = nullptr; } else { a.U_out = b.W1; }
    hipError_t e = launch_stage_adj<R>(h, MODE_STAGE1, a, ops, s);
    if (e != hipSuccess) return e;
    StageArgs<R> c = stage_args<R>(h, MODE_STAGE2);
    c.U0 = b.U; c.T_in = (const R*)b.T[1]; c.T_out = (R*)b.T[2]; c.l2_partial = nullptr;
    if (inc) { c.dW_in = (const float*)b.W1; c.dW_out = (float*)b.W2; c.U_in = nullptr; c.U_out = nullptr; }
    else { c.U_in = b.W1; c.U_out = b.W2; }
    if ((e = launch_stage_adj<R>(h, MODE_STAGE2, c, ops, s)) != hipSuccess || !U_next) return e;
    StageArgs<R> d = stage_args<R>(h, MODE_STAGE3);
    d.U0 = b.U; d.U_out = U_next; d.T_in = (const R*)b.T[2]; d.T_out = (R*)T_next; d.l2_partial = nullptr;
    if (inc) { d.dW_in = (const float*)b.W2; d.U_in = nullptr; } else { d.U_in = b.W2; }
    return launch_stage_adj<R>(h, MODE_STAGE3, d, ops, s);
}

template <typename R>
static hipError_t adjoint_stages(zj_solver* h, const OpTables<R>& ops, hipStream_t s, const StepBufs& b,
                                 bool want_dt) {
    R** Tb = (R**)h->A_Tb;
    const bool inc = std::is_same<R, float>::value;
    for (int st = 3; st >= 1; --st) {
        VjpArgs<R> a = vjp_args<R>(h);
        double* W = st == 3 ? b.W2 : b.W1;
        if (st > 1) {
            a.U0 = b.U;
            if (inc) a.dW_in = (const float*)W; else a.U_in = W;
        } else {
            a.U_in = b.U;
        }
        a.T_in = (const R*)b.T[st - 1];
        a.Ub_in = h->A_Ub;
        a.Ub_out = h->A_Ub;
        a.Tb_in = h->A_has_tb ? Tb[h->A_cur] : nullptr;
        a.Tb_out = Tb[1 - h->A_cur];
        a.U0b = h->A_U0b;
        a.dtb_partial = want_dt ? h->A_dtb + (size_t)(3 - st) * h->adj_nt : nullptr;
        if (h->viscous) a.mub_partial = h->A_mub;
        hipError_t e = h->viscous ? launch_visc_vjp<R>(h->P, st, a, ops, h->vp, s)
                                  : launch_vjp<R>(h->P, st, a, ops, h->affine, s, h->textbook);
        if (e != hipSuccess) return e;
        h->A_cur = 1 - h->A_cur;
        h->A_has_tb = true;
    }
    return hipSuccess;
}

extern "C" int zj_adjoint_set_device(zj_handle h, const double* d_ubar_aos, void* stream) {
    if (!h || !d_ubar_aos) return ZJ_ERR_ARG;
    int rc = ensure_adjoint(h);
    if (rc) return rc;
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, launch_aos_to_soa(d_ubar_aos, h->A_Ub, (int64_t)h->E * h->S, pick(h, stream), h->perm, h->S));
    h->A_has_tb = false;
    h->A_seeded = true;
    if (h->viscous)
        ZJ_HIP(h, hipMemsetAsync(h->A_mub, 0, (size_t)h->adj_nt * sizeof(double), pick(h, stream)));
    return ZJ_OK;
}

extern "C" int zj_adjoint_seed_kinetic_energy(zj_handle h, const double* d_weights, double* J, void* stream) {
    if (!h || !d_weights) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "a partial step is in progress");
    int rc = ensure_adjoint(h);
    if (rc) return rc;
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    ZJ_HIP(h, launch_kinetic_energy(h->U, d_weights, h->perm, h->E, h->S, h->A_Ub, h->partial_dt, h->nred, s));
    ZJ_HIP(h, launch_finalize(h->partial_dt, h->nred, FIN_SUM, h->d_scalars + 5, s));
    h->A_has_tb = false;
    h->A_seeded = true;
    if (h->viscous) ZJ_HIP(h, hipMemsetAsync(h->A_mub, 0, (size_t)h->adj_nt * sizeof(double), s));
    if (J) {
        ZJ_HIP(h, hipMemcpyAsync(h->h_pinned + 5, h->d_scalars + 5, sizeof(double), hipMemcpyDeviceToHost, s));
        ZJ_HIP(h, hipStreamSynchronize(s));
        *J = h->h_pinned[5];
    }
    return ZJ_OK;
}

extern "C" int zj_adjoint_set(zj_handle h, const double* ubar_aos) {
    if (!h || !ubar_aos) return ZJ_ERR_ARG;
    ZJ_HIP(h, hipSetDevice(h->device));
    const size_t n = (size_t)h->E * h->S * 4;
    ZJ_HIP(h, hipMemcpyAsync(h->aos, ubar_aos, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    int rc = zj_adjoint_set_device(h, h->aos, nullptr);
    if (rc) return rc;
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

template <typename R>
static hipError_t complete_into(zj_solver* h, const double* Ub, const void* Tb, double* out, const OpTables<R>& ops,
                                hipStream_t s) {
    CompleteArgs<R> c{h->E, Ub, (const R*)Tb, out, h->links, h->xcd_groups};
    if (h->viscous) return launch_visc_complete<R>(h->P, c, (const R*)h->Jinv_e, ops, s);
    return launch_complete<R>(h->P, c, ops, s);
}

extern "C" int zj_adjoint_get_device(zj_handle h, double* d_ubar_aos, void* stream) {
    if (!h || !d_ubar_aos) return ZJ_ERR_ARG;
    if (!h->A_seeded) return fail(h, ZJ_ERR_STATE, "no adjoint seed set");
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    const void* tb = h->A_has_tb ? h->A_Tb[h->A_cur] : nullptr;
    hipError_t e = h->prec == ZJ_PREC_FP64 ? complete_into<double>(h, h->A_Ub, tb, h->Rres, h->ops64, s)
                                           : complete_into<float>(h, h->A_Ub, tb, h->Rres, h->ops32, s);
    ZJ_HIP(h, e);
    ZJ_HIP(h, launch_soa_to_aos(h->Rres, d_ubar_aos, (int64_t)h->E * h->S, s, h->perm, h->S));
    return ZJ_OK;
}

extern "C" int zj_adjoint_get(zj_handle h, double* ubar_aos) {
    if (!h || !ubar_aos) return ZJ_ERR_ARG;
    int rc = zj_adjoint_get_device(h, h->aos, nullptr);
    if (rc) return rc;
    const size_t n = (size_t)h->E * h->S * 4;
    ZJ_HIP(h, hipMemcpyAsync(ubar_aos, h->aos, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

extern "C" int zj_rk3_step_vjp(zj_handle h, double dt, double* dt_bar, void* stream) {
    if (!h) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "a partial step is in progress");
    if (!h->A_seeded) return fail(h, ZJ_ERR_STATE, "no adjoint seed set");
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    ZJ_HIP(h, set_dt(h, dt, s));
    int rc = refresh_trace(h, s);  // traces of U^n on every edge, valid for the adjoint tile
    if (rc) return rc;
    hipError_t e;
    const StepBufs b = handle_step(h);
    if (h->prec == ZJ_PREC_FP64) {
        e = recompute_stages<double>(h, h->ops64, s, b);
        if (e == hipSuccess) e = adjoint_stages<double>(h, h->ops64, s, b, dt_bar != nullptr);
    } else {
        e = recompute_stages<float>(h, h->ops32, s, b);
        if (e == hipSuccess) e = adjoint_stages<float>(h, h->ops32, s, b, dt_bar != nullptr);
    }
    ZJ_HIP(h, e);
    if (dt_bar) {
        ZJ_HIP(h, launch_finalize(h->A_dtb, 3 * h->adj_nt, FIN_SUM, h->d_scalars + 3, s));
        ZJ_HIP(h, hipMemcpyAsync(h->h_pinned + 3, h->d_scalars + 3, sizeof(double), hipMemcpyDeviceToHost, s));
        ZJ_HIP(h, hipStreamSynchronize(s));
        *dt_bar = h->h_pinned[3];
    }
    return ZJ_OK;
}

// ------------------------------------------------- native checkpointed sweep
static size_t tape_state_bytes(const zj_solver* h) {
    return ((size_t)h->E * h->S * 4 * sizeof(double) + 255) / 256 * 256;
}
static size_t tape_trace_bytes(const zj_solver* h) {
    const size_t w = (size_t)h->E * 4 * h->N1 * (h->viscous ? 12 : 4);
    return (std::max<size_t>(w, 1) * (h->prec == ZJ_PREC_FP64 ? sizeof(double) : sizeof(float)) + 255) / 256 * 256;
}
static StepBufs tape_step(zj_solver* h, int k) {
    const size_t sb = tape_state_bytes(h), tb = tape_trace_bytes(h);
    char* base = h->tape + (size_t)k * (3 * sb + 3 * tb);
    return StepBufs{(double*)base, (double*)(base + sb), (double*)(base + 2 * sb),
                    {base + 3 * sb, base + 3 * sb + tb, base + 3 * sb + 2 * tb}};
}

static int ensure_tape(zj_solver* h, int steps) {
    if (h->tape && h->tape_steps >= steps) return ZJ_OK;
    h->rec_n = 0;
    if (h->tape) {
        ZJ_HIP(h, hipStreamSynchronize(h->stream));
        ZJ_HIP(h, hipDeviceSynchronize());
        ZJ_HIP(h, hipFree(h->tape));
        h->tape = nullptr;
        h->tape_steps = 0;
    }
    const size_t bytes = (size_t)steps * (3 * tape_state_bytes(h) + 3 * tape_trace_bytes(h));
    if (hipMalloc((void**)&h->tape, bytes) != hipSuccess) {
        h->tape = nullptr;
        return fail(h, ZJ_ERR_NOMEM, "reverse-sweep tape: device allocation failed");
    }
    h->tape_steps = steps;
    return ZJ_OK;
}

template <typename R>
static hipError_t trace_into(zj_solver* h, const double* U, void* T, const OpTables<R>& ops, hipStream_t s) {
    if (h->viscous) return launch_visc_trace<R>(h->P, h->E, U, (const R*)h->Jinv_e, (R*)T, ops, h->xcd_groups, s);
    return launch_trace<R>(h->P, TraceArgs<R>{h->E, U, (R*)T, h->xcd_groups}, ops, 0, s);
}

// m forward steps from `src` recorded on the tape from entry `base` on: the
// stage states and traces written where the VJPs read them.  With U_end,
// the last step's stage 3 too (U^{n+m} into U_end; its traces into T_end).
template <typename R>
static hipError_t record_steps(zj_solver* h, const OpTables<R>& ops, hipStream_t s, const double* src,
                               const double* dts, int m, int base, double* U_end = nullptr, void* T_end = nullptr) {
    const size_t n = (size_t)h->E * h->S * 4;
    StepBufs b0 = tape_step(h, base);
    hipError_t e = hipMemcpyAsync(b0.U, src, n * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = trace_into<R>(h, b0.U, b0.T[0], ops, s);
    for (int k = 0; k < m && e == hipSuccess; ++k) {
        if ((e = set_dt(h, dts[k], s)) != hipSuccess) break;
        const StepBufs b = tape_step(h, base + k);
        if (k + 1 < m) {
            const StepBufs nb = tape_step(h, base + k + 1);
            e = recompute_stages<R>(h, ops, s, b, nb.U, nb.T[0]);
        } else {
            e = recompute_stages<R>(h, ops, s, b, U_end, T_end);
        }
    }
    return e;
}

// One segment backwards: unless the forward pass recorded it, its checkpoint
// and its forward steps recorded on the tape (no step is computed twice);
// then the stage VJPs from the last step to the first.
template <typename R>
static hipError_t sweep_segment(zj_solver* h, const OpTables<R>& ops, hipStream_t s, const double* dts, int start,
                                int m, int slot, bool want_dt, bool recorded, int base) {
    const size_t n = (size_t)h->E * h->S * 4;
    hipError_t e = hipSuccess;
    if (!recorded) e = record_steps<R>(h, ops, s, h->slots + (size_t)slot * n, dts + start, m, base);
    for (int k = m - 1; k >= 0 && e == hipSuccess; --k) {
        if ((e = set_dt(h, dts[start + k], s)) != hipSuccess) break;
        if ((e = adjoint_stages<R>(h, ops, s, tape_step(h, base + k), want_dt)) != hipSuccess) break;
        if (want_dt) e = launch_finalize(h->A_dtb, 3 * h->adj_nt, FIN_SUM, h->d_dtb_steps + start + k, s);
    }
    return e;
}

extern "C" int zj_forward_checkpointed(zj_handle h, const double* dts, int n_steps, int every, void* stream) {
    if (!h || !dts || n_steps < 1 || every < 1) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "a partial step is in progress");
    const int n_ckpt = (n_steps + every - 1) / every;
    if (h->n_slots < n_ckpt) {
        const int rc = zj_state_slots(h, n_ckpt);
        if (rc) return rc;
    }
    h->rec_n = 0;
    // the trailing segments run at the adjoint tile, recorded on the tape
    const int n_rec = std::min(h->tape_segs, n_ckpt), c0 = n_ckpt - n_rec;
    if (n_rec > 0) {
        int rc = ensure_adjoint(h);
        if (!rc) rc = ensure_tape(h, n_rec * every);
        if (rc) return rc;
    }
    hipStream_t s = pick(h, stream);
    for (int c = 0; c < n_ckpt; ++c) {
        const int start = c * every, m = std::min(start + every, n_steps) - start;
        int rc;
        if ((rc = zj_state_save(h, c, stream))) return rc;
        if (c < c0) {
            for (int k = 0; k < m; ++k)
                if ((rc = zj_rk3_step(h, dts[start + k], stream))) return rc;
            continue;
        }
        ZJ_HIP(h, hipSetDevice(h->device));
        const int base = (c - c0) * every;
        // U^{start+m} into the handle's state; the stage-3 traces at the
        // adjoint tile land in a stage buffer, the step traces are rebuilt
        const hipError_t e =
            h->prec == ZJ_PREC_FP64
                ? record_steps<double>(h, h->ops64, s, h->U, dts + start, m, base, h->U, h->T[1])
                : record_steps<float>(h, h->ops32, s, h->U, dts + start, m, base, h->U, h->T[1]);
        ZJ_HIP(h, e);
        if ((rc = refresh_trace(h, s))) return rc;
        h->l2_blocks = 0;  // these steps leave no residual norm
    }
    if (n_rec > 0) {
        h->rec_first = c0;
        h->rec_n = n_rec;
        h->rec_every = every;
        h->rec_steps = n_steps;
        h->rec_dts.assign(dts, dts + n_steps);
    }
    return ZJ_OK;
}

extern "C" int zj_adjoint_tape_segments(zj_handle h, int n) {
    if (!h || n < 0) return ZJ_ERR_ARG;
    h->tape_segs = n;
    return ZJ_OK;
}

extern "C" int zj_adjoint_sweep(zj_handle h, const double* dts, int n_steps, int every, double* dt_bar,
                                void* stream) {
    if (!h || !dts || n_steps < 1 || every < 1) return ZJ_ERR_ARG;
    if (!h->A_seeded) return fail(h, ZJ_ERR_STATE, "no adjoint seed set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "a partial step is in progress");
    const int n_ckpt = (n_steps + every - 1) / every;
    if (h->n_slots < n_ckpt) return fail(h, ZJ_ERR_STATE, "fewer state slots than checkpoints");
    int rc = ensure_adjoint(h);
    if (rc) return rc;
    ZJ_HIP(h, hipSetDevice(h->device));
    // segments the forward pass recorded for this very trajectory
    const bool have_rec = h->rec_n > 0 && h->rec_every == every && h->rec_steps == n_steps &&
                          std::equal(dts, dts + n_steps, h->rec_dts.begin());
    if (!have_rec && (rc = ensure_tape(h, std::min(every, n_steps)))) return rc;
    if (dt_bar && h->n_dtb_steps < n_steps) {
        if (h->d_dtb_steps) {
            ZJ_HIP(h, hipDeviceSynchronize());
            ZJ_HIP(h, hipFree(h->d_dtb_steps));
            h->d_dtb_steps = nullptr;
            h->n_dtb_steps = 0;
        }
        ZJ_HIP(h, hipMalloc((void**)&h->d_dtb_steps, (size_t)n_steps * sizeof(double)));
        h->n_dtb_steps = n_steps;
    }
    hipStream_t s = pick(h, stream);
    bool overwrote = false;
    for (int c = n_ckpt - 1; c >= 0; --c) {
        const int start = c * every, m = std::min(start + every, n_steps) - start;
        // a recorded segment sits at its own place on the tape; the others
        // are recomputed into the first place, free once the recorded
        // segments (all later than them) are done
        const bool rec = have_rec && c >= h->rec_first;
        const int base = rec ? (c - h->rec_first) * every : 0;
        overwrote |= !rec;
        const hipError_t e =
            h->prec == ZJ_PREC_FP64
                ? sweep_segment<double>(h, h->ops64, s, dts, start, m, c, dt_bar != nullptr, rec, base)
                : sweep_segment<float>(h, h->ops32, s, dts, start, m, c, dt_bar != nullptr, rec, base);
        ZJ_HIP(h, e);
    }
    if (overwrote) h->rec_n = 0;
    // the handle's state is U^0 again
    const size_t n = (size_t)h->E * h->S * 4;
    ZJ_HIP(h, hipMemcpyAsync(h->U, h->slots, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    if ((rc = refresh_trace(h, s))) return rc;
    h->has_state = true;
    if (dt_bar) {
        ZJ_HIP(h, hipMemcpyAsync(dt_bar, h->d_dtb_steps, (size_t)n_steps * sizeof(double), hipMemcpyDeviceToHost, s));
        ZJ_HIP(h, hipStreamSynchronize(s));
    }
    return ZJ_OK;
}

extern "C" int zj_adjoint_get_re(zj_handle h, double* dJ_dRe) {
    if (!h || !dJ_dRe) return ZJ_ERR_ARG;
    if (!h->A_seeded) return fail(h, ZJ_ERR_STATE, "no adjoint seed set");
    if (!h->viscous) {
        *dJ_dRe = 0.0;
        return ZJ_OK;
    }
    ZJ_HIP(h, hipSetDevice(h->device));
    ZJ_HIP(h, hipDeviceSynchronize());  // the sweep may have run on a caller's stream
    ZJ_HIP(h, launch_finalize(h->A_mub, h->adj_nt, FIN_SUM, h->d_scalars + 4, h->stream));
    ZJ_HIP(h, hipMemcpyAsync(h->h_pinned + 4, h->d_scalars + 4, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    // F_v scales with mu = 1/Re: dJ/dmu = S / mu, dJ/dRe = -dJ/dmu / Re^2 = -S / Re
    *dJ_dRe = -h->h_pinned[4] * h->vp.mu;
    return ZJ_OK;
}

extern "C" int zj_residual_vjp_device(zj_handle h, const double* d_w_aos, double* d_ubar_aos, void* stream) {
    if (!h || !d_w_aos || !d_ubar_aos) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "no state set");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "residual VJP requested in the middle of a step");
    int rc = ensure_adjoint(h);
    if (rc) return rc;
    ZJ_HIP(h, hipSetDevice(h->device));
    hipStream_t s = pick(h, stream);
    ZJ_HIP(h, launch_aos_to_soa(d_w_aos, h->Rres, (int64_t)h->E * h->S, s, h->perm, h->S));
    rc = refresh_trace(h, s);
    if (rc) return rc;
    // the adjoint accumulator's pending edge cotangents sit in A_Tb[A_cur]; use the other buffer
    void* tb = h->A_Tb[1 - h->A_cur];
    hipError_t e;
    if (h->prec == ZJ_PREC_FP64) {
        VjpArgs<double> a = vjp_args<double>(h);
        a.U_in = h->U; a.T_in = (const double*)h->T[0];
        a.Ub_in = h->Rres; a.Ub_out = h->Rres; a.Tb_out = (double*)tb;
        e = h->viscous ? launch_visc_vjp<double>(h->P, MODE_RESIDUAL, a, h->ops64, h->vp, s)
                       : launch_vjp<double>(h->P, MODE_RESIDUAL, a, h->ops64, h->affine, s, h->textbook);
        if (e == hipSuccess) e = complete_into<double>(h, h->Rres, tb, h->W, h->ops64, s);
    } else {
        VjpArgs<float> a = vjp_args<float>(h);
        a.U_in = h->U; a.T_in = (const float*)h->T[0];
        a.Ub_in = h->Rres; a.Ub_out = h->Rres; a.Tb_out = (float*)tb;
        e = h->viscous ? launch_visc_vjp<float>(h->P, MODE_RESIDUAL, a, h->ops32, h->vp, s)
                       : launch_vjp<float>(h->P, MODE_RESIDUAL, a, h->ops32, h->affine, s, h->textbook);
        if (e == hipSuccess) e = complete_into<float>(h, h->Rres, tb, h->W, h->ops32, s);
    }
    ZJ_HIP(h, e);
    ZJ_HIP(h, launch_soa_to_aos(h->W, d_ubar_aos, (int64_t)h->E * h->S, s, h->perm, h->S));
    return ZJ_OK;
}

extern "C" int zj_residual_vjp(zj_handle h, const double* w_aos, double* ubar_aos) {
    if (!h || !w_aos || !ubar_aos) return ZJ_ERR_ARG;
    ZJ_HIP(h, hipSetDevice(h->device));
    const size_t n = (size_t)h->E * h->S * 4;
    ZJ_HIP(h, hipMemcpyAsync(h->aos, w_aos, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    int rc = zj_residual_vjp_device(h, h->aos, h->aos, nullptr);  // w is consumed before aos is rewritten
    if (rc) return rc;
    ZJ_HIP(h, hipMemcpyAsync(ubar_aos, h->aos, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ZJ_HIP(h, hipStreamSynchronize(h->stream));
    return ZJ_OK;
}

// ---------------------------------------------------------------- batches
// Many meshes stepped together (no reference counterpart): per stage, one
// launch per (order, metric form) group over all member tiles, the groups as
// parallel branches of one captured graph.
struct zj_batch {
    int device = 0;
    int prec = ZJ_PREC_FP64;
    hipStream_t stream = nullptr;
    std::vector<zj_solver*> m;
    struct Group {
        int P = 0;
        bool affine = false;
        bool textbook = false;
        bool viscous = false;
        int n_items = 0;
        int2* items = nullptr;
        void* insts[4] = {nullptr, nullptr, nullptr, nullptr};  // BatchInst<R>[members of the group] per mode
        hipStream_t branch = nullptr;
        hipEvent_t join = nullptr;
    };
    std::vector<Group> groups;
    std::vector<std::vector<int>> group_members;  // member indices of each group
    // single launch per stage over every member (all orders and metric forms;
    // reference-arithmetic members only)
    bool single = false;
    bool any_textbook = false;
    bool all_viscous = false;   // every member has the Navier-Stokes terms: the one-launch BR2 kernel
    int visc_lds = 0;           // its LDS bytes (largest order present)
    void* all_insts[4] = {nullptr, nullptr, nullptr, nullptr};  // BatchInst<R>[members] per mode
    int2* all_items = nullptr;  // (member, tile), larger members first
    int n_all = 0;
    int all_lds = 0;            // LDS bytes of the largest order present
    std::vector<unsigned> gens;                    // member generations the uploaded arguments match
    double* d_dt = nullptr;
    std::vector<double> dt_on_device;
    hipEvent_t fork = nullptr;
    hipEvent_t done = nullptr;  // recorded after each batched step; members' calls wait for it
    int policy = 0;  // 0: one work item per workgroup; k > 0: k workgroups per CU, persistent
    int n_cu = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // the batched step backwards: per group, work items (position in the
    // group, tile at the members' adjoint tile); the stage VJPs' instances of
    // a call ([stage 3, 2, 1][member in group order]) in one device buffer,
    // filled from pinned staging, which the next call reuses once the copy ran
    std::vector<int2*> vitems;
    std::vector<int> vn_items;
    std::vector<int> voff;          // each group's first instance within a stage's block
    std::vector<int> vorder;        // members in group order
    void* d_vinst = nullptr;
    void* h_vinst = nullptr;
    hipEvent_t vcopied = nullptr;
    std::string err;
};

static thread_local std::string g_batch_error;

static void batch_release(zj_batch* b) {
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (auto& g : b->groups) {
        if (g.items) hipFree(g.items);
        for (void* p : g.insts)
            if (p) hipFree(p);
        if (g.join) hipEventDestroy(g.join);
        if (g.branch) hipStreamDestroy(g.branch);
    }
    for (void* p : b->all_insts)
        if (p) hipFree(p);
    if (b->all_items) hipFree(b->all_items);
    for (int2* p : b->vitems)
        if (p) hipFree(p);
    if (b->d_vinst) hipFree(b->d_vinst);
    if (b->h_vinst) hipHostFree(b->h_vinst);
    if (b->vcopied) hipEventDestroy(b->vcopied);
    if (b->d_dt) hipFree(b->d_dt);
    if (b->fork) hipEventDestroy(b->fork);
    if (b->done) hipEventDestroy(b->done);
    if (b->stream) hipStreamDestroy(b->stream);
}

#define ZJ_BHIP(b, call)                                                                 \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            (b)->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return ZJ_ERR_HIP;                                                           \
        }                                                                                \
    } while (0)

template <typename R>
static int batch_upload_insts(zj_batch* b, void** insts, const std::vector<int>& members) {
    for (int mode = MODE_STAGE1; mode <= MODE_STAGE3; ++mode) {
        std::vector<BatchInst<R>> v;
        for (int idx : members) {
            zj_solver* h = b->m[idx];
            BatchInst<R> in{};
            in.args = stage_args<R>(h, mode);
            in.args.dt_ptr = b->d_dt + idx;
            in.ops = *reinterpret_cast<const OpTables<R>*>(std::is_same<R, double>::value ? (const void*)&h->ops64
                                                                                             : (const void*)&h->ops32);
            in.P = h->P;
            in.affine = h->affine ? 1 : 0;
            in.textbook = h->textbook ? 1 : 0;
            in.viscous = h->viscous ? 1 : 0;
            in.vp = h->vp;
            v.push_back(in);
        }
        if (!insts[mode]) ZJ_BHIP(b, hipMalloc(&insts[mode], v.size() * sizeof(BatchInst<R>)));
        ZJ_BHIP(b, hipMemcpy(insts[mode], v.data(), v.size() * sizeof(BatchInst<R>), hipMemcpyHostToDevice));
    }
    return ZJ_OK;
}

static int batch_upload_all(zj_batch* b) {
    std::vector<int> all(b->m.size());
    for (size_t i = 0; i < all.size(); ++i) all[i] = (int)i;
    return b->prec == ZJ_PREC_FP64 ? batch_upload_insts<double>(b, b->all_insts, all)
                                   : batch_upload_insts<float>(b, b->all_insts, all);
}

static int batch_build(zj_batch* b, const zj_handle* hs, int n) {
    b->m.assign(hs, hs + n);
    b->device = hs[0]->device;
    b->prec = hs[0]->prec;
    for (int i = 0; i < n; ++i) {
        zj_solver* h = hs[i];
        if (!h) { b->err = "null member"; return ZJ_ERR_ARG; }
        if (h->device != b->device || h->prec != b->prec) {
            b->err = "members must share the device and the precision";
            return ZJ_ERR_ARG;
        }
        if (h->epb != max_epb(h->P)) {
            b->err = "members must use the full tile of their order";
            return ZJ_ERR_UNSUPPORTED;
        }
        b->any_textbook |= h->textbook;
        for (int j = 0; j < i; ++j)
            if (hs[j] == h) { b->err = "a handle appears twice"; return ZJ_ERR_ARG; }
    }
    ZJ_BHIP(b, hipSetDevice(b->device));
    ZJ_BHIP(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    ZJ_BHIP(b, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    ZJ_BHIP(b, hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    hipDeviceProp_t prop;
    ZJ_BHIP(b, hipGetDeviceProperties(&prop, b->device));
    b->n_cu = std::max(1, prop.multiProcessorCount);
    ZJ_BHIP(b, hipMalloc((void**)&b->d_dt, (size_t)n * sizeof(double)));
    b->dt_on_device.assign(n, NAN);
    // groups by (order, metric form, scheme): reference arithmetic, textbook
    // Euler, textbook Navier-Stokes
    auto scheme = [](const zj_solver* h) { return h->viscous ? 2 : h->textbook ? 1 : 0; };
    for (int P = 1; P <= 5; ++P)
        for (int af = 0; af < 2; ++af)
            for (int sch = 0; sch < 3; ++sch) {
                std::vector<int> members;
                for (int i = 0; i < n; ++i)
                    if (hs[i]->P == P && (hs[i]->affine ? 1 : 0) == af && scheme(hs[i]) == sch) members.push_back(i);
                if (members.empty()) continue;
                // larger meshes first; tiles of one mesh stay consecutive
                std::stable_sort(members.begin(), members.end(), [&](int x, int y) { return hs[x]->E > hs[y]->E; });
                zj_batch::Group g;
                g.P = P;
                g.affine = af != 0;
                g.textbook = sch >= 1;
                g.viscous = sch == 2;
                std::vector<int2> items;
                for (size_t k = 0; k < members.size(); ++k)
                    for (int t = 0; t < hs[members[k]]->nblk; ++t) items.push_back(make_int2((int)k, t));
                g.n_items = (int)items.size();
                b->groups.push_back(g);
                zj_batch::Group& gg = b->groups.back();
                ZJ_BHIP(b, hipMalloc((void**)&gg.items, std::max<size_t>(items.size(), 1) * sizeof(int2)));
                if (!items.empty())
                    ZJ_BHIP(b, hipMemcpy(gg.items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
                ZJ_BHIP(b, hipStreamCreateWithFlags(&gg.branch, hipStreamNonBlocking));
                ZJ_BHIP(b, hipEventCreateWithFlags(&gg.join, hipEventDisableTiming));
                b->group_members.push_back(members);
                int rc = b->prec == ZJ_PREC_FP64 ? batch_upload_insts<double>(b, gg.insts, members)
                                                 : batch_upload_insts<float>(b, gg.insts, members);
                if (rc) return rc;
            }
    b->all_viscous = true;
    for (int i = 0; i < n; ++i) {
        b->all_viscous &= hs[i]->viscous;
        b->visc_lds = std::max(b->visc_lds, visc_all_lds(hs[i]->P, b->prec == ZJ_PREC_FP32BULK));
    }
    // the single-launch form: one list of every member's tiles, larger members first
    {
        std::vector<int> order(n);
        for (int i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return hs[x]->E > hs[y]->E; });
        std::vector<int2> items;
        for (int i : order) {
            for (int t = 0; t < hs[i]->nblk; ++t) items.push_back(make_int2(i, t));
            b->all_lds = std::max(b->all_lds, batched_all_lds(hs[i]->P, b->prec == ZJ_PREC_FP32BULK));
        }
        b->n_all = (int)items.size();
        ZJ_BHIP(b, hipMalloc((void**)&b->all_items, std::max<size_t>(items.size(), 1) * sizeof(int2)));
        if (!items.empty())
            ZJ_BHIP(b, hipMemcpy(b->all_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
        int rc = batch_upload_all(b);
        if (rc) return rc;
    }
    for (int i = 0; i < n; ++i) b->gens.push_back(hs[i]->gen);
    return ZJ_OK;
}

// Re-uploads the per-member kernel arguments after a member changed them
// (zj_set_boundary can move its bc_data buffer); the graph is re-captured.
static int batch_refresh(zj_batch* b) {
    bool moved = false;
    for (size_t i = 0; i < b->m.size(); ++i) moved |= b->m[i]->gen != b->gens[i];
    if (!moved) return ZJ_OK;
    ZJ_BHIP(b, hipStreamSynchronize(b->stream));
    for (size_t k = 0; k < b->groups.size(); ++k) {
        int rc = b->prec == ZJ_PREC_FP64 ? batch_upload_insts<double>(b, b->groups[k].insts, b->group_members[k])
                                         : batch_upload_insts<float>(b, b->groups[k].insts, b->group_members[k]);
        if (rc) return rc;
    }
    if (int rc = batch_upload_all(b)) return rc;
    for (size_t i = 0; i < b->m.size(); ++i) b->gens[i] = b->m[i]->gen;
    if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    return ZJ_OK;
}

static int batch_grid(const zj_batch* b, const zj_batch::Group& g) {
    return b->policy <= 0 ? g.n_items : std::min(g.n_items, b->policy * b->n_cu);
}

// The three stages of a batched step on `s`: one launch per stage over every
// member (single), or one per (order, metric form) group on parallel branches.
static hipError_t batch_enqueue(zj_batch* b, hipStream_t s) {
    hipError_t e = hipSuccess;
    if (b->single) {
        const int grid = b->policy <= 0 ? b->n_all : std::min(b->n_all, b->policy * b->n_cu);
        for (int mode = MODE_STAGE1; mode <= MODE_STAGE3 && e == hipSuccess; ++mode) {
            if (b->all_viscous && b->prec == ZJ_PREC_FP64)
                e = launch_batched_visc_all<double>(mode, (const BatchInst<double>*)b->all_insts[mode], b->all_items,
                                                    b->n_all, grid, b->visc_lds, s);
            else if (b->all_viscous)
                e = launch_batched_visc_all<float>(mode, (const BatchInst<float>*)b->all_insts[mode], b->all_items,
                                                   b->n_all, grid, b->visc_lds, s);
            else if (b->prec == ZJ_PREC_FP64)
                e = launch_batched_all<double>(mode, (const BatchInst<double>*)b->all_insts[mode], b->all_items,
                                               b->n_all, grid, b->all_lds, s);
            else
                e = launch_batched_all<float>(mode, (const BatchInst<float>*)b->all_insts[mode], b->all_items,
                                              b->n_all, grid, b->all_lds, s);
        }
        return e;
    }
    for (int mode = MODE_STAGE1; mode <= MODE_STAGE3 && e == hipSuccess; ++mode) {
        e = hipEventRecord(b->fork, s);
        for (auto& g : b->groups) {
            if (e == hipSuccess) e = hipStreamWaitEvent(g.branch, b->fork, 0);
            if (e == hipSuccess) {
                if (g.viscous && b->prec == ZJ_PREC_FP64)
                    e = launch_batched_visc<double>(g.P, mode, (const BatchInst<double>*)g.insts[mode], g.items,
                                                    g.n_items, batch_grid(b, g), g.branch);
                else if (g.viscous)
                    e = launch_batched_visc<float>(g.P, mode, (const BatchInst<float>*)g.insts[mode], g.items,
                                                   g.n_items, batch_grid(b, g), g.branch);
                else if (b->prec == ZJ_PREC_FP64)
                    e = launch_batched_stage<double>(g.P, g.affine, mode, (const BatchInst<double>*)g.insts[mode],
                                                     g.items, g.n_items, batch_grid(b, g), g.branch, g.textbook);
                else
                    e = launch_batched_stage<float>(g.P, g.affine, mode, (const BatchInst<float>*)g.insts[mode],
                                                    g.items, g.n_items, batch_grid(b, g), g.branch, g.textbook);
            }
            if (e == hipSuccess) e = hipEventRecord(g.join, g.branch);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, g.join, 0);
        }
    }
    return e;
}

static int batch_capture(zj_batch* b) {
    if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    ZJ_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t e = batch_enqueue(b, b->stream);
    hipGraph_t gr = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->stream, &gr);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (gr) hipGraphDestroy(gr);
        ZJ_BHIP(b, e != hipSuccess ? e : e2);
    }
    b->graph = gr;
    ZJ_BHIP(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    return ZJ_OK;
}

extern "C" int zj_batch_create(zj_batch** out, const zj_handle* members, int n) {
    if (!out || !members || n < 1) {
        g_batch_error = "bad arguments";
        return ZJ_ERR_ARG;
    }
    *out = nullptr;
    zj_batch* b = new (std::nothrow) zj_batch();
    if (!b) {
        g_batch_error = "out of host memory";
        return ZJ_ERR_NOMEM;
    }
    int rc;
    try {
        rc = batch_build(b, members, n);
    } catch (const std::bad_alloc&) {
        b->err = "out of host memory";
        rc = ZJ_ERR_NOMEM;
    }
    if (rc) {
        g_batch_error = b->err;
        batch_release(b);
        delete b;
        return rc;
    }
    for (zj_solver* h : b->m) ++h->in_batches;
    *out = b;
    return ZJ_OK;
}

extern "C" int zj_batch_destroy(zj_batch* b) {
    if (!b) return ZJ_ERR_ARG;
    hipSetDevice(b->device);
    hipStreamSynchronize(b->stream);
    if (b->done) hipEventSynchronize(b->done);  // a step may have run on a caller's stream
    for (zj_solver* h : b->m) {
        --h->in_batches;
        if (h->after_batch == b->done) h->after_batch = nullptr;
    }
    batch_release(b);
    delete b;
    return ZJ_OK;
}

extern "C" const char* zj_batch_last_error(zj_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

extern "C" int zj_batch_set_grid(zj_batch* b, int policy) {
    if (!b || policy < 0 || policy > 64) return ZJ_ERR_ARG;
    if (policy != b->policy) {
        b->policy = policy;
        if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
        if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    }
    return ZJ_OK;
}

extern "C" int zj_batch_set_launch(zj_batch* b, int single) {
    if (!b || (single != 0 && single != 1)) return ZJ_ERR_ARG;
    if (single && b->any_textbook && !b->all_viscous) {
        b->err = "the single launch takes reference-arithmetic members or Navier-Stokes members, not a mix "
                 "or textbook Euler members";
        return ZJ_ERR_UNSUPPORTED;
    }
    if ((single != 0) != b->single) {
        b->single = single != 0;
        if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
        if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    }
    return ZJ_OK;
}

extern "C" int zj_batch_rk3_step(zj_batch* b, const double* dts, void* stream) {
    if (!b || !dts) return ZJ_ERR_ARG;
    for (zj_solver* h : b->m) {
        if (!h->has_state || h->stage_next != 1) {
            b->err = "every member needs a state and no partial step";
            return ZJ_ERR_STATE;
        }
    }
    ZJ_BHIP(b, hipSetDevice(b->device));
    hipStream_t s = stream ? (hipStream_t)stream : b->stream;
    int rc0 = batch_refresh(b);
    if (rc0) return rc0;
    if (!b->gexec) {
        int rc = batch_capture(b);
        if (rc) return rc;
    }
    const int n = (int)b->m.size();
    bool same = true;
    for (int i = 0; i < n && same; ++i) same = dts[i] == b->dt_on_device[i];
    if (!same) {
        ZJ_BHIP(b, launch_set_many(b->d_dt, dts, n, s));
        b->dt_on_device.assign(dts, dts + n);
    }
    ZJ_BHIP(b, hipGraphLaunch(b->gexec, s));
    ZJ_BHIP(b, hipEventRecord(b->done, s));
    for (zj_solver* h : b->m) {
        h->l2_blocks = h->nblk;
        h->after_batch = b->done;
    }
    return ZJ_OK;
}

// Work items of the batched step backwards, at each member's adjoint tile,
// and the instance buffers (once; the adjoint tile depends only on the
// member's order and precision).
template <typename R>
static int batch_vjp_prepare(zj_batch* b) {
    if (b->d_vinst) return ZJ_OK;
    const int n = (int)b->m.size();
    for (size_t g = 0; g < b->groups.size(); ++g) {
        std::vector<int2> items;
        b->voff.push_back((int)b->vorder.size());
        const std::vector<int>& mem = b->group_members[g];
        for (int k = 0; k < (int)mem.size(); ++k) {
            b->vorder.push_back(mem[k]);
            for (int t = 0; t < b->m[mem[k]]->adj_nt; ++t) items.push_back(make_int2(k, t));
        }
        int2* d = nullptr;
        if (!items.empty()) {
            ZJ_BHIP(b, hipMalloc((void**)&d, items.size() * sizeof(int2)));
            ZJ_BHIP(b, hipMemcpy(d, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
        }
        b->vitems.push_back(d);
        b->vn_items.push_back((int)items.size());
    }
    if ((int)b->vorder.size() != n) {
        b->err = "batch groups do not cover the members";
        return ZJ_ERR_STATE;
    }
    const size_t bytes = (size_t)3 * n * sizeof(BatchVjpInst<R>);
    ZJ_BHIP(b, hipMalloc(&b->d_vinst, bytes));
    ZJ_BHIP(b, hipHostMalloc(&b->h_vinst, bytes, hipHostMallocDefault));
    ZJ_BHIP(b, hipEventCreateWithFlags(&b->vcopied, hipEventDisableTiming));
    return ZJ_OK;
}

// One batched step backwards: each member's forward stages 1 and 2 recomputed
// at its adjoint tile (its own launches), then per stage 3, 2, 1 one VJP launch
// per group over every member's tiles, each member's arguments as its own
// zj_rk3_step_vjp would pass them.
template <typename R>
static int batch_step_vjp(zj_batch* b, const double* dts, double* dt_bars, hipStream_t s) {
    int rc = batch_vjp_prepare<R>(b);
    if (rc) return rc;
    const int n = (int)b->m.size();
    const bool want_dt = dt_bars != nullptr;
    for (int i = 0; i < n; ++i) {
        zj_solver* h = b->m[i];
        const OpTables<R>& ops = *(const OpTables<R>*)(std::is_same<R, double>::value ? (const void*)&h->ops64
                                                                                       : (const void*)&h->ops32);
        ZJ_BHIP(b, set_dt(h, dts[i], s));
        if ((rc = refresh_trace(h, s))) {
            b->err = h->err;
            return rc;
        }
        ZJ_BHIP(b, recompute_stages<R>(h, ops, s, handle_step(h)));
    }
    // the instances of the three stages; each member's pending edge
    // cotangents alternate between its two buffers as its own calls do
    ZJ_BHIP(b, hipEventSynchronize(b->vcopied));
    BatchVjpInst<R>* I = (BatchVjpInst<R>*)b->h_vinst;
    const bool inc = std::is_same<R, float>::value;
    for (int pos = 0; pos < n; ++pos) {
        zj_solver* h = b->m[b->vorder[pos]];
        R** Tb = (R**)h->A_Tb;
        const StepBufs sb = handle_step(h);
        for (int st = 3; st >= 1; --st) {
            BatchVjpInst<R>& x = I[(size_t)(3 - st) * n + pos];
            VjpArgs<R> a = vjp_args<R>(h);
            double* W = st == 3 ? sb.W2 : sb.W1;
            if (st > 1) {
                a.U0 = sb.U;
                if (inc) a.dW_in = (const float*)W; else a.U_in = W;
            } else {
                a.U_in = sb.U;
            }
            a.T_in = (const R*)sb.T[st - 1];
            a.Ub_in = h->A_Ub;
            a.Ub_out = h->A_Ub;
            a.Tb_in = h->A_has_tb ? Tb[h->A_cur] : nullptr;
            a.Tb_out = Tb[1 - h->A_cur];
            a.U0b = h->A_U0b;
            a.dtb_partial = want_dt ? h->A_dtb + (size_t)(3 - st) * h->adj_nt : nullptr;
            if (h->viscous) a.mub_partial = h->A_mub;
            x.a = a;
            x.ops = *(const OpTables<R>*)(std::is_same<R, double>::value ? (const void*)&h->ops64
                                                                         : (const void*)&h->ops32);
            x.vp = h->vp;
            h->A_cur = 1 - h->A_cur;
            h->A_has_tb = true;
        }
    }
    ZJ_BHIP(b, hipMemcpyAsync(b->d_vinst, b->h_vinst, (size_t)3 * n * sizeof(BatchVjpInst<R>), hipMemcpyHostToDevice,
                              s));
    ZJ_BHIP(b, hipEventRecord(b->vcopied, s));
    const BatchVjpInst<R>* D = (const BatchVjpInst<R>*)b->d_vinst;
    for (int st = 3; st >= 1; --st) {
        for (size_t g = 0; g < b->groups.size(); ++g) {
            const zj_batch::Group& G = b->groups[g];
            const BatchVjpInst<R>* insts = D + (size_t)(3 - st) * n + b->voff[g];
            if (G.viscous)
                ZJ_BHIP(b, launch_batched_visc_vjp<R>(G.P, st, insts, b->vitems[g], b->vn_items[g], s));
            else
                ZJ_BHIP(b, launch_batched_vjp<R>(G.P, G.affine, G.textbook, st, insts, b->vitems[g], b->vn_items[g], s));
        }
    }
    if (want_dt) {
        for (int i = 0; i < n; ++i) {
            zj_solver* h = b->m[i];
            ZJ_BHIP(b, launch_finalize(h->A_dtb, 3 * h->adj_nt, FIN_SUM, h->d_scalars + 3, s));
            ZJ_BHIP(b, hipMemcpyAsync(h->h_pinned + 3, h->d_scalars + 3, sizeof(double), hipMemcpyDeviceToHost, s));
        }
        ZJ_BHIP(b, hipStreamSynchronize(s));
        for (int i = 0; i < n; ++i) dt_bars[i] = b->m[i]->h_pinned[3];
    }
    return ZJ_OK;
}

extern "C" int zj_batch_rk3_step_vjp(zj_batch* b, const double* dts, double* dt_bars, void* stream) {
    if (!b || !dts) return ZJ_ERR_ARG;
    for (zj_solver* h : b->m) {
        if (!h->has_state || h->stage_next != 1 || !h->A_seeded) {
            b->err = "every member needs a state, no partial step and an adjoint seed";
            return ZJ_ERR_STATE;
        }
    }
    ZJ_BHIP(b, hipSetDevice(b->device));
    hipStream_t s = stream ? (hipStream_t)stream : b->stream;
    const int rc = b->prec == ZJ_PREC_FP64 ? batch_step_vjp<double>(b, dts, dt_bars, s)
                                           : batch_step_vjp<float>(b, dts, dt_bars, s);
    if (rc) return rc;
    ZJ_BHIP(b, hipEventRecord(b->done, s));
    for (zj_solver* h : b->m) h->after_batch = b->done;
    return ZJ_OK;
}

// Times `reps` batched steps (after one untimed step) for each launch form
// (grouped branches, single launch) and grid policy (one workgroup per tile;
// 2 or 4 per CU, persistent), keeps the fastest, and puts every member's state
// back as it was: the members' states are saved on the device first and
// restored, with their traces, at the end.
extern "C" int zj_batch_autotune(zj_batch* b, const double* dts, int reps, int* single, int* policy, double* ms) {
    if (!b || !dts || reps < 1) return ZJ_ERR_ARG;
    for (zj_solver* h : b->m)
        if (!h->has_state || h->stage_next != 1) {
            b->err = "every member needs a state and no partial step";
            return ZJ_ERR_STATE;
        }
    ZJ_BHIP(b, hipSetDevice(b->device));
    if (int rc = batch_refresh(b)) return rc;
    size_t total = 0;
    for (zj_solver* h : b->m) total += (size_t)4 * h->E * h->S;
    double* save = nullptr;
    ZJ_BHIP(b, hipMalloc((void**)&save, std::max<size_t>(total, 1) * sizeof(double)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ZJ_OK;
    auto run = [&]() -> int {
        size_t off = 0;
        for (zj_solver* h : b->m) {
            const size_t n = (size_t)4 * h->E * h->S;
            ZJ_BHIP(b, hipMemcpyAsync(save + off, h->U, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
            off += n;
        }
        ZJ_BHIP(b, hipEventCreate(&e0));
        ZJ_BHIP(b, hipEventCreate(&e1));
        const int cand_single[2] = {0, 1};
        const int cand_policy[3] = {0, 2, 4};
        double best = 1e300;
        int bs = b->single ? 1 : 0, bp = b->policy;
        for (int cs : cand_single)
            for (int cp : cand_policy) {
                if (cs && b->any_textbook && !b->all_viscous) continue;
                if (int r = zj_batch_set_launch(b, cs)) return r;
                if (int r = zj_batch_set_grid(b, cp)) return r;
                if (int r = zj_batch_rk3_step(b, dts, b->stream)) return r;  // captures and warms up
                ZJ_BHIP(b, hipEventRecord(e0, b->stream));
                for (int k = 0; k < reps; ++k)
                    if (int r = zj_batch_rk3_step(b, dts, b->stream)) return r;
                ZJ_BHIP(b, hipEventRecord(e1, b->stream));
                ZJ_BHIP(b, hipEventSynchronize(e1));
                float t = 0.f;
                ZJ_BHIP(b, hipEventElapsedTime(&t, e0, e1));
                const double per = t / reps;
                if (per < best) { best = per; bs = cs; bp = cp; }
            }
        if (int r = zj_batch_set_launch(b, bs)) return r;
        if (int r = zj_batch_set_grid(b, bp)) return r;
        if (single) *single = bs;
        if (policy) *policy = bp;
        if (ms) *ms = best;
        return ZJ_OK;
    };
    rc = run();
    // restore every member's state and the traces its next step reads
    size_t off = 0;
    for (zj_solver* h : b->m) {
        const size_t n = (size_t)4 * h->E * h->S;
        if (hipMemcpyAsync(h->U, save + off, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
            refresh_trace(h, b->stream) != ZJ_OK) {
            if (!rc) { b->err = "restoring the member states failed"; rc = ZJ_ERR_HIP; }
        }
        h->l2_blocks = 0;
        off += n;
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess && !rc) {
        b->err = "restoring the member states failed";
        rc = ZJ_ERR_HIP;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(save);
    return rc;
}

extern "C" int zj_batch_synchronize(zj_batch* b) {
    if (!b) return ZJ_ERR_ARG;
    ZJ_BHIP(b, hipSetDevice(b->device));
    ZJ_BHIP(b, hipStreamSynchronize(b->stream));
    return ZJ_OK;
}

// ------------------------------------------------------------- autotuning
// Times the residual launch (which leaves the state alone) for each tile and
// XCD order and keeps the fastest.
extern "C" int zj_autotune(zj_handle h, int reps, int* epb_out, int* xcd_out, double* ms_out) {
    if (!h || reps < 1) return ZJ_ERR_ARG;
    if (!h->has_state) return fail(h, ZJ_ERR_STATE, "autotuning needs a state");
    if (h->in_batches > 0) return fail(h, ZJ_ERR_STATE, "the handle belongs to a batch, which fixes its launch");
    if (h->stage_next != 1) return fail(h, ZJ_ERR_STATE, "autotune in the middle of a step");
    ZJ_HIP(h, hipSetDevice(h->device));
    int nxcc = 1;
    if (hipDeviceGetAttribute(&nxcc, hipDeviceAttributeNumberOfXccs, h->device) != hipSuccess || nxcc < 1) nxcc = 1;
    const int M = max_epb(h->P);
    const int epbs[2] = {M, (M / 2 >= 1 && !h->textbook) ? M / 2 : M};
    const int xcds[2] = {nxcc, 1};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ZJ_HIP(h, hipEventCreate(&e0));
    ZJ_HIP(h, hipEventCreate(&e1));
    double best = 1e300;
    int best_epb = h->epb, best_x = h->xcd_groups;
    int rc = ZJ_OK;
    for (int a = 0; a < 2 && rc == ZJ_OK; ++a)
        for (int x = 0; x < 2 && rc == ZJ_OK; ++x) {
            if ((rc = zj_set_tile(h, epbs[a])) || (rc = zj_set_xcd_groups(h, xcds[x]))) break;
            hipError_t e = enqueue_stage(h, MODE_RESIDUAL, h->stream);  // warm-up
            if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
            for (int r = 0; r < reps && e == hipSuccess; ++r) e = enqueue_stage(h, MODE_RESIDUAL, h->stream);
            if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                h->err = std::string("autotune: ") + hipGetErrorString(e);
                rc = ZJ_ERR_HIP;
                break;
            }
            if (ms / reps < best) { best = ms / reps; best_epb = epbs[a]; best_x = xcds[x]; }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    if ((rc = zj_set_tile(h, best_epb)) || (rc = zj_set_xcd_groups(h, best_x))) return rc;
    if (epb_out) *epb_out = best_epb;
    if (xcd_out) *xcd_out = best_x;
    if (ms_out) *ms_out = best;
    return ZJ_OK;
}
