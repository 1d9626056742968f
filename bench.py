#!/usr/bin/env python3
"""Residual DOF-updates/s of the FR residual + SSP-RK3 step on one MI355X.

Workload (the 2D Taylor-Green vortex benchmark): periodic [0,2pi]^2,
256x256 uniform quads, P4 (25 solution points/element, 1,638,400 points,
6,553,600 DOF), Mach 0.1, the reference's executed arithmetic (DG correction,
inline Rusanov, inviscid), SSP-RK3 with the reference's positivity limiter, dt
from the CFL rule at CFL 0.5.  Synthetic initial state (closed-form TGV).

One "step" = one SSP-RK3 step = three fused stage kernels, i.e. three residual
DOF-updates per DOF.  Timed with HIP events on the stream the kernels run on,
inputs resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "residual DOF-updates/s, FP32-MFMA vs FP64 plain at stated L2 err; adjoint/fwd cost"
HBM_PEAK_GBS = 8000.0


def alg_bytes_per_stage(E, p, stage, precision, affine, cross_edges=None, faces_per_elem=2):
    """Algorithmic HBM bytes of one SSP-RK3 stage launch (stage 1, 2, 3).

    Per solution point: the FP64 state U^n (32 B) is read by every stage; stages
    2 and 3 also read the intermediate state (FP64: 32 B; FP32-bulk: an FP32
    increment over U^n, 16 B); stages 1 and 2 write the intermediate state
    (32 B or 16 B), stage 3 writes U^{n+1} (32 B). Geometry is 4 Jinv words per
    point, or per element on affine meshes (J is never read). The Fp x 4 trace
    of an element edge whose neighbour lies in another tile is written once and
    read once in the compute precision (edges inside a tile stay in LDS;
    `cross_edges` = None counts all 4E); per face a normal (2 words) plus 16 B
    of connectivity."""
    S = (p + 1) ** 2
    Fp = p + 1
    w = 8 if precision == "fp64" else 4
    inter = 32 if precision == "fp64" else 16
    per_pt = 32 + (inter if stage > 1 else 0) + (inter if stage < 3 else 32)
    geo = 4 * w * (1 if affine else S)
    ce = 4 * E if cross_edges is None else cross_edges
    return E * (S * per_pt + geo + faces_per_elem * (2 * w + 16)) + ce * Fp * 4 * w * 2


def textbook_alg_bytes_per_stage(E, p, stage, precision, affine, cross_edges, viscous=False):
    """The textbook scheme's stage launch: the reference-arithmetic bytes plus
    the per-edge factor (edge half length over J, one word per element edge).
    With the Navier-Stokes terms each cross-tile trace carries U, dU/dx and
    dU/dy (12 words per flux point instead of 4)."""
    w = 8 if precision == "fp64" else 4
    by = alg_bytes_per_stage(E, p, stage, precision, affine, cross_edges) + E * 4 * w
    if viscous:
        by += cross_edges * (p + 1) * 8 * w * 2
    return by


def roofline_of(bytes_per_step, stage_ms, kernel, **extra):
    """HBM roofline of a stage kernel: algorithmic bytes of the three launches
    over their summed kernel execution time."""
    achieved = bytes_per_step / (sum(stage_ms) * 1e-3) / 1e9
    return dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                kernel=kernel, bytes_per_launch=bytes_per_step / 3.0, avg_launch_ms=sum(stage_ms) / 3.0,
                timing="kernel execution (start/stop events recorded by the runtime at each stage kernel)",
                **extra)


def cross_tile_edges(faces, epb):
    """Element edges whose trace crosses tiles of `epb` consecutive elements."""
    f = np.asarray(faces)
    inner = f[:, 2] >= 0
    return int(2 * np.count_nonzero(f[inner, 0] // epb != f[inner, 2] // epb))


def time_adjoint(n, p, precision, steps, warmup, stream, hiprt):
    """Forward SSP-RK3 step vs its discrete adjoint (stage recompute + three
    VJP launches) on the TGV n x n, order p, FP64 state; ms per step each."""
    from zhijian_amd import adjoint, cases, solver
    mesh = cases.tgv_mesh(n)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
    sol = solver.Solver(t, precision=precision, cfl=0.5)
    U0 = cases.tgv_ic(t.pos, mach=0.1).reshape(-1)
    sol.set_state(U0)
    dt = sol.dt_cfl(0.5)
    ms_fwd, _ = time_precision(sol, U0, dt, steps, warmup, stream, hiprt)
    sol.set_state(U0)
    sol.adjoint_set(adjoint.kinetic_energy_grad(U0, t))
    for _ in range(warmup):
        sol.rk3_step_vjp(dt, stream.ptr, want_dt=False)
    stream.synchronize()
    e0, e1 = hiprt.Event(), hiprt.Event()
    e0.record(stream)
    for _ in range(steps):
        sol.rk3_step_vjp(dt, stream.ptr, want_dt=False)
    e1.record(stream)
    e1.synchronize()
    ms_adj = e0.elapsed_ms(e1) / steps
    return dict(workload=f"TGV {n}x{n} P{p}", precision=precision, fwd_ms_per_step=ms_fwd,
                adj_ms_per_step=ms_adj, adjoint_over_forward=ms_adj / ms_fwd,
                note="adjoint step = traces of U^n + stages 1-2 recomputed + 3 VJP launches")


def time_gradient(n, p, precision, steps, every, stream, hiprt, Re=None, tape_segments=1):
    """The whole checkpointed gradient of the adjoint configuration (TGV n x n,
    order p, Mach 0.1, textbook VCJH HU / Roe, with the Navier-Stokes terms at
    Re when given, kinetic energy after `steps` SSP-RK3 steps, checkpoints
    every `every` steps on the GPU, the last `tape_segments` segments recorded
    on the device tape by the forward pass) against the same `steps` forward
    steps alone (graph-replayed), wall time each."""
    from zhijian_amd import adjoint, cases, solver
    mesh = cases.tgv_mesh(n)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
    visc = dict(viscous=True, Re=Re, Pr=0.72) if Re else {}
    s = solver.Solver(t, precision=precision, compat="textbook", riemann="roe", corr="hu", positivity=False, **visc)
    U0 = cases.tgv_ic(t.pos, mach=0.1).reshape(-1)
    s.set_state(U0)
    dt = s.dt_cfl(0.5)
    dts = [dt] * steps
    for _ in range(5):
        s.rk3_step(dt, stream.ptr)
    stream.synchronize()
    s.set_state(U0)
    e0, e1 = hiprt.Event(), hiprt.Event()
    e0.record(stream)
    for k in range(steps):
        s.rk3_step(dt, stream.ptr)
    e1.record(stream)
    e1.synchronize()
    fwd_ms = e0.elapsed_ms(e1)
    # warm-up (graph capture, allocation of the slots and of the whole tape)
    adjoint.gradient(s, t, U0, dts, every=every, want_dt=False, tape_segments=tape_segments)
    hiprt.hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    adjoint.gradient(s, t, U0, dts, every=every, want_dt=False, tape_segments=tape_segments)
    hiprt.hip().hipDeviceSynchronize()
    grad_ms = (time.perf_counter() - t0) * 1e3
    physics = f"Navier-Stokes (BR2) Re {Re:g}" if Re else "Euler"
    n_seg = (steps + every - 1) // every
    rec = min(tape_segments, n_seg)
    return dict(workload=f"TGV {n}x{n} P{p}, Mach 0.1, textbook VCJH HU / Roe, {physics}, {steps} steps, "
                         f"checkpoints every {every} (device-resident), kinetic-energy objective",
                precision=precision, tape_segments=rec, forward_ms=fwd_ms, gradient_ms=grad_ms,
                gradient_over_forward=grad_ms / fwd_ms,
                note=f"gradient = forward with checkpoints (its last {rec} of {n_seg} segments recorded on the "
                     f"device tape as it steps) + per other segment a forward recompute recorded on the tape + "
                     f"stage VJPs from the tape (both passes native); host wall time around the whole call, "
                     f"including the host objective and two state transfers")


BATCH_SCHEME = dict(compat="textbook", riemann="roe", corr="hu", positivity=False, viscous=True, Pr=0.72)


def batch_members(n_inst, seed, precision, scheme="navier_stokes"):
    """Seeded batch: N ~ U{16..128} elements per side, P ~ U{1..5},
    Mach ~ U[0.1, 0.8] (PCG64 seed), Re ~ U[100, 5000] (PCG64 [seed, 1]),
    alternating Taylor-Green /
    isentropic vortex; textbook FR (VCJH HU, Roe) with the Navier-Stokes terms
    (BR2, Pr 0.72), no limiter; scheme="reference": the reference arithmetic
    (DG correction, inline Rusanov, positivity limiter) on the same meshes."""
    from zhijian_amd import cases, solver
    rng = np.random.default_rng(seed)
    re_rng = np.random.default_rng([seed, 1])  # its own stream: the meshes do not depend on it
    out = []
    for k in range(n_inst):
        n, p, mach = int(rng.integers(16, 129)), int(rng.integers(1, 6)), float(rng.uniform(0.1, 0.8))
        Re = float(re_rng.uniform(100.0, 5000.0))
        if k % 2 == 0:
            mesh = cases.tgv_mesh(n)
            t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
            U = cases.tgv_ic(t.pos, mach=mach).reshape(-1)
        else:
            mesh = cases.vortex_mesh(n)
            t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
            U = cases.vortex_ic(t.pos, u_inf=mach * math.sqrt(1.4)).reshape(-1)
        kw = dict(Re=Re, **BATCH_SCHEME) if scheme == "navier_stokes" else {}
        s = solver.Solver(t, precision=precision, cfl=0.5, **kw)
        s.set_state(U)
        out.append((s, U, dict(n=n, p=p, mach=mach, Re=Re, kind="tgv" if k % 2 == 0 else "vortex")))
    return out


def time_batch(n_inst, seed, precision, steps, warmup, stream, hiprt, err_steps=20, scheme="navier_stokes"):
    """One batched SSP-RK3 step of all members vs each member's own graph step;
    the FP32-bulk batch against an FP64 batch of the same members after
    err_steps steps."""
    from zhijian_amd import solver
    ns = scheme == "navier_stokes"
    mem = batch_members(n_inst, seed, precision, scheme)
    sols = [m[0] for m in mem]
    dts = np.array([s.dt_cfl(0.5) for s in sols])
    dof = sum(s.n for s in sols)
    by = 0
    for s, (_, _, meta) in zip(sols, mem):
        info = s.info()
        n_el = info["n_dof"] // (4 * info["n_sp"])
        if ns:
            by += sum(textbook_alg_bytes_per_stage(n_el, meta["p"], k, precision, info["affine"],
                                                   info["cross_tile_edges"], viscous=True) for k in (1, 2, 3))
        else:
            by += sum(alg_bytes_per_stage(n_el, meta["p"], k, precision, info["affine"], info["cross_tile_edges"])
                      for k in (1, 2, 3))
    batch = solver.Batch(sols)
    res = {}
    for single in (False, True):
        batch.set_launch(single)
        for policy in (0, 2, 4):
            batch.set_grid(policy)
            for _ in range(warmup):
                batch.rk3_step(dts, stream.ptr)
            stream.synchronize()
            e0, e1 = hiprt.Event(), hiprt.Event()
            e0.record(stream)
            for _ in range(steps):
                batch.rk3_step(dts, stream.ptr)
            e1.record(stream)
            e1.synchronize()
            res[("single" if single else "grouped", policy)] = e0.elapsed_ms(e1) / steps
    best = min(res, key=res.get)
    tuned = batch.autotune(dts, reps=max(2, steps // 4))
    for _ in range(warmup):
        for s, dt in zip(sols, dts):
            s.rk3_step(dt, stream.ptr)
    stream.synchronize()
    e0, e1 = hiprt.Event(), hiprt.Event()
    e0.record(stream)
    for _ in range(steps):
        for s, dt in zip(sols, dts):
            s.rk3_step(dt, stream.ptr)
    e1.record(stream)
    e1.synchronize()
    seq = e0.elapsed_ms(e1) / steps
    # one step backwards: the batched stage VJPs vs each member's own step VJP
    # (the cost does not depend on the cotangent's values)
    for s, (_, U, _) in zip(sols, mem):
        s.set_state(U)
        s.adjoint_set(U)
    k_adj = max(steps // 4, 3)
    adj = {}
    for name in ("batched", "sequential"):
        def back():
            if name == "batched":
                batch.rk3_step_vjp(dts, stream.ptr)
            else:
                for s, dt in zip(sols, dts):
                    s.rk3_step_vjp(dt, stream.ptr, want_dt=False)
        back()
        stream.synchronize()
        e0, e1 = hiprt.Event(), hiprt.Event()
        e0.record(stream)
        for _ in range(k_adj):
            back()
        e1.record(stream)
        e1.synchronize()
        adj[name] = e0.elapsed_ms(e1) / k_adj
    batch.close()
    err = None
    if err_steps > 0 and precision != "fp64":
        ref = batch_members(n_inst, seed, "fp64", scheme)
        for (s, U, _), (r, _, _) in zip(mem, ref):
            s.set_state(U)
        b32 = solver.Batch(sols)
        b64 = solver.Batch([m[0] for m in ref])
        for _ in range(err_steps):
            b32.rk3_step(dts)
            b64.rk3_step(dts)
        rel = []
        for s, (r, _, _) in zip(sols, ref):
            a, b = r.get_state(), s.get_state()
            rel.append((float(np.linalg.norm(b - a) / np.linalg.norm(a)), float(np.abs(b - a).max() / np.abs(a).max())))
        b32.close()
        b64.close()
        for r, _, _ in ref:
            r.close()
        worst = int(np.argmax([x[0] for x in rel]))
        err = dict(steps=err_steps, max_rel_l2=max(x[0] for x in rel), max_rel_linf=max(x[1] for x in rel),
                   median_rel_l2=float(np.median([x[0] for x in rel])), over="members (FP64 batch of the same meshes)",
                   worst_member=dict(index=worst, rel_l2=rel[worst][0], **mem[worst][2]),
                   members_above_1e_6=int(sum(x[0] > 1e-6 for x in rel)))
    achieved = by / (res[best] * 1e-3) / 1e9
    physics = ("Re~U[100,5000], alternating TGV/vortex; textbook FR (VCJH HU, Roe) + Navier-Stokes BR2 (Pr 0.72), "
               "no limiter" if ns else "alternating TGV/vortex; reference arithmetic (DG, inline Rusanov, "
               "positivity limiter)")
    return dict(workload=f"{n_inst} meshes, seed {seed}: N~U{{16..128}}, P~U{{1..5}}, Mach~U[0.1,0.8], {physics}, "
                         f"dt = each member's CFL 0.5 step", precision=precision, dof=dof,
                ms_per_step_by_launch={f"{k[0]}/grid{k[1]}": v for k, v in res.items()},
                launch=best[0], grid_policy=best[1],
                autotuner=dict(launch="single" if tuned[0] else "grouped", grid_policy=tuned[1], ms_per_step=tuned[2]),
                launches_per_stage=1 if best[0] == "single" else len({(m[2]["p"], s.info()["affine"])
                                                                       for s, m in zip(sols, mem)}),
                batched_ms_per_step=res[best], sequential_ms_per_step=seq, speedup=seq / res[best],
                dof_updates_per_s=3.0 * dof / (res[best] * 1e-3),
                roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=achieved / HBM_PEAK_GBS, bytes_per_step=by,
                              timing="graph-replayed batched step, launch gaps included (grouped: the "
                                     "groups' kernels run side by side on parallel graph branches; single: "
                                     "one launch per stage)"),
                adjoint=dict(batched_ms_per_step=adj["batched"], sequential_ms_per_step=adj["sequential"],
                             speedup=adj["sequential"] / adj["batched"], over_forward=adj["batched"] / res[best],
                             note="one step backwards: each member's stages 1-2 recomputed by its own launches, "
                                  "then one stage-VJP launch per (order, metric form, scheme) group"),
                fp32bulk_error_vs_fp64=err)


def time_textbook(mesh, tables, U0, steps, warmup, err_steps, stream, hiprt):
    """The textbook scheme the reference's settings name (VCJH HU correction,
    Roe flux, metric terms; stable), FP64 vs FP32-bulk: time per step and the
    FP32 error after err_steps steps from the same state (no limiter)."""
    from zhijian_amd import solver
    kw = dict(compat="textbook", riemann="roe", corr="hu", positivity=False, cfl=0.5)
    s64 = solver.Solver(tables, precision="fp64", **kw)
    s32 = solver.Solver(tables, precision="fp32bulk", **kw)
    s64.set_state(U0)
    dt = s64.dt_cfl(0.5)
    out = {"workload": "same mesh and state, textbook FR (VCJH HU, Roe, inviscid), no limiter", "dt": dt}
    for name, sol in (("fp64", s64), ("fp32bulk", s32)):
        ms, st = time_precision(sol, U0, dt, steps, warmup, stream, hiprt)
        info = sol.info()
        by = sum(textbook_alg_bytes_per_stage(tables.E, tables.p, k, name, info["affine"], info["cross_tile_edges"])
                 for k in (1, 2, 3))
        out[name] = dict(ms_per_step=ms, stage_ms=st, dof_updates_per_s=3.0 * sol.n / (ms * 1e-3),
                         roofline=roofline_of(by, st, f"stage_kernel<{'double' if name == 'fp64' else 'float'},"
                                                      f"P{tables.p},textbook>"))
    s64.set_state(U0)
    s32.set_state(U0)
    for _ in range(err_steps):
        s64.rk3_step(dt)
        s32.rk3_step(dt)
    a, b = s64.get_state(), s32.get_state()
    out["fp32bulk_error_vs_fp64"] = dict(steps=err_steps, rel_l2=float(np.linalg.norm(b - a) / np.linalg.norm(a)),
                                         rel_linf=float(np.abs(b - a).max() / np.abs(a).max()))
    out["fp32bulk_over_fp64"] = out["fp32bulk"]["dof_updates_per_s"] / out["fp64"]["dof_updates_per_s"]
    return out


def time_viscous(tables, U0, steps, warmup, err_steps, stream, hiprt, Re=1600.0):
    """The configuration's full physics: textbook FR (VCJH HU, Roe) with the
    compressible Navier-Stokes terms (BR2, Re 1600, Pr 0.72), FP64 vs
    FP32-bulk: time per step and the FP32 error after err_steps steps."""
    from zhijian_amd import solver
    kw = dict(compat="textbook", riemann="roe", corr="hu", positivity=False, cfl=0.5, viscous=True, Re=Re, Pr=0.72)
    s64 = solver.Solver(tables, precision="fp64", **kw)
    s32 = solver.Solver(tables, precision="fp32bulk", **kw)
    s64.set_state(U0)
    dt = s64.dt_cfl(0.5)
    out = {"workload": f"same mesh and state, textbook FR (VCJH HU, Roe) + Navier-Stokes BR2 (Re {Re:g}, Pr 0.72, "
                       f"penalty 2), no limiter", "dt": dt}
    for name, sol in (("fp64", s64), ("fp32bulk", s32)):
        ms, st = time_precision(sol, U0, dt, steps, warmup, stream, hiprt)
        info = sol.info()
        by = sum(textbook_alg_bytes_per_stage(tables.E, tables.p, k, name, True, info["cross_tile_edges"], True)
                 for k in (1, 2, 3))
        out[name] = dict(ms_per_step=ms, stage_ms=st, dof_updates_per_s=3.0 * sol.n / (ms * 1e-3),
                         roofline=roofline_of(by, st, f"visc_stage_kernel<{'double' if name == 'fp64' else 'float'},"
                                                      f"P{tables.p}>"))
    s64.set_state(U0)
    s32.set_state(U0)
    for _ in range(err_steps):
        s64.rk3_step(dt)
        s32.rk3_step(dt)
    a, b = s64.get_state(), s32.get_state()
    out["fp32bulk_error_vs_fp64"] = dict(steps=err_steps, rel_l2=float(np.linalg.norm(b - a) / np.linalg.norm(a)),
                                         rel_linf=float(np.abs(b - a).max() / np.abs(a).max()))
    out["fp32bulk_over_fp64"] = out["fp32bulk"]["dof_updates_per_s"] / out["fp64"]["dof_updates_per_s"]
    return out


def tri_alg_bytes_per_stage(E, p, stage, precision, cross_edges, viscous=False):
    """Algorithmic HBM bytes of one triangle stage launch: the FP64 stage state
    per solution point (stage 1 reads U^n and writes U1; stages 2 and 3 read
    U^n and the intermediate state and write one), 10 geometry words per
    element (affine metric, three face normals), 3 int2 links, and the
    (p+1) x 4 trace of every element edge whose neighbour lies in another tile,
    written once and read once in the compute precision (with the
    Navier-Stokes terms (p+1) x 12: U, dU/dx, dU/dy)."""
    Np = (p + 1) * (p + 2) // 2
    w = 8 if precision == "fp64" else 4
    per_pt = 64 if stage == 1 else 96
    tw = 12 if viscous else 4
    return E * (Np * per_pt + 10 * w + 24) + cross_edges * (p + 1) * tw * w * 2


def time_tri(n, p, steps, warmup, err_steps, stream, hiprt, cpu_seconds, Re=None):
    """Triangles (textbook FR, GA correction field, HLLC): n x n periodic
    unit square split along the cell diagonals, interior nodes jittered
    U(-0.2h, 0.2h) by PCG64(2024); Mach 0.5 at 2 degrees with 8 Fourier
    velocity modes of amplitude 0.05 (PCG64(2024)); with the Navier-Stokes
    terms (BR2) at Re when given.  FP64 vs FP32-bulk time per step, the FP32
    error after err_steps steps, and the triangle oracle (numpy, one core) on
    a bounded sample of the same workload."""
    from zhijian_amd import cases
    from zhijian_amd.tri import TriSolver, sp_coords
    m = cases.split_tri_mesh(n, jitter=0.2, seed=2024)
    U0 = cases.fourier_freestream_ic(sp_coords(m.xy, m.tris, p), seed=2024).reshape(-1)
    Np = (p + 1) * (p + 2) // 2
    dof = m.n_elem * Np * 4
    out = {"workload": f"{n}x{n} split-quad periodic triangles ({m.n_elem} elements), P{p} ({Np} points), "
                       f"jitter 0.2h seed 2024, Mach 0.5 at 2 deg + 8 Fourier modes (amp 0.05, seed 2024), "
                       f"textbook FR (GA correction, energy-stable family on triangles), HLLC, "
                       f"{f'Navier-Stokes (BR2) Re {Re:g}' if Re else 'inviscid'}, no limiter, CFL 0.5",
           "elements": m.n_elem, "dof": dof}
    sols = {}
    for name in ("fp64", "fp32bulk"):
        visc = dict(viscous=True, Re=Re) if Re else {}
        s = TriSolver(m.xy, m.tris, p, periodic=(1.0, 1.0), riemann="hllc", precision=name, corr="ga", **visc)
        s.set_state(U0)
        if name == "fp64":
            out["dt"] = dt = s.dt_cfl(0.5)
        for _ in range(warmup):
            s.rk3_step(dt, stream.ptr)
        stream.synchronize()
        e0, e1 = hiprt.Event(), hiprt.Event()
        e0.record(stream)
        for _ in range(steps):
            s.rk3_step(dt, stream.ptr)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        st = s.time_stages(dt, max(steps // 2, 5), stream.ptr)
        ce = s.info()["cross_tile_edges"]
        by = sum(tri_alg_bytes_per_stage(m.n_elem, p, k, name, ce, viscous=bool(Re)) for k in (1, 2, 3))
        kern = "tri_visc_stage_kernel" if Re else "tri_stage_kernel"
        out[name] = dict(ms_per_step=ms, stage_ms=st, dof_updates_per_s=3.0 * dof / (ms * 1e-3),
                         roofline=roofline_of(by, st, f"{kern}<{'double' if name == 'fp64' else 'float'},P{p}>",
                                              cross_tile_edge_fraction=ce / (3.0 * m.n_elem)))
        sols[name] = s
    for s in sols.values():
        s.set_state(U0)
    for _ in range(err_steps):
        for s in sols.values():
            s.rk3_step(dt)
    a, b = sols["fp64"].get_state(), sols["fp32bulk"].get_state()
    out["fp32bulk_error_vs_fp64"] = dict(steps=err_steps, rel_l2=float(np.linalg.norm(b - a) / np.linalg.norm(a)),
                                         rel_linf=float(np.abs(b - a).max() / np.abs(a).max()))
    out["fp32bulk_over_fp64"] = out["fp32bulk"]["dof_updates_per_s"] / out["fp64"]["dof_updates_per_s"]
    # one step backwards (its cost does not depend on the cotangent's values)
    for name, s in sols.items():
        s.set_state(U0)
        s.adjoint_set(U0)
        s.rk3_step_vjp(dt, stream.ptr, want_dt=False)
        stream.synchronize()
        k = max(steps // 4, 3)
        e0, e1 = hiprt.Event(), hiprt.Event()
        e0.record(stream)
        for _ in range(k):
            s.rk3_step_vjp(dt, stream.ptr, want_dt=False)
        e1.record(stream)
        e1.synchronize()
        ms_adj = e0.elapsed_ms(e1) / k
        out[name]["adjoint"] = dict(
            adj_ms_per_step=ms_adj, adjoint_over_forward=ms_adj / out[name]["ms_per_step"],
            note="adjoint step = traces of U^n " + ("and of the two recomputed stages " if Re else "")
                 + "+ stages 1-2 recomputed + 3 VJP launches")
    for s in sols.values():
        s.close()
    if cpu_seconds > 0:
        from oracle import tri_oracle as T
        from zhijian_amd.tri import tri_faces
        ns = min(n, 64)
        ms_ = cases.split_tri_mesh(ns, jitter=0.2, seed=2024)
        faces, nrm = tri_faces(ms_.xy, ms_.tris, periodic=(1.0, 1.0))
        o = T.TriTextbook(p, ms_.xy, ms_.tris, faces, nrm, riemann="hllc", corr="ga")
        if Re:
            o.set_viscous(Re)
        U = cases.fourier_freestream_ic(sp_coords(ms_.xy, ms_.tris, p), seed=2024).reshape(-1)
        U = o.advance(U, dt)
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            U = o.advance(U, dt)
            k += 1
        el = time.perf_counter() - t0
        out["cpu_port"] = dict(value=3.0 * ms_.n_elem * Np * 4 * k / el, unit="DOF-updates/s", cores=1, kind="port",
                               sample=f"{k} SSP-RK3 steps of the {ns}x{ns} mesh of the same family "
                                      f"({el:.1f} s, numpy oracle)")
    return out


def measured_traffic(n, p):
    """HBM bytes per stage launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"stage_kernel_tgv{n}_p{p}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic:"):
            f = line.split()
            return {"fp32bulk": float(f[2]), "fp64": float(f[5]),
                    "source": f"profiles/{os.path.basename(path)} (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, "
                              f"scaled by a plain copy kernel's readings; not measured in this run)"}
    return None


LIBHDF5_PY = "/opt/conda/bin/python3.9"
_H5PY_WRITE = """
import sys, time, numpy as np, h5py
sol = np.load(sys.argv[2]); nsp = np.load(sys.argv[3])
best_w = best_r = 1e30
for _ in range(3):
    t0 = time.perf_counter()
    with h5py.File(sys.argv[1], "w") as f:
        f.attrs.create("time", 0.5, dtype="f8"); f.attrs.create("iteration", 10, dtype="i4")
        f.attrs.create("poly_order", 4, dtype="i4"); f.attrs.create("gamma", 1.4, dtype="f8")
        f.create_dataset("solution", data=sol); f.create_dataset("elem_n_sp", data=nsp)
    best_w = min(best_w, time.perf_counter() - t0)
    t0 = time.perf_counter()
    with h5py.File(sys.argv[1], "r") as f:
        a = f["solution"][:]; b = f["elem_n_sp"][:]; t = float(f.attrs["time"])
    best_r = min(best_r, time.perf_counter() - t0)
print(best_w * 1e3, best_r * 1e3)
"""


def time_io(sol, mesh, p, reps=3):
    """Restart save / load of the device state (zj_restart_save / _load) and
    VTK output of it, on the bench workload.  Files go to a temporary
    directory that is removed afterwards.  With an installed libhdf5 (h5py),
    the same restart layout written and read by libhdf5 is timed beside."""
    import shutil
    import subprocess
    import tempfile
    from zhijian_amd import io
    d = tempfile.mkdtemp(prefix="zj_io_")
    out = {}
    try:
        path = os.path.join(d, "restart.h5")
        U = sol.get_state()
        nbytes = U.nbytes
        t_get = min(_timed(sol.get_state) for _ in range(reps))
        t_save = min(_timed(lambda: sol.save_restart(path, 0.5, 10)) for _ in range(reps))
        t_load = min(_timed(lambda: sol.load_restart(path)) for _ in range(reps))
        fsize = os.path.getsize(path)
        out["restart"] = dict(file_bytes=fsize, state_bytes=nbytes, get_state_ms=t_get * 1e3,
                              save_ms=t_save * 1e3, load_ms=t_load * 1e3, save_gbps=fsize / t_save / 1e9,
                              load_gbps=fsize / t_load / 1e9)
        if os.path.exists(LIBHDF5_PY):
            np.save(os.path.join(d, "sol.npy"), U)
            np.save(os.path.join(d, "nsp.npy"), np.full(mesh.elem_nodes.shape[0], (p + 1) ** 2, np.int32))
            r = subprocess.run([LIBHDF5_PY, "-c", _H5PY_WRITE, os.path.join(d, "h5py.h5"), os.path.join(d, "sol.npy"),
                                os.path.join(d, "nsp.npy")], capture_output=True, text=True, timeout=300)
            if r.returncode == 0:
                w, rd = (float(x) for x in r.stdout.split())
                out["restart"]["libhdf5_h5py"] = dict(write_ms=w, read_ms=rd,
                                                      note="same layout, host arrays, h5py 3.3 / HDF5 1.10.6")
        m = io.VisMesh(mesh.xy, mesh.elem_nodes, 1, 1)
        nsp = (p + 1) ** 2
        for name, flags in (("vtu_binary", io.BINARY), ("vtu_binary_interpolated", io.BINARY | io.INTERPOLATE),
                            ("vtu_ascii", io.VTU)):
            f = os.path.join(d, name + ".vtu")
            n_rep = 1 if name == "vtu_ascii" else reps
            t = min(_timed(lambda: io.write_vtk(f, m, nsp, U, p, 1.4, 4, flags)) for _ in range(n_rep))
            out[name] = dict(ms=t * 1e3, file_bytes=os.path.getsize(f), subdivisions=4,
                             points=mesh.elem_nodes.shape[0] * 25, mb_per_s=os.path.getsize(f) / t / 1e6)
            os.remove(f)
        out["host_threads"] = min(os.cpu_count() or 1, 16)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def time_gpu_api(tables, U0, reps, stream, hiprt):
    """The gpu:: dense operator products (interpolateToFluxPoints,
    computeReferenceGradients, computeFluxDivergence: the reference's K1, K2,
    K11 on AoS arrays) on the FP64 matrix cores and on the VALU, same mesh.
    Algorithmic bytes: the AoS inputs read and outputs written once (8-byte
    words); the operator tables are negligible.  Flops: 2 per multiply-add of
    the dense products."""
    from zhijian_amd import gpu_ops as G
    E, S, p = tables.E, tables.S, tables.p
    Fp = p + 1
    from zhijian_amd import _lib
    ops = _lib.quad_operators(p)
    DA = hiprt.DeviceArray
    U = DA(U0)
    Jinv = DA(np.tile(np.array([1.0, 0.0, 0.0, 1.0]), E * S))
    ufp = DA(E * 4 * Fp * 4)
    gx, ge, fx, fy, div = (DA(E * S * 4) for _ in range(5))
    interp, dxi, deta = DA(ops["interp"]), DA(ops["diff_xi"]), DA(ops["diff_eta"])
    G.compute_inviscid_flux_sp(U.ptr, fx.ptr, fy.ptr, 1.4, E, S, stream.ptr)
    calls = {
        "interpolateToFluxPoints": (lambda: G.interpolate_to_flux_points(U.ptr, ufp.ptr, interp.ptr, E, S, 4, Fp,
                                                                         stream.ptr),
                                    8 * E * (S * 4 + 4 * Fp * 4), 2.0 * E * 4 * Fp * S * 4),
        "computeReferenceGradients": (lambda: G.compute_reference_gradients(U.ptr, gx.ptr, ge.ptr, dxi.ptr, deta.ptr,
                                                                            E, S, stream.ptr),
                                      8 * E * S * 4 * 3, 2.0 * 2 * E * S * S * 4),
        "computeFluxDivergence": (lambda: G.compute_flux_divergence(fx.ptr, fy.ptr, div.ptr, dxi.ptr, deta.ptr,
                                                                    Jinv.ptr, E, S, stream.ptr),
                                  8 * E * S * (4 * 3 + 4), 2.0 * 4 * E * S * S * 4),
    }
    out = {}
    for eng, name in ((G.ENGINE_MFMA, "mfma_f64"), (G.ENGINE_VALU, "valu_f64")):
        G.set_operator_engine(eng)
        legs = {}
        for fn, (call, by, fl) in calls.items():
            for _ in range(3):
                call()
            e0, e1 = hiprt.Event(), hiprt.Event()
            e0.record(stream)
            for _ in range(reps):
                call()
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / reps
            legs[fn] = dict(ms=ms, gbps=by / (ms * 1e-3) / 1e9, hbm_frac=by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                            tflops=fl / (ms * 1e-3) / 1e12, bytes=by, flops=fl)
        out[name] = legs
    G.set_operator_engine(G.ENGINE_MFMA)
    return dict(workload=f"TGV {int(round(math.sqrt(E)))}x{int(round(math.sqrt(E)))} P{p} AoS arrays, FP64 "
                         f"(the reference's gpu:: layout)",
                timing="HIP events around back-to-back launches on one stream", **out)


def _timed(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_precision(sol, U0, dt, steps, warmup, stream, hiprt):
    """Graph-replayed steps and per-stage kernel times (ms)."""
    sol.set_state(U0)
    for _ in range(warmup):
        sol.rk3_step(dt, stream.ptr)
    stream.synchronize()
    e0, e1 = hiprt.Event(), hiprt.Event()
    hiprt.device_synchronize()
    e0.record(stream)
    for _ in range(steps):
        sol.rk3_step(dt, stream.ptr)
    e1.record(stream)
    e1.synchronize()
    hiprt.device_synchronize()
    ms_step = e0.elapsed_ms(e1) / steps
    # each stage kernel's own execution time: start/stop events that the
    # runtime records at the kernel's start and end (no launch gaps)
    st_ms = sol.time_stages(dt, max(steps // 2, 5), stream.ptr)
    hiprt.device_synchronize()
    return ms_step, st_ms


def time_vortex(steps, warmup, stream, hiprt):
    """The isentropic-vortex case: 10x10 quads on [-5,5]^2, P3 (6,400 DOF),
    reference arithmetic, dt = min(CFL dt, 1e-3).  0.3 MB per stage: the step
    is launch-bound, so steps/s is the figure, not a roofline fraction."""
    from zhijian_amd import cases, solver
    mesh = cases.vortex_mesh(10)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 3, periodic=(mesh.Lx, mesh.Ly))
    U0 = cases.vortex_ic(t.pos).reshape(-1)
    out = {"workload": "isentropic vortex, 10x10 quads on [-5,5]^2, P3, reference arithmetic", "dof": t.E * t.S * 4}
    sols = {}
    for name in ("fp64", "fp32bulk"):
        s = solver.Solver(t, precision=name, cfl=0.5)
        s.set_state(U0)
        dt = min(s.dt_cfl(0.5), 1e-3)
        ms, st = time_precision(s, U0, dt, steps, warmup, stream, hiprt)
        out[name] = dict(ms_per_step=ms, rk_steps_per_s=1e3 / ms, stage_ms=st,
                         dof_updates_per_s=3.0 * out["dof"] / (ms * 1e-3))
        sols[name] = s
    out["dt"] = dt

    def err_after(a, b, n):
        for s in (a, b):
            s.set_state(U0)
        errs = {}
        for k in range(1, n + 1):
            a.rk3_step(dt)
            b.rk3_step(dt)
            if k in (20, n):
                x, y = a.get_state(), b.get_state()
                errs[str(k)] = dict(rel_l2=float(np.linalg.norm(y - x) / np.linalg.norm(x)),
                                    rel_linf=float(np.abs(y - x).max() / np.abs(x).max()))
        return errs
    # the configuration's 100 steps; the reference arithmetic is linearly
    # unstable, so its FP32 error is also stated at 20 steps, before the
    # limiter's clamps engage, and the stable textbook scheme is shown beside
    out["fp32bulk_error_vs_fp64"] = err_after(sols["fp64"], sols["fp32bulk"], 100)
    kw = dict(compat="textbook", riemann="rusanov", corr="dg", positivity=False, cfl=0.5)
    out["textbook_fp32bulk_error_vs_fp64"] = err_after(solver.Solver(t, precision="fp64", **kw),
                                                       solver.Solver(t, precision="fp32bulk", **kw), 100)
    return out


def cpu_baseline(mesh, p, U0, dt, budget_s):
    """The oracle (a plain-C restatement of the reference's executed kernels)
    on the same workload, OpenMP over elements and faces on this host."""
    from oracle import pyoracle as O
    from tests import fr_case
    cores = len(os.sched_getaffinity(0))
    case = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
    o = case.solver()
    O.set_threads(cores)
    U = U0.copy()
    t0 = time.perf_counter()
    U = o.advance(U, dt, True)       # warm-up step
    t_first = time.perf_counter() - t0
    nsteps = max(1, min(200, int(budget_s / max(t_first, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(nsteps):
        U = o.advance(U, dt, True)
    el = time.perf_counter() - t0
    O.set_threads(1)
    dof = case.E * case.S * 4
    side = int(round(case.E ** 0.5))
    # the same restatement on one thread, a few steps (about a fifth of the budget)
    t0 = time.perf_counter()
    U = o.advance(U, dt, True)
    t1 = time.perf_counter() - t0
    n1 = max(1, min(50, int(0.2 * budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(n1):
        U = o.advance(U, dt, True)
    el1 = time.perf_counter() - t0
    return dict(value=3.0 * dof * nsteps / el, unit="DOF-updates/s", cores=cores, kind="port",
                sample=f"{nsteps} SSP-RK3 steps of the full {side}x{side} P{p} TGV mesh "
                       f"({el:.1f} s, after 1 warm-up step), OpenMP over elements and faces",
                rk_steps_per_s=nsteps / el,
                single_thread=dict(value=3.0 * dof * n1 / el1, unit="DOF-updates/s", cores=1,
                                   sample=f"{n1} steps of the same mesh ({el1:.1f} s, after 1 step)"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256, help="elements per side")
    ap.add_argument("--p", type=int, default=4)
    # the reference arithmetic is linearly unstable: round-off differences grow
    # exponentially, so the FP32-vs-FP64 error is stated after a short run
    ap.add_argument("--err-steps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--textbook-steps", type=int, default=100,
                    help="steps of the textbook-scheme FP32 error measurement (0: skip the textbook leg)")
    ap.add_argument("--batch", type=int, default=64,
                    help="members of the batched-step measurement (0: skip)")
    ap.add_argument("--tri-n", type=int, default=500,
                    help="cells per side of the triangle leg (each split in two; 0: skip)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the whole C5 gradient (Navier-Stokes, checkpointed and whole-tape), FP64 and FP32-bulk")
    ap.add_argument("--batch-only", action="store_true", help="only the Navier-Stokes batch (--batch members)")
    ap.add_argument("--tri-re", type=float, default=0.0,
                    help="with --tri-only: the Navier-Stokes terms at this Reynolds number (0: Euler)")
    ap.add_argument("--tri-only", action="store_true",
                    help="run only the triangle leg and print its JSON (profiling runs)")
    ap.add_argument("--io-only", action="store_true",
                    help="only the restart / VTK output timings on the bench workload")
    ap.add_argument("--no-io", dest="io", action="store_false", help="skip the restart / VTK output timings")
    ap.add_argument("--ns-only", action="store_true",
                    help="run only the Navier-Stokes (BR2) leg on the TGV mesh and print its JSON (profiling runs)")
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false",
                    help="skip the adjoint/forward cost ratio (TGV 128^2 P3)")
    ap.add_argument("--only", choices=("fp64", "fp32bulk"), default=None,
                    help="time one precision only (profiling runs)")
    ap.add_argument("--probe-bytes", type=float, default=1 << 30,
                    help="bytes per buffer of the plain copy kernel that measures achievable HBM bandwidth (0: skip)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU owning the whole mesh; --gpus {args.gpus} "
              f"(WORLD_SIZE={world}) is not supported", file=sys.stderr)
        sys.exit(2)

    from zhijian_amd import _lib, cases, hiprt, solver

    if args.tri_only:
        if args.probe_bytes > 0:
            log(f"copy kernel: {_lib.copy_probe(int(args.probe_bytes), 10):.0f} GB/s")
        tri = time_tri(args.tri_n, 3, args.steps, args.warmup, args.err_steps, hiprt.Stream(), hiprt,
                       0.0 if args.no_cpu else 5.0, Re=args.tri_re if args.tri_re > 0 else None)
        print(json.dumps({"triangles": tri}), flush=True)
        return

    if args.ns_only:
        mesh = cases.tgv_mesh(args.n)
        tables = solver.build_tables(mesh.xy, mesh.elem_nodes, args.p, periodic=(mesh.Lx, mesh.Ly))
        if args.probe_bytes > 0:
            log(f"copy kernel: {_lib.copy_probe(int(args.probe_bytes), 10):.0f} GB/s")
        ns = time_viscous(tables, cases.tgv_ic(tables.pos, mach=0.1).reshape(-1), args.steps, args.warmup,
                          args.err_steps, hiprt.Stream(), hiprt)
        print(json.dumps({"navier_stokes": ns}), flush=True)
        return

    if args.adjoint_only:
        stream = hiprt.Stream()
        adj = {}
        for prec in ("fp64", "fp32bulk"):
            adj[prec] = dict(checkpointed_gradient=time_gradient(128, 3, prec, 256, 16, stream, hiprt, Re=1600.0),
                             taped_gradient=time_gradient(128, 3, prec, 256, 16, stream, hiprt, Re=1600.0,
                                                          tape_segments=16))
        print(json.dumps({"adjoint": adj}), flush=True)
        return

    if args.batch_only:
        stream = hiprt.Stream()
        b = time_batch(args.batch, 4, "fp32bulk", max(args.steps // 5, 10), 3, stream, hiprt, err_steps=args.err_steps)
        print(json.dumps({"batched": b}), flush=True)
        return

    if args.io_only:
        mesh = cases.tgv_mesh(args.n)
        tables = solver.build_tables(mesh.xy, mesh.elem_nodes, args.p, periodic=(mesh.Lx, mesh.Ly))
        s = solver.Solver(tables, precision="fp64")
        s.set_state(cases.tgv_ic(tables.pos, mach=0.1).reshape(-1))
        print(json.dumps({"io": time_io(s, mesh, args.p)}), flush=True)
        return

    mesh = cases.tgv_mesh(args.n)
    t_build = time.perf_counter()
    tables = solver.build_tables(mesh.xy, mesh.elem_nodes, args.p, periodic=(mesh.Lx, mesh.Ly))
    log(f"tables: {tables.E} elements, {tables.faces.shape[0]} faces, {time.perf_counter() - t_build:.2f} s")
    U0 = cases.tgv_ic(tables.pos, mach=0.1).reshape(-1)
    E, S = tables.E, tables.S
    dof = E * S * 4

    s64 = solver.Solver(tables, precision="fp64", cfl=0.5)
    s32 = solver.Solver(tables, precision="fp32bulk", cfl=0.5)
    s64.set_state(U0)
    dt = s64.dt_cfl(0.5)   # A21: params.dt = 0 -> the CFL step of the initial state
    stream = hiprt.Stream()

    copy_gbps = None
    if args.probe_bytes > 0:
        copy_gbps = _lib.copy_probe(int(args.probe_bytes), 10)
        log(f"copy kernel: {copy_gbps:.0f} GB/s over 2 x {args.probe_bytes / 2**20:.0f} MiB")

    res = {}
    traffic = measured_traffic(args.n, args.p)
    for name, sol in (("fp64", s64), ("fp32bulk", s32)):
        if args.only and name != args.only:
            continue
        ms_step, st_ms = time_precision(sol, U0, dt, args.steps, args.warmup, stream, hiprt)
        w = 8 if name == "fp64" else 4
        info = sol.info()
        affine = info["affine"]
        ce = info["cross_tile_edges"]
        by = [alg_bytes_per_stage(E, args.p, st, name, affine, ce) for st in (1, 2, 3)]
        rl = roofline_of(sum(by), st_ms, f"stage_kernel<{'double' if w == 8 else 'float'},P{args.p}>",
                         affine_metric=affine, cross_tile_edge_fraction=ce / (4.0 * E),
                         step_gap_ms=ms_step - sum(st_ms))
        rl["traffic"] = None if traffic is None else traffic[name]
        rl["traffic_source"] = None if traffic is None else traffic["source"]
        res[name] = dict(ms_per_step=ms_step, stage_ms=st_ms,
                         dof_updates_per_s=3.0 * dof / (ms_step * 1e-3),
                         rk_steps_per_s=1e3 / ms_step, roofline=rl)
        log(f"{name}: {ms_step:.4f} ms/step, stage kernels {['%.4f' % x for x in st_ms]} ms, "
            f"{res[name]['dof_updates_per_s']:.3e} DOF-upd/s, {rl['achieved']:.0f} GB/s")

    if args.only:
        print(json.dumps({"only": args.only, "copy_gbps": copy_gbps, **res[args.only]}), flush=True)
        return

    # FP32-bulk error against FP64 after a fixed number of steps from the same state
    err = None
    if args.err_steps > 0:
        s64.set_state(U0); s32.set_state(U0)
        for _ in range(args.err_steps):
            s64.rk3_step(dt); s32.rk3_step(dt)
        a, b = s64.get_state(), s32.get_state()
        err = dict(steps=args.err_steps, rel_l2=float(np.linalg.norm(b - a) / np.linalg.norm(a)),
                   rel_linf=float(np.abs(b - a).max() / np.abs(a).max()))
        log(f"fp32bulk vs fp64 after {args.err_steps} steps: {err}")

    textbook = None
    if args.textbook_steps > 0:
        textbook = time_textbook(mesh, tables, U0, args.steps, args.warmup, args.textbook_steps, stream, hiprt)
        log(f"textbook: {textbook}")

    viscous = None
    if args.textbook_steps > 0:
        viscous = time_viscous(tables, U0, args.steps, args.warmup, args.textbook_steps, stream, hiprt)
        log(f"viscous: {viscous}")

    batched = None
    if args.batch > 0:
        batched = time_batch(args.batch, 4, "fp32bulk", max(args.steps // 5, 10), 3, stream, hiprt,
                             err_steps=args.err_steps)
        log(f"batched: {batched}")
        batched["reference_arithmetic"] = time_batch(args.batch, 4, "fp32bulk", max(args.steps // 5, 10), 3, stream,
                                                     hiprt, err_steps=args.err_steps, scheme="reference")
        log(f"batched, reference arithmetic: {batched['reference_arithmetic']}")

    adj = {}
    if args.adjoint:
        for prec in ("fp64", "fp32bulk"):
            adj[prec] = time_adjoint(128, 3, prec, args.steps, args.warmup, stream, hiprt)
            log(f"adjoint {prec}: {adj[prec]}")
            adj[prec]["checkpointed_gradient"] = time_gradient(128, 3, prec, 256, 16, stream, hiprt, Re=1600.0)
            log(f"gradient {prec}: {adj[prec]['checkpointed_gradient']}")
            adj[prec]["checkpointed_gradient_euler"] = time_gradient(128, 3, prec, 256, 16, stream, hiprt)
            log(f"gradient {prec} (Euler terms): {adj[prec]['checkpointed_gradient_euler']}")
            # the whole trajectory on the tape (0.1 GB per step here: 26 GB of HBM), no step computed twice
            adj[prec]["taped_gradient"] = time_gradient(128, 3, prec, 256, 16, stream, hiprt, Re=1600.0,
                                                        tape_segments=16)
            log(f"gradient {prec}, whole tape: {adj[prec]['taped_gradient']}")
            adj[prec]["taped_gradient_euler"] = time_gradient(128, 3, prec, 256, 16, stream, hiprt, tape_segments=16)
            log(f"gradient {prec} (Euler terms), whole tape: {adj[prec]['taped_gradient_euler']}")

    gpu_api = time_gpu_api(tables, U0, max(args.steps // 5, 10), stream, hiprt)
    log(f"gpu api: {gpu_api}")

    tri = None
    if args.tri_n > 0:
        tri = time_tri(args.tri_n, 3, args.steps, args.warmup, args.textbook_steps, stream, hiprt,
                       0.0 if args.no_cpu else 5.0)
        log(f"triangles: {tri}")
        tri["navier_stokes"] = time_tri(args.tri_n, 3, args.steps, args.warmup, args.textbook_steps, stream, hiprt,
                                        0.0 if args.no_cpu else 5.0, Re=5000.0)
        log(f"triangles, Navier-Stokes: {tri['navier_stokes']}")

    vortex = time_vortex(10 * args.steps, args.warmup, stream, hiprt)
    log(f"vortex: {vortex}")

    io_res = None
    if args.io:
        io_res = time_io(s64, mesh, args.p)
        log(f"io: {io_res}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, args.p, U0, dt, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    head = res["fp32bulk"]
    out = {
        "metric": METRIC,
        "value": head["dof_updates_per_s"],
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "state_dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D Taylor-Green vortex, periodic [0,2pi]^2, {args.n}x{args.n} quads, "
                               f"P{args.p}, Mach 0.1, reference arithmetic (DG correction, inline Rusanov, "
                               f"inviscid), SSP-RK3 + positivity limiter, CFL 0.5",
                   "elements": E, "solution_points": E * S, "dof": dof, "dt": dt,
                   "precision": "fp32bulk: flux, interpolation, Riemann flux, divergence and correction in FP32 "
                                "(VALU; no MFMA in the stage kernels); U^n, the RK combination, limiter, norms "
                                "and dt in FP64; the intermediate stage states stored as FP32 increments over "
                                "U^n"},
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "rk_steps_per_s": head["rk_steps_per_s"],
        "fp64_plain": res["fp64"],
        "fp32bulk_over_fp64": head["dof_updates_per_s"] / res["fp64"]["dof_updates_per_s"],
        "fp32bulk_error_vs_fp64": err,
        "adjoint_over_forward": adj.get("fp64", {}).get("checkpointed_gradient", {}).get("gradient_over_forward"),
        "adjoint_over_forward_basis": "whole checkpointed gradient (256 steps, checkpoints every 16, TGV 128x128 "
                                      "P3 textbook with the Navier-Stokes terms at Re 1600, FP64) over the same "
                                      "256 forward steps",
        "adjoint_over_forward_whole_tape": adj.get("fp64", {}).get("taped_gradient", {}).get("gradient_over_forward"),
        "adjoint": adj or None,
        "batched": batched,
        "textbook": textbook,
        "navier_stokes": viscous,
        "triangles": tri,
        "gpu_api": gpu_api,
        "vortex": vortex,
        "io": io_res,
        "copy_kernel_gbps": copy_gbps,
    }
    print(json.dumps(out), flush=True)


def reap_children():
    """Waits for any child process still attached (none should be left)."""
    try:
        import psutil
        for c in psutil.Process().children(recursive=True):
            c.wait(timeout=30)
    except Exception:
        pass


if __name__ == "__main__":
    try:
        main()
    finally:
        reap_children()
