"""Benchmark of the explicit FP64 residual path on one MI355X.

Workload (config 2): second-order Euler, HLLC, weighted least-squares
primitive gradients with the Venkatakrishnan limiter (K = 5), SSP-RK3 with
local time stepping at CFL 0.8, on a 1024 x 512 NACA 0012 O-grid split into
1,048,576 triangles (far field 50 chords, wall-normal ratio 1.05), slip wall
and far field, M 0.8 at 1.25 degrees.  Synthetic input: the mesh is
generated, the flow starts from the free stream.

One step = one call of the solver's step() followed by the residual read
back, as the reference's driver loop does (src/main.cpp:3491-3515); a
cell-update is one cell through one RK stage.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

# the OpenMP CPU baselines: one thread per core, pinned close (set before any
# OpenMP runtime is loaded; an explicit setting in the environment wins)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from finitevolume_amd import abi  # noqa: E402
from finitevolume_amd import mesh as M  # noqa: E402

METRIC = "FP64 residual Gcell-updates/s (1 GPU, % HBM roofline); adjoint/primal cost"
HBM_PEAK_GBS = 8000.0
GAMMA = 1.4

# Algorithmic bytes per cell-update (per cell per RK stage): every state array
# read and written once, the RK base state on stages 2-3, per-cell static data
# (volume 8, centroid 16, LSQ inverse 24), per-face static data (owner and
# neighbour 8, normal and length 24, midpoint 16) times faces per cell, and the
# local dt written on stage 1 and read on stages 2-3.
def b_alg_config2(faces_per_cell: float) -> float:
    stage = 32 + 32 + 8 + 16 + 24 + faces_per_cell * 48 + 8
    return stage + (32 * 2 + 8) / 3.0


def b_alg_2r(faces_per_cell: float) -> float:
    return 32 + 32 + 8 + faces_per_cell * 32


def naca_case(ni, nj, order2=True):
    m = M.naca_ogrid(ni, nj, code="0012", r_far=50.0, ratio=1.05)
    c = 1.0
    a = math.radians(1.25)
    rho, p = 1.0, 1.0 / GAMMA
    u, v = 0.8 * c * math.cos(a), 0.8 * c * math.sin(a)
    ff = M.from_primitive(rho, u, v, p, GAMMA)
    bcs = [abi.make_bc(abi.BC_WALL), abi.make_bc(abi.BC_FARFIELD, ff)]
    if order2:
        params = abi.default_params(equation=abi.EQ_EULER, flux_scheme=abi.FLUX_HLLC, cfl=0.8, order=2,
                                    limiter=abi.LIMITER_VENKATAKRISHNAN, venkat_k=5.0, rk_stages=3, local_dt=1)
    else:
        params = abi.default_params(equation=abi.EQ_EULER, flux_scheme=abi.FLUX_HLLC, cfl=0.8)
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=rho, u=u, v=v, p=p)
    return m, bcs, params, ic


def time_gpu(solver, steps, warmup, per_step=True):
    solver.step(warmup)
    solver.synchronize()
    t0 = time.perf_counter()
    if per_step:
        for _ in range(steps):
            solver.step(1)
            r = solver.residual()
            if not np.all(np.isfinite(r)):
                raise RuntimeError(f"residual not finite: {r}")
    else:
        solver.step(steps)
    solver.synchronize()
    return time.perf_counter() - t0


def timed_steps(solver, seconds):
    """Whole CPU steps until `seconds` of wall time have passed (at least one)."""
    n = 0
    t0 = time.perf_counter()
    while True:
        solver.step(1)
        n += 1
        t = time.perf_counter() - t0
        if t >= seconds:
            return n, t


def adjoint_leg(s, m, bcs, params, ic, steps_done, reps=5, n_check=32):
    """Discrete adjoint of the config-2 residual at its 1000-step state:
    device time of one VJP (state and node coordinates, lambda ~ N(0, 1) from
    default_rng(0)) against one residual evaluation with the same gradient
    and limiter recompute; and finite differences of lambda . R at 32 state
    entries and 32 node coordinates drawn by default_rng(1) among those whose
    adjoint is at least 1e-4 of the largest, each evaluated on the six-ring
    sub-mesh around the perturbed entity (the residual there is the whole
    mesh's up to entries the difference cancels), relative to the entry
    itself at the best of a sweep of steps (finitevolume_amd.fdcheck)."""
    from finitevolume_amd import abi
    from finitevolume_amd.fdcheck import fd_sweep
    from finitevolume_amd.solver import FVKSolver
    if steps_done < 1000:
        s.step(1000 - steps_done)
    U = s.field()
    nc = m.n_cells
    lam = np.random.default_rng(0).standard_normal((nc, 4))
    s.residual_eval()
    s.residual_vjp(lam, want_x=True)  # builds the adjoint's index maps
    s.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        s.residual_eval()
    wall_p = (time.perf_counter() - t0) / reps
    tp = s.kernel_times()
    s.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        Ub, Xb = s.residual_vjp(lam, want_x=True)
    wall_a = (time.perf_counter() - t0) / reps
    ta = s.kernel_times()
    s.set_kernel_timing(True)
    for _ in range(reps):
        s.residual_vjp(lam, want_x=False)
    tu = s.kernel_times()
    s.set_kernel_timing(False)
    dev_p = sum(tot for tot, _ in tp.values()) / reps
    dev_a = sum(tot for tot, _ in ta.values()) / reps
    dev_u = sum(tot for tot, _ in tu.values()) / reps

    # the adjoint solve's iteration (SSP-RK3 in pseudo time: three state-only
    # VJPs and updates at the fixed state, whose forward gradients the solve
    # forms once): the wall-time difference of a 40- and a 20-iteration solve
    # of the lift objective, so set-up, the final node-coordinate pass and
    # the copies cancel
    def solve_wall(iters):
        s.synchronize()
        t = time.perf_counter()
        s.adjoint_solve_forces(1.0, 0.5, 0.0, max_iter=iters, tol=0.0, check_every=iters)
        return time.perf_counter() - t

    solve_wall(20)
    it_ms = (solve_wall(40) - solve_wall(20)) / 20 * 1e3
    out = {
        "workload": "config4: VJP of the config-2 residual (state + node coordinates) at the 1000-step state",
        "vjp_device_ms": round(dev_a, 4),
        "primal_residual_device_ms": round(dev_p, 4),
        "adjoint_over_primal": round(dev_a / dev_p, 3),
        "vjp_state_only_device_ms": round(dev_u, 4),
        "adjoint_state_only_over_primal": round(dev_u / dev_p, 3),
        "adjoint_solve_iteration_ms": round(it_ms, 4),
        "adjoint_solve_vjp_over_primal": round(it_ms / 3 / dev_p, 3),
        "adjoint_solve_note": "per VJP of the adjoint solve: a third of one pseudo-time RK3 iteration (state-only "
                              "VJP and update, at the solve's fixed state and forward gradients), by wall time",
        "vjp_kernel_avg_ms": {k: round(tot / cnt, 5) for k, (tot, cnt) in sorted(ta.items()) if cnt},
        "vjp_wall_ms": round(wall_a * 1e3, 3),
        "primal_residual_wall_ms": round(wall_p * 1e3, 3),
        "wall_note": "wall times include the host copies and reordering of lambda, U_bar, X_bar and R",
        "fd_check": None,
    }
    if n_check <= 0:
        return out

    cut_bc = next((b for b in bcs if b.type == abi.BC_FARFIELD), abi.make_bc(abi.BC_OUTFLOW))
    sub_bcs = list(bcs) + [cut_bc]
    cut = len(bcs)
    steps = [1e-3, 1e-4, 1e-5, 1e-6, 1e-7]
    rng = np.random.default_rng(1)

    def draw(vals, n):
        idx = np.nonzero(np.abs(vals) >= 1e-4 * np.abs(vals).max())[0]
        return rng.choice(idx, min(n, idx.size), replace=False)

    def summary(errs, kinks):
        e = np.array(errs) if errs else np.array([np.nan])
        return {"entries": len(errs) + kinks, "kinks": kinks, "median_rel_err": float(f"{np.median(e):.3e}"),
                "p90_rel_err": float(f"{np.quantile(e, 0.9):.3e}"), "max_rel_err": float(f"{e.max():.3e}"),
                "within_1e-7": int(np.sum(e <= 1e-7))}

    errs, kinks = [], 0
    for k in draw(Ub.ravel(), n_check):
        c, q = divmod(int(k), 4)
        sel = m.cell_rings([c], 6)
        sub, _ = m.submesh(sel, cut)
        loc = int(np.searchsorted(sel, c))
        ls, Us = lam[sel], U[sel]
        g = FVKSolver(sub, sub_bcs, params, ic)
        g.set_field(Us)
        r0 = g.residual_eval()
        sc = max(abs(U[c, q]), 1e-3)

        def gs(h):
            Up = Us.copy()
            Up[loc, q] += h * sc
            g.set_field(Up)
            return float(np.sum(ls * (g.residual_eval() - r0))) / sc
        err, _, _, kink = fd_sweep(gs, Ub[c, q], steps)
        if kink:
            kinks += 1
        else:
            errs.append(err / abs(Ub[c, q]))
    fd = {"state": summary(errs, kinks)}
    owner = np.repeat(np.arange(nc), np.diff(m.cell_node_ptr))
    e = m.node_xy[m.face_nodes[:, 1]] - m.node_xy[m.face_nodes[:, 0]]
    edge = np.hypot(e[:, 0], e[:, 1])
    errs, kinks = [], 0
    for k in draw(Xb.ravel(), n_check):
        n, d = divmod(int(k), 2)
        sel = m.cell_rings(owner[m.cell_nodes == n], 6)
        sub, used = m.submesh(sel, cut)
        loc = int(np.searchsorted(used, n))
        ls, Us = lam[sel], U[sel]
        g = FVKSolver(sub, sub_bcs, params, ic)
        g.set_field(Us)
        r0 = g.residual_eval()
        # steps relative to the node's shortest edge: the wall layers are far
        # thinner than any absolute step
        sc = float(edge[(m.face_nodes == n).any(axis=1)].min())

        def gx(h):
            X = sub.node_xy.copy()
            X[loc, d] += h * sc
            gg = FVKSolver(sub.with_nodes(X), sub_bcs, params, ic)
            gg.set_field(Us)
            return float(np.sum(ls * (gg.residual_eval() - r0))) / sc
        err, _, _, kink = fd_sweep(gx, Xb[n, d], steps)
        if kink:
            kinks += 1
        else:
            errs.append(err / abs(Xb[n, d]))
    fd["nodes"] = summary(errs, kinks)
    fd["method"] = ("per entry, relative to the entry: best of central, Richardson and one-sided differences of "
                    "lambda . R over steps 1e-3..1e-7 (x |U| or x shortest edge) on six-ring sub-meshes (GPU "
                    "residual); kinks = one-sided slopes disagree (limiter/HLLC ties), not scored")
    out["fd_check"] = fd
    return out


def _gen_batch_case(prm):
    from finitevolume_amd import voronoi as V
    return V.naca_voronoi(prm["n_target"], prm["seed"])


def batch_leg(n_cases, max_steps, drop=1e-6, shape=None):
    """Batched airfoil cases: case k a Voronoi mesh of about N_k cells around
    NACA 0012 at its own Mach number and angle of attack (default_rng(1000 +
    k)), config-2 numerics; every case runs until its density residual has
    dropped by `drop` from its first-step value (or max_steps), first all
    together through fvk_batch (one launch per kernel for all active cases,
    steps replayed from a hipGraph), then one after the other, each alone on
    the GPU.  Meshes are generated first, on the host, before the GPU is
    used."""
    from concurrent.futures import ProcessPoolExecutor
    from finitevolume_amd import abi, mesh as M
    from finitevolume_amd import voronoi as V
    prms = [V.batch_case_params(k) for k in range(n_cases)]
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=min(8, host_cores())) as ex:
        meshes = list(ex.map(_gen_batch_case, prms))
    t_gen = time.perf_counter() - t0
    sizes = np.array([m.n_cells for m in meshes])
    print(f"[bench] batch: {n_cases} meshes, {sizes.sum()} cells, generated in {t_gen:.1f}s",
          file=sys.stderr, flush=True)
    from finitevolume_amd.batch import FVKBatch
    from finitevolume_amd.solver import FVKSolver
    gamma = 1.4
    bcs, ics = [], []
    for prm in prms:
        a = math.radians(prm["aoa"])
        rho, pinf = 1.0, 1.0 / gamma
        u, v = prm["mach"] * math.cos(a), prm["mach"] * math.sin(a)
        ff = M.from_primitive(rho, u, v, pinf, gamma)
        bcs.append([abi.make_bc(abi.BC_WALL), abi.make_bc(abi.BC_FARFIELD, ff)])
        ics.append(abi.initial_condition(abi.IC_FREESTREAM, rho=rho, u=u, v=v, p=pinf))
    params = abi.default_params(equation=abi.EQ_EULER, flux_scheme=abi.FLUX_HLLC, cfl=0.8, order=2,
                                limiter=abi.LIMITER_VENKATAKRISHNAN, venkat_k=5.0, rk_stages=3, local_dt=1)
    last = [time.perf_counter()]

    def progress(done, active, now):
        if time.perf_counter() - last[0] > 10.0:
            last[0] = time.perf_counter()
            print(f"[bench] batch: step {done}, {active} cases active, {now:.1f}s", file=sys.stderr, flush=True)

    from finitevolume_amd import autotune
    if shape:  # a given launch shape (variant, block, tile cells) instead of the autotuner's
        params.kernel_variant, params.block_size, params.tile_cells = shape
        tuned = {"variant": shape[0], "block": shape[1], "tile_cells": shape[2], "given": True}
    else:
        tuned = autotune.tune_batch(meshes, bcs, params, ics, use_cache=False,
                                    progress=lambda name, ms: print(f"[bench] batch: autotune {name} {ms:.3f} ms/step",
                                                                    file=sys.stderr, flush=True))
    print(f"[bench] batch: launch shape {tuned}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = FVKBatch(meshes, bcs, params, ics)
    b.set_graph(True)
    b_setup = time.perf_counter() - t0
    rb = b.run_to_convergence(drop, max_steps, chunk=100, progress=progress)
    del b
    print(f"[bench] batch: all cases done in {rb['wall']:.1f}s ({rb['total_steps']} steps)", file=sys.stderr,
          flush=True)
    # the sequential runs take the single-case launch shape the autotuner
    # picks on the median-size case (the fused tile kernel where it wins)
    med = int(np.argsort(sizes)[n_cases // 2])
    seq_params = abi.ParamsC.from_buffer_copy(params)
    seq_tuned = autotune.tune(meshes[med], bcs[med], seq_params, ics[med], use_cache=False)
    print(f"[bench] batch: sequential shape {seq_tuned['variant']}/{seq_tuned['block']}/{seq_tuned['tile_cells']}",
          file=sys.stderr, flush=True)
    seq_steps = np.full(n_cases, -1)
    seq_time = np.zeros(n_cases)
    seq_setup = 0.0
    t_seq0 = time.perf_counter()
    for k in range(n_cases):
        t1 = time.perf_counter()
        g = FVKSolver(meshes[k], bcs[k], seq_params, ics[k])
        seq_setup += time.perf_counter() - t1
        t1 = time.perf_counter()
        first, done = None, 0
        while done < max_steps:
            r = g.step(100)
            if first is None:
                first = r[0, 0]
            hit = np.nonzero(r[:, 0] <= drop * first)[0]
            if hit.size:
                seq_steps[k] = done + int(hit[0]) + 1
                done += 100
                break
            done += 100
        g.synchronize()
        seq_time[k] = time.perf_counter() - t1
        del g
        progress(k + 1, n_cases - k - 1, time.perf_counter() - t_seq0)
    # the sequential wall time is the cases' run times, without their solvers'
    # set-up, as the batched wall time is the batch's run without its set-up
    t_seq = float(seq_time.sum())
    conv = rb["steps"] > 0
    # cell-updates executed: a case runs until the end of the chunk in which it converged
    run = np.where(conv, np.ceil(np.maximum(rb["steps"], 1) / 100.0) * 100, rb["total_steps"])
    executed = float(np.sum(sizes * 3 * run))
    tb = rb["time"][conv]
    return {
        "workload": "config5: %d Voronoi NACA 0012 cases (%d cells), config-2 numerics, each to a %.0e drop of "
                    "its density residual (at most %d steps)" % (n_cases, int(sizes.sum()), drop, max_steps),
        "cells_total": int(sizes.sum()),
        "cells_min_max": [int(sizes.min()), int(sizes.max())],
        "converged": int(conv.sum()),
        "steps_to_converge_p50_max": [int(np.median(rb["steps"][conv])) if conv.any() else None,
                                      int(rb["steps"][conv].max()) if conv.any() else None],
        "batched_wall_s": round(rb["wall"], 2),
        "sequential_wall_s": round(t_seq, 2),
        "setup_s": {"batched": round(b_setup, 2), "sequential": round(seq_setup, 2),
                    "note": "solver set-up (layout, tiles, uploads; the batch's graph capture happens in its "
                            "first chunk), outside both wall times"},
        "batched_over_sequential_speedup": round(t_seq / rb["wall"], 2),
        "batched_time_to_converge_p50_p99_s": [round(float(np.percentile(tb, 50)), 2),
                                               round(float(np.percentile(tb, 99)), 2)] if conv.any() else None,
        "sequential_time_to_converge_p50_p99_s": [round(float(np.percentile(seq_time[conv], 50)), 3),
                                                  round(float(np.percentile(seq_time[conv], 99)), 3)]
        if conv.any() else None,
        "batched_Gcell_updates_per_s": round(executed / rb["wall"] / 1e9, 3),
        "sequential_Gcell_updates_per_s": round(executed / t_seq / 1e9, 3),
        "mesh_generation_s": round(t_gen, 1),
        "hip_graph": True,
        "autotune": tuned,
        "sequential_launch_shape": {k: seq_tuned[k] for k in ("variant", "block", "tile_cells")},
    }


def rans_case(reference_mode=False):
    """Config 3: Spalart-Allmaras RANS around NACA 4412 at Mach 0.2, Re 6e6 per
    chord, 8 degrees: a 1024-point O-grid with 145 quad layers from a 5e-6
    chord wall spacing (growth 1.1) and triangle pairs beyond, 366 layers to 50
    chords (601,088 cells).  HLLC, Green-Gauss gradients with the Jasak face
    gradient, Pr 0.72, Pr_t 0.9, R = 1/gamma (T = 1), nut0 = nut_farfield =
    3 nu; adiabatic no-slip wall.  SSP-RK3 with local dt at CFL 0.5, or (the
    reference mode) forward Euler with a global dt."""
    m = M.naca_ogrid(1024, 366, code="4412", ratio=1.1, first=5e-6, quad_layers=145)
    gamma = 1.4
    mach, re, aoa = 0.2, 6e6, math.radians(8.0)
    rho, pinf = 1.0, 1.0 / gamma
    u, v = mach * math.cos(aoa), mach * math.sin(aoa)
    mu = rho * mach * 1.0 / re
    nut = 3.0 * mu / rho
    ff = M.from_primitive(rho, u, v, pinf, gamma)
    wall = abi.make_bc(abi.BC_WALL)
    wall.is_isothermal_wall = 0
    far = abi.make_bc(abi.BC_FARFIELD, ff)
    for b in (wall, far):
        b.farfield_nut = nut
    ext = dict(rk_stages=1, local_dt=0) if reference_mode else dict(rk_stages=3, local_dt=1)
    params = abi.default_params(equation=abi.EQ_RANS_SA, flux_scheme=abi.FLUX_HLLC,
                                gradient_scheme=abi.GRAD_GREEN_GAUSS, mu=mu, prandtl=0.72, prandtl_t=0.9,
                                gas_constant=1.0 / gamma, cfl=0.5 if not reference_mode else 0.8, **ext)
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=rho, u=u, v=v, p=pinf, initial_nut=nut)
    return m, [wall, far], params, ic


def rans_leg(steps, warmup, cpu_seconds, no_cpu, args=None):
    """Config 3 throughput (SA, RK3 + local dt) and reference mode 3R on the GPU,
    with the reference solver's 3R on the host cores beside it."""
    from finitevolume_amd.solver import FVKSolver
    m, bcs, params, ic = rans_case()
    nc, fpc = m.n_cells, m.n_faces / m.n_cells
    tuned = None
    if args is not None:
        params.kernel_variant, params.block_size, params.tile_cells = args.variant, args.block, args.tile_cells
        if not args.no_tune and not (args.variant or args.block):
            from finitevolume_amd import autotune
            tuned = autotune.tune(m, bcs, params, ic, steps=10, warmup=2, use_cache=False)
            print(f"[bench] rans autotune: {tuned['table_ms_per_step']}", file=sys.stderr, flush=True)
    s = FVKSolver(m, bcs, params, ic)
    wall = time_gpu(s, steps, warmup, per_step=True)
    s.set_kernel_timing(True)
    s.step(steps)
    times = s.kernel_times()
    s.set_kernel_timing(2)  # one event pair around the launches of a step() call
    s.step(steps)
    region = s.kernel_times()["stages"]
    s.set_kernel_timing(False)
    stage_ms = region[0] / region[1]
    b_alg = 149.3 + 48.0 * fpc      # SA stage bytes: 40+40+8+16+8+48 F/N+8, +40 on stages 2-3, +8 on stage 1
    cups = nc * 3 * steps / wall
    out = {"workload": "config3: SA RANS, NACA 4412 hybrid O-grid (601,088 cells), HLLC + Green-Gauss/Jasak, "
                       "SSP-RK3 local dt CFL 0.5", "cells": nc, "faces": m.n_faces,
           "value": round(cups / 1e9, 4), "unit": "Gcell-updates/s", "ms_per_step": round(wall / steps * 1e3, 4),
           "roofline": {"bound": "hbm", "achieved": round(b_alg * nc / (stage_ms * 1e-3) / 1e9, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(b_alg * nc / (stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "b_alg_per_cell_update": round(b_alg, 1), "stage_device_ms": round(stage_ms, 4),
                        "stage_device_ms_events_per_launch": round(sum(t for t, _ in times.values()) / (steps * 3), 4),
                        "kernel_avg_ms": {k: round(tot / cnt, 5) for k, (tot, cnt) in sorted(times.items()) if cnt}},
           "launch_shape": s.launch_shape()}
    if tuned:
        out["autotune_ms_per_step"] = tuned["table_ms_per_step"]
    del s
    m3, bcs3, p3, ic3 = rans_case(reference_mode=True)
    # the library's default launch shape (the gradient, face and cell passes):
    # with forward Euler and a global dt each step needs a dt pass before the
    # fused tile pass, and the two measure slower than the split passes
    s3 = FVKSolver(m3, bcs3, p3, ic3)
    w3 = time_gpu(s3, steps, warmup, per_step=True)
    shape3 = s3.launch_shape()
    del s3
    out["gpu_3R"] = {"value": round(nc * steps / w3 / 1e9, 4), "unit": "Gcell-updates/s",
                     "ms_per_step": round(w3 / steps * 1e3, 4),
                     "roofline_frac": round((112.0 + 48.0 * fpc) * nc * steps / w3 / 1e9 / HBM_PEAK_GBS, 4),
                     "launch_shape": shape3}
    if not no_cpu:
        cores = host_cores()
        try:
            from oracle import pyref
            if pyref.available("fast"):
                r = pyref.RefSolver(m3, bcs3, p3, ic3, flavour="fast")
                r.step(1)
                n2, t2 = timed_steps(r, cpu_seconds)
                ref = nc * n2 / t2
                out["cpu_baseline_reference_3R"] = {
                    "value": round(ref / 1e9, 6), "unit": "Gcell-updates/s", "cores": cores, "kind": "reference",
                    "threading": cpu_threading(),
                    "sample": f"reference RANSTurbulenceSASolver built from its sources with its Release flags "
                              f"({cores} threads), mode 3R on the same mesh, {n2} steps ({t2:.2f} s)"}
                out["gpu_3R_over_reference_3R"] = round(out["gpu_3R"]["value"] * 1e9 / ref, 1)
                del r
        except Exception as e:
            out["cpu_baseline_reference_3R"] = {"value": None, "sample": f"unavailable: {e}"}
    return out


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    return int(omp) if omp and omp.isdigit() else n


def host_sockets():
    """Distinct physical packages among the CPUs this process may run on."""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    pk = set()
    for c in cpus:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id") as f:
                pk.add(f.read().strip())
        except OSError:
            pass
    return len(pk) or None


def cpu_threading():
    return {"cores": host_cores(), "sockets": host_sockets(), "cpu_model": cpu_model(),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "OMP_PLACES": os.environ.get("OMP_PLACES")}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--ni", type=int, default=1024)
    ap.add_argument("--nj", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="wall time of each bounded CPU-baseline sample (whole steps, at least one)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-2r", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-fd-points", type=int, default=32,
                    help="cells and nodes of the adjoint's finite-difference check (0: skip it)")
    ap.add_argument("--batch", action="store_true",
                    help="run only the batched-airfoil workload (64 Voronoi cases to convergence) and print its line")
    ap.add_argument("--batch-cases", type=int, default=64)
    ap.add_argument("--rans", action="store_true",
                    help="run only the RANS SA workload (and its reference mode) and print its line")
    ap.add_argument("--batch-max-steps", type=int, default=20000)
    ap.add_argument("--variant", type=int, default=0, help="Euler kernel variant (0 = library default)")
    ap.add_argument("--block", type=int, default=0, help="Euler workgroup size (0 = library default)")
    ap.add_argument("--tile-cells", type=int, default=0, help="cells per tile of kernel variant 3 (0 = block)")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner (library defaults)")
    ap.add_argument("--probe", action="store_true", help="run a plain 256 MiB device copy first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is a design, not a build, in this repository)", file=sys.stderr)
        sys.exit(2)

    if args.batch:
        line = batch_leg(args.batch_cases, args.batch_max_steps,
                         shape=(args.variant, args.block, args.tile_cells) if args.variant else None)
        line.update({"metric": "batched airfoil cases: wall time to convergence, batched vs sequential",
                     "n_gpus": 1, "dtype": "f64", "data": "synthetic (generated Voronoi meshes, free-stream start)"})
        print(json.dumps(line))
        return

    if args.rans:
        line = rans_leg(args.steps, args.warmup, args.cpu_seconds, args.no_cpu, args)
        line.update({"metric": "FP64 RANS SA residual Gcell-updates/s (1 GPU, % HBM roofline)", "n_gpus": 1,
                     "dtype": "f64", "data": "synthetic (generated NACA 4412 hybrid O-grid, free-stream start)"})
        print(json.dumps(line))
        return

    from finitevolume_amd.solver import FVKSolver, copy_probe

    probe = None
    if args.probe:
        n = 32 * 1024 * 1024  # doubles: 256 MiB read + 256 MiB written per copy
        ms = copy_probe(n, 5)
        probe = {"bytes_per_copy": 2 * 8 * n, "ms_per_copy": round(ms, 4), "GBps": round(2 * 8 * n / ms / 1e6, 1)}
        print(f"[bench] copy probe {probe}", file=sys.stderr, flush=True)

    m, bcs, params, ic = naca_case(args.ni, args.nj, order2=True)
    params.kernel_variant = args.variant
    params.block_size = args.block
    params.tile_cells = args.tile_cells
    nc, nf = m.n_cells, m.n_faces
    sweep, refined = {}, {}
    if not args.no_tune and not (args.variant or args.block):
        # the project's launch-shape autotuner picks the kernel variant and workgroup size
        from finitevolume_amd import autotune
        best = autotune.tune(m, bcs, params, ic, steps=30, use_cache=False)
        sweep = best["table_ms_per_step"]
        refined = best.get("refined_device_ms_per_step", {})
        print(f"[bench] autotune: variant {best['variant']} block {best['block']} ({sweep}; refined {refined})",
              file=sys.stderr, flush=True)
    fpc = nf / nc
    stages = params.rk_stages
    print(f"[bench] mesh {nc} cells {nf} faces; building solver", file=sys.stderr, flush=True)
    s = FVKSolver(m, bcs, params, ic)
    wall = time_gpu(s, args.steps, args.warmup, per_step=True)
    ms_step = wall / args.steps * 1e3
    cups = nc * stages * args.steps / wall
    value = cups / 1e9
    print(f"[bench] {args.steps} steps {wall:.3f}s {ms_step:.3f} ms/step {value:.3f} Gcell-updates/s",
          file=sys.stderr, flush=True)

    # device time per kernel with HIP events on the solver's stream, over a
    # second timed region of the same length
    s.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.step(1)
        s.residual()
    wall_ev = time.perf_counter() - t0
    times = s.kernel_times()
    per_launch = {k: tot / cnt for k, (tot, cnt) in times.items() if cnt}
    stage_ms_launch_events = sum(tot for tot, _ in times.values()) / (args.steps * stages)
    # device time of one RK stage: one event pair around the launches of
    # args.steps steps in one step() call (no event between the kernels,
    # which would add its own time to every launch), / (steps x stages)
    s.set_kernel_timing(2)
    s.step(args.steps)
    region = s.kernel_times()["stages"]
    s.set_kernel_timing(False)
    stage_ms = region[0] / region[1]
    b_alg = b_alg_config2(fpc)
    achieved = b_alg * nc / (stage_ms * 1e-3) / 1e9
    dominant = max(per_launch, key=per_launch.get)
    s_shape = s.launch_shape()
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": "one RK stage = " + " + ".join(sorted(per_launch)),
        "b_alg_per_cell_update": round(b_alg, 1),
        "stage_device_ms": round(stage_ms, 4),
        "stage_timing": "HIP events on the solver's stream around %d steps' launches in one step() call" % args.steps,
        "stage_device_ms_events_per_launch": round(stage_ms_launch_events, 4),
        "kernel_avg_ms": {k: round(v, 5) for k, v in sorted(per_launch.items())},
        "dominant_kernel": dominant,
        "wall_with_events_s": round(wall_ev, 3),
    }
    prof = os.path.join(REPO, "profiles", "traffic_config2.json")
    shape = s_shape if s_shape else {}
    key = f"v{shape.get('kernel_variant', 1)}_b{shape.get('block_size', 0)}_t{shape.get('tile_cells', 0)}"
    roofline["traffic_shape"] = key
    if os.path.exists(prof):
        with open(prof) as f:
            tr = json.load(f).get("shapes", {}).get(key)
        if tr:  # measured for this launch shape on this workload (rocprofv3 FETCH_SIZE, WRITE_SIZE)
            roofline["traffic"] = tr.get("stage_bytes")
            roofline["traffic_per_cell_update"] = tr.get("bytes_per_cell_update")
        else:
            roofline["traffic_note"] = "no counter measurement for this launch shape in profiles/traffic_config2.json"
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated NACA 0012 O-grid, free-stream start)",
        "config": {"workload": "config2: 2nd-order Euler HLLC + LSQ/Venkatakrishnan K=5, SSP-RK3, local dt CFL 0.8, "
                               "NACA 0012 O-grid 1024x512 split to triangles",
                   "cells": nc, "faces": nf, "rk_stages": stages, "ni": args.ni, "nj": args.nj,
                   "kernel_variant": params.kernel_variant, "block_size": params.block_size,
                   "tile_cells": params.tile_cells, "launch_shape": s.launch_shape()},
        "roofline": roofline,
    }
    if sweep:
        out["autotune_ms_per_step"] = sweep
        out["autotune_refined_device_ms_per_step"] = refined
    if probe:
        out["copy_probe"] = probe
    if not args.no_adjoint:
        done = 3 * args.steps + args.warmup
        out["adjoint"] = adjoint_leg(s, m, bcs, params, ic, done, n_check=args.adjoint_fd_points)
        print(f"[bench] adjoint {out['adjoint']['vjp_device_ms']} ms vs residual "
              f"{out['adjoint']['primal_residual_device_ms']} ms", file=sys.stderr, flush=True)
    del s

    # reference mode 2R on the GPU: first-order HLLC, forward Euler, global dt
    if not args.no_2r:
        m2, bcs2, p2, ic2 = naca_case(args.ni, args.nj, order2=False)
        tuned2 = None
        if not args.no_tune and not (args.variant or args.block):
            from finitevolume_amd import autotune
            tuned2 = autotune.tune(m2, bcs2, p2, ic2, steps=50, use_cache=False)
        s2 = FVKSolver(m2, bcs2, p2, ic2)
        w2 = time_gpu(s2, args.steps, args.warmup, per_step=True)
        # device time of a step (HIP events on the solver's stream), as for config 2
        s2.set_kernel_timing(True)
        for _ in range(args.steps):
            s2.step(1)
            s2.residual()
        times2 = s2.kernel_times()
        s2.set_kernel_timing(2)  # one event pair around a step() call of args.steps steps
        s2.step(args.steps)
        region2 = s2.kernel_times()["stages"]
        s2.set_kernel_timing(False)
        step2_ms = region2[0] / region2[1]
        out["gpu_2R"] = {"value": round(nc * args.steps / w2 / 1e9, 4), "unit": "Gcell-updates/s",
                         "ms_per_step": round(w2 / args.steps * 1e3, 4),
                         "step_device_ms": round(step2_ms, 4),
                         "step_device_ms_events_per_launch": round(sum(t for t, _ in times2.values()) / args.steps, 4),
                         "kernel_avg_ms": {k: round(tot / cnt, 5) for k, (tot, cnt) in sorted(times2.items()) if cnt},
                         "roofline_frac": round(b_alg_2r(fpc) * nc / (step2_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "launch_shape": s2.launch_shape()}
        if tuned2:
            out["gpu_2R"]["autotune_ms_per_step"] = tuned2["table_ms_per_step"]
        del s2

    if not args.no_cpu:
        cores = host_cores()
        try:
            from oracle import pyoracle
            o = pyoracle.OracleSolver(m, bcs, params, ic, flavour="fast")
            o.step(1)
            n_cpu, tc = timed_steps(o, args.cpu_seconds)
            cpu_cups = nc * stages * n_cpu / tc
            out["cpu_baseline"] = {"value": round(cpu_cups / 1e9, 6), "unit": "Gcell-updates/s", "cores": cores,
                                   "kind": "port", "threading": cpu_threading(),
                                   "sample": f"oracle C restatement built with the reference's Release flags "
                                             f"(-O2 -ffast-math -flto -fopenmp; {cores} threads, {cpu_model()}) on "
                                             f"the same config-2 mesh and numerics, {n_cpu} RK3 steps after "
                                             f"1 warm-up step ({tc:.2f} s)"}
            out["gpu_over_cpu"] = round(cups / cpu_cups, 1)
            del o
        except Exception as e:  # the baseline is reported, not required
            out["cpu_baseline"] = {"value": None, "unit": "Gcell-updates/s", "cores": cores, "kind": "port",
                                   "sample": f"unavailable: {e}"}
        try:
            from oracle import pyref
            if pyref.available("fast") and "gpu_2R" in out:
                m2, bcs2, p2, ic2 = naca_case(args.ni, args.nj, order2=False)
                r = pyref.RefSolver(m2, bcs2, p2, ic2, flavour="fast")
                r.step(2)
                n2, tr2 = timed_steps(r, args.cpu_seconds)
                ref_cups = nc * n2 / tr2
                out["cpu_baseline_reference_2R"] = {
                    "value": round(ref_cups / 1e9, 6), "unit": "Gcell-updates/s", "cores": cores, "kind": "reference",
                    "threading": cpu_threading(),
                    "sample": f"reference solver built from its own sources with its Release flags "
                              f"(-O2 -ffast-math -flto -fopenmp, {cores} threads), reference mode 2R "
                              f"(first-order HLLC, forward Euler, global dt) on the same mesh, {n2} steps ({tr2:.2f} s)"}
                out["gpu_2R_over_reference_2R"] = round(out["gpu_2R"]["value"] * 1e9 / ref_cups, 1)
                del r
        except Exception as e:
            out["cpu_baseline_reference_2R"] = {"value": None, "sample": f"unavailable: {e}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
