"""Benchmark of the body-fitted IGA Navier-Stokes time step on one MI355X.

Default workload (`tree_c4`): the seeded vessel tree -- binary, depth 4
(15 segments), root radius 3 mm, Murray radii (r^3 conserved), segment
lengths U[8, 12] x radius, branching angles U[25, 45] deg, quadratic C1
splines at 8.1M velocity+pressure DOFs (1.68M elements) -- with pulsatile
inflow Q(t) = 5 (1 + 0.5 sin 2 pi t) ml/s at 1 Hz, zero-traction outlets,
backward Euler with dt = 1e-3 s, rho = 1060 kg/m^3, mu = 3.5e-3 Pa s.  Every
Newton linear solve goes to a true relative residual of 1e-8 and Newton to a
relative nonlinear residual of 1e-12 (--newton-rtol).  Two methods run the same physical
steps from the same initial state (rest):
  * plain: FP64 FGMRES + FP64 block AMG, Newton from the previous step;
  * mixed + learned: FP32 SpMV/AMG inside FP64 FGMRES with iterative
    refinement, Newton from the online-trained MLP prediction over the
    increment extrapolation (residual-checked, previous step as fallback).
`value` is the mixed+learned wall time per time step (lower is better); the
plain FP64 time, the speedup, the assembly kernel's FP64-MFMA roofline and the
SpMV HBM roofline sit beside it.

`--workload ybif_c1` runs the symmetric Y-bifurcation (0.6M DOFs; steady
Q = 5 ml/s by default) with its secondary measurements: a pulsatile
comparison (Q(t) = 5 (1 + 0.6 sin 2 pi t + 0.2 sin 4 pi t) ml/s from rest,
dt = 5e-3 s, 100 untimed + 10 timed steps), the two-species transport and the configured three-field
transport on the same mesh.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, AMD public spec
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector / matrix, AMD public spec


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pulsatile(t):
    """Inflow scale of the learned-warm-start workload: Q(t)/Q0 = 1 + 0.6 sin 2 pi t + 0.2 sin 4 pi t."""
    return 1.0 + 0.6 * math.sin(2 * math.pi * t) + 0.2 * math.sin(4 * math.pi * t)


def tree_pulse(t):
    """Inflow scale of the vessel-tree workload: Q(t)/Q0 = 1 + 0.5 sin 2 pi t (1 Hz)."""
    return 1.0 + 0.5 * math.sin(2 * math.pi * t)


WAVEFORMS = {"pulsatile": (pulsatile, "pulsatile Q(t)=5(1+0.6 sin 2 pi t+0.2 sin 4 pi t) ml/s"),
             "pulse": (tree_pulse, "pulsatile Q(t)=5(1+0.5 sin 2 pi t) ml/s at 1 Hz"),
             "steady": (None, "steady Q=5 ml/s")}
DESCRIPTIONS = {"tree_c4": "seeded vessel tree (binary, depth 4, 15 segments, root radius 3 mm, Murray radii, "
                           "L/r ~ U[8,12], angles ~ U[25,45] deg, seed 20261202)",
                "ybif_c1": "symmetric Y-bifurcation (parent 2 mm, daughters 1.6 mm at +-35 deg, 20 mm branches)",
                "tube_c2": "straight tube R = 1 mm, L = 64 mm (O-grid 20 control hexes per section x 256 axial, "
                           "refined 5.5-fold)"}


def run_method(dm, case, bc, faces, rho, mu, dt, steps, warmup, precision, warm, tol, train_steps=20,
               waveform=None, smoother=None, start=None, cheb_degree=2):
    from tubularflowiga_amd import boundary, device
    vmask, pmask, vt, pt = bc
    flow = device.FlowSolver(dm, density=rho, viscosity=mu, dt=dt, nonlinear_rtol=tol["rtol"],
                             nonlinear_atol=tol["atol"], mass_rtol=tol["mass"], linear_rtol=1e-8,
                             preconditioner=device.TFI_PC_AMG, precision=precision, warm_start=warm,
                             restart=100, linear_maximum_iterations=4000, verbose=1,
                             amg_chebyshev_degree=cheb_degree,
                             amg_smoother=device.TFI_SMOOTH_AUTO if smoother is None else smoother)
    if warm == device.TFI_WARM_LEARNED:
        flow.configure_warm_start(hidden=128, train_steps=train_steps, learning_rate=1e-3, min_snapshots=2)
    scale0 = waveform(0.0) if waveform else 1.0
    if start is None:
        flow.set_state(boundary.initial_state(vmask, pmask, vt * scale0, pt))
    else:
        x0 = scale0 * np.array(start, dtype=np.float64).reshape(-1, 4)
        x0[vmask == 1, :3] = vt[vmask == 1] * scale0
        x0[pmask == 1, 3] = pt[pmask == 1]
        flow.set_state(x0.ravel())
    flow.set_boundary(vmask, pmask, vt * scale0, pt)
    flow.set_boundary_faces(*faces)
    ctx = dm.ctx
    records = []
    timed = []
    for step in range(warmup + steps):
        is_timed = step >= warmup
        if is_timed and step == warmup:
            ctx.timers(enable=True, reset=True)
        t0 = time.perf_counter()
        if waveform:
            t = (step + 1) * dt
            flow.begin_step(np.array([math.sin(2 * math.pi * t), math.cos(2 * math.pi * t),
                                      math.sin(4 * math.pi * t), math.cos(4 * math.pi * t)]))
            flow.set_boundary(vmask, pmask, vt * waveform(t), pt)
        else:
            flow.begin_step()
        rep = flow.newton()
        flow.commit_step()
        t1 = time.perf_counter()
        rec = {"step": step + 1, "wall_s": t1 - t0, "newton": rep["newton_iterations"],
               "linear": rep["linear_iterations"], "assembly_s": rep["assembly_s"], "linear_s": rep["linear_s"],
               "warm_used": rep["warm_start_used"], "warm_ratio": rep["warm_start_initial_ratio"],
               "chord": rep["chord_steps"], "blend": rep["warm_start_blend"],
               "base_ratio": rep["warm_start_base_ratio"],
               "final_residual_rel": rep["final_residual"] / max(rep["tolerance"], 1e-300) * tol["rtol"]}
        records.append(rec)
        log(f"  step {step + 1} {'timed' if is_timed else 'warmup'} wall={rec['wall_s']*1e3:.1f} ms "
            f"newton={rec['newton']} linear={rec['linear']} asm={rec['assembly_s']*1e3:.1f} ms "
            f"lin={rec['linear_s']*1e3:.1f} ms warm={rec['warm_used']} ratio={rec['warm_ratio']:.3g} "
            f"chord={rec['chord']} blend={rec['blend']:.3f} final_rel={rec['final_residual_rel']:.2e}")
        if is_timed:
            timed.append(rec)
    timers = ctx.timers(enable=False)
    state = flow.state()
    warm_stats = flow.warm_stats()
    flow.close()
    return timed, records, timers, state, warm_stats


def measured_traffic(workload):
    """HBM bytes per launch of the roofline kernels from the committed rocprofv3
    counter summary of the workload (FETCH_SIZE and WRITE_SIZE passes, scaled by
    the factors a plain copy kernel of known size gives); {} when there is none."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.txt")
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        if line.startswith("traffic "):
            fields = dict(f.split("=", 1) for f in line.split()[1:])
            out[fields["kernel"].replace("_", " ")] = int(float(fields["bytes_per_launch"]))
    return out


def network_gain(timed):
    """What the trained network adds to its base prediction, over the timed
    steps where both were evaluated: the geometric mean of the step's initial
    residual (relative to the previous-step start) of the learned guess and of
    the base prediction alone, and their ratio (< 1: the network helps)."""
    pairs = [(r["warm_ratio"], r["base_ratio"]) for r in timed if r["warm_ratio"] > 0 and r["base_ratio"] > 0]
    if not pairs:
        return None
    g = float(np.exp(np.mean([math.log(a) for a, _ in pairs])))
    b = float(np.exp(np.mean([math.log(c) for _, c in pairs])))
    return {"steps": len(pairs), "learned_initial_ratio": g, "base_initial_ratio": b, "learned_over_base": g / b}


def kernel_breakdown(timers, steps):
    """GPU milliseconds per time step of the timed kernels (HIP events)."""
    steps = max(steps, 1)
    return {"assembly_tangent": round(timers["assembly_ms"] / steps, 3),
            "assembly_residual_only": round(timers["residual_ms"] / steps, 3),
            "spmv_fp64": round(timers["spmv_ms"] / steps, 3), "spmv_fp32": round(timers["spmv32_ms"] / steps, 3)}


def method_summary(timed):
    return {"ms_per_step": round(1e3 * float(np.mean([r["wall_s"] for r in timed])), 3),
            "assembly_ms_per_step": round(1e3 * float(np.mean([r["assembly_s"] for r in timed])), 3),
            "linear_ms_per_step": round(1e3 * float(np.mean([r["linear_s"] for r in timed])), 3),
            "newton_per_step": float(np.mean([r["newton"] for r in timed])),
            "krylov_per_step": float(np.mean([r["linear"] for r in timed])),
            "chord_steps": int(sum(r["chord"] for r in timed))}


def rel_l2(a, b):
    a, b = a.reshape(-1, 4), b.reshape(-1, 4)
    return {"velocity": float(np.linalg.norm(a[:, :3] - b[:, :3]) / np.linalg.norm(b[:, :3])),
            "pressure": float(np.linalg.norm(a[:, 3] - b[:, 3]) / max(np.linalg.norm(b[:, 3]), 1e-300))}


def pulsatile_comparison(dm, case, bc, faces, rho, mu, dt, tol, train_steps, warmup=100, steps=10):
    """The learned-warm-start workload on the same mesh: pulsatile inflow
    from rest at dt = 5e-3 s, the first `warmup` steps untimed (they also
    train the predictor), then `steps` timed steps of both methods."""
    from tubularflowiga_amd import device
    log(f"pulsatile inflow: plain FP64, {warmup} + {steps} steps")
    plain, plain_all, _, plain_state, _ = run_method(
        dm, case, bc, faces, rho, mu, dt, steps, warmup, device.TFI_FP64, device.TFI_WARM_PREVIOUS, tol,
        waveform=pulsatile)
    log(f"pulsatile inflow: mixed precision + learned warm start, {warmup} + {steps} steps")
    fast, fast_all, _, fast_state, warm_stats = run_method(
        dm, case, bc, faces, rho, mu, dt, steps, warmup, device.TFI_MIXED, device.TFI_WARM_LEARNED, tol,
        train_steps, waveform=pulsatile)
    p, f = method_summary(plain), method_summary(fast)
    return {"workload": f"same mesh, pulsatile Q(t)=5(1+0.6 sin 2 pi t+0.2 sin 4 pi t) ml/s from rest, backward Euler "
                        f"dt={dt}",
            "warmup": warmup, "steps": steps, "newton_rtol": tol["rtol"], "fp64_ms_per_step": p["ms_per_step"],
            "mixed_learned_ms_per_step": f["ms_per_step"],
            "speedup_vs_fp64": round(p["ms_per_step"] / f["ms_per_step"], 3), "plain_fp64": p, "mixed_learned": f,
            "whole_run_ms_per_step": {
                "plain_fp64": round(1e3 * float(np.mean([r["wall_s"] for r in plain_all])), 3),
                "mixed_learned": round(1e3 * float(np.mean([r["wall_s"] for r in fast_all])), 3)},
            "rel_l2_mixed_vs_fp64": rel_l2(fast_state, plain_state), "warm_start": warm_stats}


def transport_measurement(dm, case, state, steps=20, warmup=2, cpu=False):
    """Two-species neuron transport on the same mesh, advected by the
    converged flow (`iga_hip transport`): assembly of L and P once, then
    `steps` timed backward-Euler steps, and the HBM roofline of the 2x2 SpMV
    (FP64 values and column indices once, x and y once per launch), timed over
    50 back-to-back launches (an event pair around each short launch of the
    solve adds a few microseconds; that per-launch figure is reported beside
    it).  The steps use the reference's preconditioner (the block-Jacobi scaling
    alone); three more with the AMG give its iteration count and time.  With
    `cpu`, the CPU
    oracle is timed on a sample (element matrices of 2000 elements, scipy CSR
    products of the full 2x2 system) and scaled to the assembly and to one
    step of that block-Jacobi solve."""
    from tubularflowiga_amd import device
    ctx = dm.ctx
    velocity = np.asarray(state).reshape(-1, 4)[:, :3]
    constrained = (np.asarray(case.point_labels) == 1).astype(np.int32)
    params = dict(diffusion=1e-3, vplus=1.0, kplus=0.5, kminus=0.1, detach_plus=0.3, dt=1e-3, artificial_diffusion=0.1)
    ctx.timers(enable=True, reset=True)
    tr = device.Transport(dm, params, velocity, constrained)
    try:
        assembly_ms = ctx.timers()["transport_assembly_ms"]
        tr.set_boundary(np.where(constrained == 1, 1.0, 0.0), np.where(constrained == 1, 0.5, 0.0))
        tr.apply_boundary()
        for k in range(warmup):
            tr.step(zero_guess=(k == 0))
        ctx.timers(reset=True)
        t0 = time.perf_counter()
        its = [tr.step()[0] for _ in range(steps)]
        wall = time.perf_counter() - t0
        tm = ctx.timers(enable=False)
        spmv_ms = tr.bench_spmv(50)
        tr.set_preconditioner(device.TFI_PC_AMG)
        tr.step()  # AMG set-up
        t0 = time.perf_counter()
        its_amg = [tr.step()[0] for _ in range(3)]
        wall_amg = time.perf_counter() - t0
    finally:
        tr.close()
    spmv_bytes = tm["spmv2_bytes"] / max(tm["spmv2_calls"], 1)
    gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    extra = {"amg": {"ms_per_step": round(1e3 * wall_amg / 3, 3), "krylov_per_step": float(np.mean(its_amg))}}
    if cpu:
        extra["cpu_baseline"] = transport_cpu_baseline(dm, case, velocity, params, float(np.mean(its)))
    out = {"workload": "two-species neuron transport on the same mesh, advected by the converged steady flow, "
                        "dt=0.001, FGMRES(50) on the block-Jacobi-scaled system to 1e-8",
            "steps": steps, "ms_per_step": round(1e3 * wall / steps, 3),
            "krylov_per_step": float(np.mean(its)), "assembly_ms": round(assembly_ms, 3),
            "spmv_roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                              "kernel": "Bsr2SpmvKernel (2x2 FP64 values)",
                              "avg_ms": round(spmv_ms, 4), "algorithmic_bytes": int(spmv_bytes),
                              "in_solve_avg_ms": round(tm["spmv2_ms"] / max(tm["spmv2_calls"], 1), 4)}}
    out.update(extra)
    return out


def solve_measurement(dm, case, state, steps=10, warmup=2):
    """Configured three-field transport (`iga_hip solve`) on the same mesh,
    advected by the converged flow: oxygen (consistent SUPG, advective-diffusive
    tau, a volume source, a Robin wall), a carrier (SUPG-test diffusion), a
    tracer (a wall flux), coupled by linear terms; assembly once, then `steps`
    timed backward-Euler steps from the previous state to 1e-8 of the initial
    residual (the reference's GMRES test), and the HBM roofline of the 3x3 SpMV
    timed over 50 back-to-back launches."""
    from tubularflowiga_amd import device
    ctx = dm.ctx
    mesh = case.mesh
    velocity = np.asarray(state).reshape(-1, 4)[:, :3]
    terms = [(0, 0, 0, 1.0), (1, 0, 0, 1e-3), (2, 0, 0, 1.0), (3, 0, 1, -0.2), (4, 0, 0, 0.3),
             (0, 1, 1, 1.0), (1, 1, 1, 5e-4), (2, 1, 1, 1.0), (3, 1, 0, -0.1),
             (0, 2, 2, 2.0), (1, 2, 2, 1e-3)]
    labels = np.asarray(mesh.boundary_labels).reshape(-1, 6)
    element, face = np.nonzero(labels == 0)
    n = len(element)
    surfaces = dict(element=element.astype(np.int32), face=face.astype(np.int32),
                    flux=np.tile([0.0, 0.0, 0.4], (n, 1)), robin=np.tile([0.7, 0.0, 0.0], (n, 1)),
                    robin_source=np.tile([0.7 * 0.25, 0.0, 0.0], (n, 1)))
    inlet = np.asarray(case.point_labels) == 1
    constrained = np.zeros((mesh.nodes, 3), np.int32)
    values = np.zeros((mesh.nodes, 3))
    constrained[inlet, :2] = 1
    values[inlet, 0], values[inlet, 1] = 1.0, 0.25
    ctx.timers(enable=True, reset=True)
    t0 = time.perf_counter()
    tr = device.ConfiguredTransport(dm, 3, terms, [3, 2, 0], [1e-3, 0.0, 0.0], 1e-3, velocity, constrained, surfaces)
    create_s = time.perf_counter() - t0
    try:
        assembly_ms = ctx.timers(enable=False)["transport_assembly_ms"]
        tr.set_boundary(values)
        tr.set_state(np.where(constrained == 1, values, 0.0))
        for k in range(warmup):
            tr.step(zero_guess=(k == 0))
        t0 = time.perf_counter()
        its = [tr.step()[0] for _ in range(steps)]
        wall = time.perf_counter() - t0
        spmv_ms = tr.bench_spmv(50)
        tr.set_preconditioner(device.TFI_PC_AMG)
        tr.step()  # AMG set-up
        t0 = time.perf_counter()
        its_amg = [tr.step()[0] for _ in range(3)]
        wall_amg = time.perf_counter() - t0
    finally:
        tr.close()
    nnzb = dm.nnzb
    spmv_bytes = nnzb * (9 * 8 + 4) + 4 * (mesh.nodes + 1) + 2 * 3 * mesh.nodes * 8
    gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    return {"workload": "configured three-field transport (iga_hip solve) on the same mesh, advected by the "
                        "converged steady flow, dt=0.001, FGMRES(50) on the block-Jacobi-scaled system to 1e-8 "
                        "of the initial residual",
            "steps": steps, "ms_per_step": round(1e3 * wall / steps, 3), "krylov_per_step": float(np.mean(its)),
            "assembly_ms": round(assembly_ms, 3), "create_s": round(create_s, 3), "wall_faces": int(n),
            "spmv_roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                              "kernel": "SolveSpmvKernel<3> (3x3 FP64 values)", "avg_ms": round(spmv_ms, 4),
                              "algorithmic_bytes": int(spmv_bytes)},
            "amg": {"ms_per_step": round(1e3 * wall_amg / 3, 3), "krylov_per_step": float(np.mean(its_amg))}}


def transport_cpu_baseline(dm, case, velocity, params, iterations, n_el=2000, products=20):
    """The oracle's transport element (C++, one element per call) on a sample
    of elements and scipy CSR products of a 2x2 system on the full pattern;
    one step = `iterations` products (block Jacobi, as the reference solves)."""
    import scipy.sparse as sp
    from oracle import transport as T
    mesh = case.mesh
    n_el = min(n_el, mesh.elements)
    t0 = time.perf_counter()
    for e in range(n_el):
        conn = mesh.element_connectivity(e)
        T.element(mesh.element_bezier(e), mesh.dense_extraction(e), velocity[conn], params)
    per_element = (time.perf_counter() - t0) / n_el
    ro, cols, _ = dm.pattern()
    A = sp.bsr_matrix((np.full((len(cols), 2, 2), 1e-3), cols, ro), shape=(2 * mesh.nodes,) * 2).tocsr()
    x = np.ones(2 * mesh.nodes)
    A @ x
    t0 = time.perf_counter()
    for _ in range(products):
        A @ x
    per_product = (time.perf_counter() - t0) / products
    return {"value": round(1e3 * per_product * iterations, 3), "unit": "ms/step", "cores": 1, "kind": "port",
            "assembly_ms": round(1e3 * per_element * mesh.elements, 1),
            "sample": f"oracle transport element on {n_el} elements ({1e3 * per_element:.3f} ms each) and "
                      f"{products} scipy CSR products of the 2x2 system ({1e3 * per_product:.2f} ms each); one step "
                      f"= {iterations:.0f} products (block-Jacobi iterations of the GPU run)"}


def cpu_baseline(case, dm, rho, mu, dt, plain_timed, budget_s=20.0):
    """CPU oracle (FP64 C++ restatement, OpenMP) on a bounded sample: element
    assembly of a slice of the mesh and FP64 SpMVs on the full pattern, scaled
    to one time step with the plain GPU run's Newton / Krylov counts and the
    AMG V-cycle's SpMV-equivalent cost."""
    from oracle import oracle as O
    mesh = case.mesh
    state = np.zeros(4 * mesh.nodes)
    n_el = min(mesh.elements, 2000)
    t_el = O.time_elements(mesh, state, rho, mu, dt, 0, n_el)
    per_element = t_el / n_el
    pattern = dm.pattern()
    vals = np.full((len(pattern[1]), 4, 4), 1e-3)
    x = np.ones(4 * mesh.nodes)
    O.bsr_spmv(pattern, vals, x)
    reps = 0
    t0 = time.perf_counter()
    while reps < 20 and time.perf_counter() - t0 < budget_s * 0.5:
        O.bsr_spmv(pattern, vals, x)
        reps += 1
    t_spmv = (time.perf_counter() - t0) / max(reps, 1)
    newton = np.mean([r["newton"] for r in plain_timed])
    linear = np.mean([r["linear"] for r in plain_timed])
    assemblies = newton + 1
    spmv_equivalents_per_iteration = 1.0 + 5.5   # outer SpMV + V-cycle (GS pre/post, residual, coarse)
    step_s = assemblies * per_element * mesh.elements + linear * spmv_equivalents_per_iteration * t_spmv
    return {"value": step_s * 1e3, "unit": "ms/step", "cores": O.num_threads(), "kind": "port",
            "sample": (f"FP64 C++ oracle: {n_el} element systems ({t_el:.2f} s) and {reps} full-pattern BSR SpMVs "
                       f"({t_spmv*1e3:.1f} ms each); one step = {assemblies:.1f} assemblies x {mesh.elements} "
                       f"elements + {linear:.0f} Krylov iterations x {spmv_equivalents_per_iteration} "
                       f"SpMV-equivalents (counts from the plain GPU run)")}


def cpu_model():
    """CPU model name and physical core count of this host (from /proc/cpuinfo)."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "model name":
                model = val
            elif key == "physical id":
                phys = val
            elif key == "core id":
                core = val
                cores.add((phys, core))
    except OSError:
        pass
    return model, len(cores) or (os.cpu_count() or 1)


def cpu_full_step(ctx, rho_unused=None):
    """A complete plain FP64 backward-Euler Newton step of the CPU oracle (OpenMP)
    on the coarse version of the tree (tree_small, the same seeded geometry),
    from rest to a relative residual of 1e-8 with FGMRES + block Jacobi, beside
    the GPU's plain step on the same case."""
    from oracle import oracle as O
    from tubularflowiga_amd import boundary, casegen, device
    case = casegen.workload("tree_small")
    rho, mu = casegen.flow_parameters(case)
    vm, pm, vt, pt = boundary.resolve_flow_boundaries(case)
    s0 = boundary.initial_state(vm, pm, vt, pt)
    pattern = O.block_pattern(case.mesh)
    t0 = time.perf_counter()
    _, newton, linear, _ = O.ns_newton_step(case.mesh, pattern, s0, s0.copy(), rho, mu, 1e-3, vm, pm, vt, pt,
                                            max_newton=30, rtol=1e-8, atol=1e-16, linear_rtol=1e-8, tangent="exact")
    cpu_s = time.perf_counter() - t0
    dm = device.DeviceMesh(ctx, case.mesh)
    flow = device.FlowSolver(dm, density=rho, viscosity=mu, dt=1e-3, nonlinear_rtol=1e-8, nonlinear_atol=1e-16,
                             mass_rtol=1e30, linear_rtol=1e-8, maximum_newton=30)
    try:
        flow.set_state(s0)
        flow.set_boundary(vm, pm, vt, pt)
        flow.set_boundary_faces(*boundary.boundary_faces(case.mesh))
        flow.begin_step()
        t0 = time.perf_counter()
        rep = flow.newton()
        gpu_s = time.perf_counter() - t0
    finally:
        flow.close()
        dm.close()
    return {"case": f"tree_small ({4 * case.mesh.nodes} DOFs, {case.mesh.elements} elements), first step from rest, "
                    "relative residual 1e-8",
            "cpu_ms": round(1e3 * cpu_s, 1), "cpu_newton": newton, "cpu_krylov": linear,
            "cpu_solver": "FGMRES(200) + block Jacobi, exact tangent (oracle)",
            "gpu_plain_fp64_ms": round(1e3 * gpu_s, 1), "gpu_newton": rep["newton_iterations"],
            "gpu_krylov": rep["linear_iterations"]}


def measured_peaks(ctx):
    """HBM (2 GiB buffers: 16-byte-vector copy, read + write bytes, and a
    read-only sum) and FP64 matrix-core (probe kernel) rates measured on this
    GPU.  The SpMV reads ~30x more than it writes, so its measured reference is
    the read rate."""
    copy_gbs, read_gbs = ctx.bench_stream(1 << 28, repetitions=10)
    return {"hbm_copy_gbs": round(copy_gbs, 1), "hbm_read_gbs": round(read_gbs, 1),
            "fp64_mfma_tflops": round(ctx.bench_mfma_fp64(3), 2)}


def spmv_only(dm, workload, repetitions=20):
    """SpMV launches on the workload's pattern (FP64 and FP32 values) and a
    plain 2 GiB device copy (CopyKernel), whose HBM counters calibrate the SpMV ones."""
    out = {"workload": workload, "nnzb": int(dm.nnzb), "nodes": int(dm.nodes)}
    for name, prec, vb in (("fp64", 0, 8), ("fp32", 1, 4)):
        ms = dm.bench_spmv(precision=prec, repetitions=repetitions)
        b = dm.spmv_bytes(vb)
        out[name] = {"ms": ms, "algorithmic_bytes": b, "gbs": b / (ms * 1e-3) / 1e9}
    n = 1 << 28                                   # 2 GiB of float64
    ms = dm.ctx.bench_copy(n, repetitions=3)
    out["copy"] = {"kernel": "CopyKernel", "bytes_read": 8 * n, "bytes_written": 8 * n, "ms": ms,
                   "gbs": 16 * n / (ms * 1e-3) / 1e9}
    print(json.dumps(out), flush=True)
    return 0


def assembly_only(dm, workload, repetitions=3):
    """Device-timed assemblies of the workload (synthetic state): exact and
    legacy tangent with residual, and the residual alone."""
    out = {"workload": workload, "elements": int(dm.host.elements)}
    for name, tangent, jac in (("exact", "exact", True), ("legacy", "legacy", True), ("residual", "exact", False)):
        ms = dm.bench_assembly(tangent, jac, repetitions)
        fl = dm.assembly_flops(tangent, jac)
        out[name] = {"ms": ms, "algorithmic_flops": fl, "tflops": fl / (ms * 1e-3) / 1e12}
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="tree_c4", help="tree_c4 (default), ybif_c1, ...")
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--newton-rtol", type=float, default=1e-12,
                    help="relative nonlinear residual of every step (both methods)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-steps", type=int, default=20, help="Adam steps of the predictor per time step")
    ap.add_argument("--inflow", choices=sorted(WAVEFORMS), default=None,
                    help="pulse: Q0 (1 + 0.5 sin 2 pi t) (tree default); steady: Q0 (Y default); "
                         "pulsatile: Q0 (1 + 0.6 sin 2 pi t + 0.2 sin 4 pi t)")
    ap.add_argument("--extras", action="store_true",
                    help="secondary measurements on the mesh (pulsatile comparison, transport, configured "
                         "transport); on by default for ybif_c1")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--pulsatile-steps", type=int, default=10,
                    help="timed steps of the pulsatile comparison after its 100 untimed ones (300: the full "
                         "400-step workload)")
    ap.add_argument("--smoother", choices=["auto", "gauss-seidel", "chebyshev", "jacobi"], default="auto",
                    help="AMG smoother of both methods (auto: Chebyshev for transient steps, Gauss-Seidel for steady)")
    ap.add_argument("--cheb-degree", type=int, default=None,
                    help="Chebyshev steps before and after each coarse correction (both methods); default 1 on the "
                         "vessel tree, 2 elsewhere")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV kernel on the workload's pattern plus a 2 GiB copy (for counter passes)")
    ap.add_argument("--assembly-only", action="store_true",
                    help="only the element assembly (full legacy / exact, residual-only) on the workload")
    ap.add_argument("--transport-only", action="store_true",
                    help="only the transport assembly and 5 steps, advected by the inlet profile (for counter passes)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: the time step runs on one GPU only (asked for {a.gpus}); "
              "the multi-GPU decomposition is not implemented", file=sys.stderr)
        return 2
    from tubularflowiga_amd import boundary, casegen, device

    inflow = a.inflow or ("pulse" if a.workload.startswith(("tree", "tube")) else "steady")
    extras = (a.extras or a.workload == "ybif_c1") and not a.no_extras
    t0 = time.perf_counter()
    case = casegen.workload(a.workload)
    rho, mu = casegen.flow_parameters(case)
    log(f"case {a.workload}: elements={case.mesh.elements} nodes={case.mesh.nodes} dofs={4*case.mesh.nodes} "
        f"generated in {time.perf_counter()-t0:.1f} s")
    ctx = device.Context(0)
    log(ctx.device_info())
    t0 = time.perf_counter()
    dm = device.DeviceMesh(ctx, case.mesh)
    log(f"mesh upload + pattern + colouring + geometry: {time.perf_counter()-t0:.2f} s; {dm.stats}")
    if a.spmv_only:
        return spmv_only(dm, a.workload)
    if a.assembly_only:
        return assembly_only(dm, a.workload)
    if a.transport_only:
        state = np.zeros((case.mesh.nodes, 4))
        state[:, :3] = case.inlet_velocity
        print(json.dumps(transport_measurement(dm, case, state, steps=5, warmup=1)), flush=True)
        print(json.dumps(solve_measurement(dm, case, state, steps=5, warmup=1)), flush=True)
        return 0
    bc = boundary.resolve_flow_boundaries(case)
    faces = boundary.boundary_faces(case.mesh)
    # Newton to a relative nonlinear residual of 1e-12 by default (both
    # methods): on the tree two states that each meet 1e-10 differ by up to
    # ~4e-8 in pressure (relative L2), at 1e-12 by less than 1e-10; every
    # Newton linear solve to a true relative residual of 1e-8
    # (the absolute floor atol * sqrt(4N) stays far below rtol ||r0|| on the 8M-DOF tree)
    tol = {"rtol": a.newton_rtol, "atol": 1e-18, "mass": 1e-6}
    wf, inflow_text = WAVEFORMS[inflow]

    # Tubes start from developed Poiseuille flow: from rest, the first Newton
    # update of the refined tube's impulsive start overshoots the inflow
    # ~24-fold next to the inlet (the oracle's direct solve on a 2 mm piece of
    # the same mesh gives the same), and the shortened Newton steps that
    # follow leave tangents the AMG-preconditioned FGMRES cannot solve
    start = casegen.developed_tube_state(case) if a.workload.startswith("tube") else None
    # Chebyshev degree: on the tree a second smoothing step saves too few
    # Krylov iterations to pay for its matrix pass (steps 4-8 from rest, plain
    # FP64: 5.77 s per step at degree 1, 6.80 s at 2, 8.16 s at 3); on the
    # Y-bifurcation degree 2 is the better one (0.48 s against 0.54 s at 1)
    cheb_degree = a.cheb_degree or (1 if a.workload.startswith("tree") else 2)
    smoother = {"gauss-seidel": device.TFI_SMOOTH_GAUSS_SEIDEL, "chebyshev": device.TFI_SMOOTH_CHEBYSHEV,
                "jacobi": device.TFI_SMOOTH_JACOBI, "auto": device.TFI_SMOOTH_AUTO}[a.smoother]
    log(f"plain FP64 (AMG with {a.smoother} smoothing, previous-step start)")
    plain, plain_all, plain_timers, plain_state, _ = run_method(
        dm, case, bc, faces, rho, mu, a.dt, a.steps, a.warmup, device.TFI_FP64, device.TFI_WARM_PREVIOUS, tol,
        waveform=wf, smoother=smoother, start=start, cheb_degree=cheb_degree)
    log("mixed precision + learned warm start")
    fast, fast_all, fast_timers, fast_state, warm_stats = run_method(
        dm, case, bc, faces, rho, mu, a.dt, a.steps, a.warmup, device.TFI_MIXED, device.TFI_WARM_LEARNED, tol,
        a.train_steps, waveform=wf, smoother=smoother, start=start, cheb_degree=cheb_degree)

    ms_plain = 1e3 * float(np.mean([r["wall_s"] for r in plain]))
    ms_fast = 1e3 * float(np.mean([r["wall_s"] for r in fast]))

    # SpMV roofline: FP64 SpMVs of the plain run's timed region (HIP events on the library stream)
    spmv_ms = plain_timers["spmv_ms"] / max(plain_timers["spmv_calls"], 1)
    spmv_bytes = plain_timers["spmv_bytes"] / max(plain_timers["spmv_calls"], 1)
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    # assembly roofline: per AssembleNsKernel launch (one element colour) of the
    # residual + exact-tangent assemblies in the plain run's timed region
    launches = max(plain_timers["assembly_launches"], 1)
    asm_launch_ms = plain_timers["assembly_ms"] / launches
    asm_launch_flops = plain_timers["assembly_flops"] / launches
    asm_tflops = asm_launch_flops / (asm_launch_ms * 1e-3) / 1e12
    asm_call_ms = plain_timers["assembly_ms"] / max(plain_timers["assembly_calls"], 1)
    res_call_ms = plain_timers["residual_ms"] / max(plain_timers["residual_calls"], 1)
    traffic = measured_traffic(a.workload)
    mesh = case.mesh
    desc = DESCRIPTIONS.get(a.workload, a.workload)

    out = {
        "metric": "NS time-step wall time @ rel. residual 1e-8: FP64 vs mixed/learned; SpMV GB/s",
        "value": round(ms_fast, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_fast, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (plain) / f32+f64 refinement (mixed); mlp bf16 inference",
        "data": f"synthetic (generated {a.workload} case; no external data)",
        "config": {"workload": f"{a.workload}: {desc}, quadratic IGA, {4*mesh.nodes} DOFs, {mesh.elements} elements, "
                               f"{inflow_text}, zero-traction outlets, backward Euler dt={a.dt} "
                               f"{'from developed Poiseuille flow' if start is not None else 'from rest'}; "
                               f"{a.warmup} untimed + {a.steps} timed steps per method",
                   "nodes": mesh.nodes, "elements": mesh.elements, "dofs": 4 * mesh.nodes,
                   "block_nonzeros": int(dm.nnzb), "newton_rtol": tol["rtol"], "linear_rtol": 1e-8,
                   "amg_smoother": a.smoother, "amg_chebyshev_degree": cheb_degree},
        "fp64_ms_per_step": round(ms_plain, 3),
        "speedup_vs_fp64": round(ms_plain / ms_fast, 3),
        "plain_fp64": method_summary(plain),
        "mixed_learned": dict(method_summary(fast),
                              correction_blend=float(np.mean([r["blend"] for r in fast])),
                              learned_guess_accepted=int(sum(r["warm_used"] for r in fast)),
                              network_gain=network_gain(fast),
                              warm_start=warm_stats),
        "whole_run_ms_per_step": {"plain_fp64": round(1e3 * float(np.mean([r["wall_s"] for r in plain_all])), 3),
                                  "mixed_learned": round(1e3 * float(np.mean([r["wall_s"] for r in fast_all])), 3),
                                  "steps": a.warmup + a.steps},
        "rel_l2_mixed_vs_fp64": rel_l2(fast_state, plain_state),
        "roofline": {"bound": "mfma", "achieved": round(asm_tflops, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(asm_tflops / FP64_PEAK_TFLOPS, 4),
                     "traffic": traffic.get("AssembleNsKernel<true, true, true>"),
                     "kernel": "AssembleNsKernel<true, true, true> (residual + exact tangent, FP64 MFMA pair phase)",
                     "avg_launch_ms": round(asm_launch_ms, 4), "algorithmic_flops_per_launch": asm_launch_flops,
                     "assembly_ms": round(asm_call_ms, 3), "residual_only_ms": round(res_call_ms, 3),
                     "basis_table_bytes": int(dm.stats["basis_table_bytes"])},
        "spmv_gbs": round(spmv_gbs, 1),
        "spmv_roofline": {"bound": "hbm", "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(spmv_gbs / HBM_PEAK_GBS, 4),
                          "traffic": traffic.get("BsrSpmvKernel<double, double, double, false>"),
                          "kernel": "BsrSpmvKernel<double> (FGMRES operator, FP64 values)",
                          "avg_ms": round(spmv_ms, 4), "algorithmic_bytes": int(spmv_bytes)},
        "gpu_time_ms_per_step": {"plain_fp64": kernel_breakdown(plain_timers, len(plain)),
                                 "mixed_learned": kernel_breakdown(fast_timers, len(fast))},
    }
    if extras:
        out["transport"] = transport_measurement(dm, case, plain_state, cpu=not a.no_cpu)
        out["transport"]["spmv_roofline"]["traffic"] = traffic.get("Bsr2SpmvKernel")
        out["solve"] = solve_measurement(dm, case, plain_state)
        if inflow == "steady":
            out["pulsatile"] = pulsatile_comparison(dm, case, bc, faces, rho, mu, 5e-3, tol, a.train_steps,
                                                    warmup=100, steps=a.pulsatile_steps)
    peaks = measured_peaks(ctx)
    out["measured_peaks"] = peaks
    out["roofline"]["frac_of_measured_peak"] = round(asm_tflops / peaks["fp64_mfma_tflops"], 4)
    out["spmv_roofline"]["frac_of_measured_peak"] = round(spmv_gbs / peaks["hbm_read_gbs"], 4)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, dm, rho, mu, a.dt, plain)
        model, cores = cpu_model()
        out["cpu_baseline"]["cpu_model"] = model
        out["cpu_baseline"]["physical_cores"] = cores
        out["cpu_baseline"]["full_step"] = cpu_full_step(ctx)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
