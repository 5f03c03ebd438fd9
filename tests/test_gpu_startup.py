"""Impulsive start at the learned-warm-start workload's time step: the
0.6M-DOF Y-bifurcation from rest with Q(t) = 5 (1 + 0.6 sin 2 pi t + 0.2 sin
4 pi t) ml/s and backward Euler with dt = 5e-3 s.  Without a limit on the
Newton update the first step overshoots into states whose tangents the
AMG-preconditioned FGMRES cannot solve; with the velocity step limit every
step converges, the first one in a few shortened steps and the later ones in
plain Newton steps.  Plain FP64 and mixed precision + learned warm start must
both converge every step and agree within 1e-10 per field."""
import math

import numpy as np
import pytest

from tubularflowiga_amd import boundary, casegen

pytestmark = pytest.mark.gpu

DT, STEPS = 5e-3, 8


def _waveform(t):
    return 1.0 + 0.6 * math.sin(2 * math.pi * t) + 0.2 * math.sin(4 * math.pi * t)


@pytest.fixture(scope="module")
def ybif():
    return casegen.workload("ybif_c1")


def _run(ctx, dm, case, precision, warm):
    from tubularflowiga_amd import device
    rho, mu = casegen.flow_parameters(case)
    vm, pm, vt, pt = boundary.resolve_flow_boundaries(case)
    flow = device.FlowSolver(dm, density=rho, viscosity=mu, dt=DT, nonlinear_rtol=1e-12, nonlinear_atol=1e-18,
                             mass_rtol=1e-6, linear_rtol=1e-8, precision=precision, warm_start=warm, restart=100,
                             linear_maximum_iterations=4000, maximum_newton=30)
    if warm == device.TFI_WARM_LEARNED:
        flow.configure_warm_start(hidden=128, train_steps=20, learning_rate=1e-3, min_snapshots=2)
    reports = []
    try:
        flow.set_state(boundary.initial_state(vm, pm, vt * _waveform(0.0), pt))
        flow.set_boundary(vm, pm, vt * _waveform(0.0), pt)
        flow.set_boundary_faces(*boundary.boundary_faces(case.mesh))
        for k in range(STEPS):
            t = (k + 1) * DT
            flow.begin_step(np.array([math.sin(2 * math.pi * t), math.cos(2 * math.pi * t),
                                      math.sin(4 * math.pi * t), math.cos(4 * math.pi * t)]))
            flow.set_boundary(vm, pm, vt * _waveform(t), pt)
            rep = flow.newton()
            steps = [r["step"] for r in rep["records"] if r["linear_iterations"] > 0]
            print(f"precision={precision} warm={warm} step {k + 1}: newton={rep['newton_iterations']} "
                  f"krylov={rep['linear_iterations']} step lengths={[round(s, 3) for s in steps]}", flush=True)
            reports.append(rep)
            flow.commit_step()
        return flow.state().reshape(-1, 4), reports
    finally:
        flow.close()


def test_impulsive_start_at_dt_5ms_converges_every_step(ybif):
    from tubularflowiga_amd import device
    ctx = device.Context(0)
    dm = device.DeviceMesh(ctx, ybif.mesh)
    try:
        plain, plain_reps = _run(ctx, dm, ybif, device.TFI_FP64, device.TFI_WARM_PREVIOUS)
        fast, fast_reps = _run(ctx, dm, ybif, device.TFI_MIXED, device.TFI_WARM_LEARNED)
    finally:
        dm.close()
    for reps in (plain_reps, fast_reps):
        assert all(r["converged"] == 1 for r in reps)
    # the first step needed the limit, the later plain steps did not
    first = [r["step"] for r in plain_reps[0]["records"] if r["linear_iterations"] > 0]
    assert min(first) < 1.0
    for rep in plain_reps[2:]:
        assert all(r["step"] == 1.0 for r in rep["records"] if r["linear_iterations"] > 0)
    v = np.linalg.norm(fast[:, :3] - plain[:, :3]) / np.linalg.norm(plain[:, :3])
    p = np.linalg.norm(fast[:, 3] - plain[:, 3]) / np.linalg.norm(plain[:, 3])
    print(f"mixed + learned vs plain after {STEPS} steps: velocity {v:.2e} pressure {p:.2e}", flush=True)
    assert v <= 1e-10 and p <= 1e-10
