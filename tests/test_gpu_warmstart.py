"""Initial-state prediction for the time step on the GPU.

Every warm-start mode must end each step at the plain method's solution (the
converged residual is the judge, so a prediction only changes the work), the
learned guess may only be used when it lowers the step's initial residual,
and the trained network must beat the method without it: its own quartic
extrapolation base and the previous-step start of iga_cuda.cu:917-920.
"""
import math

import numpy as np
import pytest

from tubularflowiga_amd import boundary, casegen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tubularflowiga_amd import device
    c = device.Context(0)
    print(c.device_info(), flush=True)
    return c


@pytest.fixture(scope="module")
def ybif():
    return casegen.workload("ybif_small")


def _pulsatile(t):
    """Inflow scale Q(t)/Q0 = 1 + 0.6 sin 2 pi t + 0.2 sin 4 pi t (period 1 s)."""
    return 1.0 + 0.6 * math.sin(2 * math.pi * t) + 0.2 * math.sin(4 * math.pi * t)


def _run(ctx, case, mode, precision, steps, dt, rtol, waveform=None, dm=None, base=0):
    from tubularflowiga_amd import device
    rho, mu = casegen.flow_parameters(case)
    dm = dm or device.DeviceMesh(ctx, case.mesh)
    vmask, pmask, vt, pt = boundary.resolve_flow_boundaries(case)
    scale0 = waveform(0.0) if waveform else 1.0
    flow = device.FlowSolver(dm, density=rho, viscosity=mu, dt=dt, maximum_newton=20, nonlinear_rtol=rtol,
                             nonlinear_atol=1e-16, mass_rtol=1e30, linear_rtol=1e-8, precision=precision,
                             warm_start=mode)
    if mode == device.TFI_WARM_LEARNED:
        flow.configure_warm_start(hidden=128, train_steps=40, learning_rate=1e-3, min_snapshots=2, base=base)
    flow.set_state(boundary.initial_state(vmask, pmask, vt * scale0, pt))
    flow.set_boundary_faces(*boundary.boundary_faces(case.mesh))
    reports = []
    for k in range(steps):
        t = (k + 1) * dt
        flow.begin_step(np.array([math.sin(2 * math.pi * t), math.cos(2 * math.pi * t)]))
        flow.set_boundary(vmask, pmask, vt * (waveform(t) if waveform else 1.0), pt)
        reports.append(flow.newton())
        flow.commit_step()
    state = flow.state()
    stats = flow.warm_stats()
    flow.close()
    return state, reports, stats


def _rel(a, b):
    a, b = a.reshape(-1, 4), b.reshape(-1, 4)
    v = np.linalg.norm(a[:, :3] - b[:, :3]) / np.linalg.norm(b[:, :3])
    p = np.linalg.norm(a[:, 3] - b[:, 3]) / max(np.linalg.norm(b[:, 3]), 1e-300)
    return v, p


MODES = ["extrapolate", "extrapolate2", "extrapolate3", "extrapolate4", "mpe", "learned"]


def _history_solver(ctx, case, mode):
    from tubularflowiga_amd import device
    rho, mu = casegen.flow_parameters(case)
    dm = device.DeviceMesh(ctx, case.mesh)
    return dm, device.FlowSolver(dm, density=rho, viscosity=mu, dt=1e-3, warm_start=mode)


def test_mpe_continues_a_sum_of_modes_exactly(ctx, ybif):
    """States u_k = u_inf + sum_j a_j lambda_j^k v_j (four real modes and one
    damped oscillation): their increments obey a linear recurrence of order 6,
    so minimal polynomial extrapolation must find rank 6 among the 12
    increments it may combine and continue the sequence to rounding, as the
    least-squares fit of numpy does."""
    from tubularflowiga_amd import device
    rng = np.random.default_rng(7)
    dm, flow = _history_solver(ctx, ybif, device.TFI_WARM_MPE)
    n = 4 * ybif.mesh.nodes
    base, modes = rng.standard_normal(n), rng.standard_normal((6, n))
    lam, rho, theta = [0.9, 0.7, -0.5, 0.95], 0.8, 0.6

    def u(k):
        s = base + sum(l ** k * modes[j] for j, l in enumerate(lam))
        return s + rho ** k * (np.cos(k * theta) * modes[4] + np.sin(k * theta) * modes[5])

    K = 15
    for k in range(K + 1):
        flow.warm_record(u(k))
    pred, order = flow.warm_predict(device.TFI_WARM_MPE)
    step = np.linalg.norm(u(K + 1) - u(K))
    err = np.linalg.norm(pred - u(K + 1)) / step
    d = np.array([u(k) - u(k - 1) for k in range(K - 12, K + 1)])  # d_{K-12} .. d_K
    A = d[-2::-1][:12].T                                          # d_{K-1} .. d_{K-12}
    c = np.linalg.lstsq(A, d[-1], rcond=None)[0]
    ref = u(K) + d[::-1][:12].T @ c
    print(f"mpe order={order} error/increment={err:.2e} vs numpy {np.linalg.norm(pred - ref) / step:.2e}", flush=True)
    assert order == 6
    assert err <= 1e-9
    assert np.linalg.norm(pred - ref) / step <= 1e-9
    flow.close()


def test_polynomial_extrapolation_is_exact_on_polynomials(ctx, ybif):
    """Order-P extrapolation reproduces a degree-P polynomial in time."""
    from tubularflowiga_amd import device
    rng = np.random.default_rng(11)
    dm, flow = _history_solver(ctx, ybif, device.TFI_WARM_EXTRAPOLATE4)
    n = 4 * ybif.mesh.nodes
    w = rng.standard_normal((5, n))
    u = lambda k: sum((0.1 * k) ** p * w[p] for p in range(5))  # noqa: E731
    for k in range(6):
        flow.warm_record(u(k))
    pred, order = flow.warm_predict(device.TFI_WARM_EXTRAPOLATE4)
    assert order == 0
    assert np.linalg.norm(pred - u(6)) <= 1e-10 * np.linalg.norm(u(6))
    flow.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_every_mode_reaches_the_plain_solution(ctx, ybif, precision):
    from tubularflowiga_amd import device
    dm = device.DeviceMesh(ctx, ybif.mesh)
    steps, dt, rtol = 12, 1e-3, 1e-10
    plain, plain_reps, _ = _run(ctx, ybif, device.TFI_WARM_PREVIOUS, 0, steps, dt, rtol, dm=dm)
    for name in MODES:
        mode = getattr(device, "TFI_WARM_" + name.upper())
        state, reps, stats = _run(ctx, ybif, mode, precision, steps, dt, rtol, dm=dm)
        v, p = _rel(state, plain)
        used = sum(r["warm_start_used"] for r in reps)
        print(f"precision={precision} {name}: used={used} newton={sum(r['newton_iterations'] for r in reps)} "
              f"(plain {sum(r['newton_iterations'] for r in plain_reps)}) rel_l2 v={v:.2e} p={p:.2e}", flush=True)
        assert all(r["converged"] == 1 for r in reps)
        assert v <= 1e-10 and p <= 1e-10
        assert used > 0


def test_learned_guess_only_used_when_it_lowers_the_residual(ctx, ybif):
    from tubularflowiga_amd import device
    _, reps, stats = _run(ctx, ybif, device.TFI_WARM_LEARNED, 1, 16, 5e-3, 1e-8, waveform=_pulsatile)
    predicted = [r for r in reps if r["warm_start_initial_ratio"] > 0.0]
    assert predicted, "the learned predictor never became ready"
    for r in predicted:
        assert r["warm_start_used"] == (1 if r["warm_start_initial_ratio"] < 1.0 else 0)
    assert stats["accepted"] == sum(r["warm_start_used"] for r in reps)
    assert stats["accepted"] + stats["rejected"] == len(predicted)
    assert stats["snapshots"] > 0 and np.isfinite(stats["last_loss"])


@pytest.mark.parametrize("inflow", ["steady", "pulsatile"])
def test_trained_network_gains_over_the_method_without_it(ctx, ybif, inflow):
    """Initial residual of the step relative to the previous-step start (1.0),
    geometric mean over the last 10 of 40 steps: the learned guess against its
    base prediction alone at the same steps of the same run (the method
    without the network), for each of its two bases.  On the quartic
    extrapolation base, from rest under a steady inflow (a developing flow),
    the network must cut the base's initial residual; under the smooth
    pulsatile inflow at dt = 5e-3 s the base is already within 1e-4 of the next
    state and the online gate must keep the correction from making it worse.
    On the minimal-polynomial-extrapolation base (the default), which already
    continues the slow modes of the step map, the gate must keep the network
    from making it worse, and the learned predictor must start closer than on
    the quartic base."""
    from tubularflowiga_amd import device
    dm = device.DeviceMesh(ctx, ybif.mesh)
    dt, wf = (1e-3, None) if inflow == "steady" else (5e-3, _pulsatile)
    out = {}
    gmean = lambda v: float(np.exp(np.mean(np.log(v))))  # noqa: E731
    for name, base in (("quartic", device.TFI_WARM_EXTRAPOLATE4), ("mpe", device.TFI_WARM_MPE)):
        _, reps, stats = _run(ctx, ybif, device.TFI_WARM_LEARNED, 0, 40, dt, 1e-8, waveform=wf, dm=dm, base=base)
        tail = reps[-10:]
        out[name] = {"learned": gmean([r["warm_start_initial_ratio"] for r in tail]),
                     "base": gmean([r["warm_start_base_ratio"] for r in tail]),
                     "newton": sum(r["newton_iterations"] for r in tail),
                     "krylov": sum(r["linear_iterations"] for r in tail), "gate": stats.get("gate"),
                     "blend": float(np.mean([r["warm_start_blend"] for r in tail]))}
        assert all(r["warm_start_used"] == 1 for r in tail)
    print(f"{inflow}: {out}", flush=True)
    q, m = out["quartic"], out["mpe"]
    assert q["learned"] < (0.8 if inflow == "steady" else 1.1) * q["base"]
    assert m["learned"] <= 1.1 * m["base"]
    assert m["learned"] < q["learned"]


def test_chord_with_earlier_tangent_reaches_the_plain_solution(ctx, ybif):
    """Once the learned guess is within the chord window of the tolerance, the
    step is closed by chord steps with the tangent set up at an earlier state
    (no new tangent): the state must still be the plain method's."""
    from tubularflowiga_amd import device
    dm = device.DeviceMesh(ctx, ybif.mesh)
    steps, dt, rtol = 100, 1e-3, 1e-9
    plain, _, _ = _run(ctx, ybif, device.TFI_WARM_PREVIOUS, 0, steps, dt, rtol, waveform=_pulsatile, dm=dm)
    state, reps, _ = _run(ctx, ybif, device.TFI_WARM_LEARNED, 0, steps, dt, rtol, waveform=_pulsatile, dm=dm)
    tangent_free = [k for k, r in enumerate(reps) if r["newton_iterations"] == 0 and r["chord_steps"] >= 1]
    v, p = _rel(state, plain)
    print(f"steps closed by a chord with an earlier tangent: {len(tangent_free)} of {steps}; "
          f"rel_l2 v={v:.2e} p={p:.2e}", flush=True)
    assert all(r["converged"] == 1 for r in reps)
    assert tangent_free
    # both runs stop somewhere below the same Newton tolerance
    assert v <= rtol and p <= rtol
