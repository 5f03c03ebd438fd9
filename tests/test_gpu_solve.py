"""The configured linear transport on the GPU (`iga_cuda solve`,
solvers/cuda/include/ConfiguredTransportHost.inc:142-471) against the CPU
oracle (oracle/solve.py): the assembled F x F blocks of L and P and the source
through the C-ABI, and backward-Euler steps, for 1, 2 and 3 coupled fields
covering every term kind, the three SUPG diffusion weightings, both tau
forms, constrained (node, field) entries, and flux / Robin faces."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import solve as S
from tubularflowiga_amd import casegen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tubularflowiga_amd import device
    c = device.Context(0)
    print(c.device_info(), flush=True)
    return c


@pytest.fixture(scope="module")
def tube():
    case = casegen.workload("tube_c0")
    pattern = O.block_pattern(case.mesh)
    return case, pattern, casegen.tube_velocity(case)


def surfaces_on(mesh, label, fields, flux, robin, exterior):
    labels = np.asarray(mesh.boundary_labels).reshape(-1, 6)
    element, face = np.nonzero(labels == label)
    n = len(element)
    return dict(element=element.astype(np.int32), face=face.astype(np.int32),
                flux=np.tile(flux, (n, 1)), robin=np.tile(robin, (n, 1)),
                robin_source=np.tile(np.asarray(robin) * np.asarray(exterior), (n, 1)))


def systems(case):
    """name -> (system, constrained fields at the inlet, inlet values, surfaces)"""
    m = case.mesh
    neuron = dict(fields=2, dt=1e-3, supg=[S.SUPG_NONE, S.SUPG_TEST], tau_diffusivity=[0.0, 0.0],
                  terms=[(S.TIME_DERIVATIVE, 0, 0, 1.0), (S.LINEAR_COUPLING, 0, 0, 0.6), (S.DIFFUSION, 0, 0, 0.05),
                         (S.LINEAR_COUPLING, 0, 1, -0.3), (S.TIME_DERIVATIVE, 1, 1, 1.0),
                         (S.LINEAR_COUPLING, 1, 1, 0.3), (S.ADVECTION, 1, 1, 1.0), (S.DIFFUSION, 1, 1, 0.125),
                         (S.LINEAR_COUPLING, 1, 0, -0.5)])
    three = dict(fields=3, dt=2e-3, supg=[S.SUPG_CONSISTENT, S.SUPG_TEST, S.SUPG_NONE],
                 tau_diffusivity=[0.05, 0.0, 0.0],
                 terms=[(S.TIME_DERIVATIVE, 0, 0, 1.0), (S.DIFFUSION, 0, 0, 0.05), (S.ADVECTION, 0, 0, 1.0),
                        (S.LINEAR_COUPLING, 0, 1, -0.2), (S.VOLUME_SOURCE, 0, 0, 0.3),
                        (S.TIME_DERIVATIVE, 1, 1, 1.0), (S.DIFFUSION, 1, 1, 0.02), (S.ADVECTION, 1, 1, 1.0),
                        (S.LINEAR_COUPLING, 1, 0, -0.1), (S.LINEAR_COUPLING, 1, 1, 0.1),
                        (S.TIME_DERIVATIVE, 2, 2, 2.0), (S.DIFFUSION, 2, 2, 0.01), (S.LINEAR_COUPLING, 2, 0, -0.05),
                        (S.VOLUME_SOURCE, 2, 2, 0.1)])
    walls = surfaces_on(m, 0, 3, flux=[0.0, 0.0, 0.4], robin=[0.7, 0.0, 0.0], exterior=[0.25, 0.0, 0.0])
    outlet = surfaces_on(m, 2, 3, flux=[0.0, 0.2, 0.0], robin=[0.0, 0.0, 1.5], exterior=[0.0, 0.0, 0.1])
    both = {k: np.concatenate([walls[k], outlet[k]]) for k in walls}
    single = dict(fields=1, dt=5e-3, supg=[S.SUPG_GALERKIN], tau_diffusivity=[0.0],
                  terms=[(S.TIME_DERIVATIVE, 0, 0, 1.0), (S.DIFFUSION, 0, 0, 0.1), (S.ADVECTION, 0, 0, 1.0)])
    return {"neuron": (neuron, [0, 1], [1.0, 0.5], None),
            "three_fields": (three, [0, 1], [1.0, 0.25, 0.0], both),
            "one_field": (single, [0], [1.0], None)}


def masks(case, sysdef, constrained_fields, inlet_values):
    F = sysdef["fields"]
    inlet = np.asarray(case.point_labels) == 1
    constrained = np.zeros((case.mesh.nodes, F), np.int32)
    values = np.zeros((case.mesh.nodes, F))
    for f in constrained_fields:
        constrained[inlet, f] = 1
        values[inlet, f] = inlet_values[f]
    return constrained, values


def make(ctx, case, vel, sysdef, constrained, surfaces):
    from tubularflowiga_amd import device
    dm = device.DeviceMesh(ctx, case.mesh)
    return device.ConfiguredTransport(dm, sysdef["fields"], sysdef["terms"], sysdef["supg"],
                                      sysdef["tau_diffusivity"], sysdef["dt"], vel, constrained, surfaces)


@pytest.mark.parametrize("name", ["neuron", "three_fields", "one_field"])
def test_assembled_blocks_match_the_oracle(ctx, tube, name):
    case, pattern, vel = tube
    sysdef, cf, iv, surfaces = systems(case)[name]
    constrained, _ = masks(case, sysdef, cf, iv)
    print(f"case {name}: fields={sysdef['fields']} terms={len(sysdef['terms'])}", flush=True)
    tr = make(ctx, case, vel, sysdef, constrained, surfaces)
    try:
        L, P, src, singular = tr.matrices()
    finally:
        tr.close()
    Lr, Pr, Sr = S.assemble(case.mesh, pattern, vel, sysdef, constrained, surfaces)
    for got, want, label in ((L, Lr, "L"), (P, Pr, "P"), (src, Sr, "source")):
        scale = np.abs(want).max()
        err = np.abs(got - want).max() / scale if scale > 0 else np.abs(got).max()
        print(f"{name} {label}: max |gpu - oracle| / max |oracle| = {err:.2e}", flush=True)
        assert err <= 1e-13
    assert singular == 0


@pytest.mark.parametrize("name,preconditioner", [("neuron", "bjacobi"), ("three_fields", "bjacobi"),
                                                 ("one_field", "bjacobi"), ("neuron", "amg"), ("three_fields", "amg")])
def test_time_steps_match_the_oracle(ctx, tube, name, preconditioner):
    """Both preconditioners (the reference's block-Jacobi scaling alone, and
    the coupled AMG through the 4x4 embedding) reach the oracle's states."""
    from tubularflowiga_amd import device
    case, pattern, vel = tube
    sysdef, cf, iv, surfaces = systems(case)[name]
    F = sysdef["fields"]
    constrained, values = masks(case, sysdef, cf, iv)
    start = np.where(constrained == 1, values, 0.0)
    print(f"case {name}: 4 steps", flush=True)
    tr = make(ctx, case, vel, sysdef, constrained, surfaces)
    try:
        tr.set_preconditioner(device.TFI_PC_AMG if preconditioner == "amg" else device.TFI_PC_BJACOBI)
        tr.set_boundary(values)
        tr.set_state(start)
        for k in range(4):
            it, res = tr.step(zero_guess=True, rtol=1e-13)
            assert res <= 1e-13
        got = tr.state()
        # the reference's warm start: from the previous state to rtol of the
        # initial residual
        it, res = tr.step(zero_guess=False, rtol=1e-8)
        assert 0 < it and res <= 1e-8
    finally:
        tr.close()
    want = S.run(case.mesh, pattern, vel, sysdef, constrained, values, 4, start, surfaces)
    # the FGMRES stopping test weighs the whole state, so a field two orders
    # smaller than the others (the tracer of three_fields, ~1e-3 against ~1)
    # carries the same absolute error: 1e-10 of the state's norm per field and
    # for the whole state, 1e-9 of each field's own norm
    whole = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"{name} state: rel L2 = {whole:.2e}", flush=True)
    assert whole <= 1e-10
    for f in range(F):
        rel = np.linalg.norm(got[:, f] - want[:, f]) / np.linalg.norm(want[:, f])
        print(f"{name} field {f}: rel L2 = {rel:.2e}", flush=True)
        assert rel <= 1e-9
        assert np.linalg.norm(got[:, f] - want[:, f]) <= 1e-10 * np.linalg.norm(want)
    mask = constrained.astype(bool)
    assert np.abs(got[mask] - values[mask]).max() <= 1e-13


def _cli_case(tmp_path):
    """The three-field case of tests/test_solve_host.py with node-label
    Dirichlet entries, and the oracle system it compiles to."""
    import copy
    import test_solve_host as H
    boundaries = copy.deepcopy(H.BOUNDARIES)
    for c in boundaries[1]["conditions"]:
        c["dirichlet_dofs"] = "node_label"
    case = casegen.workload("tube_c0")
    d = casegen.write_solve_case(case, tmp_path / "case", H.FIELDS, [H.SYSTEM], boundaries, dt=2e-3, steps=6)
    K = {"time_derivative": S.TIME_DERIVATIVE, "diffusion": S.DIFFUSION, "advection": S.ADVECTION,
         "linear_coupling": S.LINEAR_COUPLING, "volume_source": S.VOLUME_SOURCE}
    index = {"oxygen": 0, "carrier": 1, "tracer": 2}
    terms = [(K[t["operator"]], index[t["equation"]], index[t.get("trial", t["equation"])], t.get("coefficient", 1.0))
             for t in H.SYSTEM["terms"]]
    sysdef = dict(fields=3, dt=2e-3, terms=terms, supg=[S.SUPG_CONSISTENT, S.SUPG_TEST, S.SUPG_NONE],
                  tau_diffusivity=[0.06, 0.0, 0.0])
    n = case.mesh.nodes
    inlet = np.asarray(case.point_labels) == 1
    constrained = np.zeros((n, 3), np.int32)
    values = np.zeros((n, 3))
    constrained[inlet, 0], constrained[inlet, 1] = 1, 1
    values[inlet, 0], values[inlet, 1] = 1.0, 0.25
    start = np.where(constrained == 1, values, np.array([0.2, 0.05, 0.0]))
    surfaces = surfaces_on(case.mesh, 0, 3, flux=[0.0, 0.0, 0.4], robin=[0.7, 0.0, 0.0], exterior=[0.25, 0.0, 0.0])
    return case, d, sysdef, constrained, values, start, surfaces


def _read_output(path, fields):
    data = np.loadtxt(path)
    assert data.shape[1] == fields + 1 and np.array_equal(data[:, 0], np.arange(len(data)))
    return data[:, 1:]


def test_cli_solve_matches_the_oracle_and_restarts(tmp_path):
    import subprocess
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    case, d, sysdef, constrained, values, start, surfaces = _cli_case(tmp_path)
    cli = Path(__file__).resolve().parent.parent / "tubularflowiga_amd" / "iga_hip"
    out = tmp_path / "run" / "species.txt"
    print("case: iga_hip solve, tube_c0, 3 fields, 6 steps", flush=True)
    r = subprocess.run([str(cli), "solve", str(d / "case.ntiga"), str(d), "--output", str(out), "--linear-rtol", "1e-11",
                        "--output-every", "3", "--visualization-format", "vtu"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "iga_cuda_solve system=species fields=3 steps=6 run_end_step=6" in r.stdout
    assert (out.parent / "species.txt.fields").read_text().split() == ["oxygen", "carrier", "tracer"]
    assert (out.parent / "species.pvd").exists() and (out.parent / "species.step000003.vtu").exists()
    assert (out.parent / "species.step000003.txt").exists() and (out.parent / "species.vtu").exists()
    # by default the run writes the Bezier VTKHDF series instead of the VTU files:
    # the initial state and steps 3 and 6, one point array per field
    hout = tmp_path / "hdf" / "species.txt"
    hout.parent.mkdir()
    r = subprocess.run([str(cli), "solve", str(d / "case.ntiga"), str(d), "--output", str(hout), "--linear-rtol",
                        "1e-11", "--output-every", "3"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert not (hout.parent / "species.vtu").exists() and not (hout.parent / "species.pvd").exists()
    assert "bezier_geometry_points=" in r.stdout
    from test_gpu_cli import h5_values
    h5 = hout.parent / "species.vtkhdf"
    assert np.allclose(h5_values(h5, "/VTKHDF/Steps/Values"), [0.0, 6e-3, 1.2e-2], rtol=1e-12, atol=1e-15)
    npts = int(h5_values(h5, "/VTKHDF/NumberOfPoints")[0])
    for name in ("oxygen", "carrier", "tracer"):
        assert h5_values(h5, f"/VTKHDF/PointData/{name}").shape == (3 * npts,)
    manifest = (out.parent / "physiology_fields.json").read_text()
    assert '"enabled": false' in manifest and '"oxygen": {"status": "solved"}' in manifest
    got = _read_output(out, 3)
    pattern = O.block_pattern(case.mesh)
    vel = casegen.tube_velocity(case)
    want = S.run(case.mesh, pattern, vel, sysdef, constrained, values, 6, start, surfaces)
    whole = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"cli state: rel L2 = {whole:.2e}", flush=True)
    assert whole <= 1e-8
    for f in range(3):
        rel = np.linalg.norm(got[:, f] - want[:, f]) / np.linalg.norm(want[:, f])
        print(f"cli field {f}: rel L2 = {rel:.2e}", flush=True)
        assert rel <= 1e-7
    # checkpoint after 3 steps, restart to 6: the same state as the straight run
    ck = tmp_path / "ck" / "state"
    ck.parent.mkdir()
    r1 = subprocess.run([str(cli), "solve", str(d / "case.ntiga"), str(d), "--linear-rtol", "1e-11",
                         "--stop-after-step", "3", "--checkpoint", str(ck)], capture_output=True, text=True, timeout=300)
    assert r1.returncode == 0, r1.stderr
    assert "checkpoint=" in r1.stdout and "completed_step=3" in r1.stdout
    out2 = tmp_path / "restart" / "species.txt"
    r2 = subprocess.run([str(cli), "solve", str(d / "case.ntiga"), str(d), "--linear-rtol", "1e-11", "--restart", str(ck),
                         "--output", str(out2)], capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stderr
    assert "restart=" in r2.stdout and "completed_step=3" in r2.stdout
    again = _read_output(out2, 3)
    assert np.abs(again - got).max() <= 1e-12 * np.abs(got).max()


@pytest.mark.parametrize("preconditioner", ["bjacobi", "amg"])
def test_cli_solve_with_a_velocity_series(tmp_path, preconditioner):
    """A snapshot-series velocity source: the operators are re-assembled before
    every step for the velocity interpolated linearly at the step's end time
    (ConfiguredTransportHost.inc:337-376), with the block-Jacobi inverses or the
    AMG hierarchy rebuilt from each new matrix; the oracle does the same with
    its own assembly and a direct solve per step."""
    import json
    import subprocess
    import sys
    from pathlib import Path
    import scipy.sparse.linalg as spla
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    case, d, sysdef, constrained, values, start, surfaces = _cli_case(tmp_path)
    config = json.loads((d / "simulation_config.json").read_text())
    for t in config["equation_systems"][0]["terms"]:
        if t["operator"] == "advection":
            t["velocity"] = "series"
    for x in config["equation_systems"][0]["stabilization"]:
        x["velocity"] = "series"
    config["velocity_sources"] = [{"name": "series", "kind": "snapshot_series", "manifest": "velocity.csv",
                                   "interpolation": "linear", "out_of_range": "hold"}]
    config["time"]["steps"] = 4
    (d / "simulation_config.json").write_text(json.dumps(config))
    base = casegen.tube_velocity(case)
    snaps = [(0.0, base), (0.005, 2.0 * base), (0.0075, 0.5 * base)]
    for k, (_, v) in enumerate(snaps):
        np.savetxt(d / f"v{k}.txt", v, fmt="%.17g")
    (d / "velocity.csv").write_text("time,file\n" + "".join(f"{t!r},v{k}.txt\n" for k, (t, _) in enumerate(snaps)))
    cli = Path(__file__).resolve().parent.parent / "tubularflowiga_amd" / "iga_hip"
    out = tmp_path / "series" / "species.txt"
    print(f"case: iga_hip solve with a velocity series, 4 steps, {preconditioner}", flush=True)
    r = subprocess.run([str(cli), "solve", str(d / "case.ntiga"), str(d), "--output", str(out), "--linear-rtol",
                        "1e-11", "--preconditioner", preconditioner], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    got = _read_output(out, 3)
    pattern = O.block_pattern(case.mesh)
    mask = constrained.ravel().astype(bool)
    c = start.ravel().copy()
    for step in range(4):
        t = (step + 1) * 2e-3
        times = [s[0] for s in snaps]
        if t >= times[-1]:
            vel = snaps[-1][1]
        else:
            k = int(np.searchsorted(times, t))
            w = (t - times[k - 1]) / (times[k] - times[k - 1])
            vel = (1.0 - w) * snaps[k - 1][1] + w * snaps[k][1]
        L, P, Sv = S.assemble(case.mesh, pattern, vel, sysdef, constrained, surfaces)
        b = S.to_csr(pattern, P) @ c + Sv.ravel()
        b[mask] = values.ravel()[mask]
        c = spla.spsolve(S.to_csr(pattern, L).tocsc(), b)
    want = c.reshape(-1, 3)
    whole = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"series state: rel L2 = {whole:.2e}", flush=True)
    assert whole <= 1e-8
