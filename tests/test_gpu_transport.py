"""The two-species neuron transport on the GPU (`iga_cuda transport`,
solvers/cuda/src/iga_cuda.cu:621-733) against the CPU oracle
(oracle/transport.py): the assembled 2x2 blocks of L and P, the time steps
through the C-ABI, and the `iga_hip transport` drop-in end to end."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from oracle import transport as T
from tubularflowiga_amd import casegen

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "tubularflowiga_amd" / "iga_hip"
PARAMS = dict(diffusion=0.05, vplus=1.25, vminus=0.0, kplus=0.5, kminus=0.1, detach_plus=0.3, detach_minus=0.0,
              dt=1e-3, artificial_diffusion=0.1)


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
    constrained = (np.asarray(case.point_labels) == 1).astype(np.int32)
    return case, pattern, casegen.tube_velocity(case), constrained


def _transport(ctx, case, velocity, constrained, params=PARAMS):
    from tubularflowiga_amd import device
    dm = device.DeviceMesh(ctx, case.mesh)
    return dm, device.Transport(dm, params, velocity, constrained)


def test_assembled_blocks_match_the_oracle(ctx, tube):
    case, pattern, vel, constrained = tube
    dm, tr = _transport(ctx, case, vel, constrained)
    try:
        L, P, singular = tr.matrices()
    finally:
        tr.close()
    Lr, Pr = T.assemble(case.mesh, pattern, vel, PARAMS, constrained)
    for got, want, name in ((L, Lr, "L"), (P, Pr, "P")):
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"{name}: max |gpu - oracle| / max |oracle| = {err:.2e}", flush=True)
        assert err <= 1e-13
    assert singular == 0


@pytest.mark.parametrize("preconditioner", ["amg", "bjacobi"])
def test_time_steps_match_the_oracle(ctx, tube, preconditioner):
    """Both preconditioners (the AMG through the 4x4 embedding, and the
    reference's block Jacobi alone) reach the oracle's states, the
    constrained values to rounding."""
    from tubularflowiga_amd import device
    case, pattern, vel, constrained = tube
    n0 = np.where(constrained == 1, 1.0, 0.0)
    nplus = np.where(constrained == 1, 0.5, 0.0)
    dm, tr = _transport(ctx, case, vel, constrained)
    its = []
    try:
        tr.set_preconditioner(device.TFI_PC_AMG if preconditioner == "amg" else device.TFI_PC_BJACOBI)
        tr.set_boundary(n0, nplus)
        tr.apply_boundary()
        for k in range(5):
            it, res = tr.step(zero_guess=(k == 0), rtol=1e-13)
            its.append(it)
            assert res <= 1e-13
        got = tr.state()
    finally:
        tr.close()
    want = T.run(case.mesh, pattern, vel, PARAMS, constrained, n0, nplus, 5)
    rel = [np.linalg.norm(got[:, f] - want[:, f]) / np.linalg.norm(want[:, f]) for f in range(2)]
    print(f"{preconditioner} 5 steps: iterations {its} rel_l2 N0={rel[0]:.2e} Nplus={rel[1]:.2e}", flush=True)
    assert max(rel) <= 1e-10
    assert np.abs(got[constrained == 1] - np.column_stack([n0, nplus])[constrained == 1]).max() <= 1e-13


def test_cli_transport_matches_the_oracle(tmp_path, tube):
    case, pattern, vel, constrained = tube
    legacy = dict(PARAMS, n0_bc=1.0, nplus_bc=0.5, nminus_bc=0.0)
    d = casegen.write_transport_case(case, tmp_path / "case", legacy, 30, velocity=vel)
    out = tmp_path / "transport.txt"
    cmd = [str(CLI), "transport", str(d / "case.ntiga"), str(d), "30", str(out), str(d / "transport_velocity.txt"),
           "--linear-rtol", "1e-14"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert re.search(r"^transport step=25/30 iterations=\d+ residual=", r.stdout, re.M)
    assert re.search(r"^transport step=30/30 ", r.stdout, re.M)
    rec = re.search(r"^transport_cuda nodes=(\d+) elements=(\d+) steps=30 .* final_l2=(\S+) "
                    r"singular_diagonal_blocks=0 ", r.stdout, re.M)
    assert rec and int(rec.group(1)) == case.mesh.nodes
    vals = np.loadtxt(out)
    assert np.array_equal(vals[:, 0], np.arange(case.mesh.nodes))
    n0 = np.where(constrained == 1, 1.0, 0.0)
    nplus = np.where(constrained == 1, 0.5, 0.0)
    want = T.run(case.mesh, pattern, vel, PARAMS, constrained, n0, nplus, 30)
    rel = [np.linalg.norm(vals[:, 1 + f] - want[:, f]) / np.linalg.norm(want[:, f]) for f in range(2)]
    print(f"cli 30 steps: rel_l2 N0={rel[0]:.2e} Nplus={rel[1]:.2e}", flush=True)
    assert max(rel) <= 1e-10
    # printed at the stream's default 6 significant digits, as the reference prints it
    assert float(rec.group(3)) == pytest.approx(np.linalg.norm(want), rel=1e-5)
    vtk = Path(str(out) + ".vtk").read_text()
    assert "SCALARS N0 double 1" in vtk and "SCALARS Nplus double 1" in vtk
