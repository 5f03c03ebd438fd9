"""The `iga_hip` host driver's input path on the CPU (`case-check` touches no
GPU): the C++ reader, wall trace, configuration and boundary resolution agree
with the Python host mirror, and bad input fails the way the reference's
driver fails (stderr `iga_hip: <message>`, exit status 1)."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tubularflowiga_amd import boundary, casegen

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "tubularflowiga_amd" / "iga_hip"

pytestmark = pytest.mark.skipif(not CLI.exists(), reason="iga_hip is not built (run __graft_entry__.build())")


def run(*args):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=120)


def record(text, name):
    for line in text.splitlines():
        if line.startswith(name + " "):
            return dict(kv.split("=", 1) for kv in line.split()[1:])
    raise AssertionError(f"no {name} record in:\n{text}")


@pytest.fixture(scope="module")
def tube_case(tmp_path_factory):
    case = casegen.workload("tube_c0")
    d = casegen.write_case(case, tmp_path_factory.mktemp("tube") / "case", steps=3, dt=1e-3)
    return case, d


def test_case_check_matches_python_resolution(tube_case):
    case, d = tube_case
    r = run("case-check", d / "case.ntiga", d)
    assert r.returncode == 0, r.stderr
    rec = record(r.stdout, "case_check")
    vmask, pmask, vt, pt = boundary.resolve_flow_boundaries(case)
    trace = boundary.wall_trace_basis(case.mesh, case.cells, case.point_labels)
    assert int(rec["nodes"]) == case.mesh.nodes and int(rec["elements"]) == case.mesh.elements
    assert int(rec["wall_trace"]) == len(trace)
    assert int(rec["velocity_nodes"]) == int(vmask.sum())
    assert int(rec["pressure_nodes"]) == int(pmask.sum())
    assert int(rec["velocity_mask_index_sum"]) == int(np.nonzero(vmask)[0].sum())
    assert float(rec["velocity_target_sum"]) == pytest.approx(float(vt.sum()), rel=1e-13, abs=1e-13)
    rho, mu = casegen.flow_parameters(case)
    assert float(rec["density"]) == pytest.approx(rho, rel=1e-15)
    assert rec["traction_labels"] == "2:0"


def test_case_check_applies_waveform(tmp_path):
    case = casegen.workload("tube_c0")
    wave = {"name": "pulse", "kind": "sinusoid", "units": "1", "mean": 1.0, "amplitude": 0.5, "period": 1.0}
    d = casegen.write_case(case, tmp_path / "case", steps=3, dt=0.25, waveform=wave)
    rec = record(run("case-check", d / "case.ntiga", d).stdout, "case_check")
    vmask, pmask, vt, pt = boundary.resolve_flow_boundaries(case, scale=1.0 + 0.5 * np.sin(2 * np.pi * 0.25))
    assert float(rec["velocity_target_sum"]) == pytest.approx(float(vt.sum()), rel=1e-13)


@pytest.mark.parametrize("version", [3, 4, 5])
def test_case_check_reads_every_database_version(tube_case, tmp_path, version):
    case, d = tube_case
    from tubularflowiga_amd.mesh import write_ntiga
    db = tmp_path / f"v{version}.ntiga"
    write_ntiga(db, case.mesh, version=version)
    r = run("case-check", db, d)
    assert r.returncode == 0, r.stderr
    assert record(r.stdout, "case_check")["version"] == str(version)


def _edit_config(d, tmp_path, fn):
    import shutil
    out = tmp_path / "edited"
    shutil.copytree(d, out)
    cfg = json.loads((out / "simulation_config.json").read_text())
    fn(cfg)
    (out / "simulation_config.json").write_text(json.dumps(cfg))
    return out


@pytest.mark.parametrize("edit,message", [
    (lambda c: c.update(schema_version=7), "supported schema versions"),
    (lambda c: c.update(bogus=1), "unknown key 'bogus'"),
    (lambda c: c["boundaries"].pop(), "has no simulation_config.json boundary definition"),
    (lambda c: c["boundaries"][1]["conditions"][0].update(profile="other.txt"), "unsupported velocity profile"),
    (lambda c: c["equation_systems"][0].update(viscosity=-1.0), "viscosity must be positive"),
])
def test_case_check_rejects_bad_configuration(tube_case, tmp_path, edit, message):
    case, d = tube_case
    bad = _edit_config(d, tmp_path, edit)
    r = run("case-check", bad / "case.ntiga", bad)
    assert r.returncode == 1
    assert r.stderr.startswith("iga_hip: ") and message in r.stderr, r.stderr


def test_corrupt_database_and_usage_errors(tube_case, tmp_path):
    case, d = tube_case
    bad = tmp_path / "bad.ntiga"
    bad.write_bytes(b"NOTADB00" + (d / "case.ntiga").read_bytes()[8:])
    r = run("case-check", bad, d)
    assert r.returncode == 1 and "not a supported .ntiga database" in r.stderr
    r = run("navier-stokes", d / "case.ntiga")
    assert r.returncode == 1 and "usage: iga_hip navier-stokes" in r.stderr
    r = run("navier-stokes", d / "case.ntiga", d, "--max-newton", "0")
    assert r.returncode == 1 and "--max-newton requires a positive integer" in r.stderr
    r = run("navier-stokes", d / "case.ntiga", d, "--output-every", "2")
    assert r.returncode == 1 and "--output-every requires --output" in r.stderr
    r = run("navier-stokes", d / "case.ntiga", d, "--precision", "fp16")
    assert r.returncode == 1 and "--precision must be fp64|mixed" in r.stderr
    r = run("navier-stokes", d / "case.ntiga", d, "--chord-polish", "maybe")
    assert r.returncode == 1 and "--chord-polish must be on|off" in r.stderr
    r = run("navier-stokes", d / "case.ntiga", d, "--train-steps", "-3")
    assert r.returncode == 1 and "--train-steps requires a positive integer" in r.stderr
