"""The `iga_hip transport` input path on the CPU (`transport-check` touches no
GPU): the legacy parameter file, the optional case_config.json rules and the
boundary resolution agree with the reference's own readers
(tests/golden/transport_vectors.json, written through ReadTransportParameters,
ReadCaseConfiguration and ResolveBoundaryConditions of
solvers/cpu/include/CaseInput.hpp and include/CaseConfig.hpp), including their
error messages."""
import json
import subprocess
from pathlib import Path

import pytest

from tubularflowiga_amd import casegen

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "tubularflowiga_amd" / "iga_hip"
GOLDEN = json.loads((ROOT / "tests" / "golden" / "transport_vectors.json").read_text())
PARAMS = ["D", "vplus", "vminus", "kplus", "kminus", "detach_plus", "detach_minus", "dt", "nstep", "N0bc", "Nplusbc",
          "Nminusbc", "artificial_diffusion"]

pytestmark = pytest.mark.skipif(not CLI.exists(), reason="iga_hip is not built (run __graft_entry__.build())")


def run(*args):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("variant", GOLDEN["variants"], ids=[v["name"] for v in GOLDEN["variants"]])
def test_transport_inputs_resolve_like_the_reference(tmp_path, variant):
    case = casegen.workload("tube_c0")
    d = casegen.write_transport_case(case, tmp_path / "case", dict(GOLDEN["parameters"]), 5,
                                     case_config=variant["config"])
    path = d / "simulation_parameter.txt"
    text = path.read_text()
    if variant["edit"] == "artificial_diffusion_weight":
        text = text.replace("artificial_diffusion ", "artificial_diffusion_weight ")
    elif variant["edit"] == "drop_kplus":
        text = "\n".join(line for line in text.splitlines() if not line.startswith("kplus ")) + "\n"
    path.write_text(text)
    r = run("transport-check", d / "case.ntiga", d)
    if "error" in variant:
        assert r.returncode == 1 and r.stderr.startswith("iga_hip: ") and variant["error"] in r.stderr, r.stderr
        return
    assert r.returncode == 0, r.stderr
    line = [x for x in r.stdout.splitlines() if x.startswith("transport_check ")][0]
    rec = dict(kv.split("=", 1) for kv in line.split()[1:])
    assert [float(rec[k]) for k in PARAMS] == variant["parameters"]
    for k in ("transport_nodes", "constrained_index_sum", "velocity_nodes", "pressure_nodes"):
        assert int(rec[k]) == variant[k], k
    assert float(rec["n0_sum"]) == pytest.approx(variant["n0_sum"], rel=1e-15)
    assert float(rec["nplus_sum"]) == pytest.approx(variant["nplus_sum"], rel=1e-15)
    assert rec["boundary_config"] == ("case_config.json" if variant["config"] else "legacy-defaults")


def test_transport_usage_errors(tmp_path):
    case = casegen.workload("tube_c0")
    d = casegen.write_transport_case(case, tmp_path / "case", dict(GOLDEN["parameters"]), 5)
    r = run("transport", d / "case.ntiga")
    assert r.returncode == 1 and "usage: iga_hip transport" in r.stderr
    r = run("transport", d / "case.ntiga", d, "--linear-rtol", "-1")
    assert r.returncode == 1 and "--linear-rtol requires a positive number" in r.stderr
    r = run("transport", d / "case.ntiga", d, "x")
    assert r.returncode == 1 and "STEPS must be a non-negative integer" in r.stderr
    r = run("transport", d / "case.ntiga", d, "--preconditioner", "ilu")
    assert r.returncode == 1 and "--preconditioner must be amg|bjacobi" in r.stderr
    r = run("transport", d / "case.ntiga", d, "--bogus")
    assert r.returncode == 1 and "unknown option: --bogus" in r.stderr
    (d / "simulation_parameter.txt").unlink()
    r = run("transport", d / "case.ntiga", d)
    assert r.returncode == 1 and "cannot open simulation parameters" in r.stderr
