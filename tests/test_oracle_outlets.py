"""Lumped outlet models (resistance, RC and RCR Windkessel) on the CPU.

The oracle (oracle/outlets.py) and the drop-in driver's C++ models
(tubularflowiga_amd/host/outlets.hpp, through tests/outlet_host_check.cpp) are
both checked against the reference's own header outputs
(tests/golden/outlet_vectors.json, written by tests/golden/make_ref_vectors.py
through include/OutletModel.hpp), the values asserted in
solvers/cpu/tests/test_outlet_model.cpp, and closed forms; the driver's
configuration path (`iga_hip case-check`) is checked against the reference
reader's outlet models, applied tractions and error messages.
"""
import json
import shutil
import subprocess
from pathlib import Path

import pytest

from oracle import outlets as OL

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = json.loads((ROOT / "tests" / "golden" / "outlet_vectors.json").read_text())
CLI = ROOT / "tubularflowiga_amd" / "iga_hip"
KIND_INDEX = {"resistance": 0, "windkessel_rc": 1, "windkessel_rcr": 2}


def _model(args):
    kind, r, rp, rd, c, pref, pcap, p = args[:8]
    return OL.Outlet(label=2, kind=kind, resistance=r, proximal_resistance=rp, distal_resistance=rd, capacitance=c,
                     reference_pressure=pref, capacitor_pressure=pcap, pressure=p)


def _oracle_row(args):
    m = _model(args)
    dt, q = args[8], args[9]
    p, pc = OL.evaluate(m, q, dt)
    ev = OL.evaluate_coupling([m], [m.capacitor_pressure], [q], dt)
    relaxed, committed = [OL.Outlet(**vars(m))], [OL.Outlet(**vars(m))]
    OL.relax(relaxed, ev)
    OL.commit(committed, ev)
    return [p, pc, ev.maximum_pressure_change, ev.tolerance(), relaxed[0].pressure, committed[0].pressure,
            committed[0].capacitor_pressure]


def _close(a, b):
    return all(x == y or abs(x - y) <= 1e-15 * max(abs(x), abs(y)) for x, y in zip(a, b))


def test_oracle_matches_the_reference_header_outputs():
    rows = GOLDEN["evaluations"]
    assert len(rows) == 135 and sum("error" in r for r in rows) == 30
    for r in rows:
        if "error" in r:
            with pytest.raises(ValueError, match="positive transient dt"):
                _oracle_row(r["args"])
        else:
            assert _close(_oracle_row(r["args"]), r["out"]), r


def test_oracle_reproduces_the_reference_unit_test_values():
    """solvers/cpu/tests/test_outlet_model.cpp."""
    res = OL.Outlet(label=2, kind="resistance", resistance=3.0, reference_pressure=4.0)
    assert abs(OL.evaluate(res, 2.0, 0.0)[0] - 10.0) < 1e-14
    rc = OL.Outlet(label=2, kind="windkessel_rc", resistance=4.0, capacitance=2.0, reference_pressure=1.0,
                   capacitor_pressure=5.0)
    p, pc = OL.evaluate(rc, 3.0, 0.5)
    assert abs(pc - 23.25 / 4.25) < 1e-14 and abs(p - pc) < 1e-14
    rcr = OL.Outlet(**{**vars(rc), "kind": "windkessel_rcr", "proximal_resistance": 2.0, "distal_resistance": 4.0})
    p2, pc2 = OL.evaluate(rcr, 3.0, 0.5)
    assert abs(pc2 - pc) < 1e-14 and abs(p2 - (pc + 6.0)) < 1e-14
    saved = OL.Outlet(**{**vars(rcr), "flow": 2.5, "pressure": 8.0, "capacitor_pressure": 6.0})
    ev = OL.evaluate_coupling([saved], [5.0], [3.0], 0.5)
    assert ev.flow[0] == 3.0 and abs(ev.capacitor_pressure[0] - pc) < 1e-14 and abs(ev.pressure[0] - p2) < 1e-14
    models = [saved]
    OL.relax(models, ev)
    assert abs(models[0].pressure - 0.5 * (8.0 + ev.pressure[0])) < 1e-14
    OL.commit(models, ev)
    assert models[0].flow == 3.0 and models[0].pressure == ev.pressure[0]
    assert models[0].capacitor_pressure == ev.capacitor_pressure[0]
    with pytest.raises(ValueError):
        OL.evaluate(rc, 3.0, 0.0)


@pytest.mark.parametrize("kind", ["windkessel_rc", "windkessel_rcr"])
def test_windkessel_recursion_has_the_closed_form_decay(kind):
    """Backward Euler for C dPc/dt = Q - (Pc - p_ref)/R at constant Q decays to
    the steady state p_ref + R Q by exactly 1/(1 + dt/(R C)) per step."""
    r, c, pref, q, dt, pc0, rp = 0.8, 0.3, 0.25, 1.7, 0.05, 6.0, 0.4
    m = OL.Outlet(label=2, kind=kind, resistance=r, distal_resistance=r, proximal_resistance=rp, capacitance=c,
                  reference_pressure=pref, capacitor_pressure=pc0)
    steady = pref + r * q
    for n in range(1, 60):
        p, m.capacitor_pressure = OL.evaluate(m, q, dt)
        exact = steady + (pc0 - steady) / (1.0 + dt / (r * c)) ** n
        assert abs(m.capacitor_pressure - exact) <= 1e-13 * abs(exact)
        assert abs(p - (m.capacitor_pressure + (rp * q if kind == "windkessel_rcr" else 0.0))) <= 1e-15 * abs(p)


def test_coupled_step_settles_at_the_linear_fixed_point():
    """A vessel of resistance Rv fed at p_in passes Q = (p_in - p)/Rv; with a
    resistance outlet the coupling must settle at p = (p_ref + R p_in/Rv)/(1 + R/Rv)
    within the coupling tolerance, and a loop that cannot contract (R = 3 Rv:
    relaxed error factor -1) must fail after 12 passes."""
    p_in, rv = 2.0, 0.5
    for r, ok in ((0.4, True), (1.5, False)):
        models = [OL.Outlet(label=2, kind="resistance", resistance=r, reference_pressure=0.3, pressure=0.3)]
        applied = {}
        solve = applied.update
        flows = lambda labels: [(p_in - applied[lab]) / rv for lab in labels]  # noqa: E731
        if ok:
            hist = OL.coupled_step(models, 1e-3, solve, flows)
            exact = (0.3 + r * p_in / rv) / (1.0 + r / rv)
            assert abs(models[0].pressure - exact) <= 4 * hist[-1][1]
            assert hist[-1][0] <= hist[-1][1] and all(h[0] > h[1] for h in hist[:-1])
        else:
            with pytest.raises(RuntimeError, match="did not converge"):
                OL.coupled_step(models, 1e-3, solve, flows)


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    exe = tmp_path_factory.mktemp("outlets") / "outlet_host_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", str(ROOT / "tubularflowiga_amd" / "host"),
                    str(ROOT / "tests" / "outlet_host_check.cpp"), "-o", str(exe)], check=True)
    return exe


def test_driver_models_match_the_reference_header_outputs(host_check):
    """The driver's outlet bank reproduces the reference header's outputs on
    every golden case (to rounding: it evaluates each model as an affine
    response in the flow), and its error messages."""
    for r in GOLDEN["evaluations"]:
        args = r["args"]
        out = subprocess.run([str(host_check), args[0], *map(repr, args[1:])], capture_output=True, text=True)
        if "error" in r:
            assert out.returncode == 1 and r["error"] in out.stderr, (r, out.stderr)
        else:
            assert out.returncode == 0, out.stderr
            got = [float(v) for v in out.stdout.split()]
            # the bank evaluates each model as an affine response in the flow:
            # the reference's values up to rounding
            assert len(got) == len(r["out"]), r
            for g, w in zip(got, r["out"]):
                assert abs(g - w) <= 4e-15 * max(abs(w), 1.0), (r, got)


@pytest.mark.skipif(not CLI.exists(), reason="iga_hip is not built (run __graft_entry__.build())")
def test_case_check_reads_outlet_conditions_like_the_reference(tmp_path):
    from tubularflowiga_amd import casegen
    base = casegen.write_case(casegen.workload("tube_c0"), tmp_path / "case", steps=2, dt=1e-3)
    for k, rec in enumerate(GOLDEN["configs"]):
        d = tmp_path / f"c{k}"
        shutil.copytree(base, d)
        cfg = json.loads((d / "simulation_config.json").read_text())
        for b in cfg["boundaries"]:
            if b["label"] == 2:
                b["conditions"] = [rec["condition"]]
        (d / "simulation_config.json").write_text(json.dumps(cfg))
        r = subprocess.run([str(CLI), "case-check", str(d / "case.ntiga"), str(d)], capture_output=True, text=True,
                           timeout=120)
        if "error" in rec:
            assert r.returncode == 1 and rec["error"] in r.stderr, (rec, r.stderr)
            continue
        assert r.returncode == 0, r.stderr
        lines = [dict(kv.split("=", 1) for kv in line.split()[1:]) for line in r.stdout.splitlines()
                 if line.startswith("outlet ")]
        assert len(lines) == len(rec["models"])
        for got, (label, kind, pressure, capacitor) in zip(lines, rec["models"]):
            assert int(got["label"]) == int(label) and KIND_INDEX[got["kind"]] == int(kind)
            assert float(got["pressure"]) == pressure and float(got["capacitor_pressure"]) == capacitor
        check = [line for line in r.stdout.splitlines() if line.startswith("case_check ")][0]
        tractions = dict(kv.split("=", 1) for kv in check.split()[1:])["traction_labels"]
        assert [[float(v) for v in t.split(":")] for t in tractions.split(",")] == rec["tractions"]
