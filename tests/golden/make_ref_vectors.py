"""Write tests/golden/ref_vectors.json from the reference's own PETSc-free headers.

Runs oracle/_ref/ref_probe (our driver around solvers/cpu/include/{IgaDatabase,
ElementGeometry,Quadrature,BoundaryFlow,BoundarySupport}.hpp, built by
`make -C oracle ref`) on a few elements and on a generated case, and stores the
outputs as fixtures:
  * basis values and physical gradients (EvaluateBoundaryBasis) at points of a
    unit, a curved, an affine-skewed and a degree-elevated quadratic element;
  * face flows (IntegrateBoundaryFlow) and pressure-traction loads
    (IntegrateBoundaryPressureTraction) with a seeded nodal state;
  * surface Jacobian measures (SurfaceJacobianMeasure);
  * the SHA-256 of the reference reader's text dump of a generated `.ntiga`
    database (IgaDatabase.hpp) and its wall-trace basis (WallTraceBasis);
  * into transport_vectors.json: the transport parameters and boundary
    resolution of the tube case under several simulation_parameter.txt /
    case_config.json variants, or the reference's error message
    (ReadTransportParameters / ReadCaseConfiguration / ResolveBoundaryConditions,
    solvers/cpu/include/CaseInput.hpp, include/CaseConfig.hpp);
  * into outlet_vectors.json: lumped outlet model evaluations, coupling
    tolerance, relaxation and commit (include/OutletModel.hpp) over a grid of
    kinds, parameters, flows and time steps, and the outlet models, applied
    outlet tractions or configuration errors of the tube case with each outlet
    condition (InitializeOutletModels / MaterializeOutletPressures /
    ReadSimulationConfiguration, include/SimulationConfig.hpp).

Run in the build container only (the reference tree is not shipped):
    python tests/golden/make_ref_vectors.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
PROBE = os.path.join(ROOT, "oracle", "_ref", "ref_probe")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.json")
OUTLETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "outlet_vectors.json")
TRANSPORT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "transport_vectors.json")

TRANSPORT_PARAMETERS = {"diffusion": 0.05, "vplus": 1.25, "vminus": 0.0, "kplus": 0.5, "kminus": 0.1,
                        "detach_plus": 0.3, "detach_minus": 0.0, "dt": 1e-3, "n0_bc": 1.0, "nplus_bc": 0.5,
                        "nminus_bc": 0.0, "artificial_diffusion": 0.1}
# (name, parameter-file edit, case_config.json or None)
TRANSPORT_VARIANTS = [
    ("legacy", None, None),
    ("weight_key", "artificial_diffusion_weight", None),
    ("inlet_override", None, {"schema_version": 1, "boundaries": {"inherit_legacy": True, "conditions": [
        {"label": 1, "type": "inlet", "transport": {"N0": 2.5, "Nplus": 0.25}}]}}),
    ("explicit_rules", None, {"boundaries": {"inherit_legacy": False, "conditions": [
        {"label": 0, "type": "wall"}, {"label": 1, "type": "inlet", "velocity_scale": 2.0},
        {"label": 2, "type": "outlet", "pressure": 0.1}]}}),
    ("missing_key", "drop_kplus", None),
    ("outlet_transport", None, {"boundaries": {"conditions": [
        {"label": 2, "type": "outlet", "transport": {"N0": 1.0, "Nplus": 1.0}}]}}),
    ("unknown_label", None, {"boundaries": {"conditions": [{"label": 7, "type": "wall"}]}}),
    ("uncovered_label", None, {"boundaries": {"inherit_legacy": False, "conditions": [
        {"label": 0, "type": "wall"}, {"label": 1, "type": "inlet"}]}}),
]

# (resistance, proximal, distal, capacitance, reference pressure, capacitor
# pressure, applied pressure): the first row is solvers/cpu/tests/test_outlet_model.cpp's
OUTLET_PARAMETERS = [(4.0, 2.0, 4.0, 2.0, 1.0, 5.0, 8.0), (3.0, 0.5, 1.5, 0.25, 4.0, -2.0, 0.0),
                     (0.0125, 1e-3, 0.02, 40.0, 0.1, 0.2, 0.15)]
OUTLET_FLOWS = [3.0, 2.0, -1.25, 0.0, 7.5e-4]
OUTLET_STEPS = [0.5, 1e-3, 0.0]
# outlet conditions placed on the tube case's outlet label; the last ones are
# rejected by the reference's configuration reader
OUTLET_CONDITIONS = [
    {"field": "pressure", "type": "resistance", "resistance": 3.0, "reference_pressure": 4.0},
    {"field": "pressure", "type": "windkessel_rc", "resistance": 4.0, "capacitance": 2.0,
     "reference_pressure": 1.0, "initial_pressure": 5.0},
    {"field": "pressure", "type": "windkessel_rcr", "proximal_resistance": 2.0, "distal_resistance": 4.0,
     "capacitance": 2.0, "reference_pressure": 1.0, "initial_pressure": 5.0},
    {"field": "pressure", "type": "resistance", "resistance": 0.0},
    {"field": "pressure", "type": "windkessel_rc", "resistance": 4.0},
    {"field": "pressure", "type": "windkessel_rcr", "proximal_resistance": -1.0, "distal_resistance": 4.0,
     "capacitance": 2.0},
    {"field": "pressure", "type": "windkessel_rcr", "proximal_resistance": 1.0, "distal_resistance": 0.0,
     "capacitance": 2.0},
]


def element_file(path, conn, extraction, bezier, state):
    with open(path, "w") as f:
        f.write(f"{len(conn)}\n")
        f.write(" ".join(str(int(c)) for c in conn) + "\n")
        for row in extraction:
            f.write(" ".join(repr(float(v)) for v in row) + "\n")
        for p in bezier:
            f.write(" ".join(repr(float(v)) for v in p) + "\n")
        for s in state:
            f.write(" ".join(repr(float(v)) for v in s) + "\n")


def probe(*args):
    out = subprocess.run([PROBE, *map(str, args)], check=True, capture_output=True, text=True).stdout
    return [float(v) for v in out.split()]


def elements():
    from tubularflowiga_amd import casegen
    grid = np.array([[i / 3.0, j / 3.0, k / 3.0] for k in range(4) for j in range(4) for i in range(4)])
    eye = np.eye(64)
    out = {"unit": (np.arange(64), eye, grid)}
    curved = grid.copy()
    curved[:, 0] += 0.08 * grid[:, 0] * grid[:, 1]
    curved[:, 1] += 0.05 * grid[:, 1] * grid[:, 2]
    curved[:, 2] += 0.04 * grid[:, 0] * grid[:, 2]
    out["curved"] = (np.arange(64), eye, curved)
    A = np.array([[2.0, 0.3, 0.2], [0.1, 3.0, 0.4], [0.2, 0.3, 4.0]])
    out["skewed"] = (np.arange(64), eye, grid @ A.T)
    case = casegen.workload("tube_c0")
    m = case.mesh
    for e in (0, 37):
        out[f"tube_c0_e{e}"] = (m.element_connectivity(e), m.dense_extraction(e), m.element_bezier(e))
    return out


def outlet_vectors():
    from tubularflowiga_amd import casegen
    out = {"evaluations": [], "configs": []}
    for kind in ("resistance", "windkessel_rc", "windkessel_rcr"):
        for params in OUTLET_PARAMETERS:
            for dt in OUTLET_STEPS:
                for q in OUTLET_FLOWS:
                    args = [kind, *params, dt, q]
                    r = subprocess.run([PROBE, "outlet", kind, *map(repr, args[1:])], capture_output=True, text=True)
                    rec = {"args": args}
                    if r.returncode == 0:
                        rec["out"] = [float(v) for v in r.stdout.split()]
                    else:
                        rec["error"] = r.stderr.strip().removeprefix("ref_probe: ")
                    out["evaluations"].append(rec)
    with tempfile.TemporaryDirectory() as tmp:
        casegen.write_case(casegen.workload("tube_c0"), tmp, steps=2, dt=1e-3)
        path = os.path.join(tmp, "simulation_config.json")
        base = json.loads(open(path).read())
        for cond in OUTLET_CONDITIONS:
            cfg = json.loads(json.dumps(base))
            for b in cfg["boundaries"]:
                if b["label"] == 2:
                    b["conditions"] = [cond]
            with open(path, "w") as f:
                json.dump(cfg, f)
            r = subprocess.run([PROBE, "outlet-config", path], capture_output=True, text=True)
            rec = {"condition": cond}
            if r.returncode == 0:
                lines = r.stdout.split("\n")
                count = int(lines[0])
                rec["models"] = [[float(v) for v in line.split()] for line in lines[1:1 + count]]
                rec["tractions"] = [[float(v) for v in line.split()] for line in lines[1 + count:] if line]
            else:
                rec["error"] = r.stderr.strip().removeprefix("ref_probe: ")
            out["configs"].append(rec)
    with open(OUTLETS, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUTLETS, os.path.getsize(OUTLETS), "bytes")


def transport_vectors():
    from tubularflowiga_amd import casegen
    case = casegen.workload("tube_c0")
    out = []
    with tempfile.TemporaryDirectory() as tmp:
        for name, edit, config in TRANSPORT_VARIANTS:
            d = os.path.join(tmp, name)
            params = dict(TRANSPORT_PARAMETERS)
            casegen.write_transport_case(case, d, params, 5, case_config=config)
            path = os.path.join(d, "simulation_parameter.txt")
            text = open(path).read()
            if edit == "artificial_diffusion_weight":
                text = text.replace("artificial_diffusion ", "artificial_diffusion_weight ")
            elif edit == "drop_kplus":
                text = "\n".join(line for line in text.splitlines() if not line.startswith("kplus ")) + "\n"
            open(path, "w").write(text)
            r = subprocess.run([PROBE, "transport-case", d, str(case.mesh.nodes)], capture_output=True, text=True)
            rec = {"name": name, "edit": edit, "config": config}
            if r.returncode == 0:
                v = r.stdout.split()
                rec["parameters"] = [float(x) for x in v[:13]]
                rec["transport_nodes"], rec["constrained_index_sum"] = int(v[13]), int(v[14])
                rec["n0_sum"], rec["nplus_sum"] = float(v[15]), float(v[16])
                rec["velocity_nodes"], rec["pressure_nodes"] = int(v[17]), int(v[18])
            else:
                rec["error"] = r.stderr.strip().removeprefix("ref_probe: ")
            out.append(rec)
    with open(TRANSPORT, "w") as f:
        json.dump({"parameters": TRANSPORT_PARAMETERS, "variants": out}, f, indent=0)
    print("wrote", TRANSPORT, os.path.getsize(TRANSPORT), "bytes")


def main():
    if not os.path.exists(PROBE):
        raise SystemExit("build the probe first: make -C oracle ref")
    from tubularflowiga_amd import boundary, casegen
    from tubularflowiga_amd.mesh import write_ntiga
    rng = np.random.default_rng(20261202)
    points = [(0.2, 0.3, 0.4), (0.21, 0.37, 0.64), (0.9, 0.05, 0.5)]
    result = {"points": points, "elements": {}}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (conn, ext, bez) in elements().items():
            state = rng.standard_normal((len(conn), 4))
            path = os.path.join(tmp, name + ".txt")
            element_file(path, conn, ext, bez, state)
            rec = {"conn": [int(c) for c in conn], "bezier": np.asarray(bez).tolist(), "state": state.tolist(),
                   "extraction_nonzeros": [[int(a), int(p), float(ext[a, p])] for a, p in zip(*np.nonzero(ext))],
                   "basis": [], "flow": [], "traction": [], "surface": []}
            for (u, v, w) in points:
                vals = probe("basis", path, u, v, w)
                rec["basis"].append({"raw_determinant": vals[0], "inverse": vals[1:10],
                                     "value_gradient": vals[10:]})
            for face in range(6):
                rec["flow"].append(probe("flow", path, face)[0])
                rec["traction"].append(probe("traction", path, face, 1.7))
                rec["surface"].append(probe("surface", path, face))
            result["elements"][name] = rec
        # database round trip through the reference reader
        case = casegen.workload("tube_c0")
        db = os.path.join(tmp, "case.ntiga")
        write_ntiga(db, case.mesh, version=5)
        dump = subprocess.run([PROBE, "ntiga", db], check=True, capture_output=True, text=True).stdout
        result["tube_c0_ntiga_dump_sha256"] = hashlib.sha256(dump.encode()).hexdigest()
        result["tube_c0_ntiga_dump_head"] = dump.splitlines()[:2]
        casegen.write_case(case, tmp)
        trace = subprocess.run([PROBE, "walltrace", os.path.join(tmp, "case.ntiga"),
                                os.path.join(tmp, "controlmesh.vtk")], check=True, capture_output=True,
                               text=True).stdout
        result["tube_c0_wall_trace"] = [int(v) for v in trace.split()]
    with open(OUT, "w") as f:
        json.dump(result, f)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    outlet_vectors()
    transport_vectors()


if __name__ == "__main__":
    main()
