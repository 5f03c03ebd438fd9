"""Writes tests/golden/bezier_reference.json: the reference's Bezier
geometry report and the digest of its temporal VTKHDF file (h5dump text,
output directory replaced by "X") for two generated cases with two seeded
states each.  Runs oracle/_ref/ref_bezier (`make -C oracle ref-bezier`: our
driver around include/BezierVisualization.hpp and include/TemporalVtkHdf.hpp
of the reference, built against this image's HDF5).  Test infrastructure:
tests/test_bezier_output.py checks the drop-in's writer against this file."""
import hashlib
import json
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from tubularflowiga_amd import casegen  # noqa: E402

CASES = ("tube_c0", "tree_small")
H5DUMP = "/opt/conda/bin/h5dump"


def write_inputs(name, directory):
    """The case files and two seeded states ("vx vy vz p" per control point)."""
    import numpy as np
    case = casegen.workload(name)
    casegen.write_case(case, directory)
    rng = np.random.default_rng(20261203)
    states = []
    for k in range(2):
        f = Path(directory) / f"state{k}.txt"
        np.savetxt(f, rng.standard_normal((case.mesh.nodes, 4)), fmt="%.17g")
        states.append(f)
    return Path(directory) / "case.ntiga", states


def digest(directory):
    """(report text, sha256 of the h5dump text with the directory replaced by X)."""
    d = Path(directory)
    dump = subprocess.run([H5DUMP, str(d / "flow.vtkhdf")], capture_output=True, text=True, check=True).stdout
    return (d / "flow.bezier_geometry.json").read_text(), hashlib.sha256(dump.replace(str(d), "X").encode()).hexdigest()


def main():
    probe = ROOT / "oracle" / "_ref" / "ref_bezier"
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref-bezier"], check=True)
    out = {}
    for name in CASES:
        with tempfile.TemporaryDirectory() as tmp:
            db, states = write_inputs(name, Path(tmp) / "case")
            run = Path(tmp) / "ref"
            run.mkdir()
            r = subprocess.run([str(probe), str(db), str(run / "flow.txt"), *map(str, states)], capture_output=True,
                               text=True, check=True)
            report, sha = digest(run)
            out[name] = {"probe_stdout": r.stdout.strip(), "report": report, "h5dump_sha256": sha}
    (ROOT / "tests" / "golden" / "bezier_reference.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
