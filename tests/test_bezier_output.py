"""The drop-in driver's Bezier VTKHDF output (tubularflowiga_amd/host/
bezier_output.hpp, through tests/bezier_host_check.cpp) on the CPU.

* Against the reference: for the CPU-runnable tube and the coarse vessel tree,
  with two seeded states each, the geometry report is the reference's text
  and the VTKHDF file dumps (h5dump: every group, dataset, attribute and
  value) to the reference's text (tests/golden/bezier_reference.json, written
  by tests/golden/make_bezier_reference.py through the reference's
  include/BezierVisualization.hpp and include/TemporalVtkHdf.hpp).
* Closed forms: the spline space reproduces linear fields, so control values
  equal to the control points' coordinates give, at every merged Bezier
  point, that point's coordinates; the cell connectivity lists the 64 points
  of each element once, corners first.
"""
import hashlib
import json
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tubularflowiga_amd import build, casegen

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = json.loads((ROOT / "tests" / "golden" / "bezier_reference.json").read_text())
H5DUMP = "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not build.hdf5_flags() or not Path(H5DUMP).exists(),
                                reason="HDF5 library or h5dump not in this image")


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    exe = tmp_path_factory.mktemp("bezier") / "bezier_host_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", str(ROOT / "tubularflowiga_amd" / "host"),
                    str(ROOT / "tests" / "bezier_host_check.cpp"), "-o", str(exe), "-pthread", *build.hdf5_flags()],
                   check=True)
    return exe


def _inputs(name, directory, states):
    case = casegen.workload(name)
    casegen.write_case(case, directory)
    files = []
    for k, s in enumerate(states):
        f = Path(directory) / f"state{k}.txt"
        np.savetxt(f, s, fmt="%.17g")
        files.append(f)
    return case, Path(directory) / "case.ntiga", files


def _dump(path, *flags):
    return subprocess.run([H5DUMP, *flags, str(path)], capture_output=True, text=True, check=True).stdout


@pytest.mark.parametrize("name", ["tube_c0", "tree_small"])
def test_report_and_file_match_the_reference(host_check, tmp_path, name):
    case = casegen.workload(name)
    rng = np.random.default_rng(20261203)
    states = [rng.standard_normal((case.mesh.nodes, 4)) for _ in range(2)]
    _, db, files = _inputs(name, tmp_path / "case", states)
    run = tmp_path / "ours"
    run.mkdir()
    r = subprocess.run([str(host_check), str(db), str(run / "flow.txt"), *map(str, files)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    want = GOLDEN[name]
    assert r.stdout.strip() == want["probe_stdout"]
    assert (run / "flow.bezier_geometry.json").read_text() == want["report"]
    dump = _dump(run / "flow.vtkhdf").replace(str(run), "X")
    assert hashlib.sha256(dump.encode()).hexdigest() == want["h5dump_sha256"]


def _dataset(dump, name):
    """Numbers of one dataset of an h5dump text."""
    m = re.search(r'DATASET "' + re.escape(name) + r'" \{.*?DATA \{(.*?)\n\s*\}\n', dump, re.S)
    assert m, name
    body = re.sub(r"\(\d+(,\d+)*\):", " ", m.group(1))
    return np.array([float(x) for x in re.findall(r"-?\d+(?:\.\d+)?(?:e[-+]?\d+)?", body)])


def test_linear_field_is_reproduced_at_the_bezier_points(host_check, tmp_path):
    case = casegen.workload("tube_c0")
    cp = np.asarray(case.control_points)
    state = np.column_stack([cp, cp[:, 0] + 2.0 * cp[:, 1] - cp[:, 2]])
    _, db, files = _inputs("tube_c0", tmp_path / "case", [state, state])
    run = tmp_path / "ours"
    run.mkdir()
    assert subprocess.run([str(host_check), str(db), str(run / "flow.txt"), *map(str, files)]).returncode == 0
    dump = _dump(run / "flow.vtkhdf", "-m", "%.17g")
    points = _dataset(dump, "Points").reshape(-1, 3)
    lower, extent, _ = case.mesh.geometry_transform
    normalised = (points - np.asarray(lower)) / extent
    velocity = _dataset(dump, "velocity").reshape(-1, 3)
    pressure = _dataset(dump, "pressure")
    n = len(points)
    # two steps, stacked
    assert velocity.shape == (2 * n, 3) and pressure.shape == (2 * n,)
    for step in range(2):
        assert np.max(np.abs(velocity[step * n:(step + 1) * n] - normalised)) < 1e-12
        want = normalised[:, 0] + 2 * normalised[:, 1] - normalised[:, 2]
        assert np.max(np.abs(pressure[step * n:(step + 1) * n] - want)) < 1e-11
    conn = _dataset(dump, "Connectivity").astype(np.int64).reshape(-1, 64)
    assert all(len(set(row)) == 64 for row in conn)
    assert np.array_equal(_dataset(dump, "Types"), np.full(len(conn), 79))
    # corners first, in VTK's hexahedron vertex order (i, j, k) = (0,0,0) (3,0,0)
    # (3,3,0) (0,3,0), then the same at k = 3: the Bezier map interpolates its
    # corner control points, tensor index i + 4 j + 16 k
    corners = [i + 4 * j + 16 * k for k in (0, 3) for (i, j) in ((0, 0), (3, 0), (3, 3), (0, 3))]
    bez = case.mesh.bezier_points.reshape(-1, 64, 3)
    for e in range(len(conn)):
        assert np.max(np.abs(normalised[conn[e, :8]] - bez[e, corners])) < 1e-12


def test_resumed_series_appends_and_checks_the_geometry(host_check, tmp_path):
    """A restarted run opens the file and appends (TemporalVtkHdf.hpp:600-640):
    one state, then two more resumed, dumps like the three written at once; a
    time equal to the last one replaces that step; another geometry is refused."""
    case = casegen.workload("tube_c0")
    rng = np.random.default_rng(7)
    states = [rng.standard_normal((case.mesh.nodes, 4)) for _ in range(3)]
    _, db, files = _inputs("tube_c0", tmp_path / "case", states)
    whole, split = tmp_path / "whole", tmp_path / "split"
    whole.mkdir()
    split.mkdir()
    assert subprocess.run([str(host_check), str(db), str(whole / "flow.txt"), *map(str, files)]).returncode == 0
    assert subprocess.run([str(host_check), str(db), str(split / "flow.txt"), str(files[0])]).returncode == 0
    # resume at step 2 after a state written again at step 1: that step is replaced
    assert subprocess.run([str(host_check), "--resume", "1", str(db), str(split / "flow.txt"), str(files[0]),
                           str(files[1]), str(files[2])]).returncode == 0
    a = _dump(whole / "flow.vtkhdf").replace(str(whole), "X")
    b = _dump(split / "flow.vtkhdf").replace(str(split), "X")
    assert a == b
    other = casegen.make_case("tube", radius_mm=1.0, length_mm=10.0, n_core=1, m_ring=1, n_axial=12, u_max=0.1)
    casegen.write_case(other, tmp_path / "other")
    np.savetxt(tmp_path / "other" / "s.txt", np.zeros((other.mesh.nodes, 4)))
    r = subprocess.run([str(host_check), "--resume", "4", str(tmp_path / "other" / "case.ntiga"), str(split / "flow.txt"),
                        str(tmp_path / "other" / "s.txt")], capture_output=True, text=True)
    assert r.returncode != 0
    assert "existing VTKHDF geometry does not match the current .ntiga database" in r.stderr
