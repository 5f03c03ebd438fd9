"""Case generator, `.ntiga` I/O and boundary resolution (CPU only)."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tubularflowiga_amd import boundary, casegen
from tubularflowiga_amd.mesh import read_ntiga, write_ntiga

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def tube():
    return casegen.workload("tube_c0")


def _g(v):
    return "%.17g" % v


def reference_reader_dump(path):
    """The text the reference reader's probe prints for a database
    (oracle/ref_probe.cpp "ntiga"), rebuilt from our own reader."""
    m = read_ntiga(path)
    lines = [f"version 5 ranks 1 elements {m.elements} nodes {m.nodes}"]
    (o, u, s) = m.geometry_transform
    lines.append("transform " + " ".join(_g(v) for v in (*o, u, s)))
    for e in range(m.elements):
        conn = m.element_connectivity(e)
        lab = m.boundary_labels[6 * e:6 * e + 6]
        lines.append(f"element {int(m.element_ids[e])} type {int(m.element_types[e])} owner 0 nen {len(conn)} labels"
                     + "".join(f" {int(v)}" for v in lab))
        lines.append("conn" + "".join(f" {int(c)}" for c in conn))
        C = m.dense_extraction(e)
        lines.append("extraction" + "".join(f" {p}:{_g(C[a, p])}" for a in range(len(conn))
                                            for p in range(64) if C[a, p] != 0.0))
        lines.append("bezier" + "".join(f" {_g(v)}" for v in m.element_bezier(e).ravel()))
    lines.append("rank 0 required" + "".join(f" {e}" for e in range(m.elements)))
    return "\n".join(lines) + "\n"


def test_ntiga_written_file_reads_as_the_reference_reader_reads_it(tube, golden, tmp_path):
    """Our writer's file, read by our reader, prints exactly what the
    reference's IgaDatabase reader printed for the same file."""
    db = tmp_path / "case.ntiga"
    write_ntiga(db, tube.mesh, version=5)
    dump = reference_reader_dump(db)
    assert dump.splitlines()[:2] == golden["tube_c0_ntiga_dump_head"]
    assert hashlib.sha256(dump.encode()).hexdigest() == golden["tube_c0_ntiga_dump_sha256"]


@pytest.mark.parametrize("version", [3, 4, 5])
def test_ntiga_round_trip(tube, tmp_path, version):
    db = tmp_path / f"v{version}.ntiga"
    write_ntiga(db, tube.mesh, version=version)
    m = read_ntiga(db)
    a = tube.mesh
    for name in ("element_offsets", "connectivity", "extraction_offsets", "extraction_columns",
                 "extraction_values", "bezier_points"):
        assert np.array_equal(getattr(m, name), getattr(a, name)), name
    if version >= 4:
        assert np.array_equal(m.boundary_labels, a.boundary_labels)
    else:
        assert np.all(m.boundary_labels == -1)


def test_ntiga_rejects_corrupt_input(tube, tmp_path):
    db = tmp_path / "bad.ntiga"
    write_ntiga(db, tube.mesh)
    data = bytearray(db.read_bytes())
    data[0:8] = b"NOTADB00"
    db.write_bytes(bytes(data))
    with pytest.raises(ValueError):
        read_ntiga(db)


def test_wall_trace_matches_reference(tube, golden):
    trace = boundary.wall_trace_basis(tube.mesh, tube.cells, tube.point_labels)
    assert sorted(trace.tolist()) == sorted(golden["tube_c0_wall_trace"])


GAUSS4 = np.array([0.06943184420297371, 0.33000947820757187, 0.6699905217924281, 0.9305681557970262])


def _volume(m):
    _, _, _, w = O.reference_tables()
    pts = np.array([[x, y, z] for z in GAUSS4 for y in GAUSS4 for x in GAUSS4])
    return sum(wq / 8.0 * O.evaluate_basis(m.element_bezier(e), m.dense_extraction(e), *q,
                                           hessian=False)["raw_determinant"]
               for e in range(m.elements) for q, wq in zip(pts, w))


@pytest.mark.parametrize("name", ["tube_c0", "ybif_small"])
def test_generated_geometry_is_valid(name):
    """Partition of unity, positive Jacobians, and the labelled faces close the
    domain: (1/3) sum over labelled faces of the flux of u = x equals the volume
    (divergence theorem; the mesh is isoparametric, so u = x exactly)."""
    case = casegen.workload(name)
    m = case.mesh
    m.validate()
    rng = np.random.default_rng(1)
    for e in range(m.elements):
        C, B = m.dense_extraction(e), m.element_bezier(e)
        for p in rng.random((2, 3)):
            b = O.evaluate_basis(B, C, *p, hessian=False)
            assert abs(b["value"].sum() - 1.0) < 1e-12
            assert b["raw_determinant"] > 0.0
    st = np.zeros((m.nodes, 4))
    st[:, :3] = case.control_points
    e_, f_, _ = boundary.boundary_faces(m)
    surface = sum(O.boundary_flow(m.element_bezier(e), m.dense_extraction(e), st[m.element_connectivity(e)], int(f))
                  for e, f in zip(e_, f_)) / 3.0
    vol = _volume(m)
    assert abs(surface - vol) <= 1e-10 * vol


def test_tube_volume_converges_to_cylinder():
    """Quadratic B-spline sections through points on the circle cut the disc;
    the deficit shrinks under refinement (pi R^2 L, R = 1 mm, L = 10 mm)."""
    ratios = []
    for nc, mr in ((1, 1), (2, 1), (4, 2)):
        case = casegen.make_case("tube", radius_mm=1.0, length_mm=10.0, n_core=nc, m_ring=mr, n_axial=2, u_max=0.1)
        L0 = case.length_unit_m * 1e3
        ratios.append(_volume(case.mesh) / (np.pi * (1.0 / L0) ** 2 * (10.0 / L0)))
    assert ratios[0] < ratios[1] < ratios[2] < 1.0
    assert ratios[2] > 0.96


def test_boundary_resolution(tube):
    vmask, pmask, vt, pt = boundary.resolve_flow_boundaries(tube)
    lab = tube.point_labels
    assert np.all(vmask[lab == 0] == 1) and np.all(vt[lab == 0] == 0.0)
    assert np.all(vmask[lab == 1] == 1)
    assert not np.any(pmask)
    # outlet interior nodes are free
    assert np.any(vmask[lab == 2] == 0)
    s = boundary.initial_state(vmask, pmask, vt, pt).reshape(-1, 4)
    assert np.array_equal(s[vmask == 1, :3], vt[vmask == 1])
    # inlet flow of the prescribed profile is an inflow no larger than the exact
    # Poiseuille flow u_max/2 * pi R^2 (the coarse section cuts the disc and
    # the wall trace pins the rim to zero)
    e, f, labels = boundary.boundary_faces(tube.mesh)
    state = s
    q_in = sum(O.boundary_flow(tube.mesh.element_bezier(ei), tube.mesh.dense_extraction(ei),
                               state[tube.mesh.element_connectivity(ei)], int(fi))
               for ei, fi, li in zip(e, f, labels) if li == 1)
    L0 = tube.length_unit_m
    area = np.pi * (1e-3 / L0) ** 2
    u_max = 0.1 / L0
    assert -0.5 * u_max * area < q_in < -0.5 * 0.5 * u_max * area


def test_write_case_files(tube, tmp_path):
    d = casegen.write_case(tube, tmp_path / "case", steps=3)
    cfg = json.loads((d / "simulation_config.json").read_text())
    assert cfg["schema_version"] == 3
    assert cfg["time"]["steps"] == 3
    es = cfg["equation_systems"][0]
    rho, mu = casegen.flow_parameters(tube)
    assert es["density"] == rho and es["viscosity"] == mu
    assert {b["label"] for b in cfg["boundaries"]} == {0, 1, 2}
    prof = np.loadtxt(d / "initial_velocityfield.txt")
    assert prof.shape == (tube.mesh.nodes, 3)
    text = (d / "controlmesh.vtk").read_text()
    assert f"POINTS {tube.mesh.nodes} double" in text
