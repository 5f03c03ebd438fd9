"""tfi_pack, the `.ntiga` v5 packer (host code, CPU only).

* Round trip: a synthetic case directory (a structured hexahedral control
  grid with interior -1 / wall 0 / tip 1, 2 point labels, random connectivity,
  extraction rows with dropped zeros and Bezier points) is packed and read
  back; every field, the boundary face labels (lateral faces 0, bottom 1, top
  2, shared faces -1), the geometry transform (bounding-box origin, smallest
  extent, schema-4 length scale), the owners and the per-rank element lists
  (block distribution of nodes) are checked against what the case holds.
* Against the reference packer (solvers/cpu/src/iga_pack.cpp, built from its
  own source into oracle/_ref/iga_pack where the reference tree exists): the
  same database bytes, stdout and exit status for the legacy text input, the
  sparse cache, the forced legacy path and several rank counts, and the same
  error message and status for malformed inputs.
"""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tubularflowiga_amd import mesh as M

ROOT = Path(__file__).resolve().parents[1]
PACK = ROOT / "tubularflowiga_amd" / "tfi_pack"
REF_PACK = ROOT / "oracle" / "_ref" / "iga_pack"

MESH_BLOCK = {
    "smoothing": {"iterations": 2, "bifurcation_ratio": 0.5, "noise_ratio": 0.1},
    "centerline": {"target_spacing": 1.0, "max_spacing_over_diameter": 1.0, "max_turn_degrees": 12,
                   "max_diameter_change_fraction": 0.15, "maximum_curvature_radius_product": 0.8},
    "junction": {"max_spacing_over_diameter": 0.25, "upstream_clearance_over_diameter": 1.0,
                 "downstream_clearance_over_diameter": 1.5, "minimum_angle_degrees": 10,
                 "maximum_radius_ratio": 8, "optimization_iterations": 4},
    "quality": {"minimum_scaled_jacobian": 0.1, "check_self_intersection": True, "collision_safety_factor": 1.0},
}


def _config(length_scale=1e-3, **override):
    import copy
    cfg = {"schema_version": 4, "dimension": "3d",
           "geometry": {"kind": "swc_network", "file": "tree.swc", "length_scale_to_m": length_scale},
           "mesh": copy.deepcopy(MESH_BLOCK)}
    for path, value in override.items():
        node = cfg
        keys = path.split("__")
        for k in keys[:-1]:
            node = node[k]
        if value is None:
            node.pop(keys[-1])
        else:
            node[keys[-1]] = value
    return cfg


class Grid:
    """nx x ny x nz hexahedral control grid with the reference generator's
    label convention (MeshGenerator.cpp: -1 interior, 0 on the wall circle,
    positive tip labels on the end sections)."""

    def __init__(self, nx=2, ny=3, nz=4, nen=27, seed=0):
        rng = np.random.default_rng(seed)
        self.nx, self.ny, self.nz, self.nen = nx, ny, nz, nen
        i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
        i, j, k = i.transpose(2, 1, 0).ravel(), j.transpose(2, 1, 0).ravel(), k.transpose(2, 1, 0).ravel()
        self.nodes = len(i)
        self.xyz = np.stack([0.7 * i + 0.05 * np.sin(k), 0.9 * j - 0.03 * k, 1.3 * k + 0.02 * i], 1) + 10.0
        lateral = (i == 0) | (i == nx) | (j == 0) | (j == ny)
        self.label = np.full(self.nodes, -1, np.int32)
        self.label[(k == 0) & ~lateral] = 1
        self.label[(k == nz) & ~lateral] = 2
        self.label[lateral] = 0
        nid = lambda a, b, c: a + (nx + 1) * (b + (ny + 1) * c)  # noqa: E731
        cells, expect = [], []
        for c in range(nz):
            for b in range(ny):
                for a in range(nx):
                    cells.append([nid(a, b, c), nid(a + 1, b, c), nid(a + 1, b + 1, c), nid(a, b + 1, c),
                                  nid(a, b, c + 1), nid(a + 1, b, c + 1), nid(a + 1, b + 1, c + 1),
                                  nid(a, b + 1, c + 1)])
                    # faces: bottom (z-), front (y-), right (x+), back (y+), left (x-), top (z+)
                    expect.append([1 if c == 0 else -1, 0 if b == 0 else -1, 0 if a == nx - 1 else -1,
                                   0 if b == ny - 1 else -1, 0 if a == 0 else -1, 2 if c == nz - 1 else -1])
        self.cells = np.array(cells, np.int32)
        self.face_labels = np.array(expect, np.int32)
        E = len(cells)
        self.E = E
        self.conn = rng.integers(0, self.nodes, (E, nen)).astype(np.int32)
        rows = rng.standard_normal((E, nen, 64)) * 10.0 ** rng.integers(-3, 3, (E, nen, 64))
        rows[rng.random((E, nen, 64)) < 0.6] = 0.0
        rows[rng.random((E, nen, 64)) < 0.02] = -0.0
        self.rows = rows
        self.types = rng.integers(-3, 6, E).astype(np.int32)
        self.bezier = rng.standard_normal((E, 64, 3))

    def write(self, d: Path, legacy=True, cache=False, ranks=(1,), config=None, seed=1):
        d.mkdir(parents=True, exist_ok=True)
        lines = ["# vtk DataFile Version 3.0", "control mesh", "ASCII", "DATASET UNSTRUCTURED_GRID",
                 f"POINTS {self.nodes} double"]
        lines += [" ".join(repr(float(x)) for x in p) for p in self.xyz]
        lines.append(f"CELLS {self.E} {9 * self.E}")
        lines += ["8 " + " ".join(map(str, c)) for c in self.cells]
        lines.append(f"CELL_TYPES {self.E}")
        lines += ["12"] * self.E
        lines += [f"POINT_DATA {self.nodes}", "SCALARS label float 1", "LOOKUP_TABLE default"]
        lines += [str(v) for v in self.label]
        (d / "controlmesh.vtk").write_text("\n".join(lines) + "\n")
        if legacy:
            out = [str(self.E)]
            for e in range(self.E):
                out.append(f"{e} {self.nen} {self.types[e]}")
                out.append(" ".join(map(str, self.conn[e])))
                out += [" ".join(repr(float(x)) for x in row) for row in self.rows[e]]
            (d / "cmat.txt").write_text("\n".join(out) + "\n")
            bz = [str(64 * self.E)] + [" ".join(repr(float(x)) for x in p) for e in range(self.E) for p in self.bezier[e]]
            (d / "bzpt.txt").write_text("\n".join(bz) + "\n")
        if cache:
            (d / "spline_cache.igacache").write_bytes(self.cache_bytes(d))
        rng = np.random.default_rng(seed)
        for r in ranks:
            owners = rng.integers(0, r, self.E)
            (d / f"bzmeshinfo.txt.epart.{r}").write_text("\n".join(map(str, owners)) + "\n")
        if config is not None:
            import json
            (d / "simulation_config.json").write_text(json.dumps(config, indent=1))
        return d

    def cache_bytes(self, d: Path):
        h = 14695981039346656037
        for byte in (d / "controlmesh.vtk").read_bytes():
            h = ((h ^ byte) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        parts = [b"IGACACH1", struct.pack("<IQQQII", 1, self.E, self.nodes, h, 64, 0)]
        for e in range(self.E):
            parts.append(struct.pack("<QiI", e, int(self.types[e]), self.nen))
            parts.append(self.conn[e].astype("<i4").tobytes())
            for row in self.rows[e]:
                nz = [c for c in range(64) if row[c] != 0.0]
                parts.append(struct.pack("<B", len(nz)))
                for c in nz:
                    parts.append(struct.pack("<Bd", c, row[c]))
            parts.append(self.bezier[e].astype("<f8").tobytes())
        return b"".join(parts)


def _run(exe, d, ranks, out, *extra):
    return subprocess.run([str(exe), str(d), str(ranks), str(out), *extra], capture_output=True, text=True, timeout=120)


@pytest.fixture(scope="module")
def pack():
    if not PACK.exists():
        pytest.fail("tfi_pack is not built (python -m tubularflowiga_amd.build)")
    return PACK


def _owners_and_rank_lists(path):
    data = Path(path).read_bytes()
    version, ranks, E, N, bez, _r, rank_off = struct.unpack_from("<IIQQIIQ", data, 8)
    pos = 88
    owners = np.frombuffer(data, "<i4", E, pos + 8 * (E + 1))
    ro = np.frombuffer(data, "<u8", ranks + 1, rank_off)
    lists = np.frombuffer(data, "<u8", int(ro[-1]), rank_off + 8 * (ranks + 1))
    return version, ranks, owners, [lists[ro[r]:ro[r + 1]].tolist() for r in range(ranks)]


@pytest.mark.parametrize("ranks,cache", [(1, False), (3, False), (3, True)])
def test_pack_round_trip(tmp_path, pack, ranks, cache):
    g = Grid(seed=ranks)
    d = g.write(tmp_path / "case", legacy=not cache, cache=cache, ranks=(ranks,), config=_config(2.5e-4))
    out = tmp_path / "case.ntiga"
    r = _run(pack, d, ranks, out)
    assert r.returncode == 0, r.stderr
    assert ("sparse preprocessing cache" in r.stdout) == cache
    assert f"with {g.E} validated elements for {ranks} ranks" in r.stdout
    m = M.read_ntiga(out)
    assert m.nodes == g.nodes and m.elements == g.E
    assert np.array_equal(m.connectivity.reshape(g.E, g.nen), g.conn)
    assert np.array_equal(m.element_types, g.types)
    assert np.array_equal(m.bezier_points.reshape(g.E, 64, 3), g.bezier)
    assert np.array_equal(m.boundary_labels.reshape(g.E, 6), g.face_labels)
    dense = np.zeros((g.E * g.nen, 64))
    for row in range(g.E * g.nen):
        lo, hi = m.extraction_offsets[row], m.extraction_offsets[row + 1]
        dense[row, m.extraction_columns[lo:hi]] = m.extraction_values[lo:hi]
    assert np.array_equal(dense, g.rows.reshape(-1, 64))
    assert np.count_nonzero(m.extraction_values == 0.0) == 0
    origin, units, scale = m.geometry_transform
    assert np.array_equal(origin, g.xyz.min(0)) and units == (g.xyz.max(0) - g.xyz.min(0)).min()
    assert scale == 2.5e-4
    version, nranks, owners, lists = _owners_and_rank_lists(out)
    assert version == 5 and nranks == ranks
    want_owner = np.array((d / f"bzmeshinfo.txt.epart.{ranks}").read_text().split(), np.int32)
    assert np.array_equal(owners, want_owner)
    q, rem = divmod(g.nodes, ranks)
    node_rank = np.concatenate([np.full(q + 1, r) for r in range(rem)] + [np.full(q, r) for r in range(rem, ranks)])
    for r in range(ranks):
        assert lists[r] == [e for e in range(g.E) if (node_rank[g.conn[e]] == r).any()]


def _ref_or_skip():
    if not REF_PACK.exists():
        pytest.skip("reference packer not built (needs the reference tree: make -C oracle ref-pack)")
    return REF_PACK


@pytest.mark.parametrize("variant", ["legacy1", "legacy3", "cache3", "forced_legacy", "config", "config_v3", "nen64"])
def test_pack_bytes_match_reference(tmp_path, pack, variant):
    ref = _ref_or_skip()
    g = Grid(nen=64 if variant == "nen64" else 27, seed=7)
    ranks = 1 if variant in ("legacy1", "nen64") else 3
    d = g.write(tmp_path / "case", legacy=True, cache=variant in ("cache3", "forced_legacy"), ranks=(ranks,),
                config={"config": _config(1e-3), "config_v3": dict(_config(), schema_version=3, mesh=7)}.get(variant))
    extra = ["--legacy-text"] if variant == "forced_legacy" else []
    mine, theirs = _run(pack, d, ranks, tmp_path / "mine.ntiga", *extra), _run(ref, d, ranks, tmp_path / "ref.ntiga", *extra)
    assert theirs.returncode == 0, theirs.stderr
    assert mine.returncode == 0, mine.stderr
    assert mine.stdout.replace("mine.ntiga", "X") == theirs.stdout.replace("ref.ntiga", "X")
    assert (tmp_path / "mine.ntiga").read_bytes() == (tmp_path / "ref.ntiga").read_bytes()


def _corrupt(case: Path, name: str):
    """Mutations of a good legacy + cache case, one per error path."""
    def edit(fname, fn):
        p = case / fname
        p.write_text(fn(p.read_text()))
    if name == "truncated_cmat":
        edit("cmat.txt", lambda t: t[: len(t) // 2])
    elif name == "extra_bzpt":
        edit("bzpt.txt", lambda t: t + "1.0\n")
    elif name == "ids":
        edit("cmat.txt", lambda t: t.replace("\n1 27 ", "\n5 27 ", 1))
    elif name == "bad_partition":
        edit("bzmeshinfo.txt.epart.3", lambda t: "7\n" + t.split("\n", 1)[1])
    elif name == "short_partition":
        edit("bzmeshinfo.txt.epart.3", lambda t: t.split("\n", 1)[1])
    elif name == "no_partition":
        (case / "bzmeshinfo.txt.epart.3").unlink()
    elif name == "cell_count":
        edit("cmat.txt", lambda t: "23" + t[t.index("\n"):])
    elif name == "point_count":
        edit("bzpt.txt", lambda t: "5" + t[t.index("\n"):])
    elif name == "mixed_tips":
        edit("controlmesh.vtk", lambda t: t.replace("\n1\n", "\n3\n", 1))
    elif name == "unlabelled_face":
        lines = (case / "controlmesh.vtk").read_text().split("\n")
        i = lines.index("LOOKUP_TABLE default") + 1
        lines[i] = "-1"   # a wall corner of the bottom section (node 0)
        (case / "controlmesh.vtk").write_text("\n".join(lines))
    elif name == "not_hex":
        edit("controlmesh.vtk", lambda t: t.replace("\n8 ", "\n7 ", 1))
    elif name == "labels_missing":
        edit("controlmesh.vtk", lambda t: t[: t.index("POINT_DATA")])
    elif name == "conn_range":
        edit("cmat.txt", lambda t: t.replace("\n0 27 ", "\n0 27 ", 1).split("\n", 3)[0] + "\n"
             + t.split("\n", 3)[1] + "\n99999 " + t.split("\n", 3)[2].split(" ", 1)[1] + "\n" + t.split("\n", 3)[3])
    elif name == "cache_hash":
        edit("controlmesh.vtk", lambda t: t.replace("control mesh", "control mesh edited"))
    elif name == "cache_truncated":
        p = case / "spline_cache.igacache"
        p.write_bytes(p.read_bytes()[:-5])
    elif name.startswith("config_"):
        import json
        bad = {"config_unknown": _config(geometry__zeta=1, geometry__alpha=2), "config_scale": _config(0.0),
               "config_kind": _config(geometry__kind="mesh"), "config_file": _config(geometry__file="sub/t.swc"),
               "config_turn": _config(mesh__centerline__max_turn_degrees=95),
               "config_iters": _config(mesh__junction__optimization_iterations=2.5),
               "config_missing": _config(mesh__quality=None),
               "config_bool": _config(mesh__quality__check_self_intersection=1),
               "config_section": _config(mesh__cross_section={"kind": "circle", "target_size": 2}),
               "config_dimension": _config(dimension="2d")}[name]
        (case / "simulation_config.json").write_text(json.dumps(bad))
    else:
        raise KeyError(name)


ERRORS = ["truncated_cmat", "extra_bzpt", "ids", "bad_partition", "short_partition", "no_partition", "cell_count",
          "point_count", "mixed_tips", "unlabelled_face", "not_hex", "labels_missing", "conn_range", "cache_hash",
          "cache_truncated", "config_unknown", "config_scale", "config_kind", "config_file", "config_turn",
          "config_iters", "config_missing", "config_bool", "config_section", "config_dimension"]


@pytest.mark.parametrize("name", ERRORS)
def test_pack_errors_match_reference(tmp_path, pack, name):
    ref = _ref_or_skip()
    g = Grid(seed=11)
    cache = name.startswith("cache_")
    d = g.write(tmp_path / "case", legacy=True, cache=cache, ranks=(3,))
    _corrupt(d, name)
    mine, theirs = _run(pack, d, 3, tmp_path / "mine.ntiga"), _run(ref, d, 3, tmp_path / "ref.ntiga")
    print(name, "|", theirs.stderr.strip().splitlines()[-1] if theirs.stderr.strip() else "", flush=True)
    assert theirs.returncode == 1, f"{name}: the reference accepted the case"
    assert mine.returncode == theirs.returncode
    assert mine.stderr.replace("tfi_pack: ", "") == theirs.stderr.replace("iga_pack: ", "")


@pytest.mark.parametrize("args", [[], ["a", "b"], ["a", "1", "o", "--bogus"], ["a", "x", "o"], ["a", "0", "o"]])
def test_pack_usage_matches_reference(tmp_path, pack, args):
    ref = _ref_or_skip()
    mine = subprocess.run([str(pack), *args], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    theirs = subprocess.run([str(ref), *args], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert mine.returncode == theirs.returncode
    norm = lambda s: s.replace("tfi_pack", "P").replace("iga_pack", "P")  # noqa: E731
    assert norm(mine.stderr) == norm(theirs.stderr)


@pytest.mark.parametrize("name", ["tube_c0", "tree_small"])
def test_pack_of_a_generated_case_is_its_database(tmp_path, pack, name):
    """A generated case's spline data written as cmat.txt / bzpt.txt beside its
    control mesh packs into the case's own database, byte for byte apart from
    the 40-byte geometry transform (the control mesh holds normalised points,
    so the packer's bounding-box transform is origin 0, extent 1): the face
    labels the packer derives from the control cells and point labels are
    the generator's, element by element."""
    from tubularflowiga_amd import casegen
    case = casegen.workload(name)
    d = casegen.write_case(case, tmp_path / "case", steps=1)
    m = case.mesh
    E = m.elements
    out = [str(E)]
    for e in range(E):
        conn = m.element_connectivity(e)
        out.append(f"{e} {len(conn)} 0")
        out.append(" ".join(map(str, conn)))
        out += [" ".join(repr(float(x)) for x in row) for row in m.dense_extraction(e)]
    (d / "cmat.txt").write_text("\n".join(out) + "\n")
    bz = m.bezier_points.reshape(-1, 3)
    (d / "bzpt.txt").write_text(f"{64 * E}\n" + "\n".join(" ".join(repr(float(x)) for x in p) for p in bz) + "\n")
    (d / "bzmeshinfo.txt.epart.1").write_text("0\n" * E)
    r = _run(pack, d, 1, tmp_path / "packed.ntiga")
    assert r.returncode == 0, r.stderr
    mine, gen = (tmp_path / "packed.ntiga").read_bytes(), (d / "case.ntiga").read_bytes()
    assert len(mine) == len(gen) and mine[:48] == gen[:48] and mine[88:] == gen[88:]
    lo = case.control_points.min(0)
    assert np.array_equal(np.frombuffer(mine[48:88], "<f8"),
                          [*lo, (case.control_points.max(0) - lo).min(), 1.0])
    # the drop-in's host input path reads the packed database as it reads the
    # generator's: the same boundary data, wall trace and flow records
    # (also packed for three ranks: the single-GPU driver reads every element)
    (d / "bzmeshinfo.txt.epart.3").write_text("".join(f"{e % 3}\n" for e in range(E)))
    r = _run(pack, d, 3, tmp_path / "packed3.ntiga")
    assert r.returncode == 0, r.stderr
    cli = ROOT / "tubularflowiga_amd" / "iga_hip"
    checks = [subprocess.run([str(cli), "case-check", str(db), str(d)], capture_output=True, text=True, timeout=120)
              for db in (tmp_path / "packed.ntiga", tmp_path / "packed3.ntiga", d / "case.ntiga")]
    assert all(c.returncode == 0 for c in checks), [c.stderr for c in checks]
    assert checks[0].stdout == checks[2].stdout and checks[1].stdout == checks[2].stdout
    assert "case_check nodes=" in checks[0].stdout
