"""Flat IGA mesh arrays and the `.ntiga` v3/v4/v5 byte format.

``FlatMesh`` holds exactly the arrays the reference flattens a database into
(solvers/cuda/include/DeviceMesh.hpp:16-68): element offsets, connectivity,
sparse extraction rows (u8 Bezier column + f64 value) and 64 Bezier points per
element, plus the six per-face boundary labels of version >= 4.

The `.ntiga` layout follows solvers/cpu/include/IgaDatabase.hpp:64-161 (reader)
and solvers/cpu/src/iga_pack.cpp:338-409 (writer): a packed little-endian
header (magic ``NTIGADB2``, u32 version, u32 ranks, u64 elements, u64 nodes,
u32 Bezier points = 64, u32 reserved, u64 rank-index offset, and for v5 the
f64 source origin[3], units per normalised unit, length scale to metres), the
u64 element offsets[E+1] and i32 owners[E], the element records, then the rank
index (u64 rank offsets[R+1], u64 required elements).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

MAGIC = b"NTIGADB2"
BEZIER_POINTS = 64


@dataclass
class FlatMesh:
    nodes: int
    element_offsets: np.ndarray          # int32 [E+1]
    connectivity: np.ndarray             # int32 [sum nen]
    extraction_offsets: np.ndarray       # int32 [sum nen + 1]
    extraction_columns: np.ndarray       # uint8 [nnzE]
    extraction_values: np.ndarray        # float64 [nnzE]
    bezier_points: np.ndarray            # float64 [E*64*3]
    boundary_labels: np.ndarray          # int32 [E*6]
    element_ids: np.ndarray | None = None
    element_types: np.ndarray | None = None
    geometry_transform: tuple = ((0.0, 0.0, 0.0), 1.0, 1.0)
    extra: dict = field(default_factory=dict)

    @property
    def elements(self) -> int:
        return len(self.element_offsets) - 1

    @property
    def maximum_basis(self) -> int:
        return int(np.diff(self.element_offsets).max()) if self.elements else 0

    def element_nen(self, e: int) -> int:
        return int(self.element_offsets[e + 1] - self.element_offsets[e])

    def dense_extraction(self, e: int) -> np.ndarray:
        """Dense [nen, 64] extraction operator of element e."""
        b = self.element_offsets[e]
        nen = self.element_nen(e)
        out = np.zeros((nen, 64))
        for a in range(nen):
            lo, hi = self.extraction_offsets[b + a], self.extraction_offsets[b + a + 1]
            out[a, self.extraction_columns[lo:hi]] = self.extraction_values[lo:hi]
        return out

    def element_bezier(self, e: int) -> np.ndarray:
        return self.bezier_points[e * 192:(e + 1) * 192].reshape(64, 3)

    def element_connectivity(self, e: int) -> np.ndarray:
        return self.connectivity[self.element_offsets[e]:self.element_offsets[e + 1]]

    def validate(self) -> None:
        E = self.elements
        assert self.element_offsets.dtype == np.int32 and self.connectivity.dtype == np.int32
        assert self.element_offsets[0] == 0 and np.all(np.diff(self.element_offsets) > 0)
        assert len(self.connectivity) == self.element_offsets[-1]
        assert len(self.extraction_offsets) == len(self.connectivity) + 1
        assert len(self.extraction_columns) == len(self.extraction_values) == self.extraction_offsets[-1]
        assert len(self.bezier_points) == E * 192
        assert len(self.boundary_labels) == E * 6
        assert self.connectivity.min() >= 0 and self.connectivity.max() < self.nodes
        assert self.extraction_columns.max() < 64
        if self.maximum_basis > 80:
            raise ValueError("kernels support at most 80 basis functions per element")


def write_ntiga(path, mesh: FlatMesh, version: int = 5) -> None:
    """Write ``mesh`` as a one-rank `.ntiga` database (iga_pack.cpp:338-409).

    Extraction entries equal to 0.0 are not stored, as in the reference packer.
    Every element is owned by rank 0, and rank 0 requires every element.
    """
    if version not in (3, 4, 5):
        raise ValueError("version must be 3, 4 or 5")
    E = mesh.elements
    ranks = 1
    origin, units, length_scale = mesh.geometry_transform
    head = MAGIC + struct.pack("<IIQQII", version, ranks, E, mesh.nodes, BEZIER_POINTS, 0)
    # rank_index_offset placeholder, then the v5 transform
    header_len = len(head) + 8 + (40 if version >= 5 else 0)
    index_len = 8 * (E + 1) + 4 * E
    ids = mesh.element_ids if mesh.element_ids is not None else np.arange(E, dtype=np.uint64)
    types = mesh.element_types if mesh.element_types is not None else np.zeros(E, np.int32)
    records = []
    offsets = np.zeros(E + 1, np.uint64)
    pos = header_len + index_len
    eo, xo = mesh.element_offsets, mesh.extraction_offsets
    for e in range(E):
        offsets[e] = pos
        b, nen = int(eo[e]), int(eo[e + 1] - eo[e])
        parts = [struct.pack("<QiiI", int(ids[e]), int(types[e]), 0, nen)]
        if version >= 4:
            parts.append(mesh.boundary_labels[6 * e:6 * e + 6].astype("<i4").tobytes())
        parts.append(mesh.connectivity[b:b + nen].astype("<i4").tobytes())
        for a in range(nen):
            lo, hi = int(xo[b + a]), int(xo[b + a + 1])
            cols = mesh.extraction_columns[lo:hi]
            vals = mesh.extraction_values[lo:hi]
            keep = vals != 0.0
            cols, vals = cols[keep], vals[keep]
            order = np.argsort(cols, kind="stable")
            rec = np.zeros(len(cols), dtype=[("c", "u1"), ("v", "<f8")])
            rec["c"], rec["v"] = cols[order], vals[order]
            parts.append(struct.pack("<B", len(cols)))
            parts.append(rec.tobytes())
        parts.append(mesh.bezier_points[192 * e:192 * e + 192].astype("<f8").tobytes())
        blob = b"".join(parts)
        records.append(blob)
        pos += len(blob)
    offsets[E] = pos
    rank_index_offset = pos
    with open(path, "wb") as f:
        f.write(head)
        f.write(struct.pack("<Q", rank_index_offset))
        if version >= 5:
            f.write(struct.pack("<5d", *origin, units, length_scale))
        f.write(offsets.astype("<u8").tobytes())
        f.write(np.zeros(E, "<i4").tobytes())
        for blob in records:
            f.write(blob)
        f.write(np.array([0, E], "<u8").tobytes())
        f.write(np.arange(E, dtype="<u8").tobytes())


def read_ntiga(path) -> FlatMesh:
    """Read a `.ntiga` database into a FlatMesh (IgaDatabase.hpp:64-161 +
    DeviceMesh.hpp:27-57: zero extraction values are dropped, columns kept in
    ascending order)."""
    data = Path(path).read_bytes()
    if data[:8] != MAGIC:
        raise ValueError("not a supported .ntiga database")
    version, ranks, E, N, bez, _res, rank_off = struct.unpack_from("<IIQQIIQ", data, 8)
    if version not in (3, 4, 5):
        raise ValueError("unsupported database version")
    pos = 48
    transform = ((0.0, 0.0, 0.0), 1.0, 1.0)
    if version >= 5:
        vals = struct.unpack_from("<5d", data, pos)
        pos += 40
        transform = (tuple(vals[:3]), vals[3], vals[4])
    if bez != 64:
        raise ValueError("unsupported Bezier point count")
    if N == 0 or rank_off == 0:
        raise ValueError("invalid database ownership index")
    offsets = np.frombuffer(data, "<u8", E + 1, pos)
    eoff = [0]
    conn, xoff, xcol, xval, bezier, labels, ids, types = [], [0], [], [], [], [], [], []
    for e in range(E):
        p = int(offsets[e])
        eid, etype, _owner, nen = struct.unpack_from("<QiiI", data, p)
        p += 20
        lab = [-1] * 6
        if version >= 4:
            lab = list(struct.unpack_from("<6i", data, p))
            p += 24
        if nen == 0 or nen > 4096:
            raise ValueError("invalid element basis count")
        conn.extend(struct.unpack_from(f"<{nen}i", data, p))
        p += 4 * nen
        for _a in range(nen):
            k = data[p]
            p += 1
            rec = np.frombuffer(data, dtype=[("c", "u1"), ("v", "<f8")], count=k, offset=p)
            p += 9 * k
            if k and rec["c"].max() >= 64:
                raise ValueError("invalid extraction column")
            row = np.zeros(64)
            row[rec["c"]] = rec["v"]
            nz = np.nonzero(row)[0]
            xcol.extend(nz.tolist())
            xval.extend(row[nz].tolist())
            xoff.append(len(xval))
        bezier.append(np.frombuffer(data, "<f8", 192, p))
        p += 1536
        labels.extend(lab)
        ids.append(eid)
        types.append(etype)
        eoff.append(len(conn))
    return FlatMesh(
        nodes=int(N),
        element_offsets=np.array(eoff, np.int32),
        connectivity=np.array(conn, np.int32),
        extraction_offsets=np.array(xoff, np.int32),
        extraction_columns=np.array(xcol, np.uint8),
        extraction_values=np.array(xval, np.float64),
        bezier_points=np.concatenate(bezier).astype(np.float64) if bezier else np.zeros(0),
        boundary_labels=np.array(labels, np.int32),
        element_ids=np.array(ids, np.uint64),
        element_types=np.array(types, np.int32),
        geometry_transform=transform,
    )
