"""Boundary-condition resolution on the host (Python mirror of the C++ host CLI).

* ``wall_trace_basis`` -- WallTraceBasis (solvers/cpu/include/BoundarySupport.hpp:134-150):
  every basis function with a non-zero extraction coefficient on a Bezier face
  column of an external control-mesh face whose four corners are labelled wall.
* ``resolve_flow_boundaries`` -- ResolveFlowBoundaries for the generated cases
  (solvers/cpu/include/GenericCaseInput.hpp:174-258): wall label 0 -> zero
  velocity, inlet label 1 -> profile * scale, outlets -> pressure traction 0,
  then the wall trace is forced to no-slip (iga_cuda.cu:753-759).
* ``boundary_faces`` -- the labelled element faces whose flows the Newton loop
  integrates for the mass-balance test (iga_cuda.cu:520-557).
"""
from __future__ import annotations

import numpy as np

# hex corner tuples of the six faces (BoundarySupport.hpp:84-87) and the
# Bezier columns on each face (BoundarySupport.hpp:119-132)
_FACE_CORNERS = ((0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (4, 5, 6, 7))


def face_bezier_columns(face: int) -> np.ndarray:
    cols = []
    for k in range(4):
        for j in range(4):
            for i in range(4):
                on = ((face == 0 and k == 0) or (face == 1 and j == 0) or (face == 2 and i == 3)
                      or (face == 3 and j == 3) or (face == 4 and i == 0) or (face == 5 and k == 3))
                if on:
                    cols.append(i + 4 * j + 16 * k)
    return np.array(cols)


def external_faces(cells: np.ndarray):
    """(element, local face, corner nodes) of faces used once (first use kept)."""
    seen = {}
    for e, cell in enumerate(cells):
        for f, corners in enumerate(_FACE_CORNERS):
            nodes = tuple(int(cell[c]) for c in corners)
            key = tuple(sorted(nodes))
            if key in seen:
                seen[key][3] += 1
            else:
                seen[key] = [e, f, nodes, 1]
    return [(v[0], v[1], v[2]) for k, v in sorted(seen.items()) if v[3] == 1]


def wall_trace_basis(mesh, cells, labels) -> np.ndarray:
    out = set()
    for e, f, nodes in external_faces(cells):
        if not all(labels[n] == 0 for n in nodes):
            continue
        C = mesh.dense_extraction(e)
        rows = np.nonzero(np.any(C[:, face_bezier_columns(f)] != 0.0, axis=1))[0]
        conn = mesh.element_connectivity(e)
        out.update(int(conn[r]) for r in rows)
    return np.array(sorted(out), dtype=np.int64)


def resolve_flow_boundaries(case, scale: float = 1.0, wall_trace=None):
    """(velocity_mask, pressure_mask, velocity_target [N,3], pressure_target [N])."""
    labels = case.point_labels
    N = len(labels)
    vmask = np.zeros(N, np.int32)
    pmask = np.zeros(N, np.int32)
    vt = np.zeros((N, 3))
    pt = np.zeros(N)
    wall = labels == 0
    vmask[wall] = 1
    inlet = labels == 1
    vmask[inlet] = 1
    vt[inlet] = scale * case.inlet_velocity[inlet]
    if wall_trace is None:
        wall_trace = wall_trace_basis(case.mesh, case.cells, labels)
    vmask[wall_trace] = 1
    vt[wall_trace] = 0.0
    return vmask, pmask, vt, pt


def boundary_faces(mesh):
    """(element, face, label) for every element face with label >= 0."""
    lab = mesh.boundary_labels.reshape(-1, 6)
    e, f = np.nonzero(lab >= 0)
    return e.astype(np.int32), f.astype(np.int32), lab[e, f].astype(np.int32)


def initial_state(vmask, pmask, vt, pt):
    """Transient initial state: Dirichlet values on constrained nodes, zero elsewhere
    (iga_cuda.cu:879-889)."""
    N = len(vmask)
    s = np.zeros((N, 4))
    s[vmask != 0, :3] = vt[vmask != 0]
    s[pmask != 0, 3] = pt[pmask != 0]
    return s.ravel()
