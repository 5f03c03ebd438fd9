"""Synthetic meshes for the parity tests.

``split_basis`` turns a tricubic (nen = 64) mesh into one with larger elements
of the kind the reference's unstructured C2 splines produce next to
extraordinary edges (more than 64 basis functions spanning the 64 Bernstein
polynomials; nen <= 80, solvers/cuda/include/IgaCudaKernels.cuh:14,
DeviceMesh.hpp:55-56): in every ``every``-th element, ``extra`` of its basis
functions are each split into two with extraction rows lam*C_a and
(1 - lam)*C_a, the second one on a node of its own.  Each such element still
has a partition of unity and the same geometry; the mesh mixes nen = 64 and
nen = 64 + extra elements.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from tubularflowiga_amd.mesh import FlatMesh


def split_basis(mesh: FlatMesh, every: int = 3, extra: int = 16, lam: float = 0.375) -> FlatMesh:
    eoff, conn = mesh.element_offsets, mesh.connectivity
    xoff, xcol, xval = mesh.extraction_offsets, mesh.extraction_columns, mesh.extraction_values
    new_eoff = [0]
    new_conn, new_xoff, new_xcol, new_xval = [], [0], [], []
    nodes = mesh.nodes
    for e in range(mesh.elements):
        b, nen = int(eoff[e]), int(eoff[e + 1] - eoff[e])
        split = set(range(nen - extra, nen)) if (e % every == 0 and nen + extra <= 80) else set()
        rows = []
        for a in range(nen):
            lo, hi = int(xoff[b + a]), int(xoff[b + a + 1])
            cols, vals = xcol[lo:hi], xval[lo:hi]
            if a in split:
                rows.append((int(conn[b + a]), cols, lam * vals))
            else:
                rows.append((int(conn[b + a]), cols, vals))
        for a in sorted(split):
            lo, hi = int(xoff[b + a]), int(xoff[b + a + 1])
            rows.append((nodes, xcol[lo:hi], (1.0 - lam) * xval[lo:hi]))
            nodes += 1
        for node, cols, vals in rows:
            new_conn.append(node)
            new_xcol.append(cols)
            new_xval.append(vals)
            new_xoff.append(new_xoff[-1] + len(cols))
        new_eoff.append(new_eoff[-1] + len(rows))
    return dataclasses.replace(
        mesh, nodes=nodes,
        element_offsets=np.array(new_eoff, np.int32), connectivity=np.array(new_conn, np.int32),
        extraction_offsets=np.array(new_xoff, np.int32),
        extraction_columns=np.concatenate(new_xcol).astype(np.uint8),
        extraction_values=np.concatenate(new_xval).astype(np.float64), extra={})
