"""Deterministic generator of body-fitted vessel cases in the reference's input formats.

The reference builds `.ntiga` databases with a C2 tricubic unstructured-spline
extractor that needs Eigen (preprocessing/spline/kernel.cpp:360-560); none of
its inputs or outputs are checked in.  This module builds the same kind of
input from scratch:

* geometry as conforming multi-patch tensor-product B-spline volumes of degree
  ``p`` (2 = the C1 quadratic splines the workloads ask for, 3 = C2 cubic),
  C^{p-1} inside a patch and C0 across patch interfaces;
* open uniform knot vectors; control points placed at the Greville abscissae of
  an analytic patch map;
* each element's extraction rows written in the cubic Bernstein basis the
  `.ntiga` format hard-codes (64 Bezier points, IgaDatabase.hpp:20,90): for
  p = 2 the quadratic Bernstein coefficients are degree-elevated exactly
  (rational arithmetic, so structural zeros stay exact zeros);
* Bezier points = C^T X of the control points (isoparametric map);
* coordinates normalised by the smallest bounding-box extent as the reference
  pipeline does (docs/PIPELINE.md "Coordinate normalization"), with the affine
  map recorded in the v5 header.

Case files written beside the database: ``controlmesh.vtk`` (one 8-node hex
per element whose corners are control points, POINT_DATA ``label``: -1
interior, 0 wall, 1 inlet, 2.. outlets -- the labelling of
preprocessing/mesh/src/MeshGenerator.cpp), ``initial_velocityfield.txt``
(Poiseuille profile on inlet nodes) and ``simulation_config.json``
(schema 3).  Physical parameters are transformed to the normalised length
unit L0 consistently: rho' = rho L0^2, mu' = mu, u' = u / L0, p' = p, t' = t,
under which the VMS weak form (and tau_M, tau_C) are invariant.
"""
from __future__ import annotations

import functools
import json
import math
from dataclasses import dataclass, field
from fractions import Fraction
from pathlib import Path
from typing import Callable

import numpy as np
from scipy.spatial import cKDTree

from .mesh import FlatMesh, write_ntiga

# --------------------------------------------------------------------------- 1D spline algebra


def _poly_mul(a, b):
    out = [Fraction(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def _poly_add(a, b):
    n = max(len(a), len(b))
    return [(a[i] if i < len(a) else 0) + (b[i] if i < len(b) else 0) for i in range(n)]


def _bspline_element_polys(n: int, p: int):
    """Local polynomials (power basis in s in [0,1]) of the p+1 B-splines of an
    open uniform knot vector with n elements, per element: polys[e][r] is the
    coefficient list of basis e+r on element e (Cox-de Boor with exact
    rationals)."""
    knots = [Fraction(0)] * (p + 1) + [Fraction(i, n) for i in range(1, n)] + [Fraction(1)] * (p + 1)
    out = []
    for e in range(n):
        # x = (e + s)/n ; degree-0: only span e+p is active
        span = e + p
        N = {span: [Fraction(1)]}
        for d in range(1, p + 1):
            nxt = {}
            for i in range(span - d, span + 1):
                acc = [Fraction(0)]
                if i in N:
                    den = knots[i + d] - knots[i]
                    if den != 0:  # (x - t_i)/den
                        lin = [(Fraction(e, n) - knots[i]) / den, Fraction(1, n) / den]
                        acc = _poly_add(acc, _poly_mul(lin, N[i]))
                if i + 1 in N:
                    den = knots[i + d + 1] - knots[i + 1]
                    if den != 0:  # (t_{i+d+1} - x)/den
                        lin = [(knots[i + d + 1] - Fraction(e, n)) / den, -Fraction(1, n) / den]
                        acc = _poly_add(acc, _poly_mul(lin, N[i + 1]))
                nxt[i] = acc
            N = nxt
        out.append([N[e + r] for r in range(p + 1)])
    return out


def _power_to_cubic_bernstein(coeffs):
    """Cubic Bernstein coefficients of a polynomial of degree <= 3 on [0,1]:
    b_j = sum_{k<=j} C(j,k)/C(3,k) a_k (exact)."""
    a = list(coeffs) + [Fraction(0)] * (4 - len(coeffs))
    return [sum(Fraction(math.comb(j, k), math.comb(3, k)) * a[k] for k in range(j + 1)) for j in range(4)]


@functools.lru_cache(maxsize=None)
def extraction_1d(n: int, p: int) -> np.ndarray:
    """[n, p+1, 4] cubic Bernstein coefficients of the p+1 local B-splines of
    each element (degree elevation for p < 3 is implied by the conversion)."""
    if p not in (2, 3):
        raise ValueError("degree must be 2 or 3")
    polys = _bspline_element_polys(n, p)
    out = np.zeros((n, p + 1, 4))
    for e in range(n):
        for r in range(p + 1):
            out[e, r] = [float(c) for c in _power_to_cubic_bernstein(polys[e][r])]
    return out


def greville(n: int, p: int) -> np.ndarray:
    knots = [0.0] * (p + 1) + [i / n for i in range(1, n)] + [1.0] * (p + 1)
    return np.array([sum(knots[i + 1:i + p + 1]) / p for i in range(n + p)])


# --------------------------------------------------------------------------- patches


@dataclass
class Patch:
    """One tensor-product spline block.  ``fmap(xi, eta, zeta) -> (..., 3)``
    maps the unit cube (vectorised).  ``face_labels[f]`` labels patch face f in
    the element face numbering of the reference (0: zeta=0, 1: eta=0,
    2: xi=1, 3: eta=1, 4: xi=0, 5: zeta=1; IgaCudaKernels/Quadrature face
    tables); faces absent from the dict are interfaces.  ``branch`` orders
    elements for locality (branch, zeta layer, patch)."""
    name: str
    fmap: Callable
    n: tuple
    face_labels: dict = field(default_factory=dict)
    branch: int = 0


_FACE_AXIS = {0: (2, 0), 1: (1, 0), 2: (0, 1), 3: (1, 1), 4: (0, 0), 5: (2, 1)}


@dataclass
class Case:
    mesh: FlatMesh
    control_points: np.ndarray      # [N,3] normalised coordinates
    point_labels: np.ndarray        # [N]
    cells: np.ndarray               # [E,8] control-mesh hexes (VTK order)
    inlet_velocity: np.ndarray      # [N,3] normalised units
    length_unit_m: float            # L0 in metres
    manifest: dict


def build_patches(patches, degree: int = 2, scale_mm: float = 1.0, merge_tol: float = 1e-7, chunk: int = 16384):
    """Assemble patches into a FlatMesh (coordinates in the patches' units).
    The dense extraction rows and Bezier points are formed chunk by chunk in
    the final element order, so meshes of millions of elements stay within a
    few GB of host memory."""
    p = degree
    all_pts, patch_base, layouts = [], [], []
    for P in patches:
        g = [greville(k, p) for k in P.n]
        X, Y, Z = np.meshgrid(g[0], g[1], g[2], indexing="ij")
        pts = np.asarray(P.fmap(X, Y, Z), dtype=np.float64).reshape(len(g[0]), len(g[1]), len(g[2]), 3)
        patch_base.append(sum(len(a) for a in all_pts))
        all_pts.append(pts.reshape(-1, 3))
        layouts.append(pts.shape[:3])
    pts = np.concatenate(all_pts)
    # merge coincident control points across patch interfaces
    extent = np.ptp(pts, axis=0).max()
    tree = cKDTree(pts)
    pairs = tree.query_pairs(merge_tol * extent, output_type="ndarray")
    parent = np.arange(len(pts))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i
    for a, b in pairs:
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    root = np.arange(len(pts))
    for i in np.unique(pairs):
        root[i] = find(i)

    # per-patch element arrays, vectorised; element order (branch, zeta layer, patch, eta, xi)
    nen = (p + 1) ** 3
    lr = np.arange(p + 1)
    keys, conns, labs, cells_l, where = [], [], [], [], []
    ext = []

    def corner_index(c, n):
        return np.where(c < n, c, n + p - 1)
    corner_off = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])
    for pi, P in enumerate(patches):
        n1, n2, n3 = P.n
        s1, s2, s3 = layouts[pi]
        base = patch_base[pi]
        ext.append(tuple(extraction_1d(k, p) for k in P.n))
        I, J, K = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        # local basis a = r + (p+1) s + (p+1)^2 t
        R = (I[:, None, None, None] + lr[None, None, None, :])
        S = (J[:, None, None, None] + lr[None, None, :, None])
        T = (K[:, None, None, None] + lr[None, :, None, None])
        conns.append((base + (R*s2 + S)*s3 + T).reshape(-1, nen))
        lab = np.full((len(I), 6), -1, np.int32)
        for f, (axis, side) in _FACE_AXIS.items():
            if f in P.face_labels:
                idx = (I, J, K)[axis]
                hit = (idx == 0) if side == 0 else (idx == P.n[axis] - 1)
                lab[hit, f] = P.face_labels[f]
        labs.append(lab)
        ci = corner_index(I[:, None] + corner_off[None, :, 0], n1)
        cj = corner_index(J[:, None] + corner_off[None, :, 1], n2)
        ck = corner_index(K[:, None] + corner_off[None, :, 2], n3)
        cells_l.append(base + (ci*s2 + cj)*s3 + ck)
        keys.append(np.stack([np.full(len(I), P.branch), K, np.full(len(I), pi), J, I], axis=1))
        where.append(np.stack([np.full(len(I), pi), I, J, K], axis=1))
    keys = np.concatenate(keys)
    order_e = np.lexsort(keys.T[::-1])
    del keys
    conn_raw = np.concatenate(conns)[order_e]
    del conns
    labels = np.concatenate(labs)[order_e]
    cells_raw = np.concatenate(cells_l)[order_e]
    where = np.concatenate(where)[order_e]
    E = len(order_e)
    # node numbering by first appearance in element order (banded matrices)
    roots_in_order = root[conn_raw.ravel()]
    _, first = np.unique(roots_in_order, return_index=True)
    order = roots_in_order[np.sort(first)]
    del roots_in_order, first
    new_id = np.full(len(pts), -1, np.int64)
    new_id[order] = np.arange(len(order))
    node_of = new_id[root]
    if np.any(node_of[cells_raw] < 0):
        raise RuntimeError("control-mesh corner outside the element supports")
    N = len(order)
    cps = np.zeros((N, 3))
    cps[node_of] = pts
    conn = node_of[conn_raw].astype(np.int32)
    del conn_raw
    sc = np.sort(conn, axis=1)
    if np.any(sc[:, 1:] == sc[:, :-1]):
        raise RuntimeError("element connectivity repeats a node (degenerate patch)")
    del sc
    # extraction rows ex[i,r,a] ey[j,s,b] ez[k,t,c] -> [(t,s,r), (c,b,a)].  Inside a patch
    # most elements share their 1D operators, so the dense rows are formed once per
    # distinct (patch, 1D operator triple) and each element takes its template's
    # sparse rows; Bezier points C^T X are one batched product per template.
    kind_ids = []
    for pi in range(len(patches)):
        ids = []
        for d in range(3):
            flat = ext[pi][d].reshape(ext[pi][d].shape[0], -1)
            _, inv = np.unique(flat, axis=0, return_inverse=True)
            ids.append(inv.ravel())
        kind_ids.append(ids)
    combo = np.zeros((E, 4), np.int64)
    combo[:, 0] = where[:, 0]
    for pi in range(len(patches)):
        sel = where[:, 0] == pi
        for d in range(3):
            combo[sel, d + 1] = kind_ids[pi][d][where[sel, d + 1]]
    templates, tid = np.unique(combo, axis=0, return_inverse=True)
    tid = tid.ravel()
    t_counts, t_cols, t_vals, t_rows = [], [], [], []
    for pi, kx, ky, kz in templates:
        ex, ey, ez = ext[pi]
        ix = int(np.nonzero(kind_ids[pi][0] == kx)[0][0])
        iy = int(np.nonzero(kind_ids[pi][1] == ky)[0][0])
        iz = int(np.nonzero(kind_ids[pi][2] == kz)[0][0])
        rows = np.einsum("ra,sb,tc->tsrcba", ex[ix], ey[iy], ez[iz]).reshape(nen, 64)
        t_rows.append(rows)
        nz = rows != 0.0
        t_counts.append(nz.sum(axis=1))
        a_idx, p_idx = np.nonzero(nz)
        t_cols.append(p_idx.astype(np.uint8))
        t_vals.append(rows[a_idx, p_idx])
    t_counts = np.array(t_counts)                      # [templates, nen]
    t_nnz = t_counts.sum(axis=1)
    t_off = np.concatenate([[0], np.cumsum(t_nnz)])
    all_cols, all_vals = np.concatenate(t_cols), np.concatenate(t_vals)
    counts = t_counts[tid].ravel()
    xoff = np.zeros(E * nen + 1, np.int64)
    np.cumsum(counts, out=xoff[1:])
    if xoff[-1] >= 2 ** 31:
        raise RuntimeError("extraction nonzeros exceed the int32 offsets of the mesh arrays")
    e_nnz = t_nnz[tid]
    e_start = np.concatenate([[0], np.cumsum(e_nnz)[:-1]])
    gather = np.arange(int(xoff[-1]), dtype=np.int64) + np.repeat(t_off[tid] - e_start, e_nnz)
    xcol = all_cols[gather]
    xval = all_vals[gather]
    del gather
    bez = np.empty((E, 64, 3))
    order_t = np.argsort(tid, kind="stable")
    bounds = np.searchsorted(tid[order_t], np.arange(len(templates) + 1))
    for t in range(len(templates)):
        sel = order_t[bounds[t]:bounds[t + 1]]
        for c0 in range(0, len(sel), chunk):
            part = sel[c0:c0 + chunk]
            bez[part] = np.einsum("eap,eax->epx", np.broadcast_to(t_rows[t], (len(part), nen, 64)), cps[conn[part]])
    # point labels: wall (0) wins over caps; caps label their points
    plab = np.full(N, -1, np.int32)
    for pi, P in enumerate(patches):
        s1, s2, s3 = layouts[pi]
        grid = node_of[patch_base[pi]:patch_base[pi] + s1 * s2 * s3].reshape(s1, s2, s3)
        for f, lab in sorted(P.face_labels.items(), key=lambda kv: kv[1] == 0):
            axis, side = _FACE_AXIS[f]
            sl = [slice(None)] * 3
            sl[axis] = 0 if side == 0 else -1
            ids = grid[tuple(sl)].ravel()
            if lab == 0:
                plab[ids] = 0
            else:
                sel = ids[plab[ids] != 0]
                plab[sel] = lab
    mesh = FlatMesh(
        nodes=N,
        element_offsets=(np.arange(E + 1) * nen).astype(np.int32),
        connectivity=conn.ravel().astype(np.int32),
        extraction_offsets=xoff.astype(np.int32),
        extraction_columns=xcol,
        extraction_values=xval.astype(np.float64),
        bezier_points=bez.reshape(-1),
        boundary_labels=labels.ravel(),
    )
    return mesh, cps, plab, node_of[cells_raw].astype(np.int32)


def _normalise(mesh: FlatMesh, cps: np.ndarray, source_units_to_m: float):
    lower = cps.min(axis=0)
    extent = np.ptp(cps, axis=0).min()
    bez = mesh.bezier_points.reshape(-1, 3)
    mesh.bezier_points = ((bez - lower) / extent).reshape(-1)
    mesh.geometry_transform = (tuple(float(v) for v in lower), float(extent), float(source_units_to_m))
    return (cps - lower) / extent, extent * source_units_to_m


# --------------------------------------------------------------------------- cross sections
# 2D O-grid on the unit disc.  The core square has half-width A; the ring
# patches blend linearly from the square edge to the circle.
A_CORE = 0.45


def _ring(inner0, inner1, th0, th1):
    def f(s, r):
        ix = inner0[0] + (inner1[0] - inner0[0]) * s
        iy = inner0[1] + (inner1[1] - inner0[1]) * s
        th = th0 + (th1 - th0) * s
        return (1 - r) * ix + r * np.cos(th), (1 - r) * iy + r * np.sin(th)
    return f


def _core(x0, x1, y0, y1):
    def f(s, t):
        return x0 + (x1 - x0) * s, y0 + (y1 - y0) * t
    return f


def ogrid_sections(split: bool):
    """2D patch maps (s, t) -> (x, y) of the O-grid.  Each entry: (name, map,
    side) with side in {'L', 'R', ''} and 'wall' True when t=1 is the circle."""
    a = A_CORE
    pi = math.pi
    if not split:
        return [
            ("core", _core(-a, a, -a, a), "", False),
            ("top", _ring((-a, a), (a, a), 3 * pi / 4, pi / 4), "", True),
            ("right", _ring((a, a), (a, -a), pi / 4, -pi / 4), "", True),
            ("bottom", _ring((a, -a), (-a, -a), -pi / 4, -3 * pi / 4), "", True),
            ("left", _ring((-a, -a), (-a, a), -3 * pi / 4, -5 * pi / 4), "", True),
        ]
    return [
        ("coreL", _core(-a, 0.0, -a, a), "L", False),
        ("coreR", _core(0.0, a, -a, a), "R", False),
        ("topL", _ring((-a, a), (0.0, a), 3 * pi / 4, pi / 2), "L", True),
        ("topR", _ring((0.0, a), (a, a), pi / 2, pi / 4), "R", True),
        ("right", _ring((a, a), (a, -a), pi / 4, -pi / 4), "R", True),
        ("bottomR", _ring((a, -a), (0.0, -a), -pi / 4, -pi / 2), "R", True),
        ("bottomL", _ring((0.0, -a), (-a, -a), -pi / 2, -3 * pi / 4), "L", True),
        ("left", _ring((-a, -a), (-a, a), -3 * pi / 4, -5 * pi / 4), "L", True),
    ]


def _section_counts(name, n, m):
    half = name.endswith("L") or name.endswith("R")
    if name.startswith("core"):
        return (n // 2 if half else n, n)
    return (n // 2 if half else n, m)


def _orient(fmap, n, labels, probe=(0.5, 0.5, 0.5)):
    """Return a right-handed patch (flip xi if the map's Jacobian is negative)."""
    h = 1e-6
    x0 = np.array(fmap(*[np.array(v) for v in probe]), dtype=float)
    cols = []
    for d in range(3):
        q = list(probe)
        q[d] += h
        cols.append((np.array(fmap(*[np.array(v) for v in q]), dtype=float) - x0) / h)
    if np.linalg.det(np.array(cols).T) > 0:
        return fmap, n, labels
    flipped_labels = {({2: 4, 4: 2}.get(f, f)): v for f, v in labels.items()}
    return (lambda xi, eta, zeta: fmap(1.0 - xi, eta, zeta)), n, flipped_labels


def tube_patches(radius, length, n_core, m_ring, n_axial, inlet=1, outlet=2):
    """Straight tube along +z: O-grid with core n_core x n_core and four ring
    patches n_core x m_ring, n_axial elements along the axis."""
    out = []
    for name, f2, _side, wall in ogrid_sections(False):
        def fmap(xi, eta, zeta, f2=f2):
            x, y = f2(xi, eta)
            return np.stack([radius * x, radius * y, length * zeta + 0 * x], axis=-1)
        labels = {0: inlet, 5: outlet}
        if wall:
            labels[3] = 0
        n = _section_counts(name, n_core, m_ring) + (n_axial,)
        fm, n, labels = _orient(fmap, n, labels)
        out.append(Patch(name, fm, n, labels, branch=0))
    return out


def _smooth(x):
    x = np.clip(x, 0.0, 1.0)
    return x * x * (3.0 - 2.0 * x)


def y_bifurcation_patches(r_parent, r_daughter, angle_deg, l_parent, l_daughter, n_core, m_ring,
                          n_parent, n_daughter, transition=0.25):
    """Symmetric Y: parent along +z ending at z=0; two daughters whose
    cross-sections are the parent's left/right halves, rescaled to the area of a
    circle of radius r_daughter and swept along axes tilted by +-angle in the
    x-z plane.  The daughter sections are D-shaped (half-disc topology): the
    flat side becomes the inner (carina) wall."""
    if n_core % 2:
        raise ValueError("n_core must be even for the bifurcation split")
    cx = 4.0 / (3.0 * math.pi)              # centroid of the unit half disc
    lam_d = r_daughter * math.sqrt(2.0) / r_parent
    phi_d = math.radians(angle_deg)
    # frame angle and scale along the daughter (zeta in [0,1])
    zs = np.linspace(0.0, 1.0, 20001)
    phi_s = phi_d * np.clip(zs / transition, 0.0, 1.0)
    lam = lambda z: 1.0 + (lam_d - 1.0) * _smooth((np.asarray(z) - 0.5 * transition) / transition)
    sin_int = np.concatenate([[0.0], np.cumsum(0.5 * (np.sin(phi_s[1:]) + np.sin(phi_s[:-1])) * np.diff(zs))])
    cos_int = np.concatenate([[0.0], np.cumsum(0.5 * (np.cos(phi_s[1:]) + np.cos(phi_s[:-1])) * np.diff(zs))])

    def frame(zeta):
        z = np.asarray(zeta, dtype=float)
        phi = phi_d * np.clip(z / transition, 0.0, 1.0)
        return phi, np.interp(z, zs, sin_int), np.interp(z, zs, cos_int)

    out = []
    for name, f2, side, wall in ogrid_sections(True):
        cnt = _section_counts(name, n_core, m_ring)

        def parent_map(xi, eta, zeta, f2=f2):
            x, y = f2(xi, eta)
            return np.stack([r_parent * x, r_parent * y, -l_parent + l_parent * zeta + 0 * x], axis=-1)
        labels = {0: 1}
        if wall:
            labels[3] = 0
        fm, n, lab = _orient(parent_map, cnt + (n_parent,), labels)
        out.append(Patch("parent_" + name, fm, n, lab, branch=0))
        sgn = -1.0 if side == "L" else 1.0

        def daughter_map(xi, eta, zeta, f2=f2, sgn=sgn):
            x, y = f2(xi, eta)
            phi, si, co = frame(zeta)
            s = lam(zeta)
            qx = x - sgn * cx
            cxp = sgn * cx * r_parent + sgn * l_daughter * si
            czp = l_daughter * co
            e1x, e1z = np.cos(phi), -sgn * np.sin(phi)
            X = cxp + r_parent * s * qx * e1x
            Y = r_parent * s * y + 0 * qx
            Z = czp + r_parent * s * qx * e1z
            return np.stack([X + 0 * qx, Y, Z + 0 * qx], axis=-1)
        outlet = 2 if side == "L" else 3
        labels = {5: outlet}
        if wall:
            labels[3] = 0
        # the half-section's former split line is wall in the daughter
        split_face = {"coreL": 2, "topL": 2, "bottomL": 4, "coreR": 4, "topR": 4, "bottomR": 2}
        if name in split_face:
            labels[split_face[name]] = 0
        fm, n, lab = _orient(daughter_map, cnt + (n_daughter,), labels)
        out.append(Patch("daughter_" + name, fm, n, lab, branch=1 if side == "L" else 2))
    return out


# --------------------------------------------------------------------------- vessel tree
# A binary vessel tree whose segments are swept cross-sections.  Every section
# is a "bulged square": the Coons patch of four sine-bulged edges,
#     X = A xi (1 + beta sin(pi t)),  Y = A eta (1 + beta sin(pi s)),
# xi = 2s - 1, eta = 2t - 1 on the section parameters (s, t) in [0, 1]^2,
# which is smooth inside, convex, has corner angles of 2 atan(...) + 90 deg
# (about 133 deg at beta = 0.25) and a Jacobian determinant of at least
# 4 - beta^2 pi^2 A^2 > 0.  The root's section parameter square is shared by
# the whole tree: a bifurcation cuts a segment's parameter rectangle in half
# (along u at even levels, along v at odd ones) and each daughter starts from
# its half of the parent's end section, then morphs to its own bulged square
# of area pi r^2 while its axis turns away from the parent's.  For the
# daughters' end faces to conform to the parent's, every segment is split into
# patches along all the cuts of its subtree (C0 interfaces), so that the patch
# rectangles of a daughter at its start are exactly the parent's.

TREE_BETA = 0.25


def _bulged_square(s, t, half_width, beta=TREE_BETA):
    xi, eta = 2.0 * s - 1.0, 2.0 * t - 1.0
    return (half_width * xi * (1.0 + beta * np.sin(np.pi * t)),
            half_width * eta * (1.0 + beta * np.sin(np.pi * s)))


def _bulged_square_area(beta=TREE_BETA, n=400):
    s = (np.arange(n) + 0.5) / n
    S, T = np.meshgrid(s, s, indexing="ij")
    xi, eta = 2.0 * S - 1.0, 2.0 * T - 1.0
    det = (2.0 * (1.0 + beta * np.sin(np.pi * T)) * 2.0 * (1.0 + beta * np.sin(np.pi * S))
           - xi * beta * np.pi * np.cos(np.pi * T) * eta * beta * np.pi * np.cos(np.pi * S))
    return float(det.mean())


@dataclass
class TreeSegment:
    index: int
    level: int
    parent: int
    rect: tuple            # (u0, u1, v0, v1) in the root section's parameters
    radius: float          # equivalent radius (section area pi r^2)
    length: float
    angle_deg: float       # angle between this segment's and its parent's axis
    side: int              # -1 / +1: which half of the parent (0 for the root)
    axis: int              # 0: displaced along E1 (parent cut along u), 1: along E2
    start: np.ndarray      # centroid of the start section
    frame0: np.ndarray     # rows T, E1, E2 at the start
    children: list = field(default_factory=list)
    half_width: float = 0.0
    offset: np.ndarray = None   # start-section centroid in the parent's end frame (E1, E2)


def _tree_layout(seed, depth, r_root, length_range, angle_range_deg):
    rng = np.random.default_rng(seed)
    area_unit = _bulged_square_area()
    segs = [TreeSegment(0, 0, -1, (0.0, 1.0, 0.0, 1.0), r_root, float(rng.uniform(*length_range)) * r_root, 0.0, 0,
                        0, np.zeros(3), np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]))]
    segs[0].half_width = r_root * math.sqrt(math.pi / area_unit)
    segs[0].offset = np.zeros(2)
    queue = [0]
    while queue:
        k = queue.pop(0)
        P = segs[k]
        if P.level + 1 >= depth:
            continue
        cut = P.level % 2
        u0, u1, v0, v1 = P.rect
        end_c, end_f = _segment_end(P)
        for side in (-1, 1):
            if cut == 0:
                m = 0.5 * (u0 + u1)
                rect = (u0, m, v0, v1) if side < 0 else (m, u1, v0, v1)
            else:
                m = 0.5 * (v0 + v1)
                rect = (u0, u1, v0, m) if side < 0 else (u0, u1, m, v1)
            r = P.radius * 2.0 ** (-1.0 / 3.0)           # Murray: r_p^3 = r_1^3 + r_2^3
            child = TreeSegment(len(segs), P.level + 1, k, rect, r, float(rng.uniform(*length_range)) * r,
                                float(rng.uniform(*angle_range_deg)), side, cut, None, end_f)
            child.half_width = r * math.sqrt(math.pi / area_unit)
            child.offset = _section_centroid(P, rect)
            child.start = end_c + child.offset[0] * end_f[1] + child.offset[1] * end_f[2]
            P.children.append(child.index)
            segs.append(child)
            queue.append(child.index)
    return segs


def _local_st(seg, u, v):
    u0, u1, v0, v1 = seg.rect
    return (u - u0) / (u1 - u0), (v - v0) / (v1 - v0)


def _section_centroid(seg, rect, n=200):
    """Area centroid of seg's end section restricted to the parameter rectangle."""
    u = rect[0] + (rect[1] - rect[0]) * (np.arange(n) + 0.5) / n
    v = rect[2] + (rect[3] - rect[2]) * (np.arange(n) + 0.5) / n
    U, V = np.meshgrid(u, v, indexing="ij")
    X, Y = _bulged_square(*_local_st(seg, U, V), seg.half_width)
    dXu, dXv = np.gradient(X, u, v)
    dYu, dYv = np.gradient(Y, u, v)
    w = np.abs(dXu * dYv - dXv * dYu)
    return np.array([(X * w).sum() / w.sum(), (Y * w).sum() / w.sum()])


def _turn(seg, zeta, transition):
    """Angle phi(zeta) and the integrals of cos(phi), sin(phi) over [0, zeta]
    (phi grows linearly to the segment's angle over the transition)."""
    zeta = np.asarray(zeta, dtype=float)
    phi_d = math.radians(seg.angle_deg)
    if phi_d == 0.0:
        return np.zeros_like(zeta), zeta.copy(), np.zeros_like(zeta)
    tr = transition
    zc = np.minimum(zeta, tr)
    phi = phi_d * zc / tr
    ci = (tr / phi_d) * np.sin(phi) + np.maximum(zeta - tr, 0.0) * math.cos(phi_d)
    si = (tr / phi_d) * (1.0 - np.cos(phi)) + np.maximum(zeta - tr, 0.0) * math.sin(phi_d)
    return phi, ci, si


def _segment_frame(seg, zeta, transition=0.25):
    """Centre C(zeta) [...,3] and frame T, E1, E2 [...,3] along the segment."""
    phi, ci, si = _turn(seg, zeta, transition)
    T0, E10, E20 = seg.frame0
    Ea0 = E10 if seg.axis == 0 else E20
    sg = float(seg.side)
    c, s_ = np.cos(phi)[..., None], np.sin(phi)[..., None]
    T = c * T0 + sg * s_ * Ea0
    Ea = -sg * s_ * T0 + c * Ea0
    C = seg.start + seg.length * (ci[..., None] * T0 + sg * si[..., None] * Ea0)
    if seg.axis == 0:
        E1, E2 = Ea, np.broadcast_to(E20, Ea.shape)
    else:
        E1, E2 = np.broadcast_to(E10, Ea.shape), Ea
    return C, T, E1, E2


def _segment_end(seg, transition=0.25):
    C, T, E1, E2 = _segment_frame(seg, np.array(1.0), transition)
    return C, np.array([T, E1, E2])


def _segment_point(segs, seg, u, v, zeta, transition=0.25):
    """Physical point of the tree segment at root-section parameters (u, v) and
    axial parameter zeta."""
    C, T, E1, E2 = _segment_frame(seg, zeta, transition)
    X, Y = _bulged_square(*_local_st(seg, u, v), seg.half_width)
    if seg.parent >= 0:
        P = segs[seg.parent]
        Xh, Yh = _bulged_square(*_local_st(P, u, v), P.half_width)
        b = _smooth((np.asarray(zeta) - 0.5 * transition) / transition)
        X = (1.0 - b) * (Xh - seg.offset[0]) + b * X
        Y = (1.0 - b) * (Yh - seg.offset[1]) + b * Y
    return C + X[..., None] * E1 + Y[..., None] * E2


def _tree_cuts(segs, seg):
    """Sorted u and v cut positions inside seg's rectangle made by its subtree."""
    us, vs = {seg.rect[0], seg.rect[1]}, {seg.rect[2], seg.rect[3]}
    stack = [seg.index]
    while stack:
        s = segs[stack.pop()]
        for c in s.children:
            r = segs[c].rect
            us.update((r[0], r[1]))
            vs.update((r[2], r[3]))
            stack.append(c)
    return sorted(us), sorted(vs)


def tree_patches(seed=20261202, depth=4, r_root=3.0, n_section=16, axial_aspect=1.5, length_range=(8.0, 12.0),
                 angle_range_deg=(25.0, 45.0), transition=0.25):
    """Patches of a seeded binary vessel tree (units of r_root).  The root's
    section has n_section x n_section elements; its subtree's cuts split it
    into patches (C0 interfaces).  Axial element length = axial_aspect x the
    mean section element size of the segment.  Labels: 0 wall, 1 inlet (root
    start), 2.. outlets (leaf ends, in creation order)."""
    segs = _tree_layout(seed, depth, r_root, length_range, angle_range_deg)
    leaves = [s.index for s in segs if not s.children]
    outlet_of = {k: 2 + i for i, k in enumerate(leaves)}
    patches = []
    for seg in segs:
        us, vs = _tree_cuts(segs, seg)
        u0, u1, v0, v1 = seg.rect
        nu_seg = int(round(n_section * (u1 - u0)))
        nv_seg = int(round(n_section * (v1 - v0)))
        h_sec = 0.5 * (2.0 * seg.half_width / nu_seg + 2.0 * seg.half_width / nv_seg)
        nz = max(4, int(round(seg.length / (axial_aspect * h_sec))))
        seg.n_axial = nz
        for i in range(len(us) - 1):
            for j in range(len(vs) - 1):
                pu0, pu1, pv0, pv1 = us[i], us[i + 1], vs[j], vs[j + 1]
                nu = int(round(n_section * (pu1 - pu0)))
                nv = int(round(n_section * (pv1 - pv0)))
                if nu < 1 or nv < 1:
                    raise ValueError("n_section too small for the tree's cuts")

                def fmap(xi, eta, zeta, seg=seg, pu0=pu0, pu1=pu1, pv0=pv0, pv1=pv1):
                    u = pu0 + (pu1 - pu0) * np.asarray(xi)
                    v = pv0 + (pv1 - pv0) * np.asarray(eta)
                    return _segment_point(segs, seg, u, v, np.asarray(zeta) + 0.0 * u, transition)
                labels = {}
                if pu0 == u0:
                    labels[4] = 0
                if pu1 == u1:
                    labels[2] = 0
                if pv0 == v0:
                    labels[1] = 0
                if pv1 == v1:
                    labels[3] = 0
                if seg.parent < 0:
                    labels[0] = 1
                if not seg.children:
                    labels[5] = outlet_of[seg.index]
                fm, n, lab = _orient(fmap, (nu, nv, nz), labels)
                patches.append(Patch(f"seg{seg.index}_{i}{j}", fm, n, lab, branch=seg.index))
    return patches, segs


def _tree_clearance(segs, samples=41):
    """Smallest distance between the axes of two segments that are neither
    parent and child nor siblings, over the sum of their radii."""
    pts = []
    for s in segs:
        z = np.linspace(0.0, 1.0, samples)
        pts.append(_segment_frame(s, z)[0])
    worst = np.inf
    for a in segs:
        for b in segs:
            if b.index <= a.index or b.parent == a.index or a.parent == b.index:
                continue
            za = slice(samples // 2, None) if a.parent == b.parent else slice(None)
            d = np.linalg.norm(pts[a.index][za, None, :] - pts[b.index][None, za, :], axis=-1).min()
            worst = min(worst, d / (a.radius + b.radius))
    return float(worst)


def inlet_profile_tree(case_points, seg, flow):
    """Velocity of a parabolic-in-each-direction inflow (1 - xi^2)(1 - eta^2)
    over the root's bulged-square inlet, along the root axis, scaled to the
    volume flow `flow` (units of the points^3 per second); zero on the wall."""
    pts = np.asarray(case_points)
    x, y = pts[:, 0] - seg.start[0], pts[:, 1] - seg.start[1]
    A, beta = seg.half_width, TREE_BETA
    # invert the bulged square by Newton from (x, y) / A
    xi, eta = x / A, y / A
    for _ in range(30):
        st, ss = np.sin(np.pi * (eta + 1) / 2), np.sin(np.pi * (xi + 1) / 2)
        ct, cs = np.cos(np.pi * (eta + 1) / 2), np.cos(np.pi * (xi + 1) / 2)
        fx = A * xi * (1 + beta * st) - x
        fy = A * eta * (1 + beta * ss) - y
        a11, a12 = A * (1 + beta * st), A * xi * beta * ct * np.pi / 2
        a21, a22 = A * eta * beta * cs * np.pi / 2, A * (1 + beta * ss)
        det = a11 * a22 - a12 * a21
        xi = xi - (a22 * fx - a12 * fy) / det
        eta = eta - (a11 * fy - a21 * fx) / det
    shape = np.clip(1 - xi ** 2, 0, None) * np.clip(1 - eta ** 2, 0, None)
    n = 400
    s = (np.arange(n) + 0.5) / n
    S, T = np.meshgrid(s, s, indexing="ij")
    Xi, Eta = 2 * S - 1, 2 * T - 1
    det = A * A * (2 * (1 + beta * np.sin(np.pi * T)) * 2 * (1 + beta * np.sin(np.pi * S))
                   - Xi * beta * np.pi * np.cos(np.pi * T) * Eta * beta * np.pi * np.cos(np.pi * S))
    integral = float(((1 - Xi ** 2) * (1 - Eta ** 2) * det).mean())
    return shape * (flow / integral)


# --------------------------------------------------------------------------- case assembly


def make_case(kind: str, *, degree: int = 2, **kw) -> Case:
    """Build a case in memory.

    kind "tube": radius_mm, length_mm, n_core, m_ring, n_axial, u_max (m/s).
    kind "ybif": r_parent_mm, r_daughter_mm, angle_deg, l_parent_mm,
    l_daughter_mm, n_core, m_ring, n_parent, n_daughter, flow_ml_s.
    kind "tree": seed, depth, r_root_mm, n_section, axial_aspect, flow_ml_s
    (seeded binary vessel tree with Murray radii, see tree_patches).
    """
    if kind == "tube":
        R, L = kw["radius_mm"], kw["length_mm"]
        patches = tube_patches(R, L, kw["n_core"], kw["m_ring"], kw["n_axial"])
        inlet_axis, inlet_center, r_in = np.array([0.0, 0.0, 1.0]), np.zeros(3), R
        u_max = kw["u_max"]
    elif kind == "ybif":
        patches = y_bifurcation_patches(kw["r_parent_mm"], kw["r_daughter_mm"], kw["angle_deg"],
                                        kw["l_parent_mm"], kw["l_daughter_mm"], kw["n_core"],
                                        kw["m_ring"], kw["n_parent"], kw["n_daughter"])
        R = kw["r_parent_mm"]
        inlet_axis, inlet_center, r_in = np.array([0.0, 0.0, 1.0]), np.array([0, 0, -kw["l_parent_mm"]]), R
        q = kw["flow_ml_s"] * 1e-6
        u_max = 2.0 * q / (math.pi * (R * 1e-3) ** 2)
    elif kind == "tree":
        patches, segs = tree_patches(seed=kw["seed"], depth=kw["depth"], r_root=kw["r_root_mm"],
                                     n_section=kw["n_section"], axial_aspect=kw["axial_aspect"])
        clearance = _tree_clearance(segs)
        if clearance < 1.0:
            raise ValueError(f"tree seed {kw['seed']}: two segment axes come within {clearance:.2f} x the sum of "
                             "their radii")
    else:
        raise ValueError(f"unknown case kind {kind!r}")
    mesh, cps_mm, plab, cells = build_patches(patches, degree=degree)
    profile = np.zeros_like(cps_mm)
    inlet = plab == 1
    extra = {}
    if kind == "tree":
        # (1 - xi^2)(1 - eta^2) profile on the root's bulged-square inlet, volume flow Q (m/s in source frame)
        speed = inlet_profile_tree(cps_mm[inlet], segs[0], kw["flow_ml_s"] * 1e3) * 1e-3
        profile[inlet] = np.outer(speed, segs[0].frame0[0])
        u_max = float(speed.max())
        extra = {"clearance": clearance, "segments": [
            {"index": g.index, "level": g.level, "parent": g.parent, "radius_mm": g.radius, "length_mm": g.length,
             "angle_deg": g.angle_deg, "side": g.side, "axial_elements": g.n_axial,
             "start_mm": [float(v) for v in g.start], "axis": [float(v) for v in g.frame0[0]]} for g in segs]}
    else:
        # Poiseuille profile on the inlet (metres/second in source frame)
        rel = cps_mm - inlet_center
        rr = np.linalg.norm(rel - np.outer(rel @ inlet_axis, inlet_axis), axis=1) / r_in
        profile[inlet] = np.outer(u_max * np.clip(1.0 - rr[inlet] ** 2, 0.0, None), inlet_axis)
    cps, l0_m = _normalise(mesh, cps_mm, 1e-3)
    manifest = {"kind": kind, "degree": degree, **kw, "nodes": mesh.nodes, "elements": mesh.elements,
                "dofs": 4 * mesh.nodes, "length_unit_m": l0_m, "u_max_m_s": u_max,
                "generator": "tubularflowiga_amd.casegen", **extra}
    return Case(mesh, cps, plab, cells, profile / l0_m, l0_m, manifest)


def flow_parameters(case: Case, rho=1060.0, mu=3.5e-3):
    """(rho', mu') in the case's normalised length unit."""
    return rho * case.length_unit_m ** 2, mu


def write_case(case: Case, directory, *, rho=1060.0, mu=3.5e-3, dt=1e-3, steps=5,
               time_integration="backward_euler", waveform=None, outlets=None, version=5):
    """Write <dir>/case.ntiga, controlmesh.vtk, initial_velocityfield.txt,
    simulation_config.json and case_manifest.json.  ``waveform`` is a
    temporal_functions entry (dict) scaling the inlet profile."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    write_ntiga(d / "case.ntiga", case.mesh, version=version)
    N = case.mesh.nodes
    with open(d / "controlmesh.vtk", "w") as f:
        f.write("# vtk DataFile Version 3.0\ncontrol mesh\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {N} double\n")
        f.write("\n".join(" ".join(repr(float(v)) for v in p) for p in case.control_points))
        E = len(case.cells)
        f.write(f"\nCELLS {E} {9 * E}\n")
        f.write("\n".join("8 " + " ".join(str(int(v)) for v in c) for c in case.cells))
        f.write(f"\nCELL_TYPES {E}\n" + "\n".join(["12"] * E))
        f.write(f"\nPOINT_DATA {N}\nSCALARS label int 1\nLOOKUP_TABLE default\n")
        f.write("\n".join(str(int(v)) for v in case.point_labels) + "\n")
    np.savetxt(d / "initial_velocityfield.txt", case.inlet_velocity, fmt="%.17g")
    rho_n, mu_n = flow_parameters(case, rho, mu)
    labels = sorted(set(int(v) for v in case.point_labels if v >= 0))
    inlet_cond = {"field": "velocity", "type": "dirichlet", "profile": "initial_velocityfield.txt", "scale": 1.0}
    temporal = []
    if waveform:
        temporal.append(waveform)
        inlet_cond["waveform"] = waveform["name"]
    boundaries = [{"label": 0, "name": "wall",
                   "conditions": [{"field": "velocity", "type": "dirichlet", "value": [0.0, 0.0, 0.0]}]},
                  {"label": 1, "name": "inlet", "conditions": [inlet_cond]}]
    for lab in labels:
        if lab >= 2:
            cond = {"field": "pressure", "type": "pressure_traction", "value": 0.0}
            if outlets and lab in outlets:
                cond = outlets[lab]
            boundaries.append({"label": lab, "name": f"outlet{lab - 1}", "conditions": [cond]})
    config = {
        "schema_version": 3, "dimension": "3d",
        "fields": [{"name": "velocity", "kind": "vector3"}, {"name": "pressure", "kind": "pressure"}],
        "time": {"dt": dt, "steps": steps},
        "equation_systems": [{"name": "blood_flow", "kind": "navier_stokes",
                              "unknowns": ["velocity", "pressure"], "viscosity": mu_n,
                              "density": rho_n, "time_integration": time_integration}],
        "boundaries": boundaries,
    }
    if temporal:
        config["temporal_functions"] = temporal
    (d / "simulation_config.json").write_text(json.dumps(config, indent=2) + "\n")
    man = dict(case.manifest)
    man.update({"rho_si": rho, "mu_si": mu, "rho_normalised": rho_n, "mu_normalised": mu_n, "dt": dt,
                "steps": steps})
    (d / "case_manifest.json").write_text(json.dumps(man, indent=2, default=float) + "\n")
    return d


TRANSPORT_KEYS = {"diffusion": "D", "vplus": "vplus", "vminus": "vminus", "kplus": "kplus", "kminus": "kminus",
                  "detach_plus": "k'plus", "detach_minus": "k'minus", "dt": "dt", "n0_bc": "N0bc",
                  "nplus_bc": "Nplusbc", "nminus_bc": "Nminusbc", "artificial_diffusion": "artificial_diffusion"}


def tube_velocity(case: Case) -> np.ndarray:
    """[N,3] velocity over a straight tube along z: every cross-section carries
    the inlet profile (nodes matched by their (x, y) control-point position)."""
    P = np.asarray(case.control_points)
    inlet = np.nonzero(np.abs(case.inlet_velocity).sum(1) > 0)[0]
    lookup = {(round(float(P[i, 0]), 9), round(float(P[i, 1]), 9)): case.inlet_velocity[i] for i in inlet}
    out = np.zeros((case.mesh.nodes, 3))
    for n in range(case.mesh.nodes):
        out[n] = lookup.get((round(float(P[n, 0]), 9), round(float(P[n, 1]), 9)), 0.0)
    return out


def developed_tube_state(case: Case) -> np.ndarray:
    """[N, 4] developed Poiseuille flow in a straight tube along z: the inlet
    profile on every cross-section (``tube_velocity``) and the pressure
    p = G (z_max - z), G = 4 mu u_max / R^2, in the case's normalised units
    (mu = 3.5e-3 Pa s as ``flow_parameters``).  A start for pulsatile runs that
    skips the impulsive start from rest."""
    _, mu = flow_parameters(case)
    L0 = case.length_unit_m
    radius = case.manifest["radius_mm"] * 1e-3 / L0
    u_max = case.manifest["u_max"] / L0
    z = np.asarray(case.control_points)[:, 2]
    out = np.zeros((case.mesh.nodes, 4))
    out[:, :3] = tube_velocity(case)
    out[:, 3] = 4.0 * mu * u_max / (radius * radius) * (z.max() - z)
    return out


def write_transport_case(case: Case, directory, params: dict, steps: int, velocity=None, case_config=None):
    """A case directory for `transport`: the flow case files plus the legacy
    simulation_parameter.txt ("key value" lines, solvers/cpu/include/CaseInput.hpp:85-119),
    optionally a velocity file transport_velocity.txt and a case_config.json."""
    d = write_case(case, directory)
    lines = [f"{TRANSPORT_KEYS[k]} {float(v)!r}" for k, v in params.items()]
    lines.append(f"nstep {int(steps)}")
    (d / "simulation_parameter.txt").write_text("\n".join(lines) + "\n")
    if velocity is not None:
        np.savetxt(d / "transport_velocity.txt", np.asarray(velocity).reshape(-1, 3), fmt="%.17g")
    if case_config is not None:
        (d / "case_config.json").write_text(json.dumps(case_config, indent=2) + "\n")
    return d


def write_solve_case(case: Case, directory, fields, systems, boundaries, dt: float, steps: int, velocity=None,
                     extra=None):
    """A case directory for `solve`: the flow case files with a
    simulation_config.json holding the given scalar `fields` (dicts with name,
    kind, initial_value), linear_transport `systems` and `boundaries`; the
    prescribed velocity in initial_velocityfield.txt (the tube's Poiseuille
    field unless `velocity` is given).  `extra` adds top-level keys."""
    d = write_case(case, directory)
    if velocity is None:
        velocity = tube_velocity(case)
    np.savetxt(d / "initial_velocityfield.txt", np.asarray(velocity).reshape(-1, 3), fmt="%.17g")
    config = {"schema_version": 3, "dimension": "3d", "fields": list(fields),
              "time": {"dt": dt, "steps": steps}, "equation_systems": list(systems), "boundaries": list(boundaries)}
    config.update(extra or {})
    (d / "simulation_config.json").write_text(json.dumps(config, indent=2) + "\n")
    return d


# Named workloads (sizes chosen so the DOF counts match the workload text).
WORKLOADS = {
    # straight tube R=1 mm, L=10 mm, 5 hexes per section x 16 axial, Poiseuille 0.1 m/s
    "tube_c0": dict(kind="tube", radius_mm=1.0, length_mm=10.0, n_core=1, m_ring=1, n_axial=16, u_max=0.1),
    # symmetric Y, parent 2 mm, daughters 1.6 mm at +-35 deg, 20 mm branches, Q = 5 ml/s, ~0.6M DOFs
    "ybif_c1": dict(kind="ybif", r_parent_mm=2.0, r_daughter_mm=1.6, angle_deg=35.0, l_parent_mm=20.0,
                    l_daughter_mm=20.0, n_core=12, m_ring=8, n_parent=104, n_daughter=104, flow_ml_s=5.0),
    # small Y for CPU tests
    "ybif_small": dict(kind="ybif", r_parent_mm=2.0, r_daughter_mm=1.6, angle_deg=35.0, l_parent_mm=20.0,
                       l_daughter_mm=20.0, n_core=2, m_ring=1, n_parent=8, n_daughter=8, flow_ml_s=5.0),
    # refined straight tube, ~4M DOFs: the O-grid section of 2 x 2 core + 4 rings of 2 x 2 control hexes
    # (20 per section; the O-grid cannot make the 24 the workload names) and 256 axial, refined 5.5-fold
    # uniformly (11 x 11 core, rings 11 x 11, 1408 axial); R = 1 mm, L = 64 mm, Poiseuille inlet 0.1 m/s
    "tube_c2": dict(kind="tube", radius_mm=1.0, length_mm=64.0, n_core=11, m_ring=11, n_axial=1408, u_max=0.1),
    # seeded vessel tree: binary, depth 4 (15 segments), root radius 3 mm, Murray radii (r^3 conserved),
    # segment lengths U[8, 12] x radius, branching angles U[25, 45] deg, Q0 = 5 ml/s; ~8M DOFs
    "tree_c4": dict(kind="tree", seed=20261202, depth=4, r_root_mm=3.0, n_section=48, axial_aspect=1.05,
                    flow_ml_s=5.0),
    # the same tree, coarse, for parity tests against the oracle (2 x 4 elements per leaf section: with
    # fewer, every control-mesh cell of an outlet cap has four wall corners, the wall trace then pins the
    # whole cap and the pressure loses its outlet reference)
    "tree_small": dict(kind="tree", seed=20261202, depth=4, r_root_mm=3.0, n_section=8, axial_aspect=6.0,
                       flow_ml_s=5.0),
}


def workload(name: str, degree: int = 2) -> Case:
    kw = dict(WORKLOADS[name])
    kind = kw.pop("kind")
    return make_case(kind, degree=degree, **kw)
