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


def build_patches(patches, degree: int = 2, scale_mm: float = 1.0, merge_tol: float = 1e-7):
    """Assemble patches into a FlatMesh (coordinates in the patches' units)."""
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
    keys, conns, rowss, labs, cells_l = [], [], [], [], []

    def corner_index(c, n):
        return np.where(c < n, c, n + p - 1)
    corner_off = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])
    for pi, P in enumerate(patches):
        n1, n2, n3 = P.n
        s1, s2, s3 = layouts[pi]
        base = patch_base[pi]
        ex, ey, ez = (extraction_1d(k, p) for k in P.n)
        I, J, K = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        # local basis a = r + (p+1) s + (p+1)^2 t
        R = (I[:, None, None, None] + lr[None, None, None, :])
        S = (J[:, None, None, None] + lr[None, None, :, None])
        T = (K[:, None, None, None] + lr[None, :, None, None])
        conns.append((base + (R*s2 + S)*s3 + T).reshape(-1, nen))
        # extraction rows: ex[i,r,a] ey[j,s,b] ez[k,t,c] -> [(t,s,r), (c,b,a)]
        rows = np.einsum("ira,jsb,ktc->ijktsrcba", ex, ey, ez, optimize=True).reshape(n1*n2*n3, nen, 64)
        rowss.append(rows)
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
    keys = np.concatenate(keys)
    order_e = np.lexsort(keys.T[::-1])
    conn_raw = np.concatenate(conns)[order_e]
    rows = np.concatenate(rowss)[order_e]
    labels = np.concatenate(labs)[order_e]
    cells_raw = np.concatenate(cells_l)[order_e]
    E = len(order_e)
    # node numbering by first appearance in element order (banded matrices)
    roots_in_order = root[conn_raw.ravel()]
    _, first = np.unique(roots_in_order, return_index=True)
    order = roots_in_order[np.sort(first)]
    new_id = np.full(len(pts), -1, np.int64)
    new_id[order] = np.arange(len(order))
    node_of = new_id[root]
    if np.any(node_of[cells_raw] < 0):
        raise RuntimeError("control-mesh corner outside the element supports")
    N = len(order)
    cps = np.zeros((N, 3))
    cps[node_of] = pts
    conn = node_of[conn_raw].astype(np.int32)
    sc = np.sort(conn, axis=1)
    if np.any(sc[:, 1:] == sc[:, :-1]):
        raise RuntimeError("element connectivity repeats a node (degenerate patch)")
    bez = np.einsum("eap,eax->epx", rows, cps[conn])
    # sparse extraction
    nzmask = rows != 0.0
    counts = nzmask.sum(axis=2).ravel()
    xoff = np.zeros(E * nen + 1, np.int32)
    np.cumsum(counts, out=xoff[1:])
    e_idx, a_idx, p_idx = np.nonzero(nzmask)
    xcol = p_idx.astype(np.uint8)
    xval = rows[e_idx, a_idx, p_idx]
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
        extraction_offsets=xoff,
        extraction_columns=xcol,
        extraction_values=xval.astype(np.float64),
        bezier_points=bez.reshape(-1).astype(np.float64),
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


# --------------------------------------------------------------------------- case assembly


def make_case(kind: str, *, degree: int = 2, **kw) -> Case:
    """Build a case in memory.

    kind "tube": radius_mm, length_mm, n_core, m_ring, n_axial, u_max (m/s).
    kind "ybif": r_parent_mm, r_daughter_mm, angle_deg, l_parent_mm,
    l_daughter_mm, n_core, m_ring, n_parent, n_daughter, flow_ml_s.
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
    else:
        raise ValueError(f"unknown case kind {kind!r}")
    mesh, cps_mm, plab, cells = build_patches(patches, degree=degree)
    # Poiseuille profile on the inlet (metres/second in source frame)
    rel = cps_mm - inlet_center
    rr = np.linalg.norm(rel - np.outer(rel @ inlet_axis, inlet_axis), axis=1) / r_in
    profile = np.zeros_like(cps_mm)
    inlet = plab == 1
    profile[inlet] = np.outer(u_max * np.clip(1.0 - rr[inlet] ** 2, 0.0, None), inlet_axis)
    cps, l0_m = _normalise(mesh, cps_mm, 1e-3)
    manifest = {"kind": kind, "degree": degree, **kw, "nodes": mesh.nodes, "elements": mesh.elements,
                "dofs": 4 * mesh.nodes, "length_unit_m": l0_m, "u_max_m_s": u_max,
                "generator": "tubularflowiga_amd.casegen"}
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
}


def workload(name: str, degree: int = 2) -> Case:
    kw = dict(WORKLOADS[name])
    kind = kw.pop("kind")
    return make_case(kind, degree=degree, **kw)
