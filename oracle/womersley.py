"""Analytic pipe-flow references and the volume error of a discrete velocity
field.  TEST INFRASTRUCTURE ONLY: the checker for the end-to-end flow tests,
never part of the product path.

* ``axial_velocity`` restates the Womersley solution of
  solvers/cpu/include/WomersleyReference.hpp:131-199 (ComplexBesselJ0 and
  WomersleyAxialVelocity): the steady Poiseuille part G0 (R^2 - r^2) / (4 mu)
  plus, per Fourier harmonic n of the driving gradient
  G(t) = G0 + sum a_n cos(n w t) + b_n sin(n w t), the complex amplitude
  (a_n - i b_n) / (i n w rho) (1 - J0(lambda r/R) / J0(lambda)) e^{i n w t}
  with lambda = alpha e^{-i pi/4}, alpha = R sqrt(n w rho / mu).  J0 of a
  complex argument is scipy.special.jv (the reference sums the power series).
  Pinned by tests/test_oracle_womersley.py against the values asserted in
  solvers/cpu/tests/test_womersley_reference.cpp.
* ``read_configuration`` reads the reference's Womersley configuration JSON
  (examples/validation/womersley/womersley_reference.json, a copy of which is
  the fixture tests/golden/womersley_reference.json).
* ``volume_relative_l2`` integrates ||u_h - u||_2 / ||u||_2 over the mesh with
  the 4x4x4 Gauss rule of the elements (the measure of
  solvers/cpu/include/Quadrature.hpp:182-207), the discrete field evaluated
  from the extraction operator and the Bezier geometry, as
  `iga_flow_validate --womersley` does (examples/validation/womersley/README.md).
"""
from __future__ import annotations

import json
import math

import numpy as np
from scipy import special

GAUSS = np.array([0.06943184420297371, 0.33000947820757187, 0.6699905217924281, 0.9305681557970262])
GAUSS_W = np.array([0.3478548451374539, 0.6521451548625461, 0.6521451548625461, 0.3478548451374539]) / 2.0


def read_configuration(path):
    cfg = json.loads(open(path).read())
    axis = np.asarray(cfg["axis_direction"], dtype=float)
    cfg["axis_direction"] = axis / np.linalg.norm(axis)
    cfg["axis_origin"] = np.asarray(cfg["axis_origin"], dtype=float)
    n = max(len(cfg.get("cosine_pressure_gradient", [])), len(cfg.get("sine_pressure_gradient", [])))
    cfg["cosine_pressure_gradient"] = list(cfg.get("cosine_pressure_gradient", [])) + [0.0] * (
        n - len(cfg.get("cosine_pressure_gradient", [])))
    cfg["sine_pressure_gradient"] = list(cfg.get("sine_pressure_gradient", [])) + [0.0] * (
        n - len(cfg.get("sine_pressure_gradient", [])))
    return cfg


def bessel_j0(z):
    return special.jv(0, np.asarray(z, dtype=complex))


def axial_velocity(cfg, r, t):
    """Axial velocity at radial distance r (array) and time t."""
    R, rho, mu = cfg["radius"], cfg["density"], cfg["dynamic_viscosity"]
    r = np.asarray(r, dtype=float)
    frac = np.minimum(r / R, 1.0)
    u = cfg["mean_pressure_gradient"] * (R * R - r * r) / (4.0 * mu)
    w = 2.0 * math.pi / cfg["period"]
    for k, (a, b) in enumerate(zip(cfg["cosine_pressure_gradient"], cfg["sine_pressure_gradient"])):
        omega = (k + 1) * w
        alpha = R * math.sqrt(omega * rho / mu)
        wall = complex(alpha / math.sqrt(2.0), -alpha / math.sqrt(2.0))
        profile = 1.0 - bessel_j0(wall * frac) / bessel_j0(wall)
        amplitude = complex(a, -b) / (1j * omega * rho) * profile
        u = u + np.real(amplitude * np.exp(1j * omega * t))
    return u


def velocity(cfg, points, t):
    """[n, 3] Womersley velocity at the points [n, 3]."""
    d = np.asarray(points, dtype=float) - cfg["axis_origin"]
    axis = cfg["axis_direction"]
    s = d @ axis
    radial = d - np.outer(s, axis)
    return np.outer(axial_velocity(cfg, np.linalg.norm(radial, axis=1), t), axis)


def _bernstein(x):
    y = 1.0 - x
    b = np.stack([y ** 3, 3 * y * y * x, 3 * y * x * x, x ** 3])
    d = np.stack([-3 * y * y, 3 - 12 * x + 9 * x * x, 3 * (2 - 3 * x) * x, 3 * x * x])
    return b, d


def _tables():
    """[64 points, 64 Bernstein] values and the three parametric derivatives,
    point index q = i + 4 j + 16 k over (xi, eta, zeta), Bezier index the same."""
    b, d = _bernstein(GAUSS)          # [4 functions, 4 points]
    val = np.einsum("ai,bj,ck->kjicba", b, b, b).reshape(64, 64)
    dx = np.einsum("ai,bj,ck->kjicba", d, b, b).reshape(64, 64)
    dy = np.einsum("ai,bj,ck->kjicba", b, d, b).reshape(64, 64)
    dz = np.einsum("ai,bj,ck->kjicba", b, b, d).reshape(64, 64)
    w = np.einsum("i,j,k->kji", GAUSS_W, GAUSS_W, GAUSS_W).reshape(64)
    return val, dx, dy, dz, w


def volume_relative_l2(mesh, state, exact, chunk=4096):
    """||u_h - exact(x)|| / ||exact(x)|| over the volume.  state: [4N]
    (velocity in the first three fields of each node); exact: points [n, 3] ->
    [n, 3].  Elements of the mesh all carry the same number of basis
    functions."""
    val, dx, dy, dz, w = _tables()
    u_nodes = np.asarray(state, dtype=float).reshape(-1, 4)[:, :3]
    E = mesh.elements
    nen = mesh.maximum_basis
    if np.any(np.diff(mesh.element_offsets) != nen):
        raise ValueError("volume_relative_l2 expects elements of one size")
    conn = mesh.connectivity.reshape(E, nen)
    num = den = 0.0
    for e0 in range(0, E, chunk):
        e1 = min(E, e0 + chunk)
        C = np.stack([mesh.dense_extraction(e) for e in range(e0, e1)])       # [e, a, p]
        X = mesh.bezier_points.reshape(E, 64, 3)[e0:e1]                      # [e, p, 3]
        pts = np.einsum("qp,epx->eqx", val, X)
        J = np.stack([np.einsum("qp,epx->eqx", d, X) for d in (dx, dy, dz)], axis=-1)
        det = np.abs(np.linalg.det(J))                                         # [e, q]
        N = np.einsum("eap,qp->eqa", C, val)                                   # [e, q, a]
        uh = np.einsum("eqa,eax->eqx", N, u_nodes[conn[e0:e1]])
        ue = exact(pts.reshape(-1, 3)).reshape(uh.shape)
        m = det * w
        num += float(np.sum(m * np.sum((uh - ue) ** 2, axis=-1)))
        den += float(np.sum(m * np.sum(ue ** 2, axis=-1)))
    return math.sqrt(num / den)
