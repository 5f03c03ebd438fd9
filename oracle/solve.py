"""Configured linear transport of F coupled scalar fields (`iga_cuda solve`),
restated on the CPU.  TEST INFRASTRUCTURE ONLY: the checker for the HIP path
of tfi_solve_*, never part of the product path.

* ``element`` restates AssembleGenericTransportKernel
  (solvers/cuda/include/GenericTransportKernels.cuh:26-143) for one element
  with numpy, the basis from the oracle's C++ EvaluateBasis
  (TransportElement.hpp:118-205) at the 4x4x4 Gauss points.
* ``face`` restates the per-face integrals of AssembleGenericSurfaceKernel
  (GenericTransportKernels.cuh:146-215): int N_a N_b dA and int N_a dA.
* ``assemble`` scatters them into F x F node blocks and applies the
  constrained rows (ApplyGenericBoundaryKernel, :217-233).
* ``run`` restates the time loop of ConfiguredTransportHost.inc:378-400:
  b = P c + s, constrained entries set, L c' = b (sparse direct solve here).

Pinned by tests/test_oracle_solve.py: configured as the two-species neuron
model the element equals the transport oracle's (itself pinned by closed
forms); the consistent SUPG diffusion term vanishes on a linear field and
has a closed form for x^2 on the unit cube; a huge SUPG diffusivity switches
SUPG off; face integrals on the unit cube have the Bernstein closed form; a
constant state follows the reaction recurrence and a volume source adds
dt S exactly.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

TIME_DERIVATIVE, DIFFUSION, ADVECTION, LINEAR_COUPLING, VOLUME_SOURCE = 0, 1, 2, 3, 4
SUPG_NONE, SUPG_GALERKIN, SUPG_TEST, SUPG_CONSISTENT = 0, 1, 2, 3

_G, _W = np.polynomial.legendre.leggauss(4)
GP = 0.5 * (_G + 1.0)
GW = 0.5 * _W
# faces as Quadrature.hpp: fixed axis and value, then the two varying axes
FIXED_AXIS = (2, 1, 0, 1, 0, 2)
FIXED_VALUE = (0.0, 0.0, 1.0, 1.0, 0.0, 1.0)
VARYING = ((0, 1), (0, 2), (1, 2), (0, 2), (1, 2), (0, 1))


def _tau(u, grad, dt, tau_diffusivity):
    adv = grad @ u
    inverse_length = np.abs(adv).sum()
    tau_space = 1.0 / inverse_length if inverse_length > 0.0 else 0.0
    tau_time = dt / 2.0
    speed2 = float(u @ u)
    out = np.zeros(len(tau_diffusivity))
    if not tau_space > 0.0:
        return out, adv
    for f, d in enumerate(tau_diffusivity):
        inv2 = 1.0 / tau_space ** 2 + 1.0 / tau_time ** 2
        if d > 0.0:
            tau_diffusion = speed2 / (d * inverse_length ** 2)
            inv2 += 1.0 / tau_diffusion ** 2
        out[f] = 1.0 / np.sqrt(inv2)
    return out, adv


def element(bezier, extraction, velocity, fields, terms, supg, tau_diffusivity, dt):
    """(left, previous, source): [nen F][nen F] with dof F a + field, and [nen F]."""
    extraction = np.ascontiguousarray(extraction, dtype=np.float64)
    nen = extraction.shape[0]
    F = int(fields)
    velocity = np.asarray(velocity, dtype=np.float64).reshape(nen, 3)
    L = np.zeros((nen, F, nen, F))
    P = np.zeros((nen, F, nen, F))
    s = np.zeros((nen, F))
    for q in range(64):
        i, j, k = q % 4, (q // 4) % 4, q // 16
        b = O.evaluate_basis(bezier, extraction, GP[i], GP[j], GP[k], hessian=True)
        N, G = b["value"], b["gradient"]
        lap = np.trace(b["hessian"], axis1=1, axis2=2)
        dv = GW[i] * GW[j] * GW[k] * b["raw_determinant"]
        u = N @ velocity
        tau, adv = _tau(u, G, dt, tau_diffusivity)
        gdot = G @ G.T
        for kind, e, t, c in terms:
            test = N + (tau[e] * adv if supg[e] else 0.0)
            if kind == VOLUME_SOURCE:
                s[:, e] += dt * c * test * dv
                continue
            if kind == TIME_DERIVATIVE:
                m = c * np.outer(test, N) * dv
                L[:, e, :, t] += m
                P[:, e, :, t] += m
            elif kind == DIFFUSION:
                g = test[:, None] * gdot if supg[e] == SUPG_TEST else gdot.copy()
                if supg[e] == SUPG_CONSISTENT:
                    g = g - tau[e] * np.outer(adv, lap)
                L[:, e, :, t] += dt * c * g * dv
            elif kind == ADVECTION:
                L[:, e, :, t] += dt * c * np.outer(test, adv) * dv
            elif kind == LINEAR_COUPLING:
                L[:, e, :, t] += dt * c * np.outer(test, N) * dv
    return L.reshape(nen * F, nen * F), P.reshape(nen * F, nen * F), s.ravel()


def _bernstein(x):
    y = 1.0 - x
    b = np.array([y ** 3, 3 * y * y * x, 3 * y * x * x, x ** 3])
    db = np.array([-3 * y * y, 3 * y * y - 6 * y * x, 6 * y * x - 3 * x * x, 3 * x * x])
    return b, db


def face(bezier, extraction, face_index):
    """(mass [nen][nen] = int N_a N_b dA, load [nen] = int N_a dA) on one face,
    4x4 Gauss rule, dA = |t_0 x t_1| from the Bezier geometry."""
    X = np.asarray(bezier, dtype=np.float64).reshape(64, 3)
    C = np.asarray(extraction, dtype=np.float64)
    nen = C.shape[0]
    mass = np.zeros((nen, nen))
    load = np.zeros(nen)
    v0, v1 = VARYING[face_index]
    for qi in range(4):
        for qj in range(4):
            c = [0.0, 0.0, 0.0]
            c[FIXED_AXIS[face_index]] = FIXED_VALUE[face_index]
            c[v0], c[v1] = GP[qi], GP[qj]
            bs = [_bernstein(x) for x in c]
            B = np.zeros(64)
            dB = np.zeros((64, 3))
            for p in range(64):
                ii, jj, kk = p % 4, (p // 4) % 4, p // 16
                B[p] = bs[0][0][ii] * bs[1][0][jj] * bs[2][0][kk]
                dB[p] = [bs[0][1][ii] * bs[1][0][jj] * bs[2][0][kk], bs[0][0][ii] * bs[1][1][jj] * bs[2][0][kk],
                         bs[0][0][ii] * bs[1][0][jj] * bs[2][1][kk]]
            t0, t1 = X.T @ dB[:, v0], X.T @ dB[:, v1]
            m = GW[qi] * GW[qj] * np.linalg.norm(np.cross(t0, t1))
            N = C @ B
            mass += np.outer(N, N) * m
            load += N * m
    return mass, load


def assemble(mesh, pattern, velocity, system, constrained, surfaces=None):
    """Global F x F blocks [nnzb][F][F] of L and P and the source [N][F] on the
    node pattern, surfaces {element, face, flux, robin, robin_source} added,
    constrained (node, field) rows applied."""
    F = system["fields"]
    ro, cols, diag = pattern
    nnzb = len(cols)
    L = np.zeros((nnzb, F, F))
    P = np.zeros((nnzb, F, F))
    S = np.zeros((mesh.nodes, F))
    velocity = np.asarray(velocity, dtype=np.float64).reshape(-1, 3)
    lookup = {}
    for r in range(mesh.nodes):
        for k in range(ro[r], ro[r + 1]):
            lookup[(r, int(cols[k]))] = k

    def scatter(conn, le, pe=None):
        nen = len(conn)
        le = le.reshape(nen, F, nen, F)
        pe = None if pe is None else pe.reshape(nen, F, nen, F)
        for a in range(nen):
            for b in range(nen):
                k = lookup[(int(conn[a]), int(conn[b]))]
                L[k] += le[a, :, b, :]
                if pe is not None:
                    P[k] += pe[a, :, b, :]

    for e in range(mesh.elements):
        conn = mesh.element_connectivity(e)
        le, pe, se = element(mesh.element_bezier(e), mesh.dense_extraction(e), velocity[conn], F, system["terms"],
                             system["supg"], system["tau_diffusivity"], system["dt"])
        scatter(conn, le, pe)
        S[conn] += se.reshape(-1, F)
    if surfaces is not None:
        dt = system["dt"]
        for i, (e, f) in enumerate(zip(surfaces["element"], surfaces["face"])):
            conn = mesh.element_connectivity(int(e))
            m, ld = face(mesh.element_bezier(int(e)), mesh.dense_extraction(int(e)), int(f))
            robin = np.asarray(surfaces["robin"][i], dtype=np.float64)
            total = np.asarray(surfaces["flux"][i], dtype=np.float64) + np.asarray(surfaces["robin_source"][i])
            nen = len(conn)
            le = np.zeros((nen, F, nen, F))
            for fld in range(F):
                le[:, fld, :, fld] = dt * robin[fld] * m
            scatter(conn, le.reshape(nen * F, nen * F))
            S[conn] += dt * np.outer(ld, total)
    mask = np.asarray(constrained).reshape(-1, F).astype(bool)
    for r, fld in zip(*np.nonzero(mask)):
        L[ro[r]:ro[r + 1], fld, :] = 0.0
        P[ro[r]:ro[r + 1], fld, :] = 0.0
        L[diag[r], fld, fld] = 1.0
    return L, P, S


def to_csr(pattern, blocks):
    import scipy.sparse as sp
    ro, cols, _ = pattern
    F = blocks.shape[1]
    return sp.bsr_matrix((blocks, np.asarray(cols), np.asarray(ro)), shape=(F * (len(ro) - 1),) * 2).tocsr()


def run(mesh, pattern, velocity, system, constrained, values, steps, state, surfaces=None):
    """State [N][F] after `steps` backward-Euler steps from `state` with the
    prescribed `values` [N][F] at the constrained entries."""
    import scipy.sparse.linalg as spla
    F = system["fields"]
    L, P, S = assemble(mesh, pattern, velocity, system, constrained, surfaces)
    solve = spla.factorized(to_csr(pattern, L).tocsc())
    Pc = to_csr(pattern, P)
    mask = np.asarray(constrained).ravel().astype(bool)
    vals = np.asarray(values, dtype=np.float64).ravel()
    c = np.asarray(state, dtype=np.float64).ravel().copy()
    for _ in range(steps):
        b = Pc @ c + S.ravel()
        b[mask] = vals[mask]
        c = solve(b)
    return c.reshape(-1, F)
