"""Two-species neuron transport (`iga_cuda transport`), restated on the CPU.
TEST INFRASTRUCTURE ONLY: the checker for the HIP transport path, never part
of the product path.

* ``element`` calls the C++ restatement of BuildTransportElement
  (solvers/cpu/include/TransportElement.hpp:213-253; the GPU reference builds
  the same matrices in AssembleTransportKernel,
  solvers/cuda/include/IgaCudaKernels.cuh:155-237).
* ``assemble`` scatters the element matrices into 2x2 node blocks on the mesh
  pattern and applies the constrained rows (ApplyTransportBoundaryKernel,
  IgaCudaKernels.cuh:411-424).
* ``run`` restates the time loop of iga_cuda.cu:674-705: b = P c, constrained
  entries of b set to their values, L c' = b, solved here by a sparse direct
  factorisation.

Pinned by tests/test_oracle_transport.py (closed forms: the mass matrix sums
to the element volume, diffusion annihilates constants, a uniform flow's SUPG
weights, and the reference driver's constant-inflow steady state).
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

PARAMS = ("diffusion", "vplus", "vminus", "kplus", "kminus", "detach_plus", "detach_minus", "dt",
          "artificial_diffusion")


def param_vector(params: dict) -> np.ndarray:
    return np.array([float(params.get(k, 0.0)) for k in PARAMS])


def element(bezier, extraction, velocity, params):
    """oracle_transport_element: (left, previous), each [2 nen][2 nen], dof 2a + field."""
    extraction = np.ascontiguousarray(extraction, dtype=np.float64)
    nen = extraction.shape[0]
    left = np.zeros(4 * nen * nen)
    previous = np.zeros(4 * nen * nen)
    O._check(O.lib().oracle_transport_element(
        nen, np.ascontiguousarray(bezier, dtype=np.float64).ravel(), extraction.ravel(),
        np.ascontiguousarray(velocity, dtype=np.float64).ravel(), param_vector(params), left, previous))
    return left.reshape(2 * nen, 2 * nen), previous.reshape(2 * nen, 2 * nen)


def assemble(mesh, pattern, velocity, params, constrained):
    """Global 2x2 blocks [nnzb][2][2] of L and P on the node pattern
    (row_offsets, columns, diagonal), constrained rows applied."""
    ro, cols, diag = pattern
    nnzb = len(cols)
    L = np.zeros((nnzb, 2, 2))
    P = np.zeros((nnzb, 2, 2))
    velocity = np.asarray(velocity, dtype=np.float64).reshape(-1, 3)
    lookup = {}
    for r in range(mesh.nodes):
        for k in range(ro[r], ro[r + 1]):
            lookup[(r, int(cols[k]))] = k
    for e in range(mesh.elements):
        conn = mesh.element_connectivity(e)
        le, pe = element(mesh.element_bezier(e), mesh.dense_extraction(e), velocity[conn], params)
        nen = len(conn)
        le = le.reshape(nen, 2, nen, 2)
        pe = pe.reshape(nen, 2, nen, 2)
        for a in range(nen):
            for b in range(nen):
                k = lookup[(int(conn[a]), int(conn[b]))]
                L[k] += le[a, :, b, :]
                P[k] += pe[a, :, b, :]
    for r in np.nonzero(np.asarray(constrained))[0]:
        L[ro[r]:ro[r + 1]] = 0.0
        P[ro[r]:ro[r + 1]] = 0.0
        L[diag[r]] = np.eye(2)
    return L, P


def to_csr(pattern, blocks):
    import scipy.sparse as sp
    ro, cols, _ = pattern
    return sp.bsr_matrix((blocks, np.asarray(cols), np.asarray(ro)), shape=(2 * (len(ro) - 1),) * 2).tocsr()


def run(mesh, pattern, velocity, params, constrained, n0, nplus, steps, state=None):
    """State [N][2] after `steps` backward-Euler steps (iga_cuda.cu:674-705)
    from `state` (zero by default) with the prescribed values at the
    constrained nodes."""
    import scipy.sparse.linalg as spla
    L, P = assemble(mesh, pattern, velocity, params, constrained)
    Lc, Pc = to_csr(pattern, L), to_csr(pattern, P)
    solve = spla.factorized(Lc.tocsc())
    c = np.zeros(2 * mesh.nodes) if state is None else np.asarray(state, dtype=np.float64).ravel().copy()
    mask = np.asarray(constrained).astype(bool)
    c2 = c.reshape(-1, 2)
    c2[mask, 0] = np.asarray(n0)[mask]
    c2[mask, 1] = np.asarray(nplus)[mask]
    for _ in range(steps):
        b = (Pc @ c).reshape(-1, 2)
        b[mask, 0] = np.asarray(n0)[mask]
        b[mask, 1] = np.asarray(nplus)[mask]
        c = solve(b.ravel())
    return c.reshape(-1, 2)
