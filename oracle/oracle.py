"""ctypes front-end of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Every wrapper names the oracle C function it calls; the C functions cite the
reference file:line they restate.  ``ns_newton_step`` restates the Newton loop of
the reference driver (solvers/cuda/src/iga_cuda.cu:956-1046) on top of them.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build(force: bool = False) -> Path:
    """Compile oracle/liboracle.so with g++ (OpenMP)."""
    so = _HERE / "liboracle.so"
    src = _HERE / "tfi_oracle.cpp"
    if force or not so.exists() or so.stat().st_mtime < src.stat().st_mtime:
        subprocess.run(["make", "-C", str(_HERE), "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = _HERE / "liboracle.so"
        if not so.exists():
            build()
        L = ctypes.CDLL(str(so))
        L.oracle_last_error.restype = ctypes.c_char_p
        c_int, c_dbl, c_ll = ctypes.c_int, ctypes.c_double, ctypes.c_longlong
        vp = ctypes.c_void_p
        L.oracle_reference_tables.argtypes = [_dp, _dp, _dp, _dp]
        L.oracle_evaluate_basis.argtypes = [c_int, _dp, _dp, c_dbl, c_dbl, c_dbl, _dp, _dp, vp, vp,
                                            ctypes.POINTER(c_dbl)]
        L.oracle_ns_element.argtypes = [c_int, _dp, _dp, _dp, _dp, c_dbl, c_dbl, c_dbl, c_int, c_int,
                                        vp, vp, vp, _dp]
        L.oracle_transport_element.argtypes = [c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_block_pattern.argtypes = [c_int, c_int, _ip, _ip, vp, vp, vp, ctypes.POINTER(c_ll)]
        L.oracle_ns_assemble.argtypes = [c_int, c_int, _ip, _ip, _ip, _u8p, _dp, _dp, _ip, _ip, _dp, _dp,
                                         c_dbl, c_dbl, c_dbl, c_int, vp, _dp]
        L.oracle_apply_dirichlet.argtypes = [c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, vp, vp]
        L.oracle_bsr_spmv.argtypes = [c_int, _ip, _ip, _dp, _dp, _dp]
        L.oracle_block_inverse.argtypes = [c_int, _ip, _dp, _dp, ctypes.POINTER(c_int)]
        for name in ("oracle_gmres_bjacobi", "oracle_fgmres_bjacobi"):
            getattr(L, name).argtypes = [c_int, _ip, _ip, _ip, _dp, _dp, _dp, c_int, c_int, c_dbl,
                                         ctypes.POINTER(c_int), ctypes.POINTER(c_dbl)]
        L.oracle_boundary_flow.argtypes = [c_int, _dp, _dp, _dp, c_int, ctypes.POINTER(c_dbl)]
        L.oracle_pressure_traction.argtypes = [c_int, _dp, _dp, c_int, c_dbl, _dp]
        L.oracle_num_threads.restype = c_int
        L.oracle_time_elements.argtypes = [c_int, c_int, _ip, _ip, _ip, _u8p, _dp, _dp, _dp, c_dbl, c_dbl, c_dbl,
                                           c_int, c_int, ctypes.POINTER(c_dbl)]
        _LIB = L
    return _LIB


def _check(status: int) -> None:
    if status != 0:
        raise RuntimeError("oracle: " + lib().oracle_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def num_threads() -> int:
    return lib().oracle_num_threads()


# ---------------------------------------------------------------- element level
def reference_tables():
    """oracle_reference_tables: Bernstein tables of solvers/cuda/src/iga_cuda.cu:73-184."""
    basis = np.zeros(64 * 64)
    grad = np.zeros(64 * 64 * 3)
    hess = np.zeros(64 * 64 * 9)
    weight = np.zeros(64)
    _check(lib().oracle_reference_tables(basis, grad, hess, weight))
    return basis.reshape(64, 64), grad.reshape(64, 64, 3), hess.reshape(64, 64, 3, 3), weight


def evaluate_basis(bezier, extraction, u, v, w, hessian=True):
    """oracle_evaluate_basis: TransportElement.hpp:118-205."""
    bezier = np.ascontiguousarray(bezier, dtype=np.float64).reshape(64, 3)
    extraction = np.ascontiguousarray(extraction, dtype=np.float64)
    nen = extraction.shape[0]
    value = np.zeros(nen)
    grad = np.zeros(nen * 3)
    hess = np.zeros(nen * 9) if hessian else None
    inv = np.zeros(9)
    det = ctypes.c_double()
    _check(lib().oracle_evaluate_basis(nen, bezier.ravel(), extraction.ravel(), u, v, w, value, grad,
                                       _ptr(hess), _ptr(inv), ctypes.byref(det)))
    out = {"value": value, "gradient": grad.reshape(nen, 3), "inverse": inv.reshape(3, 3),
           "raw_determinant": det.value}
    if hessian:
        out["hessian"] = hess.reshape(nen, 3, 3)
    return out


def ns_element(bezier, extraction, state, previous, rho, mu, dt, tangent="legacy",
               points=None, weights=None, jacobian=True):
    """oracle_ns_element: negated residual (4nen) and tangent ((4nen)^2) of one element.

    tangent: "legacy" (tau frozen; IgaCudaKernels.cuh:380-396) or "exact"
    (NavierStokesElement.hpp:244-313).  points/weights default to the full
    4x4x4 rule with weights/8 (Quadrature.hpp:193-202).
    """
    bezier = np.ascontiguousarray(bezier, dtype=np.float64).reshape(64, 3)
    extraction = np.ascontiguousarray(extraction, dtype=np.float64)
    nen = extraction.shape[0]
    state = np.ascontiguousarray(state, dtype=np.float64).reshape(nen, 4)
    previous = np.zeros((nen, 4)) if previous is None else np.ascontiguousarray(previous, dtype=np.float64).reshape(nen, 4)
    negres = np.zeros(4 * nen)
    jac = np.zeros((4 * nen) ** 2) if jacobian else None
    nq = 0
    qp = qw = None
    if points is not None:
        qp = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        qw = np.ascontiguousarray(weights, dtype=np.float64).ravel()
        nq = qp.shape[0]
    kind = {"legacy": 0, "exact": 1}[tangent]
    _check(lib().oracle_ns_element(nen, bezier.ravel(), extraction.ravel(), state.ravel(), previous.ravel(),
                                   rho, mu, dt, kind, nq, _ptr(qp), _ptr(qw), _ptr(jac), negres))
    return (jac.reshape(4 * nen, 4 * nen) if jacobian else None), negres


def boundary_flow(bezier, extraction, state, face):
    """oracle_boundary_flow: BoundaryFlow.hpp:164-200."""
    extraction = np.ascontiguousarray(extraction, dtype=np.float64)
    nen = extraction.shape[0]
    out = ctypes.c_double()
    _check(lib().oracle_boundary_flow(nen, np.ascontiguousarray(bezier, dtype=np.float64).ravel(),
                                      extraction.ravel(),
                                      np.ascontiguousarray(state, dtype=np.float64).ravel(), face,
                                      ctypes.byref(out)))
    return out.value


def pressure_traction(bezier, extraction, face, pressure):
    """oracle_pressure_traction: BoundaryFlow.hpp:445-477."""
    extraction = np.ascontiguousarray(extraction, dtype=np.float64)
    nen = extraction.shape[0]
    local = np.zeros(4 * nen)
    _check(lib().oracle_pressure_traction(nen, np.ascontiguousarray(bezier, dtype=np.float64).ravel(),
                                          extraction.ravel(), face, pressure, local))
    return local


# ---------------------------------------------------------------- global level
def block_pattern(mesh):
    """oracle_block_pattern: BlockCsr.hpp:27-55 -> (row_offsets, columns, diagonal)."""
    L = lib()
    n = ctypes.c_longlong()
    _check(L.oracle_block_pattern(mesh.nodes, mesh.elements, mesh.element_offsets, mesh.connectivity,
                                  None, None, None, ctypes.byref(n)))
    ro = np.zeros(mesh.nodes + 1, np.int32)
    cols = np.zeros(n.value, np.int32)
    diag = np.zeros(mesh.nodes, np.int32)
    _check(L.oracle_block_pattern(mesh.nodes, mesh.elements, mesh.element_offsets, mesh.connectivity,
                                  _ptr(ro), _ptr(cols), _ptr(diag), ctypes.byref(n)))
    return ro, cols, diag


def ns_assemble(mesh, pattern, state, previous, rho, mu, dt, tangent="legacy", jacobian=True):
    """oracle_ns_assemble: global negated residual and block Jacobian (nnzb,4,4)."""
    ro, cols, _ = pattern
    negres = np.zeros(4 * mesh.nodes)
    jac = np.zeros(16 * len(cols)) if jacobian else None
    _check(lib().oracle_ns_assemble(mesh.nodes, mesh.elements, mesh.element_offsets, mesh.connectivity,
                                    mesh.extraction_offsets, mesh.extraction_columns, mesh.extraction_values,
                                    mesh.bezier_points, ro, cols,
                                    np.ascontiguousarray(state, dtype=np.float64),
                                    np.ascontiguousarray(previous, dtype=np.float64),
                                    rho, mu, dt, {"legacy": 0, "exact": 1}[tangent], _ptr(jac), negres))
    return (jac.reshape(-1, 4, 4) if jacobian else None), negres


def apply_dirichlet(pattern, vmask, pmask, vtarget, ptarget, state, jac=None, rhs=None):
    """oracle_apply_dirichlet: IgaCudaKernels.cuh:436-463 (in place)."""
    ro, _, diag = pattern
    nodes = len(diag)
    _check(lib().oracle_apply_dirichlet(nodes, ro, diag, np.ascontiguousarray(vmask, dtype=np.int32),
                                        np.ascontiguousarray(pmask, dtype=np.int32),
                                        np.ascontiguousarray(vtarget, dtype=np.float64).ravel(),
                                        np.ascontiguousarray(ptarget, dtype=np.float64),
                                        np.ascontiguousarray(state, dtype=np.float64),
                                        _ptr(jac), _ptr(rhs)))


def bsr_spmv(pattern, values, x):
    ro, cols, diag = pattern
    y = np.zeros(4 * len(diag))
    _check(lib().oracle_bsr_spmv(len(diag), ro, cols, np.ascontiguousarray(values, dtype=np.float64).ravel(),
                                 np.ascontiguousarray(x, dtype=np.float64), y))
    return y


def block_inverse(pattern, values):
    ro, cols, diag = pattern
    inv = np.zeros(16 * len(diag))
    singular = ctypes.c_int()
    _check(lib().oracle_block_inverse(len(diag), diag, np.ascontiguousarray(values, dtype=np.float64).ravel(),
                                      inv, ctypes.byref(singular)))
    return inv.reshape(-1, 4, 4), singular.value


def _krylov(name, pattern, values, b, x0, restart, maxit, rtol):
    ro, cols, diag = pattern
    x = np.array(x0 if x0 is not None else np.zeros(4 * len(diag)), dtype=np.float64)
    its = ctypes.c_int()
    res = ctypes.c_double()
    _check(getattr(lib(), name)(len(diag), ro, cols, diag, np.ascontiguousarray(values, dtype=np.float64).ravel(),
                                np.ascontiguousarray(b, dtype=np.float64), x, restart, maxit, rtol,
                                ctypes.byref(its), ctypes.byref(res)))
    return x, its.value, res.value


def gmres_bjacobi(pattern, values, b, x0=None, restart=200, maxit=5000, rtol=1e-8):
    """oracle_gmres_bjacobi: Gmres.hpp:69-213 (left-preconditioned)."""
    return _krylov("oracle_gmres_bjacobi", pattern, values, b, x0, restart, maxit, rtol)


def fgmres_bjacobi(pattern, values, b, x0=None, restart=200, maxit=5000, rtol=1e-8):
    """oracle_fgmres_bjacobi: right-preconditioned, true-residual test."""
    return _krylov("oracle_fgmres_bjacobi", pattern, values, b, x0, restart, maxit, rtol)


def ns_newton_step(mesh, pattern, state, previous, rho, mu, dt, vmask, pmask, vtarget, ptarget,
                   traction=None, max_newton=12, rtol=1e-5, atol=1e-10, linear_rtol=1e-8,
                   tangent="legacy", restart=200, maxit=20000, log=None):
    """One nonlinear (Newton) solve of a time step, restating the reference
    driver's loop (solvers/cuda/src/iga_cuda.cu:956-1046): assemble, add the
    traction load, apply Dirichlet rows, test ||r|| <= max(atol sqrt(4N),
    rtol ||r0||), solve J du = r to ``linear_rtol`` and update.  The mass
    imbalance test is the caller's (it needs boundary faces).  Returns
    (state, newton_iterations, linear_iterations, residual_history)."""
    state = np.array(state, dtype=np.float64)
    n4 = 4 * mesh.nodes
    scale = np.sqrt(n4)
    r0 = None
    history = []
    linear = 0
    for it in range(max_newton):
        jac, rhs = ns_assemble(mesh, pattern, state, previous, rho, mu, dt, tangent)
        if traction is not None:
            rhs += traction
        apply_dirichlet(pattern, vmask, pmask, vtarget, ptarget, state, jac, rhs)
        norm = float(np.linalg.norm(rhs))
        history.append(norm)
        if r0 is None:
            r0 = norm
        if norm <= max(atol * scale, rtol * r0):
            return state, it, linear, history
        du, its, res = fgmres_bjacobi(pattern, jac, rhs, None, restart, maxit, linear_rtol)
        linear += its
        if res > linear_rtol * norm * (1 + 1e-6):
            raise RuntimeError(f"oracle FGMRES did not converge ({res} vs {norm})")
        state += du
        if log:
            log(f"oracle newton={it} residual={norm:.6e} linear_iterations={its}")
    raise RuntimeError("oracle Newton reached max_newton")


def time_elements(mesh, state, rho, mu, dt, first, count):
    """oracle_time_elements: wall seconds to build `count` element systems (CPU baseline sample)."""
    sec = ctypes.c_double()
    _check(lib().oracle_time_elements(mesh.nodes, mesh.elements, mesh.element_offsets, mesh.connectivity,
                                      mesh.extraction_offsets, mesh.extraction_columns, mesh.extraction_values,
                                      mesh.bezier_points, np.ascontiguousarray(state, dtype=np.float64),
                                      rho, mu, dt, first, count, ctypes.byref(sec)))
    return sec.value
