"""Python front-end of the HIP hot path (thin wrappers over the C-ABI).

Names follow the reference's types: ``DeviceMesh`` uploads a FlatMesh and owns
the node-block pattern (DeviceMesh.hpp / BlockCsr.hpp); ``FlowSolver`` owns the
device-resident Navier-Stokes state and runs the Newton loop of
``iga_cuda navier-stokes`` (iga_cuda.cu:956-1046).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._abi import (FlowOptions, KernelTimers, MeshStats, NewtonReport, SolveSystem, SolveTerm, TransportParams, WarmOptions, TFI_FP64, TFI_MIXED,
                   TFI_PC_AMG, TFI_PC_BJACOBI, TFI_WARM_EXTRAPOLATE, TFI_WARM_EXTRAPOLATE2, TFI_WARM_EXTRAPOLATE3,
                   TFI_WARM_EXTRAPOLATE4,
                   TFI_WARM_LEARNED, TFI_WARM_MPE,
                   TFI_WARM_PREVIOUS, TFI_SMOOTH_JACOBI, TFI_SMOOTH_GAUSS_SEIDEL, TFI_SMOOTH_CHEBYSHEV,
                   TFI_SMOOTH_AUTO)

__all__ = ["Context", "DeviceMesh", "FlowSolver", "TFI_FP64", "TFI_MIXED", "TFI_PC_AMG", "TFI_PC_BJACOBI",
           "TFI_WARM_PREVIOUS", "TFI_WARM_LEARNED", "TFI_WARM_EXTRAPOLATE",
           "TFI_WARM_EXTRAPOLATE2", "TFI_WARM_EXTRAPOLATE3",
           "TFI_WARM_EXTRAPOLATE4", "TFI_WARM_MPE"]


class TfiError(RuntimeError):
    pass


class Context:
    def __init__(self, device: int = 0):
        self.lib = _abi.load()
        h = ctypes.c_void_p()
        if self.lib.tfi_ctx_create(device, ctypes.byref(h)) != 0:
            raise TfiError("tfi_ctx_create failed (no HIP device?)")
        self.handle = h

    def check(self, status: int) -> None:
        if status != 0:
            raise TfiError(self.lib.tfi_last_error(self.handle).decode())

    def bench_copy(self, count, repetitions=5):
        """Average ms of a device copy of `count` doubles (HBM counter calibration)."""
        ms = ctypes.c_double()
        self.check(self.lib.tfi_bench_copy(self.handle, int(count), repetitions, ctypes.byref(ms)))
        return ms.value

    def bench_stream(self, count, repetitions=5):
        """Measured HBM rates (GB/s): 16-byte-vector copy (read + write bytes) and read-only sum."""
        c, r = ctypes.c_double(), ctypes.c_double()
        self.check(self.lib.tfi_bench_stream(self.handle, int(count), repetitions, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def bench_mfma_fp64(self, repetitions=5):
        """Measured FP64 matrix-core TFLOP/s (probe kernel)."""
        t = ctypes.c_double()
        self.check(self.lib.tfi_bench_mfma_fp64(self.handle, repetitions, ctypes.byref(t)))
        return t.value

    def device_info(self) -> str:
        buf = ctypes.create_string_buffer(512)
        self.check(self.lib.tfi_device_info(self.handle, buf, 512))
        return buf.value.decode()

    def allocations(self):
        peak, live = ctypes.c_longlong(), ctypes.c_longlong()
        self.check(self.lib.tfi_allocation_stats(self.handle, ctypes.byref(peak), ctypes.byref(live)))
        return peak.value, live.value

    def timers(self, enable: bool | None = None, reset: bool = False) -> dict:
        if enable is not None:
            self.check(self.lib.tfi_timers_enable(self.handle, int(enable)))
        if reset:
            self.check(self.lib.tfi_timers_reset(self.handle))
        t = KernelTimers()
        self.check(self.lib.tfi_timers_get(self.handle, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in KernelTimers._fields_}

    def close(self):
        if self.handle:
            self.lib.tfi_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- test hooks (learned warm start kernels)
    def mlp_reference(self, x, target, weights, hidden):
        x = np.ascontiguousarray(x, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        w = [np.ascontiguousarray(v, np.float32).ravel() for v in weights]
        out = np.zeros((x.shape[0], 4), np.float32)
        loss = ctypes.c_float()
        grads = [np.zeros_like(v) for v in w]
        self.check(self.lib.tfi_mlp_reference(self.handle, hidden, x.shape[0], x.ravel(), target.ravel(), *w,
                                              out.ravel(), ctypes.byref(loss), *grads))
        return out, loss.value, grads

    def mlp_infer_bf16(self, x, weights, hidden):
        x = np.ascontiguousarray(x, np.float32)
        w = [np.ascontiguousarray(v, np.float32).ravel() for v in weights]
        out = np.zeros((x.shape[0], 4), np.float32)
        self.check(self.lib.tfi_mlp_infer_bf16(self.handle, hidden, x.shape[0], x.ravel(), *w, out.ravel()))
        return out


class DeviceMesh:
    """Uploads a FlatMesh and builds pattern, colouring and geometry on the GPU."""

    def __init__(self, ctx: Context, mesh):
        self.ctx = ctx
        self.lib = ctx.lib
        self.host = mesh
        h = ctypes.c_void_p()
        ctx.check(self.lib.tfi_mesh_create(
            ctx.handle, int(mesh.nodes), int(mesh.elements),
            np.ascontiguousarray(mesh.element_offsets, np.int32), np.ascontiguousarray(mesh.connectivity, np.int32),
            np.ascontiguousarray(mesh.extraction_offsets, np.int32),
            np.ascontiguousarray(mesh.extraction_columns, np.uint8),
            np.ascontiguousarray(mesh.extraction_values, np.float64),
            np.ascontiguousarray(mesh.bezier_points, np.float64), ctypes.byref(h)))
        self.handle = h
        s = MeshStats()
        ctx.check(self.lib.tfi_mesh_stats_get(h, ctypes.byref(s)))
        self.stats = {k: getattr(s, k) for k, _ in MeshStats._fields_}
        self.nodes = s.nodes
        self.nnzb = s.block_nonzeros

    def pattern(self):
        ro = np.zeros(self.nodes + 1, np.int32)
        cols = np.zeros(self.nnzb, np.int32)
        diag = np.zeros(self.nodes, np.int32)
        self.ctx.check(self.lib.tfi_mesh_pattern(self.handle, ro, cols, diag))
        return ro, cols, diag

    def basis_table(self, enable: int):
        """Build (1) or free (0) the precomputed basis table; -1 builds it if it fits
        in half of the free device memory.  Returns its size in bytes."""
        self.ctx.check(self.lib.tfi_mesh_basis_table(self.handle, int(enable)))
        s = MeshStats()
        self.ctx.check(self.lib.tfi_mesh_stats_get(self.handle, ctypes.byref(s)))
        self.stats["basis_table_bytes"] = s.basis_table_bytes
        return s.basis_table_bytes

    def colors(self):
        c = np.zeros(self.host.elements, np.int32)
        self.ctx.check(self.lib.tfi_mesh_colors(self.handle, c))
        return c

    def assemble(self, state, previous, density, viscosity, dt, jacobian=True, tangent="exact"):
        n4 = 4 * self.nodes
        negres = np.zeros(n4)
        jac = np.zeros(16 * self.nnzb) if jacobian else None
        self.ctx.check(self.lib.tfi_ns_assemble(
            self.handle, np.ascontiguousarray(state, np.float64), np.ascontiguousarray(previous, np.float64),
            density, viscosity, dt, None if jac is None else jac.ctypes.data_as(ctypes.c_void_p), negres,
            {"legacy": 0, "exact": 1}[tangent]))
        return (None if jac is None else jac.reshape(-1, 4, 4)), negres

    def spmv(self, values, x, precision=TFI_FP64):
        y = np.zeros(4 * self.nodes)
        self.ctx.check(self.lib.tfi_bsr_spmv(self.handle, np.ascontiguousarray(values, np.float64).ravel(),
                                             np.ascontiguousarray(x, np.float64), y, precision))
        return y

    def block_inverse(self, values):
        inv = np.zeros(16 * self.nodes)
        s = ctypes.c_int()
        self.ctx.check(self.lib.tfi_block_inverse(self.handle, np.ascontiguousarray(values, np.float64).ravel(),
                                                  inv, ctypes.byref(s)))
        return inv.reshape(-1, 4, 4), s.value

    def apply_dirichlet(self, vmask, pmask, vtarget, ptarget, state, jac=None, rhs=None):
        jac = None if jac is None else np.ascontiguousarray(jac, np.float64)
        rhs = None if rhs is None else np.ascontiguousarray(rhs, np.float64)
        self.ctx.check(self.lib.tfi_apply_dirichlet(
            self.handle, np.ascontiguousarray(vmask, np.int32), np.ascontiguousarray(pmask, np.int32),
            np.ascontiguousarray(vtarget, np.float64).ravel(), np.ascontiguousarray(ptarget, np.float64),
            np.ascontiguousarray(state, np.float64),
            None if jac is None else jac.ctypes.data_as(ctypes.c_void_p),
            None if rhs is None else rhs.ctypes.data_as(ctypes.c_void_p)))
        return jac, rhs

    def pressure_traction(self, element, face, pressure):
        """Traction load [4N] of faces (element, local face) at pressures (GPU face
        integrals, summed in the given face order)."""
        element = np.ascontiguousarray(element, np.int32)
        out = np.zeros(4 * self.nodes)
        self.ctx.check(self.lib.tfi_pressure_traction(
            self.handle, len(element), element, np.ascontiguousarray(face, np.int32),
            np.ascontiguousarray(pressure, np.float64), out))
        return out

    def solve(self, values, b, x0=None, preconditioner=TFI_PC_AMG, precision=TFI_FP64, restart=200,
              maxit=5000, rtol=1e-8):
        x = np.zeros(4 * self.nodes) if x0 is None else np.array(x0, np.float64)
        its, res = ctypes.c_int(), ctypes.c_double()
        self.ctx.check(self.lib.tfi_linear_solve(
            self.handle, np.ascontiguousarray(values, np.float64).ravel(), np.ascontiguousarray(b, np.float64), x,
            preconditioner, precision, restart, maxit, rtol, ctypes.byref(its), ctypes.byref(res)))
        return x, its.value, res.value

    def bench_spmv(self, precision=TFI_FP64, repetitions=50):
        ms = ctypes.c_double()
        self.ctx.check(self.lib.tfi_bench_spmv(self.handle, precision, repetitions, ctypes.byref(ms)))
        return ms.value

    def bench_assembly(self, tangent="exact", with_jacobian=True, repetitions=3):
        ms = ctypes.c_double()
        self.ctx.check(self.lib.tfi_bench_assembly(self.handle, {"legacy": 0, "exact": 1}[tangent],
                                                   1 if with_jacobian else 0, repetitions, ctypes.byref(ms)))
        return ms.value

    def assembly_flops(self, tangent="exact", with_jacobian=True):
        """Algorithmic flops of one assembly (the library's per-element count:
        the basis evaluation is not counted when the basis table supplies it)."""
        nen = np.diff(self.host.element_offsets).astype(np.float64)
        pair = (132.0 * nen * nen + 100.0 * nen if tangent == "exact" else 72.0 * nen * nen) if with_jacobian else 0.0
        basis = 0.0 if self.stats.get("basis_table_bytes", 0) else 656.0 + 42.0
        return float(np.sum(64.0 * (nen * (basis + 44.0 + 38.0) + 300.0 + pair)))

    def spmv_bytes(self, value_bytes=8):
        """Algorithmic bytes of one SpMV: nnzb (16 s_v + 4) + 4 (N+1) + 2 * 4N * 8 (x read once, y written)."""
        return self.nnzb * (16 * value_bytes + 4) + 4 * (self.nodes + 1) + 2 * 4 * self.nodes * 8

    def close(self):
        if self.handle:
            self.lib.tfi_mesh_destroy(self.handle)
            self.handle = None


class AmgPreconditioner:
    """The coupled 4x4-block aggregation AMG as a stand-alone preconditioner
    (tfi_amg_*).  Matrix values and vectors are DEVICE pointers (int or
    ctypes.c_void_p); the hierarchy keeps reading the values given to setup /
    update until the next update or close."""

    def __init__(self, dmesh: DeviceMesh, precision=TFI_FP64, smoother=TFI_SMOOTH_AUTO):
        self.dmesh, self.ctx, self.lib = dmesh, dmesh.ctx, dmesh.lib
        h = ctypes.c_void_p()
        self.ctx.check(self.lib.tfi_amg_create(dmesh.handle, precision, smoother, ctypes.byref(h)))
        self.handle = h

    def setup(self, values):
        self.ctx.check(self.lib.tfi_amg_setup(self.handle, values))

    def update(self, values):
        self.ctx.check(self.lib.tfi_amg_update(self.handle, values))

    def apply(self, v, z):
        """z = M^-1 v (one V-cycle)."""
        self.ctx.check(self.lib.tfi_amg_apply(self.handle, v, z))

    def levels(self):
        """[(rows, block_nonzeros, gauss_seidel_colours)] per level."""
        cap = 16
        n = ctypes.c_int()
        rows = (ctypes.c_int * cap)()
        nnzb = (ctypes.c_longlong * cap)()
        cols = (ctypes.c_int * cap)()
        self.ctx.check(self.lib.tfi_amg_levels(self.handle, cap, ctypes.byref(n), rows, nnzb, cols))
        return [(rows[i], nnzb[i], cols[i]) for i in range(min(n.value, cap))]

    def close(self):
        if self.handle:
            self.lib.tfi_amg_destroy(self.handle)
            self.handle = None


class FlowSolver:
    """Device-resident backward-Euler / steady Navier-Stokes solver."""

    def __init__(self, dmesh: DeviceMesh, **options):
        self.dmesh = dmesh
        self.ctx = dmesh.ctx
        self.lib = dmesh.lib
        o = FlowOptions()
        self.lib.tfi_flow_default_options(ctypes.byref(o))
        for k, v in options.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        self.options = o
        h = ctypes.c_void_p()
        self.ctx.check(self.lib.tfi_flow_create(dmesh.handle, ctypes.byref(o), ctypes.byref(h)))
        self.handle = h
        self.nodes = dmesh.nodes

    def set_state(self, state, previous=None):
        prev = None if previous is None else np.ascontiguousarray(previous, np.float64)
        self.ctx.check(self.lib.tfi_flow_set_state(self.handle, np.ascontiguousarray(state, np.float64),
                                                   None if prev is None else prev.ctypes.data_as(ctypes.c_void_p)))

    def state(self):
        out = np.zeros(4 * self.nodes)
        self.ctx.check(self.lib.tfi_flow_get_state(self.handle, out))
        return out

    def set_boundary(self, vmask, pmask, vtarget, ptarget, traction=None):
        tr = None if traction is None else np.ascontiguousarray(traction, np.float64)
        self.ctx.check(self.lib.tfi_flow_set_boundary(
            self.handle, np.ascontiguousarray(vmask, np.int32), np.ascontiguousarray(pmask, np.int32),
            np.ascontiguousarray(vtarget, np.float64).ravel(), np.ascontiguousarray(ptarget, np.float64),
            None if tr is None else tr.ctypes.data_as(ctypes.c_void_p)))

    def set_boundary_faces(self, element, face, label):
        element = np.ascontiguousarray(element, np.int32)
        self.ctx.check(self.lib.tfi_flow_set_boundary_faces(self.handle, len(element), element,
                                                            np.ascontiguousarray(face, np.int32),
                                                            np.ascontiguousarray(label, np.int32)))

    def begin_step(self, phase=None):
        ph = None if phase is None else np.ascontiguousarray(phase, np.float64)
        self.ctx.check(self.lib.tfi_flow_begin_step(self.handle,
                                                    None if ph is None else ph.ctypes.data_as(ctypes.c_void_p),
                                                    0 if ph is None else len(ph)))

    def newton(self) -> dict:
        r = NewtonReport()
        status = self.lib.tfi_flow_newton(self.handle, ctypes.byref(r))
        rep = {k: getattr(r, k) for k, _ in NewtonReport._fields_ if k != "record"}
        rep["records"] = [{k: getattr(r.record[i], k) for k, _ in r.record[i]._fields_} for i in range(r.records)]
        if status != 0:
            err = TfiError(self.lib.tfi_last_error(self.ctx.handle).decode())
            err.report = rep  # the iterations done before the failure
            raise err
        return rep

    def commit_step(self):
        self.ctx.check(self.lib.tfi_flow_commit_step(self.handle))

    def boundary_flows(self, labels):
        labels = np.ascontiguousarray(labels, np.int32)
        out = np.zeros(len(labels))
        self.ctx.check(self.lib.tfi_flow_boundary_flows(self.handle, len(labels), labels, out))
        return out

    def configure_warm_start(self, hidden=128, train_steps=40, learning_rate=1e-3, min_snapshots=3, base=0):
        """base: TFI_WARM_MPE (0: default) or TFI_WARM_EXTRAPOLATE4, the
        prediction whose error the network learns."""
        o = WarmOptions(hidden, 27, train_steps, learning_rate, min_snapshots, base)
        self.ctx.check(self.lib.tfi_flow_warm_configure(self.handle, ctypes.byref(o)))

    def amg_levels(self):
        """[(rows, block_nonzeros, gauss_seidel_colours)] of the AMG hierarchy (after the first solve)."""
        cap = 16
        n = ctypes.c_int()
        rows = (ctypes.c_int * cap)()
        nnzb = (ctypes.c_longlong * cap)()
        cols = (ctypes.c_int * cap)()
        self.ctx.check(self.lib.tfi_flow_amg_levels(self.handle, cap, ctypes.byref(n), rows, nnzb, cols))
        return [(rows[i], nnzb[i], cols[i]) for i in range(min(n.value, cap))]

    def warm_record(self, state):
        """Append a state to the warm-start history (no training)."""
        self.ctx.check(self.lib.tfi_flow_warm_record(self.handle, np.ascontiguousarray(state, np.float64)))

    def warm_train(self, state, phase=None):
        """Append a state to the history and train the learned predictor on the
        snapshot it forms (once five states precede it); returns the last
        training loss (0.0 before the first snapshot)."""
        ph = np.zeros(0) if phase is None else np.ascontiguousarray(phase, np.float64)
        loss = ctypes.c_double()
        self.ctx.check(self.lib.tfi_flow_warm_train(self.handle, np.ascontiguousarray(state, np.float64),
                                                    ph.ctypes.data_as(ctypes.c_void_p) if len(ph) else None,
                                                    len(ph), ctypes.byref(loss)))
        return loss.value

    def warm_predict(self, mode):
        """(state the predicting `mode` would start the next step from, order
        of the increment extrapolation for TFI_WARM_MPE)."""
        out = np.zeros(4 * self.nodes)
        order = ctypes.c_int()
        self.ctx.check(self.lib.tfi_flow_warm_predict(self.handle, mode, out, ctypes.byref(order)))
        return out, order.value

    def warm_stats(self):
        s, l, a, r = ctypes.c_int(), ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        self.ctx.check(self.lib.tfi_flow_warm_stats(self.handle, ctypes.byref(s), ctypes.byref(l), ctypes.byref(a),
                                                    ctypes.byref(r)))
        g = ctypes.c_double()
        self.ctx.check(self.lib.tfi_flow_warm_gate(self.handle, ctypes.byref(g)))
        return {"snapshots": s.value, "last_loss": l.value, "accepted": a.value, "rejected": r.value,
                "gate": g.value}

    def close(self):
        if self.handle:
            self.lib.tfi_flow_destroy(self.handle)
            self.handle = None


class Transport:
    """Two-species neuron transport on the mesh pattern (2x2 node blocks
    [N0, Nplus]); see tfi_transport_* in include/tfi.h."""

    def __init__(self, dmesh: DeviceMesh, params: dict, velocity, constrained):
        self.dmesh = dmesh
        self.ctx = dmesh.ctx
        self.lib = dmesh.lib
        self.nodes = dmesh.nodes
        p = TransportParams(*[float(params.get(k, 0.0)) for k, _ in TransportParams._fields_])
        h = ctypes.c_void_p()
        self.ctx.check(self.lib.tfi_transport_create(
            dmesh.handle, ctypes.byref(p), np.ascontiguousarray(velocity, np.float64).ravel(),
            np.ascontiguousarray(constrained, np.int32), ctypes.byref(h)))
        self.handle = h

    def set_preconditioner(self, preconditioner):
        """TFI_PC_BJACOBI (default) or TFI_PC_AMG."""
        self.ctx.check(self.lib.tfi_transport_set_preconditioner(self.handle, preconditioner))

    def set_boundary(self, n0, nplus):
        self.ctx.check(self.lib.tfi_transport_set_boundary(self.handle, np.ascontiguousarray(n0, np.float64),
                                                           np.ascontiguousarray(nplus, np.float64)))

    def apply_boundary(self):
        self.ctx.check(self.lib.tfi_transport_apply_boundary(self.handle))

    def set_state(self, state):
        self.ctx.check(self.lib.tfi_transport_set_state(self.handle, np.ascontiguousarray(state, np.float64).ravel()))

    def state(self):
        out = np.zeros(2 * self.nodes)
        self.ctx.check(self.lib.tfi_transport_get_state(self.handle, out))
        return out.reshape(-1, 2)

    def step(self, zero_guess=False, maximum_iterations=10000, rtol=1e-8):
        it, res = ctypes.c_int(), ctypes.c_double()
        self.ctx.check(self.lib.tfi_transport_step(self.handle, int(zero_guess), maximum_iterations, rtol,
                                                   ctypes.byref(it), ctypes.byref(res)))
        return it.value, res.value

    def bench_spmv(self, repetitions=50):
        """Average ms per launch of the 2x2 SpMV, timed back to back."""
        ms = ctypes.c_double()
        self.ctx.check(self.lib.tfi_transport_bench_spmv(self.handle, repetitions, ctypes.byref(ms)))
        return ms.value

    def matrices(self):
        nnzb = self.dmesh.nnzb
        left, prev = np.zeros(4 * nnzb), np.zeros(4 * nnzb)
        s = ctypes.c_int()
        self.ctx.check(self.lib.tfi_transport_matrices(self.handle, left.ctypes.data_as(ctypes.c_void_p),
                                                       prev.ctypes.data_as(ctypes.c_void_p), ctypes.byref(s)))
        return left.reshape(-1, 2, 2), prev.reshape(-1, 2, 2), s.value

    def close(self):
        """Release the device buffers (before the context is closed)."""
        if self.handle:
            self.lib.tfi_transport_destroy(self.handle)
            self.handle = None


class ConfiguredTransport:
    """A linear_transport system of F coupled scalar fields (F x F node
    blocks); see tfi_solve_* in include/tfi.h.

    terms: [(kind, equation, trial, coefficient)] with kind a TERM_KINDS
    value; supg, tau_diffusivity: per field; surfaces: optional dict with
    element, face and per-face [count][F] flux, robin, robin_source."""

    def __init__(self, dmesh: DeviceMesh, fields, terms, supg, tau_diffusivity, dt, velocity, constrained,
                 surfaces=None):
        self.dmesh = dmesh
        self.ctx = dmesh.ctx
        self.lib = dmesh.lib
        self.nodes = dmesh.nodes
        self.fields = int(fields)
        term_array = (SolveTerm * max(1, len(terms)))(*[SolveTerm(int(k), int(e), int(t), float(c))
                                                        for k, e, t, c in terms])
        keep = [term_array]

        def arr(values, dtype):
            a = np.ascontiguousarray(values, dtype=dtype).ravel()
            keep.append(a)
            return a.ctypes.data_as(ctypes.c_void_p)

        sysdef = SolveSystem()
        sysdef.fields = self.fields
        sysdef.term_count = len(terms)
        sysdef.terms = ctypes.cast(term_array, ctypes.c_void_p)
        sysdef.supg = arr(supg, np.int32)
        sysdef.tau_diffusivity = arr(tau_diffusivity, np.float64)
        sysdef.dt = float(dt)
        if surfaces is not None and len(surfaces["element"]):
            sysdef.surface_count = len(surfaces["element"])
            sysdef.surface_element = arr(surfaces["element"], np.int32)
            sysdef.surface_face = arr(surfaces["face"], np.int32)
            sysdef.surface_flux = arr(surfaces["flux"], np.float64)
            sysdef.surface_robin = arr(surfaces["robin"], np.float64)
            sysdef.surface_robin_source = arr(surfaces["robin_source"], np.float64)
        h = ctypes.c_void_p()
        self.ctx.check(self.lib.tfi_solve_create(
            dmesh.handle, ctypes.byref(sysdef), np.ascontiguousarray(velocity, np.float64).ravel(),
            np.ascontiguousarray(constrained, np.int32).ravel(), ctypes.byref(h)))
        self.handle = h

    def set_preconditioner(self, preconditioner):
        """TFI_PC_BJACOBI (default) or TFI_PC_AMG (at most 4 fields)."""
        self.ctx.check(self.lib.tfi_solve_set_preconditioner(self.handle, preconditioner))

    def set_velocity(self, velocity):
        """Re-assemble for a new nodal velocity [N][3]."""
        self.ctx.check(self.lib.tfi_solve_set_velocity(self.handle,
                                                       np.ascontiguousarray(velocity, np.float64).ravel()))

    def set_boundary(self, values):
        self.ctx.check(self.lib.tfi_solve_set_boundary(self.handle,
                                                       np.ascontiguousarray(values, np.float64).ravel()))

    def set_state(self, state):
        self.ctx.check(self.lib.tfi_solve_set_state(self.handle, np.ascontiguousarray(state, np.float64).ravel()))

    def state(self):
        out = np.zeros(self.fields * self.nodes)
        self.ctx.check(self.lib.tfi_solve_get_state(self.handle, out))
        return out.reshape(-1, self.fields)

    def step(self, zero_guess=False, maximum_iterations=10000, rtol=1e-8):
        it, res = ctypes.c_int(), ctypes.c_double()
        self.ctx.check(self.lib.tfi_solve_step(self.handle, int(zero_guess), maximum_iterations, rtol,
                                               ctypes.byref(it), ctypes.byref(res)))
        return it.value, res.value

    def bench_spmv(self, repetitions=50):
        """Average ms per launch of the F x F SpMV, timed back to back."""
        ms = ctypes.c_double()
        self.ctx.check(self.lib.tfi_solve_bench_spmv(self.handle, repetitions, ctypes.byref(ms)))
        return ms.value

    def matrices(self):
        F, nnzb = self.fields, self.dmesh.nnzb
        left, prev, src = np.zeros(F * F * nnzb), np.zeros(F * F * nnzb), np.zeros(F * self.nodes)
        s = ctypes.c_int()
        self.ctx.check(self.lib.tfi_solve_matrices(self.handle, left.ctypes.data_as(ctypes.c_void_p),
                                                   prev.ctypes.data_as(ctypes.c_void_p),
                                                   src.ctypes.data_as(ctypes.c_void_p), ctypes.byref(s)))
        return left.reshape(-1, F, F), prev.reshape(-1, F, F), src.reshape(-1, F), s.value

    def close(self):
        """Release the device buffers (before the context is closed)."""
        if self.handle:
            self.lib.tfi_solve_destroy(self.handle)
            self.handle = None
