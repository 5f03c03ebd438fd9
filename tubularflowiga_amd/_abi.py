"""ctypes binding of libtfi_hip.so (the C-ABI declared in include/tfi.h).

There is no CPU fallback: every entry point runs the HIP library, and loading
fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libtfi_hip.so"

TFI_PC_BJACOBI, TFI_PC_AMG = 0, 1
TFI_FP64, TFI_MIXED = 0, 1
TFI_WARM_PREVIOUS, TFI_WARM_LEARNED, TFI_WARM_EXTRAPOLATE, TFI_WARM_EXTRAPOLATE2, TFI_WARM_EXTRAPOLATE3 = 0, 1, 2, 3, 4
TFI_WARM_EXTRAPOLATE4, TFI_WARM_MPE = 5, 6
TFI_SMOOTH_JACOBI, TFI_SMOOTH_GAUSS_SEIDEL, TFI_SMOOTH_CHEBYSHEV, TFI_SMOOTH_AUTO = 0, 1, 2, 3
TFI_TANGENT_LEGACY, TFI_TANGENT_EXACT = 0, 1
TERM_KINDS = {"time_derivative": 0, "diffusion": 1, "advection": 2, "linear_coupling": 3, "volume_source": 4}
TFI_SUPG_NONE, TFI_SUPG_GALERKIN, TFI_SUPG_TEST, TFI_SUPG_CONSISTENT = 0, 1, 2, 3
MAX_NEWTON_RECORDS = 64

c_int, c_double, c_ll, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_void_p


class MeshStats(ctypes.Structure):
    _fields_ = [("nodes", c_int), ("elements", c_int), ("maximum_basis", c_int),
                ("basis_entries", c_ll), ("extraction_nonzeros", c_ll), ("block_nonzeros", c_ll),
                ("element_pairs", c_ll), ("colors", c_int), ("minimum_determinant", c_double),
                ("bad_samples", c_ll), ("basis_table_bytes", c_ll)]


class FlowOptions(ctypes.Structure):
    _fields_ = [("density", c_double), ("viscosity", c_double), ("dt", c_double),
                ("maximum_newton", c_int), ("nonlinear_rtol", c_double), ("nonlinear_atol", c_double),
                ("mass_rtol", c_double), ("linear_maximum_iterations", c_int), ("linear_rtol", c_double),
                ("restart", c_int), ("preconditioner", c_int), ("precision", c_int), ("warm_start", c_int),
                ("verbose", c_int), ("amg_smoother", c_int), ("tangent", c_int), ("amg_coarse_nodes", c_int),
                ("amg_gs_min_rows", c_int), ("amg_gs_sweeps", c_int), ("chord_polish", c_int),
                ("amg_chebyshev_degree", c_int), ("amg_half_smoother", c_int)]


class NewtonRecord(ctypes.Structure):
    _fields_ = [("residual_l2", c_double), ("continuity_l2", c_double), ("continuity_sum", c_double),
                ("net_boundary_flow", c_double), ("relative_mass_imbalance", c_double),
                ("update_l2", c_double), ("linear_residual", c_double), ("assembly_s", c_double),
                ("linear_s", c_double), ("linear_iterations", c_int), ("singular_diagonal_blocks", c_int),
                ("step", c_double)]


class NewtonReport(ctypes.Structure):
    _fields_ = [("converged", c_int), ("newton_iterations", c_int), ("linear_iterations", c_int),
                ("records", c_int), ("initial_residual", c_double), ("final_residual", c_double),
                ("tolerance", c_double), ("assembly_s", c_double), ("linear_s", c_double),
                ("total_s", c_double), ("warm_start_initial_ratio", c_double), ("warm_start_used", c_int),
                ("chord_steps", c_int), ("warm_start_blend", c_double), ("warm_start_base_ratio", c_double),
                ("record", NewtonRecord * MAX_NEWTON_RECORDS)]


class WarmOptions(ctypes.Structure):
    _fields_ = [("hidden", c_int), ("stencil", c_int), ("train_steps", c_int), ("learning_rate", c_double),
                ("min_snapshots", c_int), ("base", c_int)]


class TransportParams(ctypes.Structure):
    _fields_ = [("diffusion", c_double), ("vplus", c_double), ("vminus", c_double), ("kplus", c_double),
                ("kminus", c_double), ("detach_plus", c_double), ("detach_minus", c_double), ("dt", c_double),
                ("artificial_diffusion", c_double)]


class SolveTerm(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("equation", ctypes.c_int32), ("trial", ctypes.c_int32),
                ("coefficient", c_double)]


class SolveSystem(ctypes.Structure):
    _fields_ = [("fields", ctypes.c_int32), ("term_count", ctypes.c_int32), ("terms", c_void_p),
                ("supg", c_void_p), ("tau_diffusivity", c_void_p), ("dt", c_double),
                ("surface_count", ctypes.c_int32), ("surface_element", c_void_p), ("surface_face", c_void_p),
                ("surface_flux", c_void_p), ("surface_robin", c_void_p), ("surface_robin_source", c_void_p)]


class KernelTimers(ctypes.Structure):
    _fields_ = [("spmv_ms", c_double), ("spmv_calls", c_double), ("spmv_bytes", c_double),
                ("spmv32_ms", c_double), ("spmv32_calls", c_double), ("spmv32_bytes", c_double),
                ("assembly_ms", c_double), ("assembly_calls", c_double), ("assembly_flops", c_double),
                ("assembly_launches", c_double), ("residual_ms", c_double), ("residual_calls", c_double),
                ("residual_flops", c_double), ("spmv2_ms", c_double), ("spmv2_calls", c_double),
                ("spmv2_bytes", c_double), ("transport_assembly_ms", c_double)]


def _np(dtype):
    return np.ctypeslib.ndpointer(dtype=dtype, flags="C_CONTIGUOUS")


_d, _f, _i32, _u8 = _np(np.float64), _np(np.float32), _np(np.int32), _np(np.uint8)
P = ctypes.POINTER

# name -> (restype, argtypes); the list matches include/tfi.h
SIGNATURES = {
    "tfi_ctx_create": (c_int, [c_int, P(c_void_p)]),
    "tfi_ctx_destroy": (None, [c_void_p]),
    "tfi_last_error": (ctypes.c_char_p, [c_void_p]),
    "tfi_device_info": (c_int, [c_void_p, ctypes.c_char_p, c_int]),
    "tfi_allocation_stats": (c_int, [c_void_p, P(c_ll), P(c_ll)]),
    "tfi_mesh_create": (c_int, [c_void_p, c_int, c_int, _i32, _i32, _i32, _u8, _d, _d, P(c_void_p)]),
    "tfi_mesh_destroy": (None, [c_void_p]),
    "tfi_mesh_stats_get": (c_int, [c_void_p, P(MeshStats)]),
    "tfi_mesh_pattern": (c_int, [c_void_p, _i32, _i32, _i32]),
    "tfi_mesh_colors": (c_int, [c_void_p, _i32]),
    "tfi_mesh_basis_table": (c_int, [c_void_p, c_int]),
    "tfi_ns_assemble": (c_int, [c_void_p, _d, _d, c_double, c_double, c_double, c_void_p, _d, c_int]),
    "tfi_bsr_spmv": (c_int, [c_void_p, _d, _d, _d, c_int]),
    "tfi_block_inverse": (c_int, [c_void_p, _d, _d, P(c_int)]),
    "tfi_block_inverse_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, P(c_int)]),
    "tfi_apply_dirichlet": (c_int, [c_void_p, _i32, _i32, _d, _d, _d, c_void_p, c_void_p]),
    "tfi_pressure_traction": (c_int, [c_void_p, c_int, _i32, _i32, _d, _d]),
    "tfi_linear_solve": (c_int, [c_void_p, _d, _d, _d, c_int, c_int, c_int, c_int, c_double, P(c_int),
                                 P(c_double)]),
    "tfi_bench_spmv": (c_int, [c_void_p, c_int, c_int, P(c_double)]),
    "tfi_bench_copy": (c_int, [c_void_p, c_ll, c_int, P(c_double)]),
    "tfi_bench_mfma_fp64": (c_int, [c_void_p, c_int, P(c_double)]),
    "tfi_ns_assemble_device": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_double, c_double, c_void_p, c_void_p,
                                       c_int]),
    "tfi_bsr_spmv_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "tfi_apply_dirichlet_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    "tfi_amg_create": (c_int, [c_void_p, c_int, c_int, P(c_void_p)]),
    "tfi_amg_setup": (c_int, [c_void_p, c_void_p]),
    "tfi_amg_update": (c_int, [c_void_p, c_void_p]),
    "tfi_amg_apply": (c_int, [c_void_p, c_void_p, c_void_p]),
    "tfi_amg_levels": (c_int, [c_void_p, c_int, P(c_int), P(c_int), P(c_ll), P(c_int)]),
    "tfi_amg_destroy": (None, [c_void_p]),
    "tfi_flow_warm_train": (c_int, [c_void_p, _d, c_void_p, c_int, P(c_double)]),
    "tfi_bench_stream": (c_int, [c_void_p, c_ll, c_int, P(c_double), P(c_double)]),
    "tfi_bench_assembly": (c_int, [c_void_p, c_int, c_int, c_int, P(c_double)]),
    "tfi_flow_default_options": (None, [P(FlowOptions)]),
    "tfi_flow_create": (c_int, [c_void_p, P(FlowOptions), P(c_void_p)]),
    "tfi_flow_destroy": (None, [c_void_p]),
    "tfi_flow_set_state": (c_int, [c_void_p, _d, c_void_p]),
    "tfi_flow_get_state": (c_int, [c_void_p, _d]),
    "tfi_flow_set_boundary": (c_int, [c_void_p, _i32, _i32, _d, _d, c_void_p]),
    "tfi_flow_set_boundary_faces": (c_int, [c_void_p, c_int, _i32, _i32, _i32]),
    "tfi_flow_begin_step": (c_int, [c_void_p, c_void_p, c_int]),
    "tfi_flow_newton": (c_int, [c_void_p, P(NewtonReport)]),
    "tfi_flow_commit_step": (c_int, [c_void_p]),
    "tfi_flow_boundary_flows": (c_int, [c_void_p, c_int, _i32, _d]),
    "tfi_flow_warm_configure": (c_int, [c_void_p, P(WarmOptions)]),
    "tfi_flow_amg_levels": (c_int, [c_void_p, c_int, P(c_int), P(c_int), P(c_ll), P(c_int)]),
    "tfi_flow_warm_stats": (c_int, [c_void_p, P(c_int), P(c_double), P(c_int), P(c_int)]),
    "tfi_flow_warm_gate": (c_int, [c_void_p, P(c_double)]),
    "tfi_flow_warm_predict": (c_int, [c_void_p, c_int, _d, P(c_int)]),
    "tfi_flow_warm_record": (c_int, [c_void_p, _d]),
    "tfi_transport_create": (c_int, [c_void_p, P(TransportParams), _d, _i32, P(c_void_p)]),
    "tfi_transport_destroy": (None, [c_void_p]),
    "tfi_transport_set_boundary": (c_int, [c_void_p, _d, _d]),
    "tfi_transport_set_preconditioner": (c_int, [c_void_p, c_int]),
    "tfi_transport_set_state": (c_int, [c_void_p, _d]),
    "tfi_transport_get_state": (c_int, [c_void_p, _d]),
    "tfi_transport_apply_boundary": (c_int, [c_void_p]),
    "tfi_transport_step": (c_int, [c_void_p, c_int, c_int, c_double, P(c_int), P(c_double)]),
    "tfi_transport_matrices": (c_int, [c_void_p, c_void_p, c_void_p, P(c_int)]),
    "tfi_transport_bench_spmv": (c_int, [c_void_p, c_int, P(c_double)]),
    "tfi_solve_create": (c_int, [c_void_p, P(SolveSystem), _d, _i32, P(c_void_p)]),
    "tfi_solve_destroy": (None, [c_void_p]),
    "tfi_solve_set_boundary": (c_int, [c_void_p, _d]),
    "tfi_solve_set_velocity": (c_int, [c_void_p, _d]),
    "tfi_solve_set_preconditioner": (c_int, [c_void_p, c_int]),
    "tfi_solve_set_state": (c_int, [c_void_p, _d]),
    "tfi_solve_get_state": (c_int, [c_void_p, _d]),
    "tfi_solve_step": (c_int, [c_void_p, c_int, c_int, c_double, P(c_int), P(c_double)]),
    "tfi_solve_matrices": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, P(c_int)]),
    "tfi_solve_bench_spmv": (c_int, [c_void_p, c_int, P(c_double)]),
    "tfi_timers_get": (c_int, [c_void_p, P(KernelTimers)]),
    "tfi_timers_enable": (c_int, [c_void_p, c_int]),
    "tfi_timers_reset": (c_int, [c_void_p]),
    "tfi_mlp_reference": (c_int, [c_void_p, c_int, c_int, _f, _f, _f, _f, _f, _f, _f, _f, _f, P(ctypes.c_float),
                                  _f, _f, _f, _f, _f, _f]),
    "tfi_mlp_infer_bf16": (c_int, [c_void_p, c_int, c_int, _f, _f, _f, _f, _f, _f, _f, _f]),
}

_LIB = None


def load(path: Path | str | None = None):
    """Load libtfi_hip.so; raises if it is missing (no fallback path exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -m tubularflowiga_amd.build` "
                           "(the HIP library is the only implementation)")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib
