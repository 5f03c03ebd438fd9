"""The C-ABI library loads on a machine without a GPU and exports every entry
point include/tfi.h declares; the Python binding declares a signature for each.
No compute call is made here."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "tfi.h"
LIB = ROOT / "tubularflowiga_amd" / "libtfi_hip.so"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b(tfi_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ("tfi_ctx_create", "tfi_mesh_create", "tfi_ns_assemble", "tfi_bsr_spmv", "tfi_linear_solve",
                     "tfi_flow_create", "tfi_flow_newton", "tfi_flow_boundary_flows"):
        assert required in names


@pytest.mark.skipif(not LIB.exists(), reason="libtfi_hip.so is not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


@pytest.mark.skipif(not LIB.exists(), reason="libtfi_hip.so is not built (run __graft_entry__.build())")
def test_python_binding_covers_every_symbol():
    from tubularflowiga_amd import _abi
    _abi.load()
    missing = [n for n in declared_functions() if n not in _abi.SIGNATURES]
    assert not missing, f"no ctypes signature for {missing}"


def test_product_path_refuses_to_run_without_the_library(monkeypatch, tmp_path):
    """The binding raises when the HIP library is missing instead of falling
    back to any CPU implementation."""
    from tubularflowiga_amd import _abi
    monkeypatch.setattr(_abi, "_LIB", None, raising=False)
    monkeypatch.setattr(_abi, "LIB_PATH", tmp_path / "missing.so", raising=False)
    with pytest.raises((OSError, RuntimeError)):
        _abi.load()


@pytest.mark.skipif(not LIB.exists(), reason="libtfi_hip.so is not built (run __graft_entry__.build())")
def test_amg_handle_entry_points_refuse_null_handles():
    """The stand-alone AMG entry points return TFI_ERROR for a null mesh or
    handle (no device call is made)."""
    from tubularflowiga_amd import _abi
    lib = _abi.load()
    out = ctypes.c_void_p()
    assert lib.tfi_amg_create(None, 0, 3, ctypes.byref(out)) == 1
    assert lib.tfi_amg_setup(None, None) == 1
    assert lib.tfi_amg_update(None, None) == 1
    assert lib.tfi_amg_apply(None, None, None) == 1
    n = ctypes.c_int(-1)
    assert lib.tfi_amg_levels(None, 0, ctypes.byref(n), None, None, None) == 1 and n.value == -1
    lib.tfi_amg_destroy(None)
