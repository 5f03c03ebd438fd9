"""The C-ABI from C: include/tfi.h compiles as C99, a C program that takes the
address of every entry point the header declares links against
libtfi_hip.so, and the one entry point that needs no GPU
(tfi_flow_default_options) returns the reference defaults
(iga_cuda.cu:402-405, 1015) with this library's solver settings.  No GPU
call is made."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from test_abi import HEADER, LIB, declared_functions

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(not LIB.exists(), reason="libtfi_hip.so is not built (run __graft_entry__.build())")
@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_c_program_links_every_entry_point(tmp_path):
    names = declared_functions()
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    for n in names:
        assert re.search(r"\b" + n + r"\s*\(", text)
    table = ",\n".join(f"    (void (*)(void)){n}" for n in names)
    src = tmp_path / "abi.c"
    src.write_text(f"""#include <stdio.h>
#include "tfi.h"

static void (*const entry_points[])(void) = {{
{table}
}};

int main(void)
{{
    tfi_flow_options o;
    int i, linked = 0;
    for (i = 0; i < (int)(sizeof entry_points/sizeof entry_points[0]); ++i) linked += entry_points[i] != 0;
    tfi_flow_default_options(&o);
    printf("linked=%d density=%g viscosity=%g maximum_newton=%d nonlinear_rtol=%g nonlinear_atol=%g "
           "mass_rtol=%g linear_maximum_iterations=%d linear_rtol=%g precision=%d smoother=%d half=%d tangent=%d\\n",
           linked, o.density, o.viscosity, o.maximum_newton, o.nonlinear_rtol, o.nonlinear_atol, o.mass_rtol,
           o.linear_maximum_iterations, o.linear_rtol, o.precision, o.amg_smoother, o.amg_half_smoother, o.tangent);
    return 0;
}}
""")
    exe = tmp_path / "abi"
    pkg = LIB.parent
    build = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wno-pedantic", "-I", str(HEADER.parent), str(src), "-o", str(exe),
                            "-L", str(pkg), "-ltfi_hip", f"-Wl,-rpath,{pkg}"], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    fields = dict(kv.split("=") for kv in run.stdout.split())
    assert int(fields["linked"]) == len(names)
    assert float(fields["density"]) == 1.0 and float(fields["viscosity"]) == 0.1
    assert int(fields["maximum_newton"]) == 12
    assert float(fields["nonlinear_rtol"]) == 1e-5 and float(fields["nonlinear_atol"]) == 1e-10
    assert float(fields["mass_rtol"]) == 1e-3
    assert int(fields["linear_maximum_iterations"]) == 5000 and float(fields["linear_rtol"]) == 1e-8
    assert int(fields["precision"]) == 0 and int(fields["smoother"]) == 3 and int(fields["half"]) == 1
    assert int(fields["tangent"]) == 1
