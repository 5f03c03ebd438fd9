"""Build libtfi_hip.so (HIP kernels + C-ABI, gfx950) and the `iga_hip` host CLI in-tree.

    python -m tubularflowiga_amd.build            # library + CLI
    python -m tubularflowiga_amd.build --lib-only

Every HIP translation unit is compiled by hipcc for gfx950 only, with packed
FP32 instructions disabled (they have produced silent wrong results on these
machines); the host CLI is plain C++17 built with g++ against the C-ABI.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "libtfi_hip.so"
CLI = PKG / "iga_hip"
PACK = PKG / "tfi_pack"

SOURCES = ["tfi_core.hip", "tfi_mesh.hip", "tfi_assembly.hip", "tfi_sparse.hip", "tfi_krylov.hip",
           "tfi_amg.hip", "tfi_mlp.hip", "tfi_warmstart.hip", "tfi_flow.hip", "tfi_transport.hip",
           "tfi_solve.hip"]

HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Wall", "-Wno-unused-function",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-ffp-contract=fast", "-munsafe-fp-atomics"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_library(jobs: int | None = None, verbose: bool = False, force: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    cc = hipcc()
    jobs = jobs or min(8, os.cpu_count() or 4)
    headers = sorted(CSRC.glob("*.hpp")) + [ROOT / "include" / "tfi.h"]
    newest_header = max(h.stat().st_mtime for h in headers)

    def compile_one(src: str) -> Path:
        s = CSRC / src
        o = BUILD / (s.stem + ".o")
        if not force and o.exists() and o.stat().st_mtime > max(s.stat().st_mtime, newest_header):
            return o
        out = _run([cc, *HIPCC_FLAGS, "-c", str(s), "-o", str(o)])
        if verbose and out.strip():
            print(out)
        return o

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([cc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


HDF5_DIR = Path(os.environ.get("HDF5_DIR", "/opt/conda"))
HDF5_LINKS = PKG / "_hdf5"


def hdf5_flags() -> list[str]:
    """Compile and link flags for the HDF5 C library of the image (the Bezier
    VTKHDF output), or [] when it is absent.  The library is reached through a
    private directory holding only a link to it (tubularflowiga_amd/_hdf5), so
    that neither the linker nor the loader picks up the other libraries of its
    installation (an older libstdc++ among them)."""
    header = HDF5_DIR / "include" / "hdf5.h"
    libs = sorted((HDF5_DIR / "lib").glob("libhdf5.so.*[0-9]"))
    libs = [l for l in libs if l.name.count(".") == 2]   # libhdf5.so.<soversion>
    if not header.exists() or not libs:
        return []
    HDF5_LINKS.mkdir(exist_ok=True)
    target = libs[0]
    for name in (target.name, "libhdf5.so"):
        link = HDF5_LINKS / name
        if link.is_symlink() or link.exists():
            link.unlink()
        link.symlink_to(target)
    return ["-DTFI_HAVE_HDF5", "-I", str(HDF5_DIR / "include"), "-L", str(HDF5_LINKS), "-lhdf5",
            "-Wl,-rpath,$ORIGIN/_hdf5", "-Wl,-rpath," + str(HDF5_LINKS)]


def build_cli() -> Path | None:
    """The solver CLI iga_hip (links libtfi_hip.so) and the database packer
    tfi_pack (host code only)."""
    src = HOST / "iga_hip.cpp"
    if not src.exists():
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = CLI.with_name(CLI.name + ".tmp")
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-pthread", "-I", str(ROOT / "include"), str(src), "-o", str(tmp),
          "-L", str(PKG), "-ltfi_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + str(PKG), *hdf5_flags()])
    os.replace(tmp, CLI)
    build_pack()
    return CLI


def build_pack() -> Path:
    cxx = shutil.which("g++") or "g++"
    tmp = PACK.with_name(PACK.name + ".tmp")
    _run([cxx, "-O2", "-std=c++17", "-Wall", str(HOST / "tfi_pack.cpp"), "-o", str(tmp)])
    os.replace(tmp, PACK)
    return PACK


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    ap.add_argument("--force", action="store_true", help="recompile every translation unit")
    a = ap.parse_args(argv)
    print("built", build_library(a.jobs, a.verbose, a.force))
    if not a.lib_only:
        cli = build_cli()
        if cli:
            print("built", cli)
    return 0


if __name__ == "__main__":
    sys.exit(main())
