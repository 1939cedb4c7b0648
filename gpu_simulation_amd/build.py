"""Build the engine: librie.so (HIP, gfx950) and the pybind11 module _cuda_backend.

    python -m gpu_simulation_amd.build

* ``gpu_simulation_amd/lib/librie.so``: HIP kernels + C++ host orchestration
  behind the C ABI of ``include/rie.h``, compiled with hipcc for gfx950.
* ``gpu_simulation_amd/standalone_backend/_cuda_backend*.so``: the pybind11
  module with the reference extension's Python surface, compiled with the
  host C++ compiler against ``rie.h`` only and linked to librie.so.
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OBJ_DIR = ROOT / "_build"
LIB_DIR = PKG / "lib"
LIB_PATH = LIB_DIR / "librie.so"
EXT_DIR = PKG / "standalone_backend"
EXT_PATH = EXT_DIR / ("_cuda_backend" + (sysconfig.get_config_var("EXT_SUFFIX") or ".so"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OFFLOAD_ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={OFFLOAD_ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 VALU ops stay disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
HIP_SOURCES = ["rie_kernels.hip", "rie_dense.hip", "rie_batch.hip", "rie_mlp.hip",
               "rie_noalloc.hip", "rie_engine.cpp", "rie_batch.cpp", "rie_mlp.cpp"]


def _run_parallel(commands: list[list[str]], verbose: bool) -> None:
    procs = []
    for cmd in commands:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for cmd, proc in procs:
        out, _ = proc.communicate()
        text = out.decode(errors="replace")
        if proc.returncode != 0:
            failed.append((cmd, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + t for c, t in failed)
        raise RuntimeError(f"build failed:\n{msg}")


def build(verbose: bool = False) -> None:
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    objects = []
    commands = []
    for src in HIP_SOURCES:
        obj = OBJ_DIR / (Path(src).name.replace(".", "_") + ".o")
        objects.append(str(obj))
        commands.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)])
    _run_parallel(commands, verbose)
    tmp_lib = LIB_PATH.with_suffix(".so.tmp")
    _run_parallel([[HIPCC, f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", "-o",
                    str(tmp_lib), *objects]], verbose)
    os.replace(tmp_lib, LIB_PATH)

    import pybind11

    py_include = sysconfig.get_paths()["include"]
    tmp_ext = EXT_PATH.with_name(EXT_PATH.name + ".tmp")
    cxx = os.environ.get("CXX", "g++")
    _run_parallel([[
        cxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
        str(CSRC / "bindings.cpp"),
        f"-I{INCLUDE}", f"-I{pybind11.get_include()}", f"-I{py_include}",
        f"-L{LIB_DIR}", "-lrie", "-Wl,-rpath,$ORIGIN/../lib",
        "-o", str(tmp_ext),
    ]], verbose)
    os.replace(tmp_ext, EXT_PATH)


def is_built() -> bool:
    return LIB_PATH.exists() and EXT_PATH.exists()


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(f"built {LIB_PATH.relative_to(ROOT)} and {EXT_PATH.relative_to(ROOT)}")
