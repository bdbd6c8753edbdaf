"""In-tree build of the HIP library and its Python binding.

``build()`` compiles ``egg_amd/_lib/libeggmi.so`` (HIP kernels for gfx950 +
host session + C ABI, via ``egg_amd/csrc/Makefile`` and hipcc) and the
pybind11 module ``egg_amd/_cpp/cpp_core*.so`` (g++, linked against
libeggmi.so with an $ORIGIN rpath). Both stay in the source tree.
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "_lib" / "libeggmi.so"


def module_path() -> Path:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return PKG / "_cpp" / f"cpp_core{suffix}"


def build(jobs: int = 4, verbose: bool = False) -> None:
    env = dict(os.environ)
    out = None if verbose else subprocess.DEVNULL
    subprocess.run(["make", "-C", str(PKG / "csrc"), f"-j{jobs}"], check=True, env=env, stdout=out)
    import pybind11

    cmd = [
        "g++", "-O2", "-shared", "-fPIC", "-std=c++23",
        f"-I{sysconfig.get_paths()['include']}", f"-I{pybind11.get_include()}", f"-I{ROOT / 'include'}",
        str(PKG / "_cpp" / "cpp_core.cpp"), "-o", str(module_path()),
        f"-L{LIB.parent}", "-leggmi", "-Wl,-rpath,$ORIGIN/../_lib",
    ]
    subprocess.run(cmd, check=True, stdout=out)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
