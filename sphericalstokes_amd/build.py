"""Build recipe for libsspt.so (gfx950) and the test-support libraries.

``build_all()`` compiles, in-tree:
  * sphericalstokes_amd/libsspt.so   -- the product: HIP kernels + C ABI (hipcc, gfx950)
  * oracle/libsspt_oracle.so         -- CPU restatement used as parity checker (gcc)
  * tests/emu/libsspt_emu.so         -- host emulation of the fused kernels (g++)
  * sphericalstokes_amd/sspt_run     -- command-line driver over the C ABI (g++)

FP contraction is off everywhere (``-ffp-contract=off``) so that the device
arithmetic rounds like the reference expressions; packed FP32 instructions
are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libsspt.so"
ORACLE_DIR = ROOT / "oracle"
EMU_SRC = [ROOT / "tests" / "emu" / "fused_emu.cpp", ROOT / "tests" / "emu" / "adj_emu.cpp"]
EMU_LIB = ROOT / "tests" / "emu" / "libsspt_emu.so"

HIP_SOURCES = ["sspt_api.hip", "sspt_kernels.hip", "sspt_fused.hip", "sspt_mixed.hip", "sspt_adjoint.hip",
               "sspt_noalloc.hip", "sspt_cart.hip", "sspt_cyl.hip"]
ARCH = os.environ.get("SSPT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], cwd=cwd, check=True)


def build_lib(verbose_asm: bool = False) -> Path:
    objdir = ROOT / "_build" / "sspt"
    objdir.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-I", ROOT / "include", "-c", CSRC / src, "-o", obj]
        if verbose_asm:
            cmd += ["-Rpass-analysis=kernel-resource-usage"]
        _run(cmd)
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    _run(["make", "-C", ORACLE_DIR, "-B", "libsspt_oracle.so"])
    return ORACLE_DIR / "libsspt_oracle.so"


def build_emu() -> Path:
    tmp = EMU_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-march=x86-64-v2", *EMU_SRC, "-o", tmp])
    os.replace(tmp, EMU_LIB)
    return EMU_LIB


CLI_SRC = PKG / "cli" / "sspt_run.cpp"
CLI = PKG / "sspt_run"


def build_cli() -> Path:
    """The command-line driver over the C ABI (host code only, g++), linked
    against libsspt.so in this directory."""
    tmp = CLI.with_suffix(".tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", ROOT / "include", CLI_SRC, "-L", PKG, "-lsspt",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp])
    os.replace(tmp, CLI)
    return CLI


def build_all() -> None:
    build_oracle()
    build_emu()
    build_lib()
    build_cli()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_lib, "oracle": build_oracle, "emu": build_emu, "cli": build_cli,
         "usage": lambda: build_lib(verbose_asm=True)}[w]()
