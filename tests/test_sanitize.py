"""The CPU oracle and the host emulations of the fused forward and
r-marching adjoint kernels built with -fsanitize=address,undefined and run
over the fused kernels' tiling-edge shapes (tests/emu/sanitize_main.cpp):
no out-of-bounds access, no undefined behaviour, and the emulated
iterations and adjoint step (FP64, and the FP32 right-hand-side instance
of the mixed-precision adjoint) agree with the oracle.  Every field access
through the kernels' 32-bit byte offsets is also checked against the box
allocation it is based on."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EMU = ROOT / "tests" / "emu"
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1",
       "-ffp-contract=off", "-fno-fast-math"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc/g++")
def test_sanitized_oracle_and_emulation(tmp_path):
    objs = []
    for src in ("sspt_oracle.c", "sspt_adjoint.c"):
        o = tmp_path / (src + ".o")
        subprocess.run(["gcc", "-std=c11", *SAN, "-c", str(ROOT / "oracle" / src), "-o", str(o)], check=True)
        objs.append(o)
    for src in ("fused_emu.cpp", "adj_emu.cpp"):
        o = tmp_path / (src + ".o")
        subprocess.run(["g++", "-std=c++17", *SAN, "-include", str(EMU / "access_check.h"), "-c", str(EMU / src),
                        "-o", str(o)], check=True)
        objs.append(o)
    exe = tmp_path / "sanitize_main"
    subprocess.run(["g++", "-std=c++17", *SAN, str(EMU / "sanitize_main.cpp"), *map(str, objs), "-lm", "-o",
                    str(exe)], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout)
    print(r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all cases ok" in r.stdout
