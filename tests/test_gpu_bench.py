"""GPU: bench.py end to end on a small grid, in a child process: it prints one
JSON line with the fields the measurement contract needs (metric, value,
roofline, cpu_baseline, the mixed-precision, adjoint and twin legs, and the
streaming rates of the torch probe, which runs in a grandchild process
because torch's HIP runtime cannot share the GPU with libsspt.so's)."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

pytestmark = pytest.mark.gpu


def test_bench_json_contract():
    cmd = [sys.executable, str(ROOT / "bench.py"), "--n", "63", "--steps", "3", "--warmup", "1", "--cpu-iters", "1",
           "--ttt-n", "0", "--adjoint-n", "0", "--fp32-sweeps", "2", "--adjoint-steps", "1", "--cart-n", "31", "--cyl-n", "31"]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=str(ROOT))
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    assert len(lines) == 1, r.stdout
    out = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in out, k
    assert out["n_gpus"] == 1 and out["steps"] == 3 and out["dtype"] == "f64"
    assert out["value"] > 0 and out["ms_per_step"] > 0
    rf = out["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert 0 < rf["frac"] < 1 and rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], abs=1e-4)
    cb = out["cpu_baseline"]
    assert cb["kind"] == "port" and cb["value"] > 0 and cb["cores"] >= 1 and cb["sample"]
    assert out["mixed_precision"]["fp32_sweep_ms"] > 0
    assert out["adjoint"]["iter_ms"] > 0
    for twin in ("cartesian_twin", "cylindrical_twin"):
        ct = out[twin]
        assert ct["iter_ms"] > 0 and ct["T_eff_GBps"] > 0 and 0 < ct["compute_P"]["frac_of_peak"] < 1, twin
    hb = out["hbm_measured"]
    assert "error" not in hb, hb
    assert all(hb[k] > 0 for k in ("copy_GBps", "triad_GBps", "read_GBps", "write_GBps"))

