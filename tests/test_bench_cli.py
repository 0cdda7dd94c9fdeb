"""bench.py measures one GPU: asked for several it exits with status 2 and
says so, before touching a GPU (runs without one)."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=str(ROOT))
    assert r.returncode == 2
    assert "one GPU" in r.stderr
