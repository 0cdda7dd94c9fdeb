"""CPU: the command-line driver over the C ABI (sphericalstokes_amd/sspt_run,
built by build()) starts, prints its usage and rejects bad options, without
touching a GPU."""
import subprocess
from pathlib import Path

CLI = Path(__file__).resolve().parent.parent / "sphericalstokes_amd" / "sspt_run"


def _run(*args):
    assert CLI.exists(), "sspt_run is not built: run __graft_entry__.build()"
    return subprocess.run([str(CLI), *args], capture_output=True, text=True, timeout=60)


def test_help():
    r = _run("--help")
    assert r.returncode == 0 and "usage: sspt_run" in r.stdout


def test_bad_options_exit_2():
    assert _run("--bc", "nonsense").returncode == 2
    assert _run("--setup", "nonsense").returncode == 2
    assert _run("--frobnicate", "1").returncode == 2
    assert _run("--n", "2").returncode == 2
