#!/usr/bin/env python
"""Benchmark: T_eff of the FP64 PT iteration at 511^3 on one MI355X.

One step = one PT iteration (the fused r-marching kernel + the boundary pass)
over the plateau_sector workload (511^3 cells, FP64, free slip, power-law
crust), with all fields resident in HBM.  T_eff = A_eff / t_it with
A_eff = 26 * nr*nth*nph * 8 B (the reference's accounting,
SphericalStokes_diapir.jl:447).

Prints one JSON line.  The CPU baseline is the C/OpenMP oracle (a port of the
reference CPU path) timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "T_eff GB/s & % HBM peak at 511³; time-to-1e-8 residual, FP32-refined vs FP64"
PEAK_HBM_GBS = 8000.0
REF_TEFF_511 = 150.46   # reference, 1 GPU, 511^3, out_perf/out_SphericalStokes_mtp_volta.txt:21-25


def a_eff_bytes(n):
    return 26 * n * n * n * 8


def load_traffic(n):
    """HBM bytes per main-kernel launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    its `bytes_per_launch:` line; None if there is no summary for this size."""
    p = ROOT / "profiles" / f"fused_main_{n}_traffic.txt"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(setup, n, iters):
    from oracle.binding import Oracle, num_threads
    o = Oracle(setup, copy=False)
    o.step(1)                      # warm-up (page faults, first touch)
    t0 = time.perf_counter()
    o.step(iters)
    dt = (time.perf_counter() - t0) / iters
    return {"value": round(a_eff_bytes(n) / dt / 1e9, 3), "unit": "GB/s", "cores": num_threads(),
            "kind": "port",
            "sample": f"{iters} PT iterations of the same {n}^3 plateau_sector workload after 1 warm-up "
                      f"iteration, C/OpenMP oracle (reference kernel sequence), t_it = {dt:.3f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=511)
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain field copy (k_copy_field: n^3 doubles read + written) so that "
                         "rocprofv3 --pmc byte counters can be scaled")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the PT iteration runs one rank's block on one MI355X; "
              "multi-GPU halo exchange is not part of this build.", file=sys.stderr)
        sys.exit(2)

    from sphericalstokes_amd import setups
    from sphericalstokes_amd.solver import Solver

    n = args.n
    st = setups.plateau_sector(n)
    p = st.params
    s = Solver(n, n, n)
    s.set_grid(st.R, st.th, st.ph)
    for name in ("rho", "eta", "eta_ini", "eta_PL", "VR", "Vth", "Vph"):
        s.upload(name, st.fields[name])      # everything else starts at zero, as allocated
    if args.calibrate_copy:
        s.copy_field("DT_R", "eta")          # DT_R is rewritten by the last iteration
    s.pt_iterate(p, args.warmup)
    t0 = time.perf_counter()
    main_ms, iter_ms = s.pt_iterate_timed(p, args.steps)
    wall = time.perf_counter() - t0
    s.close()
    ms_per_step = wall / args.steps * 1e3
    A = a_eff_bytes(n)
    value = A / (ms_per_step * 1e-3) / 1e9
    achieved = A / (main_ms * 1e-3) / 1e9
    traffic = load_traffic(n)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / REF_TEFF_511, 3) if n == 511 else None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plateau_sector {n}^3 FP64 PT iteration (fused kernel + boundary pass), free slip, "
                               f"power-law crust n=3", "nr": n, "nth": n, "nph": n,
                   "A_eff_GB": round(A / 1e9, 6), "pct_hbm_peak": round(100 * value / PEAK_HBM_GBS, 2),
                   "iter_ms_events": round(iter_ms, 4),
                   "baseline": "reference T_eff 150.46 GB/s at 511^3 on NVIDIA V100/A100 (CUDA.jl)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": "k_fused_main", "kernel_ms": round(main_ms, 4)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(st, n, args.cpu_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
