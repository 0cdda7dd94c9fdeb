#!/usr/bin/env python
"""Benchmark: T_eff of the FP64 PT iteration at 511^3 on one MI355X.

One step = one PT iteration (the fused r-marching kernel + the boundary pass)
over the plateau_sector workload (511^3 cells, FP64, free slip, power-law
crust), with all fields resident in HBM.  T_eff = A_eff / t_it with
A_eff = 26 * nr*nth*nph * 8 B (the reference's accounting,
SphericalStokes_diapir.jl:447).

Prints one JSON line.  The CPU baseline is the C/OpenMP oracle (a port of the
reference CPU path) timed on a bounded sample of the same workload.

"mixed_precision" holds the second half of the metric: the cost of one FP32
correction sweep on the same 511^3 workload against one FP64 iteration, and
the time to reach an FP64 residual of 1e-8 with the plain FP64 solve and
with the FP32-inner / FP64 defect-correction solve on the mixed-precision
configuration's own field (the power-law diapir with the smoothed log-normal
matrix viscosity, etaSM spread over 5 cells) at --ttt-n^3, default 127^3:
about five minutes of GPU time (887000 FP64 iterations), the largest size at
which the FP64 solve reaches 1e-8 within minutes; --ttt-setup diapir selects
the linear diapir_sector instead.
"adjoint" holds the cost of one adjoint iteration on the same workload
against one FP64 iteration, and the wall times of a forward solve and of
the adjoint solve that gives the gradient (power-law diapir at --adjoint-n^3).
"cartesian_twin" / "cylindrical_twin" time the Cartesian and cylindrical twins
of the diapir solver at --cart-n^3 / --cyl-n^3.
"""
from __future__ import annotations

import argparse
import json
import math
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


def cpu_baseline(setup, n, iters, warmup=10):
    """The oracle (the reference's kernel sequence in C/OpenMP) timed with the
    reference's accounting (SphericalStokes_diapir.jl:411,445-450): the clock
    starts after `warmup` iterations, t_it = wall / (ittot - warmup) over the
    remaining `iters`, T_eff = A_eff / t_it."""
    from oracle.binding import Oracle, num_threads
    o = Oracle(setup, copy=False)
    o.step(warmup)
    t0 = time.perf_counter()
    o.step(iters)
    dt = (time.perf_counter() - t0) / iters
    return {"value": round(a_eff_bytes(n) / dt / 1e9, 3), "unit": "GB/s", "cores": num_threads(),
            "kind": "port",
            "sample": f"{warmup} warm-up + {iters} timed PT iterations of the same {n}^3 plateau_sector "
                      f"workload (the reference's accounting), C/OpenMP oracle (reference kernel sequence), "
                      f"t_it = {dt:.3f} s"}


def fp32_sweep_ms(s, p, sweeps):
    """Average wall time of one FP32 correction sweep on the same workload:
    one outer step of the mixed solve with `sweeps` sweeps and no checks."""
    r = s.pt_solve_mixed(p, max_outer=1, max_inner=sweeps, nout_inner=sweeps + 1, tol_inner=0.0, tol=0.0)
    return r["inner_seconds"] / sweeps * 1e3


def adjoint_iteration_ms(s, p, steps, r, radius):
    """Average wall time of one adjoint iteration (T^T applied to the adjoint
    fields, ss_adjoint_step) at the state held by s."""
    s.adjoint_prepare(p)
    s.adjoint_objective(r, radius)
    s.adjoint_zero()
    s.adjoint_step(p)                       # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        s.adjoint_step(p)
    return (time.perf_counter() - t0) / steps * 1e3


def adjoint_fp32_sweep_ms(s, p, sweeps, r, radius):
    """Average wall time of one FP32 correction sweep of the mixed-precision
    adjoint solve (BC fold, boundary power-law sums and the FP32 r-marching
    kernel with its right-hand side) at the state held by s: one outer step
    of ss_adjoint_solve_mixed with tol 0 and no inner checks, from the
    library's own timing of its sweeps."""
    m = s.adjoint_solve_mixed(p, r, radius, max_outer=1, max_inner=sweeps + 1, nout_inner=sweeps + 2,
                              tol_inner=0.0, tol=0.0, gradients=False)
    return m["inner_seconds"] / m["inner_total"] * 1e3


def adjoint_solve_ratio(n, eps, mixed=True):
    """Forward solve (reference exit test, relative change of V < eps every
    500 iterations) and adjoint solve (relative change of the adjoint < eps)
    of the power-law diapir_sector at n^3, free slip: wall times and ratio.
    With mixed, also the adjoint solve with FP32 sweeps inside the FP64
    defect correction (ss_adjoint_solve_mixed) from the same state.  The
    adjoint times are the library's own (prepare, objective, iterations,
    gradients; the downloads of the gradient arrays are left out)."""
    import numpy as np
    from sphericalstokes_amd import setups
    from sphericalstokes_amd.fields import BC_FREE_SLIP, FIELDS
    from sphericalstokes_amd.solver import Solver

    st = setups.diapir_sector(n, n_exp=3.0, bc_mode=BC_FREE_SLIP, vr=10.0)
    p = st.params
    s = Solver(n, n, n)
    s.set_grid(st.R, st.th, st.ph)
    for name in FIELDS:
        s.upload(name, st.fields[name])
    t0 = time.perf_counter()
    f = s.pt_solve(p, 4_000_000, 500, eps)
    tf = time.perf_counter() - t0
    print(f"  adjoint {n}^3: forward {f['ittot']} iterations in {tf:.2f} s", file=sys.stderr, flush=True)
    a = s.adjoint_solve(p, 1.5, 0.1 ** 2, maxiter=4_000_000, ncheck=500, tol=eps)
    ta = a["seconds"]
    print(f"  adjoint {n}^3: FP64 adjoint {a['iters']} iterations in {ta:.2f} s", file=sys.stderr, flush=True)
    out = {"workload": f"diapir_sector {n}^3, n_exp=3, free slip; J = mean VR in the inclusion", "eps": eps,
           "forward_s": round(tf, 3), "forward_iters": f["ittot"], "adjoint_s": round(ta, 3),
           "adjoint_iters": a["iters"], "adjoint_converged": a["converged"],
           "ratio_adjoint_over_forward": round(ta / tf, 4)}
    if mixed:
        m = s.adjoint_solve_mixed(p, 1.5, 0.1 ** 2, nout_inner=100, tol=eps)
        print(f"  adjoint {n}^3: mixed adjoint {m['outer']} outer steps, {m['inner_total']} FP32 sweeps in "
              f"{m['seconds']:.2f} s", file=sys.stderr, flush=True)
        diff = max(float(np.abs(m[k] - a[k]).max() / np.abs(a[k]).max()) for k in ("grad_rho", "grad_log10_eta_ini"))
        out.update({"mixed_adjoint_s": round(m["seconds"], 3), "mixed_outer": m["outer"],
                    "mixed_fp32_sweeps": m["inner_total"], "mixed_converged": m["converged"],
                    "mixed_grad_rel_diff_vs_fp64": float(f"{diff:.3e}"),
                    "ratio_mixed_adjoint_over_forward": round(m["seconds"] / tf, 4)})
    s.close()
    return out


def time_to_tolerance(n, tol, tol_inner, setup="diapir"):
    """Time-to-tolerance of the plain FP64 solve and of the FP32-inner /
    FP64 defect-correction solve at n^3, free slip (the discrete equations
    are consistent, so the residual can reach tol): the linear diapir_sector
    (setup "diapir") or the power-law diapir with the smoothed log-normal
    matrix viscosity (setup "lognormal", the mixed-precision configuration;
    the outer steps are then also Picard steps of the viscosity).  Both start
    from the same state; the FP64 solve checks its residual every 500
    iterations.  After every outer step the error of V (max over the
    components) and of the interior zero-mean P against the FP64 solution
    is measured (downloads outside the timed calls)."""
    import numpy as np
    from sphericalstokes_amd import setups
    from sphericalstokes_amd.fields import BC_FREE_SLIP, FIELDS
    from sphericalstokes_amd.solver import Solver

    if setup == "lognormal":
        st = setups.diapir_lognormal(n)
        what = (f"diapir_lognormal {n}^3 (n_exp=3, log10 eta_matrix ~ N(0, 0.5^2) seed 0, smoothed 8 cells, "
                f"contrast <= 1e4, etaSM spread {st.params.etasm_spread})")
    else:
        st = setups.diapir_sector(n, n_exp=1.0, bc_mode=BC_FREE_SLIP)
        what = f"diapir_sector {n}^3, n_exp=1"
    p = st.params

    def load():
        s = Solver(n, n, n)
        s.set_grid(st.R, st.th, st.ph)
        for name in FIELDS:
            s.upload(name, st.fields[name])
        return s

    # Both solves run in chunks (FP64: 20000 iterations, mixed: one outer
    # step per call) so that progress is reported on stderr; a chunk ends at
    # a residual check in either case, so the iterates are those of one call.
    s = load()
    t0 = time.perf_counter()
    iters, conv64 = 0, False
    while iters < 4_000_000 and not conv64:
        r64 = s.pt_solve_residual(p, 20_000, 500, tol)
        iters += r64["iter_end"] if r64["converged"] else r64["iters"]
        conv64 = r64["converged"]
        print(f"  fp64 {n}^3: {iters} iterations, residual {r64['r']:.3e}", file=sys.stderr, flush=True)
    t64 = time.perf_counter() - t0
    ref = {k: s.download(k) for k in ("VR", "Vth", "Vph", "P")}
    s.close()

    def errors(s):
        ev = max(float(np.linalg.norm(s.download(k) - ref[k]) / np.linalg.norm(ref[k])) for k in ("VR", "Vth", "Vph"))
        dp = (s.download("P") - ref["P"])[1:-1, 1:-1, 1:-1]
        pc = ref["P"][1:-1, 1:-1, 1:-1]
        ep = float(np.linalg.norm(dp - dp.mean()) / np.linalg.norm(pc - pc.mean()))
        return ev, ep

    s = load()
    tm = 0.0
    outer, inner, hist, convm, ehist = 0, 0, [], False, []
    tin_prev, r_prev = tol_inner, None
    while outer < 60 and not convm:
        # one outer step per call (progress on stderr); the inner tolerance
        # follows ss_pt_solve_mixed's own rule across its outer steps
        tin = tol_inner
        if len(hist) >= 2 and hist[-1] < hist[-2]:
            q, want = hist[-1] / hist[-2], 0.5 * (tol / hist[-1])
            if want > q:
                tin = max(tol_inner, min(1e-2, tin_prev * want / q))
        tin_prev = tin
        # an outer step stops after 100000 sweeps at most (a partial correction is
        # still a valid defect-correction step; this bounds the time per call)
        t0 = time.perf_counter()
        rm = s.pt_solve_mixed(p, max_outer=1, max_inner=100_000, nout_inner=200, tol_inner=tin, tol=tol)
        tm += time.perf_counter() - t0
        ehist.append(errors(s))
        if not hist:
            hist.append(rm["r_hist"][0])
        convm = rm["converged"]
        outer += rm["outer"]
        inner += rm["inner_total"]
        hist.append(rm["r"])
        print(f"  mixed {n}^3: outer {outer}, {inner} FP32 sweeps, residual {rm['r']:.3e}, "
              f"error vs FP64: V {ehist[-1][0]:.2e} P {ehist[-1][1]:.2e}", file=sys.stderr, flush=True)
    s.close()
    return {"workload": f"{what}, free slip, FP64 residual <= {tol:g}",
            "fp64_s": round(t64, 3), "fp64_iters": iters, "fp64_converged": conv64,
            "mixed_s": round(tm, 3), "mixed_outer": outer, "mixed_inner_sweeps": inner,
            "mixed_converged": convm, "mixed_r_hist": [float("%.3e" % x) for x in hist],
            "mixed_V_err_hist": [float("%.3e" % e[0]) for e in ehist],
            "mixed_P_err_hist": [float("%.3e" % e[1]) for e in ehist],
            "ratio_mixed_over_fp64": round(tm / t64, 4), "V_rel_err_mixed_vs_fp64": float("%.3e" % ehist[-1][0]),
            "P_rel_err_mixed_vs_fp64": float("%.3e" % ehist[-1][1])}


def cartesian_twin(n, steps, warmup=3):
    """The Cartesian twin (CartesianStokes_diapir.jl, the script's own
    parameters: n_exp = 5, relaxation 1e-3, simple shear) at n^3: one
    iteration is the script's kernel sequence (:351-379, about 20 launches).
    T_eff with the same A_eff accounting as the spherical path; the roofline
    of its dominant kernel compute_P!, whose algorithmic bytes per cell are
    19 x 8 (reads VX VY VZ eta etaSM, reads and writes P and the six
    stresses)."""
    import numpy as np
    from sphericalstokes_amd.cartesian import CART_FIELDS, CartSolver, cart_diapir

    st = cart_diapir(n)
    p = st.params
    with CartSolver(n, n, n) as s:
        s.set_y(st.y)
        for name in CART_FIELDS:          # the context starts at zero
            a = st.fields.pop(name)
            if np.any(a):
                s.upload(name, a)
            del a
        s.pt_iterate_timed(p, warmup)
        it_ms, k_ms = s.pt_iterate_timed(p, steps)
    A = a_eff_bytes(n)
    kb = 19 * 8 * n ** 3
    return {"workload": f"cartesian diapir {n}^3 FP64 (CartesianStokes_diapir.jl setup: n_exp=5, relax 1e-3, "
                        f"simple shear), per-kernel launches", "iter_ms": round(it_ms, 4),
            "T_eff_GBps": round(A / (it_ms * 1e-3) / 1e9, 2),
            "compute_P": {"kernel_ms": round(k_ms, 4), "bytes": kb,
                          "achieved_GBps": round(kb / (k_ms * 1e-3) / 1e9, 2),
                          "frac_of_peak": round(kb / (k_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}}


def cylindrical_twin(n, steps, warmup=3):
    """The cylindrical twin (CylindricalStokes_diapir.jl, the script's setup:
    thin shell r = 1000, linear phase n_exp = 1) at n^3: one iteration is the
    script's kernel sequence (:331-366, about 25 launches).  T_eff by the
    spherical path's A_eff accounting; the roofline of compute_P!, whose
    algorithmic bytes per cell are 27 x 8 (reads VR Vth VZ KBDT GSDT eta,
    reads and writes P and the six stresses, writes DIVV and the six strain
    rates)."""
    import numpy as np
    from sphericalstokes_amd.cylindrical import CYL_FIELDS, CylSolver, cyl_diapir

    st = cyl_diapir(n)
    p = st.params
    with CylSolver(n, n, n) as s:
        s.set_grid(st.R, st.th)
        for name in CYL_FIELDS:           # the context starts at zero
            a = st.fields.pop(name)
            if np.any(a):
                s.upload(name, a)
            del a
        s.pt_iterate_timed(p, warmup)
        it_ms, k_ms = s.pt_iterate_timed(p, steps)
    A = a_eff_bytes(n)
    kb = 27 * 8 * n ** 3
    return {"workload": f"cylindrical diapir {n}^3 FP64 (CylindricalStokes_diapir.jl setup, linear phase), "
                        f"per-kernel launches", "iter_ms": round(it_ms, 4),
            "T_eff_GBps": round(A / (it_ms * 1e-3) / 1e9, 2),
            "compute_P": {"kernel_ms": round(k_ms, 4), "bytes": kb,
                          "achieved_GBps": round(kb / (k_ms * 1e-3) / 1e9, 2),
                          "frac_of_peak": round(kb / (k_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}}


def hbm_measured():
    """Streaming rates of plain torch kernels, measured in a child process.

    torch brings its own HIP runtime; once libsspt.so (linked against the
    system one) holds the GPU in this process, torch's runtime finds no
    device.  The child process has only torch's runtime.  Returns the rates,
    or {"error": ...} if the child fails."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--hbm-probe-child"],
                           capture_output=True, text=True, timeout=300)
    except subprocess.TimeoutExpired:
        return {"error": "streaming-rate probe timed out"}
    if r.returncode != 0:
        return {"error": f"streaming-rate probe exited with {r.returncode}: {r.stderr.strip()[-300:]}"}
    return json.loads(r.stdout.strip().splitlines()[-1])


def hbm_probe():
    """Streaming rates this GPU reaches on plain torch kernels (1 GiB FP64
    arrays, 20 repetitions timed with events): copy (1 read + 1 write),
    triad c = a + 0.5 b (2 reads + 1 write), read-only sum, fill.  They
    place the stencil kernel's rate between the vendor peak and what simple
    streaming reaches."""
    import torch
    dev = torch.device("cuda:0")
    n = (1 << 30) // 8
    a = torch.rand(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    c = torch.rand(n, dtype=torch.float64, device=dev)

    def rate(f, nbytes, reps=20):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return round(nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9, 1)

    B = n * 8
    out = {"copy_GBps": rate(lambda: b.copy_(a), 2 * B),
           "triad_GBps": rate(lambda: torch.add(a, c, alpha=0.5, out=b), 3 * B),
           "read_GBps": rate(lambda: a.sum(), B),
           "write_GBps": rate(lambda: b.fill_(1.0), B)}
    del a, b, c
    torch.cuda.empty_cache()
    return out


def mixed_large(n, outer, sweeps):
    """The mixed-precision configuration at full size: the power-law diapir
    with the log-normal matrix viscosity at n^3 (767^3: 31 FP64 and 25 FP32
    boxes, about 157 GB of HBM).  Times FP64 iterations and runs `outer`
    outer steps of `sweeps` FP32 sweeps each, reporting the FP64 residual
    after every step (a solve to 1e-8 at this size takes hours).  The setup
    spreads etaSM over 5 cells (etasm_spread): with the reference's etaSM the
    sharp jump at the inclusion's edge diverges beyond 63^3.  A non-finite
    residual is reported as null with the solver's message."""
    from sphericalstokes_amd import setups
    from sphericalstokes_amd._lib import SSError
    from sphericalstokes_amd.fields import FIELDS
    from sphericalstokes_amd.solver import Solver

    t0 = time.perf_counter()
    st = setups.diapir_lognormal(n)
    print(f"  mixed {n}^3: setup built in {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    p = st.params
    s = Solver(n, n, n)
    s.set_grid(st.R, st.th, st.ph)
    for name in FIELDS:
        s.upload(name, st.fields[name])
    del st.fields
    print(f"  mixed {n}^3: uploaded", file=sys.stderr, flush=True)
    s.pt_iterate(p, 3)
    main_ms, it_ms = s.pt_iterate_timed(p, 10)
    hist, tin, done, note = [], 0.0, 0, None
    for k in range(outer):
        try:
            r = s.pt_solve_mixed(p, max_outer=1, max_inner=sweeps, nout_inner=sweeps + 1, tol_inner=0.0, tol=0.0)
        except SSError as e:
            note = str(e)
            break
        if not hist:
            hist.append(r["r_hist"][0])
        hist.append(r["r"])
        tin += r["inner_seconds"]
        done += 1
        print(f"  mixed {n}^3: outer {k + 1}, residual {r['r']:.3e}", file=sys.stderr, flush=True)
    s.close()
    sweep_ms = tin / (done * sweeps) * 1e3 if done else None
    return {"workload": f"diapir_lognormal {n}^3 (n_exp=3, log10 eta_matrix ~ N(0, 0.5^2) seed 0, smoothed 8 cells, "
                        f"contrast <= 1e4, etaSM spread {st.params.etasm_spread}), free slip",
            "fp64_iter_ms": round(it_ms, 4),
            "fp64_kernel_ms": round(main_ms, 4), "fp32_sweep_ms": round(sweep_ms, 4) if sweep_ms else None,
            "sweep_over_iter": round(sweep_ms / it_ms, 4) if sweep_ms else None,
            "outer_steps": done, "sweeps_per_outer": sweeps,
            "r_hist": [float("%.3e" % x) if math.isfinite(x) else None for x in hist], "error": note}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=511)
    ap.add_argument("--cpu-iters", type=int, default=20,
                    help="timed CPU-baseline iterations (after --cpu-warmup untimed ones)")
    ap.add_argument("--cpu-warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ttt-n", type=int, default=127,
                    help="size of the time-to-tolerance comparison (FP64 vs FP32-refined); 0 skips it")
    ap.add_argument("--ttt-tol", type=float, default=1e-8)
    ap.add_argument("--no-hbm-probe", action="store_true",
                    help="skip the plain streaming-rate measurement (copy/triad/read/write)")
    ap.add_argument("--mixed-large-n", type=int, default=0,
                    help="also run outer steps of the mixed solve on the log-normal power-law diapir at this "
                         "size (767 for the full configuration); 0 skips it")
    ap.add_argument("--ttt-setup", choices=["diapir", "lognormal"], default="lognormal",
                    help="time-to-tolerance workload: the linear diapir_sector, or the power-law diapir "
                         "with the smoothed log-normal matrix viscosity")
    ap.add_argument("--ttt-tol-inner", type=float, default=1e-5,
                    help="inner tolerance of the FP32 sweeps (max relative change per sweep)")
    ap.add_argument("--fp32-sweeps", type=int, default=20,
                    help="FP32 correction sweeps timed on the main workload; 0 skips them")
    ap.add_argument("--adjoint-steps", type=int, default=10,
                    help="adjoint iterations timed on the main workload; 0 skips them")
    ap.add_argument("--adjoint-n", type=int, default=63,
                    help="size of the forward-vs-adjoint solve comparison; 0 skips it")
    ap.add_argument("--adjoint-eps", type=float, default=1e-10,
                    help="exit tolerance of the forward and adjoint solves of that comparison")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain field copy (k_copy_field: n^3 doubles read + written) so that "
                         "rocprofv3 --pmc byte counters can be scaled")
    ap.add_argument("--cart-n", type=int, default=511,
                    help="size of the Cartesian twin's timing leg; 0 skips it")
    ap.add_argument("--cyl-n", type=int, default=511,
                    help="size of the cylindrical twin's timing leg; 0 skips it")
    ap.add_argument("--hbm-probe-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.hbm_probe_child:        # the child process of hbm_measured()
        print(json.dumps(hbm_probe()), flush=True)
        return
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
    sweep_ms = fp32_sweep_ms(s, p, args.fp32_sweeps) if args.fp32_sweeps > 0 else None
    # adjoint iteration on the same state (J: mean VR in a sphere of radius 0.01 in the crust)
    adj_ms = adjoint_iteration_ms(s, p, args.adjoint_steps, 0.985, 0.01 ** 2) if args.adjoint_steps > 0 else None
    adj32_ms = adjoint_fp32_sweep_ms(s, p, args.adjoint_steps, 0.985, 0.01 ** 2) if args.adjoint_steps > 0 else None
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
        "mixed_precision": {
            "fp32_sweep_ms": round(sweep_ms, 4) if sweep_ms else None,
            "fp32_sweep_over_fp64_iter": round(sweep_ms / ms_per_step, 4) if sweep_ms else None,
            "time_to_tol": None},
    }
    out["adjoint"] = {"iter_ms": round(adj_ms, 4) if adj_ms else None,
                      "iter_over_fp64_iter": round(adj_ms / ms_per_step, 4) if adj_ms else None,
                      "fp32_sweep_ms": round(adj32_ms, 4) if adj32_ms else None,
                      "fp32_sweep_over_fp64_iter": round(adj32_ms / ms_per_step, 4) if adj32_ms else None,
                      "solve": adjoint_solve_ratio(args.adjoint_n, args.adjoint_eps) if args.adjoint_n > 0 else None}
    if args.ttt_n > 0:
        out["mixed_precision"]["time_to_tol"] = time_to_tolerance(args.ttt_n, args.ttt_tol, args.ttt_tol_inner,
                                                                  args.ttt_setup)
    if args.cart_n > 0:
        out["cartesian_twin"] = cartesian_twin(args.cart_n, 10)
    if args.cyl_n > 0:
        out["cylindrical_twin"] = cylindrical_twin(args.cyl_n, 10)
    if not args.no_hbm_probe:
        out["hbm_measured"] = hbm_measured()
    if args.mixed_large_n > 0:
        out["mixed_precision"]["large"] = mixed_large(args.mixed_large_n, 3, 1000)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(st, n, args.cpu_iters, args.cpu_warmup)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
