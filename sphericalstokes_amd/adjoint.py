"""Differentiable ascent velocity: the forward PT solve and its adjoint as a
torch.autograd.Function (the counterpart of the ChainRulesCore rrule a Julia
caller would write over the same C ABI).

J(rho, log10 eta_ini) = volume-weighted mean VR over the r-faces inside the
inclusion of the converged PT solve.  forward runs the fused PT loop with the
reference exit test (ss_pt_solve) and keeps the converged state on the GPU;
backward runs the adjoint solve on that state (ss_adjoint_solve): no record
of the forward iterations is kept.

The adjoint is exact only at a fixed point of the PT iteration, so forward
raises NotConvergedError when the PT loop ends without meeting its exit
test, and backward raises it when the adjoint iteration does not reach its
tolerance (unless SolveOptions.strict is False, in which case the statuses
are only reported in SolveOptions.info).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .fields import FIELDS
from .setups import Setup
from .solver import Solver


MIXED_ADJOINT_CELLS = 127 ** 3   # SolveOptions.adj_mixed = None: the mixed backward from this size up


@dataclass
class SolveOptions:
    maxiter: int = 4_000_000
    nout: int = 500
    eps: float = 1e-12          # reference exit test: relative change of V per check
    r: float = 1.5               # inclusion centre radius (diapir_sector)
    radius: float = 0.1 ** 2     # squared inclusion radius, as initialize_inclusion! takes it
    adj_maxiter: int = 4_000_000
    adj_ncheck: int = 200
    adj_tol: float = 1e-13
    # FP32 sweeps inside an FP64 defect correction (ss_adjoint_solve_mixed);
    # None: from 127^3 cells up, where it is the faster one (measured 1.27 x
    # a forward solve against 2.09 x at 127^3; 1.86 x against 1.48 x at 63^3)
    adj_mixed: bool | None = None
    adj_max_outer: int = 20       # its outer steps
    strict: bool = True          # raise NotConvergedError on an unconverged forward or adjoint solve
    info: dict = field(default_factory=dict)


class NotConvergedError(RuntimeError):
    """A solve that the gradient relies on did not converge."""


def _solver_with(setup: Setup, rho: np.ndarray, log10_ini: np.ndarray) -> Solver:
    s = Solver(setup.nr, setup.nth, setup.nph)
    s.set_grid(setup.R, setup.th, setup.ph)
    ini = np.asfortranarray(10.0 ** log10_ini)
    f = dict(setup.fields)
    f.update({"rho": np.asfortranarray(rho), "eta_ini": ini, "eta": ini, "eta_PL": ini})
    for n in FIELDS:
        s.upload(n, f[n])
    return s


class AscentVelocity(torch.autograd.Function):
    """J = AscentVelocity.apply(rho, log10_eta_ini, setup, options); rho and
    log10_eta_ini are float64 CPU tensors of the centre shape (nr, nth, nph).

    The solve runs in libsspt.so, which uses the system HIP runtime; torch
    carries its own, and in one process only the runtime that opens the GPU
    first sees it, so the tensors here stay on the host (the fields are
    uploaded through the C ABI)."""

    @staticmethod
    def forward(ctx, rho, log10_eta_ini, setup: Setup, opts: SolveOptions):
        s = _solver_with(setup, rho.detach().cpu().numpy().astype(np.float64),
                         log10_eta_ini.detach().cpu().numpy().astype(np.float64))
        res = s.pt_solve(setup.params, opts.maxiter, opts.nout, opts.eps, raise_on=())
        opts.info["forward"] = res
        if res["status"] != "SS_OK" or res["iter_end"] == 0:
            if opts.strict or res["status"] == "SS_ERR_NONFINITE":
                s.close()
                raise NotConvergedError(f"forward PT solve: {res['status']} after {res['ittot']} iterations "
                                        f"(err {res['err']:.3e}, eps {opts.eps:.1e})")
        s.adjoint_prepare(setup.params)
        J = s.adjoint_objective(opts.r, opts.radius)
        if not any(ctx.needs_input_grad[:2]):      # no backward will come: release the device state now
            s.close()
            s = None
        ctx.solver, ctx.setup, ctx.opts = s, setup, opts
        ctx.device, ctx.dtype = rho.device, rho.dtype
        return torch.tensor(J, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, gJ):
        s, setup, o = ctx.solver, ctx.setup, ctx.opts
        try:
            mixed = o.adj_mixed if o.adj_mixed is not None else setup.nr * setup.nth * setup.nph >= MIXED_ADJOINT_CELLS
            if mixed:
                r = s.adjoint_solve_mixed(setup.params, o.r, o.radius, max_outer=o.adj_max_outer,
                                          max_inner=o.adj_maxiter, nout_inner=o.adj_ncheck, tol=o.adj_tol)
                r["iters"] = r["inner_total"]
            else:
                r = s.adjoint_solve(setup.params, o.r, o.radius, o.adj_maxiter, o.adj_ncheck, o.adj_tol)
            r["mixed"] = mixed
        finally:
            s.close()
            ctx.solver = None
        o.info["adjoint"] = {k: v for k, v in r.items() if not k.startswith("grad")}
        if o.strict and not r["converged"]:
            raise NotConvergedError(f"adjoint solve: not converged after {r['iters']} iterations "
                                    f"(err {r['err']:.3e}, tol {o.adj_tol:.1e})")
        g = float(gJ)
        to = lambda a: torch.as_tensor(np.ascontiguousarray(a) * g, dtype=ctx.dtype, device=ctx.device)
        return to(r["grad_rho"]), to(r["grad_log10_eta_ini"]), None, None


def ascent_velocity(rho, log10_eta_ini, setup: Setup, opts: SolveOptions | None = None):
    return AscentVelocity.apply(rho, log10_eta_ini, setup, opts or SolveOptions())

