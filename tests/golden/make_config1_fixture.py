"""Writes config1_63_oracle.json: the CPU oracle's converged run of the
linear rising-sphere diapir at 63^3 (diapir_sector: r in [1,2], th in
[pi/3, 2pi/3], phi in [-pi/6, pi/6], sphere of radius 0.1 at (1.5, pi/2, 0),
drho = -1, g = 1, inclusion/matrix viscosity 1e-3, n_exp = 1, free slip),
with the reference loop and exit test (SphericalStokes_diapir.jl:409-443:
err every nout = 100 iterations, break when err < 1e-8 and iter > 20, at
most 1e5 iterations).

The fixture holds the exit-test record (ittot, iter_end, err and its three
components), the ascent velocity J (volume-weighted mean VR over the r-faces
inside the inclusion, orc_adj_objective) and, per field, the SHA-256 of its
Fortran-ordered bytes and its L2 norm, so that a GPU run of the same loop
can be compared bit for bit without shipping the arrays.

    python tests/golden/make_config1_fixture.py     (about 20 minutes on 8 cores)
"""
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.binding import Oracle  # noqa: E402
from sphericalstokes_amd import setups  # noqa: E402
from sphericalstokes_amd.fields import FIELDS  # noqa: E402

N, MAXITER, NOUT, EPS = 63, 100_000, 100, 1e-8
R_INC, RADIUS = 1.5, 0.1 ** 2


def field_digest(a):
    b = np.asfortranarray(a, dtype=np.float64)
    return {"sha256": hashlib.sha256(b.tobytes(order="F")).hexdigest(), "l2": float(np.linalg.norm(b.ravel()))}


def main():
    st = setups.diapir_sector(N, n_exp=1.0)
    o = Oracle(st)
    t0 = time.time()
    res = o.solve(MAXITER, NOUT, EPS)
    secs = time.time() - t0
    J, _ = o.objective(R_INC, RADIUS)
    out = {
        "workload": "diapir_sector 63^3, n_exp=1, free slip, reference exit test nout=100 eps=1e-8 maxiter=1e5",
        "maxiter": MAXITER, "nout": NOUT, "eps": EPS, "r": R_INC, "radius": RADIUS,
        "ittot": res["ittot"], "iter_end": res["iter_end"], "err": res["err"],
        "J": J, "oracle_seconds": round(secs, 1),
        "fields": {n: field_digest(o.f[n]) for n in FIELDS},
    }
    dst = Path(__file__).resolve().parent / "config1_63_oracle.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps({k: out[k] for k in ("ittot", "iter_end", "err", "J", "oracle_seconds")}))


if __name__ == "__main__":
    main()
