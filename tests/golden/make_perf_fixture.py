"""Writes reference_perf_logs.json from the reference's single-GPU perf logs.

Source: out_perf/out_SphericalStokes_mtp_volta.txt and
out_perf/out_SphericalStokes_mtp_daint.txt (SphericalStokes_mxpu_mtp.jl
runs, no halo exchange).  Each line holds nx ny nz niter wtime A_eff[GB]
t_it[s] T_eff[GB/s]; the fixture keeps them as numbers with their file and
line, so the tests can check this build's A_eff / t_it / T_eff accounting
(SphericalStokes_diapir.jl:445-453) against the reference's own outputs.
Run in a container where the reference is mounted:
    python tests/golden/make_perf_fixture.py /root/reference
"""
import json
import sys
from pathlib import Path


def main(ref):
    rows = []
    for name, gpu in (("out_SphericalStokes_mtp_volta.txt", "volta"), ("out_SphericalStokes_mtp_daint.txt", "daint")):
        path = Path(ref) / "out_perf" / name
        for ln, line in enumerate(path.read_text().splitlines(), 1):
            f = line.split()
            if len(f) != 8:
                continue
            rows.append({"source": f"out_perf/{name}:{ln}", "gpu": gpu, "nx": int(f[0]), "ny": int(f[1]),
                         "nz": int(f[2]), "niter": int(f[3]), "wtime": float(f[4]), "A_eff_GB": float(f[5]),
                         "t_it": float(f[6]), "T_eff": float(f[7])})
    out = Path(__file__).resolve().parent / "reference_perf_logs.json"
    out.write_text(json.dumps(rows, indent=1) + "\n")
    print(f"{len(rows)} rows -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
