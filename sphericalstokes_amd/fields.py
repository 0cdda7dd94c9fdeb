"""Field names, ids and shapes of the PT state (shared by the host mirror,
the oracle binding and the tests).

Shapes follow scripts_diapir/SphericalStokes_diapir.jl:381-399 (reference
sizes, Julia column-major layout with r fastest).
"""
from __future__ import annotations

FIELDS = [
    "P", "DT_R", "tau_RR", "tau_thth", "tau_phph",
    "tau_Rth", "tau_Rph", "tau_thph",
    "VR", "Vth", "Vph", "VR_err", "Vth_err", "Vph_err",
    "eta", "eta_ini", "etaSM", "tII", "rho", "eta_PL",
]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}

# Julia-script spellings accepted as aliases
ALIASES = {
    "τ_RR": "tau_RR", "τ_θθ": "tau_thth", "τ_φφ": "tau_phph",
    "τ_Rθ": "tau_Rth", "τ_Rφ": "tau_Rph", "τ_θφ": "tau_thph",
    "Vθ": "Vth", "Vφ": "Vph", "Vθ_err": "Vth_err", "Vφ_err": "Vph_err",
    "η": "eta", "η_ini": "eta_ini", "ηSM": "etaSM", "τII": "tII", "ρ": "rho", "η_PL": "eta_PL",
}

# the PT state rewritten every iteration; the rest are inputs or diagnostics
STATE = ["P", "tau_RR", "tau_thth", "tau_phph", "tau_Rth", "tau_Rph", "tau_thph",
         "VR", "Vth", "Vph", "eta"]

BC_SHEAR, BC_FREE_SLIP, BC_PURE_SHEAR = 0, 1, 2
LAYOUT_JULIA, LAYOUT_C = 0, 1


def canon(name: str) -> str:
    name = ALIASES.get(name, name)
    if name not in FIELD_ID:
        raise KeyError(f"unknown field {name!r}")
    return name


def shape(name: str, nr: int, nth: int, nph: int) -> tuple[int, int, int]:
    name = canon(name)
    if name in ("VR", "VR_err"):
        return (nr + 1, nth, nph)
    if name in ("Vth", "Vth_err"):
        return (nr, nth + 1, nph)
    if name in ("Vph", "Vph_err"):
        return (nr, nth, nph + 1)
    if name == "tau_Rth":
        return (nr - 1, nth - 1, nph - 2)
    if name == "tau_Rph":
        return (nr - 1, nth - 2, nph - 1)
    if name == "tau_thph":
        return (nr - 2, nth - 1, nph - 1)
    return (nr, nth, nph)
