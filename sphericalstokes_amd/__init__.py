"""sphericalstokes_amd -- MI355X (gfx950) pseudo-transient Stokes solver for a
staggered spherical-sector grid, behind the C ABI of include/sspt.h.

* :mod:`._lib`     ctypes binding of libsspt.so (no fallback when missing)
* :mod:`.solver`   :class:`Solver`, the reference kernel interface (compute_P, ...)
* :mod:`.setups`   diapir / plateau model setups as numpy state
* :mod:`.fields`   field names, ids and shapes
"""
from .fields import BC_FREE_SLIP, BC_PURE_SHEAR, BC_SHEAR, FIELDS, STATE, shape  # noqa: F401
from .setups import Params, Setup, diapir, diapir_sector, plateau_sector, random_state  # noqa: F401


def Solver(*args, **kwargs):  # noqa: N802 - keeps `from sphericalstokes_amd import Solver` lazy
    from .solver import Solver as _S
    return _S(*args, **kwargs)
