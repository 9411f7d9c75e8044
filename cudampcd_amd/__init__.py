"""MPCD time step for AMD Instinct MI355X: a "hip" backend for the pympcd API.

    import cudampcd_amd as pympcd
    params = pympcd.Params()
    params.n = 10
    params.volume_size = (10, 10, 100)
    params.periodicity = (1, 1, 0)
    params.drag = 0.001
    params.delta_t = 0.02
    sim = pympcd.Simulation(params, "hip")
    sim.step(1000)
    sim.step_and_sample(10000)
    density, velocity = sim.get_mean_fields()

The native extension must be built (python cudampcd_amd/build.py); importing
without it raises ImportError rather than falling back to anything else.
"""
import sys as _sys

if "torch" in _sys.modules:
    # A process holding torch's HIP runtime and this library's must start
    # torch's first, or torch finds no GPU afterwards.
    try:
        _sys.modules["torch"].cuda.is_available()
    except Exception:  # noqa: BLE001
        pass

from ._pympcd import Ensemble, Params, Simulation, __version__  # noqa: F401,E402

from .physics import srd_viscosity, poiseuille_viscosity  # noqa: F401,E402

__all__ = ["Ensemble", "Params", "Simulation", "srd_viscosity", "poiseuille_viscosity", "__version__"]
