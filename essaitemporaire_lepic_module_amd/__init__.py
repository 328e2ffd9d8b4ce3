"""MI355X (gfx950) implementation of the LePIC 3D explicit PIC-MCC time step.

Modules mirror the reference's modules/ call surface:
  particles  -- ParticleSet, sort, movers, wall BC, MCC, loaders
  fields     -- deposition, density reduction, rho, Poisson (MG/RB-SOR), E
  simulation -- Simulation.advance_one_step
All compute runs in liblepic_hip.so (hand-written HIP, include/lepic.h).
"""
from . import config  # noqa: F401

__all__ = ["config", "particles", "fields", "simulation"]


def __getattr__(name):
    if name in ("particles", "fields", "simulation", "device"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
