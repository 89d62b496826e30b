"""MI355X (gfx950) implementation of sphere's DEM time step.

Modules: ``spherebin`` (state + .bin codec, no GPU), ``configs`` (seeded
workloads), ``dem`` (the DEM class over the C-ABI library
``libsphere_dem.so``), ``build`` (hipcc build of that library).
"""

from .spherebin import Spherebin  # noqa: F401

__all__ = ["Spherebin"]
