"""Drivers (jrystal/calc/__init__.py:17-20).  Only the all-electron total
energy driver is on this path; band-structure and norm-conserving drivers
are outside it."""
from .energy_all_electrons import GroundStateEnergyOutput  # noqa: F401
from .energy_all_electrons import calc as energy_all_electrons  # noqa: F401


def _not_on_path(name):
  def f(config):
    raise NotImplementedError(f"{name} is not part of the direct-minimisation "
                              "total-energy path")
  return f


energy_normcons = _not_on_path("energy_normcons")
band_all_electrons = _not_on_path("band_all_electrons")
band_normcons = _not_on_path("band_normcons")
