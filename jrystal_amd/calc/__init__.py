"""Drivers (jrystal/calc/__init__.py:17-20): the all-electron total energy
and band-structure drivers; the norm-conserving drivers are outside this
path."""
from .band_all_electrons import BandStructureOutput  # noqa: F401
from .band_all_electrons import calc as band_all_electrons  # noqa: F401
from .energy_all_electrons import GroundStateEnergyOutput  # noqa: F401
from .energy_all_electrons import calc as energy_all_electrons  # noqa: F401


def _not_on_path(name):
  def f(config):
    raise NotImplementedError(f"{name} is not part of the direct-minimisation "
                              "total-energy path")
  return f


energy_normcons = _not_on_path("energy_normcons")
band_normcons = _not_on_path("band_normcons")
