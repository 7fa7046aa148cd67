"""Drivers (jrystal/calc/__init__.py:17-20): total energy and band
structure, all-electron or with norm-conserving pseudopotentials (the same
drivers with use_pseudopotential set)."""
from .band_all_electrons import BandStructureOutput  # noqa: F401
from .band_all_electrons import calc as band_all_electrons  # noqa: F401
from .energy_all_electrons import GroundStateEnergyOutput  # noqa: F401
from .energy_all_electrons import calc as energy_all_electrons  # noqa: F401


def energy_normcons(config, resume=None):
  """calc_ground_state_energy_normcons.calc: the energy driver with the
  config's norm-conserving pseudopotentials."""
  config.use_pseudopotential = True
  return energy_all_electrons(config, resume=resume)


def band_normcons(config):
  """calc_band_structure_normcons.calc: the band driver with the config's
  norm-conserving pseudopotentials."""
  config.use_pseudopotential = True
  return band_all_electrons(config)
