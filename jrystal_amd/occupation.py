"""Fixed occupations (jrystal/_src/occupation.py:147-278)."""
import numpy as np


def uniform(num_k, num_electrons, spin=0, num_bands=None, spin_restricted=True):
  """occupation.py:147-190 (silently clips when Ne/2 > num_bands)."""
  num_bands = num_electrons if num_bands is None else num_bands
  occ = np.zeros([2, num_k, num_bands])
  occ[0, :, :(num_electrons + spin) // 2] = 1 / num_k
  occ[1, :, :(num_electrons - spin) // 2] = 1 / num_k
  if spin_restricted:
    return np.sum(occ, axis=0, keepdims=True)
  return occ


def gamma(num_k, num_electrons, spin=0, num_bands=None, spin_restricted=True):
  """occupation.py:193-237."""
  num_bands = num_electrons if num_bands is None else num_bands
  occ = np.zeros([2, num_k, num_bands])
  occ[0, 0, :(num_electrons + spin) // 2] = 1
  occ[1, 0, :(num_electrons - spin) // 2] = 1
  if spin_restricted:
    return np.sum(occ, axis=0, keepdims=True)
  return occ


def param_init(key, num_bands, num_electrons, num_kpts, spin=0,
               method="uniform", spin_restricted=True):
  """occupation.py:240-258 for the fixed methods on this path."""
  if method == "uniform":
    return uniform(num_kpts, num_electrons, spin, num_bands, spin_restricted)
  if method == "gamma":
    return gamma(num_kpts, num_electrons, spin, num_bands, spin_restricted)
  raise NotImplementedError(
    f"occupation method {method!r} is not on the fixed-occupation path")


def occupation(params, num_kpts, num_electrons=None, spin=0,
               method="uniform", spin_restricted=True):
  """occupation.py:261-278: fixed occupations are constants (stop_gradient).
  """
  if method in ("uniform", "gamma"):
    return params
  raise NotImplementedError(
    f"occupation method {method!r} is not on the fixed-occupation path")
