"""Fixed occupations (oracle restatement of jrystal/_src/occupation.py)."""
import numpy as np


def uniform(num_k, num_electrons, spin=0, num_bands=None, spin_restricted=True):
  """occupation.py:147-190: 1/nk on the first (Ne+spin)//2 (up) and
  (Ne-spin)//2 (down) bands; slicing silently clips at num_bands."""
  num_bands = num_electrons if num_bands is None else num_bands
  occ = np.zeros([2, num_k, num_bands])
  occ[0, :, :(num_electrons + spin) // 2] = 1 / num_k
  occ[1, :, :(num_electrons - spin) // 2] = 1 / num_k
  if spin_restricted:
    return np.sum(occ, axis=0, keepdims=True)
  return occ


def gamma(num_k, num_electrons, spin=0, num_bands=None, spin_restricted=True):
  """occupation.py:193-237: occupations only at the first k-point."""
  num_bands = num_electrons if num_bands is None else num_bands
  occ = np.zeros([2, num_k, num_bands])
  occ[0, 0, :(num_electrons + spin) // 2] = 1
  occ[1, 0, :(num_electrons - spin) // 2] = 1
  if spin_restricted:
    return np.sum(occ, axis=0, keepdims=True)
  return occ


def fermi_dirac_entropy(occupation, eps=1e-8):
  """entropy.py:19-54."""
  num_spin, num_k, _ = occupation.shape
  top = (3 - num_spin) / num_k
  return -np.sum(occupation * np.log(eps + occupation) +
                 (top - occupation) * np.log(eps + top - occupation))
