"""Fermi-Dirac entropy of fixed occupations (jrystal/_src/entropy.py:19-54).

For the fixed (uniform / gamma) occupations of this path the entropy is a
constant of the run; it is evaluated once on the host from the occupation
array, exactly as written in the reference.
"""
import numpy as np


def fermi_dirac(occupation, eps=1e-8):
  occ = np.asarray(occupation, np.float64)
  num_spin, num_k, _ = occ.shape
  top = (3 - num_spin) / num_k
  return float(-np.sum(occ * np.log(eps + occ) +
                       (top - occ) * np.log(eps + top - occ)))
