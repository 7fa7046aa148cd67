"""Fermi-Dirac entropy of occupations (jrystal/_src/entropy.py:19-54).

S(f) = -sum [f ln(eps + f) + (c - f) ln(eps + c - f)], c = (3 - ns) / nk.
For fixed (uniform / gamma) occupations it is a constant of the run,
evaluated on the host; for a torch tensor (parameterised occupations) it is
a differentiable torch scalar.
"""
import numpy as np
import torch


def fermi_dirac(occupation, eps=1e-8):
  if isinstance(occupation, torch.Tensor):
    num_spin, num_k, _ = occupation.shape
    top = (3 - num_spin) / num_k
    return -torch.sum(occupation * torch.log(eps + occupation) +
                      (top - occupation) * torch.log(eps + top - occupation))
  occ = np.asarray(occupation, np.float64)
  num_spin, num_k, _ = occ.shape
  top = (3 - num_spin) / num_k
  return float(-np.sum(occ * np.log(eps + occ) +
                       (top - occ) * np.log(eps + top - occ)))
