"""Shared helpers of the test suite (not collected: no test_ prefix)."""
import numpy as np

from oracle import adjoint as OA
from oracle import grid as OG
from oracle import occupation as OO
from oracle.crystal import Crystal as OCrystal


def oracle_problem(crystal, grid_sizes, k_grid, num_bands, cutoff=None,
                   mask_method="spherical", spin_restricted=True,
                   occupation="uniform"):
  """Oracle Problem built from the oracle's own set-up functions."""
  cr = OCrystal(charges=crystal.charges, positions=crystal.positions,
                cell_vectors=crystal.cell_vectors)
  gs = [int(x) for x in grid_sizes]
  g = OG.g_vectors(cr.cell_vectors, gs)
  if mask_method == "spherical":
    mask = OG.spherical_mask(cr.cell_vectors, gs, cutoff)
  else:
    mask = np.asarray(OG.cubic_mask(gs))
  kpts = OG.k_vectors(cr.cell_vectors, k_grid)
  nk = kpts.shape[0]
  fn = OO.uniform if occupation == "uniform" else OO.gamma
  occ = fn(nk, cr.num_electron, 0, num_bands, spin_restricted)
  return OA.Problem(mask=mask, g_vector_grid=g, kpts=kpts, vol=cr.vol,
                    occupation=occ, positions=cr.positions,
                    charges=np.asarray(cr.charges, float))


def seeded_params(shape, seed=0):
  rng = np.random.default_rng(seed)
  return {"w_re": rng.standard_normal(shape), "w_im": rng.standard_normal(shape)}


def rel_max(a, b):
  """max |a - b| / max |b|."""
  a = np.asarray(a)
  b = np.asarray(b)
  return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
