"""Seeded synthetic workloads (oracle side).

The same definitions are restated independently in
``jrystal_amd/workloads.py`` for the product; tests check that both produce
identical inputs.
"""
import numpy as np

from . import grid as _grid
from . import occupation as _occ
from .adjoint import Problem
from .crystal import Crystal
from .pw import param_init

SI_FCC_A = 10.26  # bohr


def si2_fcc(z=14.0):
  """Diamond Si primitive cell: FCC a = 10.26 bohr, Si at (0,0,0), (1/4)^3."""
  a = SI_FCC_A / 2 * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0.]])
  frac = np.array([[0, 0, 0], [0.25, 0.25, 0.25]])
  return Crystal(charges=np.full(2, z), positions=frac @ a, cell_vectors=a,
                 spin=0, symbols=["Si"] * 2)


def si8_cube(z=14.0, a=SI_FCC_A, reps=1, strain=None):
  """Conventional diamond Si cube (8 atoms), optionally repeated reps^3
  times (reps=2: the 64-atom supercell) and strained by (I + strain)."""
  fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
  basis = np.concatenate([fcc, fcc + 0.25])
  cells = np.array([[i, j, k] for i in range(reps) for j in range(reps)
                    for k in range(reps)], dtype=float)
  frac = (basis[None, :, :] + cells[:, None, :]).reshape(-1, 3) / reps
  cell = a * reps * np.eye(3)
  if strain is not None:
    cell = cell @ (np.eye(3) + strain)
  n = frac.shape[0]
  return Crystal(charges=np.full(n, z), positions=frac @ cell,
                 cell_vectors=cell, spin=0, symbols=["Si"] * n)


def make_problem(crystal, grid_sizes, k_grid, num_bands, cutoff_energy=None,
                 mask_method="spherical", occupation="uniform"):
  gs = [int(x) for x in grid_sizes]
  a = crystal.cell_vectors
  g = _grid.g_vectors(a, gs)
  if mask_method == "spherical":
    mask = _grid.spherical_mask(a, gs, cutoff_energy)
  else:
    mask = np.asarray(_grid.cubic_mask(gs))
  kpts = _grid.k_vectors(a, k_grid)
  nk = kpts.shape[0]
  ne = crystal.num_electron
  if occupation == "uniform":
    occ = _occ.uniform(nk, ne, 0, num_bands)
  else:
    occ = _occ.gamma(nk, ne, 0, num_bands)
  return Problem(mask=mask, g_vector_grid=g, kpts=kpts, vol=crystal.vol,
                 occupation=occ, positions=crystal.positions,
                 charges=np.asarray(crystal.charges, float))


def cfg1(z=14.0, seed=0):
  """Si2 FCC, 32^3, 2x2x2 MP (8 k), 8 bands, E_cut 24 Ha (G_n = 1459)."""
  cr = si2_fcc(z)
  prob = make_problem(cr, [32] * 3, [2, 2, 2], 8, 24.0)
  params = param_init(seed, 8, 8, prob.mask)
  return cr, prob, params


def cfg2(z=14.0, seed=0, nk_grid=4, strain=None):
  """Si8 cube a=10.26, 64^3, 4x4x4 MP (64 k), 32 bands, E_cut 48 Ha."""
  cr = si8_cube(z, strain=strain)
  prob = make_problem(cr, [64] * 3, [nk_grid] * 3, 32, 48.0)
  params = param_init(seed, 32, prob.kpts.shape[0], prob.mask)
  return cr, prob, params


def cfg5(z=4.0, seed=0):
  """Si64 supercell a=20.52, Gamma only, 128^3, 128 bands, E_cut 48 Ha."""
  cr = si8_cube(z, reps=2)
  prob = make_problem(cr, [128] * 3, [1, 1, 1], 128, 48.0)
  params = param_init(seed, 128, 1, prob.mask)
  return cr, prob, params


def tiny(z=4.0, seed=0, grid=(12, 12, 12), kgrid=(2, 1, 1), nb=3, ecut=6.0):
  """Small FCC Si2 case for brute-force and autograd checks."""
  cr = si2_fcc(z)
  prob = make_problem(cr, grid, kgrid, nb, ecut)
  params = param_init(seed, nb, prob.kpts.shape[0], prob.mask)
  return cr, prob, params
