"""Seeded synthetic workloads of the direct-minimisation path.

All charges are nuclear (Coulomb) charges; parameters are drawn from
numpy.random.default_rng(seed): w_re first, then w_im, standard normal.

* ``si2_fcc``  -- diamond Si primitive cell, FCC a = 10.26 bohr, 32^3 grid,
  2x2x2 Monkhorst-Pack k-mesh (8 k), 8 bands, spherical cutoff 24 Ha.
* ``si8_cube`` -- conventional diamond Si cube a = 10.26 bohr (8 atoms),
  64^3 grid, 4x4x4 k-mesh (64 k), 32 bands, spherical cutoff 48 Ha.
* ``si8_strained`` -- si8_cube under a random symmetric strain (held-out
  case of the preconditioner comparison).
* ``si64``     -- 2x2x2 supercell of the cube (a = 20.52 bohr, 64 atoms),
  Gamma only, 128^3 grid, 128 bands, cutoff 48 Ha.
"""
from dataclasses import dataclass

import numpy as np

from . import grid as _grid
from . import occupation as _occ
from .crystal import Crystal

SI_A = 10.26


@dataclass
class Workload:
  name: str
  crystal: Crystal
  grid_sizes: tuple
  k_grid: tuple
  num_bands: int
  cutoff_energy: float
  mask: np.ndarray
  kpts: np.ndarray
  occupation: np.ndarray

  @property
  def num_g(self):
    return int(self.mask.sum())

  @property
  def shape(self):
    return (self.occupation.shape[0], self.kpts.shape[0], self.num_g,
            self.num_bands)

  def params(self, seed=0):
    rng = np.random.default_rng(seed)
    w_re = rng.standard_normal(self.shape)
    w_im = rng.standard_normal(self.shape)
    return {"w_re": w_re, "w_im": w_im}


def fcc_si2(z=14.0):
  a = SI_A / 2 * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0.]])
  frac = np.array([[0, 0, 0], [0.25, 0.25, 0.25]])
  return Crystal(charges=np.full(2, z), positions=frac @ a, cell_vectors=a,
                 spin=0, symbols=["Si"] * 2)


def diamond_cube(z=14.0, a=SI_A, reps=1, strain=None):
  fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
  basis = np.concatenate([fcc, fcc + 0.25])
  shifts = np.array([[i, j, k] for i in range(reps) for j in range(reps)
                     for k in range(reps)], float)
  frac = (basis[None] + shifts[:, None]).reshape(-1, 3) / reps
  cell = a * reps * np.eye(3)
  if strain is not None:
    cell = cell @ (np.eye(3) + np.asarray(strain))
  n = frac.shape[0]
  return Crystal(charges=np.full(n, z), positions=frac @ cell,
                 cell_vectors=cell, spin=0, symbols=["Si"] * n)


def build(name, crystal, grid_sizes, k_grid, num_bands, cutoff_energy,
          occupation="uniform"):
  gs = tuple(int(x) for x in grid_sizes)
  mask = _grid.spherical_mask(crystal.cell_vectors, gs, cutoff_energy)
  kpts = _grid.k_vectors(crystal.cell_vectors, k_grid)
  occ = _occ.param_init(None, num_bands, crystal.num_electron, kpts.shape[0],
                        0, occupation)
  return Workload(name, crystal, gs, tuple(k_grid), num_bands, cutoff_energy,
                  mask, kpts, occ)


def si2_fcc(z=14.0):
  return build("si2-fcc-32^3-8k-8b", fcc_si2(z), (32,) * 3, (2, 2, 2), 8,
               24.0)


def si8_cube(z=14.0, k=4, strain=None):
  return build(f"si8-cube-64^3-{k**3}k-32b", diamond_cube(z, strain=strain),
               (64,) * 3, (k, k, k), 32, 48.0)


def strain_matrix(seed=4, amplitude=0.1):
  """Symmetric strain with entries ~ U(-amplitude, amplitude):
  e = rng.uniform(-a, a, (3, 3)), strain = (e + e^T) / 2."""
  e = np.random.default_rng(seed).uniform(-amplitude, amplitude, (3, 3))
  return (e + e.T) / 2


def si8_strained(z=4.0, k=4, seed=4, amplitude=0.1):
  """The 8-atom cube under a random symmetric lattice strain, cell
  a (I + strain), same grid, k-mesh, bands and cutoff as si8_cube."""
  wl = si8_cube(z, k, strain=strain_matrix(seed, amplitude))
  wl.name = f"si8-strained{seed}-64^3-{k**3}k-32b"
  return wl


SMOOTH_32_80 = tuple(n for n in range(32, 81) if _grid.fft_factor(n) == n)


def alias_free_cutoff(cell, n):
  """Largest integer cutoff (Ha) whose G-sphere fits |m_d| <= n_d / 4 (the
  alias-free half box of grid.cubic_mask): planes of constant m_d lie
  2 pi / |a_d| apart."""
  cell = np.asarray(cell, float)
  g = min((nd // 4) * 2 * np.pi / np.linalg.norm(cell[d])
          for d, nd in enumerate(n))
  return float(np.floor(0.5 * g * g))


def batch_cells(seed=3, n=16):
  """Uneven structures for the batched step: numpy.random.default_rng(seed)
  draws, per structure, a cubic or FCC cell of 2-16 atoms (volume per atom
  U(60, 140) bohr^3), Z from {3, 6, 11, 13, 14}, fractional positions
  U[0,1) redrawn until every minimum-image separation is >= 2.5 bohr, a
  cubic FFT grid from the 7-smooth sizes in [32, 80], a k-mesh of 1^3 to
  4^3, 4-48 bands and the alias-free cutoff of the grid.  Occupations are
  uniform (clipped at 2 per band, as in the reference)."""
  rng = np.random.default_rng(seed)
  out = []
  for i in range(n):
    lattice = ("cubic", "fcc")[int(rng.integers(2))]
    natom = int(rng.integers(2, 17))
    z = float(rng.choice([3.0, 6.0, 11.0, 13.0, 14.0]))
    vol = natom * rng.uniform(60.0, 140.0)
    if lattice == "cubic":
      cell = vol**(1 / 3) * np.eye(3)
    else:
      a = (4 * vol)**(1 / 3)
      cell = a / 2 * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0.]])
    shifts = np.array([[i1, i2, i3] for i1 in (-1, 0, 1) for i2 in (-1, 0, 1)
                       for i3 in (-1, 0, 1)], float) @ cell
    while True:
      frac = rng.random((natom, 3))
      pos = frac @ cell
      d = pos[None, :, None, :] - pos[:, None, None, :] + shifts[None, None]
      r = np.linalg.norm(d, axis=-1)
      r[np.arange(natom), np.arange(natom), 13] = np.inf  # self, no shift
      if r.min() >= 2.5:
        break
    ng = int(rng.choice(SMOOTH_32_80))
    k = int(rng.integers(1, 5))
    nb = int(rng.integers(4, 49))
    gs = (ng, ng, ng)
    cr = Crystal(charges=np.full(natom, z), positions=pos, cell_vectors=cell,
                 spin=0, symbols=["X"] * natom)
    out.append(build(f"batch{seed}-{i}-{lattice}{natom}-z{int(z)}-{ng}^3-"
                     f"{k**3}k-{nb}b", cr, gs, (k, k, k), nb,
                     alias_free_cutoff(cell, gs)))
  return out


def si64(z=4.0):
  return build("si64-128^3-gamma-128b", diamond_cube(z, reps=2), (128,) * 3,
               (1, 1, 1), 128, 48.0)


def tiny(z=4.0, grid=(12, 12, 12), kgrid=(2, 1, 1), nb=3, ecut=6.0):
  return build("tiny", fcc_si2(z), grid, kgrid, nb, ecut)
