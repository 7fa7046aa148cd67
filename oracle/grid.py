"""Reciprocal/real-space grids, masks and FFT sizes (oracle restatement).

Follows jrystal/_src/grid.py and jrystal/_src/utils.py:227-254.
"""
import numpy as np


def fft_sizes(limit: int = 2048):
  """The reference's CUFFT_FACTORS table (const.py:265-456) is exactly the
  7-smooth integers 1..2048; it is regenerated here by that rule."""
  out = []
  for n in range(1, limit + 1):
    m = n
    for p in (2, 3, 5, 7):
      while m % p == 0:
        m //= p
    if m == 1:
      out.append(n)
  return out


_SIZES = np.array(fft_sizes())


def fft_factor(n: int) -> int:
  """utils.py:227-254: smallest table size >= n; ValueError above 2048."""
  if n > 2048:
    raise ValueError(f"The grid number {n} is too large!")
  return int(_SIZES[_SIZES >= n][0])


def proper_grid_size(grid_sizes):
  """grid.py:181-210."""
  if hasattr(grid_sizes, "__len__"):
    grid_sizes = np.array(grid_sizes)
  else:
    try:
      grid_sizes = np.ones(3, dtype=int) * int(grid_sizes)
    except Exception:
      raise TypeError("mesh should be a scalar, tuple, list or np.array.")
  return np.array([fft_factor(int(i)) for i in grid_sizes])


def _vector_grid(basis, grid_sizes, normalize=False):
  """grid.py:93-119: sum_d fftfreq(n_d, 1 or 1/n_d)[i_d] * basis[d]."""
  dim = len(grid_sizes)
  basis = np.asarray(basis, dtype=np.float64)
  out = 0
  for i in range(dim):
    shape = [1] * dim + [dim]
    shape[i] = grid_sizes[i]
    f = np.fft.fftfreq(grid_sizes[i], 1 if normalize else 1 / grid_sizes[i])
    out = out + np.outer(f, basis[i]).reshape(shape)
  return out


def reciprocal_vectors(cell_vectors):
  """grid.py:148 / crystal.py:113-117: B = 2 pi A^{-T} (rows b_i)."""
  return 2 * np.pi * np.linalg.inv(np.asarray(cell_vectors, np.float64)).T


def g_vectors(cell_vectors, grid_sizes):
  """grid.py:122-149: G[i,j,k] = m_i b_1 + m_j b_2 + m_k b_3, FFT order."""
  return _vector_grid(reciprocal_vectors(cell_vectors), list(grid_sizes))


def r_vectors(cell_vectors, grid_sizes):
  """grid.py:152-178."""
  return _vector_grid(cell_vectors, list(grid_sizes), normalize=True)


def monkhorst_pack(size):
  """ASE ``ase.dft.kpoints.monkhorst_pack`` (published algorithm; ASE is a
  reference dependency absent here): C-ordered indices i, frac = (i+1/2)/n-1/2.
  """
  size = np.asarray(size)
  idx = np.indices(size).transpose((1, 2, 3, 0)).reshape((-1, 3))
  return (idx + 0.5) / size - 0.5


def k_vectors(cell_vectors, grid_sizes):
  """grid.py:239-262: monkhorst_pack(grid_sizes) @ B."""
  return monkhorst_pack(grid_sizes) @ reciprocal_vectors(cell_vectors)


def spherical_mask(cell_vectors, grid_sizes, cutoff_energy):
  """grid.py:265-293: |G|^2 <= 2 E_cut (G, not G+k; same mask for all k)."""
  g = g_vectors(cell_vectors, grid_sizes)
  g_norm = np.linalg.norm(g, axis=-1)
  return g_norm**2 <= cutoff_energy * 2


def cubic_mask(grid_sizes):
  """grid.py:296-325."""
  masks = []
  for size in grid_sizes:
    g_max = (size - 1) // 2
    lower = -g_max // 2
    upper = g_max // 2
    m = np.ones((size,), dtype=bool)
    m[upper + 1:lower] = False
    masks.append(m)
  return np.einsum("i,j,k->ijk", *masks)


def estimate_max_cutoff_energy(cell_vectors, mask):
  """grid.py:328-350."""
  g = g_vectors(cell_vectors, mask.shape)
  kin = np.linalg.norm(g, axis=-1)**2 / 2
  return float(np.max(kin * mask))


def translation_vectors(cell_vectors, cutoff=1e4):
  """grid.py:213-236: ``_vector_grid(cell_vectors, [n]*dim)`` with the
  default normalize=False (grid.py:96), i.e. fftfreq(n, 1/n) = integers m, so
  the rows are the lattice translations sum_i m_i a_i, m_i in -n/2..n/2-1,
  with n = ceil(cutoff / |a_1 + a_2 + a_3|^2)."""
  cell_vectors = np.asarray(cell_vectors, np.float64)
  dim = cell_vectors.shape[0]
  n = int(np.ceil(cutoff / np.linalg.norm(np.sum(cell_vectors, axis=0))**2))
  grid = _vector_grid(cell_vectors, [n] * dim)
  return np.reshape(grid, [-1, dim])
