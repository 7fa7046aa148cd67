"""Grids and masks (jrystal/_src/grid.py), host-side set-up.

These run once per calculation, as in the reference (which builds them with
numpy / ASE outside the jitted step).  Arrays are float64 numpy; the GPU
path takes them from here.
"""
import numpy as np

from .const import FFT_SIZES

_SIZES = np.array(FFT_SIZES)


def fft_factor(n: int) -> int:
  """utils.py:227-254."""
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


def _axis_freq(n: int, normalize: bool):
  return np.fft.fftfreq(n, 1 if normalize else 1 / n)


def _vector_grid(basis, grid_sizes, normalize=False):
  """grid.py:93-119."""
  basis = np.asarray(basis, dtype=np.float64)
  dim = len(grid_sizes)
  out = np.zeros(tuple(grid_sizes) + (dim,))
  for i in range(dim):
    shape = [1] * dim + [dim]
    shape[i] = grid_sizes[i]
    out = out + (_axis_freq(grid_sizes[i], normalize)[:, None] *
                 basis[i][None, :]).reshape(shape)
  return out


def reciprocal_vectors(cell_vectors):
  return 2 * np.pi * np.linalg.inv(np.asarray(cell_vectors, np.float64)).T


def g_vectors(cell_vectors, grid_sizes):
  """grid.py:122-149."""
  return _vector_grid(reciprocal_vectors(cell_vectors), list(grid_sizes))


def r_vectors(cell_vectors, grid_sizes):
  """grid.py:152-178."""
  return _vector_grid(cell_vectors, list(grid_sizes), normalize=True)


def monkhorst_pack(size):
  """ase.dft.kpoints.monkhorst_pack: (i + 1/2)/n - 1/2 in C order."""
  size = np.asarray(size)
  kpts = np.indices(size).transpose((1, 2, 3, 0)).reshape((-1, 3))
  return (kpts + 0.5) / size - 0.5


def k_vectors(cell_vectors, grid_sizes):
  """grid.py:239-262."""
  return monkhorst_pack(grid_sizes) @ reciprocal_vectors(cell_vectors)


def spherical_mask(cell_vectors, grid_sizes, cutoff_energy):
  """grid.py:265-293."""
  g = g_vectors(cell_vectors, grid_sizes)
  return np.linalg.norm(g, axis=-1)**2 <= cutoff_energy * 2


def cubic_mask(grid_sizes):
  """grid.py:296-325."""
  masks = []
  for size in grid_sizes:
    g_max = (size - 1) // 2
    lo, hi = -g_max // 2, g_max // 2
    m = np.ones((size,), dtype=bool)
    m[hi + 1:lo] = False
    masks.append(m)
  return masks[0][:, None, None] & masks[1][None, :, None] & \
      masks[2][None, None, :]


def estimate_max_cutoff_energy(cell_vectors, mask):
  """grid.py:328-350."""
  g = g_vectors(cell_vectors, mask.shape)
  return float(np.max(np.linalg.norm(g, axis=-1)**2 / 2 * mask))


def translation_vectors(cell_vectors, cutoff=1e4):
  """grid.py:213-236: lattice translations m_1 a_1 + m_2 a_2 + m_3 a_3 with
  m_i = fftfreq(n, 1/n) (integers -n/2 .. n/2-1), n = ceil(cutoff/|sum a|^2);
  the reference calls _vector_grid with its default normalize=False."""
  cell_vectors = np.asarray(cell_vectors, np.float64)
  n = int(np.ceil(cutoff / np.linalg.norm(np.sum(cell_vectors, axis=0))**2))
  return _vector_grid(cell_vectors, [n] * 3).reshape(-1, 3)
