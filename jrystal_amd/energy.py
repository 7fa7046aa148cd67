"""Energy functions of jrystal's public API (jrystal/_src/energy.py,
re-exported by jrystal/energy.py), on the GPU.

Same names and arguments as the reference; results are 0-dim float64 GPU
tensors, differentiable with torch autograd (hand-written HIP backward).
"""
import numpy as np
import torch

from . import _ops
from . import grid as _grid
from . import pw as _pw


def _g(g_vector_grid):
  return _ops.as_device(g_vector_grid, torch.float64)


def _require_spin_axis(x, dim=3):
  assert x.ndim == dim + 1, "density must contains spin axis"


def hartree(density_grid_reciprocal, g_vector_grid, vol, kohn_sham=False):
  """energy.py:29-82 (potential.py:24-77, braket.py:27-72)."""
  _require_spin_axis(density_grid_reciprocal)
  return _ops.Hartree.apply(density_grid_reciprocal, _g(g_vector_grid),
                            float(vol), bool(kohn_sham))


def external(density_grid_reciprocal, position, charge, g_vector_grid, vol):
  """energy.py:85-135 (potential.py:121-166, literal +1e-10)."""
  _require_spin_axis(density_grid_reciprocal)
  return _ops.External.apply(density_grid_reciprocal, _g(g_vector_grid),
                             _ops.as_device(position, torch.float64),
                             _ops.as_device(charge, torch.float64))


def kinetic(g_vector_grid, kpts, coeff_grid, occupation=None):
  """energy.py:138-182: scalar with occupation, else (spin, kpt, band)."""
  occ = None if occupation is None else _ops.as_device(occupation,
                                                        torch.float64)
  return _ops.Kinetic.apply(coeff_grid, _g(g_vector_grid),
                            _ops.as_device(kpts, torch.float64), occ)


def xc_energy(density_grid, g_vector_grid, vol, xc_type="lda_x",
              kohn_sham=False):
  """energy.py:185-211: LDA exchange with or without spin (xc.py:54-64),
  and for unpolarised densities any '+'-joined sum of lda_x, lda_c_pw,
  gga_x_pbe and gga_c_pbe (xc.py:67-106,222-253; the gradient boxes are
  formed as sigma_r_fn forms them)."""
  assert density_grid.ndim == 4, "density_grid must contains spin axis"
  if kohn_sham:
    # xc_density(kohn_sham=True) is vxc_lda (xc.py:109-124, 247-250): for
    # one channel of LDA exchange d(rho eps)/d rho = 4/3 eps pointwise, so
    # the sum of v rho is 4/3 of the exchange energy
    if xc_type != "lda_x" or density_grid.shape[0] != 1:
      raise NotImplementedError("kohn_sham xc energy: unpolarised lda_x only")
    return (4.0 / 3.0) * _ops.ExchangeLDA.apply(density_grid, float(vol))
  if xc_type == "lda_x":
    return _ops.ExchangeLDA.apply(density_grid, float(vol))
  from .objective import xc_terms
  flags = xc_terms(xc_type)
  if density_grid.shape[0] != 1:
    raise NotImplementedError(
      f"xc {xc_type!r}: spin-polarised densities support lda_x only")
  gvec = _ops.as_device(g_vector_grid, torch.float64)
  return _ops.ExchangeCorrelation.apply(density_grid, gvec, float(vol), flags)


def ewald_coulomb_repulsion(positions, charges, g_vector_grid, vol,
                            ewald_eta, ewald_grid):
  """ewald.py:21-86 (reciprocal part in structure-factor form)."""
  return _ops.ewald(positions, charges, g_vector_grid, float(vol),
                    float(ewald_eta), ewald_grid)


def nuclear_repulsion(position, charge, cell_vectors, g_vector_grid, vol,
                      ewald_eta, ewald_cutoff):
  """energy.py:214-243."""
  return ewald_coulomb_repulsion(
    position, charge, g_vector_grid, vol, ewald_eta,
    _grid.translation_vectors(cell_vectors, ewald_cutoff))


def total_energy(coefficient, position, charge, g_vector_grid, kpts, vol,
                 occupation=None, kohn_sham=False, xc="lda_x", split=False):
  """energy.py:246-307: E_kin + E_ext + E_H + E_xc (no Ewald term); the
  default occupation is all ones."""
  wave = _pw.wave_grid(coefficient, vol)
  if occupation is None:
    occupation = np.ones(tuple(coefficient.shape[:3]))
  rho = _pw.wave_to_density(wave, occupation)
  rho_rec = _ops.FFT3.apply(_ops.RealToComplex.apply(rho), 1.0)
  e_kin = kinetic(g_vector_grid, kpts, coefficient, occupation)
  e_ext = external(rho_rec, position, charge, g_vector_grid, vol)
  e_har = hartree(rho_rec, g_vector_grid, vol, kohn_sham)
  e_xc = xc_energy(rho, g_vector_grid, vol, xc, kohn_sham)
  if split:
    return e_kin, e_ext, e_har, e_xc
  return e_kin + e_ext + e_har + e_xc
