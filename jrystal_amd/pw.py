"""Plane-wave functions of jrystal's public API (jrystal/_src/pw.py,
re-exported by jrystal/pw.py), on the GPU.

Same names, arguments and array layouts as the reference:
``param_init`` -> {'w_re', 'w_im'} (spin, kpt, g, band); ``coeff`` ->
(spin, kpt, band, x, y, z) complex128; ``wave_grid``, ``density_grid``,
``density_grid_reciprocal``.  Arrays are torch tensors on the GPU (numpy
inputs are moved there); every function is differentiable with torch
autograd through hand-written HIP backward kernels.
"""
from typing import Optional

import numpy as np
import torch

from . import _ops


def _mask_index(freq_mask):
  m = np.asarray(freq_mask.detach().cpu().numpy()
                 if isinstance(freq_mask, torch.Tensor) else freq_mask,
                 dtype=bool)
  return m, _ops.as_device(np.flatnonzero(m.reshape(-1)), torch.int64)


def param_init(key, num_bands, num_kpts, freq_mask, spin_restricted=True,
               sharding=None, dist: str = "uniform"):
  """pw.py:29-91.  ``key`` is an integer seed (or numpy Generator); values
  are U[0,1) as in the reference (``dist='normal'`` for N(0,1)).  The
  reference's jax.random stream is not reproduced."""
  num_spin = 1 if spin_restricted else 2
  m = np.asarray(freq_mask.cpu() if isinstance(freq_mask, torch.Tensor)
                 else freq_mask, dtype=bool)
  shape = (num_spin, num_kpts, int(m.sum()), num_bands)
  rng = key if isinstance(key, np.random.Generator) else \
      np.random.default_rng(key)
  draw = rng.standard_normal if dist == "normal" else rng.random
  return {"w_re": _ops.as_device(draw(shape), torch.float64),
          "w_im": _ops.as_device(draw(shape), torch.float64)}


def coeff(pw_param, freq_mask, sharding=None):
  """pw.py:94-137: QR orthonormalisation over (g, band) for each (spin,
  kpt), then scatter into the FFT box."""
  m, idx = _mask_index(freq_mask)
  w_re = _ops.as_device(pw_param["w_re"], torch.float64) \
      if not isinstance(pw_param["w_re"], torch.Tensor) else pw_param["w_re"]
  w_im = _ops.as_device(pw_param["w_im"], torch.float64) \
      if not isinstance(pw_param["w_im"], torch.Tensor) else pw_param["w_im"]
  if w_re.shape[2] != int(m.sum()):
    raise ValueError("parameters do not match the number of mask points")
  return _ops.Coefficients.apply(w_re, w_im, idx, m.shape)


def unitary_matrix(params, complex=True):
  """unitary_module.py:30-81 on the GPU: Q (spin, kpt, g, band)."""
  if not complex:
    raise NotImplementedError("only the complex unitary matrix is on this "
                              "path")
  w_re = _ops.as_device(params["w_re"], torch.float64)
  w_im = _ops.as_device(params["w_im"], torch.float64)
  ng = w_re.shape[2]
  box = _ops.Coefficients.apply(w_re, w_im, _ops.as_device(
    np.arange(ng, dtype=np.int64), torch.int64), (ng,))
  return box.transpose(-1, -2)


def wave_grid(coeff, vol):
  """pw.py:140-211: psi = ifftn(C) * N / sqrt(vol)."""
  return _ops.IFFT3.apply(coeff, 1.0 / float(np.sqrt(float(vol))))


def _check_occ(coeff, occupation):
  occ = _ops.as_device(occupation, torch.float64)
  if tuple(occ.shape) != tuple(coeff.shape[:3]):
    raise ValueError(
      "Occupation should have a leading dimension that is the same as coeff. "
      f"Got occupation shape: {tuple(occ.shape)}, coeff shape: "
      f"{tuple(coeff.shape)}")
  return occ


def density_grid(coeff, vol, occupation: Optional[object] = None):
  """pw.py:214-284."""
  occ = None if occupation is None else _check_occ(coeff, occupation)
  return _ops.Density.apply(wave_grid(coeff, vol), occ)


def density_grid_reciprocal(coeff, vol, occupation=None):
  """pw.py:287-334: fftn of density_grid (unnormalised)."""
  rho = density_grid(coeff, vol, occupation)
  return _ops.FFT3.apply(_ops.RealToComplex.apply(rho), 1.0)


def wave_to_density(wave, occupation=None):
  """utils.py:161-197."""
  occ = None if occupation is None else _check_occ(wave, occupation)
  return _ops.Density.apply(wave, occ)


def wave_to_density_reciprocal(wave, occupation=None):
  """utils.py:200-224."""
  return _ops.FFT3.apply(_ops.RealToComplex.apply(
    wave_to_density(wave, occupation)), 1.0)
