"""Plane-wave coefficients, wave functions and density (oracle restatement).

Follows jrystal/_src/pw.py, jrystal/_src/unitary_module.py:30-81 and
jrystal/_src/utils.py:108-130,161-224,257-308.  All arithmetic is float64 /
complex128.
"""
import numpy as np
import scipy.fft as sfft

_WORKERS = -1  # scipy.fft threads (-1: all host cores); see set_workers


def set_workers(n: int):
  """Number of scipy.fft threads used by the oracle (CPU baseline)."""
  global _WORKERS
  _WORKERS = int(n)


def param_init(seed, num_bands, num_kpts, freq_mask, spin_restricted=True,
               dist="normal"):
  """pw.py:29-91 -> unitary_module.py:84-126.

  The reference draws w_re, w_im from jax.random (a threefry stream that is
  not reproducible without JAX); the oracle draws them from
  ``numpy.random.default_rng(seed)`` -- w_re first, then w_im -- so that both
  the oracle and the HIP path consume the same committed inputs.
  ``dist='normal'`` is N(0,1); ``dist='uniform'`` is the reference's U[0,1).
  """
  num_spin = 1 if spin_restricted else 2
  num_g = int(np.sum(freq_mask))
  shape = (num_spin, num_kpts, num_g, num_bands)
  rng = np.random.default_rng(seed)
  if dist == "normal":
    return {"w_re": rng.standard_normal(shape),
            "w_im": rng.standard_normal(shape)}
  return {"w_re": rng.random(shape), "w_im": rng.random(shape)}


def unitary_matrix(params, complex=True):
  """unitary_module.py:30-81: Q of the reduced QR of W = w_re + i w_im,
  batched over the leading (spin, kpt) axes.  numpy.linalg.qr is the LAPACK
  geqrf/ungqr family that JAX-CPU calls."""
  w = params["w_re"] + (1j * params["w_im"] if complex else 0.0)
  return np.linalg.qr(w, mode="reduced")[0]


def unitary_matrix_with_r(params):
  """Q and R of the reduced QR (used by the adjoint)."""
  w = params["w_re"] + 1j * params["w_im"]
  return np.linalg.qr(w, mode="reduced")


def expand_coefficient(coeff_compact, mask):
  """utils.py:257-281: (s,k,g,b) -> (s,k,b,x,y,z), zero-filled, mask in C
  order."""
  c = np.swapaxes(coeff_compact, -1, -2)
  out = np.zeros(c.shape[:-1] + mask.shape, dtype=c.dtype)
  out[..., mask] = c
  return out


def squeeze_coefficient(coeff, mask):
  """utils.py:284-308: (s,k,b,x,y,z) -> (s,k,g,b)."""
  return np.swapaxes(coeff[..., mask], -1, -2)


def coeff(pw_param, freq_mask):
  """pw.py:94-137: QR then scatter into the FFT box."""
  return expand_coefficient(unitary_matrix(pw_param, complex=True), freq_mask)


def ifftn3d(x):
  """spmd/fft.py:52-70 (single device: jnp.fft.ifftn over the last 3 axes)."""
  if x.ndim < 3:
    raise ValueError(f"Input must have at least 3 dimensions, got {x.ndim}")
  return sfft.ifftn(x, axes=(-3, -2, -1), workers=_WORKERS)


def fftn3d(x):
  """spmd/fft.py:37-49,73-75."""
  if x.ndim < 3:
    raise ValueError(f"Input must have at least 3 dimensions, got {x.ndim}")
  return sfft.fftn(x, axes=(-3, -2, -1), workers=_WORKERS)


def wave_grid(coeff, vol):
  """pw.py:140-211: psi = ifftn(C) * N / sqrt(vol)."""
  grid_sizes = coeff.shape[-3:]
  return ifftn3d(coeff) * (np.prod(grid_sizes) / np.sqrt(vol))


def absolute_square(a):
  """utils.py:108-130."""
  return np.real(np.conj(a) * a)


def density_grid(coeff, vol, occupation=None):
  """pw.py:214-284: rho[s] = sum_{k,b} f |psi|^2 (or per-state if f None)."""
  dens = absolute_square(wave_grid(coeff, vol))
  if occupation is not None:
    if occupation.shape != coeff.shape[:3]:
      raise ValueError(
        "Occupation should have a leading dimension that is the same as "
        f"coeff. Got occupation shape: {occupation.shape}, coeff shape: "
        f"{coeff.shape}")
    dens = np.einsum("skbxyz,skb->sxyz", dens, occupation)
  return dens


def density_grid_reciprocal(coeff, vol, occupation=None):
  """pw.py:287-334: fftn of density_grid (unnormalised)."""
  return fftn3d(density_grid(coeff, vol, occupation))


def wave_to_density(wave, occupation=None):
  """utils.py:161-197."""
  dens = absolute_square(wave)
  if occupation is not None:
    dens = np.sum(dens * occupation[..., None, None, None], axis=(1, 2))
  return dens


def wave_to_density_reciprocal(wave, occupation=None):
  """utils.py:200-224."""
  return fftn3d(wave_to_density(wave, occupation))


def wave_r_bruteforce(r, coeff_box, g_vector_grid, vol):
  """Brute-force psi(r) = vol^-1/2 sum_G C_G exp(i G.r) at explicit points
  (the known-answer method of pw_test.py:36-49).  r: (P,3) -> (..., P)."""
  g = g_vector_grid.reshape(-1, 3)
  c = coeff_box.reshape(coeff_box.shape[:-3] + (-1,))
  phase = np.exp(1j * (g @ np.asarray(r).T))  # (Ng, P)
  return (c @ phase) / np.sqrt(vol)
