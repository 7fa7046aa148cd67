"""Occupations (jrystal/_src/occupation.py).

Fixed methods (uniform, gamma) are host arrays, constants of a run.  The
parameterised methods (simplex-projector, idempotent) map parameters to
occupations with torch operations on the parameters' device, so that
autograd gives the occupation Jacobian the reference gets from JAX: the
projection differentiates through the sort, the cumulative sums and the
clamp with the active set (the integer k) held fixed, and the idempotent
map through a real QR.
"""
import numpy as np
import torch

from . import _ops


def uniform(num_k, num_electrons, spin=0, num_bands=None, spin_restricted=True):
  """occupation.py:147-190 (silently clips when Ne/2 > num_bands)."""
  num_bands = num_electrons if num_bands is None else num_bands
  occ = np.zeros([2, num_k, num_bands])
  occ[0, :, :(num_electrons + spin) // 2] = 1 / num_k
  occ[1, :, :(num_electrons - spin) // 2] = 1 / num_k
  if spin_restricted:
    return np.sum(occ, axis=0, keepdims=True)
  return occ


def gamma(num_k, num_electrons, spin=0, num_bands=None, spin_restricted=True):
  """occupation.py:193-237."""
  num_bands = num_electrons if num_bands is None else num_bands
  occ = np.zeros([2, num_k, num_bands])
  occ[0, 0, :(num_electrons + spin) // 2] = 1
  occ[1, 0, :(num_electrons - spin) // 2] = 1
  if spin_restricted:
    return np.sum(occ, axis=0, keepdims=True)
  return occ


PARAMETERISED = ("simplex-projector", "idempotent")


def _device(device):
  return torch.device(device if device is not None else
                      ("cuda" if torch.cuda.is_available() else "cpu"))


def simplex_projector_init(num_bands, num_kpts, device=None):
  """occupation.py:281-292: (arange(n) - n//2) * 0.1, n = bands * kpts,
  reshaped (kpt, band), for both spin channels."""
  n = num_bands * num_kpts
  p = ((torch.arange(n, dtype=torch.float64) - n // 2) * 0.1).reshape(
    num_kpts, num_bands).to(_device(device))
  return {"param_up": p, "param_down": p.clone()}


def proj(x, total):
  """occupation.py:295-337: the point of {0 <= y <= 1, sum y = total}
  reached from x (1-D device tensor, entries in [0, 1]) by a uniform shift
  and clipping: shift down and clip at 0 when sum x > total, else shift up
  and clip at 1.  The shift follows from the sorted entries and their
  cumulative sums with the same active-set rule as the reference (HIP
  kernel jrx_simplex_projection; differentiable, with the active set held
  fixed as JAX does for the reference)."""
  return _ops.SimplexProjection.apply(x.reshape(1, -1), [float(total)],
                                      False).reshape(x.shape)


def simplex_projector(params, num_electrons, spin=0, spin_restricted=True):
  """occupation.py:340-366: occupation (spin, kpt, band) from sigmoid of
  the parameters projected per spin channel onto sum = electrons * kpts,
  0 <= f <= 1, divided by the number of k-points (both channels in one
  launch of the projection kernel)."""
  num_kpts, num_bands = params["param_up"].shape
  m_up = (num_electrons + spin) // 2 * num_kpts
  m_down = (num_electrons - spin) // 2 * num_kpts
  theta = torch.stack([params["param_up"].reshape(-1),
                       params["param_down"].reshape(-1)])
  occ = _ops.SimplexProjection.apply(theta, [float(m_up), float(m_down)],
                                     True)
  occ = occ.reshape(2, num_kpts, num_bands) / num_kpts
  if spin_restricted:
    return occ.sum(0, keepdim=True)
  return occ


def idempotent_param_init(key, num_bands, num_electrons, num_kpts, spin=0,
                          spin_restricted=True, device=None):
  """occupation.py:27-84: real parameters U[0,1) of shape
  (bands * kpts, electrons-per-spin * kpts) per spin channel (the down
  channel starts equal to the up channel when spin-restricted).  ``key`` is
  an integer seed or numpy Generator; the reference's jax.random stream is
  not reproduced."""
  if (num_electrons + spin) % 2 != 0:
    raise ValueError("num_electrons and spin must have the same parity")
  rng = key if isinstance(key, np.random.Generator) else \
      np.random.default_rng(key)
  dev = _device(device)
  up = (num_electrons + spin) // 2
  down = (num_electrons - spin) // 2
  p_up = torch.tensor(rng.random((num_bands * num_kpts, up * num_kpts)),
                      dtype=torch.float64, device=dev)
  if spin_restricted:
    return {"param_up": {"w_re": p_up}, "param_down": {"w_re": p_up.clone()}}
  p_down = torch.tensor(rng.random((num_bands * num_kpts, down * num_kpts)),
                        dtype=torch.float64, device=dev)
  return {"param_up": {"w_re": p_up}, "param_down": {"w_re": p_down}}


def idempotent(params, num_kpts, spin_restricted=True):
  """occupation.py:87-144: f = row norms squared of the orthonormal factor
  of the real parameter matrix (a diagonal of the projector U U^T),
  reshaped (kpt, band), divided by the number of k-points."""
  def rows(p):
    u = torch.linalg.qr(p["w_re"], mode="reduced")[0]
    return (u * u).sum(1)
  nb = params["param_up"]["w_re"].shape[0] // num_kpts
  up = rows(params["param_up"]).reshape(num_kpts, nb)
  down = rows(params["param_down"]).reshape(num_kpts, nb)
  if spin_restricted:
    return ((up + down) / num_kpts).unsqueeze(0)
  return torch.stack([up, down]) / num_kpts


def param_init(key, num_bands, num_electrons, num_kpts, spin=0,
               method="uniform", spin_restricted=True, device=None):
  """occupation.py:240-258."""
  if method == "uniform":
    return uniform(num_kpts, num_electrons, spin, num_bands, spin_restricted)
  if method == "gamma":
    return gamma(num_kpts, num_electrons, spin, num_bands, spin_restricted)
  if method == "simplex-projector":
    return simplex_projector_init(num_bands, num_kpts, device)
  if method == "idempotent":
    return idempotent_param_init(key, num_bands, num_electrons, num_kpts,
                                 spin, spin_restricted, device)
  raise ValueError(f"Invalid method: {method}")


def occupation(params, num_kpts, num_electrons=None, spin=0,
               method="uniform", spin_restricted=True):
  """occupation.py:261-278: fixed occupations are constants
  (stop_gradient); parameterised ones are differentiable torch tensors."""
  if method in ("uniform", "gamma"):
    return params
  if method == "simplex-projector":
    return simplex_projector(params, num_electrons, spin, spin_restricted)
  if method == "idempotent":
    return idempotent(params, num_kpts, spin_restricted)
  raise ValueError(f"Invalid method: {method}")


def leaves(params):
  """Flat {name: tensor} view of a parameterised occupation's parameters
  (the optimiser's leaves) and a function rebuilding the nested dict from
  such a view."""
  flat = {}
  for ch, v in params.items():
    if isinstance(v, dict):
      for k, t in v.items():
        flat[f"occ.{ch}.{k}"] = t
    else:
      flat[f"occ.{ch}"] = v

  def rebuild(view):
    out = {}
    for ch, v in params.items():
      out[ch] = ({k: view[f"occ.{ch}.{k}"] for k in v}
                 if isinstance(v, dict) else view[f"occ.{ch}"])
    return out
  return flat, rebuild
