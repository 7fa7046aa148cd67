"""Energy and closed-form gradient of the direct-minimisation objective.

Restates the closed ``total_energy`` of
jrystal/calc/calc_ground_state_energy_all_electrons.py:119-137
(kinetic + hartree + external + xc for fixed occupations) and the gradient
that ``jax.value_and_grad`` returns for it (:178), written as the explicit
adjoint chain:

  v(r)      = dE/drho(r): Hartree + external + LDA-x potentials, * vol/N
  Cbar      = f [ (2/sqrt(vol)) fftn(v psi) + |G+k|^2 C ]   (torch convention:
              dE/dRe C + i dE/dIm C)
  Qbar      = Cbar gathered at the mask
  Wbar      = (Qbar + Q copyltu(M)) R^{-H},  copyltu(M) = X + X^H - Re diag X,
              X = triu(-Q^H Qbar)     (QR backward; Walter & Lehmann 2018,
              Liao et al. PRX 9 031041 (2019))
  dE/dw_re  = Re Wbar,  dE/dw_im = Im Wbar.

Works k-chunk by k-chunk so that the cfg2 shape (64 k x 32 bands x 64^3)
fits in host memory.  Pinned in tests against torch CPU autograd (an
independent AD system) and central finite differences.
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.fft as sfft

from . import energy as _en
from . import pw as _pw


@dataclass
class Problem:
  """Everything the objective needs besides the parameters."""
  mask: np.ndarray          # (x,y,z) bool
  g_vector_grid: np.ndarray  # (x,y,z,3)
  kpts: np.ndarray          # (nk,3)
  vol: float
  occupation: np.ndarray    # (s,k,b)
  positions: np.ndarray     # (atom,3)
  charges: np.ndarray       # (atom,)
  xc: str = "lda_x"
  # norm-conserving pseudopotential (oracle/pseudopotential.py): the local
  # potential box in the reference's external_reciprocal normalisation,
  # replacing the Coulomb nuclei, and the non-local projectors (k, p, g)
  v_ext_rec: Optional[np.ndarray] = None
  projectors: Optional[np.ndarray] = None

  @property
  def grid_sizes(self):
    return self.mask.shape

  def mask_g(self):
    """G vectors at the mask points in C order, (ng,3)."""
    return self.g_vector_grid[self.mask]


def hartree_external_potential(rho_rec, prob: Problem):
  """V_H + V_ext on the grid with the vol/N weight, from rho_rec = fftn(rho)
  (s,x,y,z): V_H = ifftn(4 pi rho_tot(G) / |G|^2), G = 0 dropped
  (potential.py:24-77 with kohn_sham), V_ext = ifftn of
  -4 pi sum_a Z_a e^{-iG.R_a} / (|G|^2 + 1e-10) (potential.py:121-166)."""
  g = prob.g_vector_grid
  vol = prob.vol
  n = np.prod(rho_rec.shape[-3:])
  if prob.v_ext_rec is not None:  # local pseudopotential
    v_e = (vol / n) * np.real(sfft.ifftn(prob.v_ext_rec, workers=_pw._WORKERS))
  rho_tot_rec = np.sum(rho_rec, axis=0)
  g2 = np.sum(g**2, axis=-1)
  g2h = g2.copy()
  g2h[0, 0, 0] = 1.0
  vh_rec = 4 * np.pi * rho_tot_rec / g2h
  vh_rec[0, 0, 0] = 0.0
  v_h = (vol / n) * np.real(sfft.ifftn(vh_rec, workers=_pw._WORKERS))
  if prob.v_ext_rec is None:
    s_conj = np.exp(-1j * (g @ prob.positions.T)) @ prob.charges.astype(float)
    ve_rec = s_conj / (g2 + 1e-10)
    ve_rec[0, 0, 0] = 0.0
    v_e = -4 * np.pi * np.real(sfft.ifftn(ve_rec, workers=_pw._WORKERS))
  return v_h + v_e


def potential_real(rho, prob: Problem):
  """v_s(r) = dE/drho_s(r) for E = E_H + E_ext + E_xc (vol/N included).

  Returns (v (s,x,y,z), rho_rec (s,x,y,z), energies dict)."""
  g = prob.g_vector_grid
  vol = prob.vol
  n = np.prod(rho.shape[-3:])
  rho_rec = sfft.fftn(rho, axes=(-3, -2, -1), workers=_pw._WORKERS)
  e_har = _en.hartree(rho_rec, g, vol)
  if prob.v_ext_rec is None:
    e_ext = _en.external(rho_rec, prob.positions, prob.charges, g, vol)
  else:  # energy_local (local.py:158-187)
    e_ext = float(_en.reciprocal_braket(prob.v_ext_rec, rho_rec, vol))
  e_xc = _en.xc_energy(rho, g, vol, prob.xc)
  v_he = hartree_external_potential(rho_rec, prob)
  if prob.xc != "lda_x":
    from . import xc as _xc
    v_xc = _xc.xc_potential(rho, g, vol, prob.xc)
  elif rho.shape[0] == 1:
    v_xc = (vol / n) * (4.0 / 3.0) * _en.lda_x_eps(rho)
  else:
    # E = vol/N sum eps(r) rho_tot, eps = 0.5 (e(2 rho_u) + e(2 rho_d))
    eps = _en.xc_density(rho)
    rho_tot = rho[0] + rho[1]
    v_xc = np.empty_like(rho)
    for s in range(2):
      n2 = 2 * rho[s]
      de = np.where(n2 > _en.LDA_DENS_THRESHOLD,
                    -_en.CX / 3.0 * np.cbrt(np.maximum(n2, 1e-300))**-2, 0.0)
      v_xc[s] = (vol / n) * (eps + rho_tot * de)
  v = v_he[None] + v_xc
  return v, rho_rec, {"hartree": e_har, "external": e_ext, "xc": e_xc}


def qr_backward(q, r, qbar):
  """Wbar for L(Q(W)) with R unused; batched over leading axes."""
  qh = np.conj(np.swapaxes(q, -1, -2))
  x = np.triu(-qh @ qbar)
  xh = np.conj(np.swapaxes(x, -1, -2))
  d = np.real(np.diagonal(x, axis1=-2, axis2=-1))
  m = x + xh - d[..., None] * np.eye(x.shape[-1])
  y = qbar + q @ m
  # Y R^{-H}: solve Wbar R^H = Y  <=>  R Wbar^H = Y^H
  rh_inv_t = np.linalg.solve(r, np.conj(np.swapaxes(y, -1, -2)))
  return np.conj(np.swapaxes(rh_inv_t, -1, -2))


def energy_and_grad(params, prob: Problem, k_chunk: Optional[int] = None,
                    want_grad: bool = True, extras: bool = False,
                    grad_kpts=None):
  """Returns (energies dict incl. 'total', grad dict or None, rho); with
  extras, a fourth item {'q': Q, 'qbar': Qbar} (both (s,k,g,b)).
  grad_kpts (optional sequence of k indices) restricts the gradient to those
  k-points (the density and energies still use every k); the other k-points
  of the returned gradient are NaN."""
  w = params["w_re"] + 1j * params["w_im"]
  q, r = np.linalg.qr(w, mode="reduced")  # (s,k,g,b), (s,k,b,b)
  ns, nk, ng, nb = q.shape
  mask = prob.mask
  shape = mask.shape
  n = int(np.prod(shape))
  vol = prob.vol
  occ = prob.occupation
  k_chunk = k_chunk or max(1, min(nk, 8))
  scale = n / np.sqrt(vol)
  rho = np.zeros((ns,) + shape)
  for s in range(ns):
    for k0 in range(0, nk, k_chunk):
      k1 = min(nk, k0 + k_chunk)
      box = _pw.expand_coefficient(q[s, k0:k1], mask)  # (kc,b,x,y,z)
      psi = sfft.ifftn(box, axes=(-3, -2, -1), workers=_pw._WORKERS) * scale
      rho[s] += np.einsum("kbxyz,kb->xyz", np.real(np.conj(psi) * psi),
                          occ[s, k0:k1])
  v, rho_rec, parts = potential_real(rho, prob)
  gm = prob.mask_g()  # (ng,3)
  gk2 = np.sum((gm[None] + prob.kpts[:, None, :])**2, axis=-1)  # (k,g)
  absq = np.real(np.conj(q) * q)  # (s,k,g,b)
  parts["kinetic"] = float(np.sum(
    np.einsum("kg,skgb->skb", gk2, absq) * occ) / 2)
  nl = None
  if prob.projectors is not None:  # energy_nonlocal (nloc.py:196-227)
    from . import pseudopotential as _ps
    nl = _ps.energy_nonlocal(q, prob.projectors, occ, vol, want_grad)
    parts["external_local"] = parts["external"]
    parts["external_nonlocal"] = nl[0]
    parts["external"] = parts["external"] + nl[0]
  parts["total"] = (parts["kinetic"] + parts["hartree"] + parts["external"] +
                    parts["xc"])
  if not want_grad:
    return parts, None, rho
  qbar = np.full_like(q, np.nan)
  if grad_kpts is None:
    blocks = [np.arange(k0, min(nk, k0 + k_chunk))
              for k0 in range(0, nk, k_chunk)]
  else:
    ks = np.unique(np.asarray(grad_kpts, int))
    blocks = [ks[i:i + k_chunk] for i in range(0, len(ks), k_chunk)]
  for s in range(ns):
    for kb in blocks:
      box = _pw.expand_coefficient(q[s, kb], mask)
      psi = sfft.ifftn(box, axes=(-3, -2, -1), workers=_pw._WORKERS) * scale
      hv = sfft.fftn(v[s] * psi, axes=(-3, -2, -1), workers=_pw._WORKERS)
      hv = _pw.squeeze_coefficient(hv, mask)  # (kc,g,b)
      f = occ[s, kb][:, None, :]
      qbar[s, kb] = f * (2.0 / np.sqrt(vol) * hv + gk2[kb, :, None] * q[s, kb])
      if nl is not None:
        qbar[s, kb] += nl[1][s, kb]
  wbar = np.full_like(q, np.nan)
  kall = np.concatenate(blocks)
  wbar[:, kall] = qr_backward(q[:, kall], r[:, kall], qbar[:, kall])
  grad = {"w_re": np.real(wbar), "w_im": np.imag(wbar)}
  if extras:
    return parts, grad, rho, {"q": q, "qbar": qbar}
  return parts, grad, rho


def band_energies(params, prob: Problem, k_chunk: Optional[int] = None):
  """dE/df[s,k,b] for fixed W: 1/2 sum_G |G+k|^2 |Q_Gb|^2 (kinetic) plus
  sum_r v_s(r) |psi_b(r)|^2 with v = dE/drho of potential_real (Hartree,
  external and xc, the Omega/N weight included), for every band whether
  occupied or not.  Real-space form: independent of the reciprocal-space
  gather that the gradient uses."""
  w = params["w_re"] + 1j * params["w_im"]
  q = np.linalg.qr(w, mode="reduced")[0]
  ns, nk, ng, nb = q.shape
  mask = prob.mask
  n = int(np.prod(mask.shape))
  scale = n / np.sqrt(prob.vol)
  k_chunk = k_chunk or max(1, min(nk, 8))
  occ = prob.occupation
  rho = np.zeros((ns,) + mask.shape)
  psis = []
  for s in range(ns):
    for k0 in range(0, nk, k_chunk):
      k1 = min(nk, k0 + k_chunk)
      box = _pw.expand_coefficient(q[s, k0:k1], mask)
      psi = sfft.ifftn(box, axes=(-3, -2, -1), workers=_pw._WORKERS) * scale
      a2 = np.real(np.conj(psi) * psi)
      rho[s] += np.einsum("kbxyz,kb->xyz", a2, occ[s, k0:k1])
      psis.append((s, k0, k1, a2))
  v = potential_real(rho, prob)[0]
  out = np.empty((ns, nk, nb))
  for s, k0, k1, a2 in psis:
    out[s, k0:k1] = np.einsum("xyz,kbxyz->kb", v[s], a2)
  gm = prob.mask_g()
  gk2 = np.sum((gm[None] + prob.kpts[:, None, :])**2, axis=-1)
  out += 0.5 * np.einsum("kg,skgb->skb", gk2, np.real(np.conj(q) * q))
  if prob.projectors is not None:
    from . import pseudopotential as _ps
    out += _ps.energy_nonlocal(q, prob.projectors, occ, prob.vol, False)[2]
  return out


def total_energy_literal(params, prob: Problem):
  """The driver closure calc_ground_state_energy_all_electrons.py:119-137
  composed from the reference's public functions (full-box layout)."""
  c = _pw.coeff(params, prob.mask)
  occ = prob.occupation
  rho = _pw.density_grid(c, prob.vol, occ)
  rho_rec = _pw.density_grid_reciprocal(c, prob.vol, occ)
  g = prob.g_vector_grid
  kin = _en.kinetic(g, prob.kpts, c, occ)
  har = _en.hartree(rho_rec, g, prob.vol)
  ext = _en.external(rho_rec, prob.positions, prob.charges, g, prob.vol)
  xc = _en.xc_energy(rho, g, prob.vol, prob.xc)
  return {"kinetic": kin, "hartree": har, "external": ext, "xc": xc,
          "total": kin + har + ext + xc}
