"""Kohn-Sham band mode of the all-electron band-structure driver.

TEST INFRASTRUCTURE ONLY (the rules of oracle/__init__.py apply).

For a fixed ground-state density rho0, restates:

* ``effective_potential`` -- potential.effective(..., kohn_sham=True)
  (potential.py:203-279): V_H with the full 4 pi rho(G) / |G|^2
  (potential.py:67-75), V_ext (potential.py:121-166) and the xc potential
  of xc_density(kohn_sham=True) (xc.py:222-253); for LDA exchange that is
  vxc_lda (xc.py:109-124), eps + rho_s d eps / d rho_s per spin channel.
  Returned with the vol/N weight of oracle.adjoint.potential_real, so that
  sum_r v |psi|^2 is the integral of v_eff |psi|^2.
* ``hamiltonian_trace`` -- hamiltonian_matrix_trace (hamiltonian.py:
  105-168), sum_{s,k,b} f (1/2 |G+k|^2 |q|^2 + sum_r v |psi|^2) with
  Q = QR(W) (pw.coeff), and the gradient that jax.value_and_grad takes of
  it (calc_band_structure_all_electrons.py:75-108), through the QR adjoint
  of oracle.adjoint.
* ``subspace_hamiltonian`` / ``band_eigenvalues`` -- hamiltonian_matrix
  (hamiltonian.py:171-240) as Q^H H Q and its eigvalsh
  (calc_band_structure_all_electrons.py:156-169).
* ``dense_hamiltonian`` -- the textbook plane-wave matrix over the mask,
  H_GG' = 1/2 |G+k|^2 delta_GG' + v~(G - G'), v~ = fftn(v_eff) / N indexed
  by integer frequency differences modulo the grid: built without the FFT
  application above, for brute-force checks.
* With norm-conserving pseudopotentials (prob.v_ext_rec, prob.projectors)
  the local part enters v_eff and the non-local term
  (1/vol) sum_p conj(P_p(G)) P_p(G') the Hamiltonian, as in the
  reference's normcons.hamiltonian_trace (pseudopotential/nloc.py:230-279).

xc terms other than LDA exchange use the functional derivative of the
discrete E_xc (oracle.xc.xc_potential).  For LDA correlation that is the
reference's vxc_lda; for GGA terms the reference's vxc_gga (xc.py:127-219)
is a different formula (its grad_fft transforms one axis forward and the
(y, z, component) axes back), so GGA band energies are parity unpinned
against the reference and pinned here only by the checks of
tests/test_oracle_band.py.
"""
from typing import Optional

import numpy as np
import scipy.fft as sfft

from . import adjoint as _adj
from . import energy as _en
from . import pw as _pw


def effective_potential(rho, prob: _adj.Problem):
  """v_s(r) * vol / N, (s,x,y,z), of the ground-state density rho (s,x,y,z)."""
  rho = np.asarray(rho, np.float64)
  n = np.prod(rho.shape[-3:])
  rho_rec = sfft.fftn(rho, axes=(-3, -2, -1), workers=_pw._WORKERS)
  v_he = _adj.hartree_external_potential(rho_rec, prob)
  if prob.xc == "lda_x":
    v_xc = (prob.vol / n) * _en.vxc_lda_x(rho)
  else:
    from . import xc as _xc
    v_xc = _xc.xc_potential(rho, prob.g_vector_grid, prob.vol, prob.xc)
  return v_he[None] + v_xc


def _apply(q, v, prob: _adj.Problem, k_chunk: int = 8):
  """(H q pieces) for q (s,k,g,b): returns (gk2 (k,g), hv (s,k,g,b) =
  mask(fftn(v psi)), pot (s,k,b) = sum_r v |psi|^2)."""
  ns, nk, ng, nb = q.shape
  mask = prob.mask
  n = int(np.prod(mask.shape))
  scale = n / np.sqrt(prob.vol)
  hv = np.empty_like(q)
  pot = np.empty((ns, nk, nb))
  for s in range(ns):
    for k0 in range(0, nk, k_chunk):
      k1 = min(nk, k0 + k_chunk)
      box = _pw.expand_coefficient(q[s, k0:k1], mask)
      psi = sfft.ifftn(box, axes=(-3, -2, -1), workers=_pw._WORKERS) * scale
      pot[s, k0:k1] = np.einsum("xyz,kbxyz->kb", v[s],
                                np.real(np.conj(psi) * psi))
      h = sfft.fftn(v[s] * psi, axes=(-3, -2, -1), workers=_pw._WORKERS)
      hv[s, k0:k1] = _pw.squeeze_coefficient(h, mask)
  gm = prob.mask_g()
  gk2 = np.sum((gm[None] + prob.kpts[:, None, :])**2, axis=-1)
  return gk2, hv, pot


def hamiltonian_trace(params, prob: _adj.Problem, v, want_grad: bool = True,
                      extras: bool = False):
  """(parts {'kinetic', 'potential', 'total'}, grad dict or None) of
  E = sum f <psi|T + v|psi>, f = prob.occupation; with extras a third item
  {'q', 'qbar', 'hq'} ((s,k,g,b); hq = H q)."""
  w = params["w_re"] + 1j * params["w_im"]
  q, r = np.linalg.qr(w, mode="reduced")
  occ = prob.occupation
  gk2, hv, pot = _apply(q, v, prob)
  absq = np.real(np.conj(q) * q)
  kin = 0.5 * np.einsum("kg,skgb->skb", gk2, absq)
  parts = {"kinetic": float(np.sum(occ * kin)),
           "potential": float(np.sum(occ * pot)), "nonlocal": 0.0}
  grad = None
  hq = 0.5 * gk2[None, :, :, None] * q + hv / np.sqrt(prob.vol)
  if prob.projectors is not None:  # nloc.py:230-279 (hamiltonian_trace)
    P = prob.projectors
    F = np.einsum("kpg,skgb->skpb", P, q)
    parts["nonlocal"] = float(np.sum(occ * np.sum(np.abs(F)**2, axis=2)) /
                              prob.vol)
    hq = hq + np.einsum("skpb,kpg->skgb", F, np.conj(P)) / prob.vol
  parts["total"] = parts["kinetic"] + parts["potential"] + parts["nonlocal"]
  qbar = 2.0 * occ[:, :, None, :] * hq
  if want_grad:
    wbar = _adj.qr_backward(q, r, qbar)
    grad = {"w_re": np.real(wbar), "w_im": np.imag(wbar)}
  if extras:
    return parts, grad, {"q": q, "qbar": qbar, "hq": hq}
  return parts, grad


def subspace_hamiltonian(params, prob: _adj.Problem, v):
  """q_i^H H q_j, (s,k,b,b) complex."""
  x = hamiltonian_trace(params, prob, v, want_grad=False, extras=True)[2]
  return np.conj(np.swapaxes(x["q"], -1, -2)) @ x["hq"]


def band_eigenvalues(params, prob: _adj.Problem, v):
  """Ascending eigenvalues of q_i^H H q_j, (s,k,b)."""
  return np.linalg.eigvalsh(subspace_hamiltonian(params, prob, v))


def dense_hamiltonian(prob: _adj.Problem, v, s: int = 0, k: int = 0,
                      kpt: Optional[np.ndarray] = None):
  """H_GG' over the mask points (C order), (ng, ng) complex, for spin s and
  k-point k of prob (or the explicit kpt)."""
  mask = prob.mask
  shape = np.array(mask.shape)
  n = int(np.prod(shape))
  vt = np.fft.fftn(v[s]) / n * (n / prob.vol)  # fftn(v_eff) / N
  m = np.argwhere(mask)  # FFT index of each mask point, C order
  d = (m[:, None, :] - m[None, :, :]) % shape
  h = vt[d[..., 0], d[..., 1], d[..., 2]].astype(complex)
  kk = prob.kpts[k] if kpt is None else np.asarray(kpt, float)
  gk2 = np.sum((prob.mask_g() + kk)**2, axis=-1)
  h[np.diag_indices_from(h)] += 0.5 * gk2
  if prob.projectors is not None and kpt is None:
    P = prob.projectors[k]
    h += np.conj(P).T @ P / prob.vol
  return h
