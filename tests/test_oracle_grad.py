"""The oracle's closed-form gradient against an independent AD system (torch
CPU autograd), central finite differences and the QR invariances."""
import numpy as np
import pytest

from oracle import adjoint as OA
from oracle import energy as OE
from oracle.workloads import cfg1, tiny


def _torch_energy(params, prob):
  """Energy built from torch ops with autograd (complex128)."""
  torch = pytest.importorskip("torch")
  wr = torch.tensor(params["w_re"], requires_grad=True)
  wi = torch.tensor(params["w_im"], requires_grad=True)
  q, _ = torch.linalg.qr(torch.complex(wr, wi))
  mask = torch.tensor(prob.mask)
  ns, nk, ng, nb = q.shape
  box = torch.zeros((ns, nk, nb) + prob.mask.shape, dtype=torch.complex128)
  box[..., mask] = q.transpose(-1, -2)
  n = int(np.prod(prob.mask.shape))
  vol = prob.vol
  psi = torch.fft.ifftn(box, dim=(-3, -2, -1)) * n / np.sqrt(vol)
  occ = torch.tensor(prob.occupation)
  rho = torch.einsum("skbxyz,skb->sxyz", (psi.conj() * psi).real, occ)
  rr = torch.fft.fftn(rho, dim=(-3, -2, -1))
  g = torch.tensor(prob.g_vector_grid)
  g2 = (g**2).sum(-1)
  g2h = g2.clone()
  g2h[0, 0, 0] = 1
  vh = rr.sum(0) / g2h
  vh[0, 0, 0] = 0
  eh = (2 * np.pi * vh.conj()[None] * rr).sum().real / n * vol / n
  si = torch.exp(-1j * (g @ torch.tensor(prob.positions).T))
  vi = torch.tensor(prob.charges)[None, None, None, :] / (g2[..., None] + 1e-10)
  vi[0, 0, 0] = 0
  ve = -(4 * np.pi * vi * si).sum(-1) * n / vol
  ee = (ve.conj()[None] * rr).sum().real / n * vol / n
  ex = (-OE.CX * rho[0]**(1 / 3) * rho[0]).sum() * vol / n
  kp = torch.tensor(prob.kpts)
  gk2 = ((g[None] + kp[:, None, None, None, :])**2).sum(-1)
  ek = (gk2[None, :, None] * (box.conj() * box).real *
        occ[..., None, None, None]).sum() / 2
  e = ek + eh + ee + ex
  e.backward()
  return e.item(), wr.grad.numpy(), wi.grad.numpy()


def test_gradient_matches_torch_autograd():
  cr, prob, params = tiny()
  parts, grad, _ = OA.energy_and_grad(params, prob)
  e, gr, gi = _torch_energy(params, prob)
  np.testing.assert_allclose(parts["total"], e, rtol=1e-13)
  assert np.max(np.abs(grad["w_re"] - gr)) < 1e-12 * np.max(np.abs(gr))
  assert np.max(np.abs(grad["w_im"] - gi)) < 1e-12 * np.max(np.abs(gi))


def test_gradient_central_differences():
  """Central FD (h = 1e-6) on random coordinates; rounding noise of the
  difference is ~ eps |E| / h ~ 1e-8, so the tolerance is absolute 1e-6."""
  cr, prob, params = tiny(z=14.0)
  parts, grad, _ = OA.energy_and_grad(params, prob)
  rng = np.random.default_rng(1)
  h = 1e-6
  for _ in range(8):
    key = ("w_re", "w_im")[rng.integers(2)]
    idx = tuple(rng.integers(s) for s in params[key].shape)
    vals = []
    for sgn in (1, -1):
      p = {k: v.copy() for k, v in params.items()}
      p[key][idx] += sgn * h
      vals.append(OA.energy_and_grad(p, prob, want_grad=False)[0]["total"])
    fd = (vals[0] - vals[1]) / (2 * h)
    assert abs(fd - grad[key][idx]) < 1e-6, (key, idx, fd, grad[key][idx])


def test_triu_invariant_of_qr_gradient():
  """E(W T) = E(W) for upper-triangular T implies triu(W^H Wbar) = 0,
  diagonal included (Wbar = dE/dRe W + i dE/dIm W)."""
  cr, prob, params = cfg1(z=4.0)
  _, grad, _ = OA.energy_and_grad(params, prob)
  w = params["w_re"] + 1j * params["w_im"]
  wbar = grad["w_re"] + 1j * grad["w_im"]
  m = np.conj(np.swapaxes(w, -1, -2)) @ wbar
  assert np.max(np.abs(np.triu(m))) < 1e-11 * np.max(np.abs(m))


def test_energy_invariant_under_upper_triangular_mixing():
  cr, prob, params = tiny()
  w = params["w_re"] + 1j * params["w_im"]
  rng = np.random.default_rng(7)
  nb = w.shape[-1]
  t = np.triu(rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb)))
  t += 3 * np.eye(nb)
  w2 = w @ t
  e0 = OA.energy_and_grad(params, prob, want_grad=False)[0]["total"]
  e1 = OA.energy_and_grad({"w_re": w2.real, "w_im": w2.imag}, prob,
                          want_grad=False)[0]["total"]
  np.testing.assert_allclose(e1, e0, rtol=1e-13)


def test_chunked_adjoint_equals_reference_composition():
  """The k-chunked oracle equals the literal full-box composition of the
  reference's public functions (driver closure)."""
  cr, prob, params = cfg1(z=14.0)
  a, _, _ = OA.energy_and_grad(params, prob, k_chunk=3, want_grad=False)
  b = OA.total_energy_literal(params, prob)
  for k in b:
    np.testing.assert_allclose(a[k], b[k], rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("make", [tiny])
def test_band_energies_are_the_occupation_derivative(make):
  """oracle.band_energies (real-space <psi|H|psi>) against central
  differences of the oracle energy in each occupation number, including a
  band with zero occupation."""
  import dataclasses
  _, prob, params = make()
  rng = np.random.default_rng(21)
  ns, nk, nb = prob.occupation.shape
  f = rng.uniform(0.1, 2.0 / nk, (ns, nk, nb))
  f[0, 0, nb - 1] = 0.0
  pf = dataclasses.replace(prob, occupation=f)
  eb = OA.band_energies(params, pf)
  h = 1e-5
  for idx in [(0, 0, 0), (0, nk - 1, nb - 1), (0, 0, nb - 1)]:
    e = np.zeros_like(f)
    e[idx] = h
    ep = OA.energy_and_grad(params, dataclasses.replace(prob, occupation=f + e),
                            want_grad=False)[0]["total"]
    em = OA.energy_and_grad(params, dataclasses.replace(prob, occupation=f - e),
                            want_grad=False)[0]["total"]
    fd = (ep - em) / (2 * h)
    assert abs(eb[idx] - fd) < 1e-7 * max(1.0, abs(fd)), (idx, eb[idx], fd)


def test_band_energies_match_the_gather_form():
  """sum_b f_b e_b equals Re <Q, Qbar> / 2 (the reciprocal-space form the
  gradient uses: Qbar = f (2/sqrt(vol) FFT(v psi) + |G+k|^2 Q))."""
  _, prob, params = tiny()
  parts, _, _, ex = OA.energy_and_grad(params, prob, extras=True)
  eb = OA.band_energies(params, prob)
  lhs = np.sum(prob.occupation * eb)
  rhs = 0.5 * np.real(np.sum(np.conj(ex["q"]) * ex["qbar"]))
  assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(rhs))
