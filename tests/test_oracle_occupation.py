"""Parameterised occupations (simplex projector, idempotent) and the
Fermi-Dirac entropy: the oracle against independent mathematics (Euclidean
projection by bisection, the projector diagonal W (W^T W)^-1 W^T, central
differences), and the HIP projection (values and autograd gradients,
tests marked gpu) against the oracle."""
import numpy as np
import pytest
import torch

from oracle import occupation as OO


def _bisect_projection(x, total):
  """argmin ||y - x|| over 0 <= y <= 1, sum y = total: y = clip(x - lam)."""
  lo, hi = x.min() - 2.0, x.max() + 2.0
  for _ in range(200):
    mid = 0.5 * (lo + hi)
    if np.clip(x - mid, 0.0, 1.0).sum() > total:
      lo = mid
    else:
      hi = mid
  return np.clip(x - 0.5 * (lo + hi), 0.0, 1.0)


@pytest.mark.parametrize("n,frac", [(2, 0.3), (7, 0.2), (7, 0.9), (48, 0.1),
                                     (48, 0.5), (48, 0.97), (300, 0.42)])
def test_oracle_proj_is_the_euclidean_projection(n, frac):
  rng = np.random.default_rng(n)
  x = rng.random(n)
  total = frac * n
  y = OO.proj(x, total)[0]
  assert abs(y.sum() - total) < 1e-12 * n
  assert y.min() >= 0.0 and y.max() <= 1.0
  assert np.max(np.abs(y - _bisect_projection(x, total))) < 1e-12


@pytest.mark.parametrize("n,frac", [(7, 0.2), (7, 0.9), (48, 0.5),
                                     (48, 0.05), (48, 0.95)])
def test_oracle_proj_vjp_matches_finite_differences(n, frac):
  rng = np.random.default_rng(100 + n)
  x = rng.random(n)
  total = frac * n
  ybar = rng.standard_normal(n)
  g = OO.proj_vjp(x, total, ybar)
  h = 1e-7
  fd = np.array([(ybar @ OO.proj(x + h * e, total)[0] -
                  ybar @ OO.proj(x - h * e, total)[0]) / (2 * h)
                 for e in np.eye(n)])
  assert np.max(np.abs(g - fd)) < 1e-6 * max(1.0, np.max(np.abs(fd)))


def test_simplex_projector_init_matches_reference_formula():
  from jrystal_amd import occupation as O
  p = O.simplex_projector_init(5, 3, device="cpu")
  ref = OO.simplex_projector_init(5, 3)
  for k in ("param_up", "param_down"):
    assert np.array_equal(p[k].numpy(), ref[k])
  # (arange(15) - 7) * 0.1 reshaped (k, band)
  assert ref["param_up"][0, 0] == pytest.approx(-0.7)
  assert ref["param_up"][2, 4] == pytest.approx(0.7)


@pytest.mark.gpu
@pytest.mark.parametrize("restricted", [True, False])
@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("nk,nb", [(4, 9), (64, 32), (64, 128)])
def test_simplex_projector_hip_matches_oracle(restricted, spin, nk, nb):
  """The HIP projection (sort, cumulative sums, shift rule, VJP) against
  the oracle, up to 8192 entries per channel."""
  from jrystal_amd import occupation as O
  rng = np.random.default_rng(7 + spin + nb)
  ne = 10
  p = {"param_up": rng.standard_normal((nk, nb)),
       "param_down": rng.standard_normal((nk, nb))}
  ref = OO.simplex_projector(p, ne, spin, restricted)
  pt = {k: torch.tensor(v, device="cuda", requires_grad=True)
        for k, v in p.items()}
  occ = O.simplex_projector(pt, ne, spin, restricted)
  assert np.max(np.abs(occ.detach().cpu().numpy() - ref)) < 1e-14
  # electron count: (ne +- spin) // 2 per channel, summed over k with 1/nk
  assert occ.sum().item() == pytest.approx((ne + spin) // 2 + (ne - spin) // 2)
  obar = rng.standard_normal(ref.shape)
  (occ * torch.tensor(obar, device="cuda")).sum().backward()
  vjp = OO.simplex_projector_vjp(p, ne, obar, spin, restricted)
  for k in p:
    assert np.max(np.abs(pt[k].grad.cpu().numpy() - vjp[k])) < 1e-12


def test_idempotent_oracle_is_the_projector_diagonal():
  rng = np.random.default_rng(3)
  nk, nb, ne = 2, 6, 6
  w = rng.random((nb * nk, ne // 2 * nk))
  p = {"param_up": {"w_re": w}, "param_down": {"w_re": w.copy()}}
  occ = OO.idempotent(p, nk)
  P = w @ np.linalg.solve(w.T @ w, w.T)
  assert np.max(np.abs(occ[0] - 2 * np.diag(P).reshape(nk, nb) / nk)) < 1e-13
  assert occ.sum() == pytest.approx(ne)


@pytest.mark.parametrize("restricted", [True, False])
def test_idempotent_torch_matches_oracle_and_fd(restricted):
  from jrystal_amd import occupation as O
  nk, nb, ne, spin = 2, 5, 4, 0 if restricted else 2
  p = O.idempotent_param_init(11, nb, ne, nk, spin, restricted, device="cpu")
  pn = {ch: {"w_re": v["w_re"].numpy().copy()} for ch, v in p.items()}
  ref = OO.idempotent(pn, nk, restricted)
  pt = {ch: {"w_re": v["w_re"].clone().requires_grad_()} for ch, v in p.items()}
  occ = O.idempotent(pt, nk, restricted)
  assert np.max(np.abs(occ.detach().numpy() - ref)) < 1e-14
  rng = np.random.default_rng(5)
  obar = rng.standard_normal(ref.shape)
  (occ * torch.tensor(obar)).sum().backward()
  h = 1e-6
  for ch in pn:
    w = pn[ch]["w_re"]
    g = pt[ch]["w_re"].grad.numpy()
    for (i, j) in [(0, 0), (3, 1), (w.shape[0] - 1, w.shape[1] - 1)]:
      def f(d):
        q = {c: {"w_re": v["w_re"].copy()} for c, v in pn.items()}
        q[ch]["w_re"][i, j] += d
        return np.sum(obar * OO.idempotent(q, nk, restricted))
      assert abs(g[i, j] - (f(h) - f(-h)) / (2 * h)) < 1e-7


def test_entropy_torch_value_and_gradient():
  from jrystal_amd import entropy
  rng = np.random.default_rng(9)
  occ = rng.uniform(0.01, 0.99, (1, 4, 6)) * 2 / 4
  t = torch.tensor(occ, requires_grad=True)
  s = entropy.fermi_dirac(t)
  assert s.item() == pytest.approx(OO.fermi_dirac_entropy(occ), rel=1e-14)
  assert entropy.fermi_dirac(occ) == pytest.approx(s.item(), rel=1e-14)
  s.backward()
  assert np.max(np.abs(t.grad.numpy() - OO.fermi_dirac_entropy_grad(occ))) < 1e-12
  h = 1e-7
  e = np.zeros_like(occ)
  e[0, 1, 2] = h
  fd = (OO.fermi_dirac_entropy(occ + e) - OO.fermi_dirac_entropy(occ - e)) / (2 * h)
  assert abs(OO.fermi_dirac_entropy_grad(occ)[0, 1, 2] - fd) < 1e-6


@pytest.mark.parametrize("x,total,expect", [
  ([0.3], 1.0, [1.0]),
  ([0.2, 0.2, 0.2], 1.5, [0.5, 0.5, 0.5]),
  # total = n (no empty bands): the reference's shift rule divides 0/0 at
  # the last sorted entry, never selects k = n and falls back to the
  # uniform shift, so the sum stays below the target (reproduced quirk)
  ([0.9, 0.1, 0.5], 3.0, [1.0, 0.6, 1.0]),
  # total = 0: likewise the uniform shift, sum above the target
  ([0.9, 0.1, 0.5], 0.0, [0.4, 0.0, 0.0]),
])
@pytest.mark.gpu
def test_proj_edge_cases_follow_the_reference_rule(x, total, expect):
  from jrystal_amd import occupation as O
  a = OO.proj(np.array(x), total)[0]
  b = O.proj(torch.tensor(x, dtype=torch.float64, device="cuda"),
             total).cpu().numpy()
  np.testing.assert_allclose(a, expect, atol=1e-15)
  np.testing.assert_allclose(b, expect, atol=1e-15)
