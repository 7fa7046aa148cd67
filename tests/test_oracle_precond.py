"""The oracle's kinetic (Teter-Payne-Allan) preconditioner, pinned by its
closed-form limits and by the identity E_kin = 1/2 sum occ * tb."""
import numpy as np

from oracle import adjoint as OA
from oracle import optim as OOpt
from oracle.workloads import tiny


def test_tpa_factor_limits():
  x = np.array([1e-3, 2e-3, 5e-3])
  # 1 - K = (16/27) x^4 + O(x^5): the first three Taylor terms vanish
  np.testing.assert_allclose((1 - OOpt.tpa_factor(x)) / x**4, 16 / 27,
                             rtol=2e-2)
  big = np.array([1e4, 1e5, 1e6])
  np.testing.assert_allclose(big * OOpt.tpa_factor(big), 0.5, rtol=2e-4)
  assert OOpt.tpa_factor(0.0) == 1.0
  xs = np.linspace(0, 50, 5001)
  k = OOpt.tpa_factor(xs)
  assert np.all(np.diff(k) < 0) and np.all(k > 0) and np.all(k <= 1)


def test_band_norms_give_kinetic_energy():
  _, prob, params = tiny(seed=3)
  tb = OOpt.band_norms(params, prob)
  parts, _, _ = OA.energy_and_grad(params, prob, want_grad=False)
  np.testing.assert_allclose(0.5 * np.sum(prob.occupation * tb),
                             parts["kinetic"], rtol=1e-13)
  assert np.all(tb > 0)


def test_band_norm_of_single_plane_wave():
  _, prob, params = tiny(seed=0)
  ns, nk, ng, nb = params["w_re"].shape
  w = np.zeros((ns, nk, ng, nb))
  idx = np.arange(nb) * 3 + 1
  w[:, :, idx, np.arange(nb)] = 1.0  # band b = plane wave G_idx[b]
  tb = OOpt.band_norms({"w_re": w, "w_im": np.zeros_like(w)}, prob)
  gm = prob.mask_g()
  gk2 = np.sum((gm[idx][None] + prob.kpts[:, None, :])**2, axis=-1)  # (k,b)
  np.testing.assert_allclose(tb, np.broadcast_to(gk2, tb.shape), rtol=1e-14)


def test_precondition_is_diagonal_and_factor_multiplies():
  _, prob, params = tiny(seed=1)
  rng = np.random.default_rng(7)
  g = {k: rng.standard_normal(v.shape) for k, v in params.items()}
  p1 = OOpt.tpa_precondition(g, params, prob)
  ratio = p1["w_re"] / g["w_re"]
  np.testing.assert_allclose(ratio, p1["w_im"] / g["w_im"], rtol=1e-14)
  assert np.all((ratio > 0) & (ratio <= 1))
  f = rng.uniform(0.5, 2.0, g["w_re"].shape)
  p2 = OOpt.tpa_precondition(g, params, prob, factor=f)
  np.testing.assert_allclose(p2["w_re"], f * p1["w_re"], rtol=1e-15)
