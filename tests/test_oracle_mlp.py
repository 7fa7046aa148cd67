"""The oracle's network restatement: hand-written backward against torch
CPU autograd and finite differences, loss gradients against finite
differences, and the Grassmann direction against its defining property."""
import numpy as np
import pytest

from oracle import mlp as OM


def _net(seed=0, scale=0.3):
  rng = np.random.default_rng(seed)
  p = rng.standard_normal(OM.NPARAMS) * scale
  norm = np.concatenate([rng.standard_normal(8) * 0.1,
                         rng.uniform(0.5, 2.0, 8)])
  x = rng.standard_normal((37, 8))
  return rng, p, norm, x


def test_backward_matches_torch_autograd():
  torch = pytest.importorskip("torch")
  rng, p, norm, x = _net()
  dy = rng.standard_normal((37, 2))
  g = OM.backward(p, norm, x, dy)
  pt = torch.tensor(p, requires_grad=True)
  w1 = pt[OM.W1:OM.B1].reshape(64, 8)
  b1 = pt[OM.B1:OM.W2]
  w2 = pt[OM.W2:OM.B2].reshape(64, 64)
  b2 = pt[OM.B2:OM.W3]
  w3 = pt[OM.W3:OM.B3].reshape(2, 64)
  b3 = pt[OM.B3:]
  xn = (torch.tensor(x) - torch.tensor(norm[:8])) * torch.tensor(norm[8:])
  y = torch.tanh(torch.tanh(xn @ w1.T + b1) @ w2.T + b2) @ w3.T + b3
  (y * torch.tensor(dy)).sum().backward()
  np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-12, atol=1e-14)
  np.testing.assert_allclose(OM.forward(p, norm, x), y.detach().numpy(),
                             rtol=1e-13, atol=1e-14)


def test_backward_central_differences():
  rng, p, norm, x = _net(1)
  dy = rng.standard_normal((37, 2))
  g = OM.backward(p, norm, x, dy)
  for i in rng.choice(OM.NPARAMS, 25, replace=False):
    e = np.zeros(OM.NPARAMS)
    e[i] = 1e-6
    fd = (np.sum(OM.forward(p + e, norm, x) * dy) -
          np.sum(OM.forward(p - e, norm, x) * dy)) / 2e-6
    assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_cosine_loss_gradient_and_bounds():
  rng = np.random.default_rng(2)
  ns, m = 3, 50
  y = rng.standard_normal((ns * m, 2))
  t = rng.standard_normal(ns * m) + 1j * rng.standard_normal(ns * m)
  q = rng.standard_normal(ns * m) + 1j * rng.standard_normal(ns * m)
  a = np.real(np.conj(t) * q)
  b = np.abs(q)**2
  t2 = np.sum((np.abs(t)**2).reshape(ns, m), axis=1)
  loss, dy = OM.cosine_loss(y, a, b, t2, m)
  assert np.all((loss >= 0) & (loss <= 1))
  f = lambda yy: np.mean(OM.cosine_loss(yy, a, b, t2, m)[0])  # noqa
  for i in rng.choice(ns * m, 10, replace=False):
    e = np.zeros_like(y)
    e[i, 0] = 1e-6
    fd = (f(y + e) - f(y - e)) / 2e-6
    assert abs(fd - dy[i, 0]) < 1e-8, (i, fd, dy[i, 0])
  assert np.all(dy[:, 1] == 0)
  # a factor proportional to t / q (per sample) aligns perfectly
  p_opt = np.abs(t) / np.abs(q)
  s = np.real(np.conj(t) * q) / (np.abs(t) * np.abs(q))
  y2 = y.copy()
  y2[:, 0] = np.log(p_opt)
  a2 = np.abs(t) * np.abs(q) * np.sign(s)
  l2, _ = OM.cosine_loss(y2, np.abs(a2), b, t2, m)
  np.testing.assert_allclose(l2, 0.0, atol=1e-12)


def test_mse_loss_gradient():
  rng = np.random.default_rng(3)
  y = rng.standard_normal((20, 2))
  t = rng.standard_normal(20)
  loss, dy = OM.mse_loss(y, t)
  assert abs(loss - np.mean((y[:, 1] - t)**2)) < 1e-15
  e = np.zeros_like(y)
  e[4, 1] = 1e-6
  fd = (OM.mse_loss(y + e, t)[0] - OM.mse_loss(y - e, t)[0]) / 2e-6
  assert abs(fd - dy[4, 1]) < 1e-9 and np.all(dy[:, 0] == 0)


def test_grassmann_direction_reaches_target_span():
  rng = np.random.default_rng(4)
  a = rng.standard_normal((3, 40, 5)) + 1j * rng.standard_normal((3, 40, 5))
  qt = np.linalg.qr(a)[0]
  qs = np.linalg.qr(qt + 0.3 * (rng.standard_normal(a.shape) +
                                1j * rng.standard_normal(a.shape)))[0]
  d = OM.grassmann_direction(qt, qs)
  qh = np.conj(np.swapaxes(qt, -1, -2))
  assert np.max(np.abs(qh @ d)) < 1e-12  # tangent: orthogonal to span(qt)
  q1 = np.linalg.qr(qt + d)[0]
  p1 = q1 @ np.conj(np.swapaxes(q1, -1, -2))
  ps = qs @ np.conj(np.swapaxes(qs, -1, -2))
  assert np.max(np.abs(p1 - ps)) < 1e-12


def test_features_closed_form():
  f = OM.features(np.array([0.0, 3.0]), np.array([2.0, 2.0]), 0.5, 0.25)
  np.testing.assert_allclose(f[1, :5], [np.log(4.0), np.log(2.5), np.log(2.0),
                                        np.log(0.5), np.log(2.0)], rtol=1e-15)
  assert np.all(f[:, 5:] == 0) and f[0, 0] == 0 and f[0, 1] == 0
  g = OM.features(np.array([1.0]), np.array([1.0]), 1.0, 0.0, gap=np.e - 1)
  assert abs(g[0, 5] - 1.0) < 1e-15
