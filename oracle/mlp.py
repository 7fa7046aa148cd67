"""CPU restatement of the learned-preconditioner network and its losses
(test infrastructure: checks the HIP forward/backward/loss/Adam kernels).

The network has no counterpart in the reference; this file restates the
project's own definition so that the GPU kernels are pinned by:
  * finite differences of the loss (the backward below is hand-written);
  * torch CPU autograd (an independent AD system) in the tests.

Network: x (8 raw features) -> (x - mu) * inv_sd -> tanh(W1 x + b1) ->
tanh(W2 h1 + b2) -> W3 h2 + b3, flat parameter vector
[W1 (64x8), b1 (64), W2 (64x64), b2 (64), W3 (2x64), b3 (2)].
"""
import numpy as np

IN, HID, OUT = 8, 64, 2
W1, B1 = 0, HID * IN
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + OUT * HID
NPARAMS = B3 + OUT
LOG_FACTOR = (-8.0, 8.0)
LOG_AMP = (-30.0, 8.0)


def unpack(p):
  p = np.asarray(p, dtype=np.float64)
  return (p[W1:B1].reshape(HID, IN), p[B1:W2], p[W2:B2].reshape(HID, HID),
          p[B2:W3], p[W3:B3].reshape(OUT, HID), p[B3:NPARAMS])


def features(gk2, tb, rho_mean, rho_var, gap=0.0):
  """Raw features (..., 8): log1p(|G+k|^2), log1p(|G+k|^2/tb), log tb,
  log rho_mean, log1p(rho_var / rho_mean^2), log1p(gap), 0, 0."""
  gk2, tb, gap = np.broadcast_arrays(np.asarray(gk2, float),
                                     np.asarray(tb, float),
                                     np.asarray(gap, float))
  t = np.maximum(tb, 1e-300)
  m = max(float(rho_mean), 1e-300)
  out = np.zeros(gk2.shape + (IN,))
  out[..., 0] = np.log1p(gk2)
  out[..., 1] = np.log1p(gk2 / t)
  out[..., 2] = np.log(t)
  out[..., 3] = np.log(m)
  out[..., 4] = np.log1p(max(float(rho_var), 0.0) / (m * m))
  out[..., 5] = np.log1p(np.maximum(gap, 0.0))
  return out


def band_gaps(subspace, occupation):
  """Per band (s,k,b): Rayleigh quotient of the highest occupied band of
  its (s,k) block minus its own, from Q^H Qbar (diagonal 2 f eps_b); 0 for
  empty bands."""
  occ = np.asarray(occupation, float)
  d = np.real(np.diagonal(subspace, axis1=-2, axis2=-1))
  eps = np.where(occ > 0, d / (2 * np.where(occ > 0, occ, 1.0)), -np.inf)
  top = np.max(eps, axis=-1, keepdims=True)
  return np.where(occ > 0, top - eps, 0.0)


def forward(p, norm, x, keep=False):
  """y (n, 2); with keep, also the activations for backward."""
  w1, b1, w2, b2, w3, b3 = unpack(p)
  norm = np.asarray(norm, float)
  xn = (np.asarray(x, float) - norm[:IN]) * norm[IN:]
  h1 = np.tanh(xn @ w1.T + b1)
  h2 = np.tanh(h1 @ w2.T + b2)
  y = h2 @ w3.T + b3
  if keep:
    return y, (xn, h1, h2)
  return y


def backward(p, norm, x, dy):
  """Gradient of sum_i dy_i . y_i with respect to the flat parameters."""
  w1, b1, w2, b2, w3, b3 = unpack(p)
  _, (xn, h1, h2) = forward(p, norm, x, keep=True)
  dy = np.asarray(dy, float)
  g = np.zeros(NPARAMS)
  g[W3:B3] = (dy.T @ h2).ravel()
  g[B3:] = dy.sum(0)
  d2 = (dy @ w3) * (1 - h2**2)
  g[W2:B2] = (d2.T @ h1).ravel()
  g[B2:W3] = d2.sum(0)
  d1 = (d2 @ w2) * (1 - h1**2)
  g[W1:B1] = (d1.T @ xn).ravel()
  g[B1:W2] = d1.sum(0)
  return g


def cosine_loss(y, a, b, t2, snap_size):
  """Per-snapshot loss 1 - (sum p a)^2 / (t2 sum p^2 b), p = exp(clamp y0),
  and dy of the mean loss (output 1 gets zero)."""
  y = np.asarray(y, float)
  ns = len(t2)
  y0 = y[:, 0].reshape(ns, snap_size)
  p = np.exp(np.clip(y0, *LOG_FACTOR))
  a = np.asarray(a, float).reshape(ns, snap_size)
  b = np.asarray(b, float).reshape(ns, snap_size)
  t2 = np.asarray(t2, float)
  c = np.sum(p * a, axis=1)
  U = np.sum(p * p * b, axis=1)
  loss = 1.0 - c**2 / (t2 * U)
  dldp = -2.0 * (c / (t2 * U))[:, None] * (a - (c / U)[:, None] * p * b)
  inside = (y0 > LOG_FACTOR[0]) & (y0 < LOG_FACTOR[1])
  dy = np.zeros_like(y)
  dy[:, 0] = (np.where(inside, p * dldp, 0.0) / ns).ravel()
  return loss, dy


def mse_loss(y, target):
  y = np.asarray(y, float)
  r = y[:, 1] - np.asarray(target, float)
  dy = np.zeros_like(y)
  dy[:, 1] = 2 * r / len(r)
  return float(np.mean(r**2)), dy


def adam(p, g, m, v, count, lr, b1=0.9, b2=0.999, eps=1e-8):
  """optax.adam rule; returns (p, m, v)."""
  m = b1 * m + (1 - b1) * g
  v = b2 * v + (1 - b2) * g * g
  p = p - lr * (m / (1 - b1**count)) / (np.sqrt(v / (1 - b2**count)) + eps)
  return p, m, v


def grassmann_direction(q_t, q_star):
  """Tangent direction at span(q_t) towards span(q_star), per leading
  index: D = (I - q_t q_t^H) q_star (q_t^H q_star)^{-1}, so that
  span(q_t + D) = span(q_star) exactly.  q_* are (..., G, n) orthonormal."""
  qh = np.conj(np.swapaxes(q_t, -1, -2))
  m = qh @ q_star
  proj = q_star - q_t @ m
  return proj @ np.linalg.inv(m)
