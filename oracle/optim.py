"""Optimiser restatements (oracle side): optax.adam, exponential_decay and
the convergence window of the reference driver.

optax is a reference dependency (requirements.txt:3) absent here; its
published update rule is restated:
  m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2;  t += 1
  p -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t) + eps_root) + eps)
Step-level bitwise agreement with optax is parity unpinned (rounding only).
"""
import numpy as np


class Adam:

  def __init__(self, params, learning_rate=0.01, b1=0.9, b2=0.999, eps=1e-8,
               eps_root=0.0):
    self.lr, self.b1, self.b2, self.eps, self.eps_root = (
      learning_rate, b1, b2, eps, eps_root)
    self.m = {k: np.zeros_like(v) for k, v in params.items()}
    self.v = {k: np.zeros_like(v) for k, v in params.items()}
    self.t = 0

  def step(self, params, grads):
    self.t += 1
    c1 = 1 - self.b1**self.t
    c2 = 1 - self.b2**self.t
    out = {}
    for k, p in params.items():
      g = grads[k]
      self.m[k] = self.b1 * self.m[k] + (1 - self.b1) * g
      self.v[k] = self.b2 * self.v[k] + (1 - self.b2) * g * g
      out[k] = p - self.lr * (self.m[k] / c1) / (
        np.sqrt(self.v[k] / c2 + self.eps_root) + self.eps)
    return out


def exponential_decay(init_value, transition_steps, decay_rate,
                      end_value=None):
  """optax.exponential_decay, transition_begin = 0, no staircase; constant
  init_value when transition_steps <= 0 or decay_rate == 0, as optax."""
  if transition_steps <= 0 or decay_rate == 0:
    return lambda count: init_value

  def f(count):
    v = init_value * decay_rate**(count / transition_steps)
    if end_value is not None:
      v = max(v, end_value) if decay_rate < 1 else min(v, end_value)
    return v
  return f


def converged(history, window, threshold):
  """calc/convergence.py:20-32 applied to a full history."""
  if len(history) < window:
    return False
  return bool(np.std(history[-window:]) < threshold)


def tpa_factor(x):
  """Teter-Payne-Allan kinetic preconditioner (PRB 40, 12255 (1989)):
  K(x) = p / (p + 16 x^4), p = 27 + 18x + 12x^2 + 8x^3, with
  x = (kinetic energy of the plane wave) / (kinetic energy of the band).
  K = 1 - (16/27) x^4 + O(x^5) for small x and K ~ 1/(2x) for large x."""
  x = np.asarray(x, dtype=float)
  p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
  return p / (p + 16.0 * x**4)


def band_norms(params, prob):
  """tb[s,k,b] = sum_G |G+k|^2 |Q[s,k,G,b]|^2, Q from the reduced QR of
  W = w_re + i w_im: twice the kinetic energy of each band, so that
  E_kin = 1/2 sum occ * tb."""
  w = params["w_re"] + 1j * params["w_im"]
  q = np.linalg.qr(w, mode="reduced")[0]
  gm = prob.mask_g()
  gk2 = np.sum((gm[None] + prob.kpts[:, None, :])**2, axis=-1)  # (k,g)
  return np.einsum("kg,skgb->skb", gk2, np.real(np.conj(q) * q))


def tpa_precondition(grad, params, prob, factor=None):
  """Diagonal preconditioning of a gradient dict (w_re, w_im), each
  (s,k,g,b): P[s,k,g,b] = K(|G+k|^2 / tb[s,k,b]) * factor[s,k,g,b]."""
  tb = band_norms(params, prob)
  gm = prob.mask_g()
  gk2 = np.sum((gm[None] + prob.kpts[:, None, :])**2, axis=-1)
  kx = tpa_factor(gk2[None, :, :, None] / tb[:, :, None, :])
  if factor is not None:
    kx = kx * factor
  return {k: kx * v for k, v in grad.items()}
