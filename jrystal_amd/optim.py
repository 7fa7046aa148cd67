"""Optimisers of the energy driver.

* ``Adam`` -- optax.adam semantics (opt_utils.py:153-168: lr, b1, b2 from
  the config; eps = 1e-8, eps_root = 0), one fused HIP kernel per tensor.
* ``exponential_decay`` -- optax.exponential_decay (temperature annealing
  of calc_ground_state_energy_all_electrons.py:184-194).
* ``ConvergenceChecker`` -- calc/convergence.py:13-35: converged when the
  std of the last ``window_size`` energies is below the threshold.
"""
import ctypes

import numpy as np
import torch

from . import _lib


class Adam:

  def __init__(self, params, learning_rate=0.01, b1=0.9, b2=0.999, eps=1e-8,
               eps_root=0.0, **unused):
    if unused:
      raise NotImplementedError(f"adam arguments {sorted(unused)}")
    self.lr, self.b1, self.b2 = float(learning_rate), float(b1), float(b2)
    self.eps, self.eps_root = float(eps), float(eps_root)
    self.m = {k: torch.zeros_like(v) for k, v in params.items()}
    self.v = {k: torch.zeros_like(v) for k, v in params.items()}
    self.count = 0

  def step(self, params, grads):
    """In-place update of params (dict of contiguous float64 tensors)."""
    self.count += 1
    L = _lib.lib()
    s = _lib.stream_ptr()
    for k, p in params.items():
      g = grads[k]
      _lib.check(L.jrx_adam_update(
        p.data_ptr(), g.data_ptr(), self.m[k].data_ptr(), self.v[k].data_ptr(),
        p.numel(), ctypes.c_double(self.lr), ctypes.c_double(self.b1),
        ctypes.c_double(self.b2), ctypes.c_double(self.eps),
        ctypes.c_double(self.eps_root), self.count, s))

  def state(self):
    return {"m": self.m, "v": self.v, "count": self.count}


def exponential_decay(init_value, transition_steps, decay_rate,
                      end_value=None):
  """optax.exponential_decay with transition_begin = 0, staircase = False.
  Like optax, a non-positive transition_steps or a zero decay_rate gives the
  constant schedule init_value (the driver passes epoch // 2, which is 0 for
  epoch < 2)."""
  if transition_steps <= 0 or decay_rate == 0:
    return lambda count: init_value

  def schedule(count):
    v = init_value * decay_rate**(count / transition_steps)
    if end_value is not None:
      v = max(v, end_value) if decay_rate < 1 else min(v, end_value)
    return v

  return schedule


class ConvergenceChecker:

  def __init__(self, window_size: int = 20, threshold: float = 1e-5):
    self.window_size = window_size
    self.threshold = threshold
    self.history = []

  def check(self, value: float) -> bool:
    self.history.append(value)
    if len(self.history) > self.window_size:
      self.history.pop(0)
    if len(self.history) < self.window_size:
      return False
    return bool(np.std(self.history) < self.threshold)

  def reset(self):
    self.history = []
