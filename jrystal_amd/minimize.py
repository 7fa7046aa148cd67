"""Preconditioned direct minimisation of the total energy on the GPU.

The reference minimises with optax Adam on the raw parameters
(calc_ground_state_energy_all_electrons.py:175-220).  This module adds the
classic plane-wave alternative: preconditioned conjugate gradients on W
(Teter, Payne & Allan, PRB 40, 12255 (1989)) with

* search direction d = -P g + beta d_prev (Polak-Ribiere+, restarted every
  ``restart`` iterations or when d is not a descent direction);
* P = diag(K(|G+k|^2 / tb_band)) (the TPA kinetic preconditioner), times an
  optional per-coefficient factor -- the hook of the learned preconditioner;
* a line search that fits a parabola through E(0), E'(0) and one
  energy-only trial E(l_t), then evaluates energy and gradient at the
  parabola's minimum;
* acceptance only when the energy decreases; otherwise the trial point is
  taken if it decreased, else the step shrinks and the direction restarts
  from -P g (the fall-back of the learned preconditioner uses the same
  rule, with the plain preconditioner).

Every vector operation is a HIP kernel (jrx_vec_dot, jrx_vec_axpby,
jrx_plan_precondition); the only host traffic is a handful of scalars per
iteration.
"""
import ctypes
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _lib
from .objective import EnergyPlan


@dataclass
class MinimizeResult:
  energy: float
  iterations: int
  evaluations: int
  seconds: float
  history: List[tuple] = field(default_factory=list)  # (seconds, energy)
  rejected: int = 0
  fallbacks: int = 0


class Minimizer:

  def __init__(self, plan: EnergyPlan, params, factor_fn: Optional[Callable]
               = None, restart: int = 60, step0: float = 0.5):
    self.plan = plan
    dev = plan.device
    shape = (2,) + tuple(plan.param_shape)
    self.n = int(np.prod(shape))
    self.W = torch.empty(shape, dtype=torch.float64, device=dev)
    self.W[0].copy_(torch.as_tensor(params["w_re"]))
    self.W[1].copy_(torch.as_tensor(params["w_im"]))
    mk = lambda: torch.empty(shape, dtype=torch.float64, device=dev)  # noqa
    self.g, self.h, self.d = mk(), mk(), mk()
    self.W1, self.g1, self.h1 = mk(), mk(), mk()
    self.e = torch.empty(5, dtype=torch.float64, device=dev)
    self.e1 = torch.empty(5, dtype=torch.float64, device=dev)
    self.dotv = torch.empty(4, dtype=torch.float64, device=dev)
    ws = _lib.lib().jrx_dot_workspace_bytes()
    self._dws = torch.empty(ws + 256, dtype=torch.uint8, device=dev)
    self._dws_ptr = (self._dws.data_ptr() + 255) // 256 * 256
    self.factor_fn = factor_fn
    self.restart = restart
    self.step = step0
    self.evals = 0

  # -- kernels -------------------------------------------------------------
  def _eval(self, W, e, g=None):
    out = {"energies": e}
    if g is not None:
      out["dw_re"], out["dw_im"] = g[0], g[1]
    self.plan.evaluate(W[0], W[1], grad=g is not None, out=out)
    self.evals += 1

  def _precondition(self, g, h, use_factor=True):
    factor = None
    if use_factor and self.factor_fn is not None:
      factor = self.factor_fn(self)
    self.plan.precondition(g[0], g[1], factor, h[0], h[1])

  def _dot(self, a, b, slot=0):
    _lib.check(_lib.lib().jrx_vec_dot(
      a.data_ptr(), b.data_ptr(), self.n, self.dotv[slot:].data_ptr(),
      self._dws_ptr, self._dws.numel() - 256, _lib.stream_ptr()))

  def _axpby(self, alpha, x, beta, y, out):
    _lib.check(_lib.lib().jrx_vec_axpby(
      ctypes.c_double(alpha), x.data_ptr(), ctypes.c_double(beta),
      y.data_ptr(), out.data_ptr(), self.n, _lib.stream_ptr()))

  def _restart(self, use_factor=True):
    """W <- Q of the last evaluation (same energy, gauge R reset to I),
    gradient at the new W, and the preconditioned steepest-descent
    direction.  Returns (E, g.Pg)."""
    self.plan.orthonormal_params(self.W[0], self.W[1])
    self._eval(self.W, self.e, self.g)
    self._precondition(self.g, self.h, use_factor)
    self._axpby(-1.0, self.h, 0.0, self.h, self.d)
    self._dot(self.g, self.h, 0)
    return float(self.e[4]), float(self.dotv[0])

  # -- driver --------------------------------------------------------------
  def run(self, max_iter: int = 200, e_stop: Optional[float] = None,
          time_limit: Optional[float] = None, callback=None):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    self._eval(self.W, self.e)  # Q of the starting point
    E0, gh = self._restart()
    hist = [(time.perf_counter() - t0, E0)]
    rejected = fallbacks = stalled = 0
    since_restart = 0
    it = 0
    for it in range(1, max_iter + 1):
      self._dot(self.g, self.d, 0)
      s0 = float(self.dotv[0])
      if not s0 < 0.0:  # not a descent direction: steepest descent
        self._axpby(-1.0, self.h, 0.0, self.h, self.d)
        s0 = -gh
        since_restart = 0
      lt = self.step
      # energy-only trial, then the minimum of the parabola through
      # E(0), E'(0) = s0 and E(lt)
      self._axpby(1.0, self.W, lt, self.d, self.W1)
      self._eval(self.W1, self.e1)
      Et = float(self.e1[4])
      c = (Et - E0 - s0 * lt) / (lt * lt)
      lam = -s0 / (2.0 * c) if c > 0 else 4.0 * lt
      lam = min(max(lam, 0.05 * lt), 8.0 * lt)
      self._axpby(1.0, self.W, lam, self.d, self.W1)
      self._eval(self.W1, self.e1, self.g1)
      E1 = float(self.e1[4])
      if not np.isfinite(E1) or E1 > E0:
        rejected += 1
        if np.isfinite(Et) and Et < E0:
          lam = lt
          self._axpby(1.0, self.W, lam, self.d, self.W1)
          self._eval(self.W1, self.e1, self.g1)
          E1 = float(self.e1[4])
        if not np.isfinite(E1) or E1 > E0:
          # fall back: plain preconditioner, steepest descent, shorter step
          fallbacks += 1
          stalled += 1
          self.step = lt * 0.2
          self._eval(self.W, self.e)
          E0, gh = self._restart(use_factor=False)
          since_restart = 0
          hist.append((time.perf_counter() - t0, E0))
          if stalled >= 4:
            break
          continue
      stalled = 0
      self.step = lam
      self.W, self.W1 = self.W1, self.W
      self.e, self.e1 = self.e1, self.e
      since_restart += 1
      if since_restart >= self.restart:
        self.g, self.g1 = self.g1, self.g
        E1, gh = self._restart()
        since_restart = 0
      else:
        self._precondition(self.g1, self.h1)
        self._dot(self.g1, self.h1, 0)
        self._dot(self.g, self.h1, 1)
        g1h1, g0h1 = float(self.dotv[0]), float(self.dotv[1])
        beta = max(0.0, (g1h1 - g0h1) / gh) if gh > 0 else 0.0
        self._axpby(-1.0, self.h1, beta, self.d, self.d)
        self.g, self.g1 = self.g1, self.g
        self.h, self.h1 = self.h1, self.h
        gh = g1h1
      E0 = E1
      hist.append((time.perf_counter() - t0, E0))
      if callback is not None:
        callback(self, it, E0)
      if e_stop is not None and E0 <= e_stop:
        break
      if time_limit is not None and hist[-1][0] > time_limit:
        break
    torch.cuda.synchronize()
    return MinimizeResult(energy=E0, iterations=it, evaluations=self.evals,
                          seconds=time.perf_counter() - t0, history=hist,
                          rejected=rejected, fallbacks=fallbacks)

  def params(self):
    return {"w_re": self.W[0], "w_im": self.W[1]}
