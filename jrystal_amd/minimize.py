"""Preconditioned direct minimisation of the total energy on the GPU.

The reference minimises with optax Adam on the raw parameters
(calc_ground_state_energy_all_electrons.py:175-220).  This module adds the
classic plane-wave alternative: preconditioned conjugate gradients on W
(Teter, Payne & Allan, PRB 40, 12255 (1989)) with

* search direction d = -P g + beta d_prev (Polak-Ribiere+), restarted every
  ``restart`` iterations or when d is not a descent direction;
* a gauge reset at every restart: W <- Q V, the QR factor of W (E is
  invariant under W -> W T, T upper triangular) rotated by the
  Rayleigh-Ritz eigenvectors V within each block of equally occupied
  bands (E is invariant there too), so every column is an approximate
  eigenstate with R = I;
* P = diag(K(|G+k|^2 / tb_band)) (the TPA kinetic preconditioner), times an
  optional per-coefficient factor -- the hook of the learned
  preconditioner, held fixed within a restart cycle;
* a line search with the initial step lam_prev s0_prev / s0 (Nocedal &
  Wright 3.5) that fits a parabola through E(0), E'(0) and one energy-only
  trial E(lt), then evaluates energy and gradient at its minimum (or, with
  line_search="secant", one evaluation with the gradient and a secant
  correction);
* acceptance only when the energy does not increase; otherwise up to
  ``backtracks`` parabolic backtracking steps, and if those fail too, a
  fall-back to the plain preconditioner for two restart cycles from a
  steepest-descent restart;
* a stall check: ``flat_window`` iterations without a decrease above
  rounding while the first-order decrease of a step, g.Pg times the step
  (invariant under the drift of |W| between gauge resets), is still far
  from zero (the steps have collapsed, conjugacy is lost) trigger a
  restart with the initial step size; the run stops as "converged" only
  when the energy is flat and that decrease is small, and as "stalled"
  after ``max_stalls`` stall restarts in a row.

Every vector operation is a HIP kernel (jrx_vec_dot, jrx_vec_axpby,
jrx_plan_precondition, jrx_plan_rotate); the host sees a handful of
scalars per iteration and the small subspace matrices at restarts.
"""
import ctypes
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _lib
from .objective import EnergyPlan


# restart interval of plain TPA CG: the fastest to |dE| < 1e-6 Ha of 60, 120
# and no periodic restart on three strained 8-atom cells (4x4x4 k-points)
PLAIN_RESTART = 120


@dataclass
class MinimizeResult:
  energy: float
  iterations: int
  evaluations: int
  seconds: float
  history: List[tuple] = field(default_factory=list)  # (seconds, energy)
  rejected: int = 0
  fallbacks: int = 0
  status: str = ""


class Minimizer:
  """Preconditioned CG on one plan from params {'w_re', 'w_im'} (copied to
  the device).  factor_fn(minimizer) returns the per-coefficient factor of
  the preconditioner (or None for plain TPA); restart is the gauge-reset
  interval and first_restart the length of the first cycle; line_search is
  "parabola" (default) or "secant".  run() minimises; params() returns the
  current W."""

  def __init__(self, plan: EnergyPlan, params, factor_fn: Optional[Callable]
               = None, restart: int = PLAIN_RESTART, step0: float = 0.5,
               line_search: str = "parabola", ritz: bool = True,
               first_restart: Optional[int] = None):
    if line_search not in ("secant", "parabola"):
      raise ValueError(f"unknown line search {line_search!r}")
    self.line_search = line_search
    self.ritz = ritz
    self.slope_tol = 1e-15
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
    ws = max(_lib.lib().jrx_dot_workspace_bytes(),
             _lib.lib().jrx_cg_dots_workspace_bytes())
    self._dws = torch.empty(ws + 256, dtype=torch.uint8, device=dev)
    self._dws_ptr = (self._dws.data_ptr() + 255) // 256 * 256
    self._s0_next = None  # g.d of the next iteration, from the fused pass
    self.factor_fn = factor_fn
    self.restarts = 0       # gauge resets so far
    self._plain_cycles = 0  # restart cycles left on the plain preconditioner
    self.restart = restart
    # length of the first restart cycle (default: restart)
    self.first_restart = restart if first_restart is None else first_restart
    self.step0 = step0
    self.step = step0
    self.max_stalls = 3
    self.stall_slope = 1e-10  # g.Pg step / max(1, |E|) above which flat = stall
    self.backtracks = 6
    self.flat_window = 12
    self.flat_tol = 1e-14
    self._s0_prev = None
    self._ritz_v = None
    self.evals = 0

  # -- kernels -------------------------------------------------------------
  def _eval(self, W, e, g=None):
    out = {"energies": e}
    if g is not None:
      out["dw_re"], out["dw_im"] = g[0], g[1]
    self.plan.evaluate(W[0], W[1], grad=g is not None, out=out)
    self.evals += 1

  def _precondition(self, g, h):
    factor = None
    if self.factor_fn is not None and self._plain_cycles == 0:
      factor = self.factor_fn(self)
    self.plan.precondition(g[0], g[1], factor, h[0], h[1])

  def _precondition_dots(self):
    """h1 = P g1 and dotv[0..2] = (g1.h1, g.h1, g1.d) in one pass."""
    factor = None
    if self.factor_fn is not None and self._plain_cycles == 0:
      factor = self.factor_fn(self)
    p = self.plan
    _lib.check(_lib.lib().jrx_plan_precondition_dots(
      p._plan, _lib.ptr(self.g1[0]), _lib.ptr(self.g1[1]),
      _lib.ptr(self.g[0]), _lib.ptr(self.g[1]), _lib.ptr(self.d[0]),
      _lib.ptr(self.d[1]), _lib.ptr(factor), _lib.ptr(self.h1[0]),
      _lib.ptr(self.h1[1]), _lib.ptr(self.dotv), self._dws_ptr,
      self._dws.numel() - 256, p._ws_ptr, p.workspace_bytes,
      _lib.stream_ptr(p.device)))

  def _dot(self, a, b, slot=0):
    _lib.check(_lib.lib().jrx_vec_dot(
      a.data_ptr(), b.data_ptr(), self.n, self.dotv[slot:].data_ptr(),
      self._dws_ptr, self._dws.numel() - 256, _lib.stream_ptr()))

  def _axpby(self, alpha, x, beta, y, out):
    _lib.check(_lib.lib().jrx_vec_axpby(
      ctypes.c_double(alpha), x.data_ptr(), ctypes.c_double(beta),
      y.data_ptr(), out.data_ptr(), self.n, _lib.stream_ptr()))

  def _ritz_rotation(self):
    """Unitary V (s,k,nb,nb) on the device: on each run of consecutive
    bands with equal occupation f > 0, the eigenvectors of the subspace
    Hamiltonian (Q^H Qbar) / 2f in ascending order; identity elsewhere
    (a GPU Jacobi eigensolver, no host round trip).  Needs the last
    evaluation to carry a gradient."""
    if self._ritz_v is None:
      self._ritz_v = torch.empty(
        (self.plan.ns, self.plan.nk, self.plan.nb, self.plan.nb),
        dtype=torch.complex128, device=self.plan.device)
    return self.plan.ritz_rotation(self._ritz_v)

  def _restart(self, gh_before=None):
    """Gauge reset W <- Q V of the last evaluation (same energy; R = I and,
    with ``ritz``, Rayleigh-Ritz bands within each occupation block), the
    gradient at the new W, and the preconditioned steepest-descent
    direction.  The last evaluation must carry a gradient.  Returns
    (E, g.Pg).

    The reset undoes the drift of |W| since the last one; g.Pg scales as
    1/|W|^2 and the step as |W|^2, so with ``gh_before`` (g.Pg before the
    reset) the step is rescaled by gh_before / g.Pg and the next trial
    step starts from it."""
    self._s0_next = None
    if self.ritz:
      self.plan.rotate(self._ritz_rotation(), self.W[0], self.W[1])
    else:
      self.plan.orthonormal_params(self.W[0], self.W[1])
    self._eval(self.W, self.e, self.g)
    self.restarts += 1
    self._precondition(self.g, self.h)
    self._axpby(-1.0, self.h, 0.0, self.h, self.d)
    self._dot(self.g, self.h, 0)
    gh = float(self.dotv[0])
    if gh_before is not None:
      ratio = gh_before / gh if gh > 0 else np.nan
      if np.isfinite(ratio) and ratio > 0:
        self.step *= ratio
        self._s0_prev = None
    return float(self.e[4]), gh

  # -- line searches along d from W (energy E0, slope s0 = g.d < 0) --------
  def _search_parabola(self, E0, s0, lt):
    """Energy-only trial at lt, then energy and gradient at the minimum of
    the parabola through E(0), E'(0) and E(lt).  Two evaluations."""
    self._axpby(1.0, self.W, lt, self.d, self.W1)
    self._eval(self.W1, self.e1)
    Et = float(self.e1[4])
    c = (Et - E0 - s0 * lt) / (lt * lt)
    lam = -s0 / (2.0 * c) if c > 0 else 4.0 * lt
    lam = min(max(lam, 0.05 * lt), 8.0 * lt)
    self._axpby(1.0, self.W, lam, self.d, self.W1)
    self._eval(self.W1, self.e1, self.g1)
    self._next_step = lam
    return lam, float(self.e1[4]), Et

  def _search_secant(self, E0, s0, lt):
    """Energy and gradient at the predicted step lt; the slope there gives
    the secant minimum lam* = lt s0 / (s0 - s1).  The point is kept when
    lam* is within 30% of lt (the usual case once the step has adapted),
    else one more evaluation at lam*.  lam* predicts the next step."""
    self._axpby(1.0, self.W, lt, self.d, self.W1)
    self._eval(self.W1, self.e1, self.g1)
    Et = float(self.e1[4])
    self._dot(self.g1, self.d, 2)
    s1 = float(self.dotv[2])
    if np.isfinite(Et) and np.isfinite(s1) and s1 > s0:
      lam = lt * s0 / (s0 - s1)  # root of the slope, linear in lam
    else:
      c = (Et - E0 - s0 * lt) / (lt * lt) if np.isfinite(Et) else 1.0
      lam = -s0 / (2.0 * c) if c > 0 else 4.0 * lt
    lam = min(max(lam, 0.05 * lt), 8.0 * lt)
    self._next_step = lam
    if abs(lam - lt) <= 0.3 * lt and np.isfinite(Et) and Et <= E0:
      return lt, Et, Et
    self._axpby(1.0, self.W, lam, self.d, self.W1)
    self._eval(self.W1, self.e1, self.g1)
    return lam, float(self.e1[4]), Et

  # -- driver --------------------------------------------------------------
  def run(self, max_iter: int = 200, e_stop: Optional[float] = None,
          time_limit: Optional[float] = None, callback=None):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    self._eval(self.W, self.e, self.g)  # Q of the starting point
    E0, gh = self._restart()
    hist = [(time.perf_counter() - t0, E0)]
    rejected = fallbacks = stalled = stalls = 0
    status = "max_iter"
    since_restart = 0
    window_start = 0  # the flat-energy window looks at hist[window_start:]
    it = 0
    for it in range(1, max_iter + 1):
      if self._s0_next is not None:
        s0, self._s0_next = self._s0_next, None
      else:
        self._dot(self.g, self.d, 0)
        s0 = float(self.dotv[0])
      if not (np.isfinite(s0) and np.isfinite(gh)):
        # a preconditioner factor that is not finite: plain preconditioner
        fallbacks += 1
        stalled += 1
        self._plain_cycles = 2
        self._s0_prev = None
        self._eval(self.W, self.e, self.g)
        E0, gh = self._restart()
        since_restart = 0
        hist.append((time.perf_counter() - t0, E0))
        if stalled >= 3:
          status = "stalled"
          break
        continue
      if not s0 < 0.0:  # not a descent direction: steepest descent
        self._axpby(-1.0, self.h, 0.0, self.h, self.d)
        s0 = -gh
        since_restart = 0
      # E(W) = E(W T): between gauge resets |W| drifts, g.Pg scales as
      # 1/|W|^2 and the step as |W|^2; their product, the first-order
      # decrease of a step, does not
      if gh * self.step <= self.slope_tol * max(1.0, abs(E0)):
        status = "converged"  # decrease per step at the rounding of E
        break
      # initial trial step: keep the first-order decrease lam * s0 of the
      # last accepted step (Nocedal & Wright, Numerical Optimization, 3.5)
      lt = self.step
      if self._s0_prev is not None:
        lt = self.step * min(max(self._s0_prev / s0, 1e-4), 1e4)
      if not (np.isfinite(lt) and lt > 1e-150):
        status = "stalled"  # step collapsed (rounding-level slopes)
        break
      if self.line_search == "parabola":
        lam, E1, Et = self._search_parabola(E0, s0, lt)
      else:
        lam, E1, Et = self._search_secant(E0, s0, lt)
      if not np.isfinite(E1) or E1 > E0:
        rejected += 1
        if np.isfinite(Et) and Et < E0:
          lam = lt
        for _ in range(self.backtracks):
          if np.isfinite(E1) and E1 <= E0:
            break
          if not (np.isfinite(Et) and Et < E0 and lam == lt):
            # minimum of the parabola through E(0), E'(0), E(lam), kept in
            # [0.1, 0.5] lam
            c = (E1 - E0 - s0 * lam) / (lam * lam) if np.isfinite(E1) else 0
            lam_p = -s0 / (2 * c) if c > 0 else 0.1 * lam
            lam = min(max(lam_p, 0.1 * lam), 0.5 * lam)
          Et = np.nan
          self._axpby(1.0, self.W, lam, self.d, self.W1)
          self._eval(self.W1, self.e1, self.g1)
          E1 = float(self.e1[4])
        self._next_step = lam
        if not np.isfinite(E1) or E1 > E0:
          # fall back: plain preconditioner for this restart cycle and the
          # next, steepest descent from a gauge reset
          fallbacks += 1
          stalled += 1
          self.step = lam
          self._s0_prev = None
          self._plain_cycles = 2
          self._eval(self.W, self.e, self.g)
          E0, gh = self._restart(gh_before=gh)
          since_restart = 0
          hist.append((time.perf_counter() - t0, E0))
          if stalled >= 3:
            status = "stalled"
            break
          continue
      stalled = 0
      self.step = self._next_step
      self._s0_prev = s0
      self.W, self.W1 = self.W1, self.W
      self.e, self.e1 = self.e1, self.e
      since_restart += 1
      if since_restart >= (self.first_restart if self.restarts == 1
                           else self.restart):
        self._plain_cycles = max(0, self._plain_cycles - 1)
        E1, gh = self._restart(gh_before=gh)
        since_restart = 0
      else:
        # h1 = P g1 and the dots of the PR+ update in one pass; the next
        # slope g1.d_new = -g1.h1 + beta g1.d needs no further pass
        self._precondition_dots()
        g1h1, g0h1, g1d = (float(x) for x in self.dotv[:3].tolist())
        beta = max(0.0, (g1h1 - g0h1) / gh) if gh > 0 else 0.0
        self._axpby(-1.0, self.h1, beta, self.d, self.d)
        self._s0_next = -g1h1 + beta * g1d
        self.g, self.g1 = self.g1, self.g
        self.h, self.h1 = self.h1, self.h
        gh = g1h1
      E0 = E1
      hist.append((time.perf_counter() - t0, E0))
      if callback is not None:
        callback(self, it, E0)
      if e_stop is not None and E0 <= e_stop:
        status = "e_stop"
        break
      if len(hist) - window_start > self.flat_window:
        if (hist[-self.flat_window - 1][1] - E0
            > self.flat_tol * max(1.0, abs(E0))):
          stalls = 0
        else:
          # no decrease above rounding for flat_window iterations: a
          # minimum when the first-order decrease per step is small too,
          # else a stall
          if gh * self.step <= self.stall_slope * max(1.0, abs(E0)):
            status = "converged"
            break
          stalls += 1
          if stalls > self.max_stalls:
            status = "stalled"
            break
          # the steps collapsed with the slope far from zero: restart
          # with the initial step size
          self.step = self.step0
          self._s0_prev = None
          E0, gh = self._restart()
          since_restart = 0
          window_start = len(hist)
          hist.append((time.perf_counter() - t0, E0))
      if time_limit is not None and hist[-1][0] > time_limit:
        status = "time_limit"
        break
    torch.cuda.synchronize()
    return MinimizeResult(energy=E0, iterations=it, evaluations=self.evals,
                          seconds=time.perf_counter() - t0, history=hist,
                          rejected=rejected, fallbacks=fallbacks,
                          status=status)

  def params(self):
    return {"w_re": self.W[0], "w_im": self.W[1]}
