"""Batched Adam steps over structures of different shapes.

``BatchStep`` holds one ``EnergyPlan`` per structure (uneven grids,
k-meshes, band and plane-wave counts) and the native ``jrx_batch`` that
runs one optimiser step of the reference's driver
(calc_ground_state_energy_all_electrons.py:175-220: value_and_grad of the
total energy, then optax.adam on w_re and w_im) for all of them at once:

* every member's fused energy + gradient evaluation, members spread over
  ``streams`` auxiliary HIP streams (default priority) between a fork and a
  join event;
* one grouped Adam launch over all members' parameters, the step count
  kept on the device;
* ``capture()`` records one such step into a hipGraph; ``replay()``
  launches it, so a training loop costs one graph launch per step.

Plan options (FFT variant, state chunking, x-line workgroup target) come
from ``autotune.tune`` when ``tune=True``.
"""
import ctypes
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .objective import EnergyPlan


class BatchStep:

  def __init__(self, workloads, params: Optional[List[dict]] = None,
               learning_rate=0.01, b1=0.9, b2=0.999, eps=1e-8, eps_root=0.0,
               streams: int = 4, tune: bool = False, options=None,
               device=None):
    self.device = torch.device(device if device is not None else "cuda")
    self.workloads = list(workloads)
    n = len(self.workloads)
    if params is not None and len(params) != n:
      raise ValueError("one parameter dict per workload")
    if options is None:
      options = [{} for _ in range(n)]
      if tune:
        from . import autotune
        options = [autotune.tune(wl, device=self.device)
                   for wl in self.workloads]
    self.options = options
    self.plans, self.bufs = [], []
    members = (_lib.JrxBatchMember * n)()
    for i, wl in enumerate(self.workloads):
      plan = EnergyPlan.from_workload(wl, device=self.device, **options[i])
      p = params[i] if params is not None else wl.params(0)
      z = lambda: torch.zeros(plan.param_shape, dtype=torch.float64,  # noqa
                              device=self.device)
      b = {"w_re": torch.as_tensor(np.asarray(p["w_re"])).to(self.device)
           .to(torch.float64).contiguous(),
           "w_im": torch.as_tensor(np.asarray(p["w_im"])).to(self.device)
           .to(torch.float64).contiguous(),
           "dw_re": z(), "dw_im": z(), "m_re": z(), "m_im": z(),
           "v_re": z(), "v_im": z(),
           "energies": torch.zeros(5, dtype=torch.float64,
                                   device=self.device)}
      for k in ("w_re", "w_im"):
        if tuple(b[k].shape) != plan.param_shape:
          raise ValueError(f"member {i}: {k} has shape {tuple(b[k].shape)}")
      m = members[i]
      m.plan = plan._plan.value
      for k in ("w_re", "w_im", "dw_re", "dw_im", "m_re", "m_im", "v_re",
                "v_im", "energies"):
        setattr(m, k, b[k].data_ptr())
      m.workspace = plan._ws_ptr
      m.workspace_bytes = plan.workspace_bytes
      self.plans.append(plan)
      self.bufs.append(b)
    cfg = _lib.JrxAdamConfig(float(learning_rate), float(b1), float(b2),
                             float(eps), float(eps_root))
    h = ctypes.c_void_p()
    with torch.cuda.device(self.device):
      _lib.check(_lib.lib().jrx_batch_create(ctypes.byref(h), members, n,
                                             ctypes.byref(cfg), int(streams)))
    self._b = h
    self._stream = None
    self.captured = False

  def close(self):
    if getattr(self, "_b", None) is not None and self._b.value:
      _lib.lib().jrx_batch_destroy(self._b)
      self._b = None

  def __del__(self):
    try:
      self.close()
    except Exception:
      pass

  def step(self):
    """One eager step on the current stream."""
    _lib.check(_lib.lib().jrx_batch_step(self._b,
                                         _lib.stream_ptr(self.device)))

  def capture(self):
    """Record one step into a hipGraph on a side stream (capture needs a
    stream other than the null stream); the step is not executed."""
    if self._stream is None:
      self._stream = torch.cuda.Stream(device=self.device)
    self._stream.wait_stream(torch.cuda.current_stream(self.device))
    with torch.cuda.stream(self._stream):
      _lib.check(_lib.lib().jrx_batch_capture(self._b,
                                              self._stream.cuda_stream))
    self.captured = True

  def replay(self):
    """Launch the captured step on the current stream."""
    _lib.check(_lib.lib().jrx_batch_replay(self._b,
                                           _lib.stream_ptr(self.device)))

  def set_step_count(self, count: int):
    _lib.check(_lib.lib().jrx_batch_set_step_count(
      self._b, int(count), _lib.stream_ptr(self.device)))

  def step_count(self) -> int:
    c = ctypes.c_int64()
    _lib.check(_lib.lib().jrx_batch_get_step_count(
      self._b, ctypes.byref(c), _lib.stream_ptr(self.device)))
    return int(c.value)

  def energies(self):
    """(n, 5) host array of the last step's energies (kinetic, hartree,
    external, xc, total) of every member."""
    return torch.stack([b["energies"] for b in self.bufs]).cpu().numpy()

  def params(self, i):
    return {"w_re": self.bufs[i]["w_re"], "w_im": self.bufs[i]["w_im"]}
