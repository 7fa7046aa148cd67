"""jrx_energy_grad_static (the evaluation with its workspace in a static
device array, for callers that cannot allocate) against jrx_energy_grad with
a caller workspace: two calls in a row, the second on a smaller cell, every
output equal bit for bit (same kernels, same inputs, deterministic
reductions); a plan larger than the static workspace is rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _outputs(plan):
  dev = plan.device
  shp = plan.param_shape
  return {
    "energies": torch.full((5,), np.nan, dtype=torch.float64, device=dev),
    "dw_re": torch.full(shp, np.nan, dtype=torch.float64, device=dev),
    "dw_im": torch.full(shp, np.nan, dtype=torch.float64, device=dev),
    "density": torch.full((plan.ns,) + tuple(plan.grid_sizes), np.nan,
                          dtype=torch.float64, device=dev),
    "coefficients": torch.full((plan.ns, plan.nk, plan.nb, plan.num_g),
                               np.nan, dtype=torch.complex128, device=dev),
  }


def _static_call(plan, w_re, w_im):
  from jrystal_amd import _lib
  out = _outputs(plan)
  L = _lib.lib()
  with torch.cuda.device(plan.device):
    rc = L.jrx_energy_grad_static(
      plan._plan, w_re.data_ptr(), w_im.data_ptr(),
      out["energies"].data_ptr(), out["dw_re"].data_ptr(),
      out["dw_im"].data_ptr(), out["density"].data_ptr(),
      out["coefficients"].data_ptr(), _lib.stream_ptr(plan.device))
  _lib.check(rc)
  return out


def _compare(wl, seed):
  from jrystal_amd import _lib
  from jrystal_amd.objective import EnergyPlan
  plan = EnergyPlan.from_workload(wl)
  assert plan.workspace_bytes <= _lib.STATIC_WORKSPACE_BYTES
  p = wl.params(seed)
  w_re = torch.tensor(p["w_re"], device="cuda")
  w_im = torch.tensor(p["w_im"], device="cuda")
  ref = plan.evaluate(w_re, w_im, grad=True, density=True,
                      coefficients=True)
  got = _static_call(plan, w_re, w_im)
  torch.cuda.synchronize()
  for name in ("energies", "dw_re", "dw_im", "density", "coefficients"):
    assert torch.isfinite(torch.view_as_real(got[name])
                          if got[name].is_complex() else got[name]).all(), name
    assert torch.equal(got[name], ref[name]), name
  return float(got["energies"][4])


def test_static_workspace_matches_caller_workspace_twice():
  from jrystal_amd import workloads as W
  big = W.si2_fcc(4.0)                      # 32^3, 8 k, 8 bands
  small = W.tiny()                          # 12^3, 2 k, 3 bands
  e_big = _compare(big, seed=0)
  e_small = _compare(small, seed=1)         # smaller input, same array
  assert np.isfinite(e_big) and np.isfinite(e_small)


def test_static_workspace_rejects_oversize_plan():
  from jrystal_amd import _lib
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import EnergyPlan
  plan = EnergyPlan.from_workload(W.si8_cube(4.0, k=2))  # 64^3, 8 k, 32 bands
  assert plan.workspace_bytes > _lib.STATIC_WORKSPACE_BYTES
  z = torch.zeros(plan.param_shape, dtype=torch.float64, device="cuda")
  e = torch.zeros(5, dtype=torch.float64, device="cuda")
  L = _lib.lib()
  rc = L.jrx_energy_grad_static(plan._plan, z.data_ptr(), z.data_ptr(),
                                e.data_ptr(), None, None, None, None,
                                _lib.stream_ptr())
  with pytest.raises(NotImplementedError, match="static workspace"):
    _lib.check(rc)
