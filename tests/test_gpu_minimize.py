"""Preconditioned minimisation on the GPU: the kinetic preconditioner and
band norms against the oracle, the vector kernels against numpy, and the
conjugate-gradient driver (monotone energy, end point checked by the
oracle, same minimum from two starts)."""
import numpy as np
import pytest

from helpers import oracle_problem, rel_max
from oracle import adjoint as OA
from oracle import optim as OOpt

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _si2(z=4.0):
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(z)
  prob = oracle_problem(wl.crystal, wl.grid_sizes, wl.k_grid, wl.num_bands,
                        wl.cutoff_energy)
  return wl, prob


def test_band_norms_and_tpa_match_oracle():
  from jrystal_amd.objective import EnergyPlan
  wl, prob = _si2()
  p = wl.params(2)
  plan = EnergyPlan.from_workload(wl)
  r = plan.evaluate(torch.tensor(p["w_re"], device="cuda"),
                    torch.tensor(p["w_im"], device="cuda"), grad=True)
  tb = plan.band_norms().cpu().numpy()
  tb_ref = OOpt.band_norms(p, prob)
  assert rel_max(tb, tb_ref) < 1e-12
  g = {"w_re": r["dw_re"].cpu().numpy(), "w_im": r["dw_im"].cpu().numpy()}
  rng = np.random.default_rng(5)
  f = rng.uniform(0.5, 2.0, g["w_re"].shape)
  for factor in (None, f):
    ft = None if factor is None else torch.tensor(factor, device="cuda")
    hr, hi = plan.precondition(r["dw_re"], r["dw_im"], ft)
    ref = OOpt.tpa_precondition(g, p, prob, factor)
    assert rel_max(hr.cpu().numpy(), ref["w_re"]) < 1e-12
    assert rel_max(hi.cpu().numpy(), ref["w_im"]) < 1e-12


def test_fused_precondition_dots_matches_oracle():
  """jrx_plan_precondition_dots: h1 = K factor g1 as the oracle's TPA step,
  and (g1.h1, g0.h1, g1.d) as numpy dots of the same arrays."""
  from jrystal_amd import _lib
  from jrystal_amd.objective import EnergyPlan
  wl, prob = _si2()
  p = wl.params(3)
  plan = EnergyPlan.from_workload(wl)
  plan.evaluate(torch.tensor(p["w_re"], device="cuda"),
                torch.tensor(p["w_im"], device="cuda"), grad=True)
  L = _lib.lib()
  rng = np.random.default_rng(17)
  shp = plan.param_shape
  g1, g0, d = (rng.standard_normal((2,) + shp) for _ in range(3))
  f = rng.uniform(0.5, 2.0, shp)
  ws = L.jrx_cg_dots_workspace_bytes()
  scratch = torch.empty(ws, dtype=torch.uint8, device="cuda")
  t = lambda a: torch.tensor(a, device="cuda")  # noqa
  g1t, g0t, dt = t(g1), t(g0), t(d)
  for factor in (None, f):
    ft = None if factor is None else t(factor)
    h = torch.full_like(g1t, np.nan)
    out = torch.full((4,), np.nan, dtype=torch.float64, device="cuda")
    _lib.check(L.jrx_plan_precondition_dots(
      plan._plan, g1t[0].data_ptr(), g1t[1].data_ptr(), g0t[0].data_ptr(),
      g0t[1].data_ptr(), dt[0].data_ptr(), dt[1].data_ptr(), _lib.ptr(ft),
      h[0].data_ptr(), h[1].data_ptr(), out.data_ptr(), scratch.data_ptr(),
      ws, plan._ws_ptr, plan.workspace_bytes, _lib.stream_ptr()))
    ref = OOpt.tpa_precondition({"w_re": g1[0], "w_im": g1[1]}, p, prob,
                                factor)
    href = np.stack([ref["w_re"], ref["w_im"]])
    hh = h.cpu().numpy()
    assert rel_max(hh, href) < 1e-12
    o = out.cpu().numpy()
    for got, want in ((o[0], np.sum(g1 * href)), (o[1], np.sum(g0 * href)),
                      (o[2], np.sum(g1 * d))):
      assert abs(got - want) <= 1e-12 * np.sum(np.abs(g1) * np.abs(d) + 1.0)
  with pytest.raises(ValueError):
    _lib.check(L.jrx_plan_precondition_dots(
      plan._plan, g1t[0].data_ptr(), g1t[1].data_ptr(), g0t[0].data_ptr(),
      g0t[1].data_ptr(), dt[0].data_ptr(), dt[1].data_ptr(), None,
      h[0].data_ptr(), h[1].data_ptr(), out.data_ptr(), scratch.data_ptr(),
      ws - 8, plan._ws_ptr, plan.workspace_bytes, _lib.stream_ptr()))


def test_energy_only_evaluation_sets_band_norms():
  from jrystal_amd.objective import EnergyPlan
  wl, prob = _si2()
  p = wl.params(4)
  plan = EnergyPlan.from_workload(wl)
  plan.evaluate(torch.tensor(p["w_re"], device="cuda"),
                torch.tensor(p["w_im"], device="cuda"), grad=False)
  assert rel_max(plan.band_norms().cpu().numpy(),
                 OOpt.band_norms(p, prob)) < 1e-12


@pytest.mark.parametrize("n", [1, 255, 256 * 16 + 3, 1_000_003])
def test_vec_dot_and_axpby(n):
  from jrystal_amd import _lib
  L = _lib.lib()
  rng = np.random.default_rng(n)
  a, b = rng.standard_normal(n), rng.standard_normal(n)
  at, bt = torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda")
  out = torch.zeros(1, dtype=torch.float64, device="cuda")
  ws = torch.empty(L.jrx_dot_workspace_bytes(), dtype=torch.uint8,
                   device="cuda")
  _lib.check(L.jrx_vec_dot(at.data_ptr(), bt.data_ptr(), n, out.data_ptr(),
                           ws.data_ptr(), ws.numel(), _lib.stream_ptr()))
  ref = float(np.dot(a, b))
  assert abs(float(out) - ref) <= 1e-12 * max(1.0, np.abs(a * b).sum())
  import ctypes
  _lib.check(L.jrx_vec_axpby(ctypes.c_double(0.5), at.data_ptr(),
                             ctypes.c_double(-2.0), bt.data_ptr(),
                             bt.data_ptr(), n, _lib.stream_ptr()))
  np.testing.assert_allclose(bt.cpu().numpy(), 0.5 * a - 2.0 * b, rtol=1e-15,
                             atol=1e-15)


def _run(wl, seed, iters, **kw):
  from jrystal_amd.minimize import Minimizer
  from jrystal_amd.objective import EnergyPlan
  plan = EnergyPlan.from_workload(wl)
  m = Minimizer(plan, wl.params(seed), **kw)
  res = m.run(max_iter=iters)
  return m, res


def test_cg_energy_decreases_and_end_point_matches_oracle():
  wl, prob = _si2()
  m, res = _run(wl, 0, 40)
  e = [h[1] for h in res.history]
  assert all(b <= a + 1e-12 for a, b in zip(e, e[1:])), e
  assert e[-1] < e[0] - 1.0
  p = {k: v.cpu().numpy() for k, v in m.params().items()}
  parts, _, _ = OA.energy_and_grad(p, prob, want_grad=False)
  assert abs(parts["total"] - res.energy) < 1e-8


def test_cg_reaches_same_minimum_from_two_starts():
  wl, _ = _si2()
  _, r0 = _run(wl, 0, 300)
  _, r1 = _run(wl, 1, 300)
  assert abs(r0.energy - r1.energy) < 1e-6, (r0.energy, r1.energy)
  # the last iterations change the energy by far less than the target
  assert abs(r0.history[-1][1] - r0.history[-10][1]) < 1e-7


def test_orthonormal_params_is_qr_factor_and_keeps_energy():
  from jrystal_amd.objective import EnergyPlan
  wl, prob = _si2()
  p = wl.params(6)
  plan = EnergyPlan.from_workload(wl)
  wr = torch.tensor(p["w_re"], device="cuda")
  wi = torch.tensor(p["w_im"], device="cuda")
  e0 = plan.evaluate(wr, wi, grad=False)["energies"][4].item()
  qr, qi = plan.orthonormal_params()
  q = qr.cpu().numpy() + 1j * qi.cpu().numpy()
  qn, rn = np.linalg.qr(p["w_re"] + 1j * p["w_im"], mode="reduced")
  d = np.diagonal(rn, axis1=-2, axis2=-1)
  q_pos = qn * (d / np.abs(d))[..., None, :]  # R with positive diagonal
  assert rel_max(q, q_pos) < 1e-11
  e1 = plan.evaluate(qr, qi, grad=False)["energies"][4].item()
  assert abs(e1 - e0) < 1e-10
