"""Learned preconditioner on the GPU: network kernels against the oracle
restatement, plan features / density statistics / inference against the
oracle, the Rayleigh-Ritz gauge reset, and a short training run."""
import numpy as np
import pytest

from helpers import oracle_problem, rel_max
from oracle import adjoint as OA
from oracle import mlp as OM
from oracle import optim as OOpt

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _lib():
  from jrystal_amd import _lib as L
  return L


def _net(seed=0):
  rng = np.random.default_rng(seed)
  p = (rng.standard_normal(OM.NPARAMS) * 0.3).astype(np.float32)
  norm = np.concatenate([rng.standard_normal(8) * 0.1,
                         rng.uniform(0.5, 2.0, 8)]).astype(np.float32)
  return rng, p, norm


def _dev(a):
  return torch.tensor(np.ascontiguousarray(a), device="cuda")


def test_mlp_forward_backward_match_oracle():
  L = _lib()
  rng, p, norm = _net()
  n = 1000  # 7 full tiles and a ragged one
  x = rng.standard_normal((n, 8)).astype(np.float32)
  dy = rng.standard_normal((n, 2)).astype(np.float32)
  pt, nt, xt, dyt = _dev(p), _dev(norm), _dev(x), _dev(dy)
  y = torch.empty((n, 2), dtype=torch.float32, device="cuda")
  L.check(L.lib().jrx_mlp_forward(pt.data_ptr(), nt.data_ptr(), xt.data_ptr(),
                                  n, y.data_ptr(), L.stream_ptr()))
  y_ref = OM.forward(p, norm, x)
  assert rel_max(y.cpu().numpy(), y_ref) < 2e-5
  ws = torch.empty(L.lib().jrx_mlp_backward_workspace_bytes(n),
                   dtype=torch.uint8, device="cuda")
  g = torch.empty(OM.NPARAMS, dtype=torch.float32, device="cuda")
  L.check(L.lib().jrx_mlp_backward(pt.data_ptr(), nt.data_ptr(), xt.data_ptr(),
                                   dyt.data_ptr(), n, g.data_ptr(), 0,
                                   ws.data_ptr(), ws.numel(), L.stream_ptr()))
  g_ref = OM.backward(p, norm, x, dy)
  assert rel_max(g.cpu().numpy(), g_ref) < 1e-4
  # accumulate mode adds a second gradient
  L.check(L.lib().jrx_mlp_backward(pt.data_ptr(), nt.data_ptr(), xt.data_ptr(),
                                   dyt.data_ptr(), n, g.data_ptr(), 1,
                                   ws.data_ptr(), ws.numel(), L.stream_ptr()))
  assert rel_max(g.cpu().numpy(), 2 * g_ref) < 1e-4


def test_losses_and_adam_match_oracle():
  L = _lib()
  rng = np.random.default_rng(1)
  ns, m = 5, 300
  y = rng.standard_normal((ns * m, 2)).astype(np.float32)
  a = rng.standard_normal(ns * m).astype(np.float32)
  b = rng.uniform(0.1, 2.0, ns * m).astype(np.float32)
  t2 = rng.uniform(1.0, 3.0, ns).astype(np.float32) * m
  yt, at, bt, tt = _dev(y), _dev(a), _dev(b), _dev(t2)
  dy = torch.empty_like(yt)
  loss = torch.empty(ns, dtype=torch.float32, device="cuda")
  L.check(L.lib().jrx_mlp_cosine_loss(yt.data_ptr(), at.data_ptr(),
                                      bt.data_ptr(), tt.data_ptr(), m, ns,
                                      dy.data_ptr(), loss.data_ptr(),
                                      L.stream_ptr()))
  l_ref, dy_ref = OM.cosine_loss(y, a, b, t2, m)
  np.testing.assert_allclose(loss.cpu().numpy(), l_ref, rtol=1e-4, atol=1e-6)
  assert rel_max(dy.cpu().numpy(), dy_ref) < 1e-4
  tgt = rng.standard_normal(ns * m).astype(np.float32)
  ws = torch.empty(L.lib().jrx_mlp_mse_workspace_bytes(ns * m),
                   dtype=torch.uint8, device="cuda")
  l1 = torch.empty(1, dtype=torch.float32, device="cuda")
  L.check(L.lib().jrx_mlp_mse_loss(yt.data_ptr(), _dev(tgt).data_ptr(),
                                   ns * m, dy.data_ptr(), l1.data_ptr(),
                                   ws.data_ptr(), ws.numel(), L.stream_ptr()))
  l_ref, dy_ref = OM.mse_loss(y, tgt)
  assert abs(float(l1) - l_ref) < 1e-5 * l_ref
  assert rel_max(dy.cpu().numpy(), dy_ref) < 1e-5
  import ctypes
  p = rng.standard_normal(77).astype(np.float32)
  pt = _dev(p)
  mt, vt = torch.zeros_like(pt), torch.zeros_like(pt)
  pr, mr, vr = p.astype(float), np.zeros(77), np.zeros(77)
  for c in range(1, 4):
    g = rng.standard_normal(77).astype(np.float32)
    L.check(L.lib().jrx_mlp_adam(pt.data_ptr(), _dev(g).data_ptr(),
                                 mt.data_ptr(), vt.data_ptr(), 77,
                                 ctypes.c_double(1e-2), ctypes.c_double(0.9),
                                 ctypes.c_double(0.999), ctypes.c_double(1e-8),
                                 c, L.stream_ptr()))
    pr, mr, vr = OM.adam(pr, g.astype(float), mr, vr, c, 1e-2)
  np.testing.assert_allclose(pt.cpu().numpy(), pr, rtol=1e-5, atol=1e-6)


def _si2_plan(seed=2):
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import EnergyPlan
  wl = W.si2_fcc(4.0)
  prob = oracle_problem(wl.crystal, wl.grid_sizes, wl.k_grid, wl.num_bands,
                        wl.cutoff_energy)
  p = wl.params(seed)
  plan = EnergyPlan.from_workload(wl)
  return wl, prob, p, plan


def test_density_stats_features_and_inference_match_oracle():
  from jrystal_amd import learned
  wl, prob, p, plan = _si2_plan()
  plan.evaluate(p["w_re"], p["w_im"], grad=True)
  stats = learned.density_stats(plan).cpu().numpy()
  _, _, rho = OA.energy_and_grad(p, prob, want_grad=False)
  tot = rho.sum(0)
  np.testing.assert_allclose(stats, [tot.mean(), tot.var()], rtol=1e-11)
  tb = OOpt.band_norms(p, prob)
  gm = prob.mask_g()
  gk2 = np.sum((gm[None] + prob.kpts[:, None, :])**2, axis=-1)
  ns, nk, ng, nb = plan.param_shape
  # band gaps from Q^H Qbar, against the oracle's Q and Qbar
  _, _, _, ex = OA.energy_and_grad(p, prob, extras=True)
  d = np.diagonal(np.linalg.qr(p["w_re"] + 1j * p["w_im"])[1], axis1=-2,
                  axis2=-1)
  ph = d / np.abs(d)  # column phases of the positive-diagonal QR
  sub = np.conj(np.swapaxes(ex["q"] * ph[..., None, :], -1, -2)) @ (
    ex["qbar"] * ph[..., None, :])
  gap_ref = OM.band_gaps(sub, plan.occupation)
  gap = learned.band_gaps(plan)
  assert rel_max(gap.cpu().numpy(), gap_ref) < 1e-9
  assert np.all(gap_ref[..., 4:] == 0) and np.all(gap_ref[..., :4] >= 0)
  feats_ref = OM.features(gk2[None, :, :, None], tb[:, :, None, :], *stats,
                          gap=gap_ref[:, :, None, :])
  feats_ref = feats_ref.reshape(-1, 8)
  rng = np.random.default_rng(0)
  idx = rng.integers(0, ns * nk * ng * nb, 3000)
  idx[:2] = [-1, ns * nk * ng * nb]  # out of range -> zeros
  x = learned.coefficient_features(plan, idx, learned.density_stats(plan),
                                   gap)
  x = x.cpu().numpy()
  assert np.all(x[:2] == 0)
  np.testing.assert_allclose(x[2:], feats_ref[idx[2:]], rtol=1e-6, atol=1e-6)
  _, par, norm = _net(3)
  par = par * 0.2
  net = learned.PrecondNet(par, norm)
  fac = torch.empty(plan.param_shape, dtype=torch.float64, device="cuda")
  wr = torch.tensor(p["w_re"], device="cuda")
  wi = torch.tensor(p["w_im"], device="cuda")
  learned.apply(plan, net, learned.density_stats(plan), gap=gap, factor=fac,
                w_re=wr, w_im=wi)
  y = OM.forward(par, norm, feats_ref)
  f_ref = np.exp(np.clip(y[:, 0], -8, 8)).reshape(plan.param_shape)
  a_ref = np.exp(np.clip(y[:, 1], -30, 8)).reshape(plan.param_shape)
  assert rel_max(fac.cpu().numpy(), f_ref) < 5e-5
  assert rel_max(wr.cpu().numpy(), a_ref * p["w_re"]) < 5e-5
  assert rel_max(wi.cpu().numpy(), a_ref * p["w_im"]) < 5e-5


def test_rayleigh_ritz_rotation_keeps_energy_and_diagonalises():
  from jrystal_amd.minimize import Minimizer
  wl, prob, p, plan = _si2_plan(5)
  r = plan.evaluate(p["w_re"], p["w_im"], grad=True)
  e0 = float(r["energies"][4])
  m = Minimizer(plan, p)
  V = m._ritz_rotation().cpu().numpy()
  nocc = int((plan.occupation[0, 0] > 0).sum())
  eye = np.eye(plan.nb)
  assert np.max(np.abs(np.conj(np.swapaxes(V, -1, -2)) @ V - eye)) < 1e-12
  assert np.max(np.abs(V[..., nocc:, nocc:] - eye[nocc:, nocc:])) == 0
  wr, wi = plan.rotate(torch.from_numpy(V))
  r = plan.evaluate(wr, wi, grad=True)
  assert abs(float(r["energies"][4]) - e0) < 1e-10
  S = plan.subspace_matrix().cpu().numpy()[..., :nocc, :nocc]
  off = S - np.diagonal(S, axis1=-2, axis2=-1)[..., None] * np.eye(nocc)
  assert np.max(np.abs(off)) < 1e-9 * np.max(np.abs(S))
  ev = np.real(np.diagonal(S, axis1=-2, axis2=-1))
  assert np.all(np.diff(ev, axis=-1) >= -1e-12 * np.abs(ev).max())
  # the rotated parameters are the QR factor of W times V
  q, rr = np.linalg.qr(p["w_re"] + 1j * p["w_im"])
  d = np.diagonal(rr, axis1=-2, axis2=-1)
  q = q * (d / np.abs(d))[..., None, :]
  assert rel_max(wr.cpu().numpy() + 1j * wi.cpu().numpy(), q @ V) < 1e-11


def test_trajectory_collection_and_training_reduce_loss():
  from jrystal_amd import learned
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(4.0)
  ds = learned.collect_trajectory(wl, seed=7, snap_every=4, samples=256,
                                  ws_samples=1024, max_iter=120)
  assert len(ds.t2) >= 3 and ds.x_pc.shape == (len(ds.t2) * 256, 8)
  assert np.all(np.isfinite(ds.a)) and np.all(ds.b >= 0)
  net, hist = learned.train(ds, epochs=60, lr=3e-3, val_frac=0)
  net0, hist0 = learned.train(ds, epochs=0, val_frac=0)
  assert hist[-1][1] < hist0[-1][1] and hist[-1][2] < hist0[-1][2]
  tr, va = ds.split(0.25, seed=1)
  assert len(tr.t2) + len(va.t2) == len(ds.t2) and len(va.t2) >= 1
  assert tr.x_pc.shape[0] == len(tr.t2) * ds.snap


def _strained_cell():
  from jrystal_amd import workloads as W
  return W.si8_strained(4.0, k=2, seed=11)  # not the held-out strain (4)


def test_trained_network_gains_over_plain_preconditioner():
  """The shipped model against plain TPA on a strained 8-atom cell (8 k):
  same minimum, fewer CG iterations to |dE| < 1e-6."""
  from jrystal_amd import learned
  from jrystal_amd.minimize import Minimizer
  from jrystal_amd.objective import EnergyPlan
  wl = _strained_cell()
  plan = EnergyPlan.from_workload(wl)
  net = learned.PrecondNet.load()
  plain = Minimizer(plan, wl.params(0)).run(max_iter=2000)
  lrn = Minimizer(plan, wl.params(0), factor_fn=learned.LearnedFactor(plan, net),
                  restart=learned.LEARNED_RESTART,
                  first_restart=learned.LEARNED_FIRST_RESTART).run(max_iter=2000)
  e_star = min(plain.energy, lrn.energy)
  assert abs(plain.energy - lrn.energy) < 1e-7

  def its(r):
    return next(i for i, (t, e) in enumerate(r.history) if e - e_star < 1e-6)
  print(f"iterations to 1e-6: plain {its(plain)}, learned {its(lrn)}")
  assert its(lrn) <= 0.8 * its(plain)
  e = [h[1] for h in lrn.history]
  assert all(b <= a + 1e-12 for a, b in zip(e, e[1:]))


@pytest.mark.parametrize("mode", ["learned_secant", "plain_no_restart"])
def test_other_schedules_reach_the_minimum(mode):
  """The secant line search with the learned factor (its slope at the trial
  point is often just below the slope at 0, where the secant has no
  positive root) and CG without periodic restarts (stalls restart it) end
  at the parabola-search minimum."""
  from jrystal_amd import learned
  from jrystal_amd.minimize import Minimizer
  from jrystal_amd.objective import EnergyPlan
  wl = _strained_cell()
  plan = EnergyPlan.from_workload(wl)
  ref = Minimizer(plan, wl.params(0)).run(max_iter=2000)
  if mode == "learned_secant":
    m = Minimizer(plan, wl.params(0), line_search="secant",
                  factor_fn=learned.LearnedFactor(plan,
                                                  learned.PrecondNet.load()),
                  restart=learned.LEARNED_RESTART)
  else:
    m = Minimizer(plan, wl.params(0), restart=10**9)
  r = m.run(max_iter=3000)
  print(mode, r.iterations, r.status, r.energy - ref.energy)
  assert r.status in ("converged", "stalled")
  assert abs(r.energy - ref.energy) < 1e-7
  e = [h[1] for h in r.history]
  assert all(b <= a + 1e-12 for a, b in zip(e, e[1:]))


def test_fallback_to_plain_preconditioner():
  """A factor that breaks the search direction (NaN) sends the minimiser
  back to the plain preconditioner; it still reaches the plain minimum with
  a monotone energy."""
  from jrystal_amd.minimize import Minimizer
  from jrystal_amd.objective import EnergyPlan
  wl = _strained_cell()
  plan = EnergyPlan.from_workload(wl)
  bad = torch.full(plan.param_shape, float("nan"), dtype=torch.float64,
                   device="cuda")

  def factor_fn(m):
    return bad if m.restarts >= 2 else None

  ref = Minimizer(plan, wl.params(0)).run(max_iter=2000)
  r = Minimizer(plan, wl.params(0), factor_fn=factor_fn).run(max_iter=4000)
  assert r.fallbacks >= 1
  assert abs(r.energy - ref.energy) < 1e-6
  e = [h[1] for h in r.history]
  assert all(np.isfinite(e))
  assert all(b <= a + 1e-12 for a, b in zip(e, e[1:]))
