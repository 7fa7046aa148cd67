"""Batched Adam steps over uneven structures: the native batch step
against the oracle's Adam chain per structure and against a per-structure
loop of the fused evaluation + Adam, hipGraph replay against eager steps,
stream splitting, and the autotuner."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _members():
  from jrystal_amd import workloads as W
  a = W.tiny()
  b = W.build("si2-16^3-17b", W.fcc_si2(14.0), (16, 16, 16), (1, 1, 2), 17,
              11.0)
  c = W.batch_cells(seed=3, n=1)[0]  # fcc3, 42^3, Gamma, 47 bands
  return [a, b, c]


def _params(wls):
  return [wl.params(i + 1) for i, wl in enumerate(wls)]


def test_batch_step_matches_oracle_adam_chains():
  """Three structures with different grids (12^3, 16^3, 18^3), k-meshes (2,
  2 and 1 k-points) and band counts (3, 17, 5) in one batch: three batch
  steps against three steps of the oracle's optax-Adam restatement on each
  structure's energy and gradient (energies within 1e-8 Ha of the last
  evaluation, parameters within 1e-9 relative)."""
  from helpers import oracle_problem, rel_max
  from oracle import adjoint as OA
  from oracle import optim as OOpt
  from jrystal_amd import workloads as W
  from jrystal_amd.batch import BatchStep
  wls = [W.tiny(),
         W.build("si2-16^3-17b", W.fcc_si2(14.0), (16, 16, 16), (1, 1, 2), 17,
                 11.0),
         W.build("si2-18^3-gamma-5b", W.fcc_si2(14.0), (18, 18, 18),
                 (1, 1, 1), 5, 10.0)]
  ps = _params(wls)
  bs = BatchStep(wls, ps, learning_rate=0.01, b1=0.9, b2=0.99, streams=2)
  for _ in range(3):
    bs.step()
  torch.cuda.synchronize()
  e_batch = bs.energies()
  names = ("kinetic", "hartree", "external", "xc", "total")
  for i, wl in enumerate(wls):
    print(f"member {wl.name}", flush=True)
    prob = oracle_problem(wl.crystal, wl.grid_sizes, wl.k_grid, wl.num_bands,
                          cutoff=wl.cutoff_energy)
    assert np.array_equal(prob.mask, wl.mask)
    np.testing.assert_array_equal(prob.occupation, wl.occupation)
    p = {k: np.array(v) for k, v in ps[i].items()}
    opt = OOpt.Adam(p, learning_rate=0.01, b1=0.9, b2=0.99)
    for _ in range(3):
      parts, grad, _ = OA.energy_and_grad(p, prob)
      p = opt.step(p, grad)
    for j, k in enumerate(names):
      assert abs(e_batch[i][j] - parts[k]) <= 1e-8, (wl.name, k)
    for k in ("w_re", "w_im"):
      assert rel_max(bs.params(i)[k].cpu().numpy(), p[k]) <= 1e-9, (wl.name,
                                                                   k)


def test_batch_step_matches_member_loop():
  from jrystal_amd.batch import BatchStep
  from jrystal_amd.objective import EnergyPlan
  from jrystal_amd.optim import Adam
  wls = _members()
  ps = _params(wls)
  bs = BatchStep(wls, ps, learning_rate=0.01, b1=0.9, b2=0.99, streams=2)
  for _ in range(3):
    bs.step()
  torch.cuda.synchronize()
  assert bs.step_count() == 3
  e_batch = bs.energies()
  for i, wl in enumerate(wls):
    print(f"member {wl.name}", flush=True)
    plan = EnergyPlan.from_workload(wl)
    pt = {k: torch.tensor(v, device="cuda") for k, v in ps[i].items()}
    opt = Adam(pt, learning_rate=0.01, b1=0.9, b2=0.99)
    out = {}
    for _ in range(3):
      plan.evaluate(pt["w_re"], pt["w_im"], grad=True, out=out)
      opt.step(pt, {"w_re": out["dw_re"], "w_im": out["dw_im"]})
    np.testing.assert_allclose(e_batch[i], out["energies"].cpu().numpy(),
                               rtol=1e-13, atol=1e-12)
    for k in ("w_re", "w_im"):
      np.testing.assert_allclose(bs.params(i)[k].cpu().numpy(),
                                 pt[k].cpu().numpy(), rtol=1e-12, atol=1e-14)


def test_graph_replay_matches_eager_steps():
  from jrystal_amd.batch import BatchStep
  from jrystal_amd.objective import EnergyPlan
  wls = _members()
  ps = _params(wls)
  eager = BatchStep(wls, ps, streams=3)
  graph = BatchStep(wls, ps, streams=3)
  graph.capture()
  torch.cuda.synchronize()
  assert graph.step_count() == 0  # capturing does not execute the step
  other = EnergyPlan.from_workload(wls[0])
  p0 = wls[0].params(9)
  for step in range(3):
    eager.step()
    graph.replay()
    # unrelated eager work between replays
    other.evaluate(p0["w_re"], p0["w_im"], grad=True)
    x = torch.randn(1000, device="cuda").sum()
    torch.cuda.synchronize()
    assert np.isfinite(float(x))
    np.testing.assert_array_equal(graph.energies(), eager.energies())
    for i in range(len(wls)):
      for k in ("w_re", "w_im"):
        assert torch.equal(graph.params(i)[k], eager.params(i)[k]), (step, i)
  assert graph.step_count() == 3 == eager.step_count()


def test_stream_split_is_bitwise_neutral():
  from jrystal_amd.batch import BatchStep
  wls = _members()
  ps = _params(wls)
  a = BatchStep(wls, ps, streams=0)
  b = BatchStep(wls, ps, streams=4)
  for _ in range(2):
    a.step()
    b.step()
  torch.cuda.synchronize()
  np.testing.assert_array_equal(a.energies(), b.energies())
  for i in range(len(wls)):
    assert torch.equal(a.params(i)["w_re"], b.params(i)["w_re"])


def test_autotune_variants_agree_and_cache(tmp_path):
  from jrystal_amd import autotune
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import EnergyPlan
  wl = W.build("si2-32^3-4k-16b", W.fcc_si2(4.0), (32, 32, 32), (2, 1, 2),
               16, 24.0)
  assert len(autotune.candidates(wl)) == 5  # incl. two chunked variants
  p = wl.params(4)
  ref = None
  for opt in autotune.candidates(wl):
    plan = EnergyPlan.from_workload(wl, **opt)
    r = plan.evaluate(p["w_re"], p["w_im"], grad=True)
    e = r["energies"].cpu().numpy()
    g = r["dw_re"].cpu().numpy()
    if ref is None:
      ref = (e, g)
    else:
      np.testing.assert_allclose(e, ref[0], rtol=0, atol=1e-11)
      assert np.max(np.abs(g - ref[1])) <= 1e-12 * np.max(np.abs(ref[1]))
  cache = str(tmp_path / "tune.json")
  best = autotune.tune(wl, cache_path=cache, reps=2)
  assert best in autotune.candidates(wl)
  assert autotune.tune(wl, cache_path=cache) == best  # from the cache
