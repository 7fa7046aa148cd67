"""GPU driver pieces against the oracle: the fused Adam kernel, an
optimisation trajectory, and the CLI end to end."""
import numpy as np
import pytest

from helpers import oracle_problem
from oracle import adjoint as OA
from oracle import optim as OOpt

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_adam_kernel_matches_optax_rule():
  from jrystal_amd.optim import Adam
  rng = np.random.default_rng(0)
  p0 = {"a": rng.standard_normal((7, 33)), "b": rng.standard_normal(5)}
  pt = {k: torch.tensor(v, device="cuda") for k, v in p0.items()}
  opt = Adam(pt, learning_rate=0.01, b1=0.9, b2=0.99)
  ref = OOpt.Adam(p0, learning_rate=0.01, b1=0.9, b2=0.99)
  pr = p0
  for _ in range(4):
    g = {k: rng.standard_normal(v.shape) for k, v in p0.items()}
    opt.step(pt, {k: torch.tensor(v, device="cuda") for k, v in g.items()})
    pr = ref.step(pr, g)
  for k in p0:
    np.testing.assert_allclose(pt[k].cpu().numpy(), pr[k], rtol=0, atol=1e-14)


def test_adam_trajectory_matches_oracle():
  """Ten Adam steps of the fused evaluation follow the oracle's steps."""
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import EnergyPlan
  from jrystal_amd.optim import Adam
  wl = W.si2_fcc(4.0)
  prob = oracle_problem(wl.crystal, wl.grid_sizes, wl.k_grid, wl.num_bands,
                        wl.cutoff_energy)
  p_np = wl.params(0)
  plan = EnergyPlan.from_workload(wl)
  pt = {k: torch.tensor(v, device="cuda") for k, v in p_np.items()}
  opt = Adam(pt, learning_rate=0.01, b1=0.9, b2=0.99)
  ref = OOpt.Adam(p_np, learning_rate=0.01, b1=0.9, b2=0.99)
  out = {}
  for step in range(10):
    plan.evaluate(pt["w_re"], pt["w_im"], grad=True, out=out)
    opt.step(pt, {"w_re": out["dw_re"], "w_im": out["dw_im"]})
    parts, grad, _ = OA.energy_and_grad(p_np, prob)
    p_np = ref.step(p_np, grad)
    e = float(out["energies"][4])
    assert abs(e - parts["total"]) < 1e-8, (step, e, parts["total"])


def test_cli_energy_run(tmp_path):
  from jrystal_amd.__main__ import main
  from jrystal_amd.calc.energy_all_electrons import calc_with_output
  from jrystal_amd import config
  p = tmp_path / "c.yaml"
  p.write_text("crystal: si\ngrid_sizes: 16\nk_grid_sizes: 1\n"
               "cutoff_energy: 10\nempty_bands: 2\nepoch: 30\n"
               "verbose: false\nseed: 1\n")
  assert main(["-m", "energy", "-c", str(p)]) == 0
  out = calc_with_output(config.get_config(str(p)))
  assert out.steps == 30 and len(out.total_energy_history) == 30
  rho = out.density.cpu().numpy()
  vol = out.crystal.vol
  # uniform occupation: 2/nk on ceil(28/2) = 14 of 16 bands, nk = 1
  np.testing.assert_allclose(rho.sum() * vol / rho[0].size, 28.0, rtol=1e-12)
  assert out.total_energy_history[-1] < out.total_energy_history[0]


@pytest.mark.parametrize("opt", ["cg", "cg-learned"])
def test_cli_cg_optimisers_reach_lower_energy_than_adam(tmp_path, opt):
  from jrystal_amd import config
  from jrystal_amd.calc.energy_all_electrons import calc_with_output
  base = ("crystal: si\ngrid_sizes: 16\nk_grid_sizes: 1\ncutoff_energy: 10\n"
          "empty_bands: 2\nverbose: false\nseed: 1\nsmearing: 0.0\n")
  pa, pc = tmp_path / "a.yaml", tmp_path / "c.yaml"
  pa.write_text(base + "epoch: 300\n")
  pc.write_text(base + f"epoch: 300\noptimizer: {opt}\n")
  adam = calc_with_output(config.get_config(str(pa)))
  cg = calc_with_output(config.get_config(str(pc)))
  e = cg.total_energy_history
  assert all(b <= a + 1e-12 for a, b in zip(e, e[1:]))
  assert cg.total_energy < adam.total_energy
  rho = cg.density.cpu().numpy()
  np.testing.assert_allclose(rho.sum() * cg.crystal.vol / rho[0].size, 28.0,
                             rtol=1e-12)
