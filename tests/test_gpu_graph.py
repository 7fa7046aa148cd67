"""One fused evaluation captured in a HIP graph (EnergyPlan.capture) and
replayed: every replay, with eager work and synchronisations between the
replays and new parameters copied into the captured input tensors, equals
the eager evaluation bit for bit."""
import numpy as np
import pytest

from helpers import seeded_params

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("name", ["si2", "si8_k1"])
def test_graph_replay_equals_eager(name):
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import EnergyPlan
  wl = W.si2_fcc(14.0) if name == "si2" else W.si8_cube(4.0, k=1)
  plan = EnergyPlan.from_workload(wl)
  ps = [seeded_params(wl.shape, s) for s in (1, 2, 3)]
  wr = torch.tensor(ps[0]["w_re"], device="cuda")
  wi = torch.tensor(ps[0]["w_im"], device="cuda")
  graph, out = plan.capture(wr, wi)
  for p in ps + ps[:1]:
    wr.copy_(torch.as_tensor(p["w_re"]))
    wi.copy_(torch.as_tensor(p["w_im"]))
    graph.replay()
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy().copy() for k, v in out.items()}
    # eager work in between: an unrelated evaluation into fresh tensors
    ref = plan.evaluate(torch.as_tensor(p["w_re"]).cuda(),
                        torch.as_tensor(p["w_im"]).cuda(), grad=True)
    torch.cuda.synchronize()
    for k in ("energies", "dw_re", "dw_im"):
      assert np.array_equal(got[k], ref[k].cpu().numpy()), k
  plan.close()
