"""Checkpoint and resume of the energy driver on the GPU: an Adam run
interrupted at a checkpoint and resumed with -l ends with the same
parameters, energies and history, bit for bit, as the uninterrupted run
(fixed and parameterised occupations); the learned preconditioner's
trajectories are checkpoint files that the trainer reads back."""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# smearing 0: the annealing schedule depends on the run's epoch count
# (transition steps epoch // 2, as in the reference), so only a run at zero
# temperature splits into two runs of different epoch counts unchanged
BASE = ("crystal: si\ngrid_sizes: 16\nk_grid_sizes: 1\ncutoff_energy: 10\n"
        "empty_bands: 2\nverbose: false\nseed: 1\nsmearing: 0.0\n")


@pytest.mark.parametrize("occupation", ["uniform", "simplex-projector"])
def test_resume_is_bitwise(tmp_path, occupation):
  from jrystal_amd import checkpoint, config
  from jrystal_amd.calc.energy_all_electrons import calc_with_output
  text = BASE + f"occupation: {occupation}\n"
  full = tmp_path / "full.yaml"
  full.write_text(text + "epoch: 24\n")
  part = tmp_path / "part.yaml"
  save = tmp_path / "ck"
  part.write_text(text + f"epoch: 10\nsave_dir: {save}\ncheckpoint_every: 5\n")
  a = calc_with_output(config.get_config(str(full)))
  b1 = calc_with_output(config.get_config(str(part)))
  ck = checkpoint.load(str(save / "checkpoint.npz"))
  assert ck["step"] == 10 and ck["adam"]["count"] == 10
  assert len(ck["energy_history"]) == 10
  assert np.array_equal(ck["params"]["w_re"],
                        b1.params_pw["w_re"].cpu().numpy())
  rest = tmp_path / "rest.yaml"
  rest.write_text(text + "epoch: 24\n")
  b = calc_with_output(config.get_config(str(rest)),
                       resume=str(save / "checkpoint.npz"))
  assert b.steps == a.steps == 24
  assert b.total_energy_history == a.total_energy_history
  for k in ("w_re", "w_im"):
    assert torch.equal(a.params_pw[k], b.params_pw[k]), k
  assert a.total_energy == b.total_energy


def test_resume_rejects_other_shapes(tmp_path):
  from jrystal_amd import config
  from jrystal_amd.calc.energy_all_electrons import calc_with_output
  p = tmp_path / "a.yaml"
  save = tmp_path / "ck"
  p.write_text(BASE + f"epoch: 2\nsave_dir: {save}\n")
  calc_with_output(config.get_config(str(p)))
  q = tmp_path / "b.yaml"
  q.write_text(BASE.replace("empty_bands: 2", "empty_bands: 3") + "epoch: 4\n")
  with pytest.raises(ValueError, match="shape"):
    calc_with_output(config.get_config(str(q)),
                     resume=str(save / "checkpoint.npz"))


def test_trajectory_files_feed_training(tmp_path):
  from jrystal_amd import checkpoint, learned
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(4.0)
  d = tmp_path / "traj"
  ds = learned.collect_trajectory(wl, seed=7, snap_every=4, samples=128,
                                  ws_samples=256, max_iter=60,
                                  out_dir=str(d))
  snaps = sorted(glob.glob(str(d / "snap_*.npz")))
  assert len(snaps) >= 3 and os.path.exists(d / "final.npz")
  kinds = {checkpoint.load(f)["meta"]["kind"] for f in snaps}
  assert kinds == {"snapshot"}
  again = learned.trajectory_dataset(wl, str(d / "start.npz"), snaps,
                                     str(d / "final.npz"), samples=128,
                                     ws_samples=256,
                                     rng=np.random.default_rng(7))
  for k in ("x_pc", "a", "b", "t2", "x_ws", "y_ws"):
    assert np.array_equal(getattr(ds, k), getattr(again, k)), k
  assert len(ds.t2) >= 1 and np.all(np.isfinite(ds.a))
