"""Checkpoint files (jrystal_amd.checkpoint) on the host: bitwise round trip
of every array, rejection of foreign files, no partial files left behind,
and the convergence window of a resumed run (rebuilt from the saved energy
history) deciding exactly as the uninterrupted window does."""
import os

import numpy as np
import pytest

from jrystal_amd import checkpoint
from jrystal_amd.optim import ConvergenceChecker


def _state(seed=0):
  rng = np.random.default_rng(seed)
  shp = (1, 2, 7, 3)
  w = {"w_re": rng.standard_normal(shp), "w_im": rng.standard_normal(shp)}
  adam = {"count": 11,
          "m": {k: rng.standard_normal(shp) for k in w},
          "v": {k: rng.random(shp) for k in w}}
  occ = {"occ.up": rng.standard_normal((2, 3))}
  return w, adam, occ


def test_round_trip_is_bitwise(tmp_path):
  w, adam, occ = _state()
  path = str(tmp_path / "ck" / "checkpoint.npz")
  hist = np.random.default_rng(1).standard_normal(11)
  checkpoint.save(path, step=11, params=w, occupation=np.full((1, 2, 3), .5),
                  energy_history=hist, adam=adam, occ_params=occ,
                  meta={"crystal": "si", "seed": 123, "kind": "final"})
  assert sorted(os.listdir(tmp_path / "ck")) == ["checkpoint.npz"]
  r = checkpoint.load(path)
  assert r["step"] == 11 and r["adam"]["count"] == 11
  for k in w:
    assert np.array_equal(r["params"][k], w[k])
    assert r["params"][k].dtype == np.float64
    assert np.array_equal(r["adam"]["m"][k], adam["m"][k])
    assert np.array_equal(r["adam"]["v"][k], adam["v"][k])
  assert np.array_equal(r["energy_history"], hist)
  assert np.array_equal(r["occ_params"]["occ.up"], occ["occ.up"])
  assert np.array_equal(r["occupation"], np.full((1, 2, 3), .5))
  assert r["meta"] == {"crystal": "si", "seed": 123, "kind": "final"}


def test_overwrite_and_optional_parts(tmp_path):
  w, _, _ = _state(2)
  path = str(tmp_path / "c.npz")
  checkpoint.save(path, step=1, params=w)
  w2, _, _ = _state(3)
  checkpoint.save(path, step=2, params=w2)
  r = checkpoint.load(path)
  assert r["step"] == 2 and r["adam"] is None and r["occupation"] is None
  assert np.array_equal(r["params"]["w_re"], w2["w_re"])
  assert sorted(os.listdir(tmp_path)) == ["c.npz"]


def test_foreign_file_rejected(tmp_path):
  path = str(tmp_path / "x.npz")
  np.savez(path, w_re=np.zeros(3))
  with pytest.raises(ValueError, match="checkpoint"):
    checkpoint.load(path)


def test_restored_convergence_window_decides_as_uninterrupted():
  rng = np.random.default_rng(5)
  e = np.concatenate([rng.standard_normal(30), 1e-6 * rng.standard_normal(40)])
  for cut in (0, 7, 20, 35, 55):
    full = ConvergenceChecker(20, 1e-5)
    a = [full.check(x) for x in e]
    first = ConvergenceChecker(20, 1e-5)
    b = [first.check(x) for x in e[:cut]]
    resumed = ConvergenceChecker(20, 1e-5)
    resumed.history = list(e[:cut])[-resumed.window_size:]
    b += [resumed.check(x) for x in e[cut:]]
    assert a == b, cut
