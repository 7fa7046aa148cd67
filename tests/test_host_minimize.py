"""Host-side logic of the minimiser, learned-preconditioner data path,
uneven batch generator and autotuner (CPU only, no kernels)."""
import numpy as np
import pytest

from oracle import mlp as OM


def test_alias_free_cutoff_reproduces_survey_values():
  from jrystal_amd import workloads as W
  # 32^3 FCC Si2 -> 24 Ha, 64^3 Si8 cube -> 48 Ha (the workloads' cutoffs)
  assert W.alias_free_cutoff(W.fcc_si2().cell_vectors, (32,) * 3) == 24.0
  assert W.alias_free_cutoff(W.diamond_cube().cell_vectors, (64,) * 3) == 48.0
  # the sphere fits |m_d| <= n/4 along every axis
  cell = np.diag([7.0, 9.0, 11.0])
  e = W.alias_free_cutoff(cell, (40, 40, 40))
  g = np.sqrt(2 * e)
  assert g <= 10 * 2 * np.pi / 11.0 + 1e-12


def test_batch_cells_are_seeded_valid_and_uneven():
  from jrystal_amd import workloads as W
  a = W.batch_cells(seed=3, n=16)
  b = W.batch_cells(seed=3, n=16)
  assert [x.name for x in a] == [x.name for x in b]
  grids = {x.grid_sizes[0] for x in a}
  assert len(grids) >= 8 and all(n in W.SMOOTH_32_80 for n in grids)
  for wl in a:
    ns, nk, ng, nb = wl.shape
    assert 4 <= nb <= 48 and nk in (1, 8, 27, 64) and ng > nb
    cr = wl.crystal
    assert 2 <= len(cr.positions) <= 16
    # minimum-image separation >= 2.5 bohr
    shifts = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1)
                       for k in (-1, 0, 1)], float) @ cr.cell_vectors
    d = cr.positions[None, :, None] - cr.positions[:, None, None] + shifts
    r = np.linalg.norm(d, axis=-1)
    n = len(cr.positions)
    r[np.arange(n), np.arange(n), 13] = np.inf
    assert r.min() >= 2.5
    # uniform spin-restricted occupation: 2 per band on min(N_e // 2, nb)
    # bands (the reference's occupation.uniform, which clips at nb)
    assert wl.occupation.sum() == pytest.approx(
      2 * min(int(cr.num_electron) // 2, nb))


def test_smooth_sizes_match_the_fft_table():
  from jrystal_amd import grid, workloads as W
  assert W.SMOOTH_32_80 == tuple(n for n in range(32, 81)
                                 if grid.fft_factor(n) == n)
  assert 37 not in W.SMOOTH_32_80 and 75 in W.SMOOTH_32_80


def test_strain_matrix_is_symmetric_and_bounded():
  from jrystal_amd import workloads as W
  e = W.strain_matrix(4, 0.1)
  assert np.allclose(e, e.T) and np.all(np.abs(e) <= 0.1)
  wl = W.si8_strained(4.0, k=1, seed=4)
  assert np.allclose(wl.crystal.cell_vectors, 10.26 * (np.eye(3) + e))


def test_grassmann_direction_matches_oracle():
  from jrystal_amd import learned
  rng = np.random.default_rng(0)
  a = rng.standard_normal((2, 30, 4)) + 1j * rng.standard_normal((2, 30, 4))
  qt = np.linalg.qr(a)[0]
  qs = np.linalg.qr(qt + 0.2 * rng.standard_normal(qt.shape))[0]
  np.testing.assert_allclose(learned.grassmann_direction(qt, qs),
                             OM.grassmann_direction(qt, qs), rtol=1e-13,
                             atol=1e-14)


def test_dataset_split_and_concat():
  from jrystal_amd.learned import Dataset
  rng = np.random.default_rng(1)
  snap, ns = 16, 10
  ds = Dataset(rng.random((ns * snap, 8), np.float32),
               rng.random(ns * snap, np.float32),
               rng.random(ns * snap, np.float32),
               rng.random(ns, np.float32), snap,
               rng.random((50, 8), np.float32), rng.random(50, np.float32))
  tr, va = ds.split(0.3, seed=2)
  assert len(tr.t2) + len(va.t2) == ns and len(va.t2) == 3
  assert tr.x_pc.shape[0] == len(tr.t2) * snap
  assert tr.x_ws.shape[0] + va.x_ws.shape[0] == 50
  both = Dataset.concat([tr, va])
  assert sorted(both.t2.tolist()) == sorted(ds.t2.tolist())


def test_feature_norm_standardises():
  from jrystal_amd.learned import Dataset, feature_norm
  rng = np.random.default_rng(3)
  x = rng.normal(2.0, 3.0, (400, 8)).astype(np.float32)
  x[:, 7] = 0.0  # an unused slot keeps unit scale
  ds = Dataset(x[:320], np.zeros(320, np.float32), np.ones(320, np.float32),
               np.ones(20, np.float32), 16, x[320:], np.zeros(80, np.float32))
  norm = feature_norm(ds)
  z = (x - norm[:8]) * norm[8:]
  np.testing.assert_allclose(z[:, :7].mean(0), 0.0, atol=1e-5)
  np.testing.assert_allclose(z[:, :7].std(0), 1.0, atol=1e-4)
  assert norm[15] == 1.0


def test_autotune_candidates_and_keys():
  from jrystal_amd import autotune, workloads as W
  wl = W.si8_cube(4.0, k=2)
  c = autotune.candidates(wl)
  assert c[0] == autotune.candidates_default()
  assert {"chunk_states": 0, "fft_variant": 1, "xline_blocks": 0} in c
  assert any(o["chunk_states"] > 0 for o in c)
  assert autotune.shape_key(wl) == "64x64x64-g17071-s1-k8-b32"
