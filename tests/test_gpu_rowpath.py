"""The row-buffer FFT path (cubic 64^3 and 128^3 boxes) against the oracle and
against the plane-buffer path it replaced (fft_variant 2): energies,
density, gradient, band energies, two spin channels, chunked sweeps, state
groups left empty by the split of the states, and a 128^3 Gamma-point cell
of the 64-atom supercell.

Tolerances as the other parity tests: total energy 1e-8 Ha, density
max-abs 1e-10, gradient 1e-7 relative to its largest entry; the two HIP
paths agree far tighter (1e-11)."""
import dataclasses

import numpy as np
import pytest

from helpers import oracle_problem, rel_max, seeded_params
from oracle import adjoint as OA

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

NAMES = ("kinetic", "hartree", "external", "xc", "total")


def _plan(wl, **kw):
  from jrystal_amd.objective import EnergyPlan
  return EnergyPlan.from_workload(wl, **kw)


def _run(plan, p, band=False):
  if band:
    plan.set_band_energies(True)
  out = plan.evaluate(torch.tensor(p["w_re"], device="cuda"),
                      torch.tensor(p["w_im"], device="cuda"), grad=True,
                      density=True)
  res = {k: v.cpu().numpy() for k, v in out.items()}
  if band:
    res["band"] = plan.band_energies().cpu().numpy()
  return res


def _check_oracle(out, wl, p, e_tol=1e-8, rho_tol=1e-10, g_tol=1e-7,
                  spin_restricted=True):
  prob = oracle_problem(wl.crystal, wl.grid_sizes, wl.k_grid, wl.num_bands,
                        wl.cutoff_energy, spin_restricted=spin_restricted)
  prob = dataclasses.replace(prob, occupation=wl.occupation)
  parts, grad, rho = OA.energy_and_grad(p, prob)
  for i, k in enumerate(NAMES):
    assert abs(out["energies"][i] - parts[k]) <= e_tol, (k, out["energies"][i],
                                                         parts[k])
  assert np.max(np.abs(out["density"] - rho)) <= rho_tol
  assert rel_max(out["dw_re"], grad["w_re"]) <= g_tol
  assert rel_max(out["dw_im"], grad["w_im"]) <= g_tol
  return prob


def _compare(a, b, tol=1e-11):
  assert np.max(np.abs(a["energies"] - b["energies"])) <= tol
  assert np.max(np.abs(a["density"] - b["density"])) <= tol * 1e-1
  for k in ("dw_re", "dw_im"):
    assert rel_max(a[k], b[k]) <= tol


def test_row_path_si8_64_against_oracle_and_plane_path():
  from jrystal_amd import workloads as W
  wl = W.si8_cube(14.0, k=2)
  p = seeded_params(wl.shape, 7)
  row = _plan(wl)
  assert row.row_path
  plane = _plan(wl, fft_variant=2)
  assert not plane.row_path
  a = _run(row, p, band=True)
  b = _run(plane, p, band=True)
  print(f"E row {a['energies'][4]:.12f} plane {b['energies'][4]:.12f}",
        flush=True)
  _compare(a, b)
  assert rel_max(a["band"], b["band"]) <= 1e-11
  prob = _check_oracle(a, wl, p)
  ref = OA.band_energies(p, prob)
  assert rel_max(a["band"], ref) <= 1e-9


def test_row_path_two_spins_and_chunks():
  from jrystal_amd import occupation, workloads as W
  wl = W.si8_cube(4.0, k=1)
  occ = occupation.uniform(1, wl.crystal.num_electron, 0, wl.num_bands,
                           spin_restricted=False)
  wl2 = dataclasses.replace(wl, occupation=occ)
  shape = (2, 1, wl.num_g, wl.num_bands)
  p = seeded_params(shape, 9)
  full = _run(_plan(wl2), p)
  chunked = _plan(wl2, chunk_states=5)
  assert chunked.row_path and not chunked.keep_planes
  _compare(full, _run(chunked, p))
  _check_oracle(full, wl2, p, spin_restricted=False)


def test_row_path_empty_state_groups_repeated():
  """33 states at one k-point: 32 z-plane state groups of 2 states leave
  groups 17-31 empty, whose density planes the pass writes as zeros.  Two
  evaluations with different inputs (the second overwrites the first's
  partial planes) against the plane path and the oracle."""
  from jrystal_amd import workloads as W
  wl = W.build("si8-64^3-gamma-33b", W.diamond_cube(4.0), (64,) * 3,
               (1, 1, 1), 33, 30.0)
  row = _plan(wl)
  plane = _plan(wl, fft_variant=2)
  assert row.row_path and not plane.row_path
  for seed in (3, 4):
    p = seeded_params(wl.shape, seed)
    a = _run(row, p)
    _compare(a, _run(plane, p))
  _check_oracle(a, wl, p)


def test_row_path_128_gamma_si64():
  """The 64-atom supercell (a = 20.52 bohr) at 128^3, Gamma only, 16 bands
  (the row path at 128 points: 63 active x-planes)."""
  from jrystal_amd import workloads as W
  wl = W.build("si64-128^3-gamma-16b", W.diamond_cube(4.0, reps=2),
               (128,) * 3, (1, 1, 1), 16, 48.0)
  assert wl.num_g == 137059
  plan = _plan(wl)
  assert plan.row_path and plan.active_planes == 63
  p = seeded_params(wl.shape, 11)
  out = _run(plan, p)
  _check_oracle(out, wl, p)
  with pytest.raises(NotImplementedError, match="row-buffer path"):
    _plan(wl, fft_variant=2)  # a 128 x 128 plane does not fit the LDS


def test_row_path_more_than_32_active_planes():
  """The strained 8-atom cell has 35 active x-planes at 64^3: its z-plane
  passes run 512-thread blocks (64 column lane groups)."""
  from jrystal_amd import workloads as W
  wl = W.si8_strained(4.0, k=2)
  plan = _plan(wl)
  assert plan.row_path and plan.active_planes > 32
  p = seeded_params(wl.shape, 13)
  out = _run(plan, p, band=True)
  _check_oracle(out, wl, p)
  ref = _run(_plan(wl, fft_variant=2), p, band=True)
  _compare(out, ref)
  assert rel_max(out["band"], ref["band"]) <= 1e-11
