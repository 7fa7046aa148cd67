"""Norm-conserving pseudopotential energies on the GPU against the oracle:
the local potential in place of the Coulomb nuclei
(jrx_plan_set_external_potential) and the non-local projectors
(jrx_plan_set_nonlocal) in the fused evaluation, on the plane path (Si2 and
C2 at 16^3) and the 64^3 row path; energies within 1e-8 Ha, density within
1e-10, gradient within 1e-7 relative; per-band non-local and band energies;
switching both off restores the all-electron evaluation bitwise; the
norm-conserving driver's Adam chain against the oracle's."""
import dataclasses
import os

import numpy as np
import pytest

from helpers import rel_max, seeded_params
from oracle import adjoint as OA
from oracle import grid as OG
from oracle import occupation as OO
from oracle import pseudopotential as OPP
from oracle.crystal import Crystal as OCrystal

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PPDIR = os.path.join(ROOT, "jrystal_amd", "data", "pseudopotential",
                     "normconserving")
GEOM = os.path.join(ROOT, "jrystal_amd", "geometry")
NAMES = ("kinetic", "hartree", "external", "xc", "total")


def _case(geom, gs, ecut, kgrid, nb, seed=0):
  from jrystal_amd import pseudopotential as PS
  from jrystal_amd.crystal import Crystal
  from jrystal_amd.objective import EnergyPlan
  path = geom if geom.endswith(".xyz") else os.path.join(GEOM, geom + ".xyz")
  cr = Crystal.create_from_file(path)
  ocr = OCrystal.create_from_file(path)
  g = OG.g_vectors(ocr.cell_vectors, list(gs))
  mask = OG.spherical_mask(ocr.cell_vectors, list(gs), ecut)
  kpts = OG.k_vectors(ocr.cell_vectors, list(kgrid))
  opp = OPP.create(ocr.symbols, ocr.positions, PPDIR)
  nk = kpts.shape[0]
  occ = OO.uniform(nk, sum(opp.valence_charges), 0, nb)
  prob = OA.Problem(mask=mask, g_vector_grid=g, kpts=kpts, vol=ocr.vol,
                    occupation=occ, positions=ocr.positions,
                    charges=np.asarray(ocr.charges, float),
                    v_ext_rec=OPP.local_potential_reciprocal(opp, g, ocr.vol),
                    projectors=OPP.projectors(opp, g, mask, kpts))
  pp = PS.NormConservingPseudopotential.create(cr)
  plan = EnergyPlan(cr.positions, cr.charges, cr.cell_vectors, mask, kpts, occ)
  plan.set_external_potential(PS.potential_local_reciprocal(pp, g, cr.vol))
  plan.set_nonlocal(PS.potential_nonlocal_psi_reciprocal(pp, g, mask, kpts))
  params = seeded_params((1, nk, int(mask.sum()), nb), seed)
  print(f"case {geom} {gs} ecut={ecut} k={kgrid} nb={nb} "
        f"row_path={plan.row_path} nproj={plan.num_projectors}", flush=True)
  return plan, prob, params


def _eval(plan, p, density=True):
  out = plan.evaluate(torch.tensor(p["w_re"], device="cuda"),
                      torch.tensor(p["w_im"], device="cuda"), grad=True,
                      density=density)
  torch.cuda.synchronize()
  return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("geom,gs,ecut,kgrid,nb", [
  ("si", (16, 16, 16), 12.0, (2, 1, 1), 6),
  ("diamond", (16, 16, 16), 20.0, (1, 2, 1), 5),
  ("si", (64, 64, 64), 30.0, (1, 1, 1), 8),
])
def test_normcons_energy_density_gradient(geom, gs, ecut, kgrid, nb):
  plan, prob, p = _case(geom, gs, ecut, kgrid, nb)
  res = _eval(plan, p)
  parts, grad, rho = OA.energy_and_grad(p, prob)
  print("external: local", parts["external_local"], "nonlocal",
        parts["external_nonlocal"], "gpu", res["energies"][2], flush=True)
  for i, k in enumerate(NAMES):
    assert abs(res["energies"][i] - parts[k]) <= 1e-8, (k, res["energies"][i],
                                                         parts[k])
  assert np.max(np.abs(res["density"] - rho)) <= 1e-10
  assert rel_max(res["dw_re"], grad["w_re"]) <= 1e-7
  assert rel_max(res["dw_im"], grad["w_im"]) <= 1e-7
  # per-band non-local energies, and the band energies that include them
  q = np.linalg.qr(p["w_re"] + 1j * p["w_im"])[0]
  e_nl = OPP.energy_nonlocal(q, prob.projectors, prob.occupation, prob.vol,
                             False)[2]
  assert rel_max(plan.nonlocal_energies().cpu().numpy(), e_nl) <= 1e-9
  plan.set_band_energies(True)
  _eval(plan, p, density=False)
  be = plan.band_energies().cpu().numpy()
  assert rel_max(be, OA.band_energies(p, prob)) <= 1e-9


def test_normcons_many_projectors_two_pair_blocks(tmp_path):
  """Si16 (si8 doubled along x): 64 projectors (two tiles of 32) x 40 bands
  = 2560 (projector, band) pairs, two pair blocks of the projection with
  one g slice each, and 27 chunks of 32 g (855 coefficients, one k-point)
  summed by the one-wave-per-output reduction (more than 16 chunks)."""
  lines = open(os.path.join(GEOM, "si8.xyz")).read().splitlines()
  a = float(lines[1].split('"')[1].split()[0])
  atoms = [ln.split() for ln in lines[2:] if ln.strip()]
  out = [str(2 * len(atoms)),
         f'Lattice="{2 * a} 0.0 0.0 0.0 {a} 0.0 0.0 0.0 {a}" '
         'Properties=species:S:1:pos:R:3 pbc="T T T"']
  for shift in (0.0, a):
    out += [f"{s} {float(x) + shift} {y} {z}" for s, x, y, z in atoms]
  geom = str(tmp_path / "si16.xyz")
  open(geom, "w").write("\n".join(out) + "\n")
  plan, prob, p = _case(geom, (32, 16, 16), 4.0, (1, 1, 1), 40)
  assert plan.num_projectors == 64 and int(prob.mask.sum()) > 16 * 32
  res = _eval(plan, p)
  parts, grad, rho = OA.energy_and_grad(p, prob)
  print("si16 external_nonlocal", parts["external_nonlocal"], flush=True)
  for i, k in enumerate(NAMES):
    assert abs(res["energies"][i] - parts[k]) <= 1e-8, (k, res["energies"][i],
                                                         parts[k])
  assert np.max(np.abs(res["density"] - rho)) <= 1e-10
  assert rel_max(res["dw_re"], grad["w_re"]) <= 1e-7
  assert rel_max(res["dw_im"], grad["w_im"]) <= 1e-7


def test_switching_off_restores_all_electron():
  from jrystal_amd.objective import EnergyPlan
  plan, prob, p = _case("si", (16, 16, 16), 12.0, (2, 1, 1), 6)
  _eval(plan, p)
  plan.set_external_potential(None)
  plan.set_nonlocal(None)
  a = _eval(plan, p)
  ref = EnergyPlan(prob.positions, prob.charges,
                   np.asarray(OCrystal.create_from_file(
                     os.path.join(GEOM, "si.xyz")).cell_vectors),
                   prob.mask, prob.kpts, prob.occupation)
  b = _eval(ref, p)
  assert np.array_equal(a["energies"], b["energies"])
  assert np.array_equal(a["dw_re"], b["dw_re"])


def _nc_config(tmp_path, **over):
  from jrystal_amd import config as C
  cfg = C.get_config()
  cfg.update(crystal="si", grid_sizes=16, cutoff_energy=12.0, k_grid_sizes=1,
             empty_bands=2, band_structure_empty_bands=2, epoch=6,
             smearing=0.0, occupation="uniform", use_pseudopotential=True,
             pseudopotential_type="nc", optimizer="adam",
             optimizer_args={"learning_rate": 0.01, "b1": 0.9, "b2": 0.99},
             k_path_special_points="LGX", num_kpoints=4,
             band_structure_epoch=5, k_path_fine_tuning_epoch=3,
             save_dir=str(tmp_path), verbose=False)
  cfg.update(over)
  return cfg


def _nc_oracle_problem(cfg, kpts, nb, occ=None):
  from jrystal_amd import grid as G
  ocr = OCrystal.create_from_file(os.path.join(GEOM, cfg.crystal + ".xyz"))
  gs = list(G.proper_grid_size(cfg.grid_sizes))
  g = OG.g_vectors(ocr.cell_vectors, gs)
  mask = OG.spherical_mask(ocr.cell_vectors, gs, cfg.cutoff_energy)
  opp = OPP.create(ocr.symbols, ocr.positions, PPDIR)
  if occ is None:
    occ = OO.uniform(kpts.shape[0], sum(opp.valence_charges), 0, nb)
  return OA.Problem(mask=mask, g_vector_grid=g, kpts=kpts, vol=ocr.vol,
                    occupation=occ, positions=ocr.positions,
                    charges=np.asarray(ocr.charges, float),
                    v_ext_rec=OPP.local_potential_reciprocal(opp, g, ocr.vol),
                    projectors=OPP.projectors(opp, g, mask, kpts)), opp


def test_normcons_energy_driver_adam_chain(tmp_path):
  """calc_ground_state_energy_normcons: six Adam steps of the driver follow
  the oracle's chain (valence electrons: 8, so 4 + 2 bands)."""
  from jrystal_amd.calc.energy_all_electrons import calc_with_output
  from oracle import optim as OOpt
  from oracle import pw as OP
  cfg = _nc_config(tmp_path)
  out = calc_with_output(cfg)
  kpts = OG.k_vectors(out.crystal.cell_vectors, [1, 1, 1])
  prob, opp = _nc_oracle_problem(cfg, kpts, 6)
  p = OP.param_init(cfg.seed, 6, 1, prob.mask, dist="uniform")
  opt = OOpt.Adam(p, **cfg.optimizer_args)
  for _ in range(cfg.epoch):
    p = opt.step(p, OA.energy_and_grad(p, prob)[1])
  parts = OA.energy_and_grad(p, prob, want_grad=False)[0]
  for key in ("w_re", "w_im"):
    assert rel_max(out.params_pw[key].cpu().numpy(), p[key]) <= 1e-9
  e = out.energies
  print("normcons energies", e, flush=True)
  assert abs(e["external_nonlocal"] - parts["external_nonlocal"]) <= 1e-8
  assert abs(e["external_local"] - parts["external_local"]) <= 1e-8
  assert abs(e["total"] - parts["total"]) <= 1e-8


def test_normcons_band_driver_adam_chain(tmp_path):
  """calc_band_structure_normcons: the band driver with the local part in
  the fixed potential and each k-point's projectors, against the oracle."""
  import dataclasses as dc
  import importlib
  from oracle import band as OB
  from oracle import optim as OOpt
  from oracle import pw as OP
  BD = importlib.import_module("jrystal_amd.calc.band_all_electrons")
  cfg = _nc_config(tmp_path)
  ocr = OCrystal.create_from_file(os.path.join(GEOM, "si.xyz"))
  nb = 6
  prob0, _ = _nc_oracle_problem(cfg, np.zeros((1, 3)), nb)
  rho = OA.energy_and_grad(seeded_params((1, 1, int(prob0.mask.sum()), 4),
                                         5), dc.replace(
    prob0, occupation=OO.uniform(1, 8, 0, 4)), want_grad=False)[2]

  class GS:
    density = torch.tensor(rho)
    total_energy = 0.0
  out = BD.calc_with_output(cfg, ground_state=GS())
  prob, _ = _nc_oracle_problem(cfg, out.k_path, nb,
                               occ=np.ones((1, out.k_path.shape[0], nb)))
  v = OB.effective_potential(rho, prob)
  p = OP.param_init(cfg.seed, nb, 1, prob.mask, dist="uniform")
  opt = OOpt.Adam(p, **cfg.optimizer_args)
  evs = []
  for i, k in enumerate(out.k_path):
    pk = dc.replace(prob, kpts=k[None], occupation=np.ones((1, 1, nb)),
                    projectors=prob.projectors[i:i + 1])
    for _ in range(cfg.band_structure_epoch if i == 0
                   else cfg.k_path_fine_tuning_epoch):
      p = opt.step(p, OB.hamiltonian_trace(p, pk, v)[1])
    evs.append(OB.band_eigenvalues(p, pk, v)[:, 0])
  evs = np.stack(evs, axis=1)
  print("normcons band eigenvalues at k0:", out.eigenvalues[0, 0], flush=True)
  assert rel_max(out.eigenvalues, evs) <= 1e-9
