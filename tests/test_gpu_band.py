"""Kohn-Sham band mode on the GPU against the oracle (oracle/band.py):

* fixed-potential evaluations (jrx_plan_set_fixed_density): the trace
  sum f <psi|T + v_eff|psi>, its kinetic and potential parts and its
  gradient on the runtime-radix plane path (12^3), the 32^3 plane path and
  the 64^3 row path; spin-polarised LDA (the Kohn-Sham form of vxc_lda) and
  GGA; tolerances 1e-8 Ha and 1e-7 relative, as for the ground state;
* jrx_plan_set_kpoints against a plan built at the new k-points (same
  kernels: equal to rounding);
* q_i^H H q_j and its eigenvalues / eigenvectors (jrx_plan_subspace_*)
  against the oracle's Q^H H Q and numpy's eigvalsh, 1e-9 relative;
* leaving band mode restores the ground-state evaluation bitwise;
* the driver: Adam steps, the k-path warm start with the carried Adam state
  and the per-k eigenvalues against the same chain on the oracle, and (CG)
  convergence to the lowest eigenvalues of the dense plane-wave matrix.
"""
import dataclasses
import importlib

import numpy as np
import pytest

from helpers import oracle_problem, rel_max, seeded_params
from oracle import adjoint as OA
from oracle import band as OB

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _problem(wl, xc="lda_x", spin_restricted=True):
  prob = oracle_problem(wl.crystal, wl.grid_sizes, wl.k_grid, wl.num_bands,
                        wl.cutoff_energy, spin_restricted=spin_restricted)
  return dataclasses.replace(prob, xc=xc)


def _density(prob, seed):
  """A ground-state-like density: that of seeded orbitals."""
  ns, nk, nb = prob.occupation.shape
  p = seeded_params((ns, nk, int(prob.mask.sum()), nb), seed + 100)
  return OA.energy_and_grad(p, prob, want_grad=False)[2]


def _band_plan(wl, prob, rho, occ=None, xc="lda_x"):
  from jrystal_amd.objective import EnergyPlan
  occ = prob.occupation if occ is None else occ
  plan = EnergyPlan(wl.crystal.positions, wl.crystal.charges,
                    wl.crystal.cell_vectors, wl.mask, prob.kpts, occ, xc=xc)
  plan.set_fixed_density(torch.tensor(rho, device="cuda"))
  return plan


def _eval(plan, p, grad=True):
  out = plan.evaluate(torch.tensor(p["w_re"], device="cuda"),
                      torch.tensor(p["w_im"], device="cuda"), grad=grad)
  torch.cuda.synchronize()
  return {k: v.cpu().numpy() for k, v in out.items()}


def _check_trace(res, prob, p, v, e_tol=1e-8, g_tol=1e-7):
  parts, grad = OB.hamiltonian_trace(p, prob, v)
  e = res["energies"]
  print(f"trace {e[4]:.12f} oracle {parts['total']:.12f}", flush=True)
  assert abs(e[0] - parts["kinetic"]) <= e_tol
  assert abs(e[1] - parts["potential"]) <= e_tol
  assert e[2] == 0.0 and e[3] == 0.0
  assert abs(e[4] - parts["total"]) <= e_tol
  assert rel_max(res["dw_re"], grad["w_re"]) <= g_tol
  assert rel_max(res["dw_im"], grad["w_im"]) <= g_tol


@pytest.mark.parametrize("name", ["tiny", "si2_fcc", "si8_cube"])
def test_fixed_potential_trace_and_gradient(name):
  from jrystal_amd import workloads as W
  wl = {"tiny": lambda: W.tiny(), "si2_fcc": lambda: W.si2_fcc(4.0),
        "si8_cube": lambda: W.si8_cube(4.0, k=1)}[name]()
  prob = _problem(wl)
  rho = _density(prob, 1)
  v = OB.effective_potential(rho, prob)
  plan = _band_plan(wl, prob, rho)
  print(f"case {wl.name} row_path={plan.row_path}", flush=True)
  p = seeded_params(wl.shape, 2)
  tight = name == "tiny"
  _check_trace(_eval(plan, p), prob, p, v, e_tol=1e-11 if tight else 1e-8,
               g_tol=1e-11 if tight else 1e-7)


def test_fixed_potential_spin_polarised():
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(4.0)
  prob = _problem(wl, spin_restricted=False)
  assert prob.occupation.shape[0] == 2
  rho = _density(prob, 3)
  rho[1] *= 0.7  # unequal channels: the Kohn-Sham vxc_lda form matters
  v = OB.effective_potential(rho, prob)
  plan = _band_plan(wl, prob, rho)
  p = seeded_params((2,) + wl.shape[1:], 4)
  _check_trace(_eval(plan, p), prob, p, v)


def test_fixed_potential_gga():
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(4.0)
  xc = "gga_x_pbe+gga_c_pbe"
  prob = _problem(wl, xc=xc)
  rho = _density(prob, 5)
  v = OB.effective_potential(rho, prob)
  plan = _band_plan(wl, prob, rho, xc=xc)
  p = seeded_params(wl.shape, 6)
  _check_trace(_eval(plan, p), prob, p, v)


def test_set_kpoints_matches_a_new_plan():
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(4.0)
  prob = _problem(wl)
  rho = _density(prob, 7)
  plan = _band_plan(wl, prob, rho)
  k2 = prob.kpts[::-1] * 0.5 + 0.1
  plan.set_kpoints(k2)
  prob2 = dataclasses.replace(prob, kpts=k2)
  fresh = _band_plan(wl, prob2, rho)
  p = seeded_params(wl.shape, 8)
  a, b = _eval(plan, p), _eval(fresh, p)
  assert np.array_equal(a["energies"], b["energies"])
  assert np.array_equal(a["dw_re"], b["dw_re"])
  _check_trace(a, prob2, p, OB.effective_potential(rho, prob2))


@pytest.mark.parametrize("grad", [True, False])
def test_subspace_hamiltonian_and_eigenvalues(grad):
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(4.0)
  prob = _problem(wl)
  rho = _density(prob, 9)
  v = OB.effective_potential(rho, prob)
  ones = np.ones_like(prob.occupation)
  plan = _band_plan(wl, prob, rho, occ=ones)
  p = seeded_params(wl.shape, 10)
  _eval(plan, p, grad=grad)
  h = plan.subspace_hamiltonian().cpu().numpy()
  ev, vec = plan.subspace_eigenvalues(vectors=True)
  ev, vec = ev.cpu().numpy(), vec.cpu().numpy()
  prob1 = dataclasses.replace(prob, occupation=ones)
  href = OB.subspace_hamiltonian(p, prob1, v)
  # the device's QR has R with a positive diagonal, numpy's Householder QR
  # column phases d = diag(R) / |diag(R)| away from it: Q_dev = Q d
  r = np.linalg.qr(p["w_re"] + 1j * p["w_im"], mode="reduced")[1]
  d = np.diagonal(r, axis1=-2, axis2=-1)
  d = d / np.abs(d)
  href = np.conj(d)[..., :, None] * href * d[..., None, :]
  assert rel_max(h, href) <= 1e-9
  eref = np.linalg.eigvalsh(href)
  assert rel_max(ev, eref) <= 1e-9
  # the eigenvectors diagonalise h, orthonormally
  vh = np.conj(np.swapaxes(vec, -1, -2))
  d = vh @ href @ vec
  nb = ev.shape[-1]
  assert rel_max(d, ev[..., None] * np.eye(nb)) <= 1e-9
  assert rel_max(vh @ vec, np.broadcast_to(np.eye(nb), d.shape)) <= 1e-12


def test_leaving_band_mode_restores_the_ground_state():
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import EnergyPlan
  wl = W.si8_cube(4.0, k=1)
  prob = _problem(wl)
  rho = _density(prob, 11)
  plan = _band_plan(wl, prob, rho)
  p = seeded_params(wl.shape, 12)
  _eval(plan, p)
  plan.set_fixed_density(None)
  a = _eval(plan, p)
  b = _eval(EnergyPlan.from_workload(wl), p)
  assert np.array_equal(a["energies"], b["energies"])
  assert np.array_equal(a["dw_re"], b["dw_re"])


def _tiny_config(tmp_path, **over):
  from jrystal_amd import config as C
  cfg = C.get_config()
  cfg.update(crystal="diamond", grid_sizes=12, cutoff_energy=8.0,
             k_grid_sizes=1, empty_bands=2, band_structure_empty_bands=2,
             epoch=30, optimizer="adam",
             optimizer_args={"learning_rate": 0.01, "b1": 0.9, "b2": 0.99},
             k_path_special_points="LGX", num_kpoints=5,
             band_structure_epoch=6, k_path_fine_tuning_epoch=3,
             save_dir=str(tmp_path), verbose=False)
  cfg.update(over)
  return cfg


@dataclasses.dataclass
class _GS:
  density: object
  total_energy: float = 0.0


@pytest.mark.parametrize("chunks", [1, 2])
def test_band_driver_adam_chain_matches_oracle(tmp_path, chunks):
  """chunks = 2: the reference on two devices (parallel_over_k_path),
  each half of the path from the same start."""
  from jrystal_amd import grid as G
  BD = importlib.import_module("jrystal_amd.calc.band_all_electrons")
  from oracle import optim as OO
  from oracle import pw as OP
  cfg = _tiny_config(tmp_path, num_kpoints=4, k_path_chunks=chunks)
  cr = BD.create_crystal(cfg)
  m = BD.create_freq_mask(cfg, cr)
  mask = np.asarray(m.cpu() if hasattr(m, "cpu") else m, bool)
  gs = G.proper_grid_size(cfg.grid_sizes)
  g = np.asarray(G.g_vectors(cr.cell_vectors, gs))
  nb = 6 + 2
  prob = OA.Problem(mask=mask, g_vector_grid=g, kpts=np.zeros((1, 3)),
                    vol=cr.vol, occupation=np.ones((1, 1, nb)),
                    positions=cr.positions,
                    charges=np.asarray(cr.charges, float))
  rho = _density(prob, 13)
  out = BD.calc_with_output(cfg, ground_state=_GS(torch.tensor(rho)))
  assert out.eigenvalues.shape == (1, 4, nb)
  assert np.array_equal(np.load(out.save_path), out.eigenvalues)
  # the same chain on the oracle, chunk by chunk
  v = OB.effective_potential(rho, prob)
  per = 4 // chunks
  evs, last = [], []
  for d in range(chunks):
    p = OP.param_init(cfg.seed, nb, 1, mask, dist="uniform")
    opt = OO.Adam(p, **cfg.optimizer_args)
    for i in range(per):
      pk = dataclasses.replace(prob, kpts=out.k_path[d * per + i][None])
      for _ in range(cfg.band_structure_epoch if i == 0
                     else cfg.k_path_fine_tuning_epoch):
        p = opt.step(p, OB.hamiltonian_trace(p, pk, v)[1])
      evs.append(OB.band_eigenvalues(p, pk, v)[:, 0])
    last.append(p)
  evs = np.stack(evs, axis=1)
  print("eigenvalues at k0:", out.eigenvalues[0, 0], flush=True)
  assert rel_max(out.eigenvalues, evs) <= 1e-9
  for key in ("w_re", "w_im"):
    ref = np.concatenate([p[key] for p in last], axis=1)
    assert rel_max(out.params_pw[key].cpu().numpy(), ref) <= 1e-9


def test_band_driver_cg_reaches_dense_eigenvalues(tmp_path):
  from jrystal_amd import grid as G
  BD = importlib.import_module("jrystal_amd.calc.band_all_electrons")
  cfg = _tiny_config(tmp_path, optimizer="cg", band_structure_epoch=300,
                     k_path_fine_tuning_epoch=150, num_kpoints=3)
  cr = BD.create_crystal(cfg)
  m = BD.create_freq_mask(cfg, cr)
  mask = np.asarray(m.cpu() if hasattr(m, "cpu") else m, bool)
  gs = G.proper_grid_size(cfg.grid_sizes)
  g = np.asarray(G.g_vectors(cr.cell_vectors, gs))
  nb = 8
  prob = OA.Problem(mask=mask, g_vector_grid=g, kpts=np.zeros((1, 3)),
                    vol=cr.vol, occupation=np.ones((1, 1, nb)),
                    positions=cr.positions,
                    charges=np.asarray(cr.charges, float))
  rho = _density(prob, 14)
  out = BD.calc_with_output(cfg, ground_state=_GS(torch.tensor(rho)))
  v = OB.effective_potential(rho, prob)
  for i, k in enumerate(out.k_path):
    lam = np.linalg.eigvalsh(OB.dense_hamiltonian(prob, v, kpt=k))[:nb]
    err = np.max(np.abs(out.eigenvalues[0, i] - lam))
    print(f"k{i}: max |eig - dense| = {err:.2e}", flush=True)
    assert err <= 1e-6 * max(1.0, np.max(np.abs(lam)))


@pytest.mark.parametrize("nb", [40, 70])
def test_subspace_eigenvalues_many_bands(nb):
  """The whole-block Jacobi beyond one register tile (40 bands: LDS; 70
  bands: matrices in global scratch) and the wide QR in band mode."""
  from jrystal_amd import workloads as W
  wl = W.build(f"si2-16^3-{nb}b", W.fcc_si2(4.0), (16, 16, 16), (1, 1, 1),
               nb, 15.0)
  assert wl.num_g > nb
  prob = _problem(wl)
  rho = _density(prob, 21)
  v = OB.effective_potential(rho, prob)
  ones = np.ones_like(prob.occupation)
  plan = _band_plan(wl, prob, rho, occ=ones)
  p = seeded_params(wl.shape, 22)
  _eval(plan, p, grad=False)
  ev = plan.subspace_eigenvalues().cpu().numpy()
  prob1 = dataclasses.replace(prob, occupation=ones)
  eref = OB.band_eigenvalues(p, prob1, v)
  assert rel_max(ev, eref) <= 1e-9


def test_trace_equals_kohn_sham_total_energy():
  """hamiltonian_test.py:78-102 through the public functions: the band
  trace at the density of the same orbitals (all-ones occupation) equals
  energy.total_energy(kohn_sham=True) (atol 1e-7 there, 1e-9 here)."""
  import os
  from jrystal_amd import energy, grid, pw
  from jrystal_amd.crystal import Crystal
  from jrystal_amd.hamiltonian import band_plan, hamiltonian_matrix_trace
  here = os.path.dirname(os.path.abspath(__file__))
  cr = Crystal.create_from_file(os.path.join(here, "golden", "diamond.xyz"))
  kpts = grid.k_vectors(cr.cell_vectors, [2, 2, 1])
  g = grid.g_vectors(cr.cell_vectors, [7, 8, 9])
  mask = grid.cubic_mask([7, 8, 9])
  nb = cr.num_electron
  params = pw.param_init(123, nb, kpts.shape[0], mask)
  coeff = pw.coeff(params, mask)
  occ = np.ones((1, kpts.shape[0], nb))
  rho = pw.density_grid(coeff, cr.vol, occ)
  plan = band_plan(cr.positions, cr.charges, cr.cell_vectors, mask, rho,
                   kpts, nb)
  e1 = float(hamiltonian_matrix_trace(plan, params["w_re"], params["w_im"]))
  e2 = float(energy.total_energy(coeff, cr.positions, cr.charges, g, kpts,
                                 cr.vol, kohn_sham=True))
  print(f"trace {e1:.12f} kohn-sham total {e2:.12f}", flush=True)
  assert abs(e1 - e2) < 1e-9 * max(1.0, abs(e2))
