"""GPU implementations of the reference's public functions (pw.*, energy.*)
against the oracle's functions of the same names, on the same inputs."""
import numpy as np
import pytest

from helpers import rel_max
from oracle import adjoint as OA
from oracle import energy as OE
from oracle import pw as OP
from oracle import workloads as OW

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _np(t):
  return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def case():
  cr, prob, params = OW.tiny(z=4.0, grid=(12, 12, 12), kgrid=(2, 1, 1), nb=3,
                             ecut=6.0)
  return cr, prob, params


def test_coeff_matches_up_to_column_phase(case):
  from jrystal_amd import pw
  cr, prob, params = case
  got = _np(pw.coeff(params, prob.mask))
  ref = OP.coeff(params, prob.mask)
  assert got.shape == ref.shape
  # QR is unique up to a unit phase per column; energies are invariant.
  ph = np.einsum("skbxyz,skbxyz->skb", np.conj(ref), got)
  np.testing.assert_allclose(np.abs(ph), 1.0, atol=1e-13)
  ph = ph / np.abs(ph)
  np.testing.assert_allclose(got * np.conj(ph)[..., None, None, None], ref,
                             atol=1e-13)
  u = _np(pw.unitary_matrix(params))
  np.testing.assert_allclose(np.abs(np.einsum("skgb,skgb->skb", np.conj(
    OP.unitary_matrix(params)), u)), 1.0, atol=1e-13)


def test_wave_and_density_functions(case):
  from jrystal_amd import pw
  cr, prob, params = case
  c = OP.coeff(params, prob.mask)
  ct = torch.tensor(c, device="cuda")
  np.testing.assert_allclose(_np(pw.wave_grid(ct, prob.vol)),
                             OP.wave_grid(c, prob.vol), atol=1e-12)
  np.testing.assert_allclose(_np(pw.density_grid(ct, prob.vol)),
                             OP.density_grid(c, prob.vol), atol=1e-12)
  occ = prob.occupation
  np.testing.assert_allclose(_np(pw.density_grid(ct, prob.vol, occ)),
                             OP.density_grid(c, prob.vol, occ), atol=1e-12)
  np.testing.assert_allclose(
    _np(pw.density_grid_reciprocal(ct, prob.vol, occ)),
    OP.density_grid_reciprocal(c, prob.vol, occ), atol=1e-10)
  with pytest.raises(ValueError):
    pw.density_grid(ct, prob.vol, occ[:, :1])


def test_energy_functions(case):
  from jrystal_amd import energy
  cr, prob, params = case
  c = OP.coeff(params, prob.mask)
  ct = torch.tensor(c, device="cuda")
  occ = prob.occupation
  g = prob.g_vector_grid
  rho = OP.density_grid(c, prob.vol, occ)
  rr = OP.density_grid_reciprocal(c, prob.vol, occ)
  rho_t = torch.tensor(rho, device="cuda")
  rr_t = torch.tensor(rr, device="cuda")
  checks = [
    (energy.kinetic(g, prob.kpts, ct, occ), OE.kinetic(g, prob.kpts, c, occ)),
    (energy.hartree(rr_t, g, prob.vol), OE.hartree(rr, g, prob.vol)),
    (energy.hartree(rr_t, g, prob.vol, kohn_sham=True),
     OE.hartree(rr, g, prob.vol, kohn_sham=True)),
    (energy.external(rr_t, prob.positions, prob.charges, g, prob.vol),
     OE.external(rr, prob.positions, prob.charges, g, prob.vol)),
    (energy.xc_energy(rho_t, g, prob.vol), OE.xc_energy(rho, g, prob.vol)),
  ]
  for got, ref in checks:
    np.testing.assert_allclose(_np(got), ref, rtol=1e-12, atol=1e-12)
  np.testing.assert_allclose(_np(energy.kinetic(g, prob.kpts, ct)),
                             OE.kinetic(g, prob.kpts, c), rtol=1e-12)
  tot = energy.total_energy(ct, prob.positions, prob.charges, g, prob.kpts,
                            prob.vol, occupation=occ)
  ref = OE.total_energy(c, prob.positions, prob.charges, g, prob.kpts,
                        prob.vol, occupation=occ)
  np.testing.assert_allclose(_np(tot), ref, rtol=1e-12)


def test_nuclear_repulsion_against_oracle_and_textbook():
  from jrystal_amd import energy
  cr = OW.si2_fcc(4.0)
  from oracle import grid as OG
  gv = OG.g_vectors(cr.cell_vectors, (32, 32, 32))
  got = _np(energy.nuclear_repulsion(cr.positions, cr.charges,
                                     cr.cell_vectors, gv, cr.vol, 0.1, 2e4))
  ref = OE.nuclear_repulsion(cr.positions, cr.charges, cr.cell_vectors, gv,
                             cr.vol, 0.1, 2e4)
  np.testing.assert_allclose(got, ref, rtol=1e-12)
  np.testing.assert_allclose(got, -8.40046478618, atol=1e-9)
  n = np.arange(-4, 5)
  lat = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3) @ \
      cr.cell_vectors
  gv24 = OG.g_vectors(cr.cell_vectors, (24, 24, 24))
  got = _np(energy.ewald_coulomb_repulsion(cr.positions, cr.charges, gv24,
                                           cr.vol, 0.5, lat))
  np.testing.assert_allclose(got, OE.ewald_lattice(cr.positions, cr.charges,
                                                   cr.cell_vectors),
                             rtol=1e-9)


def test_autograd_through_public_api_matches_adjoint(case):
  """The driver closure (calc_ground_state_energy_all_electrons.py:119-137)
  composed from the public GPU functions, differentiated by torch autograd
  through the hand-written HIP backward kernels."""
  from jrystal_amd import energy, pw
  cr, prob, params = case
  wr = torch.tensor(params["w_re"], device="cuda", requires_grad=True)
  wi = torch.tensor(params["w_im"], device="cuda", requires_grad=True)
  c = pw.coeff({"w_re": wr, "w_im": wi}, prob.mask)
  occ = prob.occupation
  g = prob.g_vector_grid
  rho = pw.density_grid(c, prob.vol, occ)
  rr = pw.density_grid_reciprocal(c, prob.vol, occ)
  e = (energy.kinetic(g, prob.kpts, c, occ) + energy.hartree(rr, g, prob.vol) +
       energy.external(rr, prob.positions, prob.charges, g, prob.vol) +
       energy.xc_energy(rho, g, prob.vol))
  e.backward()
  parts, grad, _ = OA.energy_and_grad(params, prob)
  np.testing.assert_allclose(_np(e), parts["total"], rtol=1e-12)
  assert rel_max(_np(wr.grad), grad["w_re"]) < 1e-10
  assert rel_max(_np(wi.grad), grad["w_im"]) < 1e-10
