"""Norm-conserving pseudopotential oracle (oracle/pseudopotential.py) and the
host set-up of the product (jrystal_amd/pseudopotential.py), on the CPU.

Pins of the oracle:
* the UPF fields of the shipped Si and C files (z_valence, projectors,
  their l, D in Ha = PP_DIJ / 2, mesh size);
* the radial transform against closed forms: S_0[exp(-r^2)] =
  sqrt(pi)/4 exp(-q^2/4), S_1[r exp(-r^2)] = sqrt(pi)/8 q exp(-q^2/4)
  (to the accuracy of the reference's rectangle rule on the UPF mesh);
* a local potential that is the bare Coulomb -z/r reproduces the
  reference's external_reciprocal exactly (its S_0 term vanishes);
* a single plane wave and one l = 0 projector: E_nl = 4 pi |D| beta(q)^2 /
  vol, also for a negative D (the reference counts it as |D|);
* invariance of the non-local energy under an orthogonal change of the
  real harmonic basis; the total energy gradient against central
  differences.
The product's tables match the oracle's: the local box to rounding, the
projectors through P^H P (the product's real-harmonic order differs, which
the energy does not see).
"""
import os

import numpy as np
import pytest

from helpers import rel_max
from oracle import adjoint as OA
from oracle import energy as OE
from oracle import grid as OG
from oracle import occupation as OO
from oracle import pseudopotential as OPP
from oracle.crystal import Crystal as OCrystal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PPDIR = os.path.join(ROOT, "jrystal_amd", "data", "pseudopotential",
                     "normconserving")
GEOM = os.path.join(ROOT, "jrystal_amd", "geometry")


def test_upf_fields():
  si = OPP.parse_upf(os.path.join(PPDIR, "Si.pz-vbc.UPF"))
  assert float(si["header"]["z_valence"]) == 4.0
  assert [l for l, _ in si["betas"]] == [0, 1]
  assert si["r"].size == 431 and si["local"].size == 431
  pp = OPP.create(["Si"], np.zeros((1, 3)), PPDIR)
  np.testing.assert_allclose(pp.d_matrix[0],
                             np.diag([1.52388501179, 3.68330413052]) / 2)
  c = OPP.create(["C"], np.zeros((1, 3)), PPDIR)
  assert c.valence_charges == [4] and list(c.angular_momentum[0]) == [0]


def test_radial_transform_closed_forms():
  r = OPP.parse_upf(os.path.join(PPDIR, "Si.pz-vbc.UPF"))["r"]
  f = np.stack([np.exp(-r * r), r * np.exp(-r * r)])
  q, s = OPP.sbt_numerical(r, f, [0, 1], 8.0)
  a0 = np.sqrt(np.pi) / 4 * np.exp(-q * q / 4)
  a1 = np.sqrt(np.pi) / 8 * q * np.exp(-q * q / 4)
  # rectangle rule with forward differences on the log mesh (dx = 0.025)
  assert np.max(np.abs(s[0] - a0)) < 0.02 * a0.max()
  assert np.max(np.abs(s[1] - a1)) < 0.02 * a1.max()
  # and exactly the stated sum
  dr = np.append(np.diff(r), 0.0)
  from scipy.special import spherical_jn
  i = 137
  assert s[1, i] == pytest.approx(
    np.sum(spherical_jn(1, q[i] * r) * f[1] * r * r * dr), rel=1e-13)


def test_bare_coulomb_local_potential_is_external_reciprocal():
  cr = OCrystal.create_from_file(os.path.join(GEOM, "diamond.xyz"))
  pp = OPP.create(cr.symbols, cr.positions, PPDIR)
  pp.local_potential = [-z / r for z, r in zip(pp.valence_charges,
                                               pp.r_grid)]
  g = OG.g_vectors(cr.cell_vectors, [12, 12, 12])
  v = OPP.local_potential_reciprocal(pp, g, cr.vol)
  ref = OE.external_reciprocal(cr.positions, np.asarray(pp.valence_charges,
                                                        float), g, cr.vol)
  ref[0, 0, 0] = 0.0
  assert rel_max(v, ref) < 1e-9


@pytest.mark.parametrize("d", [0.7, -1.3])
def test_single_plane_wave_single_projector(d):
  """One l = 0 projector with D = d and one plane wave G0: F = 4 pi
  sqrt(d) Y00 beta(|G0+k|) e^{-i(G0+k).tau}, E = 4 pi |d| beta^2 / vol."""
  cr = OCrystal.create_from_file(os.path.join(GEOM, "diamond.xyz"))
  pp = OPP.create(["C"], cr.positions[:1], PPDIR)
  pp.d_matrix = [np.array([[d]])]
  gs = [8, 8, 8]
  g = OG.g_vectors(cr.cell_vectors, gs)
  mask = np.asarray(OG.cubic_mask(gs))
  k = np.array([[0.1, -0.2, 0.05]])
  P = OPP.projectors(pp, g, mask, k)
  ng = int(mask.sum())
  j = 17
  q = np.zeros((1, 1, ng, 1), complex)
  q[0, 0, j, 0] = 1.0
  e = OPP.energy_nonlocal(q, P, np.ones((1, 1, 1)), cr.vol, False)[0]
  gk = g[mask][j] + k[0]
  qmax = max(np.max(np.linalg.norm(g + kk, axis=-1)) for kk in k)
  qq, s = OPP.sbt_numerical(pp.r_grid[0], pp.betas[0], [0], qmax)
  from scipy.interpolate import CubicSpline
  beta = CubicSpline(qq, s[0])(np.linalg.norm(gk))
  assert e == pytest.approx(4 * np.pi * abs(d) * beta**2 / cr.vol, rel=1e-12)


def _si_problem(gs=(12, 12, 12), ecut=10.0, kgrid=(2, 1, 1), nb=6, seed=0):
  cr = OCrystal.create_from_file(os.path.join(GEOM, "si.xyz"))
  pp = OPP.create(cr.symbols, cr.positions, PPDIR)
  g = OG.g_vectors(cr.cell_vectors, list(gs))
  mask = OG.spherical_mask(cr.cell_vectors, list(gs), ecut)
  kpts = OG.k_vectors(cr.cell_vectors, list(kgrid))
  nk = kpts.shape[0]
  occ = OO.uniform(nk, sum(pp.valence_charges), 0, nb)
  prob = OA.Problem(mask=mask, g_vector_grid=g, kpts=kpts, vol=cr.vol,
                    occupation=occ, positions=cr.positions,
                    charges=np.asarray(cr.charges, float),
                    v_ext_rec=OPP.local_potential_reciprocal(pp, g, cr.vol),
                    projectors=OPP.projectors(pp, g, mask, kpts))
  rng = np.random.default_rng(seed)
  shape = (1, nk, int(mask.sum()), nb)
  params = {"w_re": rng.standard_normal(shape),
            "w_im": rng.standard_normal(shape)}
  return cr, pp, prob, params


def test_nonlocal_energy_basis_invariant():
  cr, pp, prob, params = _si_problem()
  q = np.linalg.qr(params["w_re"] + 1j * params["w_im"])[0]
  e0 = OPP.energy_nonlocal(q, prob.projectors, prob.occupation, cr.vol,
                           False)[0]
  # the l = 1 rows of each atom (three consecutive rows after the l = 0 one)
  # rotated by an orthogonal matrix
  rng = np.random.default_rng(3)
  o = np.linalg.qr(rng.standard_normal((3, 3)))[0]
  P = prob.projectors.copy()
  for a in range(2):
    rows = slice(4 * a + 1, 4 * a + 4)
    P[:, rows] = np.einsum("ij,kjg->kig", o, prob.projectors[:, rows])
  e1 = OPP.energy_nonlocal(q, P, prob.occupation, cr.vol, False)[0]
  assert e0 > 0 and e1 == pytest.approx(e0, rel=1e-12)


def test_normcons_energy_gradient_central_differences():
  cr, pp, prob, params = _si_problem(nb=4)
  parts, grad, _ = OA.energy_and_grad(params, prob)
  assert parts["external_nonlocal"] > 0
  rng = np.random.default_rng(1)
  h = 1e-6
  for key in ("w_re", "w_im"):
    for _ in range(5):
      idx = tuple(rng.integers(0, n) for n in params[key].shape)
      pp_, pm = ({k: v.copy() for k, v in params.items()} for _ in range(2))
      pp_[key][idx] += h
      pm[key][idx] -= h
      fd = (OA.energy_and_grad(pp_, prob, want_grad=False)[0]["total"] -
            OA.energy_and_grad(pm, prob, want_grad=False)[0]["total"]) / (2 * h)
      assert abs(fd - grad[key][idx]) < 1e-6 * max(1.0, abs(fd)), (key, idx)


def test_product_tables_match_oracle():
  from jrystal_amd import pseudopotential as PS
  from jrystal_amd.crystal import Crystal
  cr = Crystal.create_from_file(os.path.join(GEOM, "si.xyz"))
  pp = PS.NormConservingPseudopotential.create(cr)
  opp = OPP.create(cr.symbols, cr.positions, PPDIR)
  gs = [12, 12, 12]
  g = OG.g_vectors(cr.cell_vectors, gs)
  mask = OG.spherical_mask(cr.cell_vectors, gs, 10.0)
  kpts = OG.k_vectors(cr.cell_vectors, [2, 1, 1])
  assert rel_max(PS.potential_local_reciprocal(pp, g, cr.vol),
                 OPP.local_potential_reciprocal(opp, g, cr.vol)) < 1e-12
  a = PS.potential_nonlocal_psi_reciprocal(pp, g, mask, kpts)
  b = OPP.projectors(opp, g, mask, kpts)
  assert a.shape == b.shape
  for k in range(kpts.shape[0]):
    ga = np.conj(a[k].T) @ a[k]
    gb = np.conj(b[k].T) @ b[k]
    assert rel_max(ga, gb) < 1e-12


def test_band_mode_nonlocal_dense_matches_application():
  """With the pseudopotential's local part in v_eff and the projectors in
  the Hamiltonian, Q^H H_dense Q equals the FFT/projector application."""
  from oracle import band as OB
  cr, pp, prob, params = _si_problem(gs=(10, 10, 10), ecut=8.0, nb=4)
  prob.occupation = np.ones_like(prob.occupation)
  q = np.linalg.qr(params["w_re"] + 1j * params["w_im"])[0]
  rho = OA.energy_and_grad(params, prob, want_grad=False)[2]
  v = OB.effective_potential(rho, prob)
  hs = OB.subspace_hamiltonian(params, prob, v)
  for k in range(prob.kpts.shape[0]):
    hd = OB.dense_hamiltonian(prob, v, 0, k)
    assert rel_max(hd, np.conj(hd.T)) < 1e-13
    assert rel_max(hs[0, k], np.conj(q[0, k].T) @ hd @ q[0, k]) < 1e-12
