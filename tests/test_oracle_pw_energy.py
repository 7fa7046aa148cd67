"""Oracle plane-wave and energy functions against closed forms, brute force
and invariants (CPU only).  Each check is chosen so that a plausible mistake
(dropped term, sign, index, normalisation) fails it."""
import numpy as np
import pytest

from oracle import energy as OE
from oracle import grid as OG
from oracle import pw as OP
from oracle.workloads import si2_fcc, tiny


def _setup(grid=(7, 8, 9), nb=3, nk=1, seed=0):
  cr = si2_fcc(4.0)
  mask = np.asarray(OG.cubic_mask(grid))
  params = OP.param_init(seed, nb, nk, mask)
  return cr, mask, params


def test_qr_orthonormal_and_factorises():
  cr, mask, params = _setup()
  q, r = OP.unitary_matrix_with_r(params)
  w = params["w_re"] + 1j * params["w_im"]
  eye = np.eye(q.shape[-1])
  np.testing.assert_allclose(np.conj(np.swapaxes(q, -1, -2)) @ q,
                             np.broadcast_to(eye, r.shape), atol=1e-14)
  np.testing.assert_allclose(q @ r, w, atol=1e-13)
  assert np.allclose(np.tril(r, -1), 0)


def test_qr_matches_torch_up_to_column_phase():
  torch = pytest.importorskip("torch")
  cr, mask, params = _setup(nb=5, nk=2)
  q = OP.unitary_matrix(params)
  w = torch.complex(torch.tensor(params["w_re"]), torch.tensor(params["w_im"]))
  qt = torch.linalg.qr(w)[0].numpy()
  d = np.einsum("skgb,skgb->skb", np.conj(q), qt)
  np.testing.assert_allclose(np.abs(d), 1.0, atol=1e-13)


def test_expand_squeeze_roundtrip_exact():
  cr, mask, params = _setup()
  q = OP.unitary_matrix(params)
  box = OP.expand_coefficient(q, mask)
  assert box.shape == q.shape[:2] + (q.shape[3],) + mask.shape
  assert np.array_equal(OP.squeeze_coefficient(box, mask), q)
  assert np.all(box[..., ~mask] == 0)


def test_wave_grid_matches_plane_wave_sum():
  """pw_test.py:36-49 method: FFT wave function == brute-force sum at
  r_vectors; tightened from atol 1e-8 to 1e-12."""
  cr, mask, params = _setup(grid=(7, 8, 9))
  a = cr.cell_vectors
  c = OP.coeff(params, mask)
  psi = OP.wave_grid(c, cr.vol)
  g = OG.g_vectors(a, mask.shape)
  r = OG.r_vectors(a, mask.shape).reshape(-1, 3)
  ref = OP.wave_r_bruteforce(r, c, g, cr.vol).reshape(psi.shape)
  np.testing.assert_allclose(psi, ref, atol=1e-12)


def test_parseval_and_density_integral():
  cr, mask, params = _setup(grid=(12, 12, 12), nb=4, nk=2)
  c = OP.coeff(params, mask)
  psi = OP.wave_grid(c, cr.vol)
  n = np.prod(mask.shape)
  norms = np.sum(np.abs(psi)**2, axis=(-3, -2, -1)) * cr.vol / n
  np.testing.assert_allclose(norms, 1.0, atol=1e-13)
  occ = np.random.default_rng(1).random((1, 2, 4))
  rho = OP.density_grid(c, cr.vol, occ)
  assert rho.min() >= 0
  np.testing.assert_allclose(rho.sum() * cr.vol / n, occ.sum(), rtol=1e-13)
  rho_rec = OP.density_grid_reciprocal(c, cr.vol, occ)
  np.testing.assert_allclose(rho_rec[0, 0, 0, 0], rho.sum(), rtol=1e-13)
  with pytest.raises(ValueError):
    OP.density_grid(c, cr.vol, occ[:, :1])


def _cos_density(a, grid, g_index, rho0, amp):
  """rho(r) = rho0 + amp cos(G0.r) on the grid, G0 = g_vectors[g_index]."""
  g = OG.g_vectors(a, grid)
  g0 = g[g_index]
  r = OG.r_vectors(a, grid)
  return rho0 + amp * np.cos(r @ g0), g0, g


def test_hartree_cosine_closed_form():
  """E_H = pi vol A^2 / |G0|^2 for rho0 + A cos(G0.r)."""
  cr = si2_fcc()
  grid = (10, 12, 14)
  rho, g0, g = _cos_density(cr.cell_vectors, grid, (1, 2, 3), 0.3, 0.07)
  rho_rec = np.fft.fftn(rho)[None]
  e = OE.hartree(rho_rec, g, cr.vol)
  np.testing.assert_allclose(e, np.pi * cr.vol * 0.07**2 / (g0 @ g0),
                             rtol=1e-13)


def test_hartree_equals_half_real_space_potential_energy():
  """energy_test.py:61-112 invariant: E_H = 1/2 <rho, V_H> in real space."""
  cr, mask, params = _setup(grid=(12, 12, 12), nb=4)
  c = OP.coeff(params, mask)
  occ = np.full((1, 1, 4), 2.0)
  rho = OP.density_grid(c, cr.vol, occ)
  rho_rec = np.fft.fftn(rho, axes=(-3, -2, -1))
  g = OG.g_vectors(cr.cell_vectors, mask.shape)
  vh = np.real(np.fft.ifftn(OE.hartree_reciprocal(rho_rec, g, kohn_sham=True)))
  n = np.prod(mask.shape)
  e_real = 0.5 * np.sum(vh * rho[0]) * cr.vol / n
  np.testing.assert_allclose(OE.hartree(rho_rec, g, cr.vol), e_real,
                             rtol=1e-12)


def test_external_single_atom_closed_form():
  """E_ext = -4 pi Z A cos(G0.R) / (|G0|^2 + 1e-10) for one atom and
  rho0 + A cos(G0.r) (rho0 drops out with G = 0)."""
  cr = si2_fcc()
  grid = (10, 12, 14)
  rho, g0, g = _cos_density(cr.cell_vectors, grid, (2, 1, 11), 0.4, 0.05)
  pos = np.array([[0.7, -1.1, 2.3]])
  z = np.array([5.0])
  e = OE.external(np.fft.fftn(rho)[None], pos, z, g, cr.vol)
  ref = -4 * np.pi * 5.0 * 0.05 * np.cos(g0 @ pos[0]) / (g0 @ g0 + 1e-10)
  np.testing.assert_allclose(e, ref, rtol=1e-12)


def test_external_regulariser_is_literal():
  """The +1e-10 of potential.py:161 is kept: it is visible at 1e-10 rel."""
  cr = si2_fcc()
  grid = (8, 8, 8)
  rho, g0, g = _cos_density(cr.cell_vectors, grid, (1, 0, 0), 0.0, 1.0)
  e = OE.external(np.fft.fftn(rho)[None], np.zeros((1, 3)), np.ones(1), g,
                  cr.vol)
  without = -4 * np.pi / (g0 @ g0)
  assert abs(e - without) > 0.5e-10 * abs(without) / (g0 @ g0)


def test_lda_exchange_uniform_density():
  """Dirac exchange: E_x = -(3/4)(3/pi)^(1/3) n^(4/3) vol."""
  n0 = 0.0371
  rho = np.full((1, 6, 7, 8), n0)
  e = OE.xc_energy(rho, None, 123.0, "lda_x")
  np.testing.assert_allclose(e, -0.75 * (3 / np.pi)**(1 / 3) * n0**(4 / 3) *
                             123.0, rtol=1e-14)
  # spin-polarised form of xc.py:56-59 with equal spins reduces to the same
  rho2 = np.full((2, 6, 7, 8), n0 / 2)
  np.testing.assert_allclose(OE.xc_energy(rho2, None, 123.0, "lda_x"), e,
                             rtol=1e-14)


def test_lda_threshold_parity_unpinned_region():
  """Below the density threshold the exchange energy density is zero; the
  exact libxc threshold is parity unpinned, the tested densities stay far
  above it."""
  assert OE.lda_x_eps(np.array([0.0, 1e-16]))[0] == 0.0
  assert OE.lda_x_eps(np.array([1e-3]))[0] < 0


def test_kinetic_single_plane_wave():
  """One state with C = delta_{G,G1}: E_kin = 1/2 f |G1 + k|^2."""
  cr = si2_fcc()
  grid = (8, 8, 8)
  g = OG.g_vectors(cr.cell_vectors, grid)
  kpts = np.array([[0.1, -0.2, 0.05]])
  c = np.zeros((1, 1, 1) + grid, complex)
  c[0, 0, 0, 3, 1, 6] = np.exp(0.3j)
  f = np.array([[[0.7]]])
  e = OE.kinetic(g, kpts, c, f)
  gk = g[3, 1, 6] + kpts[0]
  np.testing.assert_allclose(e, 0.5 * 0.7 * gk @ gk, rtol=1e-14)


def test_ewald_formula_matches_textbook_on_lattice_translations():
  """With true lattice translations in place of translation_vectors' grid,
  the literal ewald.py formula equals an independent textbook Ewald sum."""
  cr = si2_fcc(4.0)
  a = cr.cell_vectors
  n = np.arange(-4, 5)
  lat = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3) @ a
  g = OG.g_vectors(a, (24, 24, 24))
  e = OE.ewald_coulomb_repulsion(cr.positions, cr.charges, g, cr.vol, 0.5, lat)
  ref = OE.ewald_lattice(cr.positions, cr.charges, a)
  np.testing.assert_allclose(e, ref, rtol=1e-9)
  # textbook sum is independent of the splitting parameter
  np.testing.assert_allclose(OE.ewald_lattice(cr.positions, cr.charges, a,
                                              eta=0.8), ref, rtol=1e-10)


def test_ewald_reference_grid_value():
  """nuclear_repulsion with the reference's translation grid (integer lattice
  translations, 64^3 of them for cutoff 2e4) and eta = 0.1 equals the
  textbook Ewald energy of the Si2 cell at Z = 4."""
  cr = si2_fcc(4.0)
  g = OG.g_vectors(cr.cell_vectors, (32, 32, 32))
  e = OE.nuclear_repulsion(cr.positions, cr.charges, cr.cell_vectors, g,
                           cr.vol, 0.1, 2e4)
  ref = OE.ewald_lattice(cr.positions, cr.charges, cr.cell_vectors)
  np.testing.assert_allclose(ref, -8.40046478618, rtol=1e-10)
  np.testing.assert_allclose(e, ref, atol=1e-9)


def test_total_energy_public_composite():
  """energy.py:246-307: default occupation is all ones, no Ewald term."""
  cr, prob, params = tiny()
  c = OP.coeff(params, prob.mask)
  e = OE.total_energy(c, prob.positions, prob.charges, prob.g_vector_grid,
                      prob.kpts, prob.vol, occupation=prob.occupation)
  parts = OE.total_energy(c, prob.positions, prob.charges,
                          prob.g_vector_grid, prob.kpts, prob.vol,
                          occupation=prob.occupation, split=True)
  np.testing.assert_allclose(sum(parts), e, rtol=1e-14)
  from oracle.adjoint import energy_and_grad
  p, _, _ = energy_and_grad(params, prob, want_grad=False)
  np.testing.assert_allclose(p["total"], e, rtol=1e-13)


def test_grid_translation_invariance():
  """Shifting atoms by a grid vector t and phasing C by e^{-iG.t} rolls rho
  and leaves every energy term unchanged."""
  cr, prob, params = tiny()
  from oracle.adjoint import energy_and_grad
  p0, _, rho0 = energy_and_grad(params, prob, want_grad=False)
  shift = np.array([2, 5, 7])
  t = (shift / np.array(prob.mask.shape)) @ cr.cell_vectors
  gm = prob.mask_g()
  ph = np.exp(-1j * gm @ t)[None, None, :, None]
  w = (params["w_re"] + 1j * params["w_im"]) * ph
  prob2 = type(prob)(**{**prob.__dict__, "positions": prob.positions + t})
  p1, _, rho1 = energy_and_grad({"w_re": w.real, "w_im": w.imag}, prob2,
                                want_grad=False)
  for k in p0:
    np.testing.assert_allclose(p1[k], p0[k], rtol=1e-12, atol=1e-12)
  np.testing.assert_allclose(rho1, np.roll(rho0, shift, axis=(1, 2, 3)),
                             atol=1e-13)
