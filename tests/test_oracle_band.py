"""Oracle of the Kohn-Sham band mode (oracle/band.py) against the
reference's own test, closed forms, an independently built dense
Hamiltonian and finite differences (CPU only).

* hamiltonian_test.py:78-102: the trace at the density of the same
  orbitals with all-ones occupation equals total_energy(kohn_sham=True).
* The dense plane-wave matrix (index differences of fftn(v)) reproduces the
  FFT application; its eigenvalues bound the subspace ones (Poincare
  separation) and the trace (Ky Fan), with equality at its eigenvectors.
* v = 0 gives the free-electron energies 1/2 |G+k|^2.
* vxc_lda's forms against finite differences of the exchange energy
  density; the trace gradient against central differences, and zero at the
  minimiser.
"""
import os

import numpy as np
import pytest

from helpers import rel_max
from oracle import adjoint as OA
from oracle import band as OB
from oracle import energy as OE
from oracle import grid as OG
from oracle import pw as OP
from oracle.crystal import Crystal as OCrystal

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _diamond_problem(grid=(7, 8, 9), kgrid=(2, 2, 1), spin_restricted=True,
                     seed=123, nb=None):
  """The set-up of hamiltonian_test.py:28-51: diamond, cubic mask on a
  7 x 8 x 9 grid, a 2 x 2 x 1 k-grid, num_bands = num_electron, all-ones
  occupation and the density of those orbitals."""
  cr = OCrystal.create_from_file(os.path.join(GOLD, "diamond.xyz"))
  g = OG.g_vectors(cr.cell_vectors, list(grid))
  mask = np.asarray(OG.cubic_mask(list(grid)))
  kpts = OG.k_vectors(cr.cell_vectors, list(kgrid))
  nb = cr.num_electron if nb is None else nb
  params = OP.param_init(seed, nb, kpts.shape[0], mask,
                         spin_restricted=spin_restricted)
  ns = params["w_re"].shape[0]
  occ = np.ones((ns, kpts.shape[0], nb))
  prob = OA.Problem(mask=mask, g_vector_grid=g, kpts=kpts, vol=cr.vol,
                    occupation=occ, positions=cr.positions,
                    charges=np.asarray(cr.charges, float))
  coeff = OP.coeff(params, mask)
  rho = OP.density_grid(coeff, cr.vol, occ)
  return cr, prob, params, coeff, rho


@pytest.mark.parametrize("spin_restricted", [True, False])
def test_trace_equals_kohn_sham_total_energy(spin_restricted):
  """hamiltonian_test.py:78-102 (atol 1e-7 there, 1e-10 here); the
  polarised case adds the per-channel vxc_lda."""
  cr, prob, params, coeff, rho = _diamond_problem(
    spin_restricted=spin_restricted)
  v = OB.effective_potential(rho, prob)
  e1 = OB.hamiltonian_trace(params, prob, v, want_grad=False)[0]["total"]
  e2 = OE.total_energy(coeff, cr.positions, np.asarray(cr.charges, float),
                       prob.g_vector_grid, prob.kpts, cr.vol, kohn_sham=True)
  assert abs(e1 - e2) < 1e-10 * max(1.0, abs(e2)), (e1, e2)


def test_vxc_lda_x_forms_by_finite_differences():
  rng = np.random.default_rng(1)
  rho = rng.uniform(0.01, 2.0, (2, 3, 4, 5))
  h = 1e-6
  # one channel: d(rho eps)/d rho
  r1 = rho[:1]
  fd = ((r1 + h) * OE.lda_x_eps(r1 + h) -
        (r1 - h) * OE.lda_x_eps(r1 - h)) / (2 * h)
  np.testing.assert_allclose(OE.vxc_lda_x(r1), fd, rtol=1e-8)
  # two channels: eps + rho_s d eps / d rho_s (xc.py:114-122), eps of
  # xc_density
  eps = OE.xc_density(rho)
  out = OE.vxc_lda_x(rho)
  for s in range(2):
    dp, dm = rho.copy(), rho.copy()
    dp[s] += h
    dm[s] -= h
    deps = (OE.xc_density(dp) - OE.xc_density(dm)) / (2 * h)
    np.testing.assert_allclose(out[s], eps + rho[s] * deps, rtol=1e-8)


def test_dense_matches_fft_application():
  cr, prob, params, coeff, rho = _diamond_problem(kgrid=(2, 1, 1), nb=5)
  v = OB.effective_potential(rho, prob)
  hs = OB.subspace_hamiltonian(params, prob, v)
  q = OP.unitary_matrix(params)
  for k in range(prob.kpts.shape[0]):
    hd = OB.dense_hamiltonian(prob, v, 0, k)
    assert rel_max(hd, np.conj(hd.T)) < 1e-14
    ref = np.conj(q[0, k].T) @ hd @ q[0, k]
    assert rel_max(hs[0, k], ref) < 1e-12


def test_free_electron_energies():
  cr, prob, params, coeff, rho = _diamond_problem(kgrid=(2, 1, 1), nb=4)
  v = np.zeros_like(rho)
  k = 1
  hd = OB.dense_hamiltonian(prob, v, 0, k)
  gk2 = np.sum((prob.mask_g() + prob.kpts[k])**2, axis=-1)
  np.testing.assert_allclose(np.linalg.eigvalsh(hd), np.sort(0.5 * gk2),
                             atol=1e-12)
  # orbitals = the 4 lowest plane waves: their eigenvalues exactly
  order = np.argsort(gk2)[:4]
  w = np.zeros(params["w_re"].shape)
  for b, g in enumerate(order):
    w[:, :, g, b] = 1.0
  p = {"w_re": w, "w_im": np.zeros_like(w)}
  ev = OB.band_eigenvalues(p, prob, v)[0, k]
  np.testing.assert_allclose(ev, np.sort(0.5 * gk2[order]), atol=1e-12)


def test_minimum_at_dense_eigenvectors():
  """Ky Fan: the trace over nb orthonormal orbitals is at least the sum of
  the nb lowest eigenvalues, with equality at the eigenvectors, where the
  gradient vanishes; Poincare separation bounds the subspace eigenvalues."""
  cr, prob, params, coeff, rho = _diamond_problem(kgrid=(1, 1, 1), nb=4)
  v = OB.effective_potential(rho, prob)
  hd = OB.dense_hamiltonian(prob, v, 0, 0)
  lam, vec = np.linalg.eigh(hd)
  nb = 4
  tr = OB.hamiltonian_trace(params, prob, v, want_grad=False)[0]["total"]
  assert tr > np.sum(lam[:nb]) + 1e-6
  ev = OB.band_eigenvalues(params, prob, v)[0, 0]
  assert np.all(ev >= lam[:nb] - 1e-10)
  assert np.all(ev <= lam[-nb:] + 1e-10)
  w = vec[:, :nb][None, None]
  p = {"w_re": np.real(w).copy(), "w_im": np.imag(w).copy()}
  parts, grad = OB.hamiltonian_trace(p, prob, v)
  assert abs(parts["total"] - np.sum(lam[:nb])) < 1e-10 * np.sum(np.abs(lam[:nb]))
  np.testing.assert_allclose(OB.band_eigenvalues(p, prob, v)[0, 0], lam[:nb],
                             atol=1e-10)
  gmax = max(np.max(np.abs(grad["w_re"])), np.max(np.abs(grad["w_im"])))
  assert gmax < 1e-9, gmax


def test_trace_gradient_central_differences():
  cr, prob, params, coeff, rho = _diamond_problem(kgrid=(2, 1, 1), nb=3)
  v = OB.effective_potential(rho, prob)
  prob.occupation = np.array([[[1.0, 0.7, 0.2], [1.0, 1.0, 0.5]]])
  parts, grad = OB.hamiltonian_trace(params, prob, v)
  rng = np.random.default_rng(0)
  h = 1e-6
  for key in ("w_re", "w_im"):
    for _ in range(6):
      idx = tuple(rng.integers(0, n) for n in params[key].shape)
      pp = {k: x.copy() for k, x in params.items()}
      pm = {k: x.copy() for k, x in params.items()}
      pp[key][idx] += h
      pm[key][idx] -= h
      fd = (OB.hamiltonian_trace(pp, prob, v, want_grad=False)[0]["total"] -
            OB.hamiltonian_trace(pm, prob, v, want_grad=False)[0]["total"]) / (2 * h)
      assert abs(fd - grad[key][idx]) < 1e-6 * max(1.0, abs(fd)), (key, idx)


def test_ground_state_band_energies_are_the_diagonal():
  """With v = dE/drho of the orbitals' own density (unpolarised LDA), the
  diagonal of the subspace Hamiltonian is oracle.adjoint.band_energies."""
  cr, prob, params, coeff, rho = _diamond_problem(kgrid=(2, 1, 1), nb=4)
  v = OB.effective_potential(rho, prob)
  hs = OB.subspace_hamiltonian(params, prob, v)
  be = OA.band_energies(params, prob)
  np.testing.assert_allclose(np.real(np.diagonal(hs, axis1=-2, axis2=-1)), be,
                             rtol=1e-12, atol=1e-12)
