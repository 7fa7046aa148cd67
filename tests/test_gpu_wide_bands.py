"""Band counts above 64 (the wide QR kernels: LDS-staged row slabs for the
Gram passes, packed-triangle Cholesky, chunked apply passes) and the
128^3 / 128-band Gamma-point cell of 64 Si atoms, against the oracle.

Tolerances as in test_gpu_parity: energies 1e-8 Ha, density 1e-10 max-abs,
gradients 1e-7 relative; the small cells are held tighter.
"""
import numpy as np
import pytest

from helpers import oracle_problem, rel_max, seeded_params
from test_gpu_parity import _case, _check, _run

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("nb", [80, 97, 128])
def test_wide_band_tiles(nb):
  """5, 7 and 8 MFMA tiles of bands (97: a padded last tile)."""
  from jrystal_amd import workloads as W
  wl = W.build(f"si2-16^3-{nb}b", W.fcc_si2(14.0), (16, 16, 16), (1, 1, 2),
               nb, 11.0)
  assert wl.num_g > nb
  _case(wl, seed=nb, e_tol=1e-9, rho_tol=1e-11, g_tol=1e-9)


def test_wide_qr_matches_householder_up_to_column_phase():
  """Q of the wide CholeskyQR2 against numpy's Householder QR: equal up to
  a unit phase per column, and orthonormal to rounding."""
  from jrystal_amd import _ops
  rng = np.random.default_rng(11)
  ns, nk, ng, nb = 1, 2, 700, 128
  wr = rng.standard_normal((ns, nk, ng, nb))
  wi = rng.standard_normal((ns, nk, ng, nb))
  mask_index = torch.arange(ng, dtype=torch.int64, device="cuda")
  box = _ops.Coefficients.apply(torch.tensor(wr, device="cuda"),
                                torch.tensor(wi, device="cuda"), mask_index,
                                (ng, 1, 1))
  q = box.reshape(ns, nk, nb, ng).cpu().numpy()
  for k in range(nk):
    qr, _ = np.linalg.qr(wr[0, k] + 1j * wi[0, k])   # (ng, nb)
    got = q[0, k].T
    ph = np.einsum("gb,gb->b", np.conj(qr), got)
    np.testing.assert_allclose(np.abs(ph), 1.0, atol=1e-12)
    np.testing.assert_allclose(got * np.conj(ph / np.abs(ph)), qr,
                               atol=1e-12)
    np.testing.assert_allclose(np.conj(got.T) @ got, np.eye(nb), atol=1e-12)


@pytest.mark.parametrize("cond", ["random", "ill"])
def test_wide_adaptive_cholesky_qr(cond):
  """The second CholeskyQR pass above 64 bands (W R1^-1 written out, then its
  Gram matrix): adaptive against two passes everywhere and the oracle."""
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import EnergyPlan
  nb = 80
  wl = W.build(f"si2-16^3-{nb}b", W.fcc_si2(4.0), (16, 16, 16), (1, 1, 2),
               nb, 11.0)
  prob = oracle_problem(wl.crystal, wl.grid_sizes, wl.k_grid, wl.num_bands,
                        wl.cutoff_energy)
  params = seeded_params(wl.shape, 9)
  if cond == "ill":
    for k in ("w_re", "w_im"):
      w = params[k]
      w[..., 1:] = w[..., :1] + 1e-4 * w[..., 1:]
  outs = {}
  for qp in (0, 2):
    cr = wl.crystal
    plan = EnergyPlan(cr.positions, cr.charges, cr.cell_vectors, wl.mask,
                      wl.kpts, wl.occupation, qr_passes=qp)
    outs[qp] = _run(plan, params)
    plan.close()
  _check(outs[0], prob, params, e_tol=1e-9, rho_tol=1e-11, g_tol=1e-8)
  for k in ("energies", "dw_re", "dw_im", "density"):
    if cond == "ill":
      assert np.array_equal(outs[0][k], outs[2][k]), k
    else:
      assert rel_max(outs[0][k], outs[2][k]) < 1e-12, k


def test_si64_gamma_128_bands():
  """64 Si atoms at the Gamma point, 128^3 box, 128 bands, 48 Ha cutoff
  (137059 coefficients per band): energies, density and gradient against
  the oracle."""
  from jrystal_amd import workloads as W
  wl = W.si64()
  assert wl.num_bands == 128 and tuple(wl.grid_sizes) == (128, 128, 128)
  _case(wl, seed=0)
