"""Oracle and product set-up code against the reference's golden values and
closed forms (CPU only)."""
import os

import numpy as np
import pytest

from oracle import grid as OG
from oracle.crystal import Crystal as OCrystal

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# jrystal/_src/crystal_test.py:21-48 (geometry/diamond.xyz)
DIAMOND = dict(
  positions=np.array([[-0.84251071] * 3, [0.84251071] * 3]),
  charges=np.array([6, 6]),
  cell=np.array([[0., 3.37004284, 3.37004284], [3.37004284, 0., 3.37004284],
                 [3.37004284, 3.37004284, 0.]]),
  vol=76.5484253352856,
  scaled=np.array([[-0.125] * 3, [0.125] * 3]),
  recip=np.array([[-0.93221149, 0.93221149, 0.93221149],
                  [0.93221149, -0.93221149, 0.93221149],
                  [0.93221149, 0.93221149, -0.93221149]]),
)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_diamond_golden(impl):
  if impl == "oracle":
    c = OCrystal.create_from_file(os.path.join(GOLD, "diamond.xyz"))
  else:
    from jrystal_amd.crystal import Crystal
    c = Crystal.create_from_file(os.path.join(GOLD, "diamond.xyz"))
  np.testing.assert_almost_equal(c.positions, DIAMOND["positions"])
  np.testing.assert_almost_equal(c.charges, DIAMOND["charges"])
  np.testing.assert_almost_equal(c.cell_vectors, DIAMOND["cell"])
  # The reference asserts exact equality of JAX's LU determinant; numpy's
  # LU rounds differently in the last few ulp.
  assert abs(c.vol - DIAMOND["vol"]) < 1e-12
  np.testing.assert_almost_equal(c.scaled_positions, DIAMOND["scaled"])
  np.testing.assert_almost_equal(c.reciprocal_vectors, DIAMOND["recip"])
  assert c.num_electron == 12 and c.spin == 0


def test_si_geometries_read():
  c = OCrystal.create_from_file(os.path.join(GOLD, "si8.xyz"))
  assert c.num_atom == 8 and c.num_electron == 112
  a = 5.430957859235396 * 1.8897259886
  np.testing.assert_allclose(np.diag(c.cell_vectors), [a] * 3, rtol=1e-15)


def test_fft_size_table():
  sizes = OG.fft_sizes()
  assert len(sizes) == 190 and sizes[-1] == 2048
  for n in sizes:
    m = n
    for p in (2, 3, 5, 7):
      while m % p == 0:
        m //= p
    assert m == 1
  assert [OG.fft_factor(n) for n in (11, 13, 17, 31, 33, 64, 65, 127)] == \
      [12, 14, 18, 32, 35, 64, 70, 128]
  with pytest.raises(ValueError):
    OG.fft_factor(2049)
  np.testing.assert_array_equal(OG.proper_grid_size(31), [32, 32, 32])
  np.testing.assert_array_equal(OG.proper_grid_size([7, 11, 13]), [7, 12, 14])


def test_g_vectors_fft_order():
  a = np.array([[0, 5.13, 5.13], [5.13, 0, 5.13], [5.13, 5.13, 0]])
  b = OG.reciprocal_vectors(a)
  np.testing.assert_allclose(a @ b.T, 2 * np.pi * np.eye(3), atol=1e-14)
  g = OG.g_vectors(a, [6, 5, 4])
  np.testing.assert_allclose(g[1, 0, 0], b[0])
  np.testing.assert_allclose(g[5, 0, 0], -b[0])   # fftfreq: 0,1,2,-3,-2,-1
  np.testing.assert_allclose(g[3, 0, 0], -3 * b[0])
  np.testing.assert_allclose(g[0, 2, 0], 2 * b[1])  # 0,1,2,-2,-1
  np.testing.assert_allclose(g[0, 3, 0], -2 * b[1])
  np.testing.assert_allclose(g[2, 1, 3], 2 * b[0] + b[1] - b[2])


def test_monkhorst_pack_closed_form():
  np.testing.assert_allclose(OG.monkhorst_pack([2, 1, 1])[:, 0], [-.25, .25])
  k3 = OG.monkhorst_pack([3, 3, 3])
  assert k3.shape == (27, 3)
  np.testing.assert_allclose(k3[:3], [[-1 / 3, -1 / 3, -1 / 3],
                                      [-1 / 3, -1 / 3, 0],
                                      [-1 / 3, -1 / 3, 1 / 3]], atol=1e-15)
  np.testing.assert_allclose(k3.sum(0), 0, atol=1e-14)


@pytest.mark.parametrize("grid,ecut,count", [((32,) * 3, 24.0, 1459),
                                             ((64,) * 3, 48.0, 17071)])
def test_spherical_mask_counts(grid, ecut, count):
  """Mask sizes of the Si2 FCC and Si8 cube workloads (nearest G shell is
  far from the cut-off, so the counts are rounding-proof)."""
  from oracle.workloads import si2_fcc, si8_cube
  cr = si2_fcc() if grid[0] == 32 else si8_cube()
  m = OG.spherical_mask(cr.cell_vectors, grid, ecut)
  assert int(m.sum()) == count
  g2 = np.sum(OG.g_vectors(cr.cell_vectors, grid)**2, -1)
  assert np.min(np.abs(g2 - 2 * ecut)) > 3e-3
  # inversion symmetry G -> -G (index j -> (n - j) mod n on every axis)
  np.testing.assert_array_equal(
    m, np.roll(m[::-1, ::-1, ::-1], 1, axis=(0, 1, 2)))


def test_cubic_mask_frequencies():
  for n in (7, 8, 9, 16, 17):
    m = np.asarray(OG.cubic_mask([n, n, n]))[:, 0, 0]
    gmax = (n - 1) // 2
    lo, hi = -gmax // 2, gmax // 2
    kept = set(np.fft.fftfreq(n, 1 / n)[m].astype(int).tolist())
    assert kept == set(range(lo, hi + 1))


def test_translation_vectors_integer_lattice_grid():
  """grid.py:213-236 calls _vector_grid with its default normalize=False, so
  the rows are integer combinations m_i a_i, m_i in fftfreq(n, 1/n), with
  n = ceil(cutoff/|sum a|^2)."""
  a = 10.26 * np.eye(3)
  t = OG.translation_vectors(a, 2e4)
  n = int(np.ceil(2e4 / (3 * 10.26**2)))
  assert t.shape == (n**3, 3)
  np.testing.assert_allclose(t[1], [0, 0, 10.26])
  np.testing.assert_allclose(t[n * n], [10.26, 0, 0])
  np.testing.assert_allclose(t[-1], [-10.26, -10.26, -10.26])
  m = t / 10.26
  np.testing.assert_allclose(m, np.round(m), atol=1e-12)
  assert m.min() == -(n // 2) and m.max() == (n - 1) // 2
  fcc = 5.13 * (np.ones((3, 3)) - np.eye(3))
  tf = OG.translation_vectors(fcc, 2e4)
  frac = tf @ np.linalg.inv(fcc)
  np.testing.assert_allclose(frac, np.round(frac), atol=1e-12)


def test_product_setup_matches_oracle():
  from jrystal_amd import grid as PG
  from jrystal_amd import occupation as PO
  from jrystal_amd import workloads as PW
  from oracle import occupation as OO
  from oracle import workloads as OW
  for pw_, ow in ((PW.si2_fcc(14.0), OW.cfg1(14.0)),
                  (PW.tiny(), OW.tiny())):
    cr, prob, params = ow
    np.testing.assert_array_equal(pw_.mask, prob.mask)
    np.testing.assert_allclose(pw_.kpts, prob.kpts, rtol=0, atol=0)
    np.testing.assert_array_equal(pw_.occupation, prob.occupation)
    np.testing.assert_array_equal(pw_.crystal.positions, cr.positions)
    p = pw_.params(0)
    np.testing.assert_array_equal(p["w_re"], params["w_re"])
    np.testing.assert_array_equal(p["w_im"], params["w_im"])
  a = PW.fcc_si2().cell_vectors
  for gs in ((8, 9, 10), (12, 12, 12)):
    np.testing.assert_array_equal(PG.g_vectors(a, gs), OG.g_vectors(a, gs))
    np.testing.assert_array_equal(PG.r_vectors(a, gs), OG.r_vectors(a, gs))
    np.testing.assert_array_equal(PG.cubic_mask(gs),
                                  np.asarray(OG.cubic_mask(gs)))
  np.testing.assert_array_equal(PG.translation_vectors(a, 2e4),
                                OG.translation_vectors(a, 2e4))
  np.testing.assert_array_equal(PO.uniform(4, 28, 0, 8), OO.uniform(4, 28, 0, 8))
  np.testing.assert_array_equal(PO.gamma(4, 7, 1, 6, False),
                                OO.gamma(4, 7, 1, 6, False))


def test_uniform_occupation_clips_like_reference():
  from oracle import occupation as OO
  occ = OO.uniform(8, 28, 0, 8)  # 14 doubly occupied bands requested, 8 exist
  assert occ.shape == (1, 8, 8)
  np.testing.assert_allclose(occ, 2 / 8)
  occ = OO.uniform(8, 8, 0, 8)
  np.testing.assert_allclose(occ[0, :, :4], 0.25)
  np.testing.assert_allclose(occ[0, :, 4:], 0.0)
  occ = OO.uniform(2, 5, 1, 4, spin_restricted=False)
  np.testing.assert_allclose(occ[0, :, :3], 0.5)
  np.testing.assert_allclose(occ[1, :, :2], 0.5)
  np.testing.assert_allclose(occ[1, :, 2:], 0.0)
