"""The oracle's exchange-correlation terms (oracle/xc.py) pinned by what the
mathematics fixes, independently of the formulas they restate:

- PW92: the high-density expansion eps_c = c0 ln rs - c1 + O(rs ln rs) with
  c0 = (1 - ln 2) / pi^2 (the RPA coefficient that fixes A) and
  c1 = 0.046644 (PW92's c1 for the unpolarised gas, which fixes beta1);
  the low-density limit -a1/b4 / rs (Wigner-like decay of eq. 10).
- PBE exchange: F(s = 0) = 1 (LDA exchange), F -> 1 + kappa for s -> inf,
  F = 1 + mu s^2 + O(s^4) for small s.
- PBE correlation: H -> 0 for t -> 0 (uniform gas), eps_c + H -> 0 for
  t -> inf, H = beta t^2 + O(t^4) for small t (gradient expansion).
- The discretised functional's derivative (the oracle's potential) against
  central differences of the oracle energy on a small periodic box, for
  LDA and GGA sums.
"""
import numpy as np
import pytest

from oracle import xc as X


def _n(rs):
  return 3.0 / (4.0 * np.pi * rs**3)


def test_pw92_high_and_low_density_limits():
  c0 = (1.0 - np.log(2.0)) / np.pi**2
  assert abs(X.PW92["A"] - c0) < 1e-6 and abs(X.PW92_MOD["A"] - c0) < 1e-7
  for rs in (1e-5, 1e-6, 1e-7):
    ec = X.pw92(_n(rs))
    assert abs(ec - (X.PW92["A"] * np.log(rs) - 0.046644)) < 3e-5 * (
      1 + rs * 1e5)
  # rs -> inf: -2A (a1 rs) ln(1 + 1/(2A b4 rs^2)) -> -(a1 / b4) / rs, with
  # a relative correction of order (b3 / b4) rs^(-1/2)
  for rs in (1e4, 1e5, 1e6):  # (beyond, log(1 + x) rounds x away)
    ec = X.pw92(_n(rs))
    assert abs(ec * rs + X.PW92["a1"] / X.PW92["b4"]) < 2.0 / np.sqrt(rs)


def test_pbe_exchange_limits():
  n = np.array([1e-3, 0.05, 0.7, 20.0])
  assert np.allclose(X.pbe_x(n, 0.0 * n), X.lda_x(n), rtol=0, atol=0)
  big = X.pbe_x(n, 1e30 * n**(8 / 3)) / X.lda_x(n)
  assert np.allclose(big, 1.0 + X.PBE_KAPPA, rtol=1e-12)
  kf = np.cbrt(3 * np.pi**2 * n)
  s2 = 1e-8
  sigma = s2 * 4 * kf**2 * n**2
  f = X.pbe_x(n, sigma) / X.lda_x(n)
  assert np.allclose((f - 1.0) / s2, X.PBE_MU, rtol=1e-6)


def test_pbe_correlation_limits():
  n = np.array([1e-2, 0.2, 3.0])
  ec = X.pw92(n, X.PW92_MOD)
  assert np.allclose(X.pbe_c(n, 0.0 * n), ec, rtol=0, atol=0)
  assert np.allclose(X.pbe_c(n, 1e40 * n**(7 / 3)), 0.0, atol=1e-12)
  kf = np.cbrt(3 * np.pi**2 * n)
  t2 = 1e-9
  sigma = t2 * 16.0 * kf * n**2 / np.pi
  h = X.pbe_c(n, sigma) - ec
  assert np.allclose(h / t2, X.PBE_BETA, rtol=1e-6)


def test_unknown_terms_rejected():
  with pytest.raises(NotImplementedError):
    X.parse("lda_x+mgga_x_scan")


def _box_density(seed, shape=(6, 7, 8)):
  rng = np.random.default_rng(seed)
  # smooth positive density: a few low plane waves around a mean
  x = [np.arange(m) / m for m in shape]
  X0, X1, X2 = np.meshgrid(*x, indexing="ij")
  rho = 0.05 + 0.02 * np.cos(2 * np.pi * (X0 + X1)) + 0.01 * np.sin(
    2 * np.pi * (2 * X2 - X1)) + 0.005 * rng.random(shape)
  return rho[None]


def _gvec(shape, cell):
  b = 2 * np.pi * np.linalg.inv(cell).T
  m = np.meshgrid(*[np.fft.fftfreq(n, 1.0 / n) for n in shape],
                  indexing="ij")
  return np.stack(m, -1) @ b


@pytest.mark.parametrize("xc", ["lda_x", "lda_x+lda_c_pw", "gga_x_pbe",
                                "gga_x_pbe+gga_c_pbe", "gga_c_pbe"])
def test_potential_matches_central_differences(xc):
  shape = (6, 7, 8)
  cell = np.array([[5.0, 0.3, 0.0], [0.0, 5.5, 0.2], [0.4, 0.0, 6.0]])
  vol = abs(np.linalg.det(cell))
  g = _gvec(shape, cell)
  rho = _box_density(3, shape)
  v = X.xc_potential(rho, g, vol, xc)
  rng = np.random.default_rng(4)
  h = 1e-6
  for _ in range(12):
    i = tuple(rng.integers(0, s) for s in shape)
    rp, rm = rho.copy(), rho.copy()
    rp[(0,) + i] += h
    rm[(0,) + i] -= h
    fd = (X.xc_energy(rp, g, vol, xc) - X.xc_energy(rm, g, vol, xc)) / (2 * h)
    assert abs(fd - v[(0,) + i]) <= 1e-8 * max(1.0, abs(v[(0,) + i])), (
      xc, i, fd, v[(0,) + i])
