"""Exchange-correlation energy densities beyond LDA exchange: the oracle's
restatement of jrystal/_src/xc.py:54-106,222-253 for the functionals the
reference's configuration names (config.yaml: lda_x, lda_c_pw, gga_x_pbe,
gga_c_pbe, and '+'-joined sums of them).

The reference evaluates each term with jax_xc 0.0.8 (a port of libxc), which
is not importable here.  The forms below are the published ones:

- lda_x: Dirac exchange, eps_x = -(3/4)(3/pi)^(1/3) n^(1/3).
- lda_c_pw: Perdew & Wang, PRB 45, 13244 (1992), eq. 10 with the
  unpolarised parameters of its Table I (libxc LDA_C_PW).
- gga_x_pbe: Perdew, Burke & Ernzerhof, PRL 77, 3865 (1996): eps_x^LDA F(s),
  F = 1 + kappa - kappa / (1 + mu s^2 / kappa), kappa = 0.804,
  mu = beta pi^2 / 3.
- gga_c_pbe: PBE correlation eps_c^PW + H(rs, t) with the PW92 parameters
  of libxc's "modified" set (A = 0.0310907), beta = 0.06672455060314922,
  gamma = (1 - ln 2) / pi^2.

Only the unpolarised forms are restated: spin-polarised runs support LDA
exchange only (xc.py:56-59 scaling), as on the HIP path.  Parity with
jax_xc's exact constants and density thresholds is unpinned (jax_xc is not
importable); the tests pin the forms by their published limits (the
high-density expansion of PW92, the s = 0 and s -> infinity limits of the
PBE enhancement factor, the t -> 0 and t -> infinity limits of H) and the
functional derivatives by central differences.

Every function accepts complex input, so that derivatives are taken by
complex-step differentiation (Im f(x + i h) / h, exact to rounding), not by
hand-derived formulas.
"""
import numpy as np
import scipy.fft as sfft

CX = 0.75 * (3.0 / np.pi)**(1.0 / 3.0)
# density below which every energy density is taken as zero
DENS_THRESHOLD = 1e-15

# Perdew & Wang 1992, unpolarised row of Table I (p = 1)
PW92 = dict(A=0.031091, a1=0.21370, b1=7.5957, b2=3.5876, b3=1.6382,
            b4=0.49294)
# the same with libxc's more precise A (used by gga_c_pbe)
PW92_MOD = dict(PW92, A=0.0310907)
PBE_KAPPA = 0.804
PBE_BETA = 0.06672455060314922
PBE_MU = PBE_BETA * np.pi**2 / 3.0
PBE_GAMMA = (1.0 - np.log(2.0)) / np.pi**2

TERMS = ("lda_x", "lda_c_pw", "gga_x_pbe", "gga_c_pbe")


def parse(xc: str):
  """Terms of an xc string ('+'-joined), validated."""
  terms = [t.strip() for t in xc.split("+")]
  for t in terms:
    if t not in TERMS:
      raise NotImplementedError(f"xc term {t!r}: this path has {TERMS}")
  return terms


def is_gga(xc: str) -> bool:
  return any(t.startswith("gga") for t in parse(xc))


def _cbrt(n):
  return np.power(n, 1.0 / 3.0)


def lda_x(n):
  return -CX * _cbrt(n)


def pw92(n, par=PW92):
  """eps_c(rs), rs = (3 / (4 pi n))^(1/3), PW92 eq. 10 (p = 1)."""
  rs = _cbrt(3.0 / (4.0 * np.pi * n))
  srs = np.sqrt(rs)
  A = par["A"]
  den = 2.0 * A * (par["b1"] * srs + par["b2"] * rs + par["b3"] * rs * srs +
                   par["b4"] * rs * rs)
  return -2.0 * A * (1.0 + par["a1"] * rs) * np.log(1.0 + 1.0 / den)


def pbe_x(n, sigma):
  kf = _cbrt(3.0 * np.pi**2 * n)
  s2 = sigma / (4.0 * kf**2 * n**2)
  fx = 1.0 + PBE_KAPPA - PBE_KAPPA / (1.0 + PBE_MU * s2 / PBE_KAPPA)
  return lda_x(n) * fx


def pbe_h(n, sigma, ec):
  """PBE gradient correction H(rs, t) (phi = 1) given eps_c^unif."""
  kf = _cbrt(3.0 * np.pi**2 * n)
  t2 = sigma * np.pi / (16.0 * kf * n**2)  # t = |grad n| / (2 k_s n)
  a = (PBE_BETA / PBE_GAMMA) / (np.exp(-ec / PBE_GAMMA) - 1.0)
  at2 = a * t2
  return PBE_GAMMA * np.log(1.0 + (PBE_BETA / PBE_GAMMA) * t2 *
                            (1.0 + at2) / (1.0 + at2 + at2 * at2))


def pbe_c(n, sigma):
  ec = pw92(n, PW92_MOD)
  return ec + pbe_h(n, sigma, ec)


def eps(xc: str, n, sigma=None):
  """Energy per particle of the '+'-joined terms at density n (and |grad
  n|^2 = sigma for GGA terms); zero where n <= DENS_THRESHOLD."""
  n = np.asarray(n)
  live = np.real(n) > DENS_THRESHOLD
  ns = np.where(live, n, 1.0)
  ss = None if sigma is None else np.where(live, sigma, 0.0)
  out = np.zeros(np.shape(n), dtype=np.result_type(n, 0.0) if sigma is None
                 else np.result_type(n, sigma, 0.0))
  for t in parse(xc):
    if t == "lda_x":
      out = out + lda_x(ns)
    elif t == "lda_c_pw":
      out = out + pw92(ns)
    elif t == "gga_x_pbe":
      out = out + pbe_x(ns, ss)
    else:
      out = out + pbe_c(ns, ss)
  return np.where(live, out, 0.0)


def derivatives(xc: str, n, sigma, h=1e-30):
  """(d(n eps)/dn, d(n eps)/dsigma) by complex-step differentiation."""
  n = np.asarray(n, np.float64)
  sigma = np.asarray(sigma, np.float64)
  dn = np.imag((n + 1j * h) * eps(xc, n + 1j * h, sigma + 0j)) / h
  ds = np.imag(n * eps(xc, n + 0j, sigma + 1j * h)) / h
  return dn, ds


def gradient_boxes(rho, g_vector_grid):
  """grad_d rho = ifftn(i G_d fftn(rho)) (complex, d = x, y, z) as
  sigma_r_fn of xc.py:88-106 computes it; rho (x,y,z)."""
  rr = sfft.fftn(rho)
  return np.stack([sfft.ifftn(1j * g_vector_grid[..., d] * rr)
                   for d in range(3)])


def sigma(rho, g_vector_grid):
  """|grad rho|^2 with the complex gradient boxes (xc.py:104-106)."""
  gb = gradient_boxes(rho, g_vector_grid)
  return np.sum(np.real(np.conj(gb) * gb), axis=0)


def xc_energy(rho, g_vector_grid, vol, xc: str):
  """E_xc = vol/N sum_r eps(rho, sigma) rho (energy.py:185-211), rho
  (1,x,y,z) unpolarised."""
  if rho.shape[0] != 1 and xc != "lda_x":
    raise NotImplementedError("spin-polarised densities: lda_x only")
  r = rho[0]
  n = r.size
  s = sigma(r, g_vector_grid) if is_gga(xc) else None
  return float(np.sum(eps(xc, r, s) * r)) * vol / n


def xc_potential(rho, g_vector_grid, vol, xc: str):
  """dE_xc/drho(r) of the discretised functional (unpolarised):
  vol/N d(rho eps)/drho + 2 Re sum_d ifftn(-i G_d fftn(w grad_d rho)),
  w = vol/N d(rho eps)/dsigma (the adjoint of sigma_r_fn's FFTs)."""
  r = rho[0]
  n = r.size
  if not is_gga(xc):
    dn, _ = derivatives(xc, r, np.zeros_like(r))
    return (vol / n) * dn[None]
  gb = gradient_boxes(r, g_vector_grid)
  s = np.sum(np.real(np.conj(gb) * gb), axis=0)
  dn, ds = derivatives(xc, r, s)
  w = (vol / n) * ds
  acc = np.zeros(r.shape, complex)
  for d in range(3):
    acc += -1j * g_vector_grid[..., d] * sfft.fftn(w * gb[d])
  return ((vol / n) * dn + 2.0 * np.real(sfft.ifftn(acc)))[None]
