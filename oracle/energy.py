"""Energy terms (oracle restatement).

Follows jrystal/_src/energy.py, jrystal/_src/potential.py:24-77,121-166,
jrystal/_src/braket.py:27-72, jrystal/_src/xc.py:54-64,222-253 and
jrystal/_src/ewald.py:21-86.
"""
import numpy as np
from scipy.special import erfc

from . import pw as _pw

# Dirac/Slater exchange constant: eps_x(n) = -(3/4) (3/pi)^(1/3) n^(1/3)
CX = 0.75 * (3.0 / np.pi)**(1.0 / 3.0)
# Density at or below which the exchange energy density is taken as zero.
# libxc (through jax_xc 0.0.8) applies a functional-specific threshold of
# this order; its exact value is parity unpinned (see tests).
LDA_DENS_THRESHOLD = 1e-15


def safe_real(a, tol=1e-8):
  """utils.py:25-59."""
  if np.iscomplexobj(a):
    if np.allclose(np.imag(a), 0, atol=tol):
      return np.real(a)
    raise ValueError("Array has non-zero imaginary part")
  return a


def reciprocal_braket(bra, ket, vol):
  """braket.py:27-72: Re sum conj(bra) ket / N * vol / N, N = prod(bra.shape).
  """
  if bra.shape[-3:] != ket.shape[-3:]:
    raise ValueError(
      f"bra and ket shape are not aligned. Got {bra.shape} and {ket.shape}.")
  num_grids = np.prod(np.array(bra.shape))
  return (np.sum(np.conj(bra) * ket) / num_grids * (vol / num_grids)).real


def hartree_reciprocal(rho_rec, g_vector_grid, kohn_sham=False):
  """potential.py:24-77: 4 pi sum_s rho_s(G) / |G|^2, G=0 -> 0, halved unless
  kohn_sham."""
  assert rho_rec.ndim == g_vector_grid.shape[-1] + 1
  rho_tot = np.sum(rho_rec, axis=0)
  g2 = np.sum(g_vector_grid**2, axis=-1)
  g2[(0,) * 3] = 1
  out = rho_tot / g2
  out[(0,) * 3] = 0
  out = out * 4 * np.pi
  if not kohn_sham:
    out = out / 2
  return out


def external_reciprocal(position, charge, g_vector_grid, vol):
  """potential.py:121-166: -(N/vol) 4 pi sum_a Z_a e^{-iG.R_a}/(|G|^2+1e-10),
  G=0 -> 0."""
  g2 = np.sum(g_vector_grid**2, axis=-1)[..., None]
  si = np.exp(-1j * (g_vector_grid @ np.asarray(position).T))
  n = np.prod(g_vector_grid.shape[:-1])
  vi = np.asarray(charge, np.float64)[None, None, None, :] / (g2 + 1e-10)
  vi[0, 0, 0] = 0
  vi = vi * 4 * np.pi
  return -np.sum(vi * si, axis=-1) * n / vol


def hartree(rho_rec, g_vector_grid, vol, kohn_sham=False):
  """energy.py:29-82."""
  v = hartree_reciprocal(rho_rec, g_vector_grid, kohn_sham)[None]
  return safe_real(reciprocal_braket(v, rho_rec, vol))


def external(rho_rec, position, charge, g_vector_grid, vol):
  """energy.py:85-135."""
  v = external_reciprocal(position, charge, g_vector_grid, vol)[None]
  return safe_real(reciprocal_braket(v, rho_rec, vol))


def kinetic(g_vector_grid, kpts, coeff_grid, occupation=None):
  """energy.py:138-182: 1/2 sum f sum_G |G+k|^2 |C|^2 over the full box."""
  gk2 = np.sum((g_vector_grid[None] + kpts[:, None, None, None, :])**2, -1)
  e = np.einsum("kxyz,skbxyz->skb", gk2, _pw.absolute_square(coeff_grid))
  if occupation is not None:
    return safe_real(np.sum(e * occupation) / 2)
  return safe_real(e / 2)


def lda_x_eps(rho):
  """eps_x(rho) per particle for unpolarised LDA exchange (jax_xc lda_x,
  xc.py:54-64).  Closed form; low-density threshold parity unpinned."""
  rho = np.asarray(rho, np.float64)
  out = -CX * np.cbrt(np.maximum(rho, 0.0))
  return np.where(rho > LDA_DENS_THRESHOLD, out, 0.0)


def xc_density(density_grid, xc_type="lda_x"):
  """xc.py:222-253 for LDA exchange.  Spin-polarised input uses the
  reference's exchange scaling 0.5 * (eps(2 rho_up) + eps(2 rho_dn))
  (xc.py:56-59).  Other terms: oracle/xc.py (via xc_energy)."""
  if xc_type != "lda_x":
    raise NotImplementedError(f"xc {xc_type!r}: see oracle.xc.xc_energy")
  if density_grid.shape[0] == 2:
    return 0.5 * (lda_x_eps(2 * density_grid[0]) +
                  lda_x_eps(2 * density_grid[1]))
  return lda_x_eps(density_grid[0])


def vxc_lda_x(density_grid):
  """xc.py:109-124 for LDA exchange (xc_density with kohn_sham, xc.py:
  247-250): eps + rho_s d eps / d rho_s per spin channel, (s,x,y,z); for
  one channel d(rho eps)/d rho = 4/3 eps."""
  if density_grid.shape[0] == 1:
    return (4.0 / 3.0) * lda_x_eps(density_grid[0])[None]
  eps = xc_density(density_grid)
  out = np.empty_like(density_grid)
  for s in range(2):
    n2 = 2 * density_grid[s]
    de = np.where(n2 > LDA_DENS_THRESHOLD,
                  -CX / 3.0 * np.cbrt(np.maximum(n2, 1e-300))**-2, 0.0)
    out[s] = eps + density_grid[s] * de
  return out


def xc_energy(density_grid, g_vector_grid, vol, xc_type="lda_x",
              kohn_sham=False):
  """energy.py:185-211: sum_r eps(rho) * rho_s summed over spin, * vol/N;
  with kohn_sham (LDA exchange) the potential vxc_lda takes the place of
  eps, channel by channel."""
  assert density_grid.ndim == 4
  n = np.prod(density_grid.shape[-3:])
  if kohn_sham:
    if xc_type != "lda_x":
      raise NotImplementedError("kohn_sham xc energy: lda_x only")
    return safe_real(np.sum(vxc_lda_x(density_grid) * density_grid)) * vol / n
  if xc_type != "lda_x":
    from . import xc as _xc
    return _xc.xc_energy(density_grid, g_vector_grid, vol, xc_type)
  exc = xc_density(density_grid, xc_type)
  return safe_real(np.sum(exc[None] * density_grid)) * vol / n


def ewald_coulomb_repulsion(positions, charges, g_vector_grid, vol, ewald_eta,
                            ewald_grid):
  """ewald.py:21-86, restated pair by pair (same sums, bounded memory)."""
  positions = np.asarray(positions, np.float64)
  charges = np.asarray(charges, np.float64)
  na = positions.shape[0]
  tau = positions[None, :, :] - positions[:, None, :]  # tau[a,b] = R_b - R_a
  ovlp = np.zeros((na, na))
  for a in range(na):
    for b in range(na):
      d = tau[a, b][None, :] - ewald_grid
      r = np.sqrt(np.sum(d**2, axis=-1) + 1e-20)
      r = np.where(r <= 1e-9, 1e20, r)
      ovlp[a, b] = np.sum(erfc(ewald_eta * r) / r)
  g2 = np.sum(g_vector_grid**2, axis=-1)
  g2[0, 0, 0] = 1e16
  w = np.exp(-g2 / 4 / ewald_eta**2) / g2
  g = g_vector_grid.reshape(-1, 3)
  wf = w.reshape(-1).copy()
  wf[0] = 0.0  # the reference zeroes cos(G.tau) at G = 0
  rec = np.zeros((na, na))
  for a in range(na):
    c = np.cos(g @ tau[a].T)  # (N, na)
    rec[a] = wf @ c
  rec = rec * 4 * np.pi / vol
  ew_aa = charges @ (ovlp + rec) @ charges / 2
  ew_a = -np.sum(charges**2) * 2 * ewald_eta / np.sqrt(np.pi) / 2
  ew_a -= np.sum(charges)**2 * np.pi / ewald_eta**2 / vol / 2
  return ew_aa + ew_a


def nuclear_repulsion(position, charge, cell_vectors, g_vector_grid, vol,
                      ewald_eta, ewald_cutoff):
  """energy.py:214-243."""
  from .grid import translation_vectors
  return ewald_coulomb_repulsion(position, charge, g_vector_grid, vol,
                                 ewald_eta,
                                 translation_vectors(cell_vectors,
                                                     ewald_cutoff))


def ewald_lattice(positions, charges, cell_vectors, eta=None, rcut_sigmas=8.0):
  """Textbook Ewald energy over true lattice translations (Martin 2004,
  App. F.2), independent of the reference's grid construction.  Used to
  document how the reference constant differs from the physical one."""
  positions = np.asarray(positions, np.float64)
  charges = np.asarray(charges, np.float64)
  A = np.asarray(cell_vectors, np.float64)
  vol = abs(np.linalg.det(A))
  B = 2 * np.pi * np.linalg.inv(A).T
  if eta is None:
    eta = np.sqrt(np.pi) / vol**(1 / 3)
  rmax = rcut_sigmas / eta
  gmax = 2 * eta * rcut_sigmas
  na = [int(np.ceil(rmax * np.linalg.norm(B[i]) / (2 * np.pi))) + 1
        for i in range(3)]
  ng = [int(np.ceil(gmax * np.linalg.norm(A[i]) / (2 * np.pi))) + 1
        for i in range(3)]
  t = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in na],
                           indexing="ij"), -1).reshape(-1, 3) @ A
  gg = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in ng],
                            indexing="ij"), -1).reshape(-1, 3) @ B
  gg = gg[np.sum(gg**2, -1) > 0]
  e = 0.0
  for a in range(len(charges)):
    for b in range(len(charges)):
      d = positions[b] - positions[a] - t
      r = np.sqrt(np.sum(d**2, -1))
      m = r > 1e-9
      e += 0.5 * charges[a] * charges[b] * np.sum(erfc(eta * r[m]) / r[m])
  s = np.exp(1j * gg @ positions.T) @ charges
  g2 = np.sum(gg**2, -1)
  e += 2 * np.pi / vol * np.sum(np.abs(s)**2 * np.exp(-g2 / (4 * eta**2)) / g2)
  e -= eta / np.sqrt(np.pi) * np.sum(charges**2)
  e -= np.pi * np.sum(charges)**2 / (2 * eta**2 * vol)
  return e


def total_energy(coefficient, position, charge, g_vector_grid, kpts, vol,
                 occupation=None, kohn_sham=False, xc="lda_x", split=False):
  """energy.py:246-307 (no Ewald term; default occupation all ones)."""
  wave = _pw.wave_grid(coefficient, vol)
  occupation = (np.ones(coefficient.shape[:3]) if occupation is None
                else occupation)
  rho = _pw.wave_to_density(wave, occupation)
  rho_rec = _pw.wave_to_density_reciprocal(wave, occupation)
  e_kin = kinetic(g_vector_grid, kpts, coefficient, occupation)
  e_ext = external(rho_rec, position, charge, g_vector_grid, vol)
  e_har = hartree(rho_rec, g_vector_grid, vol, kohn_sham)
  e_xc = xc_energy(rho, g_vector_grid, vol, xc, kohn_sham)
  if split:
    return e_kin, e_ext, e_har, e_xc
  return e_kin + e_ext + e_har + e_xc
