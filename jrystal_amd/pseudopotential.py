"""Norm-conserving pseudopotentials: set-up of the tables the GPU uses.

Mirrors jrystal/pseudopotential (``NormConservingPseudopotential.create``,
``potential_local_reciprocal``, ``beta_sbt_grid`` and
``potential_nonlocal_psi_reciprocal``) for the energy of
calc_ground_state_energy_normcons.py:175-190.  Like the reference, the
tables are built once per problem on the host (UPF parsing, radial
transforms and splines with numpy/scipy); every evaluation then runs on the
GPU: the local potential replaces the Coulomb nuclei in the fused
potential stage (``EnergyPlan.set_external_potential``) and the non-local
projections, energy and gradient are HIP kernels
(``EnergyPlan.set_nonlocal``).

* UPF v2 files (pseudopotential/load.py:140-230): z_valence, the radial mesh
  (r > 0 kept, dataclass.py:126-131), PP_LOCAL / 2 and PP_DIJ / 2 (Ry to
  Ha), PP_BETA / r (the file holds r beta); an atom without projectors gets
  one zero l = 0 projector with D = 0 (dataclass.py:159-166).
* Radial transforms (sbt/sbt_numerical.py:47-81): the sum over the mesh of
  f r^2 j_l(q r) dr with forward differences, on q = linspace(1e-4, qmax,
  2 n_r), read at |G| (|G+k|) through a not-a-knot cubic spline.
* Local part (local.py:34-112): 4 pi S_0[v_loc + z/r](|G|) - 4 pi z/|G|^2,
  0 at G = 0, times exp(-i G.tau), summed over atoms, times N / vol.
* Non-local part (nloc.py:58-150): projectors 4 pi i^l exp(-i (G+k).tau)
  sum_b2 Dh[b1, b2] Y_{l_b2 m}(G+k) beta_b2(|G+k|) with Dh = V sqrt(lambda)
  (complex for a negative eigenvalue of D, as the reference takes it) and
  real spherical harmonics (l <= 3; the energy does not depend on the real
  basis while D couples equal l only, as in every shipped file).
"""
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import List

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.special import spherical_jn

DEFAULT_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                           "data", "pseudopotential", "normconserving")


def parse_upf(path: str) -> dict:
  """The UPF fields of an energy calculation (load.py:205-230)."""
  root = ET.parse(path).getroot()
  nl = root.find("PP_NONLOCAL")
  betas, dij = [], None
  if nl is not None:
    betas = [(int(e.get("angular_momentum")),
              np.array(e.text.split(), dtype=np.float64))
             for e in nl if e.tag.startswith("PP_BETA")]
    d = nl.find("PP_DIJ")
    dij = None if d is None else np.array(d.text.split(), dtype=np.float64)
  return {"PP_HEADER": dict(root.find("PP_HEADER").attrib),
          "PP_R": np.array(root.find("PP_MESH").find("PP_R").text.split(),
                           dtype=np.float64),
          "PP_LOCAL": np.array(root.find("PP_LOCAL").text.split(),
                               dtype=np.float64),
          "PP_BETA": betas, "PP_DIJ": dij}


def find_upf(dir_path: str, atom: str) -> str:
  """load.py:233-241."""
  for name in sorted(os.listdir(dir_path)):
    low = name.lower()
    if low.startswith(atom.lower() + ".") and low.endswith("upf"):
      return os.path.join(dir_path, name)
  raise ValueError(f"The pseudopotential file of element '{atom}' is not "
                   f"found in directory {dir_path}")


@dataclass
class NormConservingPseudopotential:
  """dataclass.py:44-190, one entry per atom."""
  num_atom: int
  positions: np.ndarray
  atomic_symbols: List[str]
  valence_charges: List[int]
  r_grid: List[np.ndarray] = field(default_factory=list)
  local_potential_grid: List[np.ndarray] = field(default_factory=list)
  nonlocal_beta_grid: List[np.ndarray] = field(default_factory=list)
  nonlocal_angular_momentum: List[np.ndarray] = field(default_factory=list)
  nonlocal_d_matrix: List[np.ndarray] = field(default_factory=list)

  @staticmethod
  def create(crystal, dir: str = None):
    dir = DEFAULT_DIR if dir is None else dir
    pp = NormConservingPseudopotential(
      len(crystal.symbols), np.asarray(crystal.positions, np.float64),
      list(crystal.symbols), [])
    for sym in crystal.symbols:
      upf = parse_upf(find_upf(dir, sym))
      r = upf["PP_R"]
      keep = r > 0
      pp.valence_charges.append(int(float(upf["PP_HEADER"]["z_valence"])))
      pp.r_grid.append(r[keep])
      pp.local_potential_grid.append(upf["PP_LOCAL"][keep] / 2)
      if upf["PP_BETA"]:
        nb = len(upf["PP_BETA"])
        rs = np.where(keep, r, 1.0)
        pp.nonlocal_beta_grid.append(
          np.stack([np.where(keep, v / rs, 0.0)
                    for _, v in upf["PP_BETA"]])[:, keep])
        pp.nonlocal_angular_momentum.append(
          np.array([l for l, _ in upf["PP_BETA"]]))
        pp.nonlocal_d_matrix.append(upf["PP_DIJ"].reshape(nb, nb) / 2)
      else:
        pp.nonlocal_beta_grid.append(np.zeros((1, int(keep.sum()))))
        pp.nonlocal_angular_momentum.append(np.array([0]))
        pp.nonlocal_d_matrix.append(np.zeros((1, 1)))
    return pp

  @property
  def total_valence_charge(self) -> int:
    return int(sum(self.valence_charges))


def _radial_transform(r, f, ls, qmax):
  """(q, S[len(ls), len(q)]): sum_i f(r_i) r_i^2 j_l(q r_i) dr_i."""
  dr = np.append(np.diff(r), 0.0)
  w = np.atleast_2d(f) * (r * r * dr)[None, :]
  q = np.linspace(1e-4, qmax, 2 * r.size)
  qr = np.multiply.outer(q, r)
  return q, np.stack([spherical_jn(int(l), qr) @ w[i]
                      for i, l in enumerate(ls)])


def potential_local_reciprocal(pp: NormConservingPseudopotential,
                               g_vector_grid, vol) -> np.ndarray:
  """The local pseudopotential on the G box (x, y, z), complex, in the
  normalisation of energy.external_reciprocal."""
  g = np.asarray(g_vector_grid, np.float64)
  n = int(np.prod(g.shape[:-1]))
  gn = np.linalg.norm(g, axis=-1)
  g2 = gn * gn
  g2[0, 0, 0] = 1.0
  out = np.zeros(g.shape[:-1], np.complex128)
  for a in range(pp.num_atom):
    r = pp.r_grid[a]
    z = pp.valence_charges[a]
    q, s = _radial_transform(r, pp.local_potential_grid[a] + z / r, [0],
                             float(gn.max()))
    v = 4 * np.pi * (CubicSpline(q, s[0])(gn) - z / g2)
    v[0, 0, 0] = 0.0
    out += v * np.exp(-1j * (g @ pp.positions[a]))
  return out * (n / vol)


def _real_ylm(l, u):
  """Orthonormal real spherical harmonics of degree l at unit vectors u."""
  x, y, z = u[..., 0], u[..., 1], u[..., 2]
  k = 0.5 / np.sqrt(np.pi)
  if l == 0:
    return np.full(x.shape + (1,), k)
  if l == 1:
    return np.sqrt(3.0) * k * np.stack([x, y, z], -1)
  if l == 2:
    return k * np.stack([np.sqrt(15.0) * x * y, np.sqrt(15.0) * y * z,
                         np.sqrt(5.0) / 2 * (3 * z * z - 1),
                         np.sqrt(15.0) * x * z,
                         np.sqrt(15.0) / 2 * (x * x - y * y)], -1)
  if l == 3:
    return k * np.stack([
      np.sqrt(35.0 / 8) * y * (3 * x * x - y * y),
      np.sqrt(105.0) * x * y * z, np.sqrt(21.0 / 8) * y * (5 * z * z - 1),
      np.sqrt(7.0) / 2 * z * (5 * z * z - 3),
      np.sqrt(21.0 / 8) * x * (5 * z * z - 1),
      np.sqrt(105.0) / 2 * z * (x * x - y * y),
      np.sqrt(35.0 / 8) * x * (x * x - 3 * y * y)], -1)
  raise NotImplementedError("projectors above l = 3")


def potential_nonlocal_psi_reciprocal(pp: NormConservingPseudopotential,
                                     g_vector_grid, mask, kpts) -> np.ndarray:
  """Projectors at the mask points, (k, nproj, ng) complex (rows: atom, beta,
  m; rows that vanish everywhere are left out)."""
  g = np.asarray(g_vector_grid, np.float64)
  kpts = np.asarray(kpts, np.float64).reshape(-1, 3)
  qmax = max(float(np.linalg.norm(g + k, axis=-1).max()) for k in kpts)
  gk = g[np.asarray(mask, bool)][None, :, :] + kpts[:, None, :]
  rad = np.linalg.norm(gk, axis=-1)
  unit = np.where(rad[..., None] > 0,
                  gk / np.where(rad > 0, rad, 1.0)[..., None],
                  np.array([1.0, 0.0, 0.0]))  # G+k = 0: the x direction
  lmax = int(max(int(np.max(l)) for l in pp.nonlocal_angular_momentum))
  nm = 2 * lmax + 1
  rows = []
  for a in range(pp.num_atom):
    ls = pp.nonlocal_angular_momentum[a]
    q, s = _radial_transform(pp.r_grid[a], pp.nonlocal_beta_grid[a], ls,
                             qmax)
    beta = CubicSpline(q, s, axis=1)(rad)                  # (b, k, g)
    lam, vec = np.linalg.eigh(pp.nonlocal_d_matrix[a])
    dh = vec * np.sqrt(lam.astype(np.complex128))
    phase = np.exp(-1j * (gk @ pp.positions[a]))           # (k, g)
    ang = np.zeros((len(ls),) + rad.shape + (nm,))
    for j, l in enumerate(ls):
      ang[j, ..., :2 * l + 1] = _real_ylm(int(l), unit)
    for b1, l1 in enumerate(ls):
      for m in range(nm):
        acc = np.zeros(rad.shape, np.complex128)
        for b2 in range(len(ls)):
          if dh[b1, b2] != 0:
            acc += dh[b1, b2] * ang[b2, ..., m] * beta[b2]
        if np.any(acc != 0):
          rows.append((4 * np.pi * 1j**int(l1)) * phase * acc)
  if not rows:
    return np.zeros((kpts.shape[0], 0, gk.shape[1]), np.complex128)
  return np.ascontiguousarray(np.stack(rows, axis=1))
