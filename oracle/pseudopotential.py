"""Norm-conserving pseudopotentials (oracle restatement).

TEST INFRASTRUCTURE ONLY (the rules of oracle/__init__.py apply).

Restates, for the energy of calc_ground_state_energy_normcons.py:175-190:

* ``parse_upf`` -- pseudopotential/load.py:140-230: the UPF v2 (XML) fields
  the energy uses: PP_HEADER z_valence and l_max, PP_MESH/PP_R, PP_LOCAL,
  PP_NONLOCAL/PP_BETA.* (values, angular_momentum) and PP_DIJ.
* ``create`` -- pseudopotential/dataclass.py:97-190: one entry per atom;
  the radial points r > 0 only; v_loc = PP_LOCAL / 2 and D = PP_DIJ / 2
  (Ry -> Ha); beta = PP_BETA / r (the file stores r beta); an atom without
  projectors gets one zero l = 0 projector with D = 0.
* ``sbt_numerical`` -- sbt/sbt_numerical.py:13-81: S_l[f](q) =
  sum_i f(r_i) r_i^2 j_l(q r_i) dr_i with forward differences dr_i
  (dr_last = 0) on q = linspace(1e-4, qmax, 2 n_r).
* ``local_potential_reciprocal`` -- pseudopotential/local.py:34-112 (the
  "sbt" method): per atom 4 pi S_0[v_loc + z/r], a not-a-knot cubic spline
  (interpax.CubicSpline, as scipy's) at |G|, minus 4 pi z / |G|^2, G = 0
  set to 0, times exp(-i G.tau), summed over atoms, times N / vol: the
  reference's external_reciprocal normalisation, so energy_local is
  reciprocal_braket(v_loc, rho(G), vol) (local.py:158-187).
* ``projectors`` -- pseudopotential/nloc.py:58-150 and beta.py:22-75:
  P[k, (atom, b1, m), G] = 4 pi i^{l_b1} exp(-i (G+k).tau)
  sum_b2 Dh[b1, b2] Y_{l_b2, m}(G+k) beta_b2(|G+k|), Dh = V sqrt(lambda + 0j)
  from eigh(D), beta_b(q) = scipy CubicSpline (not-a-knot) of
  S_{l_b}[beta_b] at q, m over 2 l_max + 1 slots (Y = 0 beyond 2 l + 1).
  A negative D eigenvalue gives an imaginary row of Dh, so it enters the
  energy as |lambda| (the reference's behaviour, reproduced).
* ``energy_nonlocal`` -- nloc.py:196-227: sum_{s,k,b} f (1/vol)
  sum_p |F_pb|^2, F = sum_G P(G) c_b(G) (no conjugate on c), and its
  gradient qbar += (2 f / vol) sum_p F_pb conj(P_pG) (torch convention).

Real spherical harmonics: any orthonormal real basis of each l gives the
same energy when D does not couple different l (then only
sum_m Y_lm(a) Y_lm(b) = (2l+1)/(4 pi) P_l(a.b) enters); every shipped UPF
file has a diagonal D.  At G + k = 0 the direction is (1, 0, 0), which the
reference's cartesian_to_spherical gives there (theta = 0, phi = pi/2).
"""
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass
from typing import List

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.special import spherical_jn


def parse_upf(path: str) -> dict:
  root = ET.parse(path).getroot()
  header = dict(root.find("PP_HEADER").attrib)
  mesh = root.find("PP_MESH")
  r = np.array([float(x) for x in mesh.find("PP_R").text.split()])
  vloc = np.array([float(x) for x in root.find("PP_LOCAL").text.split()])
  nl = root.find("PP_NONLOCAL")
  betas = []
  dij = None
  if nl is not None:
    for e in nl:
      if e.tag.startswith("PP_BETA"):
        betas.append((int(e.get("angular_momentum")),
                      np.array([float(x) for x in e.text.split()])))
    d = nl.find("PP_DIJ")
    if d is not None:
      dij = np.array([float(x) for x in d.text.split()])
  return {"header": header, "r": r, "local": vloc, "betas": betas,
          "dij": dij}


def find_upf(dir_path: str, symbol: str) -> str:
  """load.py:233-241: the file named '<symbol>.*upf' (case-insensitive)."""
  for name in sorted(os.listdir(dir_path)):
    low = name.lower()
    if low.startswith(symbol.lower() + ".") and low.endswith("upf"):
      return os.path.join(dir_path, name)
  raise ValueError(f"no pseudopotential file for {symbol!r} in {dir_path}")


@dataclass
class NCPP:
  positions: np.ndarray            # (atom, 3) bohr
  valence_charges: List[int]       # per atom
  r_grid: List[np.ndarray]         # per atom, r > 0
  local_potential: List[np.ndarray]
  betas: List[np.ndarray]          # per atom (nbeta, nr)
  angular_momentum: List[np.ndarray]
  d_matrix: List[np.ndarray]       # per atom (nbeta, nbeta), Ha


def create(symbols, positions, dir_path: str) -> NCPP:
  r_grid, vloc, betas, ams, dms, zv = [], [], [], [], [], []
  for sym in symbols:
    pp = parse_upf(find_upf(dir_path, sym))
    r = pp["r"]
    pos = r > 0
    zv.append(int(float(pp["header"]["z_valence"])))
    r_grid.append(r[pos])
    vloc.append(pp["local"][pos] / 2)
    if pp["betas"]:
      nb = len(pp["betas"])
      b = np.stack([np.divide(v, r, out=np.zeros_like(v), where=pos)
                    for _, v in pp["betas"]])[:, pos]
      betas.append(b)
      ams.append(np.array([l for l, _ in pp["betas"]]))
      dms.append(pp["dij"].reshape(nb, nb) / 2)
    else:
      betas.append(np.zeros((1, int(pos.sum()))))
      ams.append(np.array([0]))
      dms.append(np.zeros((1, 1)))
  return NCPP(np.asarray(positions, float), zv, r_grid, vloc, betas, ams, dms)


def sbt_numerical(r, f, l, qmax):
  """(q, S) with S[i, :] the transform of f[i] with angular momentum l[i]."""
  f = np.atleast_2d(f)
  ls = np.broadcast_to(np.atleast_1d(l), (f.shape[0],))
  dr = np.zeros_like(r)
  dr[:-1] = r[1:] - r[:-1]
  q = np.linspace(1e-4, qmax, 2 * len(r))
  qr = q[:, None] * r[None, :]
  out = np.empty((f.shape[0], len(q)))
  for i in range(f.shape[0]):
    out[i] = (spherical_jn(int(ls[i]), qr) * (f[i] * r**2 * dr)[None, :]).sum(-1)
  return q, out


def local_potential_reciprocal(pp: NCPP, g_vector_grid, vol):
  g = np.asarray(g_vector_grid, float)
  gr = np.sqrt(np.sum(g**2, axis=-1))
  n = int(np.prod(g.shape[:-1]))
  gr_safe = gr.copy()
  gr_safe[0, 0, 0] = 1e10
  total = np.zeros(g.shape[:-1], complex)
  for a in range(len(pp.r_grid)):
    r, v, z = pp.r_grid[a], pp.local_potential[a], pp.valence_charges[a]
    q, s = sbt_numerical(r, v + z / r, 0, np.max(gr))
    vk = 4 * np.pi * CubicSpline(q, s[0])(gr)
    vk = vk - 4 * np.pi * z / gr_safe**2
    vk[0, 0, 0] = 0.0
    total += vk * np.exp(-1j * (g @ pp.positions[a]))
  return total * n / vol


def real_sph_harm(l: int, x):
  """Orthonormal real spherical harmonics of degree l <= 3 at the directions
  of x (..., 3); (..., 2l + 1).  The zero vector takes the direction
  (1, 0, 0)."""
  x = np.asarray(x, float)
  r = np.linalg.norm(x, axis=-1)
  u = np.where(r[..., None] > 0, x / np.where(r > 0, r, 1.0)[..., None],
               np.array([1.0, 0.0, 0.0]))
  X, Y, Z = u[..., 0], u[..., 1], u[..., 2]
  c = 1.0 / np.sqrt(4 * np.pi)
  if l == 0:
    return np.full(X.shape + (1,), c)
  if l == 1:
    return np.sqrt(3.0) * c * np.stack([Y, Z, X], -1)
  if l == 2:
    s15 = np.sqrt(15.0) * c
    return np.stack([s15 * X * Y, s15 * Y * Z,
                     np.sqrt(5.0) * c * 0.5 * (3 * Z * Z - 1),
                     s15 * X * Z, s15 * 0.5 * (X * X - Y * Y)], -1)
  if l == 3:
    return c * np.stack([
      np.sqrt(35.0 / 8.0) * Y * (3 * X * X - Y * Y),
      np.sqrt(105.0) * X * Y * Z,
      np.sqrt(21.0 / 8.0) * Y * (5 * Z * Z - 1),
      np.sqrt(7.0) * 0.5 * Z * (5 * Z * Z - 3),
      np.sqrt(21.0 / 8.0) * X * (5 * Z * Z - 1),
      np.sqrt(105.0) * 0.5 * Z * (X * X - Y * Y),
      np.sqrt(35.0 / 8.0) * X * (X * X - 3 * Y * Y)], -1)
  raise NotImplementedError("angular momentum above 3")


def projectors(pp: NCPP, g_vector_grid, mask, kpts):
  """P (k, nproj, ng) complex at the mask points (C order), projector rows
  (atom, b1, m) with all-zero rows dropped.  The transform grid reaches
  max |G+k| over the whole box and every k, as beta.py:63-66 sets it."""
  kpts = np.asarray(kpts, float).reshape(-1, 3)
  g = np.asarray(g_vector_grid, float)
  qmax = max(float(np.max(np.sqrt(np.sum((g + k)**2, axis=-1))))
             for k in kpts)
  gm = g[np.asarray(mask, bool)]
  gk = gm[None, :, :] + kpts[:, None, :]            # (k, g, 3)
  rad = np.sqrt(np.sum(gk**2, axis=-1))
  lmax = int(max(int(np.max(a)) for a in pp.angular_momentum))
  M = 2 * lmax + 1
  rows = []
  for a in range(len(pp.r_grid)):
    r, b, ls, d = pp.r_grid[a], pp.betas[a], pp.angular_momentum[a], \
      pp.d_matrix[a]
    q, s = sbt_numerical(r, b, ls, qmax)
    bq = CubicSpline(q, s, axis=1)(rad)              # (nbeta, k, g)
    lam, vec = np.linalg.eigh(d)
    dh = vec * np.sqrt(lam + 0j)                     # (b1, b2)
    sf = np.exp(-1j * np.einsum("kgc,c->kg", gk, pp.positions[a]))
    ylm = np.zeros((len(ls), gk.shape[0], gk.shape[1], M))
    for i, l in enumerate(ls):
      ylm[i, ..., :2 * l + 1] = real_sph_harm(int(l), gk)
    # sum_b2 Dh[b1, b2] Y_{l_b2, m} beta_b2
    t = np.einsum("ij,jkgm,jkg->ikgm", dh, ylm, bq)   # (b1, k, g, m)
    for i, l in enumerate(ls):
      for m in range(M):
        row = 4 * np.pi * (1j ** int(l)) * sf * t[i, :, :, m]
        if np.any(row != 0):
          rows.append(row)
  if not rows:
    return np.zeros((kpts.shape[0], 0, gm.shape[0]), complex)
  return np.stack(rows, axis=1)


def energy_nonlocal(q, P, occupation, vol, want_grad=True):
  """q (s, k, g, b) orthonormal columns, P (k, nproj, g) -> (E, qbar,
  per-band e (s, k, b) without the occupation)."""
  F = np.einsum("kpg,skgb->skpb", P, q)
  e_band = np.sum(np.abs(F)**2, axis=2) / vol
  E = float(np.sum(occupation * e_band))
  qbar = None
  if want_grad:
    qbar = (2.0 / vol) * np.einsum("skpb,kpg->skgb", F, np.conj(P)) * \
      occupation[:, :, None, :]
  return E, qbar, e_band
