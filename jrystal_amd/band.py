"""k-paths of the band-structure driver (jrystal/_src/band/k_path.py:27-57).

The reference wraps ASE (``ase.cell.Cell.bandpath``; ASE is an unpinned
dependency of the reference, requirements.txt, absent here).  This module
restates the published algorithm of ASE 3.22 for the lattices that ASE
reports in their standard form:

* special points: the fractional coordinates (in the reciprocal basis of the
  cell) of ASE's tables for the cubic (CUB), face-centred cubic (FCC),
  body-centred cubic (BCC), tetragonal (TET), orthorhombic (ORC, a < b < c)
  and hexagonal (HEX, gamma = 120 degrees) lattices, and each lattice's
  default path (used when ``path`` is None);
* the lattice is recognised from the cell's metric (lengths and angles of
  ASE's standard cell of that lattice, to 1e-6 relative); other cells raise
  NotImplementedError (give the k-points explicitly instead);
* ``paths2kpts`` (ase/dft/kpoints.py): segment lengths |(P_{i+1} - P_i) B|,
  zero across a ',' break; segment i gets
  n_i = max(2, round(L_i (npoints - placed) / (L - x0))) points of
  linspace(0, 1, n_i) without its end, and the path's last point ends it.

The result is returned in fractional coordinates or as Cartesian k = frac B
in 1/bohr (B = 2 pi inv(cell)^T), as get_k_path does.
"""
import re
from typing import Optional

import numpy as np

_F3 = 1.0 / 3.0

SPECIAL_POINTS = {
  "CUB": {"G": [0, 0, 0], "M": [0.5, 0.5, 0], "R": [0.5, 0.5, 0.5],
          "X": [0, 0.5, 0]},
  "FCC": {"G": [0, 0, 0], "K": [0.375, 0.375, 0.75], "L": [0.5, 0.5, 0.5],
          "U": [0.625, 0.25, 0.625], "W": [0.5, 0.25, 0.75],
          "X": [0.5, 0, 0.5]},
  "BCC": {"G": [0, 0, 0], "H": [0.5, -0.5, 0.5], "P": [0.25, 0.25, 0.25],
          "N": [0, 0, 0.5]},
  "TET": {"G": [0, 0, 0], "A": [0.5, 0.5, 0.5], "M": [0.5, 0.5, 0],
          "R": [0, 0.5, 0.5], "X": [0, 0.5, 0], "Z": [0, 0, 0.5]},
  "ORC": {"G": [0, 0, 0], "R": [0.5, 0.5, 0.5], "S": [0.5, 0.5, 0],
          "T": [0, 0.5, 0.5], "U": [0.5, 0, 0.5], "X": [0.5, 0, 0],
          "Y": [0, 0.5, 0], "Z": [0, 0, 0.5]},
  "HEX": {"G": [0, 0, 0], "A": [0, 0, 0.5], "H": [_F3, _F3, 0.5],
          "K": [_F3, _F3, 0], "L": [0.5, 0, 0.5], "M": [0.5, 0, 0]},
}

DEFAULT_PATHS = {
  "CUB": "GXMGRX,MR",
  "FCC": "GXWKGLUWLK,UX",
  "BCC": "GHNGPH,PN",
  "TET": "GXMGZRAZXR,MA",
  "ORC": "GXSYGZURTZ,YT,UX,SR",
  "HEX": "GMKGALHA,LM,KH",
}


def _close(a, b, tol=1e-6):
  return abs(a - b) <= tol * max(1.0, abs(b))


def lattice_type(cell_vectors) -> str:
  """Name of the lattice whose standard (ASE) cell has this cell's metric."""
  a = np.asarray(cell_vectors, np.float64)
  ln = np.linalg.norm(a, axis=1)
  cos = [np.dot(a[1], a[2]) / (ln[1] * ln[2]),
         np.dot(a[0], a[2]) / (ln[0] * ln[2]),
         np.dot(a[0], a[1]) / (ln[0] * ln[1])]
  eq01, eq12 = _close(ln[0], ln[1]), _close(ln[1], ln[2])
  if all(_close(c, 0.0) for c in cos):
    if eq01 and eq12:
      return "CUB"
    if eq01:
      return "TET"
    if ln[0] < ln[1] < ln[2] and not eq12:
      return "ORC"
  elif eq01 and eq12 and all(_close(c, 0.5) for c in cos):
    return "FCC"
  elif eq01 and eq12 and all(_close(c, -1.0 / 3.0) for c in cos):
    return "BCC"
  elif (eq01 and _close(cos[0], 0.0) and _close(cos[1], 0.0) and
        _close(cos[2], -0.5)):
    return "HEX"
  raise NotImplementedError(
    "k-path: the cell is not in the standard form of a CUB, FCC, BCC, TET, "
    "ORC (a < b < c) or HEX (gamma = 120) lattice; give the k-points "
    "explicitly (k_path_file)")


def parse_path_string(path: str):
  """'GXWK,UX' -> [['G','X','W','K'], ['U','X']] (one uppercase letter with
  optional lowercase letters or digits per point)."""
  return [[n for n in re.split(r"([A-Z][a-z0-9]*)", p) if n]
          for p in path.split(",")]


def paths2kpts(paths, reciprocal, npoints: int):
  """Fractional k-points along the paths (lists of (3,) fractional points),
  ASE's distribution over segments; returns (kpts, x, X) with x the
  coordinate along the path and X that of the special points."""
  points = np.concatenate([np.asarray(p, np.float64).reshape(-1, 3)
                           for p in paths])
  dists = points[1:] - points[:-1]
  lengths = [float(np.linalg.norm(d)) for d in dists @ reciprocal]
  i = 0
  for p in paths[:-1]:
    i += len(p)
    lengths[i - 1] = 0.0
  length = sum(lengths)
  kpts, x, X = [], [], [0.0]
  x0 = 0.0
  for P, d, L in zip(points[:-1], dists, lengths):
    diff = length - x0
    if abs(diff) < 1e-6:
      n = 0
    else:
      n = max(2, int(round(L * (npoints - len(x)) / diff)))
    for t in np.linspace(0, 1, n)[:-1]:
      kpts.append(P + t * d)
      x.append(x0 + t * L)
    x0 += L
    X.append(x0)
  if len(points):
    kpts.append(points[-1])
    x.append(x0)
  kpts = np.array(kpts) if kpts else np.empty((0, 3))
  return kpts, np.array(x), np.array(X)


def get_k_path(cell_vectors, path: Optional[str], num: int,
               fractional: bool = False) -> np.ndarray:
  """k_path.py:27-57: the k-points of ``path`` (special-point letters, ','
  between disconnected paths; None: the lattice's default path) with
  ``num`` points, fractional or Cartesian in 1/bohr."""
  a = np.asarray(cell_vectors, np.float64)
  lat = lattice_type(a)
  special = SPECIAL_POINTS[lat]
  if path is None:
    path = DEFAULT_PATHS[lat]
  names = parse_path_string(path)
  for p in names:
    for n in p:
      if n not in special:
        raise ValueError(f"special point {n!r} is not one of the {lat} "
                         f"lattice's {sorted(special)}")
  pts = [[special[n] for n in p] for p in names]
  B = 2 * np.pi * np.linalg.inv(a).T
  kpts = paths2kpts(pts, B, int(num))[0]
  return kpts if fractional else kpts @ B
