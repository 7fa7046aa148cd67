"""Crystal container (jrystal/_src/crystal.py:33-189).

Host-side input container: charges, positions (bohr), cell vectors (bohr,
rows).  ``create_from_file`` reads the extxyz files of the reference's
geometry/ directory (Angstrom, converted with ANGSTROM2BOHR as in
crystal.py:138-141); the reference goes through ASE, which is replaced here
by a reader for the extxyz subset those files use.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .const import ANGSTROM2BOHR

_SYMBOLS = (
  "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni "
  "Cu Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I "
  "Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt "
  "Au Hg Tl Pb Bi Po At Rn"
).split()
SYMBOL2NUMBER = {s: i + 1 for i, s in enumerate(_SYMBOLS)}


@dataclass
class Crystal:
  charges: np.ndarray
  positions: np.ndarray
  cell_vectors: np.ndarray
  spin: Optional[int] = None
  symbols: Optional[List[str]] = field(default=None)

  def __post_init__(self):
    self.charges = np.asarray(self.charges)
    self.positions = np.asarray(self.positions, dtype=np.float64)
    self.cell_vectors = np.asarray(self.cell_vectors, dtype=np.float64)
    if self.positions.ndim != 2 or self.positions.shape[-1] != 3:
      raise ValueError("positions must have shape (num_atom, 3)")
    if self.cell_vectors.shape != (3, 3):
      raise ValueError("cell_vectors must have shape (3, 3)")

  @property
  def scaled_positions(self):
    return self.positions @ np.linalg.inv(self.cell_vectors).T

  @property
  def vol(self):
    return np.abs(np.linalg.det(self.cell_vectors))

  @property
  def num_atom(self):
    return self.positions.shape[0]

  @property
  def num_electron(self):
    return int(np.sum(self.charges))

  @property
  def A(self):
    return self.cell_vectors

  @property
  def reciprocal_vectors(self):
    return 2 * np.pi * np.linalg.inv(self.cell_vectors).T

  @property
  def B(self):
    return self.reciprocal_vectors

  @staticmethod
  def create_from_file(file_path: str, spin: Optional[int] = None):
    symbols, pos, cell = _read_extxyz(file_path)
    charges = np.array([SYMBOL2NUMBER[s] for s in symbols])
    if spin is None:
      spin = int(np.sum(charges)) % 2
    return Crystal(charges=charges, positions=pos * ANGSTROM2BOHR,
                   cell_vectors=cell * ANGSTROM2BOHR, spin=spin,
                   symbols=symbols)

  @staticmethod
  def create_from_symbols(symbols, positions, cell_vectors, spin=None):
    """crystal.py:155-189: positions and cell given in Angstrom."""
    syms = _split_formula(symbols) if isinstance(symbols, str) else list(symbols)
    charges = np.array([SYMBOL2NUMBER[s] for s in syms])
    if spin is None:
      spin = int(np.sum(charges)) % 2
    return Crystal(charges=charges,
                   positions=np.asarray(positions, float) * ANGSTROM2BOHR,
                   cell_vectors=np.asarray(cell_vectors, float) * ANGSTROM2BOHR,
                   spin=spin, symbols=syms)


def _split_formula(s):
  import re
  out = []
  for sym, cnt in re.findall(r"([A-Z][a-z]?)(\d*)", s):
    out += [sym] * (int(cnt) if cnt else 1)
  return out


def _read_extxyz(path):
  with open(path) as f:
    lines = f.read().splitlines()
  n = int(lines[0].split()[0])
  comment = lines[1]
  key = 'Lattice="'
  if key not in comment:
    raise ValueError(f"{path}: no Lattice in the extxyz comment line")
  i = comment.index(key) + len(key)
  cell = np.array([float(v) for v in comment[i:comment.index('"', i)].split()])
  symbols, pos = [], []
  for ln in lines[2:2 + n]:
    parts = ln.split()
    symbols.append(parts[0])
    pos.append([float(v) for v in parts[1:4]])
  return symbols, np.array(pos), cell.reshape(3, 3)
