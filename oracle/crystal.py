"""Crystal container and extxyz reader (oracle restatement).

Follows jrystal/_src/crystal.py:33-189 and jrystal/_src/const.py:17-18.
The reference reads files through ASE (crystal.py:123-153); ASE is not
available, so the extxyz subset used by geometry/*.xyz (a ``Lattice="..."``
comment line and ``symbol x y z`` rows, Angstrom) is parsed directly.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# const.py:17-18
BOHR2ANGSTROM = 0.529177249
ANGSTROM2BOHR = 1.8897259886

_SYMBOLS = (
  "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni "
  "Cu Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I "
  "Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt "
  "Au Hg Tl Pb Bi Po At Rn"
).split()
SYMBOL2NUMBER = {s: i + 1 for i, s in enumerate(_SYMBOLS)}


@dataclass
class Crystal:
  """crystal.py:33-117: charges (atom,), positions (atom,3) bohr, cell (3,3)
  bohr with lattice vectors as rows."""
  charges: np.ndarray
  positions: np.ndarray
  cell_vectors: np.ndarray
  spin: Optional[int] = None
  symbols: Optional[List[str]] = field(default=None)

  def __post_init__(self):
    self.charges = np.asarray(self.charges)
    self.positions = np.asarray(self.positions, dtype=np.float64)
    self.cell_vectors = np.asarray(self.cell_vectors, dtype=np.float64)

  @property
  def scaled_positions(self):  # crystal.py:84-89
    return self.positions @ np.linalg.inv(self.cell_vectors).T

  @property
  def vol(self):  # crystal.py:91-96
    return np.abs(np.linalg.det(self.cell_vectors))

  @property
  def num_atom(self):
    return self.positions.shape[0]

  @property
  def num_electron(self):  # crystal.py:103-106
    return int(np.sum(self.charges))

  @property
  def A(self):
    return self.cell_vectors

  @property
  def reciprocal_vectors(self):  # crystal.py:113-117
    return 2 * np.pi * np.linalg.inv(self.cell_vectors).T

  @property
  def B(self):
    return self.reciprocal_vectors

  @staticmethod
  def create_from_file(file_path: str, spin: Optional[int] = None):
    """crystal.py:123-153 (Angstrom -> bohr with ANGSTROM2BOHR)."""
    symbols, pos_ang, cell_ang = read_xyz(file_path)
    charges = np.array([SYMBOL2NUMBER[s] for s in symbols])
    if spin is None:
      spin = int(np.sum(charges)) % 2
    return Crystal(
      charges=charges,
      positions=pos_ang * ANGSTROM2BOHR,
      cell_vectors=cell_ang * ANGSTROM2BOHR,
      spin=spin,
      symbols=symbols,
    )


def read_xyz(path: str):
  """Minimal extxyz reader: returns (symbols, positions[A], cell[A])."""
  with open(path) as f:
    lines = [ln.rstrip("\n") for ln in f]
  n = int(lines[0].split()[0])
  comment = lines[1]
  key = 'Lattice="'
  i = comment.index(key) + len(key)
  j = comment.index('"', i)
  cell = np.array([float(v) for v in comment[i:j].split()]).reshape(3, 3)
  symbols, pos = [], []
  for ln in lines[2:2 + n]:
    parts = ln.split()
    symbols.append(parts[0])
    pos.append([float(v) for v in parts[1:4]])
  return symbols, np.array(pos), cell
