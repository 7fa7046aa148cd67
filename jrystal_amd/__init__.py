"""jrystal_amd: MI355X-native implementation of jrystal's direct-minimisation
total-energy path (crystal / grid / pw / energy / occupation API of
sail-sg/jrystal, HIP kernels for gfx950 behind a C ABI)."""
import os

from . import const, crystal, grid, occupation, workloads  # noqa: F401
from .crystal import Crystal  # noqa: F401

__version__ = "0.1.0"


def get_pkg_path():
  """jrystal/__init__.py:39-40."""
  return os.path.dirname(os.path.abspath(__file__))
