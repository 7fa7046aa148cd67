"""Run configuration (jrystal/config.py:21-120), YAML in the same keys.

Differences from the reference, by design: defaults are merged under a
YAML file's keys (the reference applies defaults only when no file is
given), and unknown keys produce a warning instead of passing silently.
"""
import copy
import warnings
from typing import Optional

import yaml

DEFAULT_CONFIG = {
  "crystal": "diamond",
  "crystal_file_path_path": None,
  "save_dir": None,
  "spin": 0,
  "xc": "lda_x",
  "use_pseudopotential": False,
  "pseudopotential_type": "nc",
  "pseudopotential_file_dir": None,
  "freq_mask_method": "spherical",
  "cutoff_energy": 100,
  "grid_sizes": 64,
  "k_grid_sizes": 3,
  "occupation": "uniform",
  "smearing": 0.001,
  "empty_bands": 8,
  "spin_restricted": True,
  "ewald_args": {"ewald_eta": 0.1, "ewald_cutoff": 2e4},
  "epoch": 5000,
  "optimizer": "adam",
  "optimizer_args": {"learning_rate": 0.01, "b1": 0.9, "b2": 0.99},
  "scheduler": None,
  "convergence_window_size": 20,
  "convergence_condition": 1e-4,
  "band_structure_empty_bands": 8,
  "k_path_special_points": None,
  "num_kpoints": 60,
  "k_path_file": None,
  "band_structure_epoch": 5000,
  "k_path_fine_tuning": True,
  "k_path_fine_tuning_epoch": 300,
  "seed": 123,
  "parallel_over_k_mesh": False,
  "parallel_over_k_path": True,
  "xla_preallocate": True,
  "jax_enable_x64": True,
  "jax_debug_nans": False,
  "verbose": True,
  "eps": 1e-8,
  "debug": False,
  # Keys of this implementation (not in the reference):
  "param_dist": "uniform",   # U[0,1) as the reference; or "normal"
  "chunk_states": 0,         # FFT chunk of the fused evaluation
  "checkpoint_every": 0,     # with save_dir: a checkpoint every N steps
                             # (0: only at the end of the run)
  "k_path_chunks": 1,        # band structure: the path in D chunks at once
                             # (the reference's parallel_over_k_path on D
                             # devices)
}


class JrystalConfigDict(dict):
  """Dict with attribute access, like ml_collections.ConfigDict."""

  def __getattr__(self, k):
    try:
      return self[k]
    except KeyError as e:
      raise AttributeError(k) from e

  def __setattr__(self, k, v):
    self[k] = v


def get_config(config_file: Optional[str] = None) -> JrystalConfigDict:
  cfg = copy.deepcopy(DEFAULT_CONFIG)
  if config_file is not None:
    with open(config_file) as f:
      user = yaml.safe_load(f) or {}
    for k in user:
      if k not in cfg:
        warnings.warn(f"unknown config key {k!r} (ignored by this path)")
    for k, v in user.items():
      if isinstance(v, dict) and isinstance(cfg.get(k), dict):
        cfg[k] = {**cfg[k], **v}
      else:
        cfg[k] = v
  cfg = JrystalConfigDict(cfg)
  if cfg.band_structure_empty_bands is None:
    cfg.band_structure_empty_bands = cfg.empty_bands
  return cfg
