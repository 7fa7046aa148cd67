"""Checkpoints of the energy driver and trajectories of the minimiser.

The reference declares ``save_dir`` (jrystal/config.py:25,64) and a
``GroundStateEnergyOutput`` (calc_ground_state_energy_all_electrons.py:43-60)
but writes nothing.  Here a checkpoint is one ``.npz`` file of plain arrays
(no pickled objects; read with ``allow_pickle=False``):

  format            "jrystal_amd.checkpoint/1"
  step              optimiser steps taken (int)
  w_re, w_im        plane-wave parameters (spin, kpt, g, band)
  occupation        occupation in force (spin, kpt, band)
  energy_history    total energy (without the nuclear term) of every step
  adam_count        Adam step count (absent for the CG minimisers)
  adam_m/<leaf>, adam_v/<leaf>   Adam moments of every parameter leaf
  occ/<leaf>        parameters of a parameterised occupation
  meta/<key>        scalars and strings describing the run (workload, seed,
                    energy, ...)

Writes go to a temporary file that replaces the target, so a reader never
sees a half-written checkpoint.  The same format holds the snapshots of a
minimiser trajectory (``meta/kind = "snapshot"``), which the learned
preconditioner's trainer reads back.
"""
import os
from typing import Dict, Optional

import numpy as np

FORMAT = "jrystal_amd.checkpoint/1"


def _host(x):
  if hasattr(x, "detach"):
    x = x.detach().cpu().numpy()
  return np.asarray(x)


def save(path: str, *, step: int, params: Dict, occupation=None,
         energy_history=(), adam=None, occ_params: Optional[Dict] = None,
         meta: Optional[Dict] = None) -> str:
  """Write one checkpoint.  params: {'w_re', 'w_im'} (torch or numpy);
  adam: {'count': int, 'm': {leaf: array}, 'v': {leaf: array}}."""
  arrays = {"format": np.asarray(FORMAT), "step": np.asarray(int(step)),
            "w_re": _host(params["w_re"]), "w_im": _host(params["w_im"]),
            "energy_history": np.asarray(energy_history, np.float64)}
  if occupation is not None:
    arrays["occupation"] = _host(occupation)
  if adam is not None:
    arrays["adam_count"] = np.asarray(int(adam["count"]))
    for k, v in adam["m"].items():
      arrays["adam_m/" + k] = _host(v)
    for k, v in adam["v"].items():
      arrays["adam_v/" + k] = _host(v)
  for k, v in (occ_params or {}).items():
    arrays["occ/" + k] = _host(v)
  for k, v in (meta or {}).items():
    arrays["meta/" + k] = np.asarray(v)
  d = os.path.dirname(os.path.abspath(path))
  os.makedirs(d, exist_ok=True)
  tmp = path + ".tmp.npz"
  np.savez(tmp, **arrays)
  os.replace(tmp, path)
  return path


def load(path: str) -> Dict:
  """Read a checkpoint into a dict with the keys of save(): 'step',
  'params', 'occupation', 'energy_history', 'adam' (or None), 'occ_params',
  'meta'.  Raises ValueError for a file of another format."""
  with np.load(path, allow_pickle=False) as z:
    files = set(z.files)
    if "format" not in files or str(z["format"]) != FORMAT:
      raise ValueError(f"{path} is not a {FORMAT} checkpoint")
    out = {"step": int(z["step"]),
           "params": {"w_re": z["w_re"], "w_im": z["w_im"]},
           "occupation": z["occupation"] if "occupation" in files else None,
           "energy_history": z["energy_history"],
           "adam": None, "occ_params": {}, "meta": {}}
    if "adam_count" in files:
      out["adam"] = {"count": int(z["adam_count"]), "m": {}, "v": {}}
    for k in files:
      if k.startswith("adam_m/"):
        out["adam"]["m"][k[7:]] = z[k]
      elif k.startswith("adam_v/"):
        out["adam"]["v"][k[7:]] = z[k]
      elif k.startswith("occ/"):
        out["occ_params"][k[4:]] = z[k]
      elif k.startswith("meta/"):
        v = z[k]
        out["meta"][k[5:]] = v.item() if v.ndim == 0 else v
  return out
