"""Per-shape autotuner of the fused evaluation.

For one problem shape it times a small set of plan variants and keeps the
fastest:

* FFT variant: the compile-time-radix kernels where they exist (64^3
  uses the register-resident 64-point kernels) or the runtime-radix ones;
* x-line workgroup target: how many state groups the x-line passes split
  the states into (2048, 4096 or 8192 workgroups);
* state chunking: all states in one chunk (forward planes kept for the
  adjoint sweep) or half / a quarter of them per chunk.

Every variant computes the same energy and gradient; only the schedule
changes.  Results are cached in a JSON file keyed by the shape (grid,
G-sphere size, spins, k-points, bands).
"""
import json
import os
import time

import numpy as np
import torch

from .objective import EnergyPlan

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "jrystal_amd",
                             "autotune.json")


def shape_key(wl) -> str:
  ns, nk, ng, nb = wl.shape
  gs = "x".join(str(int(x)) for x in wl.grid_sizes)
  return f"{gs}-g{ng}-s{ns}-k{nk}-b{nb}"


def candidates(wl):
  ns, nk, ng, nb = wl.shape
  S = ns * nk * nb
  out = [{"chunk_states": 0, "fft_variant": 0, "xline_blocks": 0}]
  for xb in (2048, 8192):
    out.append({"chunk_states": 0, "fft_variant": 0, "xline_blocks": xb})
  if tuple(wl.grid_sizes) == (64, 64, 64):
    out.append({"chunk_states": 0, "fft_variant": 1, "xline_blocks": 0})
  for div in (2, 4):
    c = -(-S // div)
    if c < S and c >= 16:
      out.append({"chunk_states": c, "fft_variant": 0, "xline_blocks": 0})
  return out


def _time(plan, w_re, w_im, reps):
  out = {}
  plan.evaluate(w_re, w_im, grad=True, out=out)  # warm-up
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(reps):
    plan.evaluate(w_re, w_im, grad=True, out=out)
  torch.cuda.synchronize()
  return (time.perf_counter() - t0) / reps


def tune(wl, cache_path: str = DEFAULT_CACHE, reps: int = 3, device=None,
         refresh: bool = False, log=None):
  """Fastest plan options for wl's shape (cached)."""
  key = shape_key(wl)
  cache = {}
  if cache_path and os.path.exists(cache_path) and not refresh:
    try:
      with open(cache_path) as f:
        cache = json.load(f)
    except (OSError, ValueError):
      cache = {}
  if key in cache:
    return dict(cache[key]["options"])
  p = wl.params(0)
  dev = torch.device(device if device is not None else "cuda")
  w_re = torch.tensor(p["w_re"], device=dev)
  w_im = torch.tensor(p["w_im"], device=dev)
  results = []
  for opt in candidates(wl):
    plan = EnergyPlan.from_workload(wl, device=dev, **opt)
    t = _time(plan, w_re, w_im, reps)
    plan.close()
    results.append((t, opt))
    if log:
      log(f"autotune {key} {opt}: {t * 1e3:.3f} ms")
  t_best, best = min(results, key=lambda r: r[0])
  cache[key] = {"options": best, "ms": t_best * 1e3,
                "tried": [{"options": o, "ms": t * 1e3} for t, o in results]}
  if cache_path:
    os.makedirs(os.path.dirname(cache_path), exist_ok=True)
    tmp = cache_path + ".tmp"
    with open(tmp, "w") as f:
      json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, cache_path)
  return dict(best)


def speedup_table(cache_path: str = DEFAULT_CACHE):
  """(key, default ms, tuned ms) rows of a cache file."""
  with open(cache_path) as f:
    cache = json.load(f)
  rows = []
  for key, v in sorted(cache.items()):
    default = next((t["ms"] for t in v["tried"]
                    if t["options"] == candidates_default()), np.nan)
    rows.append((key, default, v["ms"]))
  return rows


def candidates_default():
  return {"chunk_states": 0, "fft_variant": 0, "xline_blocks": 0}
