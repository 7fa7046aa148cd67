"""Writes tests/golden/oracle_golden.npz: oracle energies, density checksums
and gradient samples for small seeded workloads.  The reference (JAX) cannot
run in this environment, so these vectors come from the oracle after it has
been pinned by the closed-form, brute-force, autograd and finite-difference
tests; they guard the oracle (and the GPU path) against drift.

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import adjoint as OA  # noqa: E402
from oracle import workloads as OW  # noqa: E402

CASES = {
  "tiny_z4": lambda: OW.tiny(z=4.0),
  "si2_z14": lambda: OW.cfg1(z=14.0),
  "si2_z4": lambda: OW.cfg1(z=4.0),
}


def sample_index(shape, n=64, seed=123):
  rng = np.random.default_rng(seed)
  return np.stack([rng.integers(0, s, n) for s in shape], axis=1)


def compute(name):
  cr, prob, params = CASES[name]()
  parts, grad, rho = OA.energy_and_grad(params, prob)
  idx = sample_index(params["w_re"].shape)
  t = tuple(idx.T)
  return {
    f"{name}/energies": np.array([parts[k] for k in
                                  ("kinetic", "hartree", "external", "xc",
                                   "total")]),
    f"{name}/rho_sum": np.array([rho.sum(), (rho**2).sum(), rho.max()]),
    f"{name}/grad_idx": idx,
    f"{name}/grad_re": grad["w_re"][t],
    f"{name}/grad_im": grad["w_im"][t],
    f"{name}/grad_norm": np.array([np.linalg.norm(grad["w_re"]),
                                   np.linalg.norm(grad["w_im"])]),
  }


def main():
  out = {}
  for name in CASES:
    out.update(compute(name))
  path = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                      "oracle_golden.npz")
  np.savez(path, **out)
  print(path, sorted(out))


if __name__ == "__main__":
  main()
