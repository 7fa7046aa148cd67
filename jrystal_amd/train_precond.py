"""Train the learned preconditioner / warm-start network on trajectories of
the plain method over random cells, with the project's HIP kernels.

  python -m jrystal_amd.train_precond --out model.npz [--trajectories N]
      [--seconds S] [--epochs E] [--data dataset.npz]

Random cells (numpy.random.default_rng(seed)):
  * fcc2  -- 2-atom FCC cells, a = 10.26 (1 + U(-0.05, 0.05)) bohr, a random
    symmetric strain of amplitude 0.1, atoms displaced by N(0, 0.15) bohr,
    Z in {3, 4}; 32^3 grid, 2x2x2 k-mesh, 8 bands, cutoff 24 Ha;
  * cube8 -- 8-atom diamond cubes, a = 10.26 bohr, strain amplitude 0.1,
    displacements N(0, 0.15) bohr, Z = 4; 64^3 grid, Gamma or a 2x2x2
    k-mesh, 32 bands, cutoff 48 Ha.
  * uneven -- cells drawn like the uneven batch of workloads.batch_cells: a
    cubic or FCC cell of 2-8 atoms (volume per atom U(60, 140) bohr^3),
    Z from {3, 6, 11, 13, 14}, uniform random positions, a cubic grid from
    the 7-smooth sizes in [32, 48], a k-mesh of 1^3 or 2^3, 4-24 bands and
    the grid's alias-free cutoff.
Mixes of kinds: "small" cycles fcc2, fcc2, fcc2, cube8 (Gamma); "wide"
cycles fcc2, fcc2, cube8 (Gamma), cube8 (2x2x2); "uneven" cycles fcc2,
uneven, cube8 (Gamma), fcc2, uneven, cube8 (2x2x2).
Cells whose minimum-image atom separation is below 2.5 bohr are redrawn.
Trajectory seeds start at 1000, away from the strained held-out cell
(strain seed 4).
"""
import argparse
import time

import numpy as np

from . import learned
from . import workloads as W
from .crystal import Crystal


def _min_separation(cell, pos):
  best = np.inf
  shifts = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1)
                     for k in (-1, 0, 1)], float) @ cell
  for a in range(len(pos)):
    for b in range(len(pos)):
      d = pos[b] - pos[a] + shifts
      n = np.linalg.norm(d, axis=1)
      if a == b:
        n = n[n > 1e-9]
      best = min(best, n.min())
  return best


def uneven_cell(seed: int):
  """A cell of the uneven-batch kind, at training sizes."""
  rng = np.random.default_rng(seed)
  lattice = ("cubic", "fcc")[int(rng.integers(2))]
  natom = int(rng.integers(2, 9))
  z = float(rng.choice([3.0, 6.0, 11.0, 13.0, 14.0]))
  vol = natom * rng.uniform(60.0, 140.0)
  if lattice == "cubic":
    cell = vol**(1 / 3) * np.eye(3)
  else:
    a = (4 * vol)**(1 / 3)
    cell = a / 2 * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0.]])
  while True:
    pos = rng.random((natom, 3)) @ cell
    if _min_separation(cell, pos) >= 2.5:
      break
  ng = int(rng.choice([n for n in W.SMOOTH_32_80 if n <= 48]))
  k = int(rng.integers(1, 3))
  nb = int(rng.integers(4, 25))
  gs = (ng, ng, ng)
  cr = Crystal(charges=np.full(natom, z), positions=pos, cell_vectors=cell,
               spin=0, symbols=["X"] * natom)
  return W.build(f"uneven-{seed}", cr, gs, (k, k, k), nb,
                 W.alias_free_cutoff(cell, gs))


def random_cell(seed: int, kind: str, k: int = 1):
  if kind == "uneven":
    return uneven_cell(seed)
  rng = np.random.default_rng(seed)
  while True:
    strain = W.strain_matrix(int(rng.integers(1 << 30)), 0.1)
    if kind == "fcc2":
      z = float(rng.choice([3.0, 4.0]))
      a = W.SI_A * (1 + rng.uniform(-0.05, 0.05))
      base = W.fcc_si2(z)
      cell = base.cell_vectors * (a / W.SI_A) @ (np.eye(3) + strain)
      frac = np.array([[0, 0, 0], [0.25, 0.25, 0.25]])
      grid, kgrid, nb, ecut = (32,) * 3, (2, 2, 2), 8, 24.0
    else:
      z = 4.0
      base = W.diamond_cube(z)
      cell = base.cell_vectors @ (np.eye(3) + strain)
      frac = base.positions @ np.linalg.inv(base.cell_vectors)
      grid, kgrid, nb, ecut = (64,) * 3, (k, k, k), 32, 48.0
    pos = frac @ cell + rng.normal(0, 0.15, (len(frac), 3))
    if _min_separation(cell, pos) < 2.5:
      continue
    cr = Crystal(charges=np.full(len(pos), z), positions=pos,
                 cell_vectors=cell, spin=0, symbols=["X"] * len(pos))
    return W.build(f"{kind}-{seed}", cr, grid, kgrid, nb, ecut)


MIXES = {
  "small": (("fcc2", 1), ("fcc2", 1), ("fcc2", 1), ("cube8", 1)),
  "wide": (("fcc2", 1), ("fcc2", 1), ("cube8", 1), ("cube8", 2)),
  "uneven": (("fcc2", 1), ("uneven", 1), ("cube8", 1), ("fcc2", 1),
             ("uneven", 1), ("cube8", 2)),
}


def generate(n: int, seconds: float, seed0: int = 1000, mix: str = "small",
             max_snaps: int = 24, max_iter: int = 400, log=print):
  parts, t0 = [], time.perf_counter()
  kinds = MIXES[mix]
  for i in range(n):
    if time.perf_counter() - t0 > seconds:
      break
    kind, k = kinds[i % len(kinds)]
    wl = random_cell(seed0 + i, kind, k)
    parts.append(learned.collect_trajectory(wl, seed=seed0 + i,
                                            max_snaps=max_snaps,
                                            max_iter=max_iter))
    if log and (i % 20 == 0):
      log(f"trajectory {i} ({kind}): {len(parts[-1].t2)} snapshots, "
          f"{time.perf_counter() - t0:.1f} s")
  return learned.Dataset.concat(parts), len(parts)


def main(argv=None):
  ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
  ap.add_argument("--out", required=True)
  ap.add_argument("--data", default=None, help="save / reuse the dataset")
  ap.add_argument("--trajectories", type=int, default=2000)
  ap.add_argument("--seconds", type=float, default=600.0)
  ap.add_argument("--epochs", type=int, default=1500)
  ap.add_argument("--lr", type=float, default=2e-3)
  ap.add_argument("--lr-end", type=float, default=5e-5)
  ap.add_argument("--val-frac", type=float, default=0.1)
  ap.add_argument("--mix", default="small", choices=sorted(MIXES))
  ap.add_argument("--max-snaps", type=int, default=24)
  ap.add_argument("--max-iter", type=int, default=400)
  ap.add_argument("--no-gap", action="store_true",
                  help="train without the band-gap feature")
  ap.add_argument("--both-features", action="store_true",
                  help="train with and without the band-gap feature")
  ap.add_argument("--select-cells", default="",
                  help="strain seeds:k-mesh of validation cells, e.g. "
                       "11:2,12:2,13:2; selects the checkpoint with the "
                       "fewest CG iterations relative to plain TPA")
  ap.add_argument("--checkpoint-every", type=int, default=100)
  ap.add_argument("--also", default="",
                  help="existing model files to include in the selection")
  args = ap.parse_args(argv)
  import os
  if args.data and os.path.exists(args.data):
    ds = learned.Dataset.load(args.data)
    ntraj = -1
  else:
    ds, ntraj = generate(args.trajectories, args.seconds, mix=args.mix,
                         max_snaps=args.max_snaps, max_iter=args.max_iter)
    if args.data:
      ds.save(args.data)
  print(f"dataset: {ntraj} trajectories, {len(ds.t2)} snapshots, "
        f"{ds.x_ws.shape[0]} warm-start samples", flush=True)
  variants = [not args.no_gap]
  if args.both_features:
    variants = [False, True]
  cands, labels = [], []
  net = hist = None
  for use_gap in variants:
    ckpts = []
    net, hist = learned.train(ds, epochs=args.epochs, lr=args.lr,
                              lr_end=args.lr_end, val_frac=args.val_frac,
                              use_gap=use_gap,
                              log_every=max(1, args.epochs // 10),
                              checkpoint_every=(args.checkpoint_every
                                                if args.select_cells else 0),
                              checkpoints=ckpts)
    cands += [net] + [c for _, c in ckpts]
    labels += [f"gap={use_gap} best-val"] + \
        [f"gap={use_gap} epoch {e}" for e, _ in ckpts]
  for path in filter(None, args.also.split(",")):
    cands.append(learned.PrecondNet.load(path))
    labels.append(f"file {path}")
  chosen, label = net, labels[0] if labels else ""
  if args.select_cells:
    cells = [W.si8_strained(4.0, k=int(k), seed=int(sd)) for sd, k in
             (c.split(":") for c in args.select_cells.split(","))]
    best, scores = learned.select_by_convergence(cands, cells, log=print)
    for lab, sc in zip(labels, scores):
      print(f"  {lab}: mean iteration ratio {sc:.3f}", flush=True)
    chosen, label = cands[best], labels[best]
    print(f"selected {label}", flush=True)
  chosen.save(args.out, trajectories=ntraj, snapshots=len(ds.t2),
              history=np.asarray(hist), selected=label)
  print(f"saved {args.out}", flush=True)
  return 0


if __name__ == "__main__":
  raise SystemExit(main())
