"""All-electron ground-state energy driver on the GPU.

Mirrors jrystal/calc/calc_ground_state_energy_all_electrons.py:63-247:
build crystal / grids / mask / k-points / Ewald constant / occupation from
the config, initialise the plane-wave parameters, then run the optimiser on
the free energy F = E - T S with the temperature annealing schedule and the
convergence window.  Each step's energy and gradient come from one fused
evaluation (jrystal_amd.objective.EnergyPlan); the update is the fused Adam
kernel.  Returns the density, as the reference does.

With use_pseudopotential (norm-conserving files) the same loop is the
driver of calc_ground_state_energy_normcons.py:63-250: the valence electrons
set the bands and occupations, the local pseudopotential replaces the
nuclei in the fused evaluation and the non-local projectors join it
(jrystal_amd.pseudopotential); the Ewald term keeps the crystal's nuclear
charges, as the reference's get_ewald_coulomb_repulsion does
(opt_utils.py:187-201).
"""
import logging
import math
import os
import time
from dataclasses import dataclass, field
from typing import List

import numpy as np
import torch

from .. import checkpoint, energy, entropy, grid, occupation, pw
from ..config import JrystalConfigDict
from ..crystal import Crystal
from ..objective import ENERGY_KEYS, EnergyPlan
from ..optim import Adam, ConvergenceChecker, exponential_decay

log = logging.getLogger("jrystal_amd")

GEOMETRY_DIR = os.path.join(os.path.dirname(os.path.dirname(
  os.path.abspath(__file__))), "geometry")


@dataclass
class GroundStateEnergyOutput:
  """calc_ground_state_energy_all_electrons.py:43-60, plus the breakdown."""
  config: JrystalConfigDict
  crystal: Crystal
  params_pw: dict
  params_occ: np.ndarray
  total_energy: float
  total_energy_history: List[float] = field(default_factory=list)
  energies: dict = field(default_factory=dict)
  density: object = None
  converged: bool = False
  steps: int = 0
  seconds: float = 0.0


def create_crystal(config) -> Crystal:
  """opt_utils.py:106-114."""
  if config.crystal is not None:
    path = os.path.join(GEOMETRY_DIR, config.crystal + ".xyz")
  else:
    path = config.crystal_file_path_path
  cr = Crystal.create_from_file(path, spin=config.spin)
  if cr.num_electron % 2 != cr.spin % 2:  # utils.check_spin_number
    raise ValueError("spin number is not valid for the system. ")
  return cr


def create_pseudopotential(config, crystal):
  """opt_utils.py:117-140 for norm-conserving files: the config's
  pseudopotential_file_dir, else the shipped directory."""
  from ..pseudopotential import NormConservingPseudopotential
  d = config.pseudopotential_file_dir
  return NormConservingPseudopotential.create(crystal, d if d else None)


def apply_pseudopotential(plan, pseudo, g_vec, mask, kpts, vol):
  """The local part in place of the nuclei and the non-local projectors
  (calc_ground_state_energy_normcons.py:111-145)."""
  from .. import pseudopotential as PS
  plan.set_external_potential(PS.potential_local_reciprocal(pseudo, g_vec,
                                                            vol))
  plan.set_nonlocal(PS.potential_nonlocal_psi_reciprocal(pseudo, g_vec, mask,
                                                         kpts))


def create_freq_mask(config, crystal):
  """opt_utils.py:76-103."""
  gs = grid.proper_grid_size(config.grid_sizes)
  if config.freq_mask_method == "cubic":
    return grid.cubic_mask(gs)
  if config.freq_mask_method == "spherical":
    return grid.spherical_mask(crystal.cell_vectors, gs, config.cutoff_energy)
  raise ValueError("freq_mask_method must be either cubic or spherical.")


def _checkpoint_path(config):
  if not config.save_dir:
    return None
  return os.path.join(config.save_dir, "checkpoint.npz")


def _load_resume(path, plan, occ_keys=()):
  """A checkpoint of this driver, checked against the plan's shapes."""
  ck = checkpoint.load(path)
  for k in ("w_re", "w_im"):
    if tuple(ck["params"][k].shape) != tuple(plan.param_shape):
      raise ValueError(f"checkpoint {path}: {k} has shape "
                       f"{ck['params'][k].shape}, the run needs "
                       f"{plan.param_shape}")
  missing = [k for k in occ_keys if k not in ck["occ_params"]]
  if missing:
    raise ValueError(f"checkpoint {path} lacks occupation parameters "
                     f"{missing}")
  return ck


def calc_with_output(config: JrystalConfigDict,
                     resume: str = None) -> GroundStateEnergyOutput:
  """resume: a checkpoint written by an earlier run of the same config
  (its save_dir); the run continues from its step with the same parameters,
  occupation parameters, Adam moments and energy history, so an
  interrupted and resumed Adam run ends bitwise equal to an uninterrupted
  one.  The CG optimisers restart their search direction from the saved
  parameters."""
  if config.use_pseudopotential and config.pseudopotential_type not in (
      "normcons", "normconserving", "nc"):
    raise NotImplementedError(
      f"pseudopotential type {config.pseudopotential_type!r}: norm-conserving "
      "(nc) only")
  if config.occupation not in ("uniform", "gamma") + occupation.PARAMETERISED:
    raise ValueError(f"Invalid method: {config.occupation}")
  parameterised = config.occupation in occupation.PARAMETERISED
  if config.optimizer not in ("adam", "cg", "cg-learned"):
    raise NotImplementedError(
      f"optimizer {config.optimizer!r}: adam (the reference's), cg "
      "(preconditioned conjugate gradients) or cg-learned (with the learned "
      "preconditioner) are on this path")
  if config.optimizer != "adam" and (config.smearing > 0.0 or parameterised):
    raise NotImplementedError("the cg optimisers minimise the energy at "
                              "fixed occupations: set smearing to 0 and "
                              "occupation to uniform or gamma")
  if config.scheduler:
    raise NotImplementedError("Scheduler is not implemented yet.")
  cr = create_crystal(config)
  ne = cr.num_electron
  pseudo = None
  if config.use_pseudopotential:
    # calc_ground_state_energy_normcons.py:82-105: the valence electrons
    pseudo = create_pseudopotential(config, cr)
    ne = pseudo.total_valence_charge
  gs = grid.proper_grid_size(config.grid_sizes)
  ks = grid.proper_grid_size(config.k_grid_sizes)
  g_vec = grid.g_vectors(cr.cell_vectors, gs)
  kpts = grid.k_vectors(cr.cell_vectors, ks)
  nk = kpts.shape[0]
  num_bands = math.ceil(ne / 2) + config.empty_bands
  mask = create_freq_mask(config, cr)
  ew = float(energy.nuclear_repulsion(
    cr.positions, cr.charges, cr.cell_vectors, g_vec, cr.vol,
    config.ewald_args["ewald_eta"], config.ewald_args["ewald_cutoff"]))
  occ_params = occupation.param_init(config.seed, num_bands, ne, nk, cr.spin,
                                     config.occupation, config.spin_restricted)

  def get_occupation(p):
    return occupation.occupation(p, nk, ne, cr.spin, config.occupation,
                                 config.spin_restricted)
  if parameterised:
    occ = get_occupation(occ_params).detach().cpu().numpy()
  else:
    occ = occ_params
  etro = entropy.fermi_dirac(occ, eps=config.eps)
  plan = EnergyPlan(cr.positions, cr.charges, cr.cell_vectors, mask, kpts,
                    occ, xc=config.xc, chunk_states=config.chunk_states)
  if pseudo is not None:
    apply_pseudopotential(plan, pseudo, g_vec, mask, kpts, cr.vol)
  params = pw.param_init(config.seed, num_bands, nk, mask,
                         spin_restricted=config.spin_restricted,
                         dist=config.param_dist)
  if config.optimizer != "adam":
    if resume is not None:
      ck = _load_resume(resume, plan)
      params = {k: torch.as_tensor(v, device=plan.device)
                for k, v in ck["params"].items()}
    return _calc_cg(config, cr, plan, params, occ, ew)
  opt_args = dict(config.optimizer_args)
  if parameterised:
    # one optimiser over both parameter sets, as the reference's update
    # (calc_ground_state_energy_all_electrons.py:167-181)
    occ_flat, occ_rebuild = occupation.leaves(occ_params)
    plan.set_band_energies(True)
    leaves = dict(params, **occ_flat)
  else:
    leaves = params
  opt = Adam(leaves, **opt_args)
  history = []
  step0 = 0
  if resume is not None:
    ck = _load_resume(resume, plan, occ_flat if parameterised else ())
    if ck["adam"] is None:
      raise ValueError(f"checkpoint {resume} holds no Adam state")
    for k, v in leaves.items():
      src = ck["params"][k] if k in ck["params"] else ck["occ_params"][k]
      v.copy_(torch.as_tensor(src, device=v.device))
      opt.m[k].copy_(torch.as_tensor(ck["adam"]["m"][k], device=v.device))
      opt.v[k].copy_(torch.as_tensor(ck["adam"]["v"][k], device=v.device))
    opt.count = ck["adam"]["count"]
    history = [float(e) for e in ck["energy_history"]]
    step0 = ck["step"]
    log.info("resumed from %s at step %d", resume, step0)
  ck_path = _checkpoint_path(config)
  ck_every = int(config.checkpoint_every or 0)

  def save_checkpoint(step):
    occ_now = (get_occupation(occ_rebuild(occ_flat)).detach()
               if parameterised else occ)
    checkpoint.save(
      ck_path, step=step, params=params, occupation=occ_now,
      energy_history=history,
      adam={"count": opt.count, "m": opt.m, "v": opt.v},
      occ_params=occ_flat if parameterised else None,
      meta={"crystal": str(config.crystal), "seed": int(config.seed),
            "optimizer": "adam"})
  if config.smearing > 0.0:
    temperature = exponential_decay(100.0, config.epoch // 2,
                                    config.smearing / 100, config.smearing)
  else:
    def temperature(i):
      return 0.0
  checker = ConvergenceChecker(config.convergence_window_size,
                               config.convergence_condition)
  checker.history = history[-checker.window_size:]
  log.info("crystal %s: %d electrons, %d bands, %d k-points, G_n=%d",
           cr.symbols, ne, num_bands, nk, plan.num_g)
  converged = False
  out = {}
  t0 = time.perf_counter()
  i = step0 - 1
  for i in range(step0, config.epoch):
    temp = temperature(i)
    if parameterised:
      # F = E(W, f) - T S(f): dF/dW from the plan; dF/df = band energies
      # - T dS/df, pulled back through the occupation map by autograd
      view = {k: v.detach().requires_grad_() for k, v in occ_flat.items()}
      occ_t = get_occupation(occ_rebuild(view))
      plan.set_occupation(occ_t.detach())
      plan.evaluate(params["w_re"], params["w_im"], grad=True, out=out)
      s_t = entropy.fermi_dirac(occ_t, eps=config.eps)
      surrogate = (plan.band_energies() * occ_t).sum() - temp * s_t
      g_occ = torch.autograd.grad(surrogate, list(view.values()))
      grads = {"w_re": out["dw_re"], "w_im": out["dw_im"]}
      grads.update({k: g.contiguous() for k, g in zip(view, g_occ)})
      opt.step(leaves, grads)
      etro = float(s_t.detach())
    else:
      plan.evaluate(params["w_re"], params["w_im"], grad=True, out=out)
      opt.step(params, {"w_re": out["dw_re"], "w_im": out["dw_im"]})
    etot = float(out["energies"][4])  # one host sync per step
    history.append(etot)
    if not np.isfinite(etot):
      raise FloatingPointError(f"energy became {etot} at step {i}")
    if ck_path and ck_every > 0 and (i + 1) % ck_every == 0:
      save_checkpoint(i + 1)
    if checker.check(etot):
      converged = True
      log.info("Converged.")
      break
    if config.verbose and i % 100 == 0:
      log.info("step %d | Loss %.6f | Energy %.6f | Entropy %.4f | T %.2E",
               i, etot - temp * etro, etot + ew, etro, temp)
  seconds = time.perf_counter() - t0
  if not converged:
    log.warning("Did not converge.")
  if ck_path:
    save_checkpoint(i + 1)
  if parameterised:
    with torch.no_grad():
      plan.set_occupation(get_occupation(occ_rebuild(occ_flat)))
    return _final(config, cr, plan, params, occ_rebuild(occ_flat), ew,
                  history, converged, i + 1, seconds)
  return _final(config, cr, plan, params, occ, ew, history, converged, i + 1,
                seconds)


def _final(config, cr, plan, params, occ, ew, history, converged, steps,
           seconds):
  final = plan.evaluate(params["w_re"], params["w_im"], grad=False,
                        density=True)
  e = dict(zip(ENERGY_KEYS, final["energies"].tolist()))
  if getattr(plan, "num_projectors", 0):
    occ_t = torch.as_tensor(np.asarray(plan.occupation, np.float64),
                            device=plan.device)
    e["external_nonlocal"] = float(torch.sum(plan.nonlocal_energies() *
                                             occ_t))
    e["external_local"] = e["external"] - e["external_nonlocal"]
  e["nuclear_repulsion"] = ew
  e["total_with_nuclear"] = e["total"] + ew
  terms = (("Hartree", "hartree"), ("External", "external"))
  if "external_nonlocal" in e:
    terms = (("Hartree", "hartree"), ("External (local)", "external_local"),
             ("External (nonlocal)", "external_nonlocal"))
  for name, key in terms + (("XC", "xc"), ("Kinetic", "kinetic"),
                            ("Nuclear repulsion", "nuclear_repulsion")):
    log.info("%s Energy: %.4f Ha", name, e[key])
  log.info("Total Energy: %.4f Ha", e["total_with_nuclear"])
  return GroundStateEnergyOutput(
    config=config, crystal=cr, params_pw=params, params_occ=occ,
    total_energy=e["total_with_nuclear"], total_energy_history=history,
    energies=e, density=final["density"], converged=converged, steps=steps,
    seconds=seconds)


def _calc_cg(config, cr, plan, params, occ, ew):
  """Preconditioned CG (optimizer: cg) or CG with the learned
  preconditioner (optimizer: cg-learned) for at most `epoch` iterations,
  stopping when the energy stops decreasing above rounding."""
  from .. import learned
  from ..minimize import PLAIN_RESTART, Minimizer
  factor_fn, restart = None, PLAIN_RESTART
  if config.optimizer == "cg-learned":
    factor_fn = learned.LearnedFactor(plan, learned.PrecondNet.load(
      device=plan.device))
    restart = learned.LEARNED_RESTART
  m = Minimizer(plan, params, factor_fn=factor_fn, restart=restart,
                first_restart=(learned.LEARNED_FIRST_RESTART
                               if factor_fn is not None else None))

  def cb(mm, it, e):
    if config.verbose and it % 20 == 0:
      log.info("iteration %d | Energy %.8f", it, e + ew)
  res = m.run(max_iter=config.epoch, callback=cb)
  converged = res.status == "converged"
  log.info("cg finished: %s after %d iterations (%d evaluations, %.2f s)",
           res.status, res.iterations, res.evaluations, res.seconds)
  history = [e for _, e in res.history]
  p = {"w_re": m.params()["w_re"].clone(), "w_im": m.params()["w_im"].clone()}
  ck_path = _checkpoint_path(config)
  if ck_path:
    checkpoint.save(ck_path, step=res.iterations, params=p, occupation=occ,
                    energy_history=history,
                    meta={"crystal": str(config.crystal),
                          "seed": int(config.seed),
                          "optimizer": config.optimizer})
  return _final(config, cr, plan, p, occ, ew, history, converged,
                res.iterations, res.seconds)


def calc(config: JrystalConfigDict, resume: str = None):
  """Returns the density (spin, x, y, z), as the reference driver does."""
  return calc_with_output(config, resume=resume).density
