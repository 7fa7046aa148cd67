"""Kohn-Sham Hamiltonian of the band-structure driver, on the GPU.

Mirrors jrystal/_src/hamiltonian.py:105-240.  A plan in band mode
(``EnergyPlan.set_fixed_density``) holds the effective potential
v_eff = V_H + V_ext + v_xc of a fixed (ground-state) density; its
evaluations return

  hamiltonian_matrix_trace = sum_{s,k,b} <psi_b|T + v_eff|psi_b>

(the plan's occupation weighs each band; ``band_plan`` sets all weights to
one, as the reference's trace has none) and its gradient in one fused HIP
sequence, with Q = QR(W) as in pw.coeff.  ``hamiltonian_matrix`` gives
q_i^H H q_j over the bands of each (spin, k) from the same evaluation (the
reference forms it as the complex Hessian of the trace of a band mixture,
hamiltonian.py:209-232), and ``band_eigenvalues`` its ascending eigenvalues
(jnp.linalg.eigvalsh in calc_band_structure_all_electrons.py:156-169), by
cyclic Jacobi on the device.
"""
import numpy as np
import torch

from .objective import EnergyPlan


def band_plan(positions, charges, cell_vectors, mask, density, kpts,
              num_bands: int, xc: str = "lda_x", local_potential=None,
              projectors=None, **options) -> EnergyPlan:
  """A plan in band mode: the ground-state ``density`` (ns,nx,ny,nz) fixes
  the potential, every band has weight one, ``kpts`` (nk,3) 1/bohr.  With a
  norm-conserving pseudopotential, ``local_potential`` (the G box of
  pseudopotential.potential_local_reciprocal) replaces the nuclei in that
  potential and ``projectors`` (nk, nproj, ng) add the non-local term, as
  in calc_band_structure_normcons.py:107-160."""
  rho = torch.as_tensor(density, dtype=torch.float64)
  if rho.ndim == 3:
    rho = rho[None]
  kpts = np.asarray(kpts, np.float64).reshape(-1, 3)
  occ = np.ones((rho.shape[0], kpts.shape[0], int(num_bands)))
  plan = EnergyPlan(positions, charges, cell_vectors, mask, kpts, occ, xc=xc,
                    **options)
  if local_potential is not None:
    plan.set_external_potential(local_potential)
  plan.set_fixed_density(rho)
  if projectors is not None:
    plan.set_nonlocal(projectors)
  return plan


def _require_band_mode(plan):
  if not getattr(plan, "band_mode", False):
    raise ValueError("the plan is not in band mode (set_fixed_density)")


class _Trace(torch.autograd.Function):

  @staticmethod
  def forward(ctx, plan, w_re, w_im):
    r = plan.evaluate(w_re, w_im, grad=True)
    ctx.save_for_backward(r["dw_re"], r["dw_im"])
    return r["energies"][4].clone()

  @staticmethod
  def backward(ctx, g):
    dwr, dwi = ctx.saved_tensors
    return None, g * dwr, g * dwi


def hamiltonian_matrix_trace(plan: EnergyPlan, w_re, w_im):
  """Differentiable sum_{s,k,b} f <psi_b|T + v_eff|psi_b> (hamiltonian.py:
  105-168 with keep_kpts_axis False), a 0-d device tensor."""
  _require_band_mode(plan)
  return _Trace.apply(plan, w_re, w_im)


def hamiltonian_matrix(plan: EnergyPlan, w_re, w_im):
  """q_i^H H q_j, (ns,nk,nb,nb) complex128 on the device (hamiltonian.py:
  171-240)."""
  _require_band_mode(plan)
  plan.evaluate(w_re, w_im, grad=False)
  return plan.subspace_hamiltonian()


def band_eigenvalues(plan: EnergyPlan, w_re, w_im):
  """Ascending eigenvalues of hamiltonian_matrix, (ns,nk,nb) on the
  device."""
  _require_band_mode(plan)
  plan.evaluate(w_re, w_im, grad=False)
  return plan.subspace_eigenvalues()
