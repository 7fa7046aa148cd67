"""The Rayleigh-Ritz rotation of a gauge reset on the GPU (cyclic Jacobi,
jrx_plan_ritz_rotation) against numpy's eigh of the same subspace blocks:
unitary, identity outside the runs of equal positive occupation, the run's
subspace Hamiltonian diagonalised with ascending eigenvalues equal to
eigvalsh (1e-12 relative), eigenvectors equal up to a phase where the
eigenvalues are separated; runs above 64 bands use the global-memory form."""
import numpy as np
import pytest

from helpers import seeded_params

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _runs(f):
  out, b, nb = [], 0, len(f)
  while b < nb:
    e = b
    while e + 1 < nb and f[e + 1] == f[b]:
      e += 1
    if f[b] > 0 and e > b:
      out.append((b, e + 1))
    b = e + 1
  return out


def _check(plan):
  S = plan.subspace_matrix().cpu().numpy()
  V = plan.ritz_rotation().cpu().numpy()
  occ = plan.occupation
  ns, nk, nb = occ.shape
  for s in range(ns):
    for k in range(nk):
      v = V[s, k]
      np.testing.assert_allclose(v.conj().T @ v, np.eye(nb), atol=1e-13)
      inside = np.zeros((nb, nb), bool)
      for b, e in _runs(occ[s, k]):
        inside[b:e, b:e] = True
        h = S[s, k, b:e, b:e] / (2.0 * occ[s, k, b])
        h = 0.5 * (h + h.conj().T)
        vr = v[b:e, b:e]
        d = vr.conj().T @ h @ vr
        lam = np.linalg.eigvalsh(h)
        scale = np.max(np.abs(lam))
        assert np.max(np.abs(d - np.diag(np.diag(d)))) <= 1e-12 * scale
        np.testing.assert_allclose(np.real(np.diag(d)), lam,
                                   atol=1e-12 * scale)
        w = np.linalg.eigh(h)[1]
        gap = np.minimum(np.r_[np.inf, np.diff(lam)], np.r_[np.diff(lam), np.inf])
        sep = gap > 1e-6 * scale
        ov = np.abs(np.sum(np.conj(w) * vr, axis=0))
        np.testing.assert_allclose(ov[sep], 1.0, atol=1e-9)
      np.testing.assert_array_equal(v[~inside], np.eye(nb)[~inside])


def _plan_eval(wl, occupation=None, seed=3):
  from jrystal_amd.objective import EnergyPlan
  plan = EnergyPlan.from_workload(wl)
  if occupation is not None:
    plan.set_occupation(occupation)
  p = seeded_params(wl.shape, seed)
  plan.evaluate(torch.tensor(p["w_re"], device="cuda"),
                torch.tensor(p["w_im"], device="cuda"), grad=True)
  return plan


def test_ritz_uniform_and_mixed_runs():
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(4.0)  # 4 occupied bands of 8 at f = 2
  _check(_plan_eval(wl))
  occ = np.zeros(wl.occupation.shape)
  occ[..., :3] = 2.0
  occ[..., 3:5] = 1.0
  occ[..., 5] = 0.5
  _check(_plan_eval(wl, occ))


def test_ritz_run_above_64_bands():
  from jrystal_amd import workloads as W
  wl = W.build("si2-16^3-80b", W.fcc_si2(14.0), (16, 16, 16), (1, 1, 2), 80,
               11.0)
  occ = np.ones(wl.occupation.shape)
  occ[..., 70:] = 0.0
  _check(_plan_eval(wl, occ))
