"""GPU parity of the fused HIP evaluation against the CPU oracle.

Tolerances (stated by the task for this FP64 path): total energy within
1e-8 Ha, density max-abs within 1e-10, gradients within 1e-7 relative
(max |dg| / max |g|).  The tiny cases are held much tighter.
"""
import numpy as np
import pytest

from helpers import oracle_problem, rel_max, seeded_params
from oracle import adjoint as OA

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

E_TOL = 1e-8
RHO_TOL = 1e-10
G_TOL = 1e-7


def _plan(crystal, mask, kpts, occ, chunk=0):
  from jrystal_amd.objective import EnergyPlan
  return EnergyPlan(crystal.positions, crystal.charges, crystal.cell_vectors,
                    mask, kpts, occ, chunk_states=chunk)


def _run(plan, params):
  out = plan.evaluate(params["w_re"], params["w_im"], grad=True, density=True)
  torch.cuda.synchronize()
  return {k: v.cpu().numpy() for k, v in out.items()}


def _check(out, prob, params, e_tol=E_TOL, rho_tol=RHO_TOL, g_tol=G_TOL):
  parts, grad, rho = OA.energy_and_grad(params, prob)
  e = out["energies"]
  for i, k in enumerate(("kinetic", "hartree", "external", "xc", "total")):
    assert abs(e[i] - parts[k]) <= e_tol, (k, e[i], parts[k])
  assert np.max(np.abs(out["density"] - rho)) <= rho_tol
  assert rel_max(out["dw_re"], grad["w_re"]) <= g_tol
  assert rel_max(out["dw_im"], grad["w_im"]) <= g_tol
  return parts


def _case(wl, seed=0, chunk=0, **kw):
  prob = oracle_problem(wl.crystal, wl.grid_sizes, wl.k_grid, wl.num_bands,
                        wl.cutoff_energy)
  params = seeded_params(wl.shape, seed)
  plan = _plan(wl.crystal, wl.mask, wl.kpts, wl.occupation, chunk)
  print(f"case {wl.name} shape={wl.shape} chunk={chunk}", flush=True)
  return _check(_run(plan, params), prob, params, **kw)


def test_tiny_tight():
  from jrystal_amd import workloads as W
  _case(W.tiny(), e_tol=1e-11, rho_tol=1e-13, g_tol=1e-11)


@pytest.mark.parametrize("z", [14.0, 4.0])
def test_si2_fcc_32(z):
  from jrystal_amd import workloads as W
  _case(W.si2_fcc(z))


def test_si8_cube_64_k8():
  from jrystal_amd import workloads as W
  _case(W.si8_cube(14.0, k=2))


@pytest.mark.parametrize("nb", [17, 48, 64])
def test_band_tiles(nb):
  """Band counts that fill 2, 3 and 4 MFMA tiles, with padding."""
  from jrystal_amd import workloads as W
  wl = W.build(f"si2-16^3-{nb}b", W.fcc_si2(14.0), (16, 16, 16), (1, 1, 2),
               nb, 11.0)
  assert wl.num_g > nb
  _case(wl, seed=nb, e_tol=1e-9, rho_tol=1e-11, g_tol=1e-9)


def test_chunked_matches_unchunked():
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(14.0)
  params = seeded_params(wl.shape, 3)
  a = _run(_plan(wl.crystal, wl.mask, wl.kpts, wl.occupation, 0), params)
  b = _run(_plan(wl.crystal, wl.mask, wl.kpts, wl.occupation, 7), params)
  assert np.max(np.abs(a["energies"] - b["energies"])) < 1e-11
  assert rel_max(b["dw_re"], a["dw_re"]) < 1e-12


def test_deterministic_rerun():
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(4.0)
  plan = _plan(wl.crystal, wl.mask, wl.kpts, wl.occupation)
  params = seeded_params(wl.shape, 5)
  a = _run(plan, params)
  b = _run(plan, params)
  for k in a:
    assert np.array_equal(a[k], b[k]), k


def test_noncubic_grid_cubic_mask_unrestricted():
  """Odd radices (3, 5), non-cubic box, cubic mask, two spin channels."""
  from jrystal_amd import grid, occupation
  from jrystal_amd import workloads as W
  cr = W.fcc_si2(4.0)
  gs = (15, 18, 20)
  mask = grid.cubic_mask(gs)
  kpts = grid.k_vectors(cr.cell_vectors, (1, 2, 1))
  occ = occupation.uniform(kpts.shape[0], cr.num_electron, 0, 5,
                           spin_restricted=False)
  prob = oracle_problem(cr, gs, (1, 2, 1), 5, mask_method="cubic",
                        spin_restricted=False)
  shape = (2, kpts.shape[0], int(mask.sum()), 5)
  params = seeded_params(shape, 11)
  out = _run(_plan(cr, mask, kpts, occ), params)
  _check(out, prob, params, e_tol=1e-10, rho_tol=1e-12, g_tol=1e-10)


def test_energy_only_matches():
  from jrystal_amd import workloads as W
  wl = W.si2_fcc(14.0)
  plan = _plan(wl.crystal, wl.mask, wl.kpts, wl.occupation)
  params = seeded_params(wl.shape, 1)
  full = plan.evaluate(params["w_re"], params["w_im"], grad=True)
  eo = plan.evaluate(params["w_re"], params["w_im"], grad=False)
  torch.cuda.synchronize()
  d = (full["energies"] - eo["energies"]).abs().max().item()
  assert d < 1e-11


def test_autograd_function():
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import total_energy
  wl = W.si2_fcc(4.0)
  plan = _plan(wl.crystal, wl.mask, wl.kpts, wl.occupation)
  params = seeded_params(wl.shape, 2)
  wr = torch.tensor(params["w_re"], device="cuda", requires_grad=True)
  wi = torch.tensor(params["w_im"], device="cuda", requires_grad=True)
  e = total_energy(plan, wr, wi)
  (2.5 * e).backward()
  ref = plan.evaluate(wr.detach(), wi.detach(), grad=True)
  assert torch.allclose(wr.grad, 2.5 * ref["dw_re"], rtol=0, atol=1e-13)
  assert torch.allclose(wi.grad, 2.5 * ref["dw_im"], rtol=0, atol=1e-13)


@pytest.mark.parametrize("name", ["tiny_z4", "si2_z14", "si2_z4"])
def test_golden_vectors(name):
  """GPU path against the committed oracle vectors (tests/golden)."""
  import os
  from jrystal_amd import workloads as W
  gold = np.load(os.path.join(os.path.dirname(__file__), "golden",
                              "oracle_golden.npz"))
  wl = {"tiny_z4": lambda: W.tiny(z=4.0), "si2_z14": lambda: W.si2_fcc(14.0),
        "si2_z4": lambda: W.si2_fcc(4.0)}[name]()
  out = _run(_plan(wl.crystal, wl.mask, wl.kpts, wl.occupation),
             wl.params(0))
  assert np.max(np.abs(out["energies"] - gold[f"{name}/energies"])) < E_TOL
  rho = out["density"]
  np.testing.assert_allclose([rho.sum(), (rho**2).sum(), rho.max()],
                             gold[f"{name}/rho_sum"], rtol=1e-12)
  t = tuple(gold[f"{name}/grad_idx"].T)
  norm = gold[f"{name}/grad_norm"]
  for part, n in (("re", norm[0]), ("im", norm[1])):
    got = out[f"dw_{part}"][t]
    assert np.max(np.abs(got - gold[f"{name}/grad_{part}"])) < G_TOL * n


def test_register_fft64_matches_generic(monkeypatch):
  """64^3 register-resident kernels against the runtime-radix kernels."""
  from jrystal_amd import workloads as W
  wl = W.si8_cube(4.0, k=2)
  params = seeded_params(wl.shape, 9)
  a = _run(_plan(wl.crystal, wl.mask, wl.kpts, wl.occupation), params)
  monkeypatch.setenv("JRX_FFT_GENERIC", "1")
  b = _run(_plan(wl.crystal, wl.mask, wl.kpts, wl.occupation), params)
  assert np.max(np.abs(a["energies"] - b["energies"])) < 1e-10
  assert np.max(np.abs(a["density"] - b["density"])) < 1e-13
  assert rel_max(a["dw_re"], b["dw_re"]) < 1e-12
  assert rel_max(a["dw_im"], b["dw_im"]) < 1e-12


def test_strained_gamma_only_si8():
  """Gamma-only (one k-point), sheared cell, 32 bands: the shape of the
  strained 8-atom comparison case at nk = 1."""
  from jrystal_amd import workloads as W
  _case(W.si8_strained(4.0, k=1), seed=2)
  wl = W.si8_strained(4.0, k=1)
  plan = _plan(wl.crystal, wl.mask, wl.kpts, wl.occupation)
  params = seeded_params(wl.shape, 2)
  full = plan.evaluate(params["w_re"], params["w_im"], grad=True)
  eo = plan.evaluate(params["w_re"], params["w_im"], grad=False)
  torch.cuda.synchronize()
  assert (full["energies"] - eo["energies"]).abs().max().item() < 1e-10


@pytest.mark.parametrize("n", [36, 42, 45])
def test_compile_time_radix_sizes(n):
  """Cubic grids whose plane and x-line passes use the compile-time radix
  schedules 4-3-3, 2-3-7 and 3-3-5."""
  from jrystal_amd import workloads as W
  wl = W.build(f"si2-{n}^3", W.fcc_si2(4.0), (n, n, n), (1, 1, 2), 8,
               W.alias_free_cutoff(W.fcc_si2(4.0).cell_vectors, (n, n, n)))
  _case(wl, seed=n)


@pytest.mark.parametrize("cond", ["random", "ill"])
def test_adaptive_cholesky_qr(cond):
  """Adaptive CholeskyQR (one pass for the k-points whose scaled Gram
  matrix is within 0.5 of I in Frobenius norm, two for the rest) against
  two passes everywhere and against the oracle (Householder QR).  A random
  start takes one pass; a start whose bands 1..nb-1 lean towards band 0
  (cond(W) ~ 1e4: one pass would leave ~eps cond^2 of non-orthogonality)
  must take two, i.e. give the two-pass result bit for bit."""
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import EnergyPlan
  wl = W.si2_fcc(4.0)
  prob = oracle_problem(wl.crystal, wl.grid_sizes, wl.k_grid, wl.num_bands,
                        wl.cutoff_energy)
  params = seeded_params(wl.shape, 7)
  if cond == "ill":
    for k in ("w_re", "w_im"):
      w = params[k]
      w[..., 1:] = w[..., :1] + 1e-4 * w[..., 1:]
  outs = {}
  for qp in (0, 2):
    cr = wl.crystal
    plan = EnergyPlan(cr.positions, cr.charges, cr.cell_vectors, wl.mask,
                      wl.kpts, wl.occupation, qr_passes=qp)
    outs[qp] = _run(plan, params)
    plan.close()
  _check(outs[0], prob, params)
  for k in ("energies", "dw_re", "dw_im", "density"):
    if cond == "ill":
      assert np.array_equal(outs[0][k], outs[2][k]), k
    else:
      assert rel_max(outs[0][k], outs[2][k]) < 1e-12, k


def test_qr_passes_option_is_checked():
  from jrystal_amd import workloads as W
  from jrystal_amd.objective import EnergyPlan
  wl = W.tiny()
  cr = wl.crystal
  with pytest.raises(ValueError):
    EnergyPlan(cr.positions, cr.charges, cr.cell_vectors, wl.mask, wl.kpts,
               wl.occupation, qr_passes=1)
