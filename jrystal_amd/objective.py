"""Fused evaluation of the direct-minimisation objective on the GPU.

``EnergyPlan`` owns a ``jrx_plan`` (libjrx_hip.so) for one problem and
evaluates, in one stream-ordered sequence of HIP kernels,

  E(w) = E_kin + E_H + E_ext + E_xc     and     dE/dw_re, dE/dw_im

for the parameters w = {'w_re', 'w_im'} of layout (spin, kpt, g, band).
This is the hot path of jrystal's all-electron driver:
``total_energy`` + ``jax.value_and_grad`` in
jrystal/calc/calc_ground_state_energy_all_electrons.py:119-137,175-181, for
fixed (uniform / gamma) occupations, whose gradient is zero (occupation.py:
271-274 applies stop_gradient).

``total_energy(plan, w_re, w_im)`` is the autograd-aware form: its backward
returns the gradient that the same fused call computed.
"""
import ctypes
from typing import Optional

import numpy as np
import torch

from . import _lib

ENERGY_KEYS = ("kinetic", "hartree", "external", "xc", "total")


def _require_cuda():
  if not torch.cuda.is_available():
    raise RuntimeError("the jrystal_amd GPU path needs a ROCm GPU "
                       "(no CPU fallback)")


# JRX_XC_* flags of include/jrx.h for the reference's xc term names
XC_TERMS = {"lda_x": 1, "lda_c_pw": 2, "gga_x_pbe": 4, "gga_c_pbe": 8}


def xc_terms(xc: str) -> int:
  """Flags of a '+'-joined xc string (xc.py:222-253 splits it the same
  way); terms outside XC_TERMS raise NotImplementedError."""
  flags = 0
  for t in xc.split("+"):
    t = t.strip()
    if t not in XC_TERMS:
      raise NotImplementedError(
        f"xc term {t!r}: this path has {sorted(XC_TERMS)}")
    flags |= XC_TERMS[t]
  return flags


class EnergyPlan:
  """Plan for one (crystal, mask, k-points, occupation) problem.

  Args:
    positions: (natom, 3) bohr.  charges: (natom,).  cell_vectors: (3, 3)
    bohr, rows.  mask: (nx, ny, nz) bool.  kpts: (nk, 3).  occupation:
    (ns, nk, nb).  chunk_states: states per FFT chunk (0 = automatic).
    fft_variant: 0 automatic, 1 runtime-radix FFT kernels everywhere.
    xline_blocks: target workgroups of an x-line pass (0 = 4096).
    plane_pad: extra elements per x-plane of the plane buffer (-1 =
    automatic).  qr_passes: 0 = adaptive CholeskyQR (one pass for the
    k-points whose Gram matrix is close to diagonal, two for the rest), 2 =
    always two passes.
  """

  def __init__(self, positions, charges, cell_vectors, mask, kpts,
               occupation, xc: str = "lda_x", chunk_states: int = 0,
               device=None, fft_variant: int = 0, xline_blocks: int = 0,
               plane_pad: int = -1, qr_passes: int = 0):
    _require_cuda()
    xc_flags = xc_terms(xc)
    self.device = torch.device(device if device is not None else "cuda")
    self._lib = _lib.lib()
    mask = np.ascontiguousarray(np.asarray(mask, dtype=bool))
    if mask.ndim != 3:
      raise ValueError("mask must be a 3D boolean array")
    occ = np.ascontiguousarray(np.asarray(occupation, np.float64))
    if occ.ndim != 3:
      raise ValueError("occupation must have shape (spin, kpt, band)")
    kpts = np.ascontiguousarray(np.asarray(kpts, np.float64).reshape(-1, 3))
    if kpts.shape[0] != occ.shape[1]:
      raise ValueError(
        "Occupation should have a leading dimension that is the same as "
        f"coeff. Got occupation shape: {occ.shape}, kpts: {kpts.shape}")
    pos = np.ascontiguousarray(np.asarray(positions, np.float64).reshape(-1, 3))
    chg = np.ascontiguousarray(np.asarray(charges, np.float64).reshape(-1))
    cell = np.ascontiguousarray(np.asarray(cell_vectors, np.float64))
    self._keep = (mask.astype(np.uint8), occ, kpts, pos, chg, cell)
    m8 = self._keep[0]
    d = _lib.JrxProblem()
    d.ns, d.nk, d.nb = occ.shape
    d.nx, d.ny, d.nz = mask.shape
    d.natom = pos.shape[0]
    d.mask = m8.ctypes.data
    d.kpts = kpts.ctypes.data
    d.cell = cell.ctypes.data
    d.positions = pos.ctypes.data
    d.charges = chg.ctypes.data
    d.occupation = occ.ctypes.data
    d.xc = xc_flags
    self.xc = xc
    d.chunk_states = int(chunk_states)
    d.fft_variant = int(fft_variant)
    d.xline_blocks = int(xline_blocks)
    d.plane_pad = int(plane_pad)
    d.qr_passes = int(qr_passes)
    self.options = {"chunk_states": int(chunk_states),
                    "fft_variant": int(fft_variant),
                    "xline_blocks": int(xline_blocks),
                    "plane_pad": int(plane_pad),
                    "qr_passes": int(qr_passes)}
    h = ctypes.c_void_p()
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_create(ctypes.byref(h), ctypes.byref(d)))
    self._plan = h
    info = _lib.JrxPlanInfo()
    _lib.check(self._lib.jrx_plan_get_info(h, ctypes.byref(info)))
    self.num_g = info.ng
    self.active_planes = info.active_planes
    self.chunk_states = info.chunk_states
    self.keep_planes = bool(info.keep_planes)
    self.row_path = bool(info.row_path)
    self.num_rows = info.rows
    self.ns, self.nk, self.nb = occ.shape
    self.occupation = occ.copy()
    self.grid_sizes = mask.shape
    self.param_shape = (self.ns, self.nk, self.num_g, self.nb)
    self.workspace_bytes = info.workspace_bytes
    # torch's caching allocator returns >= 512-byte aligned blocks.
    self._ws = torch.empty(self.workspace_bytes + 256, dtype=torch.uint8,
                           device=self.device)
    self._ws_ptr = (self._ws.data_ptr() + 255) // 256 * 256
    self.band_mode = False

  @classmethod
  def from_workload(cls, wl, chunk_states=0, device=None, **options):
    return cls(wl.crystal.positions, wl.crystal.charges,
               wl.crystal.cell_vectors, wl.mask, wl.kpts, wl.occupation,
               chunk_states=chunk_states, device=device, **options)

  def close(self):
    if getattr(self, "_plan", None) is not None and self._plan.value:
      self._lib.jrx_plan_destroy(self._plan)
      self._plan = None

  def __del__(self):
    try:
      self.close()
    except Exception:
      pass

  def _as_param(self, x, name):
    if isinstance(x, np.ndarray):
      x = torch.from_numpy(x)
    if tuple(x.shape) != self.param_shape:
      raise ValueError(f"{name} has shape {tuple(x.shape)}, expected "
                       f"{self.param_shape}")
    return x.to(device=self.device, dtype=torch.float64).contiguous()

  def evaluate(self, w_re, w_im, grad: bool = True, density: bool = False,
               coefficients: bool = False, out: Optional[dict] = None):
    """Run one evaluation on the current stream.  Returns a dict with
    'energies' (device tensor [5]: kinetic, hartree, external, xc, total),
    and when requested 'dw_re', 'dw_im', 'density' (ns,nx,ny,nz) and
    'coefficients' (ns,nk,nb,ng complex128)."""
    w_re = self._as_param(w_re, "w_re")
    w_im = self._as_param(w_im, "w_im")
    out = {} if out is None else out
    dev = self.device
    e = out.get("energies")
    if e is None:
      e = out["energies"] = torch.empty(5, dtype=torch.float64, device=dev)
    dwr = dwi = rho = q = None
    if grad:
      dwr = out.get("dw_re")
      if dwr is None:
        dwr = out["dw_re"] = torch.empty_like(w_re)
      dwi = out.get("dw_im")
      if dwi is None:
        dwi = out["dw_im"] = torch.empty_like(w_im)
    if density:
      rho = out["density"] = torch.empty((self.ns,) + tuple(self.grid_sizes),
                                         dtype=torch.float64, device=dev)
    if coefficients:
      q = out["coefficients"] = torch.empty(
        (self.ns, self.nk, self.nb, self.num_g), dtype=torch.complex128,
        device=dev)
    with torch.cuda.device(dev):
      rc = self._lib.jrx_energy_grad(
        self._plan, _lib.ptr(w_re), _lib.ptr(w_im), _lib.ptr(e),
        _lib.ptr(dwr), _lib.ptr(dwi), _lib.ptr(rho), _lib.ptr(q),
        self._ws_ptr, self.workspace_bytes, _lib.stream_ptr(dev))
    _lib.check(rc)
    return out

  def capture(self, w_re, w_im, grad: bool = True, out: Optional[dict] = None):
    """One evaluation recorded as a HIP graph (torch.cuda.CUDAGraph): the
    evaluation of the tensors w_re, w_im (read again at every replay) into
    the tensors of ``out``.  Returns (graph, out); graph.replay() re-runs
    every kernel with one launch from the host.  The evaluation launches no
    memset and makes no host synchronisation, so it records as is; plan
    timing must be off."""
    w_re = self._as_param(w_re, "w_re")
    w_im = self._as_param(w_im, "w_im")
    out = {} if out is None else out
    side = torch.cuda.Stream(device=self.device)
    side.wait_stream(torch.cuda.current_stream(self.device))
    with torch.cuda.stream(side):  # allocate out and load every kernel
      self.evaluate(w_re, w_im, grad=grad, out=out)
    torch.cuda.current_stream(self.device).wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
      self.evaluate(w_re, w_im, grad=grad, out=out)
    return graph, out

  def precondition(self, g_re, g_im, factor=None, out_re=None, out_im=None):
    """Kinetic (Teter-Payne-Allan) preconditioning of a gradient:
    out = K(|G+k|^2 / tb) * factor * g, with the band norms tb of the last
    evaluation on this plan (K(x) = p / (p + 16x^4),
    p = 27 + 18x + 12x^2 + 8x^3).  ``factor`` is an optional (s,k,g,b)
    float64 tensor of per-coefficient multipliers."""
    g_re = self._as_param(g_re, "g_re")
    g_im = self._as_param(g_im, "g_im")
    if factor is not None:
      factor = self._as_param(factor, "factor")
    out_re = torch.empty_like(g_re) if out_re is None else out_re
    out_im = torch.empty_like(g_im) if out_im is None else out_im
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_precondition(
        self._plan, _lib.ptr(g_re), _lib.ptr(g_im), _lib.ptr(factor),
        _lib.ptr(out_re), _lib.ptr(out_im), self._ws_ptr,
        self.workspace_bytes, _lib.stream_ptr(self.device)))
    return out_re, out_im

  def band_norms(self, out=None):
    """tb[s,k,b] = sum_G |G+k|^2 |Q_Gb|^2 of the last evaluation (twice the
    kinetic energy of each band), device tensor (ns,nk,nb)."""
    if out is None:
      out = torch.empty((self.ns, self.nk, self.nb), dtype=torch.float64,
                        device=self.device)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_band_norms(
        self._plan, _lib.ptr(out), self._ws_ptr, self.workspace_bytes,
        _lib.stream_ptr(self.device)))
    return out

  def set_band_energies(self, on: bool = True):
    """Switch on (or off) the band energies of energy+gradient
    evaluations (band_energies())."""
    _lib.check(self._lib.jrx_plan_set_band_energies(self._plan, int(bool(on))))

  def band_energies(self, out=None):
    """dE/df[s,k,b] = <psi_b|H|psi_b> (kinetic + v_H + v_ext + v_xc) of the
    last energy+gradient evaluation, for every band, device tensor
    (ns,nk,nb); needs set_band_energies(True) before that evaluation."""
    if out is None:
      out = torch.empty((self.ns, self.nk, self.nb), dtype=torch.float64,
                        device=self.device)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_band_energies(
        self._plan, _lib.ptr(out), self._ws_ptr, self.workspace_bytes,
        _lib.stream_ptr(self.device)))
    return out

  def set_occupation(self, occupation):
    """Replace the occupation (ns,nk,nb) used by later evaluations."""
    occ = torch.as_tensor(occupation, dtype=torch.float64)
    if tuple(occ.shape) != (self.ns, self.nk, self.nb):
      raise ValueError(f"occupation must have shape {(self.ns, self.nk, self.nb)}"
                       f", got {tuple(occ.shape)}")
    occ = occ.to(self.device).contiguous()
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_set_occupation(
        self._plan, _lib.ptr(occ), _lib.stream_ptr(self.device)))
    self._occ_dev = occ  # keep the source alive until the copy has run
    self._occ_host = None  # host copy made on first access (no sync here)

  @property
  def occupation(self):
    """The occupation in force (ns,nk,nb), a host numpy array; after
    set_occupation it is copied from the device on first access."""
    if self._occ_host is None:
      self._occ_host = self._occ_dev.detach().cpu().numpy()
    return self._occ_host

  @occupation.setter
  def occupation(self, value):
    self._occ_host = np.asarray(value, np.float64)

  def orthonormal_params(self, out_re=None, out_im=None):
    """Q of the last evaluation in the (s,k,g,b) parameter layout (real and
    imaginary parts).  E(Q) = E(W), and Q has R = I."""
    shape = self.param_shape
    if out_re is None:
      out_re = torch.empty(shape, dtype=torch.float64, device=self.device)
    if out_im is None:
      out_im = torch.empty(shape, dtype=torch.float64, device=self.device)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_orthonormal_params(
        self._plan, _lib.ptr(out_re), _lib.ptr(out_im), self._ws_ptr,
        self.workspace_bytes, _lib.stream_ptr(self.device)))
    return out_re, out_im

  def ritz_rotation(self, out=None):
    """Rayleigh-Ritz rotation of the last evaluation with a gradient,
    (ns,nk,nb,nb) complex128 on the device: on each run of consecutive
    bands with equal occupation f > 0 the eigenvectors (ascending) of the
    subspace Hamiltonian (Q^H Qbar) / (2 f), the identity elsewhere
    (jrx_plan_ritz_rotation, cyclic Jacobi on the GPU)."""
    if out is None:
      out = torch.empty((self.ns, self.nk, self.nb, self.nb),
                        dtype=torch.complex128, device=self.device)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_ritz_rotation(
        self._plan, _lib.ptr(out), self._ws_ptr, self.workspace_bytes,
        _lib.stream_ptr(self.device)))
    return out

  def subspace_matrix(self, out=None):
    """Q^H Qbar of the last evaluation with a gradient, (ns,nk,nb,nb)
    complex128 on the device; for bands of equal occupation f > 0 it is
    2 f q_i^H H q_j."""
    if out is None:
      out = torch.empty((self.ns, self.nk, self.nb, self.nb),
                        dtype=torch.complex128, device=self.device)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_subspace_matrix(
        self._plan, _lib.ptr(out), self._ws_ptr, self.workspace_bytes,
        _lib.stream_ptr(self.device)))
    return out

  def set_fixed_density(self, density):
    """Kohn-Sham band mode: later evaluations use the effective potential
    v_eff = V_H + V_ext + v_xc of ``density`` ((ns,nx,ny,nz), the
    ground-state density; Kohn-Sham form of potential.py:203-279) and
    return E = sum_{s,k,b} f <psi|T + v_eff|psi> (energies = kinetic
    part, potential part, 0, 0, E) and its gradient; density=None returns
    to the ground-state objective.  Graphs captured before a switch replay
    the old objective."""
    ptr = None
    if density is not None:
      rho = torch.as_tensor(density, dtype=torch.float64)
      shape = (self.ns,) + tuple(self.grid_sizes)
      if tuple(rho.shape) != shape:
        raise ValueError(f"density must have shape {shape}, got "
                         f"{tuple(rho.shape)}")
      rho = rho.to(self.device).contiguous()
      self._fixed_rho = rho  # alive until the potential has been formed
      ptr = _lib.ptr(rho)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_set_fixed_density(
        self._plan, ptr, self._ws_ptr, self.workspace_bytes,
        _lib.stream_ptr(self.device)))
    self.band_mode = density is not None

  def set_external_potential(self, v_rec):
    """Replace the Coulomb nuclei by a local potential given on the G box
    (nx,ny,nz) complex, in the normalisation of energy.external_reciprocal
    (the reference's potential_local_reciprocal, local.py:34-112);
    None restores the nuclei."""
    ptr = None
    if v_rec is not None:
      v = torch.as_tensor(v_rec).to(device=self.device,
                                    dtype=torch.complex128).contiguous()
      if tuple(v.shape) != tuple(self.grid_sizes):
        raise ValueError(f"external potential has shape {tuple(v.shape)}, "
                         f"expected {tuple(self.grid_sizes)}")
      self._vloc = v
      ptr = _lib.ptr(v)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_set_external_potential(
        self._plan, ptr, _lib.stream_ptr(self.device)))

  def set_nonlocal(self, projectors):
    """Non-local pseudopotential projectors (nk, nproj, ng) complex at the
    mask points (potential_nonlocal_psi_reciprocal restricted to the mask);
    later evaluations add sum f sum_p |F_pb|^2 / vol, F = P c_b, to the
    external energy, and its gradient.  None or nproj = 0 switches it off."""
    if projectors is None:
      n, ptr = 0, None
    else:
      pr = torch.as_tensor(projectors).to(device=self.device,
                                          dtype=torch.complex128).contiguous()
      if pr.ndim != 3 or pr.shape[0] != self.nk or pr.shape[2] != self.num_g:
        raise ValueError(f"projectors must have shape (nk={self.nk}, nproj, "
                         f"ng={self.num_g}), got {tuple(pr.shape)}")
      self._proj = pr
      n, ptr = int(pr.shape[1]), _lib.ptr(pr)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_set_nonlocal(
        self._plan, n, ptr, _lib.stream_ptr(self.device)))
    self.num_projectors = n

  def nonlocal_energies(self, out=None):
    """Per-band non-local energies sum_p |F_pb|^2 / vol (ns,nk,nb) of the
    last evaluation (without the occupation)."""
    if out is None:
      out = torch.empty((self.ns, self.nk, self.nb), dtype=torch.float64,
                        device=self.device)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_nonlocal_energies(
        self._plan, _lib.ptr(out), _lib.stream_ptr(self.device)))
    return out

  def set_kpoints(self, kpts):
    """Replace the k-points (nk,3) [1/bohr] of later evaluations."""
    k = torch.as_tensor(np.asarray(kpts, np.float64).reshape(-1, 3))
    if k.shape[0] != self.nk:
      raise ValueError(f"expected {self.nk} k-points, got {k.shape[0]}")
    k = k.to(self.device).contiguous()
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_set_kpoints(
        self._plan, _lib.ptr(k), _lib.stream_ptr(self.device)))
    self._kpts_dev = k  # alive until the copy has run

  def subspace_hamiltonian(self, out=None):
    """q_i^H H q_j of the last evaluation (with a gradient, or any
    evaluation in band mode), (ns,nk,nb,nb) complex128 on the device; NaN
    where an occupation is <= 0."""
    if out is None:
      out = torch.empty((self.ns, self.nk, self.nb, self.nb),
                        dtype=torch.complex128, device=self.device)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_subspace_hamiltonian(
        self._plan, _lib.ptr(out), self._ws_ptr, self.workspace_bytes,
        _lib.stream_ptr(self.device)))
    return out

  def subspace_eigenvalues(self, vectors: bool = False, out=None):
    """Ascending eigenvalues (ns,nk,nb) of subspace_hamiltonian(), and with
    ``vectors`` its eigenvectors (ns,nk,nb,nb) complex128 (columns); cyclic
    Jacobi on the device."""
    if out is None:
      out = torch.empty((self.ns, self.nk, self.nb), dtype=torch.float64,
                        device=self.device)
    vec = None
    if vectors:
      vec = torch.empty((self.ns, self.nk, self.nb, self.nb),
                        dtype=torch.complex128, device=self.device)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_subspace_eigenvalues(
        self._plan, _lib.ptr(out), _lib.ptr(vec), self._ws_ptr,
        self.workspace_bytes, _lib.stream_ptr(self.device)))
    return (out, vec) if vectors else out

  def rotate(self, v, out_re=None, out_im=None):
    """Q V of the last evaluation in the (s,k,g,b) parameter layout;
    v is (ns,nk,nb,nb) complex128."""
    v = v.to(device=self.device, dtype=torch.complex128).contiguous()
    if tuple(v.shape) != (self.ns, self.nk, self.nb, self.nb):
      raise ValueError(f"rotation has shape {tuple(v.shape)}")
    shape = self.param_shape
    if out_re is None:
      out_re = torch.empty(shape, dtype=torch.float64, device=self.device)
    if out_im is None:
      out_im = torch.empty(shape, dtype=torch.float64, device=self.device)
    with torch.cuda.device(self.device):
      _lib.check(self._lib.jrx_plan_rotate(
        self._plan, _lib.ptr(v), _lib.ptr(out_re), _lib.ptr(out_im),
        self._ws_ptr, self.workspace_bytes, _lib.stream_ptr(self.device)))
    return out_re, out_im

  def set_timing(self, enable: bool):
    _lib.check(self._lib.jrx_plan_set_timing(self._plan, 1 if enable else 0))

  def read_timing(self):
    """Per-kernel-class (ms, launches) since the last read (synchronises)."""
    n = len(_lib.KERNEL_CLASSES)
    ms = (ctypes.c_double * n)()
    cnt = (ctypes.c_int * n)()
    _lib.check(self._lib.jrx_plan_read_timing(self._plan, ms, cnt))
    return {k: (ms[i], cnt[i]) for i, k in enumerate(_lib.KERNEL_CLASSES)}

  def value_and_grad(self, params):
    """(energies dict of floats, {'w_re': dE/dw_re, 'w_im': dE/dw_im})."""
    r = self.evaluate(params["w_re"], params["w_im"], grad=True)
    vals = r["energies"].tolist()
    return dict(zip(ENERGY_KEYS, vals)), {"w_re": r["dw_re"],
                                          "w_im": r["dw_im"]}


class _TotalEnergy(torch.autograd.Function):

  @staticmethod
  def forward(ctx, plan, w_re, w_im):
    r = plan.evaluate(w_re, w_im, grad=True)
    ctx.save_for_backward(r["dw_re"], r["dw_im"])
    return r["energies"][4].clone()

  @staticmethod
  def backward(ctx, g):
    dwr, dwi = ctx.saved_tensors
    return None, g * dwr, g * dwi


def total_energy(plan: EnergyPlan, w_re, w_im):
  """Differentiable total electronic energy (no Ewald term)."""
  return _TotalEnergy.apply(plan, w_re, w_im)
