"""ctypes binding of libjrx_hip.so (the HIP implementation, C ABI in
include/jrx.h).  There is no CPU fallback: importing a GPU entry point
without the built library raises."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjrx_hip.so")

JRX_OK = 0
JRX_ERR_INVALID = 1
JRX_ERR_UNSUPPORTED = 2
JRX_ERR_HIP = 3
JRX_ERR_NOMEM = 4
# Size of the static device workspace of jrx_energy_grad_static (jrx.h).
STATIC_WORKSPACE_BYTES = 64 << 20

_lib = None


class JrxProblem(ctypes.Structure):
  _fields_ = [
    ("ns", ctypes.c_int), ("nk", ctypes.c_int), ("nb", ctypes.c_int),
    ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
    ("natom", ctypes.c_int),
    ("mask", ctypes.c_void_p), ("kpts", ctypes.c_void_p),
    ("cell", ctypes.c_void_p), ("positions", ctypes.c_void_p),
    ("charges", ctypes.c_void_p), ("occupation", ctypes.c_void_p),
    ("xc", ctypes.c_int), ("chunk_states", ctypes.c_int),
    ("fft_variant", ctypes.c_int), ("xline_blocks", ctypes.c_int),
    ("plane_pad", ctypes.c_int), ("qr_passes", ctypes.c_int),
  ]


class JrxBatchMember(ctypes.Structure):
  _fields_ = [("plan", ctypes.c_void_p),
              ("w_re", ctypes.c_void_p), ("w_im", ctypes.c_void_p),
              ("dw_re", ctypes.c_void_p), ("dw_im", ctypes.c_void_p),
              ("m_re", ctypes.c_void_p), ("m_im", ctypes.c_void_p),
              ("v_re", ctypes.c_void_p), ("v_im", ctypes.c_void_p),
              ("energies", ctypes.c_void_p), ("workspace", ctypes.c_void_p),
              ("workspace_bytes", ctypes.c_size_t)]


class JrxAdamConfig(ctypes.Structure):
  _fields_ = [("lr", ctypes.c_double), ("b1", ctypes.c_double),
              ("b2", ctypes.c_double), ("eps", ctypes.c_double),
              ("eps_root", ctypes.c_double)]


class JrxPlanInfo(ctypes.Structure):
  _fields_ = [("ng", ctypes.c_int), ("active_planes", ctypes.c_int),
              ("chunk_states", ctypes.c_int), ("keep_planes", ctypes.c_int),
              ("workspace_bytes", ctypes.c_size_t),
              ("row_path", ctypes.c_int), ("rows", ctypes.c_int)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
_SZ = ctypes.c_size_t

# name -> argtypes (restype is int unless listed in _RESTYPES)
SIGNATURES = {
  "jrx_last_error": [],
  "jrx_version": [],
  "jrx_fftn3d": [_P, _P, _I64, _I, _I, _I, _D, _P],
  "jrx_ifftn3d": [_P, _P, _I64, _I, _I, _I, _D, _P],
  "jrx_plan_create": [ctypes.POINTER(ctypes.c_void_p),
                      ctypes.POINTER(JrxProblem)],
  "jrx_plan_destroy": [_P],
  "jrx_plan_workspace_bytes": [_P, ctypes.POINTER(_SZ)],
  "jrx_plan_get_info": [_P, ctypes.POINTER(JrxPlanInfo)],
  "jrx_plan_param_count": [_P, ctypes.POINTER(_I64)],
  "jrx_batch_create": [ctypes.POINTER(ctypes.c_void_p),
                       ctypes.POINTER(JrxBatchMember), _I,
                       ctypes.POINTER(JrxAdamConfig), _I],
  "jrx_batch_destroy": [_P],
  "jrx_batch_set_step_count": [_P, _I64, _P],
  "jrx_batch_get_step_count": [_P, ctypes.POINTER(_I64), _P],
  "jrx_batch_step": [_P, _P],
  "jrx_batch_capture": [_P, _P],
  "jrx_batch_replay": [_P, _P],
  "jrx_energy_grad": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P],
  "jrx_energy_grad_static": [_P, _P, _P, _P, _P, _P, _P, _P, _P],
  "jrx_qr_workspace_bytes": [_I64, _I, _I],
  "jrx_qr": [_P, _P, _I64, _I, _I, _P, _P, _P, _SZ, _P],
  "jrx_qr_backward": [_P, _P, _P, _I64, _I, _I, _P, _P, _P, _SZ, _P],
  "jrx_expand_coefficient": [_P, _P, _P, _I64, _I, _I64, _P],
  "jrx_squeeze_coefficient": [_P, _P, _P, _I64, _I, _I64, _P],
  "jrx_density": [_P, _P, _P, _I, _I64, _I64, _P],
  "jrx_density_backward": [_P, _P, _P, _P, _I, _I64, _I64, _P],
  "jrx_kinetic_workspace_bytes": [_I64, _I64],
  "jrx_kinetic_box": [_P, _P, _P, _P, _I, _I, _I, _I64, _P, _P, _SZ, _P],
  "jrx_kinetic_box_backward": [_P, _P, _P, _P, _P, _I, _I, _I, _I64, _P, _P],
  "jrx_reduce_workspace_bytes": [_I64],
  "jrx_hartree": [_P, _P, _I, _I64, _D, _I, _P, _P, _P, _P, _SZ, _P],
  "jrx_external": [_P, _P, _P, _P, _I, _I, _I64, _P, _P, _P, _P, _SZ, _P],
  "jrx_xc_lda": [_P, _I, _I64, _D, _P, _P, _P, _P, _SZ, _P],
  "jrx_xc_workspace_bytes": [_I, _I, _I, _I],
  "jrx_xc": [_P, _P, _I, _I, _I, _D, _I, _P, _P, _P, _P, _SZ, _P],
  "jrx_adam_update": [_P, _P, _P, _P, _I64, _D, _D, _D, _D, _D, _I64, _P],
  "jrx_simplex_projection": [_P, _I64, _I, _P, _I, _P, _P, _P, _P],
  "jrx_real_to_complex": [_P, _P, _I64, _P],
  "jrx_complex_real": [_P, _P, _I64, _P],
  "jrx_scale": [_P, _P, _P, _I64, _P],
  "jrx_ewald_workspace_bytes": [_I64, _I64],
  "jrx_ewald": [_P, _P, _I, _P, _I64, _D, _D, _P, _I64, _P, _P, _SZ, _P],
  "jrx_plan_precondition": [_P, _P, _P, _P, _P, _P, _P, _SZ, _P],
  "jrx_plan_band_norms": [_P, _P, _P, _SZ, _P],
  "jrx_plan_set_band_energies": [_P, _I],
  "jrx_plan_band_energies": [_P, _P, _P, _SZ, _P],
  "jrx_plan_set_occupation": [_P, _P, _P],
  "jrx_plan_orthonormal_params": [_P, _P, _P, _P, _SZ, _P],
  "jrx_plan_subspace_matrix": [_P, _P, _P, _SZ, _P],
  "jrx_plan_rotate": [_P, _P, _P, _P, _P, _SZ, _P],
  "jrx_plan_ritz_rotation": [_P, _P, _P, _SZ, _P],
  "jrx_plan_set_fixed_density": [_P, _P, _P, _SZ, _P],
  "jrx_plan_set_external_potential": [_P, _P, _P],
  "jrx_plan_set_nonlocal": [_P, _I, _P, _P],
  "jrx_plan_nonlocal_energies": [_P, _P, _P],
  "jrx_plan_set_kpoints": [_P, _P, _P],
  "jrx_plan_subspace_hamiltonian": [_P, _P, _P, _SZ, _P],
  "jrx_plan_subspace_eigenvalues": [_P, _P, _P, _P, _SZ, _P],
  "jrx_plan_density_stats": [_P, _P, _P, _SZ, _P],
  "jrx_plan_features": [_P, _P, _I64, _P, _P, _P, _P, _SZ, _P],
  "jrx_plan_mlp_apply": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P],
  "jrx_mlp_num_params": [],
  "jrx_mlp_forward": [_P, _P, _P, _I64, _P, _P],
  "jrx_mlp_backward_workspace_bytes": [_I64],
  "jrx_mlp_backward": [_P, _P, _P, _P, _I64, _P, _I, _P, _SZ, _P],
  "jrx_mlp_cosine_loss": [_P, _P, _P, _P, _I64, _I64, _P, _P, _P],
  "jrx_mlp_mse_workspace_bytes": [_I64],
  "jrx_mlp_mse_loss": [_P, _P, _I64, _P, _P, _P, _SZ, _P],
  "jrx_mlp_adam": [_P, _P, _P, _P, _I64, _D, _D, _D, _D, _I64, _P],
  "jrx_cg_dots_workspace_bytes": [],
  "jrx_plan_precondition_dots": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                 _P, _SZ, _P, _SZ, _P],
  "jrx_dot_workspace_bytes": [],
  "jrx_vec_dot": [_P, _P, _I64, _P, _P, _SZ, _P],
  "jrx_vec_axpby": [_D, _P, _D, _P, _P, _I64, _P],
  "jrx_plan_set_timing": [_P, _I],
  "jrx_plan_read_timing": [_P, ctypes.POINTER(_D), ctypes.POINTER(_I)],
}

# Kernel classes of jrx_plan_read_timing (enum JRX_K_* in include/jrx.h).
KERNEL_CLASSES = (
  "qr_gram", "qr_small", "qr_apply", "plane_ifft", "xline_density",
  "potential", "xline_vpsi", "plane_fft_gather", "qr_bwd_gram",
  "qr_bwd_apply", "other",
)
_RESTYPES = {"jrx_last_error": ctypes.c_char_p,
             "jrx_version": ctypes.c_char_p,
             "jrx_qr_workspace_bytes": ctypes.c_size_t,
             "jrx_xc_workspace_bytes": ctypes.c_size_t,
             "jrx_kinetic_workspace_bytes": ctypes.c_size_t,
             "jrx_reduce_workspace_bytes": ctypes.c_size_t,
             "jrx_ewald_workspace_bytes": ctypes.c_size_t,
             "jrx_dot_workspace_bytes": ctypes.c_size_t,
             "jrx_cg_dots_workspace_bytes": ctypes.c_size_t,
             "jrx_mlp_num_params": ctypes.c_size_t,
             "jrx_mlp_backward_workspace_bytes": ctypes.c_size_t,
             "jrx_mlp_mse_workspace_bytes": ctypes.c_size_t}


def lib():
  """Load the library (once).  Raises RuntimeError when it is not built."""
  global _lib
  if _lib is not None:
    return _lib
  if not os.path.exists(LIB_PATH):
    raise RuntimeError(
      f"{LIB_PATH} is missing: build it with `python -m jrystal_amd.build` "
      "(the HIP path has no CPU fallback)")
  h = ctypes.CDLL(LIB_PATH)
  for name, args in SIGNATURES.items():
    fn = getattr(h, name)
    fn.argtypes = args
    fn.restype = _RESTYPES.get(name, ctypes.c_int)
  _lib = h
  return h


def check(rc):
  """Map a jrx status code to a Python exception."""
  if rc == JRX_OK:
    return
  msg = lib().jrx_last_error().decode()
  if rc in (JRX_ERR_INVALID,):
    raise ValueError(msg)
  if rc == JRX_ERR_UNSUPPORTED:
    raise NotImplementedError(msg)
  if rc == JRX_ERR_NOMEM:
    raise MemoryError(msg)
  raise RuntimeError(msg)


def ptr(t):
  """Device (or host numpy) pointer of a tensor/array, None -> NULL."""
  if t is None:
    return None
  if isinstance(t, np.ndarray):
    return t.ctypes.data
  return t.data_ptr()


def stream_ptr(device=None):
  import torch
  return torch.cuda.current_stream(device).cuda_stream
