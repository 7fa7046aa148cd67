"""torch.autograd Functions over the stand-alone HIP operators.

Every forward and backward runs HIP kernels of libjrx_hip.so on the current
stream; torch only allocates the outputs and workspaces.  Complex tensors
are complex128 (interleaved), real tensors float64, all on the GPU.
"""
import ctypes

import numpy as np
import torch

from . import _lib

_D = ctypes.c_double


def _L():
  return _lib.lib()


def _s():
  return _lib.stream_ptr()


def _p(t):
  return None if t is None else t.data_ptr()


def _dev():
  if not torch.cuda.is_available():
    raise RuntimeError("the jrystal_amd GPU path needs a ROCm GPU "
                       "(no CPU fallback)")
  return torch.device("cuda")


def as_device(x, dtype):
  """Host or device array -> contiguous device tensor of dtype."""
  dev = _dev()
  if isinstance(x, torch.Tensor):
    return x.to(device=dev, dtype=dtype).contiguous()
  return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=dev)


def workspace(nbytes):
  """256-byte aligned device workspace: (tensor keeping it alive, ptr)."""
  buf = torch.empty(int(nbytes) + 256, dtype=torch.uint8, device=_dev())
  return buf, (buf.data_ptr() + 255) // 256 * 256


def _fft_call(fn, x, scale):
  if x.ndim < 3:
    raise ValueError(f"Input must have at least 3 dimensions, got {x.ndim}")
  x = x.contiguous()
  nx, ny, nz = x.shape[-3:]
  batch = int(np.prod(x.shape[:-3])) if x.ndim > 3 else 1
  y = torch.empty_like(x)
  _lib.check(fn(_p(x), _p(y), batch, nx, ny, nz, _D(scale), _s()))
  return y


class IFFT3(torch.autograd.Function):
  """y = scale * unnormalised inverse DFT over the last three axes.  Its
  adjoint is scale * forward DFT (spmd/fft.py:88-101 uses the transpose
  rule of the same linear map)."""

  @staticmethod
  def forward(ctx, x, scale):
    ctx.scale = scale
    return _fft_call(_L().jrx_ifftn3d, x, scale)

  @staticmethod
  def backward(ctx, g):
    return _fft_call(_L().jrx_fftn3d, g, ctx.scale), None


class FFT3(torch.autograd.Function):
  """y = scale * forward DFT over the last three axes (complex input)."""

  @staticmethod
  def forward(ctx, x, scale):
    ctx.scale = scale
    return _fft_call(_L().jrx_fftn3d, x, scale)

  @staticmethod
  def backward(ctx, g):
    return _fft_call(_L().jrx_ifftn3d, g, ctx.scale), None


class RealToComplex(torch.autograd.Function):

  @staticmethod
  def forward(ctx, x):
    x = x.contiguous()
    y = torch.empty(x.shape, dtype=torch.complex128, device=x.device)
    _lib.check(_L().jrx_real_to_complex(_p(x), _p(y), x.numel(), _s()))
    return y

  @staticmethod
  def backward(ctx, g):
    g = g.contiguous()
    y = torch.empty(g.shape, dtype=torch.float64, device=g.device)
    _lib.check(_L().jrx_complex_real(_p(g), _p(y), g.numel(), _s()))
    return y


def scale_by(x, alpha):
  """alpha[0] * x for a device scalar alpha (HIP)."""
  x = x.contiguous()
  y = torch.empty_like(x)
  n = x.numel() * (2 if x.is_complex() else 1)
  _lib.check(_L().jrx_scale(_p(x), _p(alpha.reshape(1).contiguous()), _p(y),
                            n, _s()))
  return y


class Coefficients(torch.autograd.Function):
  """QR of W = w_re + i w_im over (s, k) then scatter into the FFT box
  (pw.coeff, pw.py:94-137).  Saves Q and R^-1 for the adjoint."""

  @staticmethod
  def forward(ctx, w_re, w_im, mask_index, box_shape):
    ns, nk, ng, nb = w_re.shape
    P = ns * nk
    dev = w_re.device
    w_re = w_re.contiguous()
    w_im = w_im.contiguous()
    q = torch.empty((ns, nk, nb, ng), dtype=torch.complex128, device=dev)
    rinv = torch.empty((P, nb, nb), dtype=torch.complex128, device=dev)
    ws, wp = workspace(_L().jrx_qr_workspace_bytes(P, ng, nb))
    _lib.check(_L().jrx_qr(_p(w_re), _p(w_im), P, ng, nb, _p(q), _p(rinv),
                           wp, ws.numel() - 256, _s()))
    N = int(np.prod(box_shape))
    box = torch.empty((ns, nk, nb) + tuple(box_shape), dtype=torch.complex128,
                      device=dev)
    _lib.check(_L().jrx_expand_coefficient(_p(q), _p(mask_index), _p(box),
                                           P * nb, ng, N, _s()))
    ctx.save_for_backward(q, rinv, mask_index)
    ctx.dims = (ns, nk, ng, nb, N)
    return box

  @staticmethod
  def backward(ctx, gbox):
    q, rinv, mask_index = ctx.saved_tensors
    ns, nk, ng, nb, N = ctx.dims
    P = ns * nk
    gbox = gbox.contiguous()
    gq = torch.empty_like(q)
    _lib.check(_L().jrx_squeeze_coefficient(_p(gbox), _p(mask_index), _p(gq),
                                            P * nb, ng, N, _s()))
    dwr = torch.empty((ns, nk, ng, nb), dtype=torch.float64, device=q.device)
    dwi = torch.empty_like(dwr)
    ws, wp = workspace(_L().jrx_qr_workspace_bytes(P, ng, nb))
    _lib.check(_L().jrx_qr_backward(_p(q), _p(gq), _p(rinv), P, ng, nb,
                                    _p(dwr), _p(dwi), wp, ws.numel() - 256,
                                    _s()))
    return dwr, dwi, None, None


class Density(torch.autograd.Function):
  """sum_{k,b} occ |psi|^2 per spin (occ given) or |psi|^2 per state."""

  @staticmethod
  def forward(ctx, psi, occ):
    psi = psi.contiguous()
    ns, nk, nb = psi.shape[:3]
    grid = psi.shape[3:]
    N = int(np.prod(grid))
    if occ is not None:
      out = torch.empty((ns,) + tuple(grid), dtype=torch.float64,
                        device=psi.device)
    else:
      out = torch.empty(psi.shape, dtype=torch.float64, device=psi.device)
    _lib.check(_L().jrx_density(_p(psi), _p(occ), _p(out), ns, nk * nb, N,
                                _s()))
    ctx.save_for_backward(psi, occ)
    return out

  @staticmethod
  def backward(ctx, g):
    psi, occ = ctx.saved_tensors
    ns, nk, nb = psi.shape[:3]
    N = int(np.prod(psi.shape[3:]))
    gpsi = torch.empty_like(psi)
    _lib.check(_L().jrx_density_backward(_p(psi), _p(occ), _p(g.contiguous()),
                                         _p(gpsi), ns, nk * nb, N, _s()))
    return gpsi, None


class Kinetic(torch.autograd.Function):

  @staticmethod
  def forward(ctx, coeff, gvec, kpts, occ):
    coeff = coeff.contiguous()
    ns, nk, nb = coeff.shape[:3]
    N = int(np.prod(coeff.shape[3:]))
    S = ns * nk * nb
    out = torch.empty((() if occ is not None else (ns, nk, nb)),
                      dtype=torch.float64, device=coeff.device)
    ws, wp = workspace(_L().jrx_kinetic_workspace_bytes(S, N))
    _lib.check(_L().jrx_kinetic_box(_p(coeff), _p(gvec), _p(kpts), _p(occ),
                                    ns, nk, nb, N, _p(out), wp,
                                    ws.numel() - 256, _s()))
    ctx.save_for_backward(coeff, gvec, kpts, occ)
    return out

  @staticmethod
  def backward(ctx, g):
    coeff, gvec, kpts, occ = ctx.saved_tensors
    ns, nk, nb = coeff.shape[:3]
    N = int(np.prod(coeff.shape[3:]))
    gc = torch.empty_like(coeff)
    _lib.check(_L().jrx_kinetic_box_backward(
      _p(coeff), _p(gvec), _p(kpts), _p(occ), _p(g.contiguous()), ns, nk, nb,
      N, _p(gc), _s()))
    return gc, None, None, None


def _reduce_ws(N):
  return workspace(_L().jrx_reduce_workspace_bytes(N))


class Hartree(torch.autograd.Function):

  @staticmethod
  def forward(ctx, rho_rec, gvec, vol, kohn_sham):
    rho_rec = rho_rec.contiguous()
    ns = rho_rec.shape[0]
    N = int(np.prod(rho_rec.shape[1:]))
    out = torch.empty((), dtype=torch.float64, device=rho_rec.device)
    ws, wp = _reduce_ws(N)
    _lib.check(_L().jrx_hartree(_p(rho_rec), _p(gvec), ns, N, _D(vol),
                                int(kohn_sham), _p(out), None, None, wp,
                                ws.numel() - 256, _s()))
    ctx.save_for_backward(rho_rec, gvec)
    ctx.args = (vol, int(kohn_sham))
    return out

  @staticmethod
  def backward(ctx, g):
    rho_rec, gvec = ctx.saved_tensors
    vol, ks = ctx.args
    ns = rho_rec.shape[0]
    N = int(np.prod(rho_rec.shape[1:]))
    grad = torch.empty_like(rho_rec)
    scratch = torch.empty((), dtype=torch.float64, device=rho_rec.device)
    ws, wp = _reduce_ws(N)
    _lib.check(_L().jrx_hartree(_p(rho_rec), _p(gvec), ns, N, _D(vol), ks,
                                _p(scratch), _p(grad), _p(g.contiguous()), wp,
                                ws.numel() - 256, _s()))
    return grad, None, None, None


class External(torch.autograd.Function):

  @staticmethod
  def forward(ctx, rho_rec, gvec, pos, chg):
    rho_rec = rho_rec.contiguous()
    ns = rho_rec.shape[0]
    N = int(np.prod(rho_rec.shape[1:]))
    out = torch.empty((), dtype=torch.float64, device=rho_rec.device)
    ws, wp = _reduce_ws(N)
    _lib.check(_L().jrx_external(_p(rho_rec), _p(gvec), _p(pos), _p(chg),
                                 pos.shape[0], ns, N, _p(out), None, None, wp,
                                 ws.numel() - 256, _s()))
    ctx.save_for_backward(rho_rec, gvec, pos, chg)
    return out

  @staticmethod
  def backward(ctx, g):
    rho_rec, gvec, pos, chg = ctx.saved_tensors
    ns = rho_rec.shape[0]
    N = int(np.prod(rho_rec.shape[1:]))
    grad = torch.empty_like(rho_rec)
    scratch = torch.empty((), dtype=torch.float64, device=rho_rec.device)
    ws, wp = _reduce_ws(N)
    _lib.check(_L().jrx_external(_p(rho_rec), _p(gvec), _p(pos), _p(chg),
                                 pos.shape[0], ns, N, _p(scratch), _p(grad),
                                 _p(g.contiguous()), wp, ws.numel() - 256,
                                 _s()))
    return grad, None, None, None


class ExchangeLDA(torch.autograd.Function):

  @staticmethod
  def forward(ctx, rho, vol):
    rho = rho.contiguous()
    ns = rho.shape[0]
    N = int(np.prod(rho.shape[1:]))
    out = torch.empty((), dtype=torch.float64, device=rho.device)
    ws, wp = _reduce_ws(N)
    _lib.check(_L().jrx_xc_lda(_p(rho), ns, N, _D(vol), _p(out), None, None,
                               wp, ws.numel() - 256, _s()))
    ctx.save_for_backward(rho)
    ctx.vol = vol
    return out

  @staticmethod
  def backward(ctx, g):
    rho, = ctx.saved_tensors
    ns = rho.shape[0]
    N = int(np.prod(rho.shape[1:]))
    grad = torch.empty_like(rho)
    scratch = torch.empty((), dtype=torch.float64, device=rho.device)
    ws, wp = _reduce_ws(N)
    _lib.check(_L().jrx_xc_lda(_p(rho), ns, N, _D(ctx.vol), _p(scratch),
                               _p(grad), _p(g.contiguous()), wp,
                               ws.numel() - 256, _s()))
    return grad, None


class SimplexProjection(torch.autograd.Function):
  """Rows of x (channels, n) projected onto {0 <= y <= 1, sum y = total_c}
  by the reference's shift-and-clip rule (occupation.py:295-337), after a
  sigmoid when ``sigmoid``; jrx_simplex_projection forward and VJP."""

  @staticmethod
  def forward(ctx, x, totals, sigmoid):
    x = x.contiguous()
    c, n = x.shape
    tot = torch.as_tensor(totals, dtype=torch.float64).to(x.device)
    y = torch.empty_like(x)
    _lib.check(_L().jrx_simplex_projection(_p(x), n, c, _p(tot),
                                           int(bool(sigmoid)), _p(y), None,
                                           None, _s()))
    ctx.save_for_backward(x, tot)
    ctx.sigmoid = bool(sigmoid)
    return y

  @staticmethod
  def backward(ctx, g):
    x, tot = ctx.saved_tensors
    c, n = x.shape
    y = torch.empty_like(x)
    xb = torch.empty_like(x)
    _lib.check(_L().jrx_simplex_projection(_p(x), n, c, _p(tot),
                                           int(ctx.sigmoid), _p(y),
                                           _p(g.contiguous()), _p(xb), _s()))
    return xb, None, None


class ExchangeCorrelation(torch.autograd.Function):
  """E_xc of the JRX_XC_* terms in `flags` on an unpolarised density
  (1, nx, ny, nz) with the G grid (nx, ny, nz, 3): jrx_xc forward and
  backward."""

  @staticmethod
  def forward(ctx, rho, gvec, vol, flags):
    rho = rho.contiguous()
    gvec = gvec.contiguous()
    nx, ny, nz = rho.shape[1:]
    out = torch.empty((), dtype=torch.float64, device=rho.device)
    ws, wp = workspace(_L().jrx_xc_workspace_bytes(nx, ny, nz, flags))
    _lib.check(_L().jrx_xc(_p(rho), _p(gvec), nx, ny, nz, _D(vol), flags,
                           _p(out), None, None, wp, ws.numel() - 256, _s()))
    ctx.save_for_backward(rho, gvec)
    ctx.vol, ctx.flags = vol, flags
    return out

  @staticmethod
  def backward(ctx, g):
    rho, gvec = ctx.saved_tensors
    nx, ny, nz = rho.shape[1:]
    grad = torch.empty_like(rho)
    scratch = torch.empty((), dtype=torch.float64, device=rho.device)
    ws, wp = workspace(_L().jrx_xc_workspace_bytes(nx, ny, nz, ctx.flags))
    _lib.check(_L().jrx_xc(_p(rho), _p(gvec), nx, ny, nz, _D(ctx.vol),
                           ctx.flags, _p(scratch), _p(grad),
                           _p(g.contiguous()), wp, ws.numel() - 256, _s()))
    return grad, None, None, None


def ewald(positions, charges, gvec, vol, eta, translations):
  pos = as_device(positions, torch.float64).reshape(-1, 3)
  chg = as_device(charges, torch.float64).reshape(-1)
  g = as_device(gvec, torch.float64).reshape(-1, 3)
  t = as_device(translations, torch.float64).reshape(-1, 3)
  out = torch.empty((), dtype=torch.float64, device=pos.device)
  ws, wp = workspace(_L().jrx_ewald_workspace_bytes(t.shape[0], g.shape[0]))
  _lib.check(_L().jrx_ewald(_p(pos), _p(chg), pos.shape[0], _p(g), g.shape[0],
                            _D(vol), _D(eta), _p(t), t.shape[0], _p(out), wp,
                            ws.numel() - 256, _s()))
  return out
