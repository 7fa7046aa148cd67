"""GPU FFT entry points (reference operator boundary ifftn3d / fftn3d,
jrystal/_src/spmd/fft.py:37-64) against numpy.fft."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("shape", [(2, 8, 8, 8), (3, 12, 15, 20),
                                   (1, 32, 32, 32), (2, 36, 45, 49),
                                   (1, 64, 64, 64), (1, 80, 42, 27),
                                   (1, 7, 128, 6)])
@pytest.mark.parametrize("inverse", [False, True])
def test_fft3d_matches_numpy(shape, inverse):
  from jrystal_amd import _lib
  L = _lib.lib()
  rng = np.random.default_rng(sum(shape))
  x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
  xt = torch.tensor(x, device="cuda")
  yt = torch.empty_like(xt)
  n = shape[1] * shape[2] * shape[3]
  fn = L.jrx_ifftn3d if inverse else L.jrx_fftn3d
  scale = 1.0 / n if inverse else 1.0
  print(f"fft shape={shape} inverse={inverse}", flush=True)
  _lib.check(fn(xt.data_ptr(), yt.data_ptr(), shape[0], shape[1], shape[2],
                shape[3], ctypes.c_double(scale), _lib.stream_ptr()))
  ref = (np.fft.ifftn if inverse else np.fft.fftn)(x, axes=(-3, -2, -1))
  got = yt.cpu().numpy()
  assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_fft3d_rejects_bad_size():
  from jrystal_amd import _lib
  L = _lib.lib()
  xt = torch.zeros((1, 11, 4, 4), dtype=torch.complex128, device="cuda")
  rc = L.jrx_fftn3d(xt.data_ptr(), xt.data_ptr(), 1, 11, 4, 4,
                    ctypes.c_double(1.0), _lib.stream_ptr())
  assert rc == _lib.JRX_ERR_INVALID
  with pytest.raises(ValueError):
    _lib.check(rc)
