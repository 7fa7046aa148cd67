"""pw.param_init (pw.py:29-91 -> unitary_module.py:84-130): the reference's
layout (spin, kpt, g, band) with g = the mask's point count, one or two
spin channels, independent real and imaginary parts drawn U[0,1) (the
reference's jax.random.uniform; its threefry stream is not reproduced, the
seed drives numpy), deterministic per seed, and a start that pw.coeff turns
into orthonormal bands."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _mask():
  from jrystal_amd import workloads as W
  return np.asarray(W.si2_fcc(14.0).mask)


@pytest.mark.parametrize("restricted", [True, False])
def test_layout_and_distribution(restricted):
  from jrystal_amd import pw
  mask = _mask()
  p = pw.param_init(7, 8, 27, mask, spin_restricted=restricted)
  ns = 1 if restricted else 2
  shape = (ns, 27, int(mask.sum()), 8)
  for k in ("w_re", "w_im"):
    t = p[k]
    assert tuple(t.shape) == shape and t.dtype == torch.float64 and t.is_cuda
    x = t.cpu().numpy()
    assert x.min() >= 0.0 and x.max() < 1.0
    # U[0,1): mean 1/2, variance 1/12 (n > 3e5: 5-sigma bands ~3e-3)
    assert abs(x.mean() - 0.5) < 5e-3
    assert abs(x.var() - 1.0 / 12.0) < 3e-3
  a = p["w_re"].cpu().numpy().ravel()
  b = p["w_im"].cpu().numpy().ravel()
  assert abs(np.corrcoef(a, b)[0, 1]) < 1e-2


def test_seed_determinism_and_normal_option():
  from jrystal_amd import pw
  mask = _mask()
  a = pw.param_init(3, 8, 8, mask)
  b = pw.param_init(3, 8, 8, mask)
  c = pw.param_init(4, 8, 8, mask)
  assert torch.equal(a["w_re"], b["w_re"]) and torch.equal(a["w_im"], b["w_im"])
  assert not torch.equal(a["w_re"], c["w_re"])
  n = pw.param_init(3, 8, 8, mask, dist="normal")["w_re"].cpu().numpy()
  assert abs(n.mean()) < 2e-2 and abs(n.std() - 1.0) < 2e-2


def test_start_orthonormalises():
  from jrystal_amd import pw
  mask = _mask()
  p = pw.param_init(5, 8, 8, mask)
  q = pw.unitary_matrix(p).cpu().numpy()  # (spin, kpt, g, band)
  gram = np.einsum("skgi,skgj->skij", np.conj(q), q)
  np.testing.assert_allclose(gram, np.broadcast_to(np.eye(8), gram.shape),
                             atol=1e-12)
