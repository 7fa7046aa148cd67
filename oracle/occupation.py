"""Fixed occupations (oracle restatement of jrystal/_src/occupation.py)."""
import numpy as np


def uniform(num_k, num_electrons, spin=0, num_bands=None, spin_restricted=True):
  """occupation.py:147-190: 1/nk on the first (Ne+spin)//2 (up) and
  (Ne-spin)//2 (down) bands; slicing silently clips at num_bands."""
  num_bands = num_electrons if num_bands is None else num_bands
  occ = np.zeros([2, num_k, num_bands])
  occ[0, :, :(num_electrons + spin) // 2] = 1 / num_k
  occ[1, :, :(num_electrons - spin) // 2] = 1 / num_k
  if spin_restricted:
    return np.sum(occ, axis=0, keepdims=True)
  return occ


def gamma(num_k, num_electrons, spin=0, num_bands=None, spin_restricted=True):
  """occupation.py:193-237: occupations only at the first k-point."""
  num_bands = num_electrons if num_bands is None else num_bands
  occ = np.zeros([2, num_k, num_bands])
  occ[0, 0, :(num_electrons + spin) // 2] = 1
  occ[1, 0, :(num_electrons - spin) // 2] = 1
  if spin_restricted:
    return np.sum(occ, axis=0, keepdims=True)
  return occ


def fermi_dirac_entropy(occupation, eps=1e-8):
  """entropy.py:19-54."""
  num_spin, num_k, _ = occupation.shape
  top = (3 - num_spin) / num_k
  return -np.sum(occupation * np.log(eps + occupation) +
                 (top - occupation) * np.log(eps + top - occupation))


# -------------------------------------------- parameterised occupations ----
def sigmoid(x):
  return 1.0 / (1.0 + np.exp(-x))


def simplex_projector_init(num_bands, num_kpts):
  """occupation.py:281-292."""
  n = num_bands * num_kpts
  p = ((np.arange(n) - n // 2) * 0.1).reshape(num_kpts, num_bands)
  return {"param_up": p, "param_down": p.copy()}


def proj(x, total):
  """occupation.py:295-337, restated: returns (y, branch, k, idx) where
  branch is "down" (sum x > total: y = max(x - lam, 0)) or "up"
  (y = min(x + lam, 1)), k the number of sorted entries left out of the
  shift's average and idx the sort order used."""
  x = np.asarray(x, np.float64)
  n = x.size
  ar = np.arange(n)
  s = x.sum()
  with np.errstate(divide="ignore", invalid="ignore"):
    if s > total:
      idx = np.argsort(x, kind="stable")
      xs = x[idx]
      cs = np.cumsum(xs)
      rho = (s - total - cs) / (n - ar - 1)
      k = int(np.argmax(xs > rho))
      lam = (s - total) / n if k == 0 else (s - total - cs[k - 1]) / (n - k)
      return np.maximum(x - lam, 0.0), "down", k, idx
    idx = np.argsort(-x, kind="stable")
    xs = x[idx]
    cs = np.cumsum(xs)
    rho = (s - total + (ar + 1 - cs)) / (n - ar - 1)
    k = int(np.argmax(xs - rho < 1.0))
    lam = (total - s) / n if k == 0 else (total - s - (k - cs[k - 1])) / (n - k)
    return np.minimum(x + lam, 1.0), "up", k, idx


def proj_vjp(x, total, ybar):
  """Vector-Jacobian product of proj, derived by hand: with the active set
  fixed, lam depends on x through sum x minus the k left-out sorted
  entries, so d lam / d x_j = +-[j not left out] / (n - k) (k = 0: all j,
  1/n); entries clipped at the bound have zero derivative."""
  y, branch, k, idx = proj(x, total)
  n = x.size
  free = np.ones(n, bool)
  free[idx[:k]] = False  # the k left-out entries (smallest, or largest)
  if branch == "down":
    live = y > 0.0
    dl = free.astype(np.float64) / (n - k)
    return ybar * live - np.sum(ybar * live) * dl
  live = y < 1.0
  dl = -free.astype(np.float64) / (n - k)
  return ybar * live + np.sum(ybar * live) * dl


def simplex_projector(params, num_electrons, spin=0, spin_restricted=True):
  """occupation.py:340-366."""
  nk, nb = params["param_up"].shape
  m_up = (num_electrons + spin) // 2 * nk
  m_down = (num_electrons - spin) // 2 * nk
  up = proj(sigmoid(params["param_up"]).ravel(), m_up)[0].reshape(nk, nb)
  down = proj(sigmoid(params["param_down"]).ravel(), m_down)[0].reshape(nk, nb)
  occ = np.stack([up, down]) / nk
  return occ.sum(0, keepdims=True) if spin_restricted else occ


def simplex_projector_vjp(params, num_electrons, occ_bar, spin=0,
                          spin_restricted=True):
  """d/d params of sum(occ_bar * simplex_projector(params))."""
  nk, nb = params["param_up"].shape
  out = {}
  for ch, m, c in (("param_up", (num_electrons + spin) // 2 * nk, 0),
                   ("param_down", (num_electrons - spin) // 2 * nk, 1)):
    s = sigmoid(params[ch]).ravel()
    ob = (occ_bar[0] if spin_restricted else occ_bar[c]).ravel() / nk
    out[ch] = (proj_vjp(s, m, ob) * s * (1 - s)).reshape(nk, nb)
  return out


def idempotent(params, num_kpts, spin_restricted=True):
  """occupation.py:87-144: diagonal of U U^T for the orthonormal factor U
  of each channel's real parameter matrix."""
  def rows(w):
    u = np.linalg.qr(w, mode="reduced")[0]
    return (u * u).sum(1)
  nb = params["param_up"]["w_re"].shape[0] // num_kpts
  up = rows(params["param_up"]["w_re"]).reshape(num_kpts, nb)
  down = rows(params["param_down"]["w_re"]).reshape(num_kpts, nb)
  if spin_restricted:
    return ((up + down) / num_kpts)[None]
  return np.stack([up, down]) / num_kpts


def fermi_dirac_entropy_grad(occupation, eps=1e-8):
  """dS/df of entropy.py:19-54, by hand."""
  num_spin, num_k, _ = occupation.shape
  top = (3 - num_spin) / num_k
  f = occupation
  return -(np.log(eps + f) + f / (eps + f)
           - np.log(eps + top - f) - (top - f) / (eps + top - f))
