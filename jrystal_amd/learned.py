"""Learned diagonal preconditioner and warm start.

A small FP32 network (8 features -> 64 -> 64 -> 2, tanh; HIP kernels in
csrc/mlp.hip and csrc/plan.hip) maps, for every plane-wave coefficient
(s,k,G,b), the features

  log1p(|G+k|^2), log1p(|G+k|^2 / tb_b), log tb_b,
  log rho_mean, log1p(rho_var / rho_mean^2), log1p(gap_b)

(tb_b = twice the band's kinetic energy, gap_b = its Rayleigh-quotient
energy below the highest occupied band of its k-point, rho_mean / rho_var
= statistics of the current total density) to

  output 0: log of a multiplier of the Teter-Payne-Allan preconditioner,
  output 1: log amplitude of the warm-start coefficients.

Training data are trajectories of the plain method itself (preconditioned
CG with the TPA preconditioner and Rayleigh-Ritz gauge resets):

* preconditioner head -- at snapshots W_t of a trajectory that converges
  to Q*, the direction to the minimum is the Grassmann direction
  D_t = (I - Q_t Q_t^H) Q* (Q_t^H Q*)^{-1}, mapped to W by R_t.  The head
  is trained to maximise the cosine between exp(y0) * (P_TPA g_t) and D_t
  over sampled occupied coefficients (loss 1 - cos^2 per snapshot);
* warm-start head -- regression of 1/2 log(w(k,G) ng / n_occ) with
  w = sum_{b occupied} |Q*_{kGb}|^2, the weight of plane wave G+k in the
  converged occupied subspace, from the features of the random start.

The minimiser (minimize.Minimizer) falls back to the plain TPA
preconditioner whenever a line search fails to decrease the energy.
"""
import ctypes
import os
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import _lib, checkpoint
from .objective import EnergyPlan

NFEAT = 8
# CG restart interval (iterations) used with the learned factor, which is
# refreshed at every restart.  Chosen on training-distribution cells
# (strained 8-atom cubes, strain seeds 11-13, 4x4x4 k), with one plain first
# cycle: 45 beat 30 and 60 (time to |dE| < 1e-6 Ha); the plain method's best
# interval on the same cells is minimize.PLAIN_RESTART.
LEARNED_RESTART = 45
# length of the plain first cycle before the factor is used (same cells:
# 30 beat 10, 20, 45 and 60)
LEARNED_FIRST_RESTART = 30
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                             "data", "precond_mlp.npz")


def num_params() -> int:
  return int(_lib.lib().jrx_mlp_num_params())


class PrecondNet:
  """Network parameters (flat FP32) and feature normalisation on a device."""

  def __init__(self, params, norm, device=None, use_gap=True):
    self.device = torch.device(device if device is not None else "cuda")
    self.use_gap = bool(use_gap)  # feed the band-gap feature (slot 5)
    self.params = torch.as_tensor(np.asarray(params, np.float32)).to(
      self.device).contiguous()
    self.norm = torch.as_tensor(np.asarray(norm, np.float32)).to(
      self.device).contiguous()
    if self.params.numel() != num_params() or self.norm.numel() != 2 * NFEAT:
      raise ValueError("network parameter / normalisation size mismatch")

  @classmethod
  def initial(cls, seed=0, device=None, use_gap=True):
    """Glorot-uniform hidden layers, a small output layer (outputs start
    near 0: factor 1, amplitude 1), identity normalisation."""
    rng = np.random.default_rng(seed)
    h, nin = 64, NFEAT
    parts = []
    lim1 = np.sqrt(6.0 / (nin + h))
    parts.append(rng.uniform(-lim1, lim1, (h, nin)))
    parts.append(np.zeros(h))
    lim2 = np.sqrt(6.0 / (2 * h))
    parts.append(rng.uniform(-lim2, lim2, (h, h)))
    parts.append(np.zeros(h))
    parts.append(rng.uniform(-0.01, 0.01, (2, h)))
    parts.append(np.zeros(2))
    p = np.concatenate([x.ravel() for x in parts]).astype(np.float32)
    norm = np.concatenate([np.zeros(nin), np.ones(nin)]).astype(np.float32)
    return cls(p, norm, device, use_gap)

  @classmethod
  def load(cls, path=DEFAULT_MODEL, device=None):
    with np.load(path, allow_pickle=False) as z:
      use_gap = bool(z["use_gap"]) if "use_gap" in z else False
      return cls(z["params"], z["norm"], device, use_gap)

  def save(self, path, **meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=self.params.cpu().numpy(),
             norm=self.norm.cpu().numpy(), use_gap=np.int32(self.use_gap),
             **{k: np.asarray(v) for k, v in meta.items()})

  def forward(self, x):
    """y (n, 2) for raw features x (n, 8) (device float32)."""
    x = x.to(device=self.device, dtype=torch.float32).contiguous()
    y = torch.empty((x.shape[0], 2), dtype=torch.float32, device=self.device)
    _lib.check(_lib.lib().jrx_mlp_forward(
      _lib.ptr(self.params), _lib.ptr(self.norm), _lib.ptr(x), x.shape[0],
      _lib.ptr(y), _lib.stream_ptr(self.device)))
    return y


# --------------------------------------------------------- plan helpers ----
def density_stats(plan: EnergyPlan, out=None):
  """(mean, variance) of the total density of the last evaluation."""
  if out is None:
    out = torch.empty(2, dtype=torch.float64, device=plan.device)
  _lib.check(_lib.lib().jrx_plan_density_stats(
    plan._plan, _lib.ptr(out), plan._ws_ptr, plan.workspace_bytes,
    _lib.stream_ptr(plan.device)))
  return out


def band_gaps(plan: EnergyPlan):
  """Per band (s,k,b), from the last evaluation with a gradient: the
  Rayleigh quotient of the highest occupied band of its (s,k) block minus
  its own (Q^H Qbar has diagonal 2 f eps_b); 0 for empty bands.  Device
  float64 tensor."""
  d = torch.real(torch.diagonal(plan.subspace_matrix(), dim1=-2, dim2=-1))
  occ = torch.as_tensor(plan.occupation, device=plan.device)
  occupied = occ > 0
  eps = torch.where(occupied, d / (2 * torch.where(occupied, occ, 1.0)),
                    torch.full_like(d, -np.inf))
  top = eps.max(dim=-1, keepdim=True).values
  return torch.where(occupied, top - eps, torch.zeros_like(d)).contiguous()


def coefficient_features(plan: EnergyPlan, idx, stats, gap=None):
  """Raw features (n, 8) float32 of flat coefficient indices idx."""
  idx = torch.as_tensor(idx, dtype=torch.int64).to(plan.device).contiguous()
  x = torch.empty((idx.numel(), NFEAT), dtype=torch.float32,
                  device=plan.device)
  _lib.check(_lib.lib().jrx_plan_features(
    plan._plan, _lib.ptr(idx), idx.numel(), _lib.ptr(stats), _lib.ptr(gap),
    _lib.ptr(x), plan._ws_ptr, plan.workspace_bytes,
    _lib.stream_ptr(plan.device)))
  return x


def apply(plan: EnergyPlan, net: PrecondNet, stats, gap=None, factor=None,
          w_re=None, w_im=None):
  """Per-coefficient inference: factor = exp(y0), w *= exp(y1)."""
  _lib.check(_lib.lib().jrx_plan_mlp_apply(
    plan._plan, _lib.ptr(net.params), _lib.ptr(net.norm), _lib.ptr(stats),
    _lib.ptr(gap), _lib.ptr(factor), _lib.ptr(w_re), _lib.ptr(w_im),
    plan._ws_ptr, plan.workspace_bytes, _lib.stream_ptr(plan.device)))


def warm_start(plan: EnergyPlan, net: PrecondNet, params):
  """Initial guess: the seeded start W0 reshaped coefficient-wise by the
  network's amplitude, W = exp(y1(features of W0)) * W0."""
  shape = (2,) + tuple(plan.param_shape)
  W = torch.empty(shape, dtype=torch.float64, device=plan.device)
  W[0].copy_(torch.as_tensor(params["w_re"]))
  W[1].copy_(torch.as_tensor(params["w_im"]))
  plan.evaluate(W[0], W[1], grad=False)
  apply(plan, net, density_stats(plan), w_re=W[0], w_im=W[1])
  return {"w_re": W[0], "w_im": W[1]}


class LearnedFactor:
  """factor_fn of minimize.Minimizer: the network's preconditioner factor,
  recomputed at every gauge reset (from the evaluation that reset made) and
  held fixed between resets, so that CG sees one preconditioner per
  cycle."""

  def __init__(self, plan: EnergyPlan, net: PrecondNet, first_cycle: int = 1):
    self.plan, self.net = plan, net
    self.first_cycle = first_cycle  # restart cycles before the factor is used
    self.factor = torch.empty(plan.param_shape, dtype=torch.float64,
                              device=plan.device)
    self.stats = torch.empty(2, dtype=torch.float64, device=plan.device)
    self._stamp = None

  def __call__(self, minimizer):
    # restarts counts the gauge reset that starts a run, so cycle c runs
    # with restarts == c; the first first_cycle cycles use plain TPA, while
    # the iterate is still far from the states the network was trained on
    if minimizer.restarts <= self.first_cycle:
      return None
    if minimizer.restarts != self._stamp:
      density_stats(self.plan, self.stats)
      gap = band_gaps(self.plan) if self.net.use_gap else None
      apply(self.plan, self.net, self.stats, gap=gap, factor=self.factor)
      self._stamp = minimizer.restarts
    return self.factor


# ------------------------------------------------------------- training ----
@dataclass
class Dataset:
  """Samples of the two heads (host arrays)."""
  x_pc: np.ndarray    # (nsnap * snap, 8) features, preconditioner head
  a: np.ndarray       # Re(conj(D) P_TPA g)
  b: np.ndarray       # |P_TPA g|^2
  t2: np.ndarray      # (nsnap,) sum |D|^2 over the snapshot's samples
  snap: int
  x_ws: np.ndarray    # (m, 8) features of the random start, warm-start head
  y_ws: np.ndarray    # (m,) target log amplitude

  @staticmethod
  def concat(parts: List["Dataset"]) -> "Dataset":
    snap = parts[0].snap
    assert all(p.snap == snap for p in parts)
    return Dataset(np.concatenate([p.x_pc for p in parts]),
                   np.concatenate([p.a for p in parts]),
                   np.concatenate([p.b for p in parts]),
                   np.concatenate([p.t2 for p in parts]), snap,
                   np.concatenate([p.x_ws for p in parts]),
                   np.concatenate([p.y_ws for p in parts]))

  def save(self, path):
    np.savez(path, x_pc=self.x_pc, a=self.a, b=self.b, t2=self.t2,
             snap=self.snap, x_ws=self.x_ws, y_ws=self.y_ws)

  def split(self, frac: float, seed: int = 0):
    """(train, validation): a random fraction of the snapshots and of the
    warm-start samples held out."""
    rng = np.random.default_rng(seed)
    ns = len(self.t2)
    val = np.zeros(ns, bool)
    val[rng.permutation(ns)[:max(1, int(round(frac * ns)))]] = True
    rows = np.repeat(val, self.snap)
    wval = rng.random(len(self.y_ws)) < frac

    def pick(m, r, w):
      return Dataset(self.x_pc[r], self.a[r], self.b[r], self.t2[m],
                     self.snap, self.x_ws[w], self.y_ws[w])
    return pick(~val, ~rows, ~wval), pick(val, rows, wval)

  @staticmethod
  def load(path):
    with np.load(path, allow_pickle=False) as z:
      return Dataset(z["x_pc"], z["a"], z["b"], z["t2"], int(z["snap"]),
                     z["x_ws"], z["y_ws"])


def _occupied_block(occ):
  """Number of leading occupied bands, required equal for every (s,k)."""
  pos = occ > 0
  nocc = pos.sum(-1)
  n = int(nocc.flat[0])
  if not np.all(nocc == n) or not np.all(pos[..., :n]):
    raise NotImplementedError("trajectory sampling needs the same leading "
                              "block of occupied bands at every (s,k)")
  return n


def _host_complex(re, im):
  return re.cpu().numpy() + 1j * im.cpu().numpy()


def grassmann_direction(q_t, q_star):
  """D = (I - q_t q_t^H) q_star (q_t^H q_star)^{-1} per leading index: the
  tangent at span(q_t) with span(q_t + D) = span(q_star) (q_* (..., G, n)
  with orthonormal columns)."""
  m = np.conj(np.swapaxes(q_t, -1, -2)) @ q_star
  return (q_star - q_t @ m) @ np.linalg.inv(m)


def collect_trajectory(wl, seed: int, snap_every: int = 8,
                       samples: int = 512, ws_samples: int = 4096,
                       max_iter: int = 800, max_snaps: int = 60,
                       rng: Optional[np.random.Generator] = None,
                       device=None, out_dir: Optional[str] = None) -> Dataset:
  """One plain-method trajectory of workload wl from wl.params(seed), stored
  as checkpoint files (jrystal_amd.checkpoint: the start, every snap_every-th
  iterate, the converged end) in out_dir (a temporary directory when None)
  and turned into training samples of both heads by trajectory_dataset."""
  import tempfile
  from .minimize import Minimizer
  tmp = None
  if out_dir is None:
    tmp = tempfile.TemporaryDirectory(prefix="jrx_traj_")
    out_dir = tmp.name
  try:
    plan = EnergyPlan.from_workload(wl, device=device)
    meta = {"workload": wl.name, "seed": int(seed)}
    p0 = wl.params(seed)
    start = checkpoint.save(os.path.join(out_dir, "start.npz"), step=0,
                            params=p0, occupation=plan.occupation,
                            meta=dict(meta, kind="start"))
    snaps = []

    def cb(m, it, e):
      if it % snap_every or len(snaps) >= max_snaps:
        return
      snaps.append(checkpoint.save(
        os.path.join(out_dir, f"snap_{it:05d}.npz"), step=it,
        params={"w_re": m.W[0], "w_im": m.W[1]}, occupation=plan.occupation,
        energy_history=[e], meta=dict(meta, kind="snapshot")))

    m = Minimizer(plan, p0)
    r = m.run(max_iter=max_iter, callback=cb)
    final = checkpoint.save(
      os.path.join(out_dir, "final.npz"), step=r.iterations,
      params={"w_re": m.W[0], "w_im": m.W[1]}, occupation=plan.occupation,
      energy_history=[e for _, e in r.history],
      meta=dict(meta, kind="final", status=r.status))
    plan.close()
    return trajectory_dataset(wl, start, snaps, final, samples=samples,
                              ws_samples=ws_samples,
                              rng=rng if rng is not None
                              else np.random.default_rng(seed),
                              device=device)
  finally:
    if tmp is not None:
      tmp.cleanup()


def trajectory_dataset(wl, start: str, snapshots: List[str], final: str,
                       samples: int = 512, ws_samples: int = 4096,
                       rng: Optional[np.random.Generator] = None,
                       device=None) -> Dataset:
  """Training samples from a trajectory stored as checkpoint files.

  Preconditioner head, per snapshot W_t (samples random occupied
  coefficients): features at W_t, a = Re(conj(D) P_TPA g) and
  b = |P_TPA g|^2 with g = dE/dW at W_t and D the Grassmann direction from
  span(Q_t) to the converged span(Q*), carried to W_t's gauge (D R_oo).
  Warm-start head: features at the start, target 1/2 log of the converged
  occupied-subspace weight of each plane wave."""
  rng = rng if rng is not None else np.random.default_rng(0)
  plan = EnergyPlan.from_workload(wl, device=device)
  ns, nk, ng, nb = plan.param_shape
  nocc = _occupied_block(plan.occupation)
  dev = plan.device

  def sample_idx(n):
    s = rng.integers(0, ns, n)
    k = rng.integers(0, nk, n)
    g = rng.integers(0, ng, n)
    b = rng.integers(0, nocc, n)
    return ((s * nk + k) * ng + g) * nb + b

  def params_of(path):
    ck = checkpoint.load(path)
    return {k: torch.as_tensor(v, device=dev)
            for k, v in ck["params"].items()}

  p0 = params_of(start)
  plan.evaluate(p0["w_re"], p0["w_im"], grad=False)
  idx_ws = sample_idx(ws_samples)
  x_ws = coefficient_features(plan, idx_ws, density_stats(plan)).cpu().numpy()

  pf = params_of(final)
  plan.evaluate(pf["w_re"], pf["w_im"], grad=False)
  qr, qi = plan.orthonormal_params()
  qs = _host_complex(qr, qi)[..., :nocc]

  xs, as_, bs, ts = [], [], [], []
  for path in snapshots:
    w_t = params_of(path)
    out = plan.evaluate(w_t["w_re"], w_t["w_im"], grad=True)
    hr, hi = plan.precondition(out["dw_re"], out["dw_im"])
    qr, qi = plan.orthonormal_params()
    q = _host_complex(qr, qi)[..., :nocc]
    w = _host_complex(w_t["w_re"], w_t["w_im"])[..., :nocc]
    idx = sample_idx(samples)
    it = torch.as_tensor(idx, device=dev)
    h = (hr.reshape(-1)[it].cpu().numpy() +
         1j * hi.reshape(-1)[it].cpu().numpy())
    x = coefficient_features(plan, idx, density_stats(plan),
                             band_gaps(plan)).cpu().numpy()
    ov = np.linalg.svd(np.conj(np.swapaxes(q, -1, -2)) @ qs, compute_uv=False)
    if ov.min() < 0.2:  # far from the minimum: D is ill-defined
      continue
    d = grassmann_direction(q, qs)                   # (s,k,g,nocc)
    r = np.conj(np.swapaxes(q, -1, -2)) @ w          # R_oo
    dw = (d @ r).reshape(ns, nk, ng, nocc)
    b_ = idx % nb
    sk_g = idx // nb
    t = dw.reshape(-1, nocc)[sk_g, b_]
    xs.append(x)
    as_.append(np.real(np.conj(t) * h))
    bs.append(np.abs(h)**2)
    ts.append(np.sum(np.abs(t)**2))
  plan.close()
  wgt = np.sum(np.abs(qs)**2, axis=-1)                # (s,k,g)
  wnorm = wgt * ng / nocc
  sk_g = idx_ws // nb
  y_ws = 0.5 * np.log(np.maximum(wnorm.reshape(-1)[sk_g], 1e-26))
  f32 = np.float32
  if not ts:
    xs, as_, bs, ts = [np.zeros((0, NFEAT))], [np.zeros(0)], [np.zeros(0)], []
  return Dataset(np.concatenate(xs).astype(f32),
                 np.concatenate(as_).astype(f32),
                 np.concatenate(bs).astype(f32),
                 np.asarray(ts, f32), samples, x_ws.astype(f32),
                 np.clip(y_ws, -30.0, 8.0).astype(f32))


def feature_norm(ds: Dataset):
  x = np.concatenate([ds.x_pc, ds.x_ws]).astype(np.float64)
  mu = x.mean(0)
  sd = x.std(0)
  inv = np.where(sd > 1e-6, 1.0 / np.maximum(sd, 1e-6), 1.0)
  return np.concatenate([mu, inv]).astype(np.float32)


class Trainer:
  """Full-batch training of both heads with the HIP kernels:
  loss = mean_snapshots(1 - cos^2) + ws_weight * mse(warm start)."""

  def __init__(self, net: PrecondNet, ds: Dataset, lr=1e-3, b1=0.9,
               b2=0.999, ws_weight=0.1):
    dev = net.device
    self.net, self.lr, self.b1, self.b2 = net, lr, b1, b2
    self.ws_weight = ws_weight
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(dev)  # noqa
    self.x_pc, self.a, self.b, self.t2 = f(ds.x_pc), f(ds.a), f(ds.b), f(ds.t2)
    self.snap, self.nsnap = ds.snap, len(ds.t2)
    self.x_ws, self.y_ws = f(ds.x_ws), f(ds.y_ws)
    n = max(self.x_pc.shape[0], self.x_ws.shape[0])
    L = _lib.lib()
    self.ws = torch.empty(max(L.jrx_mlp_backward_workspace_bytes(n),
                              L.jrx_mlp_mse_workspace_bytes(n)) // 4 + 64,
                          dtype=torch.float32, device=dev)
    self.grad = torch.zeros(num_params(), dtype=torch.float32, device=dev)
    self.m = torch.zeros_like(self.grad)
    self.v = torch.zeros_like(self.grad)
    self.loss_pc = torch.empty(max(self.nsnap, 1), dtype=torch.float32,
                               device=dev)
    self.loss_ws = torch.empty(1, dtype=torch.float32, device=dev)
    self.dy_pc = torch.empty((self.x_pc.shape[0], 2), dtype=torch.float32,
                             device=dev)
    self.dy_ws = torch.empty((self.x_ws.shape[0], 2), dtype=torch.float32,
                             device=dev)
    self.count = 0

  def loss_and_grad(self):
    """Fills self.grad; returns (cosine loss, warm-start mse) tensors."""
    L, net, s = _lib.lib(), self.net, _lib.stream_ptr(self.net.device)
    wsb = self.ws.numel() * 4
    acc = 0
    if self.nsnap:
      y = net.forward(self.x_pc)
      _lib.check(L.jrx_mlp_cosine_loss(
        _lib.ptr(y), _lib.ptr(self.a), _lib.ptr(self.b), _lib.ptr(self.t2),
        self.snap, self.nsnap, _lib.ptr(self.dy_pc), _lib.ptr(self.loss_pc),
        s))
      _lib.check(L.jrx_mlp_backward(
        _lib.ptr(net.params), _lib.ptr(net.norm), _lib.ptr(self.x_pc),
        _lib.ptr(self.dy_pc), self.x_pc.shape[0], _lib.ptr(self.grad), 0,
        _lib.ptr(self.ws), wsb, s))
      acc = 1
    if self.x_ws.shape[0] and self.ws_weight > 0:
      y = net.forward(self.x_ws)
      _lib.check(L.jrx_mlp_mse_loss(
        _lib.ptr(y), _lib.ptr(self.y_ws), self.x_ws.shape[0],
        _lib.ptr(self.dy_ws), _lib.ptr(self.loss_ws), _lib.ptr(self.ws), wsb,
        s))
      self.dy_ws.mul_(self.ws_weight)
      _lib.check(L.jrx_mlp_backward(
        _lib.ptr(net.params), _lib.ptr(net.norm), _lib.ptr(self.x_ws),
        _lib.ptr(self.dy_ws), self.x_ws.shape[0], _lib.ptr(self.grad), acc,
        _lib.ptr(self.ws), wsb, s))
    return self.loss_pc[:self.nsnap].mean(), self.loss_ws[0]

  def step(self):
    lpc, lws = self.loss_and_grad()
    self.count += 1
    _lib.check(_lib.lib().jrx_mlp_adam(
      _lib.ptr(self.net.params), _lib.ptr(self.grad), _lib.ptr(self.m),
      _lib.ptr(self.v), self.grad.numel(), ctypes.c_double(self.lr),
      ctypes.c_double(self.b1), ctypes.c_double(self.b2),
      ctypes.c_double(1e-8), self.count, _lib.stream_ptr(self.net.device)))
    return lpc, lws


def train(ds: Dataset, epochs=1500, lr=2e-3, lr_end=5e-5, seed=0,
          ws_weight=0.1, val_frac=0.1, use_gap=True, device=None,
          log_every=0, checkpoint_every=0, checkpoints=None):
  """Train a fresh network on ds with Adam and a cosine learning-rate
  schedule lr -> lr_end; the parameters of the epoch with the lowest
  validation loss (cosine + ws_weight * mse on a held-out fraction of
  snapshots and warm-start samples) are returned.  Returns (net, history
  of (epoch, train cos, train mse, val cos, val mse)).  With
  checkpoint_every > 0, (epoch, PrecondNet copy) pairs are appended to the
  list `checkpoints` along the way."""
  if not use_gap:  # the band-gap slot reads 0, as at inference
    ds = Dataset(ds.x_pc.copy(), ds.a, ds.b, ds.t2, ds.snap, ds.x_ws,
                 ds.y_ws)
    ds.x_pc[:, 5] = 0.0
  net = PrecondNet.initial(seed, device, use_gap)
  net.norm.copy_(torch.as_tensor(feature_norm(ds)))
  if val_frac > 0 and len(ds.t2) >= 10:
    ds_tr, ds_val = ds.split(val_frac, seed)
  else:
    ds_tr, ds_val = ds, ds
  tr = Trainer(net, ds_tr, lr=lr, ws_weight=ws_weight)
  ev = Trainer(net, ds_val, lr=lr, ws_weight=ws_weight)
  hist = []
  best, best_params = np.inf, net.params.clone()

  def validate(ep, lpc, lws):
    nonlocal best
    vpc, vws = ev.loss_and_grad()
    vpc, vws = float(vpc), float(vws)
    score = vpc + ws_weight * vws
    if score < best:
      best = score
      best_params.copy_(net.params)
    hist.append((ep, float(lpc), float(lws), vpc, vws))
    return vpc, vws

  for ep in range(epochs):
    tr.lr = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * ep / epochs))
    lpc, lws = tr.step()
    if checkpoint_every and checkpoints is not None and \
        (ep + 1) % checkpoint_every == 0:
      checkpoints.append((ep + 1, PrecondNet(net.params.cpu().numpy(),
                                             net.norm.cpu().numpy(),
                                             net.device, net.use_gap)))
    if ep % 10 == 0 or ep == epochs - 1:
      vpc, vws = validate(ep, lpc, lws)
      if log_every and (ep % log_every == 0 or ep == epochs - 1):
        print(f"epoch {ep}: train cos {float(lpc):.5f} mse {float(lws):.4f}"
              f" | val cos {vpc:.5f} mse {vws:.4f}", flush=True)
  lpc, lws = tr.loss_and_grad()
  validate(epochs, lpc, lws)
  net.params.copy_(best_params)
  return net, hist


def iterations_to(result, e_star, tol=1e-6):
  """First CG iteration of a MinimizeResult within tol of e_star (None if
  never)."""
  return next((i for i, (t, e) in enumerate(result.history)
               if e - e_star < tol), None)


def select_by_convergence(nets, cells, tol=1e-6, max_iter=3000, log=None):
  """Model selection on validation cells: for each network, the mean over
  the cells of (CG iterations to |E - E*| < tol with the network's factor)
  / (the same with the plain preconditioner).  Returns (best index,
  scores)."""
  from .minimize import Minimizer
  per_cell = []
  for wl in cells:
    plan = EnergyPlan.from_workload(wl)
    p0 = wl.params(0)
    plain = Minimizer(plan, p0).run(max_iter=max_iter)
    runs = []
    for n, net in enumerate(nets):
      runs.append(Minimizer(plan, p0, factor_fn=LearnedFactor(plan, net),
                            restart=LEARNED_RESTART,
                            first_restart=LEARNED_FIRST_RESTART)
                  .run(max_iter=max_iter))
      if log:
        log(f"{wl.name}: candidate {n}: {runs[-1].iterations} iterations")
    e_star = min([plain.energy] + [r.energy for r in runs])
    ip = iterations_to(plain, e_star, tol)
    ratios = []
    for r in runs:
      i = iterations_to(r, e_star, tol)
      ratios.append(np.inf if i is None or ip is None else i / ip)
    per_cell.append(ratios)
    if log:
      log(f"{wl.name}: plain {ip} iterations; ratios "
          + " ".join(f"{x:.3f}" for x in ratios))
    plan.close()
  scores = np.mean(np.asarray(per_cell), axis=0)
  return int(np.argmin(scores)), scores
