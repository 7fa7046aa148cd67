"""Host-side parts of bench.py: the algorithmic-bytes model of each kernel
class, time-to-tolerance, host facts, and the one-GPU guard."""
import os
import subprocess
import sys
from types import SimpleNamespace

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _plan(keep=True):
  # the Si8 cube: 64 k, 32 bands, G_n = 17071, 64^3, 31 active planes
  return SimpleNamespace(ns=1, nk=64, nb=32, num_g=17071,
                         grid_sizes=(64, 64, 64), active_planes=31,
                         keep_planes=keep)


def test_algorithmic_bytes_match_the_byte_model():
  p = _plan()
  S = 64 * 32
  T = S * 31 * 64 * 64 * 16
  Q = S * 17071 * 16
  W = 2 * S * 17071 * 8
  rho = 64 ** 3 * 8
  assert bench.algorithmic_bytes("xline_vpsi", p, 1) == 2 * T + rho
  assert bench.algorithmic_bytes("xline_density", p, 1) == T + 2 * rho
  assert bench.algorithmic_bytes("plane_ifft", p, 1) == Q + T
  assert bench.algorithmic_bytes("plane_ifft", _plan(False), 1) == 2 * (Q + T)
  assert bench.algorithmic_bytes("plane_fft_gather", p, 1) == T + 2 * Q
  assert bench.algorithmic_bytes("qr_apply", p, 1) == W + Q
  assert bench.algorithmic_bytes("qr_bwd_apply", p, 1) == 2 * Q + W
  assert bench.algorithmic_bytes("potential", p, 1) is None
  # the benchmark's dominant launch, as bench.py reports it
  assert bench.algorithmic_bytes("xline_vpsi", p, 1) == 8323596288


def test_time_to_tolerance():
  hist = [(0.0, 10.0), (1.0, 1.0), (2.0, 1e-7), (3.0, 0.0)]
  assert bench.time_to(hist, 0.0, 1e-6) == (2.0, 2)
  assert bench.time_to(hist, -1.0, 1e-6) == (None, None)


def test_cpu_host_facts():
  h = bench.cpu_host()
  assert h["nproc"] == os.cpu_count()
  assert "cpu_model" in h


def test_more_than_one_gpu_is_refused():
  r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                      "--gpus", "2"], capture_output=True, text=True,
                     timeout=120)
  assert r.returncode == 2
  assert "one GPU" in r.stderr
