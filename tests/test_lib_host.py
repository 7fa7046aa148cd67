"""The C ABI library loads, exports every symbol include/jrx.h declares, and
rejects bad arguments on the host (no GPU call is made by these tests)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jrx.h")


def _declared():
  text = open(HEADER).read()
  text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
  return sorted(set(re.findall(r"\b(jrx_[a-z0-9_]+)\s*\(", text)))


def _lib():
  from jrystal_amd import _lib
  if not os.path.exists(_lib.LIB_PATH):
    pytest.skip("libjrx_hip.so not built (run __graft_entry__.build())")
  return _lib, _lib.lib()


def test_header_symbols_exported():
  _, L = _lib()
  names = _declared()
  assert "jrx_energy_grad" in names and "jrx_ifftn3d" in names
  for n in names:
    assert hasattr(L, n), n


def test_ctypes_signatures_cover_header():
  from jrystal_amd import _lib
  text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
  for name in _declared():
    m = re.search(name + r"\s*\(([^)]*)\)", text)
    params = [p for p in m.group(1).split(",") if p.strip() not in ("", "void")]
    assert name in _lib.SIGNATURES, name
    assert len(_lib.SIGNATURES[name]) == len(params), name


def test_version_and_error_mapping():
  lm, L = _lib()
  assert b"gfx950" in L.jrx_version()
  rc = L.jrx_plan_create(None, None)
  assert rc == lm.JRX_ERR_INVALID
  with pytest.raises(ValueError):
    lm.check(rc)
  with pytest.raises(NotImplementedError):
    lm.check(lm.JRX_ERR_UNSUPPORTED)


def test_invalid_problem_rejected_before_device_use():
  lm, L = _lib()
  d = lm.JrxProblem()
  d.ns, d.nk, d.nb, d.nx, d.ny, d.nz = 3, 1, 1, 4, 4, 4
  h = ctypes.c_void_p()
  rc = L.jrx_plan_create(ctypes.byref(h), ctypes.byref(d))
  assert rc == lm.JRX_ERR_INVALID
  assert b"ns" in L.jrx_last_error()
  d.ns = 1
  rc = L.jrx_plan_create(ctypes.byref(h), ctypes.byref(d))  # null arrays
  assert rc == lm.JRX_ERR_INVALID


def test_fft_rejects_non_smooth_size_on_host():
  lm, L = _lib()
  rc = L.jrx_fftn3d(None, None, 1, 11, 8, 8, ctypes.c_double(1.0), None)
  assert rc == lm.JRX_ERR_INVALID
  rc = L.jrx_fftn3d(None, None, 1, 8, 8, 8, ctypes.c_double(1.0), None)
  assert rc == lm.JRX_ERR_INVALID  # null buffers


def test_bench_rejects_multi_gpu():
  r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                      "--gpus", "2"], capture_output=True, text=True,
                     timeout=120)
  assert r.returncode == 2
  assert "one GPU" in r.stderr


def test_gpu_entry_points_fail_loudly_without_gpu():
  torch = pytest.importorskip("torch")
  if torch.cuda.is_available():
    pytest.skip("a GPU is present")
  from jrystal_amd import workloads
  from jrystal_amd.objective import EnergyPlan
  with pytest.raises(RuntimeError):
    EnergyPlan.from_workload(workloads.tiny())
