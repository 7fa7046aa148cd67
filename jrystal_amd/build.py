"""Build libjrx_hip.so in-tree for gfx950 (hipcc cross-compiles without a
GPU).  Usage: python -m jrystal_amd.build [--clean]"""
import os
import subprocess
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(clean: bool = False, jobs: int = 4, force: bool = True) -> str:
  """Compile every source (force=True re-runs hipcc even when the library
  is up to date)."""
  if clean:
    subprocess.run(["make", "-C", CSRC, "clean"], check=True)
  cmd = ["make", "-C", CSRC, f"-j{jobs}"] + (["-B"] if force else [])
  subprocess.run(cmd, check=True)
  out = os.path.join(_HERE, "libjrx_hip.so")
  if not os.path.exists(out):
    raise RuntimeError("build finished without producing libjrx_hip.so")
  return out


if __name__ == "__main__":
  print(build(clean="--clean" in sys.argv))
