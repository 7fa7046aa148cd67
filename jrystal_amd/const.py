"""Constants (jrystal/_src/const.py:17-18,265-456)."""
BOHR2ANGSTROM = 0.529177249
ANGSTROM2BOHR = 1.8897259886
EV2HARTREE = 0.0367492929
HARTREE2EV = 27.211407953


def _smooth_sizes(limit=2048):
  out = []
  for n in range(1, limit + 1):
    m = n
    for p in (2, 3, 5, 7):
      while m % p == 0:
        m //= p
    if m == 1:
      out.append(n)
  return out


# The reference's CUFFT_FACTORS table: every 7-smooth integer up to 2048
# (the table is that set exactly); the HIP FFT handles these radices.
FFT_SIZES = tuple(_smooth_sizes())
