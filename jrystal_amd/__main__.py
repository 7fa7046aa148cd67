"""Command line: python -m jrystal_amd -m energy -c config.yaml  (main.py)."""
import argparse
import logging
import sys

from . import calc, config


def main(argv=None):
  parser = argparse.ArgumentParser(prog="jrystal_amd",
                                   description="Command for jrystal_amd.")
  parser.add_argument("-m", "--mode", choices=["energy", "band"],
                      default="energy")
  parser.add_argument("-c", "--config", default="config.yaml")
  parser.add_argument("-l", "--load",
                      help="energy mode: resume from a checkpoint .npz that "
                      "an earlier run wrote under its save_dir")
  args = parser.parse_args(argv)
  cfg = config.get_config(args.config)
  logging.basicConfig(level=logging.INFO if cfg.verbose else logging.WARNING,
                      format="%(message)s")
  if args.mode == "energy":
    if cfg.use_pseudopotential:
      calc.energy_normcons(cfg, resume=args.load)
    else:
      calc.energy_all_electrons(cfg, resume=args.load)
  else:
    if cfg.use_pseudopotential:
      calc.band_normcons(cfg)
    else:
      calc.band_all_electrons(cfg)
  return 0


if __name__ == "__main__":
  sys.exit(main())
