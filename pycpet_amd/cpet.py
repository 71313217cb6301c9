"""Command line `python -m pycpet_amd.cpet -o options.json [-i in] [-p out]
[-m method] [-d]`, with the reference CLI's flags and override rules
(CPET/source/scripts/cpet.py:10-94)."""
from __future__ import annotations

import argparse
import json
import logging
import os

from .CPET import CPET


def main(argv=None):
    parser = argparse.ArgumentParser(
        description="CPET hot paths on MI355X: electric fields, potentials and field topology")
    parser.add_argument("-d", "--debug", action="store_const", dest="loglevel",
                        const=logging.DEBUG, default=logging.INFO,
                        help="Print lots of debugging statements")
    parser.add_argument("-o", type=str, default="./options/options.json", help="Options for CPET")
    parser.add_argument("-i", type=str, default=None, help="Input path (overrides options)")
    parser.add_argument("-p", type=str, default=None, help="Output path (overrides options)")
    parser.add_argument("-m", type=str, default=None, help="Method (overrides options)")
    args = parser.parse_args(argv)
    logging.basicConfig(level=args.loglevel,
                        format="%(asctime)s %(levelname)s [%(name)s] %(message)s",
                        datefmt="%H:%M:%S")
    if not os.path.exists(args.o):
        raise FileNotFoundError(f"Options File {args.o} not found!")
    with open(args.o, "r") as fh:
        options = json.load(fh)
    if args.i:
        options["inputpath"] = args.i
    if args.p:
        options["outputpath"] = args.p
    if args.m:
        options["CPET_method"] = args.m
    CPET(options).run()


if __name__ == "__main__":
    main()
