"""Write the basis-set data files this package ships (H, C, N, O only).

The data are the Basis Set Exchange JSON records (schema "complete") for
cc-pVDZ, def2-SVP, cc-pVTZ and aug-cc-pVTZ and the fitting sets cc-pVDZ-RIFIT and
def2-SVP-RIFIT.  This script copies the
``electron_shells`` entries of the four elements verbatim (exponents and
coefficients stay strings, exactly as published) from a directory of BSE JSON
files, so that the loader sees the same text the Basis Set Exchange serves.

    python make_subset.py /root/reference/share/basis_sets

The reference reads the same records through
``src/shared/data/bse_json_parser.cpp:27-87``.
"""
import json
import os
import sys

ELEMENTS = ("1", "6", "7", "8")
NAMES = ("cc-pvdz", "def2-svp", "cc-pvtz", "aug-cc-pvtz", "cc-pvdz-rifit", "def2-svp-rifit")


def main(src_dir):
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for name in NAMES:
        with open(os.path.join(src_dir, name + ".json")) as f:
            full = json.load(f)
        sub = {
            "molssi_bse_schema": full["molssi_bse_schema"],
            "name": full["name"],
            "version": full.get("version", ""),
            "elements": {
                z: {"electron_shells": full["elements"][z]["electron_shells"]}
                for z in ELEMENTS
            },
        }
        with open(os.path.join(out_dir, name + ".json"), "w") as f:
            json.dump(sub, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/share/basis_sets")
