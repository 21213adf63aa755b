"""Collect the reference's own test data into tests/golden/.

Run in a container that has the reference checkout (default
/root/reference).  It copies data only -- the regression decks, their
expected xyz frames and logs -- and extracts the 1000-entry uniform table
that RunType 'T' cycles through (mulskips-source/mcevent/PickTreeEvent.f:
77-1076) as plain numbers.  The values are default-REAL literals in the
reference, so they are stored here rounded to single precision and widened
back to double, exactly as the Fortran program sees them.
"""
import gzip
import os
import re
import shutil
import struct
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = ["cube", "cube-zigzag", "sphere", "surface", "inverted-pyramid",
         "finfet", "Si-SL", "CVD-Si-SiH2"]


def numberslist():
    src = open(os.path.join(REF, "mulskips-source/mcevent/PickTreeEvent.f")).read()
    body = src[src.index("numbers = (/"):src.index("/)", src.index("numbers = (/"))]
    vals = re.findall(r"[0-9]+\.[0-9]+E[+-][0-9]+", body)
    assert len(vals) == 1000, len(vals)
    out = []
    for v in vals:
        f32 = struct.unpack("f", struct.pack("f", float(v)))[0]
        out.append(repr(f32))
    with open(os.path.join(HERE, "numberslist.txt"), "w") as f:
        f.write("\n".join(out) + "\n")


def regression():
    src = os.path.join(REF, "regression-tests/input-and-references-files")
    dst = os.path.join(HERE, "regression")
    os.makedirs(dst, exist_ok=True)
    for d in DECKS:
        shutil.copy(os.path.join(src, f"start-{d}.dat"), dst)
        shutil.copy(os.path.join(src, f"log-{d}.txt"), dst)
        for fr in ("I00000000", "I00000001"):
            with open(os.path.join(src, f"{fr}-{d}.xyz"), "rb") as fi, \
                    gzip.open(os.path.join(dst, f"{fr}-{d}.xyz.gz"), "wb", 9) as fo:
                fo.write(fi.read())


if __name__ == "__main__":
    numberslist()
    regression()
