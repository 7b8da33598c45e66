"""Copies the reference's example inputs (x64/Debug) into tests/golden/ as data
fixtures and records their SHA-256.  The two reference rasters
redwood_edgecorrection.tif and redwood_densityestimation.tif are byte-identical;
one copy is kept.  Run in a container where the reference tree is mounted."""
import hashlib
import os
import shutil
import sys

SRC = os.environ.get("MKDE_REFERENCE", "/root/reference") + "/x64/Debug"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {"pntsRedwood.csv": "pntsRedwood.csv", "redwood_edgecorrection.tif": "redwood_mask.tif"}

if __name__ == "__main__":
    for src, dst in FILES.items():
        p = os.path.join(SRC, src)
        if not os.path.exists(p):
            sys.exit("missing %s" % p)
        shutil.copyfile(p, os.path.join(HERE, dst))
        print(dst, hashlib.sha256(open(p, "rb").read()).hexdigest())
