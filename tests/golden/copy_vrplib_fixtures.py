"""Copy the reference's benchmark instances used as golden fixtures.

The reference ships no tests; its ``dataset/`` directory (VRPLIB ``.vrp``
instances with best-known ``.sol``/``.opt`` solutions) is the only golden
data it holds.  This script copies a small, fixed selection into
``tests/golden/vrplib/`` so the tests run where the reference is absent.
Run it from the repository root with the reference checked out at
``/root/reference`` (or pass another path).
"""
import os
import shutil
import sys

FILES = [
    "X/X-n101-k25.vrp", "X/X-n101-k25.sol",
    "X/X-n148-k46.vrp", "X/X-n148-k46.sol",
    "X/X-n242-k48.vrp", "X/X-n242-k48.sol", "X/X-n242-k48.opt",
    "X/X-n480-k70.vrp", "X/X-n480-k70.sol",
    "X/X-n1001-k43.vrp", "X/X-n1001-k43.sol",
    "CMT/CMT1.vrp", "CMT/CMT1.sol",
    "CMT/CMT6.vrp", "CMT/CMT6.sol",
    "CMT/CMT11.vrp", "CMT/CMT11.sol",
    "Golden/Golden_1.vrp", "Golden/Golden_1.sol",
    "Li/Li_21.vrp", "Li/Li_21.sol",
]


def main(ref="/root/reference"):
    here = os.path.dirname(os.path.abspath(__file__))
    dst = os.path.join(here, "vrplib")
    os.makedirs(dst, exist_ok=True)
    for rel in FILES:
        shutil.copyfile(os.path.join(ref, "dataset", rel), os.path.join(dst, os.path.basename(rel)))
    print(f"copied {len(FILES)} files to {dst}")


if __name__ == "__main__":
    main(*sys.argv[1:])
