"""Write tests/golden/b1_layout.json: the byte layout (sizeof and offsetof)
of the reference's GPU-backend structs compiled from the reference's own
headers at every grid size (oracle/ref_layout.c, built by oracle/ref.mk into
oracle/_ref/).  Runs only where the reference sources exist.

Usage:  python tests/golden/make_b1_layout.py
"""
from __future__ import annotations

import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
SIZES = (64, 128, 256, 512)


def main():
    subprocess.check_call(["make", "-C", os.path.join(REPO, "oracle"), "ref"])
    out = {}
    for n in SIZES:
        exe = os.path.join(REPO, "oracle", "_ref", f"ref_layout_n{n}")
        rows = {}
        for line in subprocess.check_output([exe], text=True).splitlines():
            t, f, off = line.split()
            rows.setdefault(t, {})["" if f == "-" else f] = int(off)
        out[str(n)] = rows
    with open(os.path.join(HERE, "b1_layout.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
