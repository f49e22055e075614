"""Writes tests/golden/readme_c1.json from the console log of the one
published run in the reference (README.md:99-115): 20x20x20 Laplacian,
interval [0.6, 1.2] of [0, 12], 5 slices.  Records the slice boundaries and,
per slice, the eigenvalue count found (and the exact count printed beside
it), the filter degree and the Lanczos iteration count.

Run in a checkout that has the reference at /root/reference; the fixture
itself is plain data and is what the tests read.
"""
import json
import os
import re
import sys

SRC = "/root/reference/README.md"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "readme_c1.json")


def main():
    lines = open(SRC).read().splitlines()
    sl = []
    for ln in lines:
        m = re.match(r"\s*Slice (\d+): \[([-0-9.e+]+), ([-0-9.e+]+)\]", ln)
        if m:
            sl.append((int(m.group(1)), float(m.group(2)), float(m.group(3))))
    sli = [sl[0][1]] + [b for (_, _, b) in sl]
    groups = []
    for ln in lines:
        m = re.match(r"\s*Group\s+(\d+) \(NP=(\d+)\): \[.*\], eval computed (\d+) \((\d+)\), "
                     r"polydeg\s+(\d+), its (\d+)", ln)
        if m:
            groups.append({"slice": int(m.group(1)), "np": int(m.group(2)),
                           "found": int(m.group(3)), "exact": int(m.group(4)),
                           "deg": int(m.group(5)), "its": int(m.group(6))})
    m = [re.search(r"Eigenvalue count ([0-9.]+)", ln) for ln in lines]
    ecount = [float(x.group(1)) for x in m if x][0]
    out = {"source": "README.md:99-115", "nx": 20, "ny": 20, "nz": 20,
           "a": 0.6, "b": 1.2, "lmin": 0.0, "lmax": 12.0, "nslices": 5,
           "ecount_kpm_10ranks": ecount, "sli": sli, "slices": groups}
    json.dump(out, open(OUT, "w"), indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    sys.exit(main())
