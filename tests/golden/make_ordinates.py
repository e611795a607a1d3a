"""Extract the ordinate tables that the reference's run logs print.

The reference prints its discrete-ordinates set to ``lb.out`` with 16 decimals
(``src/angular/AngularSpace.f90:52-118``).  Three such tables pin the product
quadrature (ascending Gauss-Legendre polar cosines, polar index fastest,
Chebyshev azimuths ``(2j-1) pi / Nc``):

* L = 7 (128 directions):  ``examples/three_beam_water_tank/lb.out:1296-1426``
* L = 11 (288 directions): ``validation/photons_incident/water_tank_1MeV/lb.out:2714-3004``
* slab L = 15 (16 cosines): ``validation/electrons_incident/Al_521keV_0deg/lb.out:1819-1837``

Run from the repository root with the reference checkout at /root/reference;
writes ``tests/golden/ordinates.json`` (data only).
"""
import json
import os
import re

REF = "/root/reference"
SOURCES = [
    ("L7_3d", "examples/three_beam_water_tank/lb.out", 1296, 7),
    ("L11_3d", "validation/photons_incident/water_tank_1MeV/lb.out", 2714, 11),
    ("L15_slab", "validation/electrons_incident/Al_521keV_0deg/lb.out", 1819, 15),
]
NUM = r"(-?\d+\.\d+)"


def parse(path, start):
    with open(path, encoding="utf-8") as fh:
        lines = fh.read().split("\n")
    rows = []
    for line in lines[start + 2:]:        # skip the header and the rule line
        m3 = re.match(r"\s*(\d+) \)\s+\(\s*" + NUM + r",\s*" + NUM + r",\s*" + NUM + r"\s*\)", line)
        m1 = re.match(r"\s*(\d+) \)\s+" + NUM + r"\s*$", line)
        if m3:
            rows.append([float(m3.group(i)) for i in (2, 3, 4)])
        elif m1:
            rows.append([float(m1.group(2))])
        else:
            break
    return rows


def main():
    out = {}
    for name, rel, line, L in SOURCES:
        rows = parse(os.path.join(REF, rel), line)
        out[name] = {"L": L, "source": f"{rel}:{line}", "k": rows}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ordinates.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=0)
    for name, v in out.items():
        print(name, len(v["k"]))


if __name__ == "__main__":
    main()
