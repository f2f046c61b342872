#!/usr/bin/env python3
"""Write tests/golden/ref_vectors.json from the reference's own code.

Runs oracle/_ref/ref_harness, which is the reference's gll_mod,
interpolation_mod, coordinates_mod, coordinate_utilities_mod,
source_signature_mod and models/ak135_f_mod compiled unmodified from
/root/reference with the reference's own test doubles (real_kind = REAL32),
and models/prem_mod in its STAND_ALONE mode (REAL64). Build it first with
`make -C oracle ref`. The JSON holds only inputs and the harness's outputs.

The Earth-model records are taken at the GLL radii of MODEL_GRIDS (computed
as the Fortran host computes them) and at the PREM region boundaries.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
sys.path.insert(0, os.path.join(ROOT, "ses3d-nt_amd"))
from ses3d_amd import synth  # noqa: E402

# (nz, lpd, rad_max, rad_min) [m]: a crust/upper-mantle box, the whole
# mantle and core, and a box whose top and bottom are PREM discontinuities
MODEL_GRIDS = [(16, 4, 6371000.0, 5371000.0), (10, 3, 6371000.0, 1000000.0),
               (4, 2, 6368000.0, 5701000.0)]
PREM_BOUNDARIES = [0.0, 1221500.0, 3480000.0, 3630000.0, 5600000.0, 5701000.0,
                   5771000.0, 5971000.0, 6151000.0, 6291000.0, 6346600.0,
                   6356000.0, 6368000.0, 6371000.0]


def model_radii():
    rs = set(PREM_BOUNDARIES)
    for nz, p, top, bottom in MODEL_GRIDS:
        rs.update(float(v) for v in synth.coords_1d(p, nz, top, bottom).ravel())
    return sorted(rs)


def commands():
    cmds = []
    for p in range(1, 8):
        cmds.append(("gll", "gll %d" % p, {"p": p}))
    # theta of lat -5..5 (12 elements, p=4); phi of 0.25..0.9 rad (7, p=3);
    # r of 6000..5000 m (5, p=6) and 6371..4971 km (96, p=4), surface first
    for elms, p, lo, hi in [(12, 4, 1.4835298641951802, 1.6580627893946132),
                            (7, 3, 0.25, 0.9), (5, 6, 6000.0, 5000.0),
                            (96, 4, 6371000.0, 4971000.0)]:
        cmds.append(("coords", "coords %d %d %r %r" % (elms, p, lo, hi),
                     {"elms": elms, "p": p, "lo": lo, "hi": hi}))
    for p in range(2, 8):
        for x in (-0.83, -0.3, 0.0, 0.123, 0.77, 1.0):
            cmds.append(("lagrange", "lagrange %d %r" % (p, x), {"p": p, "x": x}))
    for p, x, y, z in [(4, 0.1, -0.6, 0.35), (6, -0.9, 0.2, -1.0)]:
        cmds.append(("vandermonde", "vandermonde %d %r %r %r" % (p, x, y, z),
                     {"p": p, "x": x, "y": y, "z": z}))
    for name in ("RICKER", "GAUSS", "SIN**3", "BRUNE", "DELTA"):
        for t in (0.0, 10.5, 31.25, 40.0, 47.5, 60.0, 95.0):
            cmds.append(("wavelet", "wavelet %s %r 40.0 20.0" % (name, t),
                         {"name": name, "t": t, "onset": 40.0, "width": 20.0}))
    for lat in (-20.0, 5.0, 67.0):
        cmds.append(("colat", "colat %r" % lat, {"lat": lat}))
    for r in model_radii():
        for nocrust, noocean in ((False, True), (False, False), (True, True)):
            cmds.append(("prem", "prem %r %s %s" % (r, "TF"[not nocrust], "TF"[not noocean]),
                         {"r": r, "nocrust": nocrust, "noocean": noocean}))
        cmds.append(("premti", "premti %r" % r, {"r": r}))
        if r > 0.0:
            cmds.append(("ak135", "ak135 %r" % (r / 1000.0), {"x": r / 1000.0}))
    return cmds


def main():
    cmds = commands()
    text = "\n".join(c[1] for c in cmds) + "\nend\n"
    out = subprocess.run([HARNESS], input=text, capture_output=True, text=True, check=True).stdout
    lines = out.split("\n")
    pos = 0
    records = []
    for kind, _, args in cmds:
        if kind == "gll":
            vals = {}
            for key, n in (("knots", args["p"] + 1), ("weights", args["p"] + 1),
                           ("dl", (args["p"] + 1) ** 2)):
                assert lines[pos].strip() == key, (lines[pos], key)
                vals[key] = [float(v) for v in lines[pos + 1:pos + 1 + n]]
                pos += 1 + n
            records.append({"kind": kind, "args": args, "out": vals})
            continue
        n = {"coords": args.get("elms", 0) * (args.get("p", 0) + 1),
             "lagrange": args.get("p", 0) + 1,
             "vandermonde": (args.get("p", 0) + 1) ** 3,
             "wavelet": 1, "colat": 1, "prem": 3, "premti": 6, "ak135": 3}[kind]
        assert lines[pos].strip() == kind, (lines[pos], kind)
        vals = [float(v) for v in lines[pos + 1:pos + 1 + n]]
        pos += 1 + n
        records.append({"kind": kind, "args": args, "out": vals})
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump({"source": "reference src/gll_mod.f90, src/interpolation_mod.f90, "
                             "src/coordinates_mod.f90, src/misc/coordinate_utilities_mod.f90, "
                             "src/misc/source_signature_mod.f90, "
                             "src/models/ak135_f_mod.F90 built with "
                             "src/examples/dummy_parameters.f90 (REAL32); "
                             "src/models/prem_mod.F90 built with -DSTAND_ALONE (REAL64)",
                   "records": records}, f, indent=0)
    print("wrote %d records" % len(records))


if __name__ == "__main__":
    main()
