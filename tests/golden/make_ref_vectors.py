#!/usr/bin/env python3
"""Write tests/golden/ref_vectors.json from the reference's own code.

Runs oracle/_ref/ref_harness, which is the reference's gll_mod,
interpolation_mod, coordinates_mod, coordinate_utilities_mod and
source_signature_mod compiled unmodified from /root/reference with the
reference's own test doubles (real_kind = REAL32). Build it first with
`make -C oracle ref`. The JSON holds only inputs and the harness's outputs.
"""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


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
             "wavelet": 1, "colat": 1}[kind]
        assert lines[pos].strip() == kind, (lines[pos], kind)
        vals = [float(v) for v in lines[pos + 1:pos + 1 + n]]
        pos += 1 + n
        records.append({"kind": kind, "args": args, "out": vals})
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump({"source": "reference src/gll_mod.f90, src/interpolation_mod.f90, "
                             "src/coordinates_mod.f90, src/misc/coordinate_utilities_mod.f90, "
                             "src/misc/source_signature_mod.f90 built with "
                             "src/examples/dummy_parameters.f90 (REAL32)",
                   "records": records}, f, indent=0)
    print("wrote %d records" % len(records))


if __name__ == "__main__":
    main()
