#!/usr/bin/env python3
"""Write tests/golden/model_known_answers.json: the known answers of the
reference's own model tests, src/examples/ak135_f_test.f90 (values at six
radii, compared with ==) and src/examples/prem_test.f90 (intervals
[v_1, v_2] at eight radii), read as data from /root/reference. Only the radii
and expected values are kept."""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/examples"
NUM = r"(-?[0-9.]+)(?:_real_kind)?"


def blocks(path, start_key):
    """Split a test program into blocks that begin with `start_key = value`."""
    out, cur = [], None
    for line in open(path):
        line = line.split("!")[0].strip()
        m = re.match(r"(\w+)\s*=\s*" + NUM + r"\s*$", line)
        if not m:
            continue
        k, v = m.group(1), float(m.group(2))
        if k == start_key:
            cur = {start_key: v}
            out.append(cur)
        elif cur is not None:
            cur[k] = v
    return out


def main():
    ak = [{"x_km": b["x"], "rho": b["rho_expect"], "vp": b["vp_expect"], "vs": b["vs_expect"]}
          for b in blocks(os.path.join(SRC, "ak135_f_test.f90"), "x")]
    prem = []
    for b in blocks(os.path.join(SRC, "prem_test.f90"), "r"):
        rec = {"r_m": b["r"]}
        for q in ("rho", "vpv", "vph", "vsv", "vsh", "eta"):
            if q + "_1" in b:
                rec[q] = sorted([b[q + "_1"], b[q + "_2"]])
            elif q + "_" in b:
                rec[q] = [b[q + "_"], b[q + "_"]]
        prem.append(rec)
    with open(os.path.join(HERE, "model_known_answers.json"), "w") as f:
        json.dump({"source": "reference src/examples/ak135_f_test.f90 and src/examples/prem_test.f90",
                   "ak135_f": ak, "prem": prem}, f, indent=1)
    print(len(ak), len(prem))


if __name__ == "__main__":
    main()
