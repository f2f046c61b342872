"""The FP64 oracle against the reference's own code.

tests/golden/ref_vectors.json holds outputs of the reference's gll_mod,
interp_mod, coordinates_mod, coordinate_utilities_mod and source_signature_mod
compiled unmodified with the reference's test doubles (REAL32; written by
tests/golden/make_ref_vectors.py from oracle/_ref/ref_harness). The oracle
computes in FP64, so agreement is to single precision: a few REAL32 ulps.
When oracle/_ref/ref_harness is built, it is also run live here.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from ses3d_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "ref_vectors.json")))["records"]
F32 = 2.0 ** -23


def close32(a, b, scale=None, ulps=8):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.max(np.abs(b)) if scale is None else scale
    return np.max(np.abs(a - b)) <= ulps * F32 * max(s, 1e-30)


def recs(kind):
    return [r for r in VEC if r["kind"] == kind]


@pytest.mark.parametrize("p", range(1, 8))
def test_gll_tables_match_reference(p):
    r = [x for x in recs("gll") if x["args"]["p"] == p][0]["out"]
    k, w, d = oracle.gll(p)
    assert close32(k, r["knots"], scale=1.0)
    assert close32(w, r["weights"], scale=1.0)
    # reference prints dl(i, j) with i fastest; oracle.gll returns D[j, i]
    ref = np.array(r["dl"]).reshape(p + 1, p + 1)
    assert close32(d, ref, ulps=64)


def test_lagrange_matches_reference():
    for r in recs("lagrange"):
        p, x = r["args"]["p"], r["args"]["x"]
        assert close32(oracle.lagrange(p, x), r["out"], scale=1.0, ulps=16), (p, x)


def _grid_for(elms, p, lo, hi):
    """An oracle whose theta, phi or r axis is coordinates_1d(elms, p, lo, hi)."""
    if hi < lo and lo > 1000.0:   # r axis, surface first
        g = synth.Grid(2, 2, elms, p, -1.0, 1.0, -1.0, 1.0, hi, lo, 0, 1.0, 0.1, 1)
        return g, 2
    if lo > 1.0:                  # theta: colatitude lo..hi
        lat_max = 90.0 - math.degrees(lo)
        lat_min = 90.0 - math.degrees(hi)
        return synth.Grid(elms, 2, 2, p, lat_min, lat_max, -1.0, 1.0, 6000e3, 6300e3, 0, 1.0, 0.1, 1), 0
    g = synth.Grid(2, elms, 2, p, -1.0, 1.0, math.degrees(lo), math.degrees(hi), 6000e3, 6300e3, 0, 1.0, 0.1, 1)
    return g, 1


def test_coordinates_match_reference():
    for r in recs("coords"):
        a = r["args"]
        g, axis = _grid_for(a["elms"], a["p"], a["lo"], a["hi"])
        o = oracle.Oracle(g)
        P1 = a["p"] + 1
        o.set_model(oracle.RHOINV, np.full(o.shape, 1 / 3000.0))
        o.set_model(oracle.LAMBDA, np.full(o.shape, 3e10))
        o.set_model(oracle.MU, np.full(o.shape, 3e10))
        o.init()
        c = o.coords_1d()[axis]               # (P1, elms), elm fastest
        ref = np.array(r["out"]).reshape(P1, a["elms"])
        assert close32(c, ref, ulps=4), a
        # seamless element boundaries (reference src/examples/coordinates_test.f90:29-65)
        assert np.array_equal(c[-1, :-1], c[0, 1:])


def test_vandermonde_matches_reference():
    for r in recs("vandermonde"):
        a = r["args"]
        p = a["p"]
        lx, ly, lz = (oracle.lagrange(p, a[q]) for q in "xyz")
        vm = lz[:, None, None] * ly[None, :, None] * lx[None, None, :]   # [k, j, i]
        assert close32(vm.ravel(), r["out"], scale=1.0, ulps=16)


@pytest.mark.parametrize("name", ["RICKER", "GAUSS", "SIN**3", "BRUNE", "DELTA"])
def test_wavelets_match_reference(name):
    for r in recs("wavelet"):
        a = r["args"]
        if a["name"] != name:
            continue
        # t = (i-1) dt with dt = 0.25 hits every sample time used
        i = int(round(a["t"] / 0.25))
        s = oracle.stf(name, a["onset"], a["width"], 0.25, i + 1)[i]
        assert close32(s, r["out"][0], scale=1.0, ulps=16), (name, a["t"], s, r["out"][0])


def test_colatitude_conversion_matches_reference():
    for r in recs("colat"):
        lat = r["args"]["lat"]
        g = synth.Grid(2, 2, 2, 2, lat, lat + 1.0, 0.0, 1.0, 6000e3, 6300e3, 0, 1.0, 0.1, 1)
        o = oracle.Oracle(g)
        o.set_model(oracle.RHOINV, np.full(o.shape, 1 / 3000.0))
        o.set_model(oracle.LAMBDA, np.full(o.shape, 3e10))
        o.set_model(oracle.MU, np.full(o.shape, 3e10))
        o.init()
        th = o.coords_1d()[0]
        # theta at the southern edge (largest colatitude) is deg2rad(90 - lat)
        assert close32(th.max(), r["out"][0], scale=1.0, ulps=2)


HARNESS = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_harness")


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="partial reference build absent")
def test_live_reference_harness_reproduces_fixture():
    r = recs("gll")[3]
    out = subprocess.run([HARNESS], input="gll 4\nend\n", capture_output=True,
                         text=True, check=True).stdout.split()
    assert [float(x) for x in out[1:6]] == r["out"]["knots"]
