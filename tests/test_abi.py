"""The C ABI library: it loads and exports every entry point the headers in
include/ declare. No compute calls here (no GPU); the host-only helpers
(GLL tables, wavelets) are checked against the reference vectors."""
import ctypes
import glob
import json
import os
import re

import numpy as np
import pytest

from ses3d_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(s3d_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(api.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(api.EXPORTS) == names


def test_gll_helper_matches_reference_vectors():
    vec = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_vectors.json")))["records"]
    for r in vec:
        if r["kind"] != "gll" or r["args"]["p"] < 2:
            continue
        p = r["args"]["p"]
        k, w, d = api.gll(p)
        assert np.max(np.abs(k - r["out"]["knots"])) < 8 * 2 ** -23
        assert np.max(np.abs(w - r["out"]["weights"])) < 8 * 2 ** -23
        ref = np.array(r["out"]["dl"]).reshape(p + 1, p + 1)
        assert np.max(np.abs(d - ref)) < 64 * 2 ** -23 * np.abs(ref).max()


def test_wavelet_helper_matches_reference_vectors():
    vec = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_vectors.json")))["records"]
    for r in vec:
        if r["kind"] != "wavelet":
            continue
        a = r["args"]
        i = int(round(a["t"] / 0.25))
        s = api.wavelet(a["name"], a["onset"], a["width"], 0.25, i + 1)[i]
        assert abs(s - r["out"][0]) < 16 * 2 ** -23


def test_create_without_device_fails_loudly():
    if api.device_count() > 0:
        pytest.skip("a device is present")
    from ses3d_amd import synth
    g = synth.homogeneous().grid
    with pytest.raises(api.S3dError, match="no HIP device"):
        api.Simulation(g)


def test_bad_arguments_are_rejected():
    from ses3d_amd import synth
    g = synth.homogeneous().grid
    bad = synth.Grid(**{**g.__dict__, "lpd": 9})
    with pytest.raises(api.S3dError):
        api.Simulation(bad)
    # the kernels index unique nodes in 32 bits: 1000^3 p=4 blocks hold 6.4e10
    huge = synth.Grid(**{**g.__dict__, "nx": 1000, "ny": 1000, "nz": 1000})
    with pytest.raises(api.S3dError, match="mesh too large"):
        api.Simulation(huge)


def test_fortran_binding_declares_every_entry_point():
    """The ISO_C_BINDING module a Fortran host uses binds every entry point of
    the C header, under the header's name."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "ses3d_hip.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(s3d_\w+)\(", hdr, re.M))
    f90 = open(os.path.join(root, "ses3d-nt_amd", "fortran", "s3d_c_binding.f90")).read()
    bound = set(re.findall(r"BIND\(C, NAME='(s3d_\w+)'\)", f90))
    assert declared and declared == bound, (declared - bound, bound - declared)
