"""Built-in 1-D Earth models (model_type 1, 2/15, 3/13, 11, 12, 14, 21) of
the Fortran host against the reference's own prem_mod (STAND_ALONE, REAL64)
and ak135_f_mod (REAL32) through tests/golden/ref_vectors.json.

The host writes the model with workflow 'model' (raw REAL32 files in the
reference's global layout), so every check reads those files: PREM agrees
to the last REAL32 bit of the reference's FP64 values, ak135-f to the
reference's single-precision interpolation.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from helpers import write_config
from ses3d_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ses3d-nt_amd", "bin", "ses3d_nt")
HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "ref_vectors.json")))["records"]

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")

# the grids the fixture's radii were taken on: (nz, lpd, rad_max, rad_min)
MODEL_GRIDS = [(16, 4, 6371000.0, 5371000.0), (10, 3, 6371000.0, 1000000.0),
               (4, 2, 6368000.0, 5701000.0)]
PARAMS = ("rhoinv", "lambda", "mu", "a", "b", "c")


def _index():
    prem, premti, ak = {}, {}, {}
    for r in VEC:
        a = r["args"]
        if r["kind"] == "prem":
            prem[(a["r"], a["nocrust"], a["noocean"])] = r["out"]
        elif r["kind"] == "premti":
            premti[a["r"]] = r["out"]
        elif r["kind"] == "ak135":
            ak[a["x"]] = r["out"]
    return prem, premti, ak


PREM, PREMTI, AK = _index()


def iso(rho, vp, vs):
    """Moduli of src/model_paras.F90:262-273 in FP64."""
    rho, vp, vs = 1000.0 * rho, 1000.0 * vp, 1000.0 * vs
    mu = rho * vs ** 2
    return [1.0 / rho, rho * vp ** 2 - 2.0 * mu, mu, 0.0, 0.0, 0.0]


def expected(model_type, r, ez):
    """rhoinv, lambda, mu, a, b, c of the reference's make_model_parameters
    (src/model_paras.F90:219-336) from the reference's own model values."""
    if model_type == 1:
        return [1.0 / 3000.0, float(np.float32(1.1e11)), float(np.float32(7.0e10)), 0.0, 0.0, 0.0]
    if model_type == 11:
        rho, vpv, vph, vsv, vsh, eta = PREMTI[r]
        rho, vph, vsv, vsh = 1000.0 * rho, 1000.0 * vph, 1000.0 * vsv, 1000.0 * vsh
        mu = rho * vsh ** 2
        lam = rho * vph ** 2 - 2.0 * mu
        a = rho * vph ** 2 - lam - 2.0 * mu
        return [1.0 / rho, lam, mu, a, rho * vsv ** 2 - mu, eta * (lam + a) - lam]
    if model_type == 21:
        return iso(*AK[r / 1000.0])
    flags = {3: (False, True), 13: (False, True), 12: (False, False), 14: (True, True),
             2: (True, True), 15: (True, True)}[model_type]
    rho, vp, vs = PREM[(r, flags[0], flags[1])]
    if model_type in (2, 15) and ez == 0:
        rho, vp, vs = (float(np.float32(v)) for v in (3.00, 7.20, 4.10))
    return iso(rho, vp, vs)


def run_model_workflow(tmp_path, model_type, nz, lpd, top, bottom, params=PARAMS, extra=None):
    case = synth.homogeneous(nt=10, nx=4, ny=4, nz=nz, lpd=lpd, n_rec=1)
    case.grid.rad_max, case.grid.rad_min = top, bottom
    case.grid.taper_width = 1
    model = {"model_type": model_type}
    for k in params:
        model[k] = os.path.join(str(tmp_path), k + ".out")
    model.update(extra or {})
    cfg = write_config(case, str(tmp_path), model=model, workflow="model")
    r = subprocess.run([EXE, cfg], capture_output=True, text=True, timeout=120)
    return case, r


@pytest.mark.parametrize("grid", range(len(MODEL_GRIDS)))
@pytest.mark.parametrize("model_type", [1, 2, 3, 11, 12, 13, 14, 15, 21])
def test_builtin_model_matches_reference(tmp_path, model_type, grid):
    nz, lpd, top, bottom = MODEL_GRIDS[grid]
    case, r = run_model_workflow(tmp_path, model_type, nz, lpd, top, bottom)
    assert r.returncode == 0, r.stderr
    assert "successfully completed" in r.stdout
    g = case.grid
    p1 = lpd + 1
    radii = synth.coords_1d(lpd, nz, top, bottom)          # (p+1, nz)
    for pi, name in enumerate(PARAMS):
        raw = np.fromfile(os.path.join(str(tmp_path), name + ".out"), dtype=np.float32)
        assert raw.size == g.nx * g.ny * nz * p1 ** 3
        f = raw.reshape(p1, p1, p1, nz, g.ny, g.nx)          # k, j, i, ez, ey, ex
        # a 1-D model: constant over x, y elements and the i, j GLL indices
        assert np.all(f == f[:, :1, :1, :, :1, :1])
        got = f[:, 0, 0, :, 0, 0]
        want = np.array([[expected(model_type, float(radii[k, ez]), ez)[pi]
                          for ez in range(nz)] for k in range(p1)])
        if model_type == 21:
            # REAL32 interpolation in the reference against FP64 here
            scale = np.maximum(np.abs(want), 1e-6 * np.max(np.abs(want)) + 1e-30)
            assert np.all(np.abs(got - want) <= 2e-6 * scale), (name, np.max(np.abs(got - want) / scale))
        else:
            # the reference's FP64 values rounded once to REAL32
            assert np.array_equal(got, want.astype(np.float32)), (
                name, np.argwhere(got != want.astype(np.float32))[:4])


def test_prem_regions_at_boundaries():
    """The region test is `lower < r <= upper`: at each PREM boundary the
    reference returns the deeper region's values (pins the comparisons)."""
    # Moho: the LID formula, not the lower crust's constants
    rho, vp, vs = PREM[(6346600.0, False, True)]
    x = 6346600.0 / 6371000.0
    assert abs(rho - (2.691 + 0.6924 * x)) < 1e-6
    # ocean bottom with ocean: upper crust; surface with ocean: water
    assert PREM[(6368000.0, False, False)][1] == float(np.float32(5.8))
    assert PREM[(6371000.0, False, False)][2] == 0.0
    assert PREM[(6371000.0, False, True)][2] == float(np.float32(3.2))
    # CMB: D'' mantle values; just below it the fluid outer core
    assert PREM[(3480000.0, False, True)][2] == 0.0
    assert PREM[(3630000.0, False, True)][2] > 7.0


def test_ak135_discontinuity_takes_deeper_node():
    """At a listed discontinuity (the ocean floor, r = 6368 km) the reference
    lookup returns the node below it; at the surface the water values."""
    assert AK[6371.0] == pytest.approx([1.02, 1.45, 0.0], rel=1e-6)
    assert AK[6368.0] == pytest.approx([2.0, 1.65, 1.0], rel=1e-6)


def test_model_workflow_needs_override(tmp_path):
    nz, lpd, top, bottom = MODEL_GRIDS[0]
    _, r = run_model_workflow(tmp_path, 3, nz, lpd, top, bottom, params=("mu",))
    assert r.returncode == 0, r.stderr
    _, r = run_model_workflow(tmp_path, 3, nz, lpd, top, bottom, params=("mu",))
    assert r.returncode != 0 and "override" in r.stderr
    _, r = run_model_workflow(tmp_path, 3, nz, lpd, top, bottom, params=("mu",),
                              extra={"override": ".TRUE."})
    assert r.returncode == 0, r.stderr


def test_q_written_as_elastic(tmp_path):
    nz, lpd, top, bottom = MODEL_GRIDS[2]
    _, r = run_model_workflow(tmp_path, 3, nz, lpd, top, bottom, params=("q",))
    assert r.returncode == 0, r.stderr
    q = np.fromfile(os.path.join(str(tmp_path), "q.out"), dtype=np.float32)
    assert np.all(q == np.finfo(np.float32).max)


@pytest.mark.parametrize("model,msg", [
    ({"model_type": -1, "mu": "x"}, "model_type different from -1"),
    ({"model_type": 3}, "at least one model-parameter-filename"),
    ({"model_type": 7, "mu": "x"}, "Unknown model_type"),
])
def test_model_workflow_checks(tmp_path, model, msg):
    case = synth.homogeneous(nt=10, nx=4, ny=4, nz=4, n_rec=1)
    case.grid.taper_width = 1
    model = {k: (os.path.join(str(tmp_path), v) if k == "mu" else v) for k, v in model.items()}
    cfg = write_config(case, str(tmp_path), model=model, workflow="model")
    r = subprocess.run([EXE, cfg], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and msg in r.stderr


def test_forward_rejects_model_files_with_model_type(tmp_path):
    case = synth.homogeneous(nt=10, nx=4, ny=4, nz=4, n_rec=1)
    case.grid.taper_width = 1
    cfg = write_config(case, str(tmp_path), model={"model_type": 3,
                                                   "mu": os.path.join(str(tmp_path), "mu")})
    r = subprocess.run([EXE, cfg], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "can't be specified at the same time" in r.stderr


KAT = json.load(open(os.path.join(HERE, "golden", "model_known_answers.json")))


def _top_values(tmp_path, model_type, r_top):
    """Moduli written by workflow 'model' at the top GLL point (r = rad_max
    exactly) of a small grid."""
    names = ("rhoinv", "lambda", "mu", "b", "c")
    case, r = run_model_workflow(tmp_path, model_type, 4, 2, r_top, r_top - 50e3, params=names)
    assert r.returncode == 0, r.stderr
    out = {}
    for n in names:
        f = np.fromfile(os.path.join(str(tmp_path), n + ".out"), dtype=np.float32).astype(np.float64)
        out[n] = f[0]   # k = 0, ez = 0: r = rad_max
    rho = 1.0 / out["rhoinv"]
    lam, mu = out["lambda"], out["mu"]
    return {"rho": rho / 1000.0, "vsh": np.sqrt(mu / rho) / 1000.0,
            "vph": np.sqrt((lam + 2.0 * mu) / rho) / 1000.0,
            "vsv": np.sqrt(max(out["b"] + mu, 0.0) / rho) / 1000.0,
            "eta": 1.0 + out["c"] / lam}


@pytest.mark.parametrize("k", range(len(KAT["ak135_f"])))
def test_ak135_reference_known_answers(tmp_path, k):
    """The six radii of src/examples/ak135_f_test.f90 (surface, Lehmann, 410,
    660, CMB, ICB), through model_type 21."""
    kat = KAT["ak135_f"][k]
    v = _top_values(tmp_path, 21, kat["x_km"] * 1000.0)
    assert v["rho"] == pytest.approx(kat["rho"], rel=1e-6)
    assert v["vph"] == pytest.approx(kat["vp"], rel=1e-6)
    assert v["vsh"] == pytest.approx(kat["vs"], rel=1e-6, abs=1e-6)


@pytest.mark.parametrize("k", [k for k, p in enumerate(KAT["prem"]) if p["r_m"] > 0.0])
def test_prem_reference_known_answers(tmp_path, k):
    """The radii of src/examples/prem_test.f90 10 m either side of the CMB,
    the 670 and the Moho, and the surface, through model_type 11 (the
    reference's transversely isotropic prem()); each value must lie in the
    test's interval (5 decimals, so widened by 1e-5). The centre r = 0 is not
    a radius a mesh can start at (rad_max > rad_min >= 1)."""
    kat = KAT["prem"][k]
    v = _top_values(tmp_path, 11, kat["r_m"])
    for q in ("rho", "vph", "vsv", "vsh", "eta"):
        lo, hi = kat[q]
        assert lo - 1e-5 <= v[q] <= hi + 1e-5, (q, v[q], kat[q])
