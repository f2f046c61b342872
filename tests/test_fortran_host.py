"""The Fortran host program (ses3d-nt_amd/bin/ses3d_nt): NAMELIST config
checks on the CPU, and an end-to-end forward run on the GPU compared with the
FP64 oracle through the SAC files it writes."""
import os
import subprocess

import numpy as np
import pytest

from helpers import read_sac, rel_l2, write_config
from ses3d_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ses3d-nt_amd", "bin", "ses3d_nt")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")


def run(cfg, timeout=300):
    return subprocess.run([EXE, cfg], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("edit,msg", [
    (("nt = 60", "nt = 60000"), "Non-conforming nt"),
    (("nx = 6", "nx = 3"), "greater than 3"),
    (("lpd = 4", "lpd = 9"), "between 2 and 7"),
    (("taper_width = 2", "taper_width = 4"), "too wide"),
    (("precision = 'fp32'", "precision = 'fp16'"), "precision must be"),
])
def test_config_checks_abort(tmp_path, edit, msg):
    case = synth.homogeneous(nt=60, nx=6, ny=6, nz=5, n_rec=2)
    cfg = write_config(case, str(tmp_path))
    text = open(cfg).read()
    assert edit[0] in text
    open(cfg, "w").write(text.replace(edit[0], edit[1], 1))
    r = run(cfg)
    assert r.returncode != 0
    assert msg in r.stderr


def test_missing_model_file_aborts(tmp_path):
    case = synth.homogeneous(nt=60, nx=6, ny=6, nz=5, n_rec=2)
    cfg = write_config(case, str(tmp_path))
    os.remove(os.path.join(str(tmp_path), "mu"))
    r = run(cfg)
    assert r.returncode != 0
    assert "does not exist" in r.stderr or "no HIP device" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("fp32", 1e-5)])
def test_forward_run_matches_oracle(tmp_path, precision, tol):
    import oracle
    nt = 200
    case = synth.homogeneous(nt=nt, nx=8, ny=8, nz=5, n_rec=5)
    snap = ("\n&output_raw\n timestep_start = 100\n timestep_increment = 100\n"
            " timestep_end = 200\n attributes = 'vz' 'pressure'\n override = .TRUE.\n"
            " prefix = '%s/'\n/\n" % str(tmp_path))
    cfg = write_config(case, str(tmp_path), precision=precision, extra=snap)
    r = run(cfg)
    print(r.stdout[-600:], r.stderr[-600:])
    assert r.returncode == 0
    assert "successfully completed" in r.stdout
    o, recs = oracle.build_case(case)
    o.step(100)
    dtype = "<f4" if precision == "fp32" else "<f8"
    vz100 = np.fromfile(os.path.join(str(tmp_path), "vz_100.vol"), dtype=dtype)
    assert rel_l2(vz100, o.get(oracle.VZ).ravel()) < tol
    o.step(nt - 100)
    p200 = np.fromfile(os.path.join(str(tmp_path), "pressure_200.vol"), dtype=dtype)
    pref = -(o.get(oracle.SXX) + o.get(oracle.SYY) + o.get(oracle.SZZ)).ravel() / 3.0
    assert rel_l2(p200, pref) < 10 * tol
    for ri, rec in enumerate(recs):
        for ci, ch in enumerate(rec.channels):
            f, i, data = read_sac(os.path.join(str(tmp_path), "XX.00.%s.%s.sac" % (rec.station, ch)))
            assert i[9] == nt and abs(f[0] - case.grid.dt) < 1e-6 * case.grid.dt
            # SAC stores float32: the floor is single-precision rounding
            assert rel_l2(data, o.trace(ri, ci)) < tol, (rec.station, ch)


@pytest.mark.gpu
def test_forward_builtin_prem_matches_oracle(tmp_path):
    """model_type 3 (PREM, crust without ocean) built by the host and uploaded
    in FP64, against the oracle run on the model that workflow 'model' writes
    for the same grid (REAL32 files: agreement to single precision)."""
    import oracle
    nt = 200
    case = synth.homogeneous(nt=nt, nx=8, ny=8, nz=5, n_rec=5)
    case.grid.dt = synth.cfl_dt(case.grid, lambda r: np.full_like(r, 10000.0), 0.3)
    names = ("rhoinv", "lambda", "mu")
    mdir = tmp_path / "model"
    mdir.mkdir()
    cfg = write_config(case, str(mdir), workflow="model",
                       model=dict([("model_type", 3)] +
                                  [(k, os.path.join(str(mdir), k)) for k in names]))
    r = run(cfg)
    assert r.returncode == 0, r.stderr
    g = case.grid
    shp = (g.lpd + 1,) * 3 + (g.nz, g.ny, g.nx)
    m = {k: np.fromfile(os.path.join(str(mdir), k), dtype="<f4").astype(np.float64).reshape(shp)
         for k in names}
    case.model_full = lambda: m
    cfg = write_config(case, str(tmp_path), precision="fp64", model={"model_type": 3})
    r = run(cfg)
    print(r.stdout[-600:], r.stderr[-600:])
    assert r.returncode == 0
    o, recs = oracle.build_case(case)
    o.step(nt)
    for ri, rec in enumerate(recs):
        for ci, ch in enumerate(rec.channels):
            _, _, data = read_sac(os.path.join(str(tmp_path), "XX.00.%s.%s.sac" % (rec.station, ch)))
            assert rel_l2(data, o.trace(ri, ci)) < 1e-5, (rec.station, ch)
