"""The displacement gradient (reference K3, src/evolution.f90:179-187) and
the snapshot fields exx .. ezz built from it (src/volume_snapshot.f90:
220-236), against the oracle, which integrates the reference's nine
components (pinned by a closed form in test_oracle_physics.py)."""
import numpy as np
import pytest

import oracle
from helpers import random_model, rel_l2, small_grid
from ses3d_amd import api, synth
from test_gpu_parity import _pair, _random_state

pytestmark = pytest.mark.gpu


def _symmetric(o):
    d = [o.get(oracle.DXUX + q) for q in range(9)]
    return {api.EXX: d[0], api.EXY: (d[1] + d[3]) / 2, api.EXZ: (d[2] + d[6]) / 2,
            api.EYY: d[4], api.EYZ: (d[5] + d[7]) / 2, api.EZZ: d[8]}


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_displacement_gradient_matches_oracle(precision, tol):
    grid = small_grid(nx=4, ny=3, nz=3, taper_width=1)
    rng = np.random.default_rng(21)
    o, s = _pair(grid, random_model(grid, rng, visco=True), precision=precision)
    _random_state(o, s, grid, rng, visco=True)
    o.enable_displacement_gradient()
    s.enable_displacement_gradient()
    for _ in range(4):
        o.step(5)
        s.step(5)
    for f, ref in _symmetric(o).items():
        assert np.max(np.abs(ref)) > 0, f
        assert rel_l2(s.get(f), ref) < tol, f


def test_displacement_gradient_in_a_batch_matches_solo():
    case = synth.layered(seed=4, nt=12, nx=6, ny=5, nz=4, lpd=4, n_rec=2)
    solo, _ = api.build_case(case, precision="fp32", kahan_sigma="bf16")
    solo.enable_displacement_gradient()
    solo.step(12)
    member, _ = api.build_case(case, precision="fp32", kahan_sigma="bf16")
    member.enable_displacement_gradient()
    b = api.Batch([member])
    b.step(12)
    b.synchronize()
    with pytest.raises(api.S3dError, match="batch"):
        member.enable_displacement_gradient()
    b.close()
    for f in (api.EXX, api.EXY, api.EZZ):
        assert np.array_equal(member.get(f), solo.get(f)), f


def test_displacement_gradient_must_be_enabled():
    s, _ = api.build_case(synth.homogeneous(nt=4, nx=4, ny=4, nz=3, n_rec=1), precision="fp32")
    with pytest.raises(api.S3dError, match="not integrated"):
        s.get(api.EXX)
    s.enable_displacement_gradient()
    s.enable_displacement_gradient()   # a second call keeps the sums
    s.step(2)
    assert np.all(np.isfinite(s.get(api.EYZ)))
