"""Compensated velocity integration with BF16 compensation words (kahan_v 2):
the node kernel keeps the Kahan term of each velocity component in BF16, as
the stress already can. Checked here: a velocity field reads back as set to
far below FP32 rounding (the term carries the residual of the FP32 value),
every reader of the compensated value (receivers, curl channels, field
reads) agrees with FP32 words to FP32 accuracy, and batches group the
compensation kinds apart and equal solo runs bit for bit. The 1e-5 bar
against FP64 is in test_gpu_parity.py::test_precision_modes_against_fp64."""
from dataclasses import replace

import numpy as np
import pytest

from helpers import random_model, random_unique, rel_l2, small_grid, unique_to_local
from ses3d_amd import api, synth

pytestmark = pytest.mark.gpu


def test_velocity_field_reads_back_as_set():
    grid = small_grid(lpd=4, nx=3, ny=3, nz=2)
    rng = np.random.default_rng(21)
    model = random_model(grid, rng)
    params = {"rhoinv": api.RHOINV, "lambda": api.LAMBDA, "mu": api.MU}
    out = {}
    for kv in (False, True, "bf16"):
        s = api.Simulation(grid, "fp32", kahan_v=kv, kahan_sigma="bf16")
        for k, arr in model.items():
            if k in params:
                s.set_model(params[k], arr)
        s.init()
        v = unique_to_local(random_unique(grid, np.random.default_rng(5), 1e-3), grid)
        s.set(api.VY, v)
        out[kv] = rel_l2(s.get(api.VY), v)
        s.close()
    print(out)
    assert out[False] > 1e-9            # plain FP32
    assert out[True] < 1e-13            # FP32 words hold the residual
    assert out["bf16"] < 1e-9           # BF16 words: the residual to 8 bits


def _case(nt=40):
    case = synth.layered(seed=5, nt=nt, nx=6, ny=5, nz=4, lpd=4, n_rec=3)
    src = case.sources[0]
    extra = [synth.Receiver(src.lat + 0.5, src.lon - 0.4, 10.0, "P1", ("pressure", "rot_vz", "vz"))]
    return replace(case, receivers=extra + case.receivers)


def test_bf16_words_follow_fp32_words():
    """Traces (including pressure and a curl channel) and fields of a run
    with BF16 compensation words against FP32 words."""
    case = _case()
    a, recs = api.build_case(case, precision="fp32", kahan_v=True, kahan_sigma="bf16")
    b, _ = api.build_case(case, precision="fp32", kahan_v="bf16", kahan_sigma="bf16")
    a.step(40)
    b.step(40)
    for r in range(len(recs)):
        for ch in range(3):
            ref = a.trace(r, ch)
            if np.linalg.norm(ref) > 0:
                assert rel_l2(b.trace(r, ch), ref) < 1e-6, (r, ch)
    for f in (api.VX, api.VZ, api.SXY):
        assert rel_l2(b.get(f), a.get(f)) < 1e-6, f
    a.close()
    b.close()


def test_batch_groups_compensation_kinds():
    case = _case()
    kinds = [("bf16", 2), (True, 2), ("bf16", 1)]
    solo = []
    for kv, nbuf in kinds:
        s, recs = api.build_case(case, precision="fp32", kahan_v=kv, kahan_sigma="bf16", state_buffers=nbuf)
        s.step(37)
        solo.append(s)
    members = [api.build_case(case, precision="fp32", kahan_v=kv, kahan_sigma="bf16", state_buffers=nbuf)[0]
               for kv, nbuf in kinds]
    b = api.Batch(members)
    b.step(37)
    b.synchronize()
    b.close()
    for q, (s, m) in enumerate(zip(solo, members)):
        for r in range(len(recs)):
            for ch in range(3):
                assert np.array_equal(s.trace(r, ch), m.trace(r, ch)), (q, r, ch)
        for f in (api.VX, api.SXX):
            assert np.array_equal(s.get(f), m.get(f)), (q, f)


def test_kahan_v_option_checked():
    grid = small_grid(nx=2, ny=2, nz=2)
    with pytest.raises(api.S3dError, match="kahan_v"):
        api.Simulation(grid, "fp32", kahan_v=3)
