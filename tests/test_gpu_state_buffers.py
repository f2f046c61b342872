"""Two state buffers (the element kernel reads the stress, its compensation
and the memory variables from one block and writes them to the other,
swapped every step) against one block updated in place: the arithmetic is
the same, so traces and fields must agree bit for bit, through graph chunks
that start on either block, odd step counts, s3d_set_field between steps,
receivers that read the stress (pressure) and batches whose members start on
different blocks."""
from dataclasses import replace

import numpy as np
import pytest

from ses3d_amd import api, synth

pytestmark = pytest.mark.gpu

STATE = [api.VX, api.VY, api.VZ] + list(range(api.SXX, api.SZZ + 1)) + [api.MXX, api.MXX + 7, api.MXX + 17]


def _case(nt=60):
    case = synth.layered(seed=5, nt=nt, nx=6, ny=5, nz=4, lpd=4, n_rec=3)
    src = case.sources[0]
    extra = [synth.Receiver(src.lat + 0.5, src.lon - 0.4, 10.0, "P1", ("pressure", "vz", "n"))]
    return replace(case, receivers=extra + case.receivers)


def _same(a, b, nrec, what):
    for r in range(nrec):
        for ch in range(3):
            assert np.array_equal(a.trace(r, ch), b.trace(r, ch)), (what, r, ch)
    for f in STATE:
        assert np.array_equal(a.get(f), b.get(f)), (what, f)


@pytest.mark.parametrize("precision,kw", [("fp32", {"kahan_sigma": "bf16"}), ("fp32", {"kahan_sigma": True}),
                                          ("fp64", {})])
def test_two_state_buffers_match_one(precision, kw):
    case = _case()
    one, recs = api.build_case(case, precision=precision, state_buffers=1, **kw)
    two, _ = api.build_case(case, precision=precision, state_buffers=2, **kw)
    for n in (5, 16, 3, 17):       # graph chunks from both blocks, odd remainders
        one.step(n)
        two.step(n)
    # a field set between steps lands in the block the next step reads
    sxy = one.get(api.SXY) * 1.5
    one.set(api.SXY, sxy)
    two.set(api.SXY, sxy)
    one.step(19)
    two.step(19)
    _same(one, two, len(recs), precision)


def test_batch_members_on_either_block_match_solo_runs():
    case = _case()
    solo = []
    for pre, nbuf in ((3, 2), (0, 2), (0, 1)):
        s, recs = api.build_case(case, precision="fp32", kahan_sigma="bf16", state_buffers=nbuf)
        s.step(pre + 37)
        solo.append(s)
    members = []
    for pre, nbuf in ((3, 2), (0, 2), (0, 1)):
        s, _ = api.build_case(case, precision="fp32", kahan_sigma="bf16", state_buffers=nbuf)
        s.step(pre)                # the first member starts the batch on its second block
        members.append(s)
    b = api.Batch(members)
    b.step(17)                     # one graph chunk and an odd remainder
    b.step(20)                     # a chunk from the other phase, then four steps
    b.synchronize()
    b.close()
    for q, (s, m) in enumerate(zip(solo, members)):
        _same(s, m, len(recs), "member %d" % q)


def test_shared_sources_with_one_and_two_buffers_agree():
    ms = synth.multi_source(n_src=3, nt=40, nx=6, ny=6, nz=4, n_rec=4)
    runs = []
    for nbuf in (1, 2):
        sims, recs = api.build_shared_sources(ms, kahan_sigma="bf16", state_buffers=nbuf)
        b = api.Batch(sims)
        b.step(40)
        b.synchronize()
        b.close()
        runs.append(sims)
    for a, c in zip(*runs):
        _same(a, c, len(recs), "shared")


def test_state_buffers_option_and_memory():
    case = synth.homogeneous(nt=4, nx=8, ny=8, nz=4, n_rec=1)
    with pytest.raises(api.S3dError, match="state_buffers"):
        api.Simulation(case.grid, "fp32", state_buffers=3)
    one, _ = api.build_case(case, precision="fp32", state_buffers=1)
    two, _ = api.build_case(case, precision="fp32")
    g = case.grid
    nn = g.nx * g.ny * g.nz * (g.lpd + 1) ** 3
    # elastic, no stress compensation: the state block is the six stresses
    assert two.device_bytes() - one.device_bytes() == 6 * 4 * nn
