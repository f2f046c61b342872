"""s3d_vrange_noalloc against s3d_vrange through the C ABI: the allocation-free
variant keeps its partial results in a fixed per-device scratch array and must
give the same min and max, on two calls in a row (the second on a smaller
mesh, so that stale partials of the first call would show), in FP32 and FP64."""
import numpy as np
import pytest

from ses3d_amd import api, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_vrange_noalloc_matches_vrange(precision):
    big = synth.layered(seed=2, nt=30, nx=20, ny=18, nz=8, n_rec=2)
    small = synth.homogeneous(seed=3, nt=30, nx=5, ny=4, nz=4, n_rec=1)
    results = []
    for case, steps in ((big, 30), (small, 25)):
        c = replace_onset(case)
        sim, _ = api.build_case(c, precision=precision)
        sim.step(steps)
        want = sim.vrange()
        got = sim.vrange_noalloc()
        print(precision, case.grid.nx, case.grid.ny, case.grid.nz, want, got, flush=True)
        assert want[0] < 0 < want[1]
        assert np.array_equal(got, want)
        # again on the same state: the result does not depend on earlier calls
        assert np.array_equal(sim.vrange_noalloc(), want)
        results.append(got)
        sim.close()
    # the smaller mesh's range differs from the larger one's
    assert not np.array_equal(results[0], results[1])


def replace_onset(case):
    """A source-time function that starts at once, so the first steps move."""
    from dataclasses import replace
    return replace(case, sources=[replace(s, onset=1.0, width=1.0) for s in case.sources])
