"""Five memory words per relaxation mechanism (memory_vars 5: the deviatoric
part xx xy xz yy yz of the memory variables, zz = -(xx + yy)) against the
six of the reference's form. The stress rate sees the memory variables only
as Md - tr(Md)/3 I (src/evolution.f90:216-284, loop form), and every
component relaxes at the same rate with the same forcing (:415-442), so the
deviatoric part evolves on its own and the stress and velocity follow the
same equations. Checked against the FP64 oracle, which keeps all six words
with their trace, at FP64 rounding; on the timed FP32 mode against FP64 at
the 1e-5 trace bar; in batches and kernel variants bit for bit against solo
runs; and for the field interface of the form."""
import numpy as np
import pytest

import oracle
from helpers import random_model, random_unique, rel_l2, small_grid, unique_to_local
from ses3d_amd import api, synth

pytestmark = pytest.mark.gpu

PARAMS = {"rhoinv": (oracle.RHOINV, api.RHOINV), "lambda": (oracle.LAMBDA, api.LAMBDA),
          "mu": (oracle.MU, api.MU), "a": (oracle.A, api.A), "b": (oracle.B, api.B),
          "c": (oracle.C_, api.C_), "q": (oracle.Q, api.Q)}
XX, XY, XZ, YY, YZ, ZZ = range(6)
FIELDS = [api.VX, api.VY, api.VZ] + list(range(api.SXX, api.SZZ + 1))


def _pair(grid, model, precision="fp64", **kw):
    o = oracle.Oracle(grid)
    s = api.Simulation(grid, precision, memory_vars=5, **kw)
    for k, arr in model.items():
        o.set_model(PARAMS[k][0], arr)
        s.set_model(PARAMS[k][1], arr)
    o.init()
    s.init()
    return o, s


def _random_state(o, s, grid, rng):
    """Random velocity and stress, and random memory variables WITH a trace
    for the oracle; the simulation gets their deviatoric part."""
    for f in (api.VX, api.VY, api.VZ):
        v = unique_to_local(random_unique(grid, rng, 1e-3), grid)
        o.set(f, v)
        s.set(f, v)
    P1 = grid.lpd + 1
    shp = (P1, P1, P1, grid.nz, grid.ny, grid.nx)
    for f in range(api.SXX, api.SZZ + 1):
        sg = 1e5 * rng.standard_normal(shp)
        o.set(f, sg)
        s.set(f, sg)
    for m in range(3):
        M = [1e-9 * rng.standard_normal(shp) for _ in range(6)]
        tr3 = (M[XX] + M[YY] + M[ZZ]) / 3.0
        for c in range(6):
            o.set(oracle.MXX + 6 * m + c, M[c])
            if c in (XX, YY):
                s.set(api.MXX + 6 * m + c, M[c] - tr3)
            elif c != ZZ:
                s.set(api.MXX + 6 * m + c, M[c])


def _deviatoric_memory(o, m):
    M = [o.get(oracle.MXX + 6 * m + c) for c in range(6)]
    tr3 = (M[XX] + M[YY] + M[ZZ]) / 3.0
    return [M[XX] - tr3, M[XY], M[XZ], M[YY] - tr3, M[YZ], M[ZZ] - tr3]


@pytest.mark.parametrize("aniso", [False, True])
def test_fp64_five_words_follow_the_oracle(aniso):
    """The oracle's memory variables carry a random trace that the five-word
    form never stores: velocity and stress still agree to FP64 rounding, and
    the stored words equal the deviatoric part of the oracle's."""
    grid = small_grid(nx=3, ny=3, nz=2)
    rng = np.random.default_rng(17)
    o, s = _pair(grid, random_model(grid, rng, visco=True, aniso=aniso))
    assert o.visco
    _random_state(o, s, grid, rng)
    for _ in range(3):
        o.step(1)
        s.step(1)
    for f in FIELDS:
        assert rel_l2(s.get(f), o.get(f)) < 1e-12, f
    for m in range(3):
        dev = _deviatoric_memory(o, m)
        for c in range(6):
            assert rel_l2(s.get(api.MXX + 6 * m + c), dev[c]) < 1e-12, (m, c)
    s.close()


@pytest.mark.parametrize("lpd", [2, 4, 7])
def test_fp64_five_words_forward_run_matches_oracle(lpd):
    case = synth.layered(seed=3, nt=160, nx=6, ny=6, nz=4, lpd=lpd, n_rec=6)
    o, orecs = oracle.build_case(case)
    o.step(case.grid.nt)
    sim, recs = api.build_case(case, precision="fp64", memory_vars=5)
    assert len(recs) == len(orecs) > 0
    sim.step(case.grid.nt)
    errs = [rel_l2(sim.trace(r, ch), o.trace(r, ch)) for r in range(len(recs)) for ch in range(3)]
    print("p=%d max relL2 %.3e" % (lpd, max(errs)))
    assert max(errs) < 1e-10
    sim.close()


def _traces(sim, nrec):
    return np.array([[sim.trace(r, ch) for ch in range(3)] for r in range(nrec)])


def test_timed_mode_five_words_against_fp64():
    """FP32 with compensated velocity and BF16-compensated stress, five
    memory words, against the FP64 path in the reference's six-word form:
    within 1e-5 per trace, and no worse than twice the six-word FP32 run."""
    case = synth.layered(seed=7, nt=400, nx=10, ny=10, nz=5, lpd=4, n_rec=8)
    ref, recs = api.build_case(case, precision="fp64")
    ref.step(400)
    t64 = _traces(ref, len(recs))
    ref.close()
    norms = np.linalg.norm(t64, axis=2)
    keep = norms > 1e-3 * norms.max()
    out = {}
    for vm in (6, 5):
        sim, _ = api.build_case(case, precision="fp32", kahan_sigma="bf16", memory_vars=vm)
        sim.step(400)
        err = np.linalg.norm(_traces(sim, len(recs)) - t64, axis=2)[keep] / norms[keep]
        out[vm] = (float(np.median(err)), float(err.max()))
        sim.close()
    print(out)
    assert out[5][1] < 1e-5
    assert out[5][0] < 2 * out[6][0] + 1e-8


@pytest.mark.parametrize("ks", [False, True, "bf16"])
def test_five_word_state_block_fields(ks):
    """The state block holds 21, 27 or 24 words per node (stress, its
    compensation, 15 memory words): every stored field reads back as set,
    zz reads as -(xx + yy) and cannot be set, and a few steps follow the
    FP64 oracle. A wrong stride or offset gives O(1) errors."""
    grid = small_grid(nx=3, ny=3, nz=2)
    rng = np.random.default_rng(11)
    o, s = _pair(grid, random_model(grid, rng, visco=True), precision="fp32", kahan_v=True,
                 kahan_sigma=ks)
    _random_state(o, s, grid, rng)
    for f in range(api.SXX, api.SZZ + 1):
        assert rel_l2(s.get(f), o.get(f)) < 1e-7, f
    for m in range(3):
        dev = _deviatoric_memory(o, m)
        for c in range(6):
            assert rel_l2(s.get(api.MXX + 6 * m + c), dev[c]) < 1e-6, (m, c)
    with pytest.raises(api.S3dError, match="cannot be set"):
        s.set(api.MXX + 6 + ZZ, np.zeros(o.shape))
    for _ in range(3):
        o.step(1)
        s.step(1)
    for f in FIELDS:
        assert rel_l2(s.get(f), o.get(f)) < 1e-4, f
    s.close()


def test_memory_vars_option_checked():
    grid = small_grid(nx=2, ny=2, nz=2)
    with pytest.raises(api.S3dError, match="memory_vars"):
        api.Simulation(grid, "fp32", memory_vars=4)


def _case(nt=40):
    return synth.layered(seed=5, nt=nt, nx=6, ny=5, nz=4, lpd=4, n_rec=3)


def _same_traces(a, b, nrec, what):
    for r in range(nrec):
        for ch in range(3):
            assert np.array_equal(a.trace(r, ch), b.trace(r, ch)), (what, r, ch)


def test_kernel_variants_and_tiles_agree_bit_for_bit():
    """Planar, node-interleaved and early-store variants and the element
    tiles of the five-word kernel give identical traces."""
    case = _case()
    base, recs = api.build_case(case, precision="fp32", kahan_sigma="bf16", memory_vars=5)
    base.step(case.grid.nt)
    for tile, layout, early in ((1, 0, 0), (2, 1, 0), (4, 0, 1), (2, 0, 1)):
        s, _ = api.build_case(case, precision="fp32", kahan_sigma="bf16", memory_vars=5)
        s.set_tile(tile)
        s.set_lds_layout(layout)
        s.set_store_order(early)
        s.step(case.grid.nt)
        _same_traces(base, s, len(recs), (tile, layout, early))
        s.close()
    base.close()


def test_batch_with_both_forms_matches_solo_runs():
    """A batch groups its members by memory form: five- and six-word
    members stepped together equal their solo runs bit for bit."""
    case = _case()
    kinds = [(5, 2), (6, 2), (5, 1)]
    solo = []
    for vm, nbuf in kinds:
        s, recs = api.build_case(case, precision="fp32", kahan_sigma="bf16", memory_vars=vm,
                                 state_buffers=nbuf)
        s.step(37)
        solo.append(s)
    members = [api.build_case(case, precision="fp32", kahan_sigma="bf16", memory_vars=vm,
                              state_buffers=nbuf)[0] for vm, nbuf in kinds]
    b = api.Batch(members)
    b.step(37)
    b.synchronize()
    b.close()
    for q, (s, m) in enumerate(zip(solo, members)):
        _same_traces(s, m, len(recs), "member %d" % q)
        for f in (api.VX, api.SXX, api.MXX + 7):
            assert np.array_equal(s.get(f), m.get(f)), (q, f)


@pytest.mark.parametrize("precision,kw,tol", [("fp32", {"bf16_static": True}, 1e-5),
                                              ("fp64", {}, 1e-11)])
def test_other_kernels_with_five_words(precision, kw, tol):
    """The BF16-static FP32 kernel and the FP64 kernel at every tile with
    five memory words follow the six-word runs of the same kind."""
    case = _case(nt=60)
    six, recs = api.build_case(case, precision=precision, memory_vars=6, **kw)
    six.step(60)
    tiles = (0,) if kw else (0, 1, 4)
    for tile in tiles:
        five, _ = api.build_case(case, precision=precision, memory_vars=5, **kw)
        five.set_tile(tile)
        five.step(60)
        for r in range(len(recs)):
            for ch in range(3):
                ref = six.trace(r, ch)
                if np.linalg.norm(ref) > 0:
                    assert rel_l2(five.trace(r, ch), ref) < tol, (tile, r, ch)
        five.close()
    six.close()
