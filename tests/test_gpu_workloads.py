"""The five benchmark workloads at their own sizes on the GPU.

* homogeneous elastic 12x12x8, p = 4, 1000 steps, 16 receivers: FP64 and the
  timed FP32 mode against the FP64 oracle.
* layered visco-elastic 96x96x32, p = 4, 4000 steps, 256 receivers: the
  timed FP32 mode against the FP64 GPU path (traces and a snapshot).
* the same model at 192x192x96 (442 M GLL nodes): the timed FP32 mode against
  the FP64 GPU path over 1500 steps, with receivers on rings around the
  source so that the whole wavelet reaches them inside the run.
* 64 uneven regional simulations (8x8x8 .. 48x48x24, p = 4 or 6) advanced as
  one batch: members equal the same simulations stepped alone, bit for bit.
* 32 sources on one 128x128x48, p = 4 model, as two batches of 16 (the state
  of 32 sources exceeds the 288 GB of HBM): each batched source equals the
  same source stepped alone on the shared model, bit for bit.

Misfits are relative L2 per trace over the traces with signal (norm above
1e-3 of the largest), the quantity the benchmark's 1e-5 bound is stated in.
"""
import time
from dataclasses import replace

import numpy as np
import pytest

import oracle
from helpers import rel_l2
from ses3d_amd import api, synth

pytestmark = pytest.mark.gpu

# the mode bench.py times: FP32, compensated velocity, BF16-compensated
# stress, memory variables held as their deviatoric part (five words per
# mechanism); the FP64 runs it is checked against keep all six
TIMED = dict(precision="fp32", kahan_v=True, kahan_sigma="bf16", memory_vars=5)


def log(*a):
    print("[workloads]", *a, flush=True)


def _traces(sim, recs):
    return np.array([[sim.trace(r, c) for c in range(3)] for r in range(len(recs))])


def _misfit(t, ref):
    norms = np.linalg.norm(ref, axis=2)
    keep = norms > 1e-3 * norms.max()
    err = np.linalg.norm(t - ref, axis=2)[keep] / norms[keep]
    return float(np.median(err)), float(err.max()), int(keep.sum())


def test_homogeneous_12x12x8_1000_steps_against_oracle():
    case = synth.homogeneous()                      # 12x12x8, p=4, 1000 steps, 16 receivers
    g = case.grid
    assert (g.nx, g.ny, g.nz, g.lpd, g.nt, len(case.receivers)) == (12, 12, 8, 4, 1000, 16)
    t0 = time.time()
    o, orecs = oracle.build_case(case)
    o.step(g.nt)
    ref = np.array([[o.trace(r, c) for c in range(3)] for r in range(len(orecs))])
    log("homogeneous 12x12x8 oracle %.1f s" % (time.time() - t0))
    for name, kw, tol in (("fp64", dict(precision="fp64"), 1e-11), ("fp32 timed", TIMED, 1e-5)):
        sim, recs = api.build_case(case, **kw)
        assert len(recs) == len(orecs) == 16
        sim.step(g.nt)
        med, mx, n = _misfit(_traces(sim, recs), ref)
        log("homogeneous 12x12x8 %s: trace rel L2 median %.2e max %.2e over %d traces" % (name, med, mx, n))
        assert mx < tol
        sim.close()


def _fp32_vs_fp64(case, tag, snap_tol=1e-5):
    g = case.grid
    out = {}
    for name, kw in (("fp64", dict(precision="fp64")), ("fp32", TIMED)):
        t0 = time.time()
        sim, recs = api.build_case(case, slab=4, **kw)
        log("%s %s set up in %.0f s, %d receivers" % (tag, name, time.time() - t0, len(recs)))
        t0 = time.time()
        done = 0
        while done < g.nt:
            n = min(500, g.nt - done)
            sim.step(n)
            sim.synchronize()
            done += n
        log("%s %s %d steps in %.1f s" % (tag, name, g.nt, time.time() - t0))
        out[name] = (_traces(sim, recs), sim.get(api.VZ))
        sim.close()
    med, mx, n = _misfit(out["fp32"][0], out["fp64"][0])
    snap = rel_l2(out["fp32"][1], out["fp64"][1])
    log("%s timed FP32 vs FP64: trace rel L2 median %.2e max %.2e over %d traces, "
        "vz snapshot %.2e" % (tag, med, mx, n, snap))
    assert n >= 16
    assert mx < 1e-5
    assert snap < snap_tol


def test_layered_visco_96x96x32_4000_steps():
    case = synth.layered()                          # 96x96x32, p=4, 4000 steps, 256 receivers
    g = case.grid
    assert (g.nx, g.ny, g.nz, g.nt, len(case.receivers)) == (96, 96, 32, 4000, 256)
    _fp32_vs_fp64(case, "layered 96x96x32")


def test_layered_visco_192x192x96_1500_steps():
    case = synth.layered(nx=192, ny=192, nz=96, nt=1500, n_rec=32)
    src = case.sources[0]
    ring = []
    for d in (0.5, 1.0, 2.0, 3.0):
        for az in range(0, 360, 45):
            a = np.radians(az)
            ring.append(synth.Receiver(src.lat + d * np.cos(a), src.lon + d * np.sin(a), 0.0,
                                       "R%d_%d" % (int(10 * d), az)))
    case = replace(case, receivers=ring + case.receivers)
    _fp32_vs_fp64(case, "layered 192x192x96")


def _early(case, onset=2.0, width=2.0):
    """The same case with a source-time function that starts at once, so that a
    few steps already carry signal."""
    return replace(case, sources=[replace(s, onset=onset, width=width) for s in case.sources])


def test_uneven_batch_of_64_equals_solo_runs():
    steps = 20
    cases = [_early(c) for c in synth.uneven_batch(seed=4, n_sims=64, nt=steps)]
    assert len(cases) == 64 and {c.grid.lpd for c in cases} == {4, 6}
    t0 = time.time()
    sims = [api.build_case(c, **TIMED)[0] for c in cases]
    log("uneven batch: 64 simulations, %d GLL nodes, set up in %.0f s"
        % (sum(c.grid.nodes for c in cases), time.time() - t0))
    b = api.Batch(sims)
    b.step(3)
    b.step(16)
    b.step(1)
    b.synchronize()
    b.close()
    sizes = [c.grid.nodes for c in cases]
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)),
                   next(q for q, c in enumerate(cases) if c.grid.lpd == 6), 31})
    for q in pick:
        solo, recs = api.build_case(cases[q], **TIMED)
        solo.step(steps)
        vz = solo.get(api.VZ)
        assert np.abs(vz).max() > 0
        assert np.array_equal(sims[q].get(api.VZ), vz), q
        for r in range(len(recs)):
            for ch in range(3):
                assert np.array_equal(sims[q].trace(r, ch), solo.trace(r, ch)), (q, r, ch)
        solo.close()
    log("uneven batch: members %s equal their solo runs" % pick)
    for s in sims:
        s.close()


def test_32_sources_on_128x128x48_as_two_batches_of_16():
    steps = 20
    case = _early(synth.multi_source(n_src=32, nt=steps))
    g = case.grid
    assert (g.nx, g.ny, g.nz, g.lpd, len(case.sources)) == (128, 128, 48, 4, 32)
    t0 = time.time()
    # one state buffer per source (updated in place): 16 sources and the
    # model's owner then fit on one GPU (224 GB; two buffers would need 400 GB)
    owner, recs = api.build_case(replace(case, sources=[]), slab=4, state_buffers=1, **TIMED)
    log("32 sources: model set up in %.0f s" % (time.time() - t0))
    batched = {}
    for half in (case.sources[:16], case.sources[16:]):
        sims = []
        for s in half:
            sim = api.Simulation.shared(owner)
            api._add_channels(sim, case, [s])
            sims.append(sim)
        b = api.Batch(sims)
        b.step(4)
        b.step(16)
        b.synchronize()
        b.close()
        log("32 sources: batch of %d sources, device %.0f GB" %
            (len(sims), (owner.device_bytes() + sum(x.device_bytes() for x in sims)) / 1e9))
        for s, sim in zip(half, sims):
            q = case.sources.index(s)
            if q in (5, 20):
                batched[q] = (_traces(sim, recs), sim.get(api.VZ))
            sim.close()
    for q, (tr, vz) in batched.items():
        lone = api.Simulation.shared(owner)
        api._add_channels(lone, case, [case.sources[q]])
        lone.step(steps)
        v = lone.get(api.VZ)
        assert np.abs(v).max() > 0
        assert np.array_equal(vz, v), q
        assert np.array_equal(tr, _traces(lone, recs)), q
        lone.close()
    log("32 sources: sources %s equal their solo runs" % sorted(batched))
    owner.close()
