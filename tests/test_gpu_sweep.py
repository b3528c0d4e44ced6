"""Batched sweep of runs of different resolution (mlg_sweep_*): every run's
state after the sweep's steps is bitwise the state of the same run stepped
alone (mlg_time_steps), for the ungrouped graph, the grouped RTRAN backward
launch with every tile shape, and the tuned shapes; the tile tuner's report.

The reference runs one case per job and has no sweep; the runs' own steps
are checked against the FP64 oracle in test_gpu_parity.py, so bitwise
equality with them carries that parity over to the sweep.
"""
import math

import numpy as np
import pytest

from mlegs_amd import hip as H
from mlegs_amd.configs import CFG1, LAMB_OSEEN_Q, cfg4_runs, seeded_noise

pytestmark = pytest.mark.gpu

# small runs of uneven grids: NR, NTH = NX, truncations and parameters differ
SMALL_RUNS = [
    dict(CFG1, nr=32, nth=16, nx=16, nrchop=22, ntchop=6, nxchop=5, ell=1.0, omega=0.5, bv0=1.0),
    dict(CFG1, nr=48, nth=32, nx=16, nrchop=32, ntchop=11, nxchop=6, ell=2.2, omega=1.4, bv0=0.3),
    dict(CFG1, nr=64, nth=16, nx=32, nrchop=42, ntchop=6, nxchop=11, ell=0.7, omega=0.0, bv0=3.1),
    dict(CFG1, nr=96, nth=32, nx=32, nrchop=64, ntchop=11, nxchop=11, ell=1.6, omega=1.9, bv0=0.0),
    dict(CFG1, nr=200, nth=16, nx=16, nrchop=140, ntchop=6, nxchop=6, ell=2.9, omega=0.2, bv0=2.2),
]


class Run:
    """One run: grid, solver and its started state."""

    def __init__(self, cfg):
        self.g = H.Grid(cfg)
        self.f = [self.g.scalar() for _ in range(3)]
        psi, chi, b = self.f
        H.qvortex(self.g, LAMB_OSEEN_Q, psi, chi)
        p, _, ln = psi.download()
        psi.upload(p + seeded_noise(cfg, 1), H.FFF, ln)
        chi.upload(seeded_noise(cfg, 2), H.FFF, 0.0)
        b.upload(seeded_noise(cfg, 3), H.FFF, 0.0)
        self.s = H.Solver(self.g)
        self.s.initialize(*self.f)

    def state(self):
        return [f.download() for f in self.f]

    def close(self):
        for f in self.f:
            f.close()
        self.s.close()
        self.g.close()


def same_state(x, y):
    for (xa, _, xl), (ya, _, yl) in zip(x, y):
        if not (np.array_equal(xa, ya) and xl == yl):
            return False
    return True


def check_sweep(cfgs, steps, **kw):
    a = [Run(c) for c in cfgs]
    b = [Run(c) for c in cfgs]
    sw = None
    try:
        sw = H.Sweep([(r.s, *r.f) for r in a], **kw)
        done = 0
        for n in steps:
            print(f"sweep {kw} steps {n}", flush=True)
            sw.steps(n)
            done += n
            for r in b:
                r.s.time_stepping(*r.f, n)
            for i, (ra, rb) in enumerate(zip(a, b)):
                assert same_state(ra.state(), rb.state()), f"run {i} differs after {done} steps"
                assert ra.s.nstep == rb.s.nstep and abs(ra.s.t - rb.s.t) < 1e-15
        return sw.info()
    finally:
        if sw:
            sw.close()
        for r in a + b:
            r.close()


@pytest.mark.parametrize("kw", [dict(grouped=False), dict(grouped=True, shape=0), dict(grouped=True, shape=1),
                                dict(grouped=True, shape=2), dict(grouped=True, tune=True),
                                dict(grouped=False, streams=True), dict(grouped=True, tune=True, streams=True),
                                dict(grouped=True, tune=True, streams=True, by_grid=True),
                                dict(grouped=True, shape=1, streams=True, by_grid=True)],
                         ids=["ungrouped", "grouped_64x64", "grouped_32x64", "grouped_64x32", "grouped_tuned",
                              "ungrouped_streams", "grouped_tuned_streams", "grouped_tuned_by_grid_streams",
                              "grouped_32x64_by_grid_streams"])
def test_sweep_equals_runs_alone(kw):
    # first call: eager step, record, replays; second call: replays only
    info = check_sweep(SMALL_RUNS, [4, 3], **kw)
    assert info["runs"] == len(SMALL_RUNS)
    if kw["grouped"]:
        assert info["grouped_positions"] >= 1 and info["grouped_jobs"] >= len(SMALL_RUNS)
        if "shape" in kw:
            shape = H.BWD_TILE_SHAPES[kw["shape"]]
            assert info["jobs_per_shape"][shape] == info["grouped_jobs"]
    else:
        assert info["grouped_positions"] == 0 and info["grouped_jobs"] == 0


def test_sweep_rerecords_after_buffer_swap():
    """A transform of a run's field between calls moves its buffers: the next
    call records the sweep again and still matches the run stepped alone."""
    a = [Run(c) for c in SMALL_RUNS[:3]]
    b = [Run(c) for c in SMALL_RUNS[:3]]
    sw = None
    try:
        sw = H.Sweep([(r.s, *r.f) for r in a], grouped=True, tune=True)
        sw.steps(3)
        for r in b:
            r.s.time_stepping(*r.f, 3)
        for r in (a[1], b[1]):
            H.tofp(r.f[2])
            H.toff(r.f[2])
        sw.steps(2)
        for r in b:
            r.s.time_stepping(*r.f, 2)
        for ra, rb in zip(a, b):
            assert same_state(ra.state(), rb.state())
    finally:
        if sw:
            sw.close()
        for r in a + b:
            r.close()


def test_sweep_rejects_bad_runs():
    r1, r2 = Run(SMALL_RUNS[0]), Run(SMALL_RUNS[1])
    try:
        with pytest.raises(H.MlegsError):
            H.Sweep([(r1.s, *r1.f), (r1.s, *r1.f)])  # one run per context
        with pytest.raises(H.MlegsError):
            H.Sweep([(r1.s, *r1.f), (r2.s, r2.f[0], r1.f[1], r2.f[2])])  # a field of another run
        with pytest.raises(H.MlegsError):
            H.Sweep([(r1.s, *r1.f), (r2.s, *r2.f)], by_grid=True)  # per-grid launches need streams mode
        r2.g.legendre_precision = "split32"
        with pytest.raises(H.MlegsError):
            H.Sweep([(r1.s, *r1.f), (r2.s, *r2.f)], grouped=True)  # grouping needs FP64 products
        sw = H.Sweep([(r1.s, *r1.f), (r2.s, *r2.f)], grouped=False)
        sw.steps(2)
        sw.close()
    finally:
        r1.close()
        r2.close()


def test_tuner_reports_every_bucket():
    g = H.Grid(dict(CFG1, nr=200, nth=256, nx=8, nrchop=140, ntchop=120, nxchop=3))
    try:
        rep = H.tune_rtran_bwd(g, 3)
        # NRCHOPS(m) = 140 - m over m < 120: buckets K > 128, 65..128, 33..64, <= 32
        assert sorted(rep) == [0, 1, 2, 3]
        for b, (best, us) in rep.items():
            assert best in H.BWD_TILE_SHAPES
            assert all(math.isfinite(t) and t > 0 for t in us.values())
            assert us[best] == min(us.values())
    finally:
        g.close()


def bits(state):
    return [np.ascontiguousarray(a).view(np.int64) for a, _, _ in state]


def test_cfg4_divergence_matches_oracle_and_sweep_keeps_stepping():
    """Run 3 of the uneven sweep (192x32x32, ELL=0.65) fails the per-step
    DIAGNOST divergence check (code 12) at step 6 at dt=1e-3: the FP64 oracle
    at the same step as the GPU.  In a sweep the failing run is reported
    (mlg_sweep_codes) and keeps stepping with the others, bit for bit as
    when stepped alone."""
    import _oracle as O

    cfg = cfg4_runs()[3]
    o = O.OGrid(cfg, threads=8)
    try:
        psi, chi = o.field(O.FFF), o.field(O.FFF)
        o.call("qvortex", LAMB_OSEEN_Q, psi, chi)
        psi.set(psi.arr() + seeded_noise(cfg, 1), O.FFF, psi.ln)
        chi.set(seeded_noise(cfg, 2), O.FFF)
        b = o.field(O.FFF, np.asfortranarray(seeded_noise(cfg, 3)))
        s = O.OSolver(o)
        s.initialize(psi, chi, b)
        ostep = next(st for st in range(2, 20) if s.step(psi, chi, b) == 12)
    finally:
        o.close()
    r = Run(cfg)
    try:
        gstep = None
        for st in range(2, 20):
            try:
                r.s.time_step(*r.f)
            except H.MlegsError as e:
                assert e.code == 12
                gstep = st
                break
    finally:
        r.close()
    assert ostep == gstep == 6

    cfgs = [cfg4_runs()[0], cfg]
    a = [Run(c) for c in cfgs]
    b = [Run(c) for c in cfgs]
    sw = None
    try:
        sw = H.Sweep([(x.s, *x.f) for x in a], grouped=True, tune=True)
        codes = sw.steps(9, check=False)
        assert codes == [0, 12]
        for x in b:
            with pytest.raises(H.MlegsError) if x is b[1] else _null():
                x.s.time_stepping(*x.f, 9, graph=True)
        for x, y in zip(a, b):
            assert all(np.array_equal(p, q) for p, q in zip(bits(x.state()), bits(y.state())))
        with pytest.raises(H.MlegsError):
            sw.steps(1)  # check=True raises the failing run's code
    finally:
        if sw:
            sw.close()
        for x in a + b:
            x.close()


class _null:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


@pytest.mark.timeout(300)
def test_cfg4_sweep_three_steps():
    """The 48 runs of the uneven sweep, grouped and tuned: three steps equal
    every run stepped alone, bit for bit."""
    info = check_sweep(cfg4_runs(), [3], grouped=True, tune=True)
    assert info["runs"] == 48
