"""The split hi/lo FP32 Legendre products (mlg_set_legendre_precision) against
the FP64 oracle, through the C ABI.

Metric as in the FP64 parity tests: max |gpu - oracle| / max |oracle| per
field.  Single transforms on random data (every coefficient O(1), the least
favourable case for a scale-relative split) are held to 5e-10.  Steps
(the modes m < 5 keep FP64 products) are held to 1e-10 per field, or to the
FP64 parity bar (6 times the oracle's own change under reversed Legendre
summation) where that is larger: chi, the inverse Laplacian of a projected
field, has an order sensitivity of ~1e-9 on the NR=256 grid.
"""
import numpy as np
import pytest

import _oracle as O
from mlegs_amd import hip as H
from test_gpu_parity import SMALL, SMALL_ETD, TOL, Pair, oracle_trajectory, relerr, state

pytestmark = pytest.mark.gpu

SPLIT_TOL = 1e-10
RANDOM_TOL = 5e-10
# several K chunks in every product: 64 node pairs, 85 coefficient pairs
DEEP = dict(SMALL, nr=256, nrchop=170)


@pytest.fixture(params=["small", "deep"])
def P(request):
    p = Pair({"small": SMALL, "deep": DEEP}[request.param])
    p.g.legendre_precision = "split32"
    yield p
    p.close()


def check(P, g, o, what, tol=RANDOM_TOL):
    ga, gsp, gln = g.download()
    assert gsp == o.space
    e = relerr(P.region(ga, gsp), P.region(o.arr(), gsp))
    eln = abs(gln - o.ln) / max(1.0, abs(o.ln))
    print(f"{what}: split rel err {e:.2e}, LN {eln:.1e}", flush=True)
    assert e <= tol, f"{what}: rel err {e:.3e} > {tol:.1e}"
    assert eln <= tol
    return e


def test_precision_switch(P):
    assert P.g.legendre_precision == "split32"
    P.g.legendre_precision = "fp64"
    assert P.g.legendre_precision == "fp64"
    P.g.legendre_precision = "split32"
    with pytest.raises(ValueError):
        P.g.legendre_precision = "fp16"


@pytest.mark.parametrize("iof", [0, 3])
def test_split_rtran_backward(P, iof):
    P.g.chopset(iof)
    P.o.chopset(iof)
    try:
        g, o = P.both(P.rand(O.FFF, 21 + iof), O.FFF, ln=0.37)
        H.rtran(g, 1)
        P.o.call("rtran", o, 1)
        check(P, g, o, f"RTRAN backward (chop {iof})")
    finally:
        P.g.chopset(-iof)
        P.o.chopset(-iof)


def test_split_rtran_forward(P):
    g, o = P.both(P.rand(O.PFF, 25), O.PFF)
    H.rtran(g, -1)
    P.o.call("rtran", o, -1)
    check(P, g, o, "RTRAN forward")


def test_split_project_pff(P):
    fs = [P.both(P.rand(O.PFF, 40 + i), O.PFF) for i in range(3)]
    gpsi, opsi = P.both(P.noise(7), O.FFF, ln=0.11)
    gchi, ochi = P.empty()
    H.project(*[f[0] for f in fs], gpsi, gchi, 0.11)
    P.o.call("project", *[f[1] for f in fs], opsi, ochi, 1, 0.11)
    check(P, gpsi, opsi, "PROJECT psi")
    # del^2 chi is formed from k, m-weighted products: its scale is set by
    # the largest weighted product, so the split error is held relative to it
    check(P, gchi, ochi, "PROJECT del2 chi")


def split_steps(cfg, nsteps, amp=1e-4):
    """Split GPU steps vs oracle; returns per-step (psi, chi, b) errors."""
    p = Pair(cfg)
    try:
        p.g.legendre_precision = "split32"
        st = state(p, amp)
        ref = oracle_trajectory(cfg, st, nsteps)
        rev = oracle_trajectory(cfg, st, nsteps, reverse=True)
        gf = [p.g.scalar(O.FFF, a, ln) for a, ln in st]
        gs = H.Solver(p.g)
        gs.initialize(*gf)
        smax = [0.0, 0.0, 0.0]
        lines, worst = [], [0.0, 0.0, 0.0]
        for step in range(nsteps + 1):
            if step:
                gs.time_step(*gf)
            for f, (nm, g, (ra, rl), (va, _)) in enumerate(zip(("psi", "chi", "b"), gf, ref[step], rev[step])):
                ga, _, gl = g.download()
                e, sens = relerr(ga, ra), relerr(va, ra)
                smax[f] = max(smax[f], sens)
                eln = abs(gl - rl) / max(1.0, abs(rl))
                worst[f] = max(worst[f], e)
                lines.append(f"step {step} {nm}: split err {e:.2e} (oracle order sensitivity {sens:.2e}), LN {eln:.1e}")
                bar = max(SPLIT_TOL, 6 * smax[f])
                assert e <= bar, f"step {step} {nm}: err {e:.3e} > {bar:.1e}\n" + "\n".join(lines)
                assert eln <= max(SPLIT_TOL, 6 * smax[f]), f"step {step} {nm}: LN err {eln:.2e}\n" + "\n".join(lines)
        print("\n".join(lines), flush=True)
        gs.close()
        return worst
    finally:
        p.close()


@pytest.mark.parametrize("cfg", [SMALL, SMALL_ETD, dict(DEEP, adams=0, hyper_cn=1, nup=1e-3, kappap=1e-3, visc_p=4)],
                         ids=["ab2", "etd2", "etd2_deep"])
def test_split_steps(cfg):
    split_steps(cfg, 3)


def test_split_graph_equals_eager():
    """The captured step is captured again after a precision switch, and its
    replays equal eager split steps bit for bit."""
    p = Pair(SMALL_ETD)
    try:
        st = state(p)
        a = [p.g.scalar(O.FFF, x, ln) for x, ln in st]
        b = [p.g.scalar(O.FFF, x, ln) for x, ln in st]
        s1, s2 = H.Solver(p.g), H.Solver(p.g)
        s1.initialize(*a)
        s2.initialize(*b)
        s2.time_stepping(*b, 2, graph=True)  # captured in FP64
        s1.time_stepping(*a, 2)
        p.g.legendre_precision = "split32"
        s1.time_stepping(*a, 3)
        s2.time_stepping(*b, 3, graph=True)  # captured again, split
        for x, y in zip(a, b):
            xa, _, xl = x.download()
            ya, _, yl = y.download()
            assert np.array_equal(xa, ya) and xl == yl
        s1.close()
        s2.close()
    finally:
        p.close()


def test_fp64_unchanged_after_split_use():
    """Switching to split and back leaves the FP64 path's results bit for bit."""
    p = Pair(SMALL_ETD)
    try:
        st = state(p)
        a = [p.g.scalar(O.FFF, x, ln) for x, ln in st]
        s1 = H.Solver(p.g)
        s1.initialize(*a)
        s1.time_stepping(*a, 2)
        ref = [x.download() for x in a]
        p.g.legendre_precision = "split32"
        b = [p.g.scalar(O.FFF, x, ln) for x, ln in st]
        s2 = H.Solver(p.g)
        s2.initialize(*b)
        s2.time_stepping(*b, 1)
        p.g.legendre_precision = "fp64"
        c = [p.g.scalar(O.FFF, x, ln) for x, ln in st]
        s3 = H.Solver(p.g)
        s3.initialize(*c)
        s3.time_stepping(*c, 2)
        for (ra, _, rl), y in zip(ref, c):
            ya, _, yl = y.download()
            assert np.array_equal(ra, ya) and rl == yl
        for s in (s1, s2, s3):
            s.close()
    finally:
        p.close()
