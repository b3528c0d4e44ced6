"""GPU vs oracle: vorticity removal, the axial sponge and the adaptive
hyperviscosity (src/mod_march.f90: RINGCUT, REMOVE, SPNGINIT, SPONGE,
CALC_BOUSSI_ETD_SPONGE, ENEMON, HYP3ADJ + HYP3SET) and TIME_STEPPING's
cadence for them.  The oracle's functions are pinned in
tests/test_oracle_remove.py."""
import numpy as np
import pytest

import _oracle as O
from mlegs_amd import hip as H
from test_gpu_parity import SMALL, SMALL_ETD, Pair, relerr, run_steps, state

pytestmark = pytest.mark.gpu

# AB2-CN with the exponential filter (VISC%SW = 3), whose coefficients
# HYP3ADJ adjusts
AB_FILTER = dict(SMALL, visc_sw=3, nup=1e-3, kappap=1e-3, visc_p=4)


@pytest.mark.parametrize("divxm2", [0, 1])
def test_ringcut_matches_oracle_bitwise(divxm2):
    p = Pair(SMALL)
    try:
        a = p.rand(O.PFF, 11 + divxm2)
        gs, of = p.both(a, O.PFF)
        nr = SMALL["nr"]
        ns = nr * 2 // 3
        ns0 = min(ns + 4, nr)
        H.ringcut(gs, ns0, ns, divxm2)
        p.o.call("ringcut", of, ns0, ns, divxm2)
        ga = gs.download()[0]
        assert np.array_equal(ga[:nr], of.arr()[:nr])
    finally:
        p.close()


@pytest.mark.parametrize("with_b", [True, False])
def test_remove_matches_oracle(with_b):
    p = Pair(SMALL)
    try:
        st = state(p, 1e-2)
        g = [p.g.scalar(O.FFF, a, ln) for a, ln in st]
        o = [p.o.field(O.FFF, a, ln) for a, ln in st]
        H.remove(g[0], g[1], g[2] if with_b else None)
        assert O.lib().og_remove(p.o.g, o[0].ptr, o[1].ptr, o[2].ptr if with_b else None) == 0
        for gf, of, nm in zip(g, o, ("psi", "chi", "b")):
            # chi passes through IDEL2, whose m = 0, k = 0 solve amplifies
            # round-off (measured 1e-11 of the field's maximum)
            p.check(gf, of, tol=1e-10, what=nm)
        if not with_b:
            assert np.array_equal(g[2].download()[0], st[2][0])
    finally:
        p.close()


def test_sponge_table_and_sponge_match_oracle():
    cfg = dict(SMALL, nx=32, nxchop=11)
    p = Pair(cfg)
    try:
        gs, os_ = H.Solver(p.g), O.OSolver(p.o)
        gs.set_march_controls(rmv_sw=2)
        os_.set_march_controls(rmv_sw=2)
        assert np.array_equal(p.g.sponge_table(), p.o.table("tophat_x"))
        st = state(p, 1e-2)
        g = [p.g.scalar(O.FFF, a, ln) for a, ln in st]
        o = [p.o.field(O.FFF, a, ln) for a, ln in st]
        for dt in (cfg["dt"], 0.5 * cfg["dt"], 0.3 * cfg["dt"]):
            H.sponge(g[0], g[1], g[2], dt)
            assert O.lib().og_sponge(p.o.g, o[0].ptr, o[1].ptr, o[2].ptr, dt) == 0
        for gf, of, nm in zip(g, o, ("psi", "chi", "b")):
            p.check(gf, of, tol=1e-12, what=nm)
        gs.close()
        os_.close()
    finally:
        p.close()


def test_enemon_matches_oracle():
    cfg = dict(SMALL, ntchop=8)
    p = Pair(cfg)
    try:
        st = state(p, 1e-2)
        g = [p.g.scalar(O.FFF, a, ln) for a, ln in st]
        o = [p.o.field(O.FFF, a, ln) for a, ln in st]
        # at NRCHOP+2 (HYP3ADJ's) and +3: DEL2 needs chi's top m = 1 row free
        for iof in (2, 3):
            p.g.chopset(iof)
            O.lib().og_chopset(p.o.g, iof)
            em, ek = H.enemon(g[0], g[1])
            oem, oek = p.o.enemon(o[0], o[1])
            p.g.chopset(-iof)
            O.lib().og_chopset(p.o.g, -iof)
            assert relerr(em, oem) <= 1e-12 and relerr(ek, oek) <= 1e-12
    finally:
        p.close()


def test_hyp3adj_matches_oracle():
    """Three calls on a changing state: the controllers' histories carry
    from call to call."""
    cfg = dict(AB_FILTER, ntchop=8, nth=16)
    p = Pair(cfg)
    try:
        st = state(p, 1e-2)
        g = [p.g.scalar(O.FFF, a, ln) for a, ln in st]
        o = [p.o.field(O.FFF, a, ln) for a, ln in st]
        assert np.array_equal(p.g.hyperv_coeffs, p.o.hyperv_coeffs())
        for call in range(3):
            c = H.hyp3adj(p.g, *g)
            assert p.o.call("hyp3adj", *o) == 0
            oc = p.o.hyperv_coeffs()
            assert np.abs(c - oc).max() <= 1e-12 * np.abs(oc).max(), f"call {call}: {c} vs {oc}"
            # the next call sees a rescaled state (the spectra's shapes stay)
            for gf, of in zip(g, o):
                a, _, ln = gf.download()
                gf.upload(a * 1.7, O.FFF, ln * 1.7)
                of.set(of.arr() * 1.7, O.FFF, of.ln * 1.7)
        # set: the filters of the given coefficients (HYP3SET), as in the oracle
        p.g.hyperv_coeffs = [0.9, 1.1, 1.3, 0.8, 0.85, 0.95]
        assert np.array_equal(p.g.hyperv_coeffs, [0.9, 1.1, 1.3, 0.8, 0.85, 0.95])
    finally:
        p.close()


def test_etd2_steps_with_axial_sponge():
    """ETD2 start and steps with RMV%SW = 2 (no REMOVE cadence): the sponge
    damps u_z and b after every ETD update."""
    run_steps(dict(SMALL_ETD, nx=32, nxchop=11), 4, controls=(2, 0, 0, 0))


def test_ab2_cn_steps_with_remove_and_hyp3adj():
    """AB2-CN with the SW=3 filter: REMOVE and the Richardson restart every 3
    steps, HYP3ADJ and HYP3SET every 2 (TIME_STEPPING's cadence)."""
    run_steps(AB_FILTER, 7, controls=(1, 3, 1, 2))


def test_etd2_steps_with_sponge_remove_and_hyp3adj():
    run_steps(dict(SMALL_ETD, nx=32, nxchop=11), 7, controls=(2, 3, 1, 2))


def test_cadence_under_graph_replays_equals_eager():
    """mlg_time_steps_graph stops its replays at every REMOVE / HYP3ADJ step
    and captures the step again when REMOVE moved b's storage: the same bits
    as eager stepping (two grids: HYP3ADJ's controller state is the
    context's)."""
    cfg = dict(SMALL_ETD, nx=32, nxchop=11)
    p = Pair(cfg)
    g2 = H.Grid(cfg)
    try:
        st = state(p)
        a = [p.g.scalar(O.FFF, x, ln) for x, ln in st]
        b = [g2.scalar(O.FFF, x, ln) for x, ln in st]
        s1, s2 = H.Solver(p.g), H.Solver(g2)
        for s in (s1, s2):
            s.set_march_controls(2, 3, 1, 2)
        s1.initialize(*a)
        s2.initialize(*b)
        s1.time_stepping(*a, 9)
        s2.time_stepping(*b, 9, graph=True)
        for x, y in zip(a, b):
            xa, _, xl = x.download()
            ya, _, yl = y.download()
            assert np.array_equal(xa, ya) and xl == yl
        assert np.array_equal(p.g.hyperv_coeffs, g2.hyperv_coeffs)
        assert s1.nstep == s2.nstep
        s1.close()
        s2.close()
    finally:
        g2.close()
        p.close()
