"""Pins the oracle's grid and basis tables against independent sources.

References restated: GAUSS_LEGENDRE/LEG_ZERO (src/mod_lin_legendre.f90:79-175),
LEG_LOG_NORMT (:306-377), LEG_TBL3 (:526-644), LEGINIT (src/mod_fft.f90:180-367).
Independent checks: numpy's Gauss-Legendre rule, mpmath's associated Legendre
functions at 40 digits, quadrature orthonormality, parity, closed-form norms.
"""
import math

import mpmath as mp
import numpy as np
import pytest

import _oracle as O
from mlegs_amd.configs import CFG1


def tiny_cfg(**kw):
    c = dict(CFG1)
    c.update(kw)
    return c


@pytest.fixture(scope="module")
def g1():
    g = O.OGrid(CFG1)
    yield g
    g.close()


@pytest.fixture(scope="module")
def gbig():
    # production radial/azimuthal sizes (384 nodes, m <= 85, n < 263)
    g = O.OGrid(tiny_cfg(nr=384, nth=128, nx=8, nrchop=256, ntchop=86, nxchop=3, ell=2.0))
    yield g
    g.close()


@pytest.mark.parametrize("n", [16, 64, 384])
def test_gauss_legendre_matches_numpy(n):
    g = O.OGrid(tiny_cfg(nr=n, nth=8, nx=8, nrchop=min(12, n - 6), ntchop=4, nxchop=3))
    x, w = g.table("x"), g.table("w")
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 4e-16
    # numpy's own weights carry up to ~3e-10 relative error at n=384, so the
    # exact reference is mpmath at 40 digits on a sample of nodes; the
    # reference's stopping rule (weight from the previous Newton iterate)
    # leaves ~3e-12 at the outermost node of n=384
    assert np.max(np.abs(w - wr) / wr) < 1e-9
    mp.mp.dps = 40
    for i in range(0, n // 2, max(1, n // 16)):
        z = mp.findroot(lambda t: mp.legendre(n, t), mp.mpf(xr[i]))
        dp = mp.diff(lambda t: mp.legendre(n, t), z)
        we = 2 / ((1 - z ** 2) * dp ** 2)
        assert abs(x[i] - float(z)) < 3e-16
        assert abs(w[i] - float(we)) / float(we) < 1e-11
    assert abs(w.sum() - 2.0) < 1e-13
    # exact for polynomials of degree <= 2n-1
    for d in (0, 2, 2 * n - 2):
        assert abs(np.sum(w * x ** d) - 2.0 / (d + 1)) < 1e-12
    r = g.table("r")
    assert np.allclose(r, g.cfg["ell"] * np.sqrt((1 + x) / (1 - x)), rtol=1e-15)
    g.close()


def test_lognorm_closed_form(g1):
    ln = g1.table("lognorm").reshape(g1.nnorm, g1.ntchopdim, order="F")
    for mm in range(g1.ntchopdim):
        for nn in range(g1.nnorm):
            n = mm + nn
            exact = 0.5 * (math.log((2 * n + 1) / 2.0) + math.lgamma(n - mm + 1) - math.lgamma(n + mm + 1))
            assert abs(ln[nn, mm] - exact) < 1e-12 * max(1.0, abs(exact))


def _pbar_mp(n, m, x):
    norm = mp.sqrt(mp.mpf(2 * n + 1) / 2 * mp.factorial(n - m) / mp.factorial(n + m))
    return norm * mp.legenp(n, m, x)


def _pbar_exact(n, m, x):
    """Orthonormal P_n^m(x) (Condon-Shortley phase) from the explicit finite
    sum P_n(x) = 2^-n sum_k (-1)^k C(n,k) C(2n-2k,n) x^(n-2k), differentiated
    m times term by term; 400-digit arithmetic absorbs the cancellation."""
    with mp.workdps(400):
        x = mp.mpf(x)
        s = mp.mpf(0)
        for k in range(0, n // 2 + 1):
            p = n - 2 * k
            if p < m:
                continue
            s += (-1) ** k * mp.binomial(n, k) * mp.binomial(2 * n - 2 * k, n) * mp.factorial(p) / mp.factorial(p - m) * x ** (p - m)
        val = (-1) ** m * (1 - x * x) ** (mp.mpf(m) / 2) * s / mp.mpf(2) ** n
        norm = mp.sqrt(mp.mpf(2 * n + 1) / 2 * mp.factorial(n - m) / mp.factorial(n + m))
        return +(norm * val)


def test_pbar_table_vs_mpmath_small(g1):
    mp.mp.dps = 40
    pf = g1.pf()
    x = g1.table("x")
    for mm in (0, 1, 3, 10):
        for nn in (0, 1, 2, 7, 20, 37):
            if nn >= g1.npf:
                continue
            for i in (0, 5, g1.nrh - 1):
                ref = float(_pbar_mp(mm + nn, mm, mp.mpf(x[i])))
                assert abs(pf[i, nn, mm] - ref) <= 2e-13 * abs(ref) + 1e-300


def test_pbar_table_vs_mpmath_production(gbig):
    """m up to 85 at 384 nodes, including entries near |x| -> 1 that need
    extended precision (P_n^m ~ (1-x^2)^{m/2})."""
    mp.mp.dps = 60
    pf = gbig.pf()
    x = gbig.table("x")
    worst = 0.0
    for mm in (0, 17, 60, 85):
        for nn in (0, 1, 50, 170, 262 - mm if 262 - mm < gbig.npf else gbig.npf - 1):
            for i in (0, 3, 60, gbig.nrh - 1):
                ref = _pbar_exact(mm + nn, mm, x[i])
                v = pf[i, nn, mm]
                if abs(ref) < 1e-300:
                    assert v == 0.0 or abs(v) < 1e-290
                    continue
                rel = abs((v - float(ref)) / float(ref))
                worst = max(worst, rel)
    assert worst < 5e-12, worst


def test_pbar_orthonormal(gbig):
    """sum_i w_i Pbar_n^m(x_i) Pbar_n'^m(x_i) = delta over the full node set
    (lower half from the table, upper half by parity (-1)^(n-m))."""
    pf = gbig.pf()
    w = gbig.table("w")[: gbig.nrh]
    for mm in (0, 1, 40, 85):
        nmax = 256 - mm
        P = pf[:, :nmax, mm]
        sgn = np.array([1.0 if (k % 2 == 0) else -1.0 for k in range(nmax)])
        G = (P * w[:, None]).T @ P
        G = G + (G * np.outer(sgn, sgn))
        err = np.max(np.abs(G - np.eye(nmax)))
        assert err < 5e-12, (mm, err)


def test_value_at_ends(g1):
    at1, at0 = g1.table("at1"), g1.table("at0")
    n = np.arange(len(at1))
    assert np.allclose(at1, np.sqrt((2 * n + 1) / 2.0), rtol=1e-13)
    assert np.allclose(at0, (-1.0) ** n * np.sqrt((2 * n + 1) / 2.0), rtol=1e-13)


def test_axial_wavenumbers(g1):
    ak = g1.table("ak").reshape(g1.ntchopdim, g1.nxchopdim, order="F")
    nxc = CFG1["nxchop"]
    k = np.array([kk if kk < nxc else kk - g1.nxchopdim for kk in range(g1.nxchopdim)])
    assert np.allclose(ak[0], 2 * np.pi / CFG1["zlen"] * k)


def test_etd_operator_is_expm():
    """E = exp(L dt), Phi = L^{-1}(exp(L dt) - I) for the rotating-stratified
    linear operator on (r u_r, r u_theta, u_z, b)."""
    from scipy.linalg import expm

    cfg = tiny_cfg(omega=0.7, bv0=1.3)
    g = O.OGrid(cfg)
    dt = 0.05
    E, PHI = g.etd_op(dt)
    Om, N = cfg["omega"], cfg["bv0"]
    L = np.array([[0, 2 * Om, 0, 0], [-2 * Om, 0, 0, 0], [0, 0, 0, -1.0], [0, 0, N * N, 0]])
    assert np.allclose(E, expm(L * dt), atol=1e-14)
    assert np.allclose(PHI, np.linalg.solve(L, expm(L * dt) - np.eye(4)), atol=1e-14)
    g.close()


def test_hyperviscosity_filter_special_values():
    """HYP3SET/HYP3MASK (src/mod_march.f90:3320-3480): H = exp(-nu_p mask
    (K^2/K^2_max)^(P/2)) with a tanh ramp mask that is 0 below the ramp
    start, 1/2 at mid-ramp and 1 at the top wavenumber (ramp starts: 1/2 of
    the top radial index, 2/3 of the top m and |k|)."""
    cfg = dict(nr=32, nth=32, nx=32, nrchop=21, ntchop=13, nxchop=13, minc=1, mklink=0, ell=1.0,
               zlen=2 * np.pi, nu=1e-3, kappa=1e-3, nup=0.3, kappap=0.2, visc_sw=3, visc_p=4,
               omega=0.5, bv0=1.0, adams=1, hyper_cn=0, dt=1e-3)
    g = O.OGrid(cfg)
    try:
        nu, ph = 0.3, 2
        ht, hx = g.table("hv_t"), g.table("hv_x")
        hr = g.table("hv_r").reshape(21, 13, order="F")
        # azimuthal: m = 0..12, top 12, ramp from 8, midpoint 10
        assert np.all(ht[:9] == 1.0)
        assert ht[12] == pytest.approx(np.exp(-nu), rel=1e-14)
        assert ht[10] == pytest.approx(np.exp(-nu * 0.5 * (100 / 144) ** ph), rel=1e-14)
        assert np.all(np.diff(ht) <= 0)
        # axial: |k| = 0..12 (both signs)
        kk = np.arange(25)
        kabs = np.where(kk < 13, kk, 25 - kk)
        for k in range(25):
            if kabs[k] <= 8:
                assert hx[k] == 1.0
        assert hx[12] == pytest.approx(np.exp(-nu), rel=1e-14) and hx[13] == hx[12]
        assert hx[10] == pytest.approx(np.exp(-nu * 0.5 * (100 / 144) ** ph), rel=1e-14)
        # radial at m = 0: i = 0..20, top 20, ramp from 10, midpoint 15; K^2 ~ i(i+1)
        assert np.all(hr[:10, 0] == 1.0)
        assert hr[20, 0] == pytest.approx(np.exp(-nu), rel=1e-14)
        assert hr[15, 0] == pytest.approx(np.exp(-nu * 0.5 * (15 * 16 / (20 * 21)) ** ph), rel=1e-14)
        # diffusivity filter uses kappa_p
        hd = g.table("hd_r").reshape(21, 13, order="F")
        assert hd[20, 0] == pytest.approx(np.exp(-0.2), rel=1e-14)
    finally:
        g.close()
