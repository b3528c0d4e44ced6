"""mlg_vel2pln_noalloc (VEL2PLN with its scratch in a fixed per-GPU array)
against mlg_vel2pln, through the C ABI.

Both run the same kernels; only where the NRCHOP projections live differs,
so results must agree bit for bit.  Two calls in a row share the GPU's one
array, the second on a smaller grid (fewer projections), so stale entries
of the first call must not reach the second.  NRCHOP above the array's 1024
entries is refused with MLG_E_SIZE before anything is launched.
"""
import numpy as np
import pytest

from mlegs_amd import hip as H
from mlegs_amd.configs import CFG1

pytestmark = pytest.mark.gpu

LARGE = dict(CFG1)  # 64x32x32, NRCHOP 48
SMALL = dict(CFG1, nr=32, nth=16, nx=16, nrchop=22, ntchop=6, nxchop=5)
# NRCHOP at the array's limit, then 3 beyond it
AT_LIMIT = dict(CFG1, nr=1032, nth=8, nx=8, nrchop=1024, ntchop=3, nxchop=3)


def pff_data(g, seed):
    rng = np.random.default_rng(seed)
    shape = g.shape(H.PFF)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def compare_once(cfg, seed, ln):
    g = H.Grid(cfg)
    fields = []
    try:
        data = pff_data(g, seed)
        ra, rb = g.scalar(H.PFF, data), g.scalar(H.PFF, data)
        pa, pb = g.scalar(H.FFF, None, 0.3), g.scalar(H.FFF, None, -0.7)
        fields = [ra, rb, pa, pb]
        H.vel2pln(ra, pa, ln)
        H.vel2pln_noalloc(rb, pb, ln)
        g.synchronize()
        ea, sa, _ = ra.download()
        eb, sb, _ = rb.download()
        assert sa == sb == H.PFF
        assert np.array_equal(ea, eb), "r u_theta differs"
        assert not np.array_equal(ea, data), "VEL2PLN left r u_theta unchanged"
        la, lb = pa.ln, pb.ln
        assert la == lb, f"psi LN {la!r} vs {lb!r}"
        if ln is not None:
            assert la == ln
        assert pa.space == pb.space == H.FFF
        return la
    finally:
        for f in fields:
            f.close()
        g.close()


@pytest.mark.parametrize("ln", [None, 0.25])
def test_noalloc_matches_vel2pln_twice_smaller_second(ln):
    print(f"vel2pln_noalloc: 64x32x32 then 32x16x16, ln={ln}", flush=True)
    first = compare_once(LARGE, 11, ln)
    second = compare_once(SMALL, 12, ln)
    if ln is None:
        assert first != second


def test_noalloc_size_limit():
    print("vel2pln_noalloc: NRCHOP 1024, then 1027", flush=True)
    g = H.Grid(AT_LIMIT)
    try:
        data = pff_data(g, 5)
        ra, rb = g.scalar(H.PFF, data), g.scalar(H.PFF, data)
        pa, pb = g.scalar(H.FFF), g.scalar(H.FFF)
        H.vel2pln(ra, pa)
        H.vel2pln_noalloc(rb, pb)
        g.synchronize()
        assert np.array_equal(ra.e, rb.e)
        assert pa.ln == pb.ln
        g.chopset(3)
        try:
            with pytest.raises(H.MlegsError) as ei:
                H.vel2pln_noalloc(rb, pb)
            assert ei.value.code == 104
        finally:
            g.chopset(-3)
        for f in (ra, rb, pa, pb):
            f.close()
    finally:
        g.close()
