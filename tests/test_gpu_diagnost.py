"""GPU vs oracle: DIAGNOST's every-IT1 quantities (src/mod_march.f90:
1797-1961 with CALCXY :1967-2047; the oracle is pinned in
tests/test_oracle_diagnost.py)."""
import math

import numpy as np
import pytest

import _oracle as O
from mlegs_amd import hip as H
from test_gpu_parity import SMALL, Pair, state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg", [SMALL, dict(SMALL, nr=48, nrchop=40, ell=1.0, zlen=2 * math.pi)],
                         ids=["small", "nr48"])
def test_diagnost_it1_matches_oracle(cfg):
    p = Pair(cfg)
    try:
        st = state(p, 1e-2)
        g = [p.g.scalar(O.FFF, a, ln) for a, ln in st[:2]]
        o = [p.o.field(O.FFF, a, ln) for a, ln in st[:2]]
        d = H.diagnost_it1(*g)
        ref = dict(zip(H.DIAGNOST_IT1_KEYS, p.o.diagnost_it1(*o)))
        lines = [f"{k}: GPU {d[k]!r} oracle {ref[k]!r}" for k in H.DIAGNOST_IT1_KEYS]
        print("\n".join(lines))
        for k in H.DIAGNOST_IT1_KEYS:
            scale = max(abs(ref[k]), 1e-300)
            # spreads and centre coordinates are differences of O(1)
            # quantities: compared against the size of what they come from
            if k in ("psiat1_spread", "x", "y"):
                scale = max(scale, abs(ref["psiat1"]) if k == "psiat1_spread" else ref["r"])
            if k == "chiat1_spread":
                scale = max(scale, abs(ref["chiat1"]))
            assert abs(d[k] - ref[k]) <= 1e-11 * scale, "\n".join(lines)
        # the fields are not changed
        for gf, (a, ln) in zip(g, st[:2]):
            ga, _, gl = gf.download()
            assert np.array_equal(ga, a) and gl == ln
    finally:
        p.close()
