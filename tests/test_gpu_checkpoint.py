"""Device fields saved in MPISAVEX format and loaded back (one process)
are bit-identical, space and LN included; files from a 2x2-process run
load into the same array."""
import numpy as np
import pytest

import _oracle as O
from mlegs_amd import checkpoint as ck
from mlegs_amd import hip as H
from test_gpu_parity import SMALL, Pair, state

pytestmark = pytest.mark.gpu


def test_save_load_round_trip(tmp_path):
    p = Pair(SMALL)
    try:
        st = state(p)
        a = p.g.scalar(O.FFF, *st[0])
        fn = str(tmp_path / "psi")
        ck.save(a, fn)
        b = p.g.scalar()
        info = ck.load(fn, b)
        assert info["space"] == O.FFF
        xa, sa, la = a.download()
        xb, sb, lb = b.download()
        assert np.array_equal(xa, xb) and sa == sb and la == lb
        # a PPP field
        H.tofp(a)
        ck.save(a, fn + "p")
        ck.load(fn + "p", b)
        assert np.array_equal(a.download()[0], b.download()[0]) and b.space == O.PPP
        # the same field written by a 2x2-process run
        info = ck.read_info(fn + ".info")
        ck.write_raw(fn + "_4", info, xa, (2, 2))
        ck.write_info(fn + "_4.info", (info["ndimr"], info["ndimth"], info["ndimx"], info["nrchopdim"],
                                       info["ntchopdim"], info["nxchopdim"]), info["space"], info["ln"],
                      info["zlen"], info["ell"])
        ck.load(fn + "_4", b, procs=(2, 2))
        assert np.array_equal(b.download()[0], xa)
    finally:
        p.close()
