"""The Fortran binding (mlegs_amd/fortran/mlegs_hip.f90, ISO_C_BINDING) drives
the same library as the Python binding: a Fortran driver in the shape of the
reference's bsnsq_test runs INITIALIZE + TIME_STEPPING on a seeded state and
its result must equal the ctypes path bit for bit, and the oracle within the
step bar of test_gpu_parity."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from mlegs_amd import hip as H
from test_gpu_parity import SMALL, Pair, relerr, state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mlegs_amd", "fortran", "_build", "drive_bsnsq")

pytestmark = pytest.mark.gpu


def write_state(path, fields):
    with open(path, "wb") as f:
        for a, ln in fields:
            f.write(np.asarray(a, dtype=np.complex128).ravel(order="F").tobytes())
            f.write(np.float64(ln).tobytes())


def read_state(path, shape):
    n = int(np.prod(shape))
    raw = open(path, "rb").read()
    out, off = [], 0
    for _ in range(3):
        a = np.frombuffer(raw, dtype=np.complex128, count=n, offset=off).reshape(shape, order="F")
        off += 16 * n
        ln = float(np.frombuffer(raw, dtype=np.float64, count=1, offset=off)[0])
        off += 8
        out.append((a.copy(), ln))
    return out


def test_fortran_driver_matches_ctypes_and_oracle(tmp_path):
    assert os.path.exists(DRIVER), "build with `make -C mlegs_amd/fortran`"
    nsteps = 4
    p = Pair(SMALL)
    try:
        st = state(p)
        fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
        write_state(fin, st)
        r = subprocess.run([DRIVER, str(fin), str(fout), str(nsteps)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        fort = read_state(fout, p.o.shape(O.FFF))
        gf = [p.g.scalar(O.FFF, a, ln) for a, ln in st]
        s = H.Solver(p.g)
        s.initialize(*gf)
        s.time_stepping(*gf, nsteps)
        of = [p.o.field(O.FFF, a, ln) for a, ln in st]
        osv = O.OSolver(p.o)
        osv.initialize(*of)
        for _ in range(nsteps):
            assert osv.step(*of) == 0
        for (fa, fl), g, o in zip(fort, gf, of):
            ga, _, gl = g.download()
            assert np.array_equal(fa, ga) and fl == gl
            assert relerr(fa, o.arr()) <= 1e-11
    finally:
        p.close()
