"""The drop-in boundary on the GPU: field files through the C ABI
(MPISAVEX / MPILOADX, src/mod_fft.f90:2304-2642) against the format's
Python implementation (mlegs_amd/checkpoint.py, pinned byte for byte
against Fortran-written files), and a run of tests/fortran/
drive_bsnsq_new.f90 -- the call sequence of the reference's bsnsq_new over
the drop-in modules -- against the oracle: read.input set-up, MLOAD of the
initial fields, INITIALIZE, TIME_STEPPING to TIM%LIMIT with the OUTPUTS
snapshot, especData energies and INSPECT output."""
import math
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from mlegs_amd import checkpoint as ck
from mlegs_amd import hip as H
from test_gpu_parity import SMALL, Pair, oracle_trajectory, relerr, state

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mlegs_amd", "fortran", "_build", "drive_bsnsq_new")


@pytest.mark.parametrize("space", [O.FFF, O.PFF, O.PPP])
def test_field_save_load_against_python_format(tmp_path, space):
    p = Pair(SMALL)
    try:
        rng = np.random.default_rng(3)
        shape = p.g.shape(space)
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        g = p.g.scalar(space, a, 0.75)
        fn = str(tmp_path / "f.dat")
        g.save(fn)
        info = ck.read_info(fn + ".info")
        assert info["space"] == space and info["ln"] == 0.75
        assert np.array_equal(ck.read_raw(fn, info), a)
        # files of a run on 2 x 3 processes, written by the Python layout
        fn2 = str(tmp_path / "g.dat")
        info2 = dict(info)
        ck.write_info(fn2 + ".info", (info["ndimr"], info["ndimth"], info["ndimx"], info["nrchopdim"],
                                      info["ntchopdim"], info["nxchopdim"]), space, -1.25, info["zlen"], info["ell"])
        b = np.asfortranarray(a * (0.5 - 0.25j))
        ck.write_raw(fn2, info2, b, procs=(2, 3))
        h = p.g.scalar(O.FFF).load(fn2, procs=(2, 3))
        ha, hs, hl = h.download()
        assert hs == space and hl == -1.25
        assert np.array_equal(ha, b)
    finally:
        p.close()


def test_field_load_rejects_another_grid(tmp_path):
    p, q = Pair(SMALL), Pair(dict(SMALL, nrchop=20))
    try:
        fn = str(tmp_path / "f.dat")
        p.g.scalar(O.FFF, noise(p), 0.0).save(fn)
        with pytest.raises(H.MlegsError) as ei:
            q.g.scalar(O.FFF).load(fn)
        assert ei.value.code == 32
    finally:
        p.close()
        q.close()


def noise(p):
    from mlegs_amd.configs import seeded_noise
    return seeded_noise(p.cfg, 5, 1.0)


READ_INPUT = """DIMENSION
#     NR     NTH      NX
     {nr}     {nth}     {nx}
CHOPS
     {nrchop}     {ntchop}     {nxchop}
ELL
      {ell}
ZLEN
   {zlen:.15f}
MINC
       1
MKLINK
       0
DT
#     DT   LIMIT
      {dt}  {limit}
VISCOSITY
       1  {nu}  0.0       0
BOUSSINESQ
      {bv0}     {omega}  {kappa}   0.0       1
DIAGNOST
       2      40
SAVEDIR
       ./out/
PSI0
        psi0.dat
CHI0
        chi0.dat
B0
          b0.dat
OUTPUTS
       1
      {limit}   psi_end.dat   chi_end.dat   b_end.dat
"""


def test_bsnsq_new_call_sequence_against_oracle(tmp_path):
    assert os.path.exists(DRIVER), "build with `make -C mlegs_amd/fortran`"
    cfg = SMALL
    nsteps = 3
    limit = 3.5 * cfg["dt"]  # int(LIMIT/DT) = 3 steps after the start
    (tmp_path / "read.input").write_text(READ_INPUT.format(limit=limit, **cfg))
    out = tmp_path / "out"
    out.mkdir()
    p = Pair(cfg)
    try:
        st = state(p)
        for (a, ln), nm in zip(st, ("psi0.dat", "chi0.dat", "b0.dat")):
            info = dict(ndimr=p.g.ndimr, ndimth=p.g.ndimth, ndimx=p.g.ndimx, nrchopdim=p.g.nrchopdim,
                        ntchopdim=p.g.ntchopdim, nxchopdim=p.g.nxchopdim, space=O.FFF, ln=ln, zlen=cfg["zlen"],
                        ell=cfg["ell"])
            ck.write_info(str(out / nm) + ".info", (info["ndimr"], info["ndimth"], info["ndimx"], info["nrchopdim"],
                                                    info["ntchopdim"], info["nxchopdim"]), O.FFF, ln, cfg["zlen"],
                          cfg["ell"])
            ck.write_raw(str(out / nm), info, a)
        print(f"drive_bsnsq_new: {cfg['nr']}x{cfg['nth']}x{cfg['nx']}, {nsteps} steps", flush=True)
        r = subprocess.run([DRIVER], input="\n", capture_output=True, text=True, cwd=tmp_path, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "Finalized after" in r.stdout
        ref = oracle_trajectory(cfg, st, nsteps)
        rev = oracle_trajectory(cfg, st, nsteps, reverse=True)
        for nm, (ra, rl), (va, _) in zip(("psi", "chi", "b"), ref[nsteps], rev[nsteps]):
            info = ck.read_info(str(out / f"{nm}_end.dat") + ".info")
            ga = ck.read_raw(str(out / f"{nm}_end.dat"), info)
            e, s = relerr(ga, ra), relerr(va, ra)
            assert e <= max(1e-12, 6 * s), (nm, e, s)
            assert abs(info["ln"] - rl) <= 1e-12 * max(1.0, abs(rl))
        # INITIALIZE saved the initial fields under the same names
        info = ck.read_info(str(out / "psi0.dat") + ".info")
        assert np.array_equal(ck.read_raw(str(out / "psi0.dat"), info), st[0][0])
        # especData: before and after INITIALIZE, then every DIAGNOST IT1 = 2 steps (n = 2, 4)
        lines = (out / "especData_K.dat").read_text().splitlines()
        ke = [ln for ln in lines if ln.startswith(" KE:")]
        assert len(ke) == 4, lines
        oke, _, _ = p.o.energy(*[p.o.field(O.FFF, a, ln) for a, ln in st])
        total0 = sum(float(x) for x in ke[0].split()[2:])
        assert abs(total0 - oke.sum()) <= 1e-12 * abs(oke.sum())
        of = [p.o.field(O.FFF, a, ln) for a, ln in ref[nsteps]]
        okf, _, _ = p.o.energy(*of)
        totalf = sum(float(x) for x in ke[-1].split()[2:])
        assert abs(totalf - okf.sum()) <= 1e-10 * abs(okf.sum())
        # DIAGNOST's IT1 section at n = 0 (the driver's call before INITIALIZE)
        # and in TIME_STEPPING at n = 2, 4: printed, and CALCXY's line in coresize.dat
        assert r.stdout.count("<<< DIAGNOSTICS >>>") == 3 and "FRACTIONAL CHANGE=" in r.stdout
        core = [[float(x) for x in ln.split()] for ln in (out / "coresize.dat").read_text().splitlines()]
        assert len(core) == 3
        d0 = dict(zip(H.DIAGNOST_IT1_KEYS, p.o.diagnost_it1(*[p.o.field(O.FFF, a, ln) for a, ln in st[:2]])))
        assert core[0][0] == 0.0 and abs(core[0][3] - d0["r"]) <= 1e-5 * d0["r"]
        # INSPECT: three fields per step, each a time line and one k line
        ins = (out / "inspect.output").read_text().splitlines()
        assert len(ins) == 3 * nsteps * 2
        assert math.isclose(float(ins[-2]), (1 + nsteps) * cfg["dt"], rel_tol=1e-12)
    finally:
        p.close()
