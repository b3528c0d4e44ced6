"""Field files in the reference's formats (MPISAVEX/MPILOADX .info + raw
blocks, MSAVE/MLOADX).  The record layout is pinned against files written
and read by a Fortran compiler (amdflang) with the reference's WRITE
sequence; the process-block layout against DECOMPOSE's closed form."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mlegs_amd import checkpoint as ck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/llvm/bin/amdflang"


@pytest.fixture(scope="module")
def fortran_exe(tmp_path_factory):
    if not os.path.exists(FC):
        pytest.skip("amdflang not available")
    d = tmp_path_factory.mktemp("ftn")
    exe = d / "ckpt_info"
    subprocess.check_call([FC, "-O0", os.path.join(ROOT, "tests", "fortran", "ckpt_info.f90"), "-o", str(exe)],
                          cwd=d)
    return exe


def test_info_records_match_fortran(fortran_exe, tmp_path):
    ck.write_info(str(tmp_path / "p.info"), (131, 65, 64, 102, 22, 43), 3, 0.125, 6.283185307179586, 1.5, 1, 0)
    out = subprocess.check_output([str(fortran_exe), str(tmp_path)], text=True).split()
    assert [int(x) for x in out[:7]] == [131, 65, 64, 102, 22, 43, 3]
    assert [float(x) for x in out[7:10]] == [0.125, 6.283185307179586, 1.5]
    assert [int(x) for x in out[10:]] == [1, 0]
    info = ck.read_info(str(tmp_path / "f.info"))
    assert (info["ndimr"], info["ndimth"], info["ndimx"]) == (387, 257, 256)
    assert (info["nrchopdim"], info["ntchopdim"], info["nxchopdim"]) == (262, 86, 171)
    assert info["space"] == 1 and info["ln"] == -0.0795774715459476
    assert info["zlen"] == 12.566370614359172 and info["ell"] == 2.0 and (info["minc"], info["mklink"]) == (1, 0)
    # byte-identical when written back
    ck.write_info(str(tmp_path / "g.info"), (387, 257, 256, 262, 86, 171), 1, -0.0795774715459476,
                  12.566370614359172, 2.0, 1, 0)
    assert open(tmp_path / "g.info", "rb").read() == open(tmp_path / "f.info", "rb").read()


def test_msave_matches_fortran(fortran_exe, tmp_path):
    ck.write_info(str(tmp_path / "p.info"), (1, 1, 1, 1, 1, 1), 0, 0.0, 1.0, 1.0)
    subprocess.check_output([str(fortran_exe), str(tmp_path)])
    info, arr = ck.read_mload(str(tmp_path / "f.msave"))
    i, j, k = np.meshgrid(np.arange(1, 4), np.arange(1, 3), np.arange(1, 3), indexing="ij")
    assert np.array_equal(arr, (i + 10 * j + 100 * k) - 1j * i)
    assert info["space"] == 1 and info["ln"] == 0.5 and info["zlen"] == 6.0
    ck.write_msave(str(tmp_path / "g.msave"), info, arr)
    assert open(tmp_path / "g.msave", "rb").read() == open(tmp_path / "f.msave", "rb").read()


@pytest.mark.parametrize("n,p", [(171, 4), (86, 3), (387, 5), (7, 7), (10, 1)])
def test_decompose_closed_form(n, p):
    sizes, starts = zip(*[ck.decompose(n, p, r) for r in range(p)])
    assert sum(sizes) == n and max(sizes) - min(sizes) <= 1
    assert list(sizes) == sorted(sizes, reverse=True)       # the first MOD(N,P) ranks get one more
    assert list(starts) == list(np.cumsum((0,) + sizes[:-1]))


@pytest.mark.parametrize("space", [ck.FFF, ck.PFF, ck.PPP])
@pytest.mark.parametrize("procs", [(1, 1), (2, 3), (4, 1), (3, 2)])
def test_raw_blocks_round_trip_and_cover(tmp_path, space, procs):
    info = dict(ndimr=35, ndimth=17, ndimx=16, nrchopdim=28, ntchopdim=6, nxchopdim=9, space=space, ln=0.0)
    shape, _, _ = ck._split_axes(info, space)
    rng = np.random.default_rng(3)
    a = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    fn = str(tmp_path / "x")
    ck.write_raw(fn, info, a, procs)
    assert np.array_equal(ck.read_raw(fn, info, procs), a)
    seen = np.zeros(shape, dtype=int)
    offs = []
    for off, sl, loc in ck._blocks(info, space, *procs):
        seen[sl] += 1
        offs.append((off, int(np.prod(loc))))
    assert np.all(seen == 1)
    offs.sort()
    for (o1, n1), (o2, _) in zip(offs, offs[1:]):
        assert o1 + n1 <= o2                                   # blocks do not overlap
    if procs == (1, 1):
        assert open(fn, "rb").read() == a.ravel(order="F").tobytes()
