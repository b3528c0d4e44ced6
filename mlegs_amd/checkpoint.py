"""Field files in the reference's on-disk formats.

MPISAVEX / MPILOADX (src/mod_fft.f90:2304-2642): two files per field,

* ``FN.info`` -- Fortran unformatted sequential records (4-byte little-endian
  record markers around each record):
  ``NDIMR,NDIMTH,NDIMX`` / ``NRCHOPDIM,NTCHOPDIM,NXCHOPDIM`` / ``SPACE,LN`` /
  ``ZLEN,ELL`` / ``MINC,MKLINK`` (INTEGER = int32, REAL(P8) = float64);
* ``FN`` -- raw complex128 coefficients.  A run on N1 x N2 processes writes
  one block per process at offset ``INDEX*ARRAYSIZE`` (elements), each
  block the process's local array E(...) in column-major order; local
  extents follow DECOMPOSE (src/mod_fft.f90:1400-1431) over the two
  sub-communicators SUBCOMM_1, SUBCOMM_2 (src/mod_fft.f90:333-342).  One
  process (the GPU) writes the whole array as one block.

MSAVE / MLOADX (src/mod_fft.f90:2420-2490): one unformatted file holding the
same five records, then ``SIZE1,SIZE2,SIZE3`` and the array as one record.

``save``/``load`` move a device field (mlegs_amd.hip.Scalar) to and from
these files; everything else here is host-side byte work.
"""
import math
import os
import struct

import numpy as np

PPP, FFF, PFP, PFF = 0, 1, 2, 3


# ---------------------------------------------------------------- Fortran records
def _write_records(f, records):
    for payload in records:
        n = len(payload)
        f.write(struct.pack("<i", n))
        f.write(payload)
        f.write(struct.pack("<i", n))


def _read_record(f):
    """One logical record; gfortran/flang split records over 2 GiB into
    sub-records whose leading marker is negative while more follow."""
    out = b""
    while True:
        head = f.read(4)
        if len(head) < 4:
            raise EOFError("truncated Fortran record")
        n = struct.unpack("<i", head)[0]
        more = n < 0
        data = f.read(abs(n))
        tail = struct.unpack("<i", f.read(4))[0]
        if abs(tail) != abs(n):
            raise ValueError("inconsistent Fortran record markers")
        out += data
        if not more:
            return out


def _info_records(dims, space, ln, zlen, ell, minc, mklink):
    ndimr, ndimth, ndimx, nrchopdim, ntchopdim, nxchopdim = dims
    return [struct.pack("<3i", ndimr, ndimth, ndimx),
            struct.pack("<3i", nrchopdim, ntchopdim, nxchopdim),
            struct.pack("<id", space, ln),
            struct.pack("<2d", zlen, ell),
            struct.pack("<2i", minc, mklink)]


def write_info(fn, dims, space, ln, zlen, ell, minc=1, mklink=0):
    """FN.info of MPISAVEX (src/mod_fft.f90:2335-2340)."""
    with open(fn, "wb") as f:
        _write_records(f, _info_records(dims, space, ln, zlen, ell, minc, mklink))


def read_info(fn):
    with open(fn, "rb") as f:
        r = [_read_record(f) for _ in range(5)]
    ndimr, ndimth, ndimx = struct.unpack("<3i", r[0])
    nrchopdim, ntchopdim, nxchopdim = struct.unpack("<3i", r[1])
    space, ln = struct.unpack("<id", r[2])
    zlen, ell = struct.unpack("<2d", r[3])
    minc, mklink = struct.unpack("<2i", r[4])
    return dict(ndimr=ndimr, ndimth=ndimth, ndimx=ndimx, nrchopdim=nrchopdim, ntchopdim=ntchopdim,
                nxchopdim=nxchopdim, space=space, ln=ln, zlen=zlen, ell=ell, minc=minc, mklink=mklink)


# ---------------------------------------------------------------- process blocks
def decompose(nsize, nprocs, rank):
    """(local size, start index) of DECOMPOSE: the first MOD(N,P) ranks get one more."""
    q, r = divmod(nsize, nprocs)
    if r > rank:
        return q + 1, (q + 1) * rank
    return q, q * rank + r


def _split_axes(info, space):
    """For a space: full shape, and which axis SUBCOMM_1 / SUBCOMM_2 split."""
    if space == FFF:
        shape = (info["nrchopdim"], info["ntchopdim"], info["nxchopdim"])
        return shape, 2, 1      # NX over SUBCOMM_1, NT over SUBCOMM_2
    if space == PFF:
        shape = (info["ndimr"], info["ntchopdim"], info["nxchopdim"])
        return shape, 2, 1
    shape = (info["ndimr"], info["ndimth"], info["ndimx"])
    return shape, 0, 2          # NDIMR over SUBCOMM_1, NDIMX over SUBCOMM_2


def _blocks(info, space, n1, n2):
    """Yield (file element offset, slices into the full array, local shape)."""
    shape, ax1, ax2 = _split_axes(info, space)
    arraysize = 1
    for ax, n in enumerate(shape):
        arraysize *= math.ceil(n / n1) if ax == ax1 else math.ceil(n / n2) if ax == ax2 else n
    for p1 in range(n1):
        for p2 in range(n2):
            if info["nxchopdim"] < info["ntchopdim"]:
                index = p1 + p2 * n1
            else:
                index = p2 + p1 * n2
            sl, loc = [], []
            for ax, n in enumerate(shape):
                if ax == ax1:
                    sz, st = decompose(n, n1, p1)
                elif ax == ax2:
                    sz, st = decompose(n, n2, p2)
                else:
                    sz, st = n, 0
                sl.append(slice(st, st + sz))
                loc.append(sz)
            yield index * arraysize, tuple(sl), tuple(loc)


def write_raw(fn, info, arr, procs=(1, 1)):
    """FN of MPISAVEX for a run on procs=(N1, N2) processes."""
    n1, n2 = procs
    arr = np.asarray(arr, dtype=np.complex128)
    blocks = list(_blocks(info, info["space"], n1, n2))
    with open(fn, "wb") as f:
        for off, sl, loc in blocks:
            f.seek(off * 16)
            f.write(np.asfortranarray(arr[sl]).ravel(order="F").tobytes())


def read_raw(fn, info, procs=(1, 1)):
    """The full array from FN written by a run on procs=(N1, N2) processes."""
    n1, n2 = procs
    shape, _, _ = _split_axes(info, info["space"])
    out = np.zeros(shape, dtype=np.complex128, order="F")
    with open(fn, "rb") as f:
        for off, sl, loc in _blocks(info, info["space"], n1, n2):
            f.seek(off * 16)
            n = int(np.prod(loc))
            data = np.frombuffer(f.read(16 * n), dtype=np.complex128, count=n)
            out[sl] = data.reshape(loc, order="F")
    return out


# ---------------------------------------------------------------- MSAVE / MLOADX
def write_msave(fn, info, arr):
    arr = np.asarray(arr, dtype=np.complex128)
    recs = _info_records((info["ndimr"], info["ndimth"], info["ndimx"], info["nrchopdim"], info["ntchopdim"],
                          info["nxchopdim"]), info["space"], info["ln"], info["zlen"], info["ell"],
                         info.get("minc", 1), info.get("mklink", 0))
    recs.append(struct.pack("<3i", *arr.shape))
    recs.append(np.asfortranarray(arr).ravel(order="F").tobytes())
    with open(fn, "wb") as f:
        _write_records(f, recs)


def read_mload(fn):
    """MLOADX: (info dict, array)."""
    with open(fn, "rb") as f:
        r = [_read_record(f) for _ in range(7)]
    info = {}
    info["ndimr"], info["ndimth"], info["ndimx"] = struct.unpack("<3i", r[0])
    info["nrchopdim"], info["ntchopdim"], info["nxchopdim"] = struct.unpack("<3i", r[1])
    info["space"], info["ln"] = struct.unpack("<id", r[2])
    info["zlen"], info["ell"] = struct.unpack("<2d", r[3])
    info["minc"], info["mklink"] = struct.unpack("<2i", r[4])
    s1, s2, s3 = struct.unpack("<3i", r[5])
    arr = np.frombuffer(r[6], dtype=np.complex128, count=s1 * s2 * s3).reshape((s1, s2, s3), order="F").copy()
    return info, arr


# ---------------------------------------------------------------- device fields
def _grid_info(grid, space, ln):
    cfg = grid.cfg
    return dict(ndimr=grid.ndimr, ndimth=grid.ndimth, ndimx=grid.ndimx, nrchopdim=grid.nrchopdim,
                ntchopdim=grid.ntchopdim, nxchopdim=grid.nxchopdim, space=space, ln=ln, zlen=cfg["zlen"],
                ell=cfg["ell"], minc=cfg.get("minc", 1) or 1, mklink=cfg.get("mklink", 0))


def save(scalar, fn):
    """MPISAVEX(A, FN) of one process: FN.info and FN."""
    arr, space, ln = scalar.download()
    info = _grid_info(scalar.grid, space, ln)
    write_info(fn + ".info", (info["ndimr"], info["ndimth"], info["ndimx"], info["nrchopdim"], info["ntchopdim"],
                              info["nxchopdim"]), space, ln, info["zlen"], info["ell"], info["minc"], info["mklink"])
    write_raw(fn, info, arr)


def load(fn, scalar, procs=(1, 1)):
    """MPILOADX(FN, A): fills the device field from FN.info + FN written on
    procs=(N1, N2) processes.  Grid mismatches raise (the reference prints a
    warning and continues; a GPU field of the wrong shape cannot be filled)."""
    info = read_info(fn + ".info")
    g = scalar.grid
    mine = _grid_info(g, info["space"], info["ln"])
    for k in ("ndimr", "ndimth", "ndimx", "nrchopdim", "ntchopdim", "nxchopdim"):
        if info[k] != mine[k]:
            raise ValueError(f"{os.path.basename(fn)}: {k} {info[k]} in the file, {mine[k]} in the grid")
    arr = read_raw(fn, info, procs)
    scalar.upload(arr, info["space"], info["ln"])
    return info
