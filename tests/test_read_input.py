"""read.input ingestion (mlg_read_input): READIN0's blocks and READCOM1's
line syntax (src/mod_init.f90:165-710), without a GPU."""
import math
import os

import pytest

from mlegs_amd import hip as H

SAMPLE = """DIMENSION
#     NR     NTH      NX
      64,     32	      32
CHOPS
# NRCHOP  NTCHOP  NXCHOP
      48      11      11
ELL
      1.D0
ZLEN
      0.5
! echo this line is a shell command in the reference and is skipped here
VORTEXPAIR
       2
       0.      0.      1.      0.      1.
       1.      0.     -1.      0.      1.
VORTEXPAIR_CROW
      0.      0.      0.       1
      0.      0.      0.       1
VORTEXPAIR_SECTION
      4.      0.      0.       1
      4.      0.      0.       1
DT
      1.0E-3  2.5
VISCOSITY
#     SW      NU     NUP       P
       1    1000.    0.        4
BOUSSINESQ
#     BV0    OMEGA    KAPPA   KAPPAP    ABSW
      2.0     0.5    1.d-3     0.        1
SAVEDIR
       ./output/
OUTPUTS
       2
         2.0        psi2.dat        chi2.dat        b2.dat
         4.0        psi4.dat        chi4.dat        b4.dat
MINC
       1
%END
DIMENSION
       8 8 8
"""


def test_sample_blocks_and_quirks(tmp_path):
    fn = tmp_path / "read.input"
    fn.write_text(SAMPLE)
    p, lim = H.read_input(str(fn))
    assert (p["nr"], p["nth"], p["nx"]) == (64, 32, 32)          # comma and tab separators
    assert (p["nrchop"], p["ntchop"], p["nxchop"]) == (48, 11, 11)
    assert p["ell"] == 1.0                                        # D exponent
    assert p["zlen"] == 2 * math.pi / 0.5                        # ZLEN <= 1 is a wavenumber
    assert p["dt"] == 1e-3 and lim == 2.5
    assert p["visc_sw"] == 1 and p["nu"] == 1e-3 and p["nup"] == 0.0 and p["visc_p"] == 4  # NU >= 1 is Re
    assert p["bv0"] == 2.0 and p["omega"] == 0.5 and p["kappa"] == 1e-3 and p["adams"] == 1
    assert p["minc"] == 1 and p["mklink"] == 0                   # blocks after OUTPUTS still read; %END stops


def test_reference_read_input_if_present():
    ref = "/root/reference/read.input"
    if not os.path.exists(ref):
        pytest.skip("reference tree not mounted")
    p, lim = H.read_input(ref)
    assert (p["nr"], p["nth"], p["nx"], p["nrchop"], p["ntchop"], p["nxchop"]) == (408, 204, 608, 400, 136, 304)
    assert p["ell"] == 8.0 and p["zlen"] == pytest.approx(18.849555921538759, rel=1e-15)
    assert p["visc_sw"] == 3 and p["nu"] == 2e-7 and p["nup"] == 1.0 and p["visc_p"] == 8
    assert p["bv0"] == 0.1 and p["omega"] == 0.0 and p["kappa"] == 2e-7 and p["adams"] == 0
    assert p["dt"] == 0.1 and lim == 300.0


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(H.MlegsError):
        H.read_input(str(tmp_path / "nope"))


def test_run_settings(tmp_path):
    """READIN0's name, OUTPUTS, DIAGNOST, REMOVE, NUP_ADJUST and MONITOR_MK
    blocks (src/mod_init.f90:551-660) and its defaults (:428-447)."""
    f = tmp_path / "read.input"
    f.write_text("DIMENSION\n 32 16 16\nSAVEDIR\n  ./run/\nPSI0\n  p0.dat\nOUTPUTS\n 2\n# t psi chi b\n 1.5 a1 b1 c1\n"
                 " 3.0, 'a 2', b2, c2\nDIAGNOST\n 7 9\nNUP_ADJUST\n 1 50\nMONITOR_MK\n 2\n 1 -3\n 0 4\n 5. 10.\n%END\n")
    s = H.read_run_settings(str(f))
    assert s["savedir"] == "./run/" and s["psi0"] == "p0.dat" and s["chi0"] == ""
    assert s["outputs"] == [(1.5, "a1", "b1", "c1"), (3.0, "a 2", "b2", "c2")]
    assert (s["it1"], s["it2"], s["adj_sw"], s["adj_int"]) == (7, 9, 1, 50)
    assert (s["rmv_sw"], s["rmv_int"]) == (0, 10000)                    # defaults
    assert s["monitor_mk"] == [(1, -3), (0, 4)] and (s["monitor_start"], s["monitor_interval"]) == (5, 10)
    d = tmp_path / "empty.input"
    d.write_text("DIMENSION\n 8 8 8\n")
    s = H.read_run_settings(str(d))
    assert (s["it1"], s["it2"], s["outputs"], s["monitor_mk"]) == (40, 40, [], [])


def test_reference_run_settings_if_present():
    ref = "/root/reference/read.input"
    if not os.path.exists(ref):
        pytest.skip("reference tree not mounted")
    s = H.read_run_settings(ref)
    assert s["savedir"] == "./output/" and (s["psi0"], s["chi0"], s["b0"]) == ("psi0.dat", "chi0.dat", "b0.dat")
    assert len(s["outputs"]) == 150 and s["outputs"][0] == (2.0, "psi2.dat", "chi2.dat", "b2.dat")
    assert (s["rmv_sw"], s["rmv_int"], s["adj_sw"], s["adj_int"]) == (2, 1000, 1, 100)
    assert s["monitor_mk"] == [(2, -10), (1, 12), (3, 2), (0, 0)]
