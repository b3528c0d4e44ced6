"""The drop-in Fortran modules (mlegs_amd/fortran/mlegs_dropin.f90): the
reference's own driver f90/bsnsq_new.f90 compiles and links unchanged
against them and libmlegs_hip.so.  It is compiled where it lies (read only)
into a temporary directory and not run here (running needs a GPU; the GPU
tests run a driver with the same call sequence, tests/fortran/
drive_bsnsq_new.f90).  Nothing of the reference is kept."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/llvm/bin/amdflang"
BUILD = os.path.join(ROOT, "mlegs_amd", "fortran", "_build")
REF_DRIVER = "/root/reference/f90/bsnsq_new.f90"
LIB = os.path.join(ROOT, "mlegs_amd", "lib")


def test_dropin_modules_built():
    for m in ("mpi", "mod_misc", "mod_bandmat", "mod_eig", "mod_fd", "mod_lin_legendre", "mod_scalar3", "mod_fft",
              "mod_layout", "mod_legops", "mod_boussinesq", "mod_march", "mod_diagnostics", "mod_init", "mod_evp"):
        assert os.path.exists(os.path.join(BUILD, m + ".mod")), f"{m}.mod: build with `make -C mlegs_amd/fortran`"


@pytest.mark.skipif(not os.path.exists(REF_DRIVER) or not os.path.exists(FC), reason="reference driver or amdflang absent")
def test_reference_bsnsq_new_compiles_and_links(tmp_path):
    exe = tmp_path / "bsnsq_new"
    r = subprocess.run([FC, "-O1", "-fopenmp", "-I" + BUILD, "-module-dir", str(tmp_path), REF_DRIVER,
                        os.path.join(BUILD, "libmlegs_hip_f.a"), "-L" + LIB, "-lmlegs_hip", "-Wl,-rpath," + LIB,
                        "-o", str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = subprocess.check_output(["ldd", str(exe)], text=True)
    line = [l for l in out.splitlines() if "libmlegs_hip.so" in l]
    assert line and "not found" not in line[0]
    syms = subprocess.check_output(["nm", "-u", str(exe)], text=True)
    for s in ("mlg_field_load", "mlg_time_step", "mlg_diagnost", "mlg_energy", "mlg_read_input"):
        assert s in syms, s
