"""The reference's own shipped case (read.input: 408x204x608, NRCHOP 400,
NTCHOP 136, NXCHOP 304, ETD with the SW=3 exponential filter, P = 8) runs
on the GPU: its theta and z lengths are not powers of two, so every
transform takes the direct-DFT kernels.  Parameters come from
tests/golden/reference_read_input.json (make_reference_input.py, the
library's parser on the reference's file).  The case is too large for the
oracle; the DFT kernels are held to the oracle at the same lengths on small
grids in test_gpu_parity.py (THETA204, Z608).  Here: a Lamb-Oseen vortex
plus noise through the Richardson start and three steps (t = 0.4) under the
file's own run settings (REMOVE with RMV%SW = 2: the axial sponge in every
ETD update; REMOVE every 1000 and HYP3ADJ every 100 steps, not reached
here), every field finite, DIAGNOST silent, and the kinetic energy within
1e-3 of its start (measured without the sponge: +1.3e-4, the vortex
exchanging energy with the weak stratification, N = 0.1, while nu = 2e-7
and the filter acts on the top modes only).  Then one HYP3ADJ and one
REMOVE at this size, with finite results."""
import json
import os

import numpy as np
import pytest

from mlegs_amd import hip as H
from mlegs_amd.configs import LAMB_OSEEN_Q, seeded_noise

pytestmark = pytest.mark.gpu


def test_reference_shipped_case_runs():
    here = os.path.dirname(os.path.abspath(__file__))
    ref = json.load(open(os.path.join(here, "golden", "reference_read_input.json")))
    cfg = dict(ref["params"], noise=1e-6)
    g = H.Grid(cfg)
    try:
        psi, chi, b = g.scalar(), g.scalar(), g.scalar()
        H.qvortex(g, LAMB_OSEEN_Q, psi, chi)
        p0, _, ln = psi.download()
        psi.upload(p0 + seeded_noise(cfg, 1), H.FFF, ln)
        chi.upload(seeded_noise(cfg, 2), H.FFF, 0.0)
        b.upload(seeded_noise(cfg, 3), H.FFF, 0.0)
        ke0 = H.calc_boussi_energy(psi, chi, b)[0].sum()
        s = H.Solver(g)
        rs = ref["run_settings"]
        s.set_march_controls(rs["rmv_sw"], rs["rmv_int"], rs["adj_sw"], rs["adj_int"])
        s.initialize(psi, chi, b)
        s.time_stepping(psi, chi, b, 3)
        H.diagnost(psi, chi)
        for f in (psi, chi, b):
            a, _, fl = f.download()
            assert np.all(np.isfinite(a)) and np.isfinite(fl)
        ke = H.calc_boussi_energy(psi, chi, b)[0].sum()
        print(f"KE {ke0:.15e} -> {ke:.15e} after the start and 3 steps", flush=True)
        assert abs(ke - ke0) <= 1e-3 * abs(ke0)
        c = H.hyp3adj(g, psi, chi, b)
        print(f"HYP3ADJ: NUP_R/T/X, KAPPAP_R/T/X = {c}", flush=True)
        assert np.all(np.isfinite(c)) and np.all(c >= 0.8)
        H.remove(psi, chi, b)
        for f in (psi, chi, b):
            a, _, fl = f.download()
            assert np.all(np.isfinite(a)) and np.isfinite(fl)
        s.close()
    finally:
        g.close()
