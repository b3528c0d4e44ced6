"""Writes adjoint_fd.json: FP64 CPU reference derivatives of the perturbation
kinetic energy J (sum over m >= 1 of CALC_BOUSSI_ENERGY's KE) after an
AB2-CN march (INITIALIZE, then STEP_BOUSSI_AB_CN), computed by the oracle
(oracle/mlegs_oracle.c, FP64 C) with central differences:

* "cfg5": the 128x64x64 workload (40 steps) from the Lamb-Oseen vortex plus
  seeded noise 1e-5 restricted to m = 1, along the 16 directions
  mlegs_amd.configs.fd_direction(cfg, i), e = 1e-6;
* "tiny": a 16x8x8 grid (NRCHOP 10, NTCHOP 3, NXCHOP 3, 6 steps) from the
  vortex plus seeded noise 1e-3 on every field, the full gradient: one
  central difference per real degree of freedom (every stored coefficient's
  real and imaginary part within the truncation, and psi%LN, chi%LN),
  e = 1e-7 times the state's largest coefficient.

Run from the repository root: python tests/golden/make_adjoint_fd.py
(a few minutes on 8 cores).
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

import _oracle as O  # noqa: E402
from mlegs_amd.configs import CFG5, LAMB_OSEEN_Q, fd_direction, seeded_noise  # noqa: E402

TINY = dict(CFG5, nr=16, nth=8, nx=8, nrchop=10, ntchop=3, nxchop=3, nsteps=6, noise=1e-3)


def initial(o, cfg, m_only):
    psi, chi = o.field(O.FFF), o.field(O.FFF)
    o.call("qvortex", LAMB_OSEEN_Q, psi, chi)
    amp = cfg["noise"]
    return [psi.arr() + seeded_noise(cfg, 1, amp, m_only=m_only), seeded_noise(cfg, 2, amp, m_only=m_only),
            seeded_noise(cfg, 3, amp, m_only=m_only)], [psi.ln, 0.0]


def energy_after(o, cfg, arrs, lns, nsteps):
    fs = [o.field(O.FFF, a, ln) for a, ln in zip(arrs, list(lns) + [0.0])]
    sv = O.OSolver(o)
    sv.initialize(*fs)
    for _ in range(nsteps - 1):
        rc = sv.step(*fs)
        if rc:
            raise RuntimeError(f"oracle step returned {rc}")
    ke, _, _ = o.energy(*fs)
    sv.close()
    return float(ke[1:, :].sum())


def cfg5_fd():
    cfg = CFG5
    o = O.OGrid(cfg)
    arrs, lns = initial(o, cfg, 1)
    eps = 1e-6
    J = energy_after(o, cfg, arrs, lns, cfg["nsteps"])
    out = []
    t0 = time.time()
    for i in range(16):
        d = fd_direction(cfg, i)
        jp = energy_after(o, cfg, [a + eps * x for a, x in zip(arrs, d)], lns, cfg["nsteps"])
        jm = energy_after(o, cfg, [a - eps * x for a, x in zip(arrs, d)], lns, cfg["nsteps"])
        out.append((jp - jm) / (2 * eps))
        print(f"cfg5 direction {i}: {out[-1]:.15e} ({time.time() - t0:.0f} s)", flush=True)
    o.close()
    return {"J": J, "eps": eps, "dJ": out}


def tiny_gradient():
    cfg = TINY
    o = O.OGrid(cfg)
    arrs, lns = initial(o, cfg, None)
    scale = max(float(np.max(np.abs(a))) for a in arrs)
    eps = 1e-7 * scale
    nst = cfg["nsteps"]
    J = energy_after(o, cfg, arrs, lns, nst)
    shape = arrs[0].shape
    grads = [np.zeros(shape, dtype=np.complex128) for _ in range(3)]
    nrchop, ntd = cfg["nrchop"], shape[1]
    for f in range(3):
        for mm in range(ntd):
            nn = max(min(nrchop, nrchop - mm), 0)
            for k in range(shape[2]):
                for n in range(nn):
                    for part in (1.0, 1j):
                        ap = [a.copy() for a in arrs]
                        am = [a.copy() for a in arrs]
                        ap[f][n, mm, k] += eps * part
                        am[f][n, mm, k] -= eps * part
                        g = (energy_after(o, cfg, ap, lns, nst) - energy_after(o, cfg, am, lns, nst)) / (2 * eps)
                        grads[f][n, mm, k] += g * part
        print(f"tiny field {f} done", flush=True)
    gln = []
    for f in range(2):
        lp, lm = list(lns), list(lns)
        lp[f] += eps
        lm[f] -= eps
        gln.append((energy_after(o, cfg, arrs, lp, nst) - energy_after(o, cfg, arrs, lm, nst)) / (2 * eps))
    o.close()
    return {"J": J, "eps": eps, "shape": list(shape),
            "grad_re": [g.real.reshape(-1, order="F").tolist() for g in grads],
            "grad_im": [g.imag.reshape(-1, order="F").tolist() for g in grads], "grad_ln": gln}


def main():
    O.lib().og_set_threads(os.cpu_count() or 1)
    out = {"source": "oracle/mlegs_oracle.c (FP64), tests/golden/make_adjoint_fd.py",
           "objective": "sum over m >= 1 and all k of KE(m, k) after INITIALIZE and nsteps-1 AB2-CN steps",
           "tiny_config": TINY, "tiny": tiny_gradient(), "cfg5": cfg5_fd()}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_fd.json"), "w") as fh:
        json.dump(out, fh)


if __name__ == "__main__":
    main()
