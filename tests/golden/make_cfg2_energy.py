"""Writes cfg2_energy.json: the oracle's kinetic energy and potential
(buoyancy-variance) energy totals of the 384x256x256 ETD2 workload after the
Richardson start and every 10 steps, with per-field norms, for up to 100
steps.  The workload does not last that long: the per-step DIAGNOST check
returns 12 (divergence) at step 68, and the series ends with that step and
code.  The same series is written a second time with the oracle's
Legendre sums in reversed order (its own round-off sensitivity): the
trajectory is unstable, so two FP64 summation orders drift apart, and that
drift is the scale a GPU run is held to.  A third series starts from the same
state with every initial coefficient's real and imaginary parts multiplied by
1 + 2^-50 u (u seeded uniform in [-1, 1]): a perturbation of every
coefficient at the round-off level, which unlike the reversed Legendre sums
reaches the trajectory's unstable modes from the first step, as a different
implementation's round-off does.

The oracle is the FP64 C restatement in oracle/ (test infrastructure); the
initial state is the workload's Lamb-Oseen vortex plus seeded noise 1e-3,
built exactly as tests/test_gpu_parity.py builds it.  Run from the repository
root: python tests/golden/make_cfg2_energy.py  (about 10-20 minutes per series
on 8 cores); `--perturbed-only` adds the third series to an existing file.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

import _oracle as O  # noqa: E402
from mlegs_amd.configs import CFG2, LAMB_OSEEN_Q, seeded_noise  # noqa: E402

NSTEPS, EVERY = 100, 10


PERTURB_SEED, PERTURB_REL = 7, 2.0 ** -50


def perturbed(a):
    rng = np.random.default_rng(PERTURB_SEED)
    u = rng.uniform(-1.0, 1.0, size=(2,) + a.shape)
    return np.asfortranarray(a.real * (1 + PERTURB_REL * u[0]) + 1j * (a.imag * (1 + PERTURB_REL * u[1])))


def series(reverse, perturb=False):
    O.lib().og_set_sum_order(1 if reverse else 0)
    cfg = CFG2
    o = O.OGrid(cfg)
    psi, chi = o.field(O.FFF), o.field(O.FFF)
    o.call("qvortex", LAMB_OSEEN_Q, psi, chi)
    amp = cfg["noise"]
    st = [(psi.arr() + seeded_noise(cfg, 1, amp), psi.ln), (seeded_noise(cfg, 2, amp), 0.0), (seeded_noise(cfg, 3, amp), 0.0)]
    if perturb:
        st = [(perturbed(a), ln) for a, ln in st]
    fs = [o.field(O.FFF, a, ln) for a, ln in st]
    sv = O.OSolver(o)
    sv.initialize(*fs)
    rows = []

    def record(step):
        ke, pe, _ = o.energy(*fs)
        arrs = [f.get() for f in fs]
        rows.append({"step": step, "ke": float(ke.sum()), "pe": float(pe.sum()),
                     "max_abs": [float(np.max(np.abs(a))) for a in arrs],
                     "l2": [float(np.sqrt(np.sum(np.abs(a) ** 2))) for a in arrs],
                     "ln": [float(f.ln) for f in fs]})
        print(f"step {step}: KE {rows[-1]['ke']:.15e} PE {rows[-1]['pe']:.15e} ({time.time() - t0:.0f} s)", flush=True)

    t0 = time.time()
    print("reversed summation order" if reverse else "perturbed start" if perturb else "oracle summation order",
          flush=True)
    record(0)
    for s in range(1, NSTEPS + 1):
        rc = sv.step(*fs)
        if rc:
            a = fs[0].get()
            print(f"step {s}: oracle returned {rc}; max |psi| n<5, 1<=m<6, k<5: "
                  f"{np.max(np.abs(a[:5, 1:6, :5])):.3e}", flush=True)
            rows.append({"step": s, "error_code": int(rc)})
            break
        if s % EVERY == 0:
            record(s)
    sv.close()
    o.close()
    O.lib().og_set_sum_order(0)
    return rows


def main():
    O.lib().og_set_threads(os.cpu_count() or 1)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg2_energy.json")
    if "--perturbed-only" in sys.argv:
        out = json.load(open(path))
        out["rows_perturbed_start"] = series(False, perturb=True)
        with open(path, "w") as fh:
            json.dump(out, fh, indent=1)
        return
    out = {"workload": "384x256x256 ETD2 (CFG2), Lamb-Oseen + seeded noise 1e-3, after the Richardson start",
           "source": "oracle/mlegs_oracle.c (FP64), tests/golden/make_cfg2_energy.py",
           "fields": ["psi", "chi", "b"], "rows": series(False), "rows_reversed_order": series(True),
           "rows_perturbed_start": series(False, perturb=True)}
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
