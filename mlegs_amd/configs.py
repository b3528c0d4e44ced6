"""Workload configurations and seeded initial conditions.

The five workloads of the Boussinesq benchmark suite, written as plain
dictionaries whose keys match the fields of ``mlg_params``:

* ``CFG1`` -- 64x32x32 AB2-CN run (the reference's own CPU-sized case).
* ``CFG2`` -- 384x256x256 ETD2 production case (the headline workload).
* ``CFG3`` -- CFG2 with split hi/lo FP32 Legendre products (``split=1``:
  ``mlegs_amd.hip.Grid`` selects ``legendre_precision = "split32"``).
* ``cfg4_runs()`` -- the 48 runs of uneven resolution of the batched sweep.
* ``CFG5`` -- 128x64x64 AB2-CN case of the adjoint (gradient of the
  perturbation kinetic energy after 40 steps; mlegs_amd.adjoint).

Mapping of the production case's viscosity: the reference offers CN
viscosity (``VISC%SW=1``, mod_march.f90:1233-1265) or the exponential
hyperviscosity filter (``VISC%SW=3``, mod_march.f90:3521-3556), not both.
CFG2 uses SW=1 with ``hyper_cn=1``, which also applies the SW=3 filter
(built from nu_p=kappa_p=1e-14, P=4) at the ETD filter call sites
(mod_march.f90:846,867).
"""
import math

import numpy as np

CFG1 = dict(nr=64, nth=32, nx=32, nrchop=48, ntchop=11, nxchop=11, minc=1, mklink=0,
            ell=1.0, zlen=2 * math.pi, nu=1e-3, kappa=1e-3, nup=0.0, kappap=0.0,
            visc_sw=1, visc_p=0, omega=0.5, bv0=1.0, adams=1, hyper_cn=0, dt=1e-3,
            nsteps=10, noise=1e-4)

CFG2 = dict(nr=384, nth=256, nx=256, nrchop=256, ntchop=86, nxchop=86, minc=1, mklink=0,
            ell=2.0, zlen=4 * math.pi, nu=1e-4, kappa=1e-4, nup=1e-14, kappap=1e-14,
            visc_sw=1, visc_p=4, omega=1.0, bv0=2.0, adams=0, hyper_cn=1, dt=5e-4,
            nsteps=100, noise=1e-3)
# CFG2's state does not last its 100 steps: the per-step DIAGNOST check
# stops it with code 12 (divergence) at step 68, in the FP64 oracle as on
# the GPU (tests/golden/cfg2_energy.json); measurements stop at step 60.

CFG3 = dict(CFG2, split=1)

CFG5 = dict(nr=128, nth=64, nx=64, nrchop=96, ntchop=22, nxchop=22, minc=1, mklink=0,
            ell=1.0, zlen=2 * math.pi, nu=1e-3, kappa=1e-3, nup=0.0, kappap=0.0,
            visc_sw=1, visc_p=0, omega=0.5, bv0=1.0, adams=1, hyper_cn=0, dt=1e-3,
            nsteps=40, noise=1e-5)

CONFIGS = {"cfg1": CFG1, "cfg2": CFG2, "cfg3": CFG3, "cfg5": CFG5}

# CFG4: the batched uneven sweep -- 48 AB2-CN runs of 50 steps, stepped
# together (mlegs_amd.hip.Sweep).  NR cycles through CFG4_NR and NTH = NX
# through CFG4_NTH (each of the 12 grids 4 times), NRCHOP = floor(2 NR / 3),
# NTCHOP = NXCHOP = floor(NTH / 3) + 1; ELL, OMEGA and the buoyancy frequency
# BV0 are seeded uniform draws on [0.5, 3], [0, 2], [0, 4].
CFG4_NR = (64, 96, 128, 192)
CFG4_NTH = (32, 64, 128)
CFG4_SEED = 20261204
CFG4_BASE = dict(CFG1, nu=5e-4, kappa=5e-4, adams=1, dt=1e-3, nsteps=50, noise=1e-4)


def cfg4_runs(nruns=48, seed=CFG4_SEED):
    """The sweep's run configurations (list of dicts)."""
    rng = np.random.default_rng(seed)
    runs = []
    for i in range(nruns):
        nr, nth = CFG4_NR[i % len(CFG4_NR)], CFG4_NTH[(i // len(CFG4_NR)) % len(CFG4_NTH)]
        ell, omega, bv0 = rng.uniform(0.5, 3.0), rng.uniform(0.0, 2.0), rng.uniform(0.0, 4.0)
        runs.append(dict(CFG4_BASE, nr=nr, nth=nth, nx=nth, nrchop=(2 * nr) // 3, ntchop=nth // 3 + 1,
                         nxchop=nth // 3 + 1, ell=float(ell), omega=float(omega), bv0=float(bv0)))
    return runs

LAMB_OSEEN_Q = 1.0 / (2.0 * math.pi)   # Gamma = 1, reference Q = Gamma/(2 pi)
NOISE_SEED = 20261202


def dims(cfg):
    """Array dimensions (reference LEGINIT, mod_fft.f90:222-235)."""
    nr, nth, nx = cfg["nr"], cfg["nth"], cfg["nx"]
    ndimth = nth + 1 if nth > 1 else 1
    return dict(ndimr=nr + 3, ndimth=ndimth, ndimx=nx,
                nrchopdim=min(cfg["nrchop"] + 6, nr), ntchopdim=min(cfg["ntchop"], nth),
                nxchopdim=2 * cfg["nxchop"] - 1)


def _mix64(x):
    """splitmix64 finaliser on uint64 arrays (counter-based generator)."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return x ^ (x >> np.uint64(31))


def seeded_noise(cfg, field_id, amplitude=None, seed=NOISE_SEED, m_only=None, nrchop=None):
    """Complex Gaussian noise on every kept (n,m,k) of an FFF field.

    Value at (n, m, k) depends only on (seed, field_id, n, m, k): a
    counter-based draw, independent of array layout.  n is the 0-based
    index within azimuthal mode m (degree m+n).  Amplitude envelope
    A*exp(-(n/16)^2-(m/4)^2-(k/4)^2).  m=0 is made Hermitian in k and
    c(n,0,0) real; entries beyond the triangular truncation are zero.
    ``nrchop`` overrides the truncation (not the array shape), e.g. to fill
    the NRCHOP+3 rows used inside the nonlinear term.
    """
    d = dims(cfg)
    amp = cfg.get("noise", 1e-4) if amplitude is None else amplitude
    nrd, ntd, nxd, nxc = d["nrchopdim"], d["ntchopdim"], d["nxchopdim"], cfg["nxchop"]
    n = np.arange(nrd, dtype=np.uint64)[:, None, None]
    m = np.arange(ntd, dtype=np.uint64)[None, :, None]
    kk = np.arange(nxd)
    kval = np.where(kk < nxc, kk, kk - nxd)
    k = (kval + 4096).astype(np.uint64)[None, None, :]
    key = (np.uint64(seed) * np.uint64(1000003) + np.uint64(field_id)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    base = _mix64(key ^ (n << np.uint64(40)) ^ (m << np.uint64(24)) ^ k)
    u1 = (_mix64(base) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u2 = (_mix64(base ^ np.uint64(0x5851F42D4C957F2D)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u1 = np.maximum(u1, 2.0 ** -60)
    rad = np.sqrt(-2.0 * np.log(u1))
    z = rad * np.cos(2 * np.pi * u2) + 1j * rad * np.sin(2 * np.pi * u2)
    nf = np.arange(nrd)[:, None, None].astype(np.float64)
    mf = np.arange(ntd)[None, :, None].astype(np.float64) * cfg.get("minc", 1)
    kf = np.abs(kval)[None, None, :].astype(np.float64)
    env = amp * np.exp(-(nf / 16.0) ** 2 - (mf / 4.0) ** 2 - (kf / 4.0) ** 2)
    a = z * env
    # triangular truncation NRCHOPS(m) = max(min(NRCHOP, NRCHOP - m), 0)
    nrchop = cfg["nrchop"] if nrchop is None else nrchop
    for mm in range(ntd):
        nn = max(min(nrchop, nrchop - mm * cfg.get("minc", 1)), 0)
        a[nn:, mm, :] = 0
    if m_only is not None:
        keep = np.zeros(ntd, dtype=bool)
        keep[m_only] = True
        a[:, ~keep, :] = 0
    # m = 0 Hermitian in k, k = 0 real
    a[:, 0, 0] = a[:, 0, 0].real
    for kpos in range(1, nxc):
        a[:, 0, nxd - kpos] = np.conj(a[:, 0, kpos])
    return np.asfortranarray(a)


FD_SEED = 20261203


def fd_direction(cfg, i):
    """The i-th seeded direction (psi, chi, b) of the adjoint's finite-difference
    checks: seeded noise over every kept coefficient (all m), with the
    amplitude, envelope, truncation and m = 0 symmetry of the workload's
    initial noise, so that a step e along it is a relative change e of the
    perturbation."""
    return tuple(seeded_noise(cfg, 3 * i + f + 1, seed=FD_SEED) for f in range(3))
