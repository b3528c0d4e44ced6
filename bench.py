#!/usr/bin/env python3
"""Steps/s of the FP64 Boussinesq ETD2 step at 384x256x256 on one MI355X.

Workload: NR=384, NTH=256, NX=256, NRCHOP=256, NTCHOP=NXCHOP=86, ELL=2,
ZLEN=4 pi, nu=kappa=1e-4, hyperviscosity filter P=4 (nu_p=kappa_p=1e-14)
beside CN viscosity, Omega=1, Nbar=2, ETD2 stepping, dt=5e-4.  Initial state:
Lamb-Oseen vortex (QVORTEX, Gamma=1) plus seeded complex Gaussian noise
1e-3 exp(-(n/16)^2-(m/4)^2-(k/4)^2) on psi, chi and b (synthetic data).

A step is one full ETD2-AB/CN step (mod_march.f90:811-901) including the
per-step DIAGNOST checks.  The Richardson start and W warm-up steps are
untimed; K steps are timed between device synchronisations, as replays of
the captured step graph (--eager: launched kernel by kernel).  The
per-kernel figures come from a second timed region of K eager steps, in
which the library records HIP events around every launch of its kernel
classes on its own stream; the class with the largest summed time is
reported in `roofline` with its algorithmic work per launch.  The CPU
baseline times the FP64 C restatement (OpenMP) on this host's cores over a
bounded sample of steps of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "steps/s at 384\u00d7256\u00d7256: split-prec vs FP64 GPU (\u22641e-10 err); adjoint/fwd cost"
PEAK_FP64 = 78.6e12      # MI355X FP64 matrix, FLOP/s
PEAK_HBM = 8.0e12        # MI355X HBM3E, B/s
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_kernels_cfg2_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps; with the warm-up at most 60 (the workload diverges at step 68)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps in the CPU baseline sample (0: skip)")
    ap.add_argument("--no-timer", action="store_true", help="time without per-kernel events")
    ap.add_argument("--eager", action="store_true",
                    help="time eager launches instead of replays of the captured step graph")
    ap.add_argument("--no-split", action="store_true",
                    help="skip the split hi/lo FP32 Legendre measurement (timing and error against FP64)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint measurement (128x64x64 AB2-CN, 40-step gradient)")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the batched uneven sweep measurement (48 AB2-CN runs)")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only run the sweep measurement (the modes of --sweep-modes) and print its JSON")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the adjoint measurement (128x64x64 AB2-CN step and its transpose) and print its JSON")
    ap.add_argument("--sweep-steps", type=int, default=50, help="steps per run of the sweep (2 untimed)")
    ap.add_argument("--sweep-modes", default=",".join(SWEEP_MODES),
                    help="comma-separated sweep modes to time (--sweep-only)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 20 copies of a 384x257x256 complex field (407 MB each way), "
                         "to scale HBM byte counters")
    return ap.parse_args()


def calibrate_copy(H, cfg):
    """20 launches of the library's copy kernel on a full physical-space field:
    each reads and writes NDIMR*NDIMTH*NDIMX*16 bytes once."""
    g = H.Grid(cfg)
    a, b = g.scalar(H.PPP), g.scalar(H.PPP)
    a.zero(H.PPP)
    for _ in range(20):
        H.copy(b, a)
    g.synchronize()
    n = g.size(H.PPP) * 16
    print(json.dumps({"copy_bytes_each_way": n, "launches": 20}), flush=True)
    a.close()
    b.close()
    g.close()


def initial_state(H, g, cfg):
    from mlegs_amd.configs import LAMB_OSEEN_Q, seeded_noise

    psi, chi, b = g.scalar(), g.scalar(), g.scalar()
    H.qvortex(g, LAMB_OSEEN_Q, psi, chi)
    p, _, ln = psi.download()
    psi.upload(p + seeded_noise(cfg, 1), H.FFF, ln)
    chi.upload(seeded_noise(cfg, 2), H.FFF, 0.0)
    b.upload(seeded_noise(cfg, 3), H.FFF, 0.0)
    return psi, chi, b


def field_errors(H, xs, ys):
    """max|x - y| / max|y| per field (psi, chi, b) and |LN_x - LN_y| / max(1, |LN_y|)."""
    import numpy as np

    out = {}
    for nm, x, y in zip(("psi", "chi", "b"), xs, ys):
        xa, _, xl = x.download()
        ya, _, yl = y.download()
        out[nm] = float(np.max(np.abs(xa - ya)) / np.max(np.abs(ya)))
        out[nm + "_ln"] = abs(xl - yl) / max(1.0, abs(yl))
    return {k: float(f"{v:.3e}") for k, v in out.items()}


def split_measurement(H, cfg, steps, warmup, ms_fp64):
    """The same workload with the Legendre products in split hi/lo FP32: its
    steps/s (graph replay, K steps after W warm-up steps), its per-kernel
    Legendre times, and its difference from the FP64 GPU step after the
    start + 1 and + 60 steps (per field, and in total kinetic and potential
    energy).  60, not 100: the workload's per-step DIAGNOST check stops it
    with code 12 (divergence) at step 68, in the FP64 oracle as here.  Two
    contexts, one per precision, march from the same state."""
    import torch

    gd, gs = H.Grid(cfg), H.Grid(cfg)
    gs.legendre_precision = "split32"
    fd, fs = initial_state(H, gd, cfg), initial_state(H, gs, cfg)
    sd, ss = H.Solver(gd), H.Solver(gs)
    sd.initialize(*fd)
    ss.initialize(*fs)
    sd.time_stepping(*fd, 1, graph=False)
    ss.time_stepping(*fs, 1, graph=False)
    err1 = field_errors(H, fs, fd)
    sd.time_stepping(*fd, 59, graph=True)
    ss.time_stepping(*fs, 59, graph=True)
    err100 = field_errors(H, fs, fd)
    ked, ped, _ = H.calc_boussi_energy(*fd)
    kes, pes, _ = H.calc_boussi_energy(*fs)
    err100["ke_total"] = float(f"{abs(kes.sum() - ked.sum()) / abs(ked.sum()):.3e}")
    err100["pe_total"] = float(f"{abs(pes.sum() - ped.sum()) / abs(ped.sum()):.3e}")
    sd.close()
    gd.close()
    # timing: the split context only, from a fresh start (the timed steps
    # must end before the workload's divergence at step 68)
    ss.close()
    fs = initial_state(H, gs, cfg)
    ss = H.Solver(gs)
    ss.initialize(*fs)
    ss.time_stepping(*fs, max(warmup, 2), graph=True)
    torch.cuda.synchronize()
    gs.synchronize()
    t0 = time.perf_counter()
    ss.time_stepping(*fs, steps, graph=True)
    torch.cuda.synchronize()
    gs.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    gs.timer(True)
    ss.time_stepping(*fs, steps)
    gs.synchronize()
    gs.timer(False)
    stages = {}
    for k in ("rtran_bwd", "rtran_fwd", "project"):
        t, n, work, unit = gs.timer_read(k)
        if n:
            stages[k] = dict(ms_per_step=round(t / steps, 4), avg_us=round(t / n * 1e3, 2),
                             fp64_equiv_tflops=round(work / n / (t / n * 1e-3) / 1e12, 2))
    ss.close()
    gs.close()
    return dict(ms_per_step=round(ms, 3), steps_per_s=round(1e3 / ms, 4), vs_fp64=round(ms_fp64 / ms, 4),
                err_vs_fp64_gpu_after_start_plus_1=err1, err_vs_fp64_gpu_after_start_plus_60=err100,
                legendre_stages=stages,
                note="split/FP64 steps/s ratio of the same graph-replayed step; errors are max|split-fp64|/max|fp64| "
                     "per field (the FP64 path itself is held to the oracle by the parity tests)")


def adjoint_measurement(H):
    """Adjoint/forward cost on the 128x64x64 AB2-CN workload (CFG5): the
    gradient of the perturbation kinetic energy after 40 steps (mlegs_amd.
    adjoint).  Per step: the library's forward step (mlg_time_step) and its
    adjoint (mlg_step_adjoint, which recomputes the state's products), each
    timed over 40 calls on the same state; per trajectory: the checkpointed
    gradient (forward with a state kept every 5 steps, reverse sweep with
    each segment recomputed) against the plain 40-step forward march."""
    import torch
    from mlegs_amd import adjoint as A
    from mlegs_amd.configs import CFG5, LAMB_OSEEN_Q, seeded_noise

    cfg = CFG5
    g = H.Grid(cfg)
    m = A.March(g)
    try:
        psi, chi = g.scalar(), g.scalar()
        H.qvortex(g, LAMB_OSEEN_Q, psi, chi)
        p0, _, ln = psi.download()
        amp = cfg["noise"]
        x0 = m.state_from(p0 + seeded_noise(cfg, 1, amp, m_only=1), seeded_noise(cfg, 2, amp, m_only=1),
                          seeded_noise(cfg, 3, amp, m_only=1), psi_ln=ln)
        n = cfg["nsteps"]
        A.gradient(m, x0, n, every=5)  # warm-up: pools, factor sets
        x1 = m.step(x0, True)
        # per step, on the library's fields directly
        m.load(x1)
        ybar = torch.zeros_like(x1)
        ybar[: 2 * m.nfff] = 1.0
        m.load(ybar, adj=True)

        def timed(fn, reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / reps * 1e3

        m.load(x1)
        fwd = timed(lambda: m.solver.time_step(*m.u), n)
        m.load(x1)
        adj = timed(lambda: m.solver.step_adjoint(*m.u, m.ub, m.hb), n)
        run_ms = timed(lambda: A.run(m, x0, n), 3)
        grad_ms = timed(lambda: A.gradient(m, x0, n, every=5), 3)
        J, _ = A.gradient(m, x0, n, every=5)
        return dict(workload="128x64x64 AB2-CN (NRCHOP=96, NTCHOP=NXCHOP=22, dt=1e-3), 40 steps, "
                             "Lamb-Oseen + m=1 noise 1e-5; J = perturbation KE (m >= 1)",
                    fwd_ms_per_step=round(fwd, 4), adjoint_ms_per_step=round(adj, 4),
                    adjoint_over_forward_step=round(adj / fwd, 3),
                    forward_40_steps_ms=round(run_ms, 3), gradient_40_steps_ms=round(grad_ms, 3),
                    gradient_over_forward_trajectory=round(grad_ms / run_ms, 3), J=J,
                    note="adjoint step includes recomputing the state's velocity, vorticity and buoyancy "
                         "gradient; the trajectory gradient adds the checkpointed forward and the "
                         "segment recomputation (every 5 steps); both trajectories include the flat-tensor "
                         "state copies of mlegs_amd.adjoint")
    finally:
        m.close()
        g.close()


SWEEP_MODES = ("alone", "ungrouped", "grouped", "ungrouped_streams", "grouped_streams", "grouped_by_grid_streams")


def sweep_measurement(H, nsteps, warm=2, modes=SWEEP_MODES):
    """The batched uneven sweep (48 AB2-CN runs of mlegs_amd.configs.
    cfg4_runs, 50 steps each): aggregate run-steps/s of the last
    nsteps - warm steps of every run, (a) each run alone, replaying its own
    captured step graph, runs one after another; (b) one sweep graph per step
    for all runs, the runs' segments independent; (c) the same with the RTRAN
    backward products of all runs grouped into one persistent launch per place
    in the step, tile shapes from the tuner.  The first `warm` steps (eager
    step, recording, tuning) are untimed.  Every mode starts from the same
    initial states, and its states after the last step are compared with
    mode (a)'s (bitwise equal)."""
    import numpy as np
    import torch
    from mlegs_amd.configs import LAMB_OSEEN_Q, cfg4_runs, seeded_noise

    cfgs = cfg4_runs()
    grids, fields, init = [], [], []
    for cfg in cfgs:
        g = H.Grid(cfg)
        f = [g.scalar() for _ in range(3)]
        H.qvortex(g, LAMB_OSEEN_Q, f[0], f[1])
        p, _, ln = f[0].download()
        init.append([(p + seeded_noise(cfg, 1), ln), (seeded_noise(cfg, 2), 0.0), (seeded_noise(cfg, 3), 0.0)])
        grids.append(g)
        fields.append(f)

    def start():
        solvers = []
        for g, f, st in zip(grids, fields, init):
            for x, (a, ln) in zip(f, st):
                x.upload(a, H.FFF, ln)
            s = H.Solver(g)
            s.initialize(*f)
            solvers.append(s)
        return solvers

    def alone(s, f, n, codes, i):
        # a run whose DIAGNOST check fails has still taken its n steps
        try:
            s.time_stepping(*f, n, graph=True)
        except H.MlegsError as e:
            if e.code >= 100:
                raise
            codes[i] = codes[i] or e.code

    res, final, info, codes = {}, {}, None, {}
    timed = nsteps - warm
    for mode in modes:
        solvers = start()
        sw = None
        codes[mode] = [0] * len(cfgs)
        if mode == "alone":
            for i, (s, f) in enumerate(zip(solvers, fields)):
                alone(s, f, warm, codes[mode], i)
        else:
            sw = H.Sweep([(s, *f) for s, f in zip(solvers, fields)], grouped=mode.startswith("grouped"), tune=True,
                         streams=mode.endswith("streams"), by_grid="by_grid" in mode)
            sw.steps(warm, check=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if sw is None:
            for i, (s, f) in enumerate(zip(solvers, fields)):
                alone(s, f, timed, codes[mode], i)
        else:
            codes[mode] = sw.steps(timed, check=False)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[mode] = len(cfgs) * timed / el
        final[mode] = [[np.ascontiguousarray(x.download()[0]).view(np.int64) for x in f] for f in fields]
        if sw is not None:
            if mode.startswith("grouped") and info is None:
                info = dict(sw.info(), mode=mode)
            sw.close()
        for s in solvers:
            s.close()
        print(f"bench: sweep {mode}: {res[mode]:.1f} run-steps/s", file=sys.stderr, flush=True)
    same = {m: all(np.array_equal(a, b) for fa, fb in zip(final[m], final["alone"]) for a, b in zip(fa, fb))
            for m in final if m != "alone"} if "alone" in final else {}
    for g, f in zip(grids, fields):
        for x in f:
            x.close()
        g.close()
    out = dict(workload="48 AB2-CN runs, NR in {64,96,128,192} x NTH=NX in {32,64,128}, NRCHOP=2NR/3, "
                        "NTCHOP=NXCHOP=NTH/3+1, seeded ELL/OMEGA/BV0, nu=kappa=5e-4, dt=1e-3; "
                        f"{nsteps} steps per run, the last {timed} timed",
               unit="run-steps/s")
    swept = {m: v for m, v in res.items() if m != "alone"}
    if swept:
        best = max(swept, key=swept.get)
        out["run_steps_per_s"] = round(swept[best], 1)
        out["best_mode"] = best
    for m in res:
        out[f"{m}_run_steps_per_s"] = round(res[m], 1)
    if "alone" in res and swept:
        out["best_over_alone"] = round(max(swept.values()) / res["alone"], 3)
    if "alone" in res and "grouped" in res:
        out["grouped_over_alone"] = round(res["grouped"] / res["alone"], 3)
    out["bitwise_equal_to_alone"] = same
    out["grouped_graph"] = info
    last = codes[modes[-1]]
    out["runs_failing_diagnost"] = [i for i, c in enumerate(last) if c]
    out["modes"] = ("alone: each run's captured step graph replayed, runs one after another; ungrouped/grouped: "
                    "one graph launch per step for all runs (grouped: the RTRAN backward products of all runs "
                    "in one persistent launch per place in the step, tuned tiles); *_streams: each run's "
                    "segments launched on its own stream (4 hardware queues), grouped launches joined by "
                    "events; by_grid: one grouped launch per class of runs of one grid")
    out["diagnost_codes_same_in_every_mode"] = all(c == last for c in codes.values())
    out["note"] = ("runs with small ELL or large NTH fail the per-step DIAGNOST divergence check (code 12) "
                   "within 50 steps at dt=1e-3 -- the FP64 oracle fails them at the same step -- and keep "
                   "stepping with the others; the timing is of all 48 runs")
    return out


def cpu_baseline(cfg, nsteps):
    """The oracle (FP64 C, OpenMP) on this host: steps/s over `nsteps` steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import _oracle as O
    from mlegs_amd.configs import LAMB_OSEEN_Q, seeded_noise

    # the host's CPUs as the scheduler reports them; one GPU's share of the
    # host is 16 cores, so the oracle runs on at most that many threads
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = min(16, nproc)
    o = O.OGrid(cfg, threads=cores)
    try:
        psi, chi = o.field(O.FFF), o.field(O.FFF)
        o.call("qvortex", LAMB_OSEEN_Q, psi, chi)
        psi.set(psi.arr() + seeded_noise(cfg, 1), O.FFF, psi.ln)
        chi.set(seeded_noise(cfg, 2), O.FFF)
        b = o.field(O.FFF, np.asfortranarray(seeded_noise(cfg, 3)))
        s = O.OSolver(o)
        s.prepare()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            rc = s.step(psi, chi, b)
            if rc:
                raise RuntimeError(f"oracle step returned {rc}")
        dt = time.perf_counter() - t0
        s.close()
    finally:
        o.close()
    return dict(value=nsteps / dt, unit="steps/s", cores=cores, host_nproc=nproc, kind="port",
                sample=f"{nsteps} ETD2 steps of the 384x256x256 workload from the seeded state "
                       f"(ETD operators set, no Richardson start), FP64 C restatement with OpenMP, "
                       f"{cores} threads (host nproc {nproc}, one GPU's share 16 cores), {dt:.1f} s")


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 "
              "(the multi-GPU decomposition is a design note, not built)", file=sys.stderr)
        sys.exit(2)
    import torch

    sys.path.insert(0, ROOT)
    from mlegs_amd import hip as H
    from mlegs_amd.configs import CFG2

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    cfg = dict(CFG2)
    if a.warmup + a.steps > 60:
        print("bench.py: --warmup + --steps must be at most 60: the workload's per-step DIAGNOST check stops it "
              "with code 12 (divergence) at step 68, in the FP64 oracle too", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy(H, cfg)
        return
    if a.sweep_only:
        print(json.dumps(sweep_measurement(H, a.sweep_steps, modes=tuple(a.sweep_modes.split(",")))), flush=True)
        return
    if a.adjoint_only:
        print(json.dumps(adjoint_measurement(H)), flush=True)
        return
    g = H.Grid(cfg)
    psi, chi, b = initial_state(H, g, cfg)
    s = H.Solver(g)
    s.initialize(psi, chi, b)
    use_graph = not a.eager
    if a.warmup:
        s.time_stepping(psi, chi, b, max(a.warmup, 2 if use_graph else 1), graph=use_graph)
    torch.cuda.synchronize()
    g.synchronize()
    print(f"bench: initialised, warm-up done; timing {a.steps} steps ({'graph' if use_graph else 'eager'})",
          file=sys.stderr, flush=True)
    if not a.no_timer and not use_graph:
        g.timer(True)
    t0 = time.perf_counter()
    s.time_stepping(psi, chi, b, a.steps, graph=use_graph)
    torch.cuda.synchronize()
    g.synchronize()
    el = time.perf_counter() - t0
    ms_step = el / a.steps * 1e3
    if not a.no_timer and use_graph:
        # per-kernel HIP events need eager launches: a second timed region of
        # the same K steps, launched one kernel at a time
        print("bench: eager timed region for the per-kernel events", file=sys.stderr, flush=True)
        # from a fresh start, so that no trajectory reaches the divergence at step 68
        s.close()
        psi, chi, b = initial_state(H, g, cfg)
        s = H.Solver(g)
        s.initialize(psi, chi, b)
        s.time_stepping(psi, chi, b, a.warmup)
        g.synchronize()
        g.timer(True)
        t1 = time.perf_counter()
        s.time_stepping(psi, chi, b, a.steps)
        torch.cuda.synchronize()
        g.synchronize()
        eager_ms = (time.perf_counter() - t1) / a.steps * 1e3
    else:
        eager_ms = ms_step
    g.timer(False)

    stages, dom = {}, None
    if not a.no_timer:
        for k in H.TIMED_KERNELS:
            ms, n, work, unit = g.timer_read(k)
            if n == 0:
                continue
            stages[k] = dict(ms_per_step=round(ms / a.steps, 4), launches_per_step=n / a.steps,
                             avg_us=round(ms / n * 1e3, 2), work_per_launch=work / n, unit=unit)
            if dom is None or ms > stages[dom]["ms_per_step"] * a.steps:
                dom = k
    roof = None
    if dom:
        st = stages[dom]
        rate = st["work_per_launch"] / (st["avg_us"] * 1e-6)
        flop = st["unit"] == "flop"
        peak = PEAK_FP64 if flop else PEAK_HBM
        traffic = None
        if os.path.exists(TRAFFIC_FILE):
            traffic = json.load(open(TRAFFIC_FILE)).get(dom)
        roof = dict(bound="mfma" if flop else "hbm", kernel=dom,
                    achieved=round(rate / (1e12 if flop else 1e9), 3), peak=(78.6 if flop else 8000.0),
                    unit="TFLOP/s" if flop else "GB/s", frac=round(rate / peak, 4), traffic=traffic,
                    work_per_launch=st["work_per_launch"], avg_launch_us=st["avg_us"])
    s.close()
    g.close()
    split = None
    if not a.no_split:
        print("bench: split hi/lo FP32 Legendre products", file=sys.stderr, flush=True)
        split = split_measurement(H, cfg, a.steps, a.warmup, ms_step)
    adjoint = None
    if not a.no_adjoint:
        print("bench: adjoint/forward cost (128x64x64 AB2-CN)", file=sys.stderr, flush=True)
        adjoint = adjoint_measurement(H)
    sweep = None
    if not a.no_sweep:
        print("bench: batched uneven sweep (48 runs)", file=sys.stderr, flush=True)
        sweep = sweep_measurement(H, a.sweep_steps)
    cpu = None
    if a.cpu_steps > 0:
        print("bench: timing the CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(cfg, a.cpu_steps)
    out = {
        "metric": METRIC,
        "value": round(a.steps / el, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Lamb-Oseen vortex + seeded noise 1e-3)",
        "config": {"workload": "384x256x256 ETD2 FP64 step (NRCHOP=256, NTCHOP=NXCHOP=86, dt=5e-4)",
                   "nr": cfg["nr"], "nth": cfg["nth"], "nx": cfg["nx"], "nrchop": cfg["nrchop"],
                   "ntchop": cfg["ntchop"], "nxchop": cfg["nxchop"], "scheme": "ETD2-AB/CN + P=4 filter",
                   "precision": "fp64"},
        "roofline": roof,
        "launch": "hip graph replay of the captured step" if use_graph else "eager",
        "eager_ms_per_step": round(eager_ms, 3),
        "cpu_baseline": cpu,
        "stages": stages,
        "split_precision": split,
        "adjoint": adjoint,
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
