"""Optimal-perturbation gradients of the AB2-CN march through the library's
discrete adjoint (mlg_step_adjoint, mlg_start_adjoint, mlg_ke_gradient).

The objective is the perturbation kinetic energy after ``nsteps`` steps:

    J(psi0, chi0, b0) = sum over m >= 1 and all k of KE(m, k) of the final state,

KE as CALC_BOUSSI_ENERGY computes it (src/mod_boussinesq.f90:112-232); the
march is INITIALIZE (the Richardson start, src/mod_march.f90:535-606) then
``nsteps - 1`` STEP_BOUSSI_AB_CN (src/mod_march.f90:609-661), as
SOLVER_T%TIME_STEPPING runs it.  The reference has no adjoint.

State layout: one flat float64 device tensor per march state,
``[psi | chi | b | N_psi | N_chi | N_b | psi%LN | chi%LN | N_psi%LN | N_chi%LN]``,
each field its FFF coefficients as (re, im) pairs in the library's layout.

Two ways to the gradient, with the same numbers:

* ``gradient(x0)``: forward march keeping a state every ``every`` steps,
  then the reverse sweep, recomputing each segment's states from its
  checkpoint;
* ``torch.autograd``: ``StepFunction`` is one step (forward: the library's
  step; backward: its adjoint), ``EnergyFunction`` the objective, and
  ``trajectory_energy`` chains them with ``torch.utils.checkpoint`` over
  segments of ``every`` steps.

Every device call runs on PyTorch's current stream (the grid is moved onto
it), so library kernels and tensor copies are ordered.
"""
import numpy as np

from . import hip as H

FFF = H.FFF


def _view(ptr, n):
    """A float64 torch tensor over n doubles of device memory at ptr."""
    import torch

    class _Iface:
        __cuda_array_interface__ = {"shape": (n,), "typestr": "<f8", "data": (ptr, False), "version": 2, "strides": None}

    return torch.as_tensor(_Iface(), device="cuda")


class March:
    """Library fields and solver for one grid, with flat-tensor state I/O."""

    def __init__(self, grid):
        import torch

        self.grid = grid
        grid.set_stream(torch.cuda.current_stream().cuda_stream)
        self.solver = H.Solver(grid)
        self.u = [grid.scalar() for _ in range(3)]
        self.h = list(self.solver.history())
        self.ub = [grid.scalar() for _ in range(3)]
        self.hb = [grid.scalar() for _ in range(3)]
        self.nfff = grid.size(FFF)
        self.size = 12 * self.nfff + 4
        self._ev = {}

    def close(self):
        for f in self.u + self.ub + self.hb:
            f.close()
        self.solver.close()

    # ---- flat tensor <-> fields
    def _fields(self, adj):
        return (self.ub + self.hb) if adj else (self.u + self.h)

    def _views(self, adj):
        key = adj
        fs = self._fields(adj)
        ptrs = [f.device_ptrs() for f in fs]
        if self._ev.get(key, (None,))[0] != ptrs:
            es = [_view(e, 2 * self.nfff) for e, _ in ptrs]
            lns = [_view(ln, 1) for _, ln in ptrs]
            self._ev[key] = (ptrs, es, lns)
        return self._ev[key][1], self._ev[key][2]

    def load(self, x, adj=False, state_only=False):
        es, lns = self._views(adj)
        n2 = 2 * self.nfff
        for i in range(3 if state_only else 6):
            es[i].copy_(x[i * n2:(i + 1) * n2])
        o = 6 * n2
        lns[0].copy_(x[o:o + 1])
        lns[1].copy_(x[o + 1:o + 2])
        lns[2].zero_()
        if not state_only:
            lns[3].copy_(x[o + 2:o + 3])
            lns[4].copy_(x[o + 3:o + 4])
            lns[5].zero_()

    def dump(self, adj=False):
        import torch

        es, lns = self._views(adj)
        return torch.cat(es + [lns[0], lns[1], lns[3], lns[4]])

    # ---- one step forward and backward
    def step(self, x, start):
        """The march state after one step from x (start: the Richardson start)."""
        self.load(x, state_only=start)
        if start:
            self.solver.initialize(*self.u)
        else:
            self.solver.time_step(*self.u)
        return self.dump()

    def step_adjoint(self, x, ybar, start):
        """Adjoint of step(x, start) applied to ybar."""
        self.load(x, state_only=True)
        self.load(ybar, adj=True)
        if start:
            self.solver.start_adjoint(*self.u, self.ub, self.hb)
        else:
            self.solver.step_adjoint(*self.u, self.ub, self.hb)
        return self.dump(adj=True)

    def energy(self, x):
        """J: the sum over m >= 1 of KE(m, k) of the state in x."""
        self.load(x, state_only=True)
        ke, _, _ = H.calc_boussi_energy(*self.u)
        return float(ke[1:, :].sum())

    def energy_gradient(self, x):
        """dJ/dx (the state part; the history part is zero)."""
        import torch

        self.load(x, state_only=True)
        H.ke_gradient(self.u[0], self.u[1], self.ub[0], self.ub[1], 1)
        es, lns = self._views(True)
        g = torch.zeros(self.size, dtype=torch.float64, device="cuda")
        n2 = 2 * self.nfff
        g[0:n2] = es[0]
        g[n2:2 * n2] = es[1]
        return g

    # ---- state construction
    def state_from(self, psi, chi, b, psi_ln=0.0):
        """Flat initial state from host arrays (FFF, Fortran order); history zero."""
        import torch

        x = torch.zeros(self.size, dtype=torch.float64, device="cuda")
        n2 = 2 * self.nfff
        for i, a in enumerate((psi, chi, b)):
            a = np.asfortranarray(np.asarray(a, dtype=np.complex128))
            x[i * n2:(i + 1) * n2] = torch.from_numpy(a.reshape(-1, order="F").view(np.float64).copy()).cuda()
        x[6 * n2] = float(psi_ln)
        return x

    def fields_of(self, x):
        """Host arrays (psi, chi, b) and psi%LN, chi%LN of the state in x."""
        shape = self.grid.shape(FFF)
        h = x.detach().cpu().numpy()
        n2 = 2 * self.nfff
        out = [h[i * n2:(i + 1) * n2].view(np.complex128).reshape(shape, order="F") for i in range(3)]
        return out, (h[6 * n2], h[6 * n2 + 1])


def run(march, x0, nsteps):
    """The state after the march of nsteps steps (the first is the start)."""
    x = x0
    for s in range(nsteps):
        x = march.step(x, start=(s == 0))
    return x


def gradient(march, x0, nsteps, every=5):
    """(J, dJ/dx0): forward with a checkpoint every `every` steps, then the
    reverse sweep recomputing each segment from its checkpoint."""
    ckpt = {0: x0.clone()}
    x = x0
    for s in range(nsteps):
        x = march.step(x, start=(s == 0))
        if (s + 1) % every == 0 and s + 1 < nsteps:
            ckpt[s + 1] = x.clone()
    J = march.energy(x)
    xbar = march.energy_gradient(x)
    for seg0 in sorted(ckpt, reverse=True):
        seg1 = min(seg0 + every, nsteps)
        states = [ckpt[seg0]]
        for s in range(seg0, seg1 - 1):
            states.append(march.step(states[-1], start=(s == 0)))
        for s in range(seg1 - 1, seg0 - 1, -1):
            xbar = march.step_adjoint(states[s - seg0], xbar, start=(s == 0))
    return J, xbar


# ---------------------------------------------------------------- torch.autograd
def _functions():
    import torch

    class StepFunction(torch.autograd.Function):
        """One march step; backward is the library's adjoint of that step."""

        @staticmethod
        def forward(ctx, x, march, start):
            ctx.save_for_backward(x)
            ctx.march, ctx.start = march, start
            return march.step(x, start)

        @staticmethod
        def backward(ctx, ybar):
            (x,) = ctx.saved_tensors
            return ctx.march.step_adjoint(x, ybar.contiguous(), ctx.start), None, None

    class EnergyFunction(torch.autograd.Function):
        """J of a state (perturbation kinetic energy, m >= 1)."""

        @staticmethod
        def forward(ctx, x, march):
            ctx.save_for_backward(x)
            ctx.march = march
            return torch.tensor(march.energy(x), dtype=torch.float64, device=x.device)

        @staticmethod
        def backward(ctx, jbar):
            (x,) = ctx.saved_tensors
            return ctx.march.energy_gradient(x) * jbar, None

    return StepFunction, EnergyFunction


_FN = None


def functions():
    global _FN
    if _FN is None:
        _FN = _functions()
    return _FN


def trajectory_energy(march, x0, nsteps, every=5):
    """J after nsteps as a differentiable function of x0 (torch.autograd),
    the steps in checkpointed segments of `every` steps."""
    from torch.utils.checkpoint import checkpoint

    StepFunction, EnergyFunction = functions()

    def segment(x, s0, s1):
        for s in range(s0, s1):
            x = StepFunction.apply(x, march, s == 0)
        return x

    x = x0
    for s0 in range(0, nsteps, every):
        x = checkpoint(segment, x, s0, min(s0 + every, nsteps), use_reentrant=False)
    return EnergyFunction.apply(x, march)
