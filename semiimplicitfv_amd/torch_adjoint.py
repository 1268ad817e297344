"""torch.autograd.Function over the GPU discrete adjoint.

    U_N = ExplicitRun.apply(U0, gamma, spec)

U0: float64 device tensor (4, totalCells) of conservative fields (rho, rhoU,
rhoV, rhoE; padded layout, ghosts ignored); gamma: 0-d float64 tensor; spec:
RunSpec (mesh, config with explicitParams.constDt > 0, steps, checkpoint
interval).  Forward runs the fixed-dt WENO5 + Rusanov + SSP-RK3 steps on the
GPU; backward applies the hand-written discrete adjoint
(sifv_adjoint_backward) and returns d/dU0 and d/dgamma.  The extension
module (_sifv_torch, built by semiimplicitfv_amd.build) provides the
zero-copy state views; there is no CPU path.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from .solvers import ExplicitSolver
from .state import IdealGasEOS, RusanovSolver


def _ext():
    try:
        from . import _sifv_torch
    except ImportError as e:  # pragma: no cover - build problem
        raise RuntimeError("torch extension _sifv_torch is not built: run semiimplicitfv_amd.build") from e
    return _sifv_torch


@dataclass
class RunSpec:
    """One fixed-dt run shape.  The solver handle it owns records the
    checkpoints of one forward run at a time; `_recorded` names that run
    (a counter over forwards), so a backward whose run has since been
    replaced (chained or repeated applies) records its own run again from the
    U0 and gamma its forward saved."""
    mesh: object
    config: object
    nsteps: int
    checkpoint_every: int = 20
    _solver: object = None
    _gamma: float = None
    _runs: int = 0
    _recorded: int = -1

    def solver(self, gamma: float):
        if self._solver is None or self._gamma != gamma:
            if self._solver is not None:
                self._solver.close()
            eos = IdealGasEOS(gamma, 1.0, self.config)
            self._solver = ExplicitSolver(self.mesh, RusanovSolver(eos), eos, None, self.config)
            self._gamma = gamma
            self._recorded = -1
        return self._solver

    def record(self, U0, gamma: float, run_id: int):
        """Runs the forward steps from U0 and keeps their checkpoints."""
        ext = _ext()
        s = self.solver(gamma)
        ext.load_state(s._h.value, U0.detach().contiguous())
        s.adjoint_forward(self.config, self.nsteps, self.checkpoint_every)
        self._recorded = run_id
        return s


class ExplicitRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, gamma, spec: RunSpec):
        ext = _ext()
        spec._runs += 1
        ctx.run_id = spec._runs
        ctx.gamma = float(gamma)
        s = spec.record(U0, ctx.gamma, ctx.run_id)
        UN = ext.resident_state(s._h.value, U0).clone()
        ctx.save_for_backward(U0.detach())
        ctx.spec = spec
        ctx.gamma_shape = gamma.shape
        return UN

    @staticmethod
    def backward(ctx, gUN):
        ext = _ext()
        spec = ctx.spec
        (U0,) = ctx.saved_tensors
        if spec._recorded != ctx.run_id or spec._gamma != ctx.gamma:
            # another forward replaced this run's checkpoints: record it again
            s = spec.record(U0, ctx.gamma, ctx.run_id)
        else:
            s = spec.solver(ctx.gamma)
        grad, dg = ext.adjoint_backward(s._h.value, gUN.contiguous())
        return grad, dg.reshape(ctx.gamma_shape).to(gUN.device), None
