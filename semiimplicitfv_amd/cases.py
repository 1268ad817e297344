"""Initial conditions of the measured workloads (the role of the reference's
examples/ drivers).  Each builder returns (config, mesh, state, eos, solver
object) for one workload and fills interior primitive and conservative
fields the way the reference drivers do (setPrimitiveState +
setConservativeState(eos.toConservative(W))).
"""
from __future__ import annotations

import math

import numpy as np

from .config import BoundaryCondition, ReconstructionOrder, SimulationConfig
from .mesh import RectilinearMesh, uniform_mesh
from .state import HLLCSolver, IdealGasEOS, LFSolver, RusanovSolver, SolutionState


def _fill(mesh: RectilinearMesh, state: SolutionState, eos, rho, u, p):
    """Write interior primitive + conservative fields from (nz,ny,nx) arrays."""
    r, mom, E = eos.toConservative(rho, u, p)
    mesh.interior(state.rho)[...] = r
    mesh.interior(state.pres)[...] = p
    vel = [state.velU, state.velV, state.velW]
    mo = [state.rhoU, state.rhoV, state.rhoW]
    for d in range(mesh.dim()):
        mesh.interior(vel[d])[...] = u[d]
        mesh.interior(mo[d])[...] = mom[d]
    mesh.interior(state.rhoE)[...] = E


def riemann_by_name(name: str, eos):
    return {"hllc": HLLCSolver, "rusanov": RusanovSolver, "lf": LFSolver}[name.lower()](eos)


def sod_1d(n: int = 1024, riemann: str = "hllc", recon=ReconstructionOrder.WENO5, cfl: float = 0.6):
    """Sod tube: [0,1], n cells, (1,0,1) | (0.125,0,0.1) at x=0.5, outflow,
    ideal gas 1.4, SSP-RK3, CFL 0.6 (examples/1D_sod_shocktube set-up with
    WENO5 + HLLC)."""
    cfg = SimulationConfig(dim=1, nGhost=3, RKOrder=3, reconOrder=recon)
    cfg.explicitParams.cfl = cfl
    cfg.validate()
    mesh = uniform_mesh(cfg, n, 0.0, 1.0)
    mesh.setBoundaryCondition(RectilinearMesh.XLow, BoundaryCondition.Outflow)
    mesh.setBoundaryCondition(RectilinearMesh.XHigh, BoundaryCondition.Outflow)
    st = SolutionState()
    st.allocate(mesh.totalCells(), cfg)
    eos = IdealGasEOS(1.4, 287.0, cfg)
    xc, _, _ = mesh.centroids()
    left = xc < 0.5
    shape = (1, 1, n)
    rho = np.where(left, 1.0, 0.125).reshape(shape)
    p = np.where(left, 1.0, 0.1).reshape(shape)
    u = [np.zeros(shape)]
    _fill(mesh, st, eos, rho, u, p)
    return cfg, mesh, st, eos, riemann_by_name(riemann, eos)


def isentropic_vortex_2d(n: int = 1024, riemann: str = "hllc", recon=ReconstructionOrder.WENO5, cfl: float = 0.6,
                         beta: float = 5.0, L: float = 10.0, center=(5.0, 5.0), gamma: float = 1.4,
                         periodic: bool = True):
    """Periodic isentropic vortex on [0,L]^2 with freestream rho=p=u=v=1:
    T = 1 - (g-1) beta^2/(8 g pi^2) e^{1-r^2}, rho = T^{1/(g-1)},
    p = T^{g/(g-1)}, (u,v) = (1,1) + beta/(2 pi) e^{(1-r^2)/2} (-dy, dx).
    periodic=False gives the same initial state between outflow faces."""
    cfg = SimulationConfig(dim=2, nGhost=3, RKOrder=3, reconOrder=recon)
    cfg.explicitParams.cfl = cfl
    cfg.explicitParams.maxDt = 10.0
    cfg.validate()
    mesh = uniform_mesh(cfg, n, 0.0, L, n, 0.0, L, periods=(1, 1, 0) if periodic else (0, 0, 0))
    st = SolutionState()
    st.allocate(mesh.totalCells(), cfg)
    eos = IdealGasEOS(gamma, 1.0, cfg)
    rho, u, p = vortex_fields(mesh, beta, center, gamma)
    _fill(mesh, st, eos, rho, u, p)
    return cfg, mesh, st, eos, riemann_by_name(riemann, eos)


def vortex_fields(mesh: RectilinearMesh, beta=5.0, center=(5.0, 5.0), gamma=1.4, shift=(0.0, 0.0), L=None):
    """Primitive vortex fields on the interior (1, ny, nx); `shift` advects the
    centre (with periodic wrap when L is given) for exact-solution checks."""
    xc, yc, _ = mesh.centroids()
    X, Y = np.meshgrid(xc, yc)
    cx, cy = center[0] + shift[0], center[1] + shift[1]
    dx, dy = X - cx, Y - cy
    if L is not None:
        dx -= L * np.round(dx / L)
        dy -= L * np.round(dy / L)
    r2 = dx * dx + dy * dy
    g = gamma
    T = 1.0 - (g - 1.0) * beta * beta / (8.0 * g * math.pi * math.pi) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (g - 1.0))
    p = T ** (g / (g - 1.0))
    f = np.exp(0.5 * (1.0 - r2))
    u = 1.0 - beta / (2.0 * math.pi) * dy * f
    v = 1.0 + beta / (2.0 * math.pi) * dx * f
    sh = (1,) + rho.shape
    return rho.reshape(sh), [u.reshape(sh), v.reshape(sh)], p.reshape(sh)


def shock_tube(n: int, left, right, x0: float = 0.5, riemann: str = "hllc", cfl: float = 0.6, gamma: float = 1.4,
               maxDt: float = 1.0):
    """Riemann problem on [0,1] with n cells: (rho, u, p) = left for cell
    centroids x < x0, right otherwise; outflow ends, ideal gas, WENO5 +
    SSP-RK3 (the 1D_sod_shocktube set-up with arbitrary states)."""
    cfg = SimulationConfig(dim=1, nGhost=3, RKOrder=3, reconOrder=ReconstructionOrder.WENO5)
    cfg.explicitParams.cfl = cfl
    cfg.explicitParams.maxDt = maxDt
    cfg.validate()
    mesh = uniform_mesh(cfg, n, 0.0, 1.0)
    mesh.setBoundaryCondition(RectilinearMesh.XLow, BoundaryCondition.Outflow)
    mesh.setBoundaryCondition(RectilinearMesh.XHigh, BoundaryCondition.Outflow)
    st = SolutionState()
    st.allocate(mesh.totalCells(), cfg)
    eos = IdealGasEOS(gamma, 287.0, cfg)
    xc, _, _ = mesh.centroids()
    lft = xc < x0
    shape = (1, 1, n)
    rho = np.where(lft, left[0], right[0]).reshape(shape)
    u = [np.where(lft, left[1], right[1]).reshape(shape)]
    p = np.where(lft, left[2], right[2]).reshape(shape)
    _fill(mesh, st, eos, rho, u, p)
    return cfg, mesh, st, eos, riemann_by_name(riemann, eos)


ENSEMBLE_SEED = 20261202


def riemann_ensemble_params(ntubes: int = 4096, seed: int = ENSEMBLE_SEED):
    """Batched Riemann ensemble table: cells uniform over {256, 512, ..., 4096},
    rho_L, rho_R log-uniform in [0.1, 10], p_L, p_R log-uniform in [0.01, 100],
    u_L, u_R uniform in [-1, 1], interface x0 uniform in [0.3, 0.7]
    (numpy default_rng(seed), drawn in this column order)."""
    rng = np.random.default_rng(seed)
    n = 256 * rng.integers(1, 17, size=ntubes)
    rho = 10.0 ** rng.uniform(-1.0, 1.0, size=(ntubes, 2))
    p = 10.0 ** rng.uniform(-2.0, 2.0, size=(ntubes, 2))
    u = rng.uniform(-1.0, 1.0, size=(ntubes, 2))
    x0 = rng.uniform(0.3, 0.7, size=ntubes)
    return {"ncells": n.astype(np.int64), "rho": rho, "u": u, "p": p, "x0": x0}


ADJOINT_SEED = 20261203


def perturbed_vortex_2d(n: int = 512, seed: int = ADJOINT_SEED, amplitude: float = 1e-3, ncenters: int = 8,
                        width: float = 0.5, riemann: str = "rusanov", cfl: float = 0.6):
    """Differentiable-run set-up: the isentropic vortex of
    isentropic_vortex_2d with its density multiplied by
    1 + amplitude * sum_m exp(-|x - c_m|^2 / width^2) (velocity and pressure
    unchanged; conservative fields rebuilt), centres c_m uniform in the
    domain from numpy default_rng(seed).  The time step is fixed (constDt)
    at the CFL dt of the initial state, as the adjoint needs a dt that does
    not depend on the state."""
    cfg, mesh, st, eos, rs = isentropic_vortex_2d(n, riemann=riemann, cfl=cfl)
    rho, u, p = vortex_fields(mesh, 5.0, (5.0, 5.0), 1.4)
    rng = np.random.default_rng(seed)
    cen = rng.uniform(0.0, 10.0, size=(ncenters, 2))
    xc, yc, _ = mesh.centroids()
    Y, X = np.meshgrid(yc, xc, indexing="ij")
    bump = np.zeros_like(X)
    for cx, cy in cen:
        bump += np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / width ** 2)
    rho = rho * (1.0 + amplitude * bump.reshape(rho.shape))
    _fill(mesh, st, eos, rho, u, p)
    # constDt = CFL dt of the initial state (computeAcousticTimeStep form)
    I = mesh.interior
    c = np.sqrt(1.4 * np.maximum(I(st.pres), 1e-14) / np.maximum(I(st.rho), 1e-14))
    h = 10.0 / n
    dtc = np.minimum(h / (np.abs(I(st.velU)) + c), h / (np.abs(I(st.velV)) + c)).min()
    cfg.explicitParams.constDt = float(min(cfg.explicitParams.maxDt, cfl * dtc))
    return cfg, mesh, st, eos, rs, cen
