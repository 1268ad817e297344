"""Forward-mode (tangent) restatement of the explicit 2D/1D step for
gradient checks.  Test infrastructure only: the checker of the discrete
adjoint, never on the product path.

The primal part follows the same algorithm as the C oracle (sifv_oracle.c):
cons -> prim (SolutionState::convertConservativeToPrimitiveVariables,
src/SolutionState.cpp:187-205; IdealGasEOS::toPrimitive,
src/IdealGasEOS.cpp:46-64), WENO5 reconstruction of rho, u, v, p
(Reconstructor::weno5Left/Right, src/Reconstruction.cpp:119-168), Rusanov
flux (computeRusanovFlux, src/RusanovSolver.cpp:7-83), flux divergence
(ExplicitSolver::computeRHS, src/ExplicitSolver.cpp:270-351) and the SSP-RK3
update with a fixed dt (src/ExplicitSolver.cpp:93-190) on a doubly periodic
uniform mesh.  Every quantity is carried as a (value, tangent) pair of
numpy arrays, so one run gives the exact directional derivative of the
final state along a direction in (U0, gamma).

Pins: the primal part equals the C oracle's step to rounding
(tests/test_adjoint_cpu.py), and the tangent equals central differences of
the primal part to truncation order.
"""
from __future__ import annotations

import numpy as np


class D:
    """Dual number over numpy arrays: value v, tangent t."""
    __slots__ = ("v", "t")

    def __init__(self, v, t=None):
        self.v = np.asarray(v, dtype=np.float64)
        self.t = np.zeros_like(self.v) if t is None else np.asarray(t, dtype=np.float64)

    def __add__(self, o):
        o = _d(o)
        return D(self.v + o.v, self.t + o.t)

    __radd__ = __add__

    def __sub__(self, o):
        o = _d(o)
        return D(self.v - o.v, self.t - o.t)

    def __rsub__(self, o):
        return _d(o) - self

    def __mul__(self, o):
        o = _d(o)
        return D(self.v * o.v, self.t * o.v + self.v * o.t)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = _d(o)
        q = self.v / o.v
        return D(q, (self.t - q * o.t) / o.v)

    def __rtruediv__(self, o):
        return _d(o) / self

    def __neg__(self):
        return D(-self.v, -self.t)

    def shift(self, s, axis):
        return D(np.roll(self.v, s, axis), np.roll(self.t, s, axis))


def _d(x):
    return x if isinstance(x, D) else D(x)


def dsqrt(x):
    r = np.sqrt(x.v)
    return D(r, 0.5 * x.t / r)


def dabs(x):
    s = np.where(x.v < 0, -1.0, 1.0)
    return D(np.abs(x.v), s * x.t)


def dmax(a, b):
    a, b = _d(a), _d(b)
    m = a.v < b.v  # std::max(a, b) returns b when a < b
    return D(np.where(m, b.v, a.v), np.where(m, b.t, a.t))


def weno5_plus(v0, v1, v2, v3, v4, eps):
    """Left-biased WENO5 value at the cell's high face (weno5Left)."""
    b0 = (13.0 / 12.0) * (v0 - 2.0 * v1 + v2) * (v0 - 2.0 * v1 + v2) + 0.25 * (v0 - 4.0 * v1 + 3.0 * v2) * (
        v0 - 4.0 * v1 + 3.0 * v2)
    b1 = (13.0 / 12.0) * (v1 - 2.0 * v2 + v3) * (v1 - 2.0 * v2 + v3) + 0.25 * (v1 - v3) * (v1 - v3)
    b2 = (13.0 / 12.0) * (v2 - 2.0 * v3 + v4) * (v2 - 2.0 * v3 + v4) + 0.25 * (3.0 * v2 - 4.0 * v3 + v4) * (
        3.0 * v2 - 4.0 * v3 + v4)
    a0 = 0.1 / ((eps + b0) * (eps + b0))
    a1 = 0.6 / ((eps + b1) * (eps + b1))
    a2 = 0.3 / ((eps + b2) * (eps + b2))
    p0 = (1.0 / 3.0) * v0 - (7.0 / 6.0) * v1 + (11.0 / 6.0) * v2
    p1 = -(1.0 / 6.0) * v1 + (5.0 / 6.0) * v2 + (1.0 / 3.0) * v3
    p2 = (1.0 / 3.0) * v2 + (5.0 / 6.0) * v3 - (1.0 / 6.0) * v4
    return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2)


def prims(U, gamma, dim):
    rho = U[0]
    rs = dmax(rho, 1e-14)
    u = [U[1 + d] / rs for d in range(dim)]
    ke = u[0] * u[0]
    for d in range(1, dim):
        ke = ke + u[d] * u[d]
    ke = ke * (0.5 * rs)
    p = dmax((gamma - 1.0) * (U[dim + 1] - ke), 1e-14)
    return rho, u, p


def rusanov(L, R, n, gamma, dim):
    """L, R = (rho, [u], p) face states; n = normal axis."""
    rL, uL, pL = L
    rR, uR, pR = R
    cL = dsqrt(gamma * dmax(pL, 1e-14) / dmax(rL, 1e-14))
    cR = dsqrt(gamma * dmax(pR, 1e-14) / dmax(rR, 1e-14))
    S = dmax(dabs(uL[n]) + cL, dabs(uR[n]) + cR)

    def energy(r, u, p):
        q = u[0] * u[0]
        for d in range(1, dim):
            q = q + u[d] * u[d]
        return p / (gamma - 1.0) + 0.5 * r * q

    EL, ER = energy(rL, uL, pL), energy(rR, uR, pR)
    F = [0.5 * (rL * uL[n] + rR * uR[n]) - 0.5 * S * (rR - rL)]
    for i in range(dim):
        fl = rL * uL[i] * uL[n]
        fr = rR * uR[i] * uR[n]
        if i == n:
            fl = fl + pL
            fr = fr + pR
        F.append(0.5 * (fl + fr) - 0.5 * S * (rR * uR[i] - rL * uL[i]))
    F.append(0.5 * ((EL * uL[n] + pL * uL[n]) + (ER * uR[n] + pR * uR[n])) - 0.5 * S * (ER - EL))
    return F


def rhs(U, gamma, dx, eps, dim):
    """Flux divergence on a periodic uniform grid; arrays are (ny, nx) or (nx,)."""
    rho, u, p = prims(U, gamma, dim)
    W = [rho] + u + [p]
    R = [D(np.zeros_like(U[0].v)) for _ in range(dim + 2)]
    for n in range(dim):
        axis = (U[0].v.ndim - 1) - n  # x is the last axis
        sh = [w for w in W]
        st = [[w.shift(2 - k, axis) for k in range(5)] for w in sh]  # v_k = w[i - 2 + k]
        plus = [weno5_plus(*s, eps) for s in st]
        minus = [weno5_plus(*s[::-1], eps) for s in st]
        # face i+1/2: left = plus of cell i, right = minus of cell i+1
        Rm = [m.shift(-1, axis) for m in minus]
        F = rusanov((plus[0], plus[1:1 + dim], plus[1 + dim]), (Rm[0], Rm[1:1 + dim], Rm[1 + dim]), n, gamma, dim)
        c = 1.0 / dx[n]
        for q in range(dim + 2):
            # cell i: + c F(i-1/2) - c F(i+1/2)
            R[q] = R[q] + (c * F[q].shift(1, axis) - c * F[q])
    return R


def rk3_step(U, gamma, dt, dx, eps, dim):
    R0 = rhs(U, gamma, dx, eps, dim)
    U1 = [U[q] + dt * R0[q] for q in range(dim + 2)]
    R1 = rhs(U1, gamma, dx, eps, dim)
    U2 = [(U1[q] + 3.0 * U[q] + dt * R1[q]) / 4.0 for q in range(dim + 2)]
    R2 = rhs(U2, gamma, dx, eps, dim)
    return [(2.0 * U2[q] + U[q] + 2.0 * dt * R2[q]) / 3.0 for q in range(dim + 2)]


def kinetic_energy(U, dx, dim):
    """J = sum over cells of 0.5 |rho u|^2 / rho * dV."""
    m2 = U[1] * U[1]
    for d in range(1, dim):
        m2 = m2 + U[1 + d] * U[1 + d]
    dV = 1.0
    for d in range(dim):
        dV *= dx[d]
    e = 0.5 * m2 / U[0] * dV
    return D(e.v.sum(), e.t.sum())


def run(U0, dU0, gamma, dgamma, dt, nsteps, dx, eps=1e-6):
    """U0, dU0: lists of dim+2 interior arrays.  Returns (final state as
    duals, J as a dual)."""
    dim = U0[0].ndim
    U = [D(a, t) for a, t in zip(U0, dU0)]
    g = D(np.float64(gamma), np.float64(dgamma))
    for _ in range(nsteps):
        U = rk3_step(U, g, dt, dx, eps, dim)
    return U, kinetic_energy(U, dx, dim)
