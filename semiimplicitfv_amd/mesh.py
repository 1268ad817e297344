"""Host-side mirror of RectilinearMesh (include/RectilinearMesh.hpp:30-162).

Geometry comes from node arrays exactly as the reference builds it
(extended nodes mirror the edge cell widths into the ghost region,
src/RectilinearMesh.cpp:91-114), so cell widths and face/volume
coefficients agree bit for bit with the reference.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .config import BoundaryCondition, SimulationConfig


def runtime_linspace(a: float, b: float, n: int) -> np.ndarray:
    """Runtime's node spacing (src/Runtime.cpp:12-17): a + (b-a)*i/(n-1)."""
    i = np.arange(n, dtype=np.float64)
    return a + (b - a) * i / (n - 1)


def _extended(phys: np.ndarray, n: int, ng: int) -> np.ndarray:
    e = np.zeros(n + 2 * ng + 1)
    e[ng:ng + n + 1] = phys
    for g in range(1, ng + 1):
        m = min(g - 1, n - 1)
        e[ng - g] = e[ng - g + 1] - (phys[m + 1] - phys[m])
    for g in range(1, ng + 1):
        m = max(n - g, 0)
        e[ng + n + g] = e[ng + n + g - 1] + (phys[m + 1] - phys[m])
    return e


class RectilinearMesh:
    XLow, XHigh, YLow, YHigh, ZLow, ZHigh = range(6)

    def __init__(self, config: SimulationConfig, xNodes, yNodes=(0.0, 1.0), zNodes=(0.0, 1.0),
                 periods=(0, 0, 0)):
        dim = config.dim
        if dim < 1 or dim > 3:
            raise ValueError("RectilinearMesh: dim must be 1, 2, or 3")
        if config.nGhost < 0:
            raise ValueError("RectilinearMesh: nGhost must be non-negative")
        self.xNodes = np.ascontiguousarray(xNodes, dtype=np.float64)
        self.yNodes = np.ascontiguousarray(yNodes if dim >= 2 else (0.0, 1.0), dtype=np.float64)
        self.zNodes = np.ascontiguousarray(zNodes if dim >= 3 else (0.0, 1.0), dtype=np.float64)
        if len(self.xNodes) < 2:
            raise ValueError("RectilinearMesh: xNodes must have at least 2 entries")
        self.dim_ = dim
        self.nGhost_ = config.nGhost
        self.nx_ = len(self.xNodes) - 1
        self.ny_ = len(self.yNodes) - 1
        self.nz_ = len(self.zNodes) - 1
        self.ngx_ = self.nGhost_
        self.ngy_ = self.nGhost_ if dim >= 2 else 0
        self.ngz_ = self.nGhost_ if dim >= 3 else 0
        self.xe = _extended(self.xNodes, self.nx_, self.ngx_)
        self.ye = _extended(self.yNodes, self.ny_, self.ngy_)
        self.ze = _extended(self.zNodes, self.nz_, self.ngz_)
        self.bc = [BoundaryCondition.Outflow] * 6
        self.periods = [int(p) for p in periods]

    # ---- topology (include/RectilinearMesh.hpp:48-73)
    def dim(self): return self.dim_
    def nx(self): return self.nx_
    def ny(self): return self.ny_
    def nz(self): return self.nz_
    def nGhost(self): return self.nGhost_
    def ngx(self): return self.ngx_
    def ngy(self): return self.ngy_
    def ngz(self): return self.ngz_
    def nxTotal(self): return self.nx_ + 2 * self.ngx_
    def nyTotal(self): return self.ny_ + 2 * self.ngy_
    def nzTotal(self): return self.nz_ + 2 * self.ngz_
    def totalCells(self): return self.nxTotal() * self.nyTotal() * self.nzTotal()

    def index(self, i, j=0, k=0):
        return (i + self.ngx_) + self.nxTotal() * ((j + self.ngy_) + self.nyTotal() * (k + self.ngz_))

    def dx(self, i): return self.xe[i + self.ngx_ + 1] - self.xe[i + self.ngx_]
    def dy(self, j): return self.ye[j + self.ngy_ + 1] - self.ye[j + self.ngy_]
    def dz(self, k): return self.ze[k + self.ngz_ + 1] - self.ze[k + self.ngz_]

    def cellCentroidX(self, i): return 0.5 * (self.xe[i + self.ngx_] + self.xe[i + self.ngx_ + 1])
    def cellCentroidY(self, j): return 0.5 * (self.ye[j + self.ngy_] + self.ye[j + self.ngy_ + 1])
    def cellCentroidZ(self, k): return 0.5 * (self.ze[k + self.ngz_] + self.ze[k + self.ngz_ + 1])

    def centroids(self):
        """Interior cell centroids as 1D arrays (x, y, z)."""
        xc = 0.5 * (self.xe[self.ngx_:self.ngx_ + self.nx_] + self.xe[self.ngx_ + 1:self.ngx_ + self.nx_ + 1])
        yc = 0.5 * (self.ye[self.ngy_:self.ngy_ + self.ny_] + self.ye[self.ngy_ + 1:self.ngy_ + self.ny_ + 1])
        zc = 0.5 * (self.ze[self.ngz_:self.ngz_ + self.nz_] + self.ze[self.ngz_ + 1:self.ngz_ + self.nz_ + 1])
        return xc, yc, zc

    def shape3(self):
        """Padded array shape as (nzT, nyT, nxT) for numpy views."""
        return (self.nzTotal(), self.nyTotal(), self.nxTotal())

    def interior(self, a: np.ndarray) -> np.ndarray:
        """Interior view (nz, ny, nx) of a padded flat field."""
        v = a.reshape(self.shape3())
        return v[self.ngz_:self.ngz_ + self.nz_, self.ngy_:self.ngy_ + self.ny_, self.ngx_:self.ngx_ + self.nx_]

    # ---- boundary conditions
    def setBoundaryCondition(self, face: int, bc: BoundaryCondition):
        """Runtime::setBoundaryCondition (src/Runtime.cpp:123-127): a no-op on
        faces that are periodic through `periods` (not physical boundaries)."""
        if face < 0 or face > 5:
            raise IndexError("RectilinearMesh::setBoundaryCondition: face must be 0-5")
        if self.periods[face // 2]:
            return
        self.bc[face] = BoundaryCondition(bc)

    def boundaryCondition(self, face: int):
        return self.bc[face]

    # ---- C ABI view
    def to_c(self) -> _abi.Mesh_C:
        m = _abi.Mesh_C()
        m.dim = self.dim_
        m.nx, m.ny, m.nz = self.nx_, self.ny_, self.nz_
        m.nGhost = self.nGhost_
        P = C.POINTER(C.c_double)
        m.xNodes = self.xNodes.ctypes.data_as(P)
        m.yNodes = self.yNodes.ctypes.data_as(P)
        m.zNodes = self.zNodes.ctypes.data_as(P)
        for f in range(6):
            m.bc[f] = int(self.bc[f])
        for d in range(3):
            m.periods[d] = int(self.periods[d])
        m._keep = (self.xNodes, self.yNodes, self.zNodes)
        return m


def uniform_mesh(config: SimulationConfig, nx: int, xMin: float, xMax: float, ny: int = 1, yMin: float = 0.0,
                 yMax: float = 1.0, nz: int = 1, zMin: float = 0.0, zMax: float = 1.0,
                 periods=(0, 0, 0)) -> RectilinearMesh:
    """Runtime::createUniformMesh on one rank (src/Runtime.cpp:39-119)."""
    xn = runtime_linspace(xMin, xMax, nx + 1)
    yn = runtime_linspace(yMin, yMax, ny + 1) if config.dim >= 2 else np.array([0.0, 1.0])
    zn = runtime_linspace(zMin, zMax, nz + 1) if config.dim >= 3 else np.array([0.0, 1.0])
    return RectilinearMesh(config, xn, yn, zn, periods=periods)
