"""Grid2D / Grid3D -- mirror of Grid.jl (host side).

Node arrays live on the GPU during a solve; these objects keep the grid
geometry, the Dirichlet mask `fixedNodes` (d x nodeCount, Grid.jl:130,223) and
the face node lists, and receive the node arrays back after a solve.  The
dense 2D mass matrix `M` (Grid.jl:118) is not allocated: USL never reads it.
Node indices here are 0-based (the reference's are 1-based).
"""
import numpy as np


class Grid2D:
    """Grid.jl:104-158"""

    def __init__(self, xmin, xmax, ymin, ymax, iN_x, iN_y):
        self.xmin, self.ymin = float(xmin), float(ymin)
        self.xmax, self.ymax = float(xmax), float(ymax)
        self.lx, self.ly = float(xmax) - float(xmin), float(ymax) - float(ymin)
        self.dx = (float(xmax) - float(xmin)) / float(iN_x - 1.0)
        self.dy = (float(ymax) - float(ymin)) / float(iN_y - 1.0)
        self.dxI = 1.0 / self.dx
        self.dyI = 1.0 / self.dy
        self.nodeCountX, self.nodeCountY = int(iN_x), int(iN_y)
        self.nodeCount = self.nodeCountX * self.nodeCountY
        self.cellCount = (iN_x - 1) * (iN_y - 1)
        i = np.arange(iN_x, dtype=np.float64)
        j = np.arange(iN_y, dtype=np.float64)
        jj, ii = np.meshgrid(j, i, indexing="ij")
        self.pos = np.stack([self.xmin + ii.ravel() * self.dx, self.ymin + jj.ravel() * self.dy], axis=1)
        n = self.nodeCount
        self.mass = np.zeros(n)
        self.momentum0 = np.zeros((n, 2))
        self.momentum = np.zeros((n, 2))
        self.momentum2 = np.zeros((n, 2))
        self.force = np.zeros((n, 2))
        self.fixedNodes = np.zeros((2, n), dtype=np.int64)
        # Grid.jl:134-144 (0-based)
        self.bottomNodes = np.arange(iN_x)
        self.topNodes = np.arange(iN_x) + iN_x * (iN_y - 1)
        self.rightNodes = iN_x * np.arange(1, iN_y + 1) - 1
        self.leftNodes = iN_x * np.arange(iN_y)

    @property
    def dim(self):
        return 2

    @property
    def lo(self):
        return (self.xmin, self.ymin)

    @property
    def hi(self):
        return (self.xmax, self.ymax)

    @property
    def nodes(self):
        return (self.nodeCountX, self.nodeCountY)


class Grid3D:
    """Grid.jl:198-267"""

    def __init__(self, xmin, xmax, ymin, ymax, zmin, zmax, iN_x, iN_y, iN_z):
        self.xmin, self.ymin, self.zmin = float(xmin), float(ymin), float(zmin)
        self.xmax, self.ymax, self.zmax = float(xmax), float(ymax), float(zmax)
        self.lx, self.ly, self.lz = self.xmax - self.xmin, self.ymax - self.ymin, self.zmax - self.zmin
        self.dx = (self.xmax - self.xmin) / float(iN_x - 1.0)
        self.dy = (self.ymax - self.ymin) / float(iN_y - 1.0)
        self.dz = (self.zmax - self.zmin) / float(iN_z - 1.0)
        self.dxI, self.dyI, self.dzI = 1.0 / self.dx, 1.0 / self.dy, 1.0 / self.dz
        self.nodeCountX, self.nodeCountY, self.nodeCountZ = int(iN_x), int(iN_y), int(iN_z)
        self.nodeCountXY = self.nodeCountX * self.nodeCountY
        self.nodeCount = self.nodeCountXY * self.nodeCountZ
        n = self.nodeCount
        self.mass = np.zeros(n)
        self.momentum0 = np.zeros((n, 3))
        self.momentum = np.zeros((n, 3))
        self.momentum2 = np.zeros((n, 3))
        self.force = np.zeros((n, 3))
        self.fixedNodes = np.zeros((3, n), dtype=np.int64)
        nx, ny, nz = iN_x, iN_y, iN_z
        k = np.arange(nz)
        # Grid.jl:226-249 (0-based)
        self.bottomNodes = (np.arange(nx)[None, :] + (k * nx * ny)[:, None]).ravel()
        self.topNodes = (np.arange(nx)[None, :] + nx * (ny - 1) + (k * nx * ny)[:, None]).ravel()
        self.rightNodes = ((nx * np.arange(1, ny + 1) - 1)[None, :] + (k * nx * ny)[:, None]).ravel()
        self.leftNodes = ((nx * np.arange(ny))[None, :] + (k * nx * ny)[:, None]).ravel()
        self.frontNodes = np.arange(nx * ny)
        self.backNodes = np.arange(nx * ny) + (nz - 1) * nx * ny

    @property
    def pos(self):
        i = np.arange(self.nodeCountX, dtype=np.float64)
        j = np.arange(self.nodeCountY, dtype=np.float64)
        k = np.arange(self.nodeCountZ, dtype=np.float64)
        kk, jj, ii = np.meshgrid(k, j, i, indexing="ij")
        return np.stack([self.xmin + ii.ravel() * self.dx, self.ymin + jj.ravel() * self.dy,
                         self.zmin + kk.ravel() * self.dz], axis=1)

    @property
    def dim(self):
        return 3

    @property
    def lo(self):
        return (self.xmin, self.ymin, self.zmin)

    @property
    def hi(self):
        return (self.xmax, self.ymax, self.zmax)

    @property
    def nodes(self):
        return (self.nodeCountX, self.nodeCountY, self.nodeCountZ)


def to_1D_index(i, j, grid, k=None):
    """Grid.jl:275-291, 0-based indices."""
    if k is None:
        return grid.nodeCountX * j + i
    return grid.nodeCountXY * k + grid.nodeCountX * j + i
