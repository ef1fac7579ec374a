"""fix_Dirichlet_grid -- mirror of Dirichlet.jl:28-82.

data["dirichlet_grid"] = [("bottom", (0, 1, 0)), ...] sets grid.fixedNodes on
the named faces (0-based node lists from the Grid constructors)."""
import numpy as np

from .grid import Grid2D, Grid3D


def fix_Dirichlet_grid(grid, data, ghostcell=False):
    if "dirichlet_grid" not in data:
        return
    d = grid.dim
    for bnd, fix in data["dirichlet_grid"]:
        fix = np.asarray(fix, dtype=np.int64)[:d].reshape(d, 1)
        if isinstance(grid, Grid3D):
            faces = {"bottom": grid.bottomNodes, "top": grid.topNodes, "left": grid.leftNodes,
                     "right": grid.rightNodes, "front": grid.frontNodes, "back": grid.backNodes}
            if bnd not in faces:
                raise ValueError("Not recognized grid faces!!!")
            grid.fixedNodes[:, faces[bnd]] = fix
        elif isinstance(grid, Grid2D):
            if bnd == "bottom":
                grid.fixedNodes[:, grid.bottomNodes] = fix
                if ghostcell:
                    grid.fixedNodes[:, grid.bottomNodes + grid.nodeCountX] = fix
            elif bnd == "top":
                grid.fixedNodes[:, grid.topNodes] = fix
                if ghostcell:
                    grid.fixedNodes[:, grid.topNodes - grid.nodeCountX] = fix
            elif bnd == "left":
                grid.fixedNodes[:, grid.leftNodes] = fix
                if ghostcell:
                    grid.fixedNodes[:, grid.leftNodes + 1] = fix
            elif bnd == "right":
                grid.fixedNodes[:, grid.rightNodes] = fix
                if ghostcell:
                    grid.fixedNodes[:, grid.rightNodes - 1] = fix
            else:
                raise ValueError("Not recognized grid faces!!!")
        else:
            raise TypeError("unknown grid type")
