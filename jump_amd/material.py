"""Material types -- mirror of Material.jl (host side).

Constructors derive the Lame constants exactly as the reference does; history
arrays (alpha, vmStr, pstrain, alphadot) receive the device state after a
solve.  `to_c()` gives the ABI struct.
"""
import math

import numpy as np

from . import _ffi
from .engine import make_material


class MaterialType:
    kind = None


class RigidMaterial(MaterialType):
    """Material.jl:29-31 holds only the density, but the 2D USL solver reads
    mat.vx, mat.vy and mat.fixed (SolveExplicitSolidMechanics2D.jl:453-461) and
    the drivers call RigidMaterial(vx, vy, density)
    (Main_LateralCompression_Cellular.jl:58-59): this type follows the call
    sites."""

    kind = _ffi.MAT_RIGID

    def __init__(self, *args, fixed=False):
        if len(args) == 1:
            self.vx, self.vy, self.density = 0.0, 0.0, float(args[0])
        elif len(args) == 3:
            self.vx, self.vy, self.density = float(args[0]), float(args[1]), float(args[2])
        else:
            raise TypeError("RigidMaterial(density) or RigidMaterial(vx, vy, density)")
        self.fixed = bool(fixed)

    def to_c(self):
        return make_material(self.kind, density=self.density, vx=self.vx, vy=self.vy, fixed=self.fixed)


def _lame(E, nu):
    lam = E * nu / (1 + nu) / (1 - 2 * nu)
    mu = E / 2 / (1 + nu)
    return lam, mu


class ElasticMaterial(MaterialType):
    """Material.jl:37-64.  ElasticMaterial(E, nu, density, Gf, l, parCount) or
    ElasticMaterial(E, nu, density, parCount)."""

    kind = _ffi.MAT_ELASTIC

    def __init__(self, E, nu, density, *rest):
        if len(rest) == 3:
            Gf, l0, parCount = rest
        elif len(rest) == 1:
            Gf, l0, parCount = 0.0, 0.0, rest[0]
        else:
            raise TypeError("ElasticMaterial(E, nu, density, [Gf, l,] parCount)")
        self.E, self.nu, self.density = float(E), float(nu), float(density)
        self.lambda_, self.mu = _lame(self.E, self.nu)
        setattr(self, "lambda", self.lambda_)
        self.Gf, self.l0, self.fac = float(Gf), float(l0), 0.5 / self.E
        self.vmStr = np.zeros(int(parCount))
        self.dam = np.zeros(int(parCount))
        self.c_dil = math.sqrt((self.lambda_ + 2 * self.mu) / self.density)

    def to_c(self):
        return make_material(self.kind, E=self.E, nu=self.nu, density=self.density)


class NeoHookeanMaterial(MaterialType):
    """Material.jl:100-117"""

    kind = _ffi.MAT_NEOHOOKEAN

    def __init__(self, E, nu, density, parCount):
        self.E, self.nu, self.density = float(E), float(nu), float(density)
        self.lambda_, self.mu = _lame(self.E, self.nu)
        setattr(self, "lambda", self.lambda_)
        self.vmStr = np.zeros(int(parCount))

    def to_c(self):
        return make_material(self.kind, E=self.E, nu=self.nu, density=self.density)


class ElastoPlasticMaterial(MaterialType):
    """Material.jl:128-171 (J2 with linear isotropic hardening)."""

    kind = _ffi.MAT_ELASTOPLASTIC

    def __init__(self, E, nu, density, fy, k1, parCount):
        self.E, self.nu, self.density = float(E), float(nu), float(density)
        self.lambda_, self.mu = _lame(self.E, self.nu)
        setattr(self, "lambda", self.lambda_)
        self.kappa = self.lambda_ + self.mu
        self.fy, self.k1 = float(fy), float(k1)
        n = int(parCount)
        self.alpha = np.zeros(n)
        self.vmStr = np.zeros(n)
        self.pstrain = np.zeros((n, 3))
        self.pstrain_3D = np.zeros((n, 3, 3))

    def to_c(self):
        return make_material(self.kind, E=self.E, nu=self.nu, density=self.density, fy=self.fy, k1=self.k1)


class JohnsonCookMaterial(MaterialType):
    """Material.jl:280-330"""

    kind = _ffi.MAT_JOHNSONCOOK

    def __init__(self, E, nu, density, A, B, C, n, eps0dot, m, chi, Cp, cellsize, parCount):
        self.E, self.nu, self.density = float(E), float(nu), float(density)
        self.A, self.B, self.C, self.n = float(A), float(B), float(C), float(n)
        self.epsdot0, self.m, self.chi, self.Cp = float(eps0dot), float(m), float(chi), float(Cp)
        self.lambda_, self.mu = _lame(self.E, self.nu)
        setattr(self, "lambda", self.lambda_)
        self.kappa = self.lambda_ + self.mu
        self.signal_velocity = math.sqrt((self.lambda_ + 2 * self.mu) / self.density)
        self.cellsize = float(cellsize)
        self.delta = self.chi / (self.density * self.Cp)
        k = int(parCount)
        self.alpha = np.zeros(k)
        self.alphadot = np.zeros(k)
        self.vmStr = np.zeros(k)
        self.c_dil = self.signal_velocity

    def to_c(self):
        return make_material(self.kind, E=self.E, nu=self.nu, density=self.density, A=self.A, B=self.B, C=self.C,
                             n=self.n, epsdot0=self.epsdot0, m=self.m, chi=self.chi, Cp=self.Cp,
                             cellsize=self.cellsize)
