"""Body forces -- mirror of BodyForce.jl: the constant force functors
(:56-60, :86-92) and the manufactured vortex force VortexBodyForce2D
(:34-53, :94-121).  The step evaluates them on the device
(jump_set_body_force); __call__ here evaluates the same formula on the host,
for callers that want the value."""
import math

import numpy as np


class BodyForceType:
    pass


class ConstantBodyForce2D(BodyForceType):
    def __init__(self, g):
        self.g = np.asarray(g, dtype=np.float64).reshape(2)

    def __call__(self, body, x, t):
        body[:] = self.g


class ConstantBodyForce3D(BodyForceType):
    def __init__(self, g):
        self.g = np.asarray(g, dtype=np.float64).reshape(3)

    def __call__(self, body, x, t):
        body[:] = self.g


class VortexBodyForce2D(BodyForceType):
    """VortexBodyForce2D(G, T, Ri, Ro, rho0, mu, xcenter, ycenter): the body
    force under which an annulus Ri < r < Ro of a Neo-Hookean solid rotates
    by alpha = G sin(pi t / T) h(r) about (xcenter, ycenter)
    (BodyForce.jl:34-53, :94-121; the exact map of Main_Vortex_FEMP.jl)."""

    def __init__(self, G, T, Ri, Ro, rho0, mu, xcenter, ycenter):
        self.G, self.T, self.Ri, self.Ro = float(G), float(T), float(Ri), float(Ro)
        self.rho0, self.mu, self.xcenter, self.ycenter = float(rho0), float(mu), float(xcenter), float(ycenter)
        self.R = (self.Ri + self.Ro) / 2

    def params(self):
        return [self.G, self.T, self.Ri, self.Ro, self.rho0, self.mu, self.xcenter, self.ycenter]

    def rotation(self, X, t):
        """The exact map: the positions at time t of material points X (n, 2)."""
        X = np.asarray(X, dtype=np.float64)
        x, y = X[:, 0] - self.xcenter, X[:, 1] - self.ycenter
        r = np.sqrt(x * x + y * y)
        s = (r - self.R) / (self.Ri - self.Ro)
        alpha = self.G * math.sin(math.pi * t / self.T) * (1 - 8 * s ** 2 + 16 * s ** 4)
        c, sn = np.cos(alpha), np.sin(alpha)
        return np.stack([x * c - y * sn + self.xcenter, x * sn + y * c + self.ycenter], axis=1)

    def __call__(self, body, xp, t):
        G, T, R, Ri, Ro = self.G, self.T, self.R, self.Ri, self.Ro
        x, y = xp[0] - self.xcenter, xp[1] - self.ycenter
        r = math.sqrt(x * x + y * y)
        theta = math.atan2(y, x)
        h = 1 - 8 * ((r - R) / (Ri - Ro)) ** 2 + 16 * ((r - R) / (Ri - Ro)) ** 4
        hp = -16 * (r - R) / (Ri - Ro) ** 2 + 16 * 4 * (r - R) ** 3 / (Ri - Ro) ** 4
        hpp = -16 / (Ri - Ro) ** 2 + 16 * 4 * 3 * (r - R) ** 2 / (Ri - Ro) ** 4
        g = G * math.sin(math.pi * t / T)
        gp = G * math.pi / T * math.cos(math.pi * t / T)
        gpp = -math.pi ** 2 / T ** 2 * g
        alpha = g * h
        mdr = self.mu / self.rho0
        br = (mdr * (3 * g * hp + r * g * hpp) - r * gpp * h) * math.sin(alpha) + \
            (mdr * r * (g * hp) ** 2 - r * (gp * h) ** 2) * math.cos(alpha)
        bt = (-mdr * (3 * g * hp + r * g * hpp) + r * gpp * h) * math.cos(alpha) + \
            (mdr * r * (g * hp) ** 2 + r * (gp * h) ** 2) * math.sin(alpha)
        body[0] = br * math.cos(theta) - bt * math.sin(theta)
        body[1] = br * math.sin(theta) + bt * math.cos(theta)
