"""Body forces -- mirror of BodyForce.jl:56-91 (the constant force functors
used on this path).  The constant value is passed to the GPU step."""
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
