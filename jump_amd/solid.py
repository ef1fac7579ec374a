"""Solid2D / Solid3D and particle builders -- mirror of Solid.jl (host side).

Per-particle arrays use the reference's AoS memory layout so they cross the C
ABI without conversion: pos/velocity (n, d); deformationGradient, strain and
stress are stored (n, d*d) column-major per particle and exposed as (n, d, d)
views indexed [p, a, b] = M[a, b].

Float ranges follow Julia's `start:step:stop` for Float64 (rational
reconstruction of the endpoints, TwicePrecision element arithmetic); element
values are the correctly rounded exact values, which is what the TwicePrecision
computation yields except at rare last-ulp ties.
"""
import math
from fractions import Fraction

import numpy as np

from .material import RigidMaterial


# ---------------------------------------------------------------- Julia ranges
def _rat(x):
    """Base.rat: continued-fraction approximation a/b with Float64(a/b) == x."""
    y = x
    a = d = 1
    b = c = 0
    m = 16777216  # maxintfloat(Float32, Int)
    while abs(y) <= m:
        f = int(y)  # trunc
        y -= f
        a, c = f * a + c, a
        b, d = f * b + d, b
        if max(abs(a), abs(b)) > m:
            return c, d
        if float(a) / float(b) == x:
            break
        if y == 0.0:
            break
        y = 1.0 / y
    return a, b


def _tdiv(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def _between(a, x, b):
    return (a <= x <= b) or (b <= x <= a)


def julia_range(start, step, stop):
    """Elements of Julia's `start:step:stop` for Float64 arguments."""
    start, step, stop = float(start), float(step), float(stop)
    if step == 0.0:
        raise ValueError("range step cannot be zero")
    step_n, step_d = _rat(step)
    if step_d != 0 and step_n / step_d == step:
        start_n, start_d = _rat(start)
        stop_n, stop_d = _rat(stop)
        if start_d != 0 and stop_d != 0 and start_n / start_d == start and stop_n / stop_d == stop:
            den = math.lcm(start_d, step_d)
            m = 2.0 ** 53
            if den != 0 and abs(start * den) <= m and abs(step * den) <= m and den % start_d == 0 \
                    and den % step_d == 0:
                s_n = round(start * den)
                st_n = round(step * den)
                ln = max(0, _tdiv(den * stop_n - stop_d * s_n + st_n * stop_d, st_n * stop_d))
                if _between(start, start + (ln - 1) * step, stop + step / 2) and \
                        not _between(start, start + ln * step, stop):
                    return np.array([float(Fraction(s_n + i * st_n, den)) for i in range(ln)])
    lf = (stop - start) / step
    if lf < 0:
        ln = 0
    elif lf == 0:
        ln = 1
    else:
        ln = int(round(lf)) + 1
        stop2 = start + (ln - 1) * step
        ln -= int(start < stop < stop2) + int(start > stop > stop2)
    fs, fst = Fraction(start), Fraction(step)
    return np.array([float(fs + i * fst) for i in range(ln)])


# ---------------------------------------------------------------- solids
class _SolidBase:
    dim = 2

    def _init_common(self, coords, mat):
        d = self.dim
        coords = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).reshape(-1, d))
        n = coords.shape[0]
        self.parCount = n
        self.mat = mat
        self.mass = np.full(n, float(mat.density))
        self.volumeInitial = np.zeros(n)
        self.volume = np.zeros(n)
        self.pos = coords.copy()
        self.velocity = np.zeros((n, d))
        self._F = np.zeros((n, d * d))
        self._F[:, ::d + 1] = 1.0
        self._strain = np.zeros((n, d * d))
        self._stress = np.zeros((n, d * d))
        # material history (kept per solid; the reference keeps it in the
        # material indexed by particle, Material.jl:140-155)
        self.alpha = np.zeros(n)
        self.alphadot = np.zeros(n)
        self.vmStr = np.zeros(n)
        self.pstrain = np.zeros((n, 3 if d == 2 else 9))
        self.lp = None

    def _view(self, raw):
        d = self.dim
        return raw.reshape(-1, d, d).transpose(0, 2, 1)

    @property
    def deformationGradient(self):
        return self._view(self._F)

    @property
    def strain(self):
        return self._view(self._strain)

    @property
    def stress(self):
        return self._view(self._stress)


class Solid2D(_SolidBase):
    """Solid.jl:67-259.  Solid2D(coords, mat); Solid2D(nodes, elems, mat) for
    CPDI-Q4 particles (corners (nc, 2), elems (n, 4) 0-based);
    Solid2D(fileName, mat) for CPDI-Q4 particles of a Gmsh mesh
    (Solid.jl:162-222: elements oriented counter-clockwise first)."""

    dim = 2

    def __init__(self, *args, lp=None):
        if len(args) == 2 and isinstance(args[0], (str, bytes)):
            from .mesh import loadGMSH, orient_q4
            coords3, elems1 = loadGMSH(args[0])
            nodes2 = np.ascontiguousarray(coords3[:2])
            elems1 = orient_q4(nodes2, elems1)
            args = (nodes2.T, elems1 - 1, args[1])
        if len(args) == 2:
            coords, mat = args
            self._init_common(coords, mat)
            self.nodes = None
            self.elems = None
        elif len(args) == 3:
            nodes, elems, mat = args
            nodes = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).reshape(-1, 2))
            elems = np.ascontiguousarray(np.asarray(elems, dtype=np.int64).reshape(-1, 4))
            # Solid.jl:143-153: area (shoelace), mass, centre
            c = nodes[elems]  # (n, 4, 2)
            x, y = c[:, :, 0], c[:, :, 1]
            a = 0.5 * (x[:, 0] * y[:, 1] - x[:, 1] * y[:, 0] + x[:, 1] * y[:, 2] - x[:, 2] * y[:, 1]
                       + x[:, 2] * y[:, 3] - x[:, 3] * y[:, 2] + x[:, 3] * y[:, 0] - x[:, 0] * y[:, 3])
            self._init_common(c.mean(axis=1), mat)
            self.volume[:] = a
            self.volumeInitial[:] = a
            self.mass[:] = a * mat.density
            self.nodes = nodes
            self.elems = elems
        else:
            raise TypeError("Solid2D(coords, mat) or Solid2D(nodes, elems, mat)")
        self.pos0 = self.pos.copy()
        self.rigid = isinstance(self.mat, RigidMaterial)
        self.lp = None if lp is None else np.broadcast_to(np.asarray(lp, dtype=np.float64), (2,)).copy()


class Solid3D(_SolidBase):
    """Solid.jl:267-297"""

    dim = 3

    def __init__(self, coords, mat, lp=None):
        if isinstance(mat, RigidMaterial):
            raise TypeError("Solid3D has no rigid support in the reference (Solid.jl:267-297)")
        self._init_common(coords, mat)
        self.rigid = False
        self.nodes = None
        self.elems = None
        self.lp = None if lp is None else np.broadcast_to(np.asarray(lp, dtype=np.float64), (3,)).copy()


# ---------------------------------------------------------------- builders
def buildParticleForSegment(length, fOffset):
    """Solid.jl:307-315"""
    return julia_range(0.5 * fOffset, fOffset, length - 0.5 * fOffset)


def buildParticleForCircle(fCenter, fRadius, fOffset):
    """Solid.jl:322-337"""
    fRadius = math.floor(fRadius / fOffset) * fOffset
    r = julia_range(-fRadius + 0.5 * fOffset, fOffset, +fRadius - 0.5 * fOffset)
    fy, fx = np.meshgrid(r, r, indexing="ij")
    fx, fy = fx.ravel(), fy.ravel()
    keep = fx * fx + fy * fy < fRadius * fRadius
    return np.stack([fCenter[0] + fx[keep], fCenter[1] + fy[keep]], axis=1)


def buildParticleForRing(fCenter, innerRad, outerRad, fOffset):
    """Solid.jl:339-354"""
    fRadius = math.floor(outerRad / fOffset) * fOffset
    r = julia_range(-fRadius + 0.5 * fOffset, fOffset, +fRadius - 0.5 * fOffset)
    fy, fx = np.meshgrid(r, r, indexing="ij")
    fx, fy = fx.ravel(), fy.ravel()
    d2 = fx * fx + fy * fy
    keep = (d2 < outerRad * outerRad) & (d2 > innerRad * innerRad)
    return np.stack([fCenter[0] + fx[keep], fCenter[1] + fy[keep]], axis=1)


def buildParticleForSphere(fCenter, fRadius, fOffset):
    """Solid.jl:356-373"""
    fRadius = math.floor(fRadius / fOffset) * fOffset
    r = julia_range(-fRadius + 0.5 * fOffset, fOffset, +fRadius - 0.5 * fOffset)
    fz, fy, fx = np.meshgrid(r, r, r, indexing="ij")
    fx, fy, fz = fx.ravel(), fy.ravel(), fz.ravel()
    keep = fx * fx + fy * fy + fz * fz < fRadius * fRadius
    return np.stack([fCenter[0] + fx[keep], fCenter[1] + fy[keep], fCenter[2] + fz[keep]], axis=1)


def buildParticleForCylinder(fCenter, fRadius, fLength, fOffsetX, fOffsetZ):
    """Solid.jl:447-465 (axis along y)."""
    fRadius = math.floor(fRadius / fOffsetX) * fOffsetX
    ry = julia_range(-fLength + 0.5 * fOffsetZ, fOffsetZ, +fLength - 0.5 * fOffsetZ)
    rx = julia_range(-fRadius + 0.5 * fOffsetX, fOffsetX, +fRadius - 0.5 * fOffsetX)
    fy, fz, fx = np.meshgrid(ry, rx, rx, indexing="ij")
    fx, fy, fz = fx.ravel(), fy.ravel(), fz.ravel()
    keep = fx * fx + fz * fz < fRadius * fRadius
    return np.stack([fCenter[0] + fx[keep], fCenter[1] + fy[keep], fCenter[2] + fz[keep]], axis=1)


def buildParticleForRectangle(fCenter, fWidth, fHeight, fOffset):
    """Solid.jl:471-486"""
    fWidth = math.floor(fWidth / fOffset) * fOffset
    fHeight = math.floor(fHeight / fOffset) * fOffset
    ry = julia_range(-0.5 * fHeight + 0.5 * fOffset, fOffset, +0.5 * fHeight - 0.5 * fOffset)
    rx = julia_range(-0.5 * fWidth + 0.5 * fOffset, fOffset, +0.5 * fWidth - 0.5 * fOffset)
    fy, fx = np.meshgrid(ry, rx, indexing="ij")
    return np.stack([fCenter[0] + fx.ravel(), fCenter[1] + fy.ravel()], axis=1)


def buildParticleForBlock(fCenter, fWidth, fHeight, fThickness, fOffset):
    """Solid.jl:492-509 (the thickness is not floored, as in the reference)."""
    fWidth = math.floor(fWidth / fOffset) * fOffset
    fHeight = math.floor(fHeight / fOffset) * fOffset
    rz = julia_range(-0.5 * fThickness + 0.5 * fOffset, fOffset, +0.5 * fThickness - 0.5 * fOffset)
    ry = julia_range(-0.5 * fHeight + 0.5 * fOffset, fOffset, +0.5 * fHeight - 0.5 * fOffset)
    rx = julia_range(-0.5 * fWidth + 0.5 * fOffset, fOffset, +0.5 * fWidth - 0.5 * fOffset)
    fz, fy, fx = np.meshgrid(rz, ry, rx, indexing="ij")
    return np.stack([fCenter[0] + fx.ravel(), fCenter[1] + fy.ravel(), fCenter[2] + fz.ravel()], axis=1)


def assign_velocity(solid, v0):
    """Solid.jl:532-537"""
    solid.velocity[:] = np.asarray(v0, dtype=np.float64).reshape(1, -1)


def move_cpdi(solid, dx):
    """Solid.jl:539-548: moves the corner nodes and the particles."""
    d = np.asarray(dx, dtype=np.float64).reshape(1, -1)
    solid.nodes += d
    solid.pos += d


def move(solid, dx):
    """Solid.jl:550-555"""
    solid.pos += np.asarray(dx, dtype=np.float64).reshape(1, -1)


def rotate(solid, alpha):
    """Solid.jl:557-568"""
    beta = math.radians(alpha)
    xp, yp = solid.pos[:, 0].copy(), solid.pos[:, 1].copy()
    solid.pos[:, 0] = xp * math.cos(beta) - yp * math.sin(beta)
    solid.pos[:, 1] = yp * math.cos(beta) + xp * math.sin(beta)


def make_rectangular_pattern(coords, nx, ny, dx=0.0, dy=0.0):
    """Solid.jl:574-591"""
    coords = np.asarray(coords, dtype=np.float64).reshape(-1, 2)
    out = []
    for j in range(ny):
        for i in range(nx):
            out.append(np.stack([coords[:, 0] + i * dx, coords[:, 1] + j * dy], axis=1))
    return np.concatenate(out, axis=0)
