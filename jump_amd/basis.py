"""Basis tags -- mirror of Basis.jl:37-48, plus uGIMP (absent from the
reference).  The shape functions themselves run on the GPU
(csrc/mpm_device.hpp); these objects select them."""
from . import _ffi


class BasisType:
    code = None


class LinearBasis(BasisType):
    code = _ffi.BASIS_LINEAR


class QuadBsplineBasis(BasisType):
    code = _ffi.BASIS_QUAD_BSPLINE


class CubicBsplineBasis(BasisType):
    code = _ffi.BASIS_CUBIC_BSPLINE


class CPDIQ4Basis(BasisType):
    code = _ffi.BASIS_CPDI_Q4


class GIMPBasis(BasisType):
    """Uniform GIMP (Bardenhagen & Kober 2004).  Each solid supplies its
    particle half-width `lp` (default: half the particle spacing implied by
    its volume)."""
    code = _ffi.BASIS_GIMP
