"""Algorithm tags -- mirror of Algorithm.jl:16-39."""
from . import _ffi


class AlgorithmType:
    code = None


class USL(AlgorithmType):
    """Update-stress-last; nodes with mass <= tolerance are skipped in G2P."""
    code = _ffi.ALG_USL

    def __init__(self, tolerance):
        self.tolerance = float(tolerance)


class MUSL(AlgorithmType):
    """Modified USL with FLIP/PIC blend alpha."""
    code = _ffi.ALG_MUSL

    def __init__(self, alpha):
        self.alpha = float(alpha)
