"""Output objects -- the part of Output.jl the solvers read on this path:
the dump `interval` (SolveExplicitSolidMechanics2D.jl:574).  Writing LAMMPS /
VTK dumps (Output.jl:146-335) is not part of the GPU step."""


class OutputType:
    def __init__(self, interval, *args, **kwargs):
        self.interval = int(interval)


class OvitoOutput(OutputType):
    pass


class PyPlotOutput(OutputType):
    pass


class VTKOutput(OutputType):
    pass
