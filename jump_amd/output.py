"""Output objects of Output.jl.  The solvers read the dump `interval`
(SolveExplicitSolidMechanics2D.jl:574); OvitoOutput also names the
directory prefix and the extra columns of the LAMMPS particle dumps the
library writes at each output point (plotParticles_2D/3D, Output.jl:146-288:
Engine.dump_lammps).  PyPlot figures and VTK files are not produced."""
import glob
import os


class OutputType:
    def __init__(self, interval, *args, **kwargs):
        self.interval = int(interval)


class OvitoOutput(OutputType):
    """OvitoOutput(interval, dir, outs) (Output.jl:31-45): `dir` is a prefix
    (the reference concatenates it with the file name, so it normally ends
    with a slash); existing *.LAMMPS files in it are removed, or it is
    created."""

    def __init__(self, interval, dir="", outs=()):
        super().__init__(interval)
        self.dir = str(dir)
        self.outs = list(outs)
        if self.dir:
            if os.path.isdir(self.dir):
                for f in glob.glob(self.dir + "*.LAMMPS"):
                    os.remove(f)
            else:
                os.mkdir(self.dir)

    def dump_path(self, counter):
        return f"{self.dir}dump_p.{int(counter)}.LAMMPS"


class PyPlotOutput(OutputType):
    pass


class VTKOutput(OutputType):
    pass
