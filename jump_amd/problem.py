"""solve_explicit_dynamics_2D / _3D -- drop-in for module Problem's solvers
(Problem.jl:163-166):

  solve_explicit_dynamics_2D(grid, solids, basis, alg::USL, output, fixes, data)
      SolveExplicitSolidMechanics2D.jl:335-585
  solve_explicit_dynamics_2D(grid, solids, basis, alg::MUSL, output, fixes, data)
      SolveExplicitSolidMechanics2D.jl:25-329 (blended consistent mass)
  solve_explicit_dynamics_3D(grid, solids, basis, alg::MUSL, output, fixes, data)
      SolveExplicitSolidMechanics3D.jl:14-223

Same arguments, same `data` keys ("total_time", "dt", "time", "bodyforce",
"ghostcell", "dirichlet_grid"), same loop semantics: `while t < Tf`, output
and fixes after the step when counter % interval == 0, then t += dt.  The
steps between two output points run on the GPU in one call; particle and grid
arrays are copied back into the host objects when the solve returns.
"""
import sys

import numpy as np

from . import _ffi
from .algorithm import MUSL, USL
from .basis import GIMPBasis
from .bodyforce import VortexBodyForce2D
from .dirichlet import fix_Dirichlet_grid
from .engine import Engine, make_params
from .fix import EnergiesFix, closeFile, compute
from .output import OvitoOutput


def _fix_list(fixes):
    if fixes is None:
        return []
    return list(fixes) if isinstance(fixes, (list, tuple)) else [fixes]


def build_engine(grid, solids, basis, alg, data, device=0, engine=None):
    """The configured context of a solve: parameters from the grid, basis,
    algorithm and data; solids added in order; the Dirichlet mask of
    fix_Dirichlet_grid.  `engine(params)` makes the context (default: the
    HIP library's Engine on `device`)."""
    dim = grid.dim
    dt = float(data["dt"])
    bf = None
    if dim == 2:
        bf = data["bodyforce"]  # required by the 2D USL solver (:340)
        body = tuple(np.asarray(bf.g, dtype=np.float64).tolist()) + (0.0,) if hasattr(bf, "g") else (0.0, 0.0, 0.0)
    else:
        body = (0.0, 0.0, 0.0)  # the 3D MUSL step applies no body force
    ghost = "ghostcell" in data
    params = make_params(dim, basis.code, alg.code, dt, grid.lo, grid.hi, grid.nodes,
                         tolerance=getattr(alg, "tolerance", 0.0), alpha=getattr(alg, "alpha", 1.0), body=body,
                         ghostcell=ghost)
    eng = Engine(params, device=device) if engine is None else engine(params)
    if isinstance(bf, VortexBodyForce2D):  # evaluated per particle and step on the device
        eng.set_body_force(_ffi.BODY_VORTEX2D, bf.params())
    for s in solids:
        lp = s.lp
        if isinstance(basis, GIMPBasis) and lp is None and not s.rigid:
            # uGIMP default: half the particle spacing implied by its volume
            lp = 0.5 * float(s.volumeInitial[0]) ** (1.0 / dim)
        eng.add_solid(s.mat.to_c(), s.pos, s.velocity, s.mass, s.volume, s.volumeInitial, s._F, s._strain,
                      s._stress, corners=s.nodes, elems=s.elems, lp=lp)
    fix_Dirichlet_grid(grid, data, ghostcell=ghost)
    eng.set_fixed(np.ascontiguousarray(grid.fixedNodes.T))
    return eng


def _pull(eng, grid, solids):
    for i, s in enumerate(solids):
        r = eng.get_solid(i)
        s.pos[:] = r["pos"]
        s.velocity[:] = r["vel"]
        s.mass[:] = r["mass"]
        s.volume[:] = r["volume"]
        s._F[:] = r["F"]
        s._strain[:] = r["strain"]
        s._stress[:] = r["stress"]
        st = eng.get_state(i)
        s.alpha[:] = st["alpha"]
        s.alphadot[:] = st["alphadot"]
        s.vmStr[:] = st["vm"]
        s.pstrain[:] = st["pstrain"]
        m = s.mat
        for attr, src in (("alpha", st["alpha"]), ("alphadot", st["alphadot"]), ("vmStr", st["vm"])):
            arr = getattr(m, attr, None)
            if isinstance(arr, np.ndarray) and arr.shape[0] >= s.parCount:
                arr[:s.parCount] = src
    g = eng.get_grid()
    grid.mass[:] = g["mass"]
    grid.momentum0[:] = g["momentum0"]
    grid.momentum[:] = g["momentum"]
    grid.momentum2[:] = g["momentum2"]
    grid.force[:] = g["force"]


def _time_loop(eng, output, fixes, data):
    Tf = float(data["total_time"])
    dt = float(data["dt"])
    t = float(data["time"])
    interval = max(1, int(output.interval))
    fl = _fix_list(fixes)
    for f in fl:
        if isinstance(f, EnergiesFix):
            f._source = eng.energies
    counter = 0
    while t < Tf:
        k, tc, out = 0, t, False
        while True:
            k += 1
            if (counter + k - 1) % interval == 0:
                out = True
                break
            if not (tc + dt < Tf):
                break
            tc = tc + dt
        t_new, rc = eng.step(k, t)
        if rc == _ffi.JUMP_ERR_NUMERIC:
            flags = eng.flags()
            sys.stderr.write(f"jump: particle error flags 0x{flags:x} near t={tc}\n")
        if out:
            if isinstance(output, OvitoOutput):  # plotParticles_2D/3D (:575-576)
                eng.dump_lammps(output.dump_path(counter + k - 1), output.outs)
            compute(fl, tc)
        t = t_new
        counter += k
    eng.dump_wait()
    data["time_end"] = t
    return counter


def solve_explicit_dynamics_2D(grid, solids, basis, alg, output, fixes, data):
    if not isinstance(alg, (USL, MUSL)):
        raise TypeError("alg must be USL or MUSL")
    eng = build_engine(grid, solids, basis, alg, data)
    try:
        _time_loop(eng, output, fixes, data)
        _pull(eng, grid, solids)
    finally:
        closeFile(_fix_list(fixes))
        eng.close()


def solve_explicit_dynamics_3D(grid, solids, basis, alg, output, fixes, data):
    if not isinstance(alg, MUSL):
        raise NotImplementedError("3D USL is unfinished in the reference (SolveExplicitSolidMechanics3D.jl:226)")
    eng = build_engine(grid, solids, basis, alg, data)
    try:
        _time_loop(eng, output, fixes, data)
        _pull(eng, grid, solids)
    finally:
        closeFile(_fix_list(fixes))
        eng.close()
