"""jump_amd -- MI355X-native explicit MPM step for jump's Grid/Solid/Material/
Basis API.  The step runs in libjumphip.so (hand-written HIP kernels for
gfx950) behind the C ABI in include/jump_mpm.h; this package is the host-side
mirror of the reference's Julia interface."""
from .algorithm import MUSL, USL
from .basis import CPDIQ4Basis, CubicBsplineBasis, GIMPBasis, LinearBasis, QuadBsplineBasis
from .bodyforce import ConstantBodyForce2D, ConstantBodyForce3D
from .dirichlet import fix_Dirichlet_grid
from .engine import Engine, JumpError, make_material, make_params
from .fix import EnergiesFix, closeFile, compute
from .grid import Grid2D, Grid3D, to_1D_index
from .material import (ElasticMaterial, ElastoPlasticMaterial, JohnsonCookMaterial, NeoHookeanMaterial,
                       RigidMaterial)
from .output import OvitoOutput, PyPlotOutput, VTKOutput
from .problem import solve_explicit_dynamics_2D, solve_explicit_dynamics_3D
from .solid import (Solid2D, Solid3D, assign_velocity, buildParticleForBlock, buildParticleForCircle,
                    buildParticleForCylinder, buildParticleForRectangle, buildParticleForRing, buildParticleForSegment,
                    buildParticleForSphere, julia_range, make_rectangular_pattern, move, rotate)

__all__ = [k for k in dir() if not k.startswith("_")]
