"""Synthetic workloads of the benchmark configurations, built with the
reference's own particle builders (Solid.jl:322-509) and driver set-ups.

A Scene holds ABI parameters and per-solid arrays; `load(engine)` pushes them
into any Engine (the HIP library, or the oracle in tests).

  two_disks      Main_TwoDisks.jl:40-116 at 20x20 cells, linear basis
  disks_contact  two disks colliding head-on or obliquely, with or without
                 multi-body contact; spheres_contact: the same in 3D (MUSL)
  impact3d       3D high-velocity impact: elastic sphere on a J2 plate,
                 quadratic B-splines, MUSL (Main_HighVelocity_Impact_3D.jl)
  taylor_bar     3D Taylor bar (Main_TaylorBar.jl), J2 copper, GIMP
  rings2d        2D rings between rigid platens (Main_TwoRings.jl /
                 Main_LateralCompression_Cellular.jl), any 2D basis
  cpdi_tubes     2D lateral compression of a tube array with CPDI-Q4 domains
  two_rings      2D two-rings impact (Main_TwoRings.jl:40-130) at a given
                 resolution; two_rings_instance: one seeded instance of the
                 batched workload (drawn grid, size, material, speed, basis)
"""
import math

import numpy as np

from . import _ffi
from .engine import make_material, make_params
from .solid import (buildParticleForBlock, buildParticleForCircle, buildParticleForCylinder, buildParticleForRing,
                    buildParticleForRectangle, buildParticleForSphere)


class Scene:
    def __init__(self, name, params, solids, fixed=None, steps=1, meta=None):
        self.name = name
        self.params = params
        self.solids = solids
        self.fixed = fixed
        self.steps = steps
        self.meta = meta or {}

    @property
    def n_particles(self):
        return int(sum(s["pos"].shape[0] for s in self.solids))

    def load(self, eng):
        for s in self.solids:
            eng.add_solid(s["material"], s["pos"], s["vel"], s["mass"], s["volume"], s["volume0"],
                          s.get("F"), s.get("strain"), s.get("stress"), corners=s.get("corners"),
                          elems=s.get("elems"), lp=s.get("lp"))
        if self.fixed is not None:
            eng.set_fixed(self.fixed)
        return eng


def _solid(material, pos, vel, vol, density, lp=None, **kw):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    v = np.zeros((n, d))
    v[:] = np.asarray(vel, dtype=np.float64).reshape(1, d)
    vol = np.broadcast_to(np.asarray(vol, dtype=np.float64), (n,)).copy()
    out = dict(material=material, pos=pos, vel=v, mass=density * vol, volume=vol.copy(), volume0=vol.copy(), lp=lp)
    out.update(kw)
    return out


def _fixed_faces(nodes, faces, dirs=None):
    """Dirichlet mask (nn, d) for whole faces, Dirichlet.jl:28-82."""
    d = len(nodes)
    nn = int(np.prod(nodes))
    m = np.zeros((nn, d), dtype=np.uint8)
    idx = np.arange(nn)
    i = idx % nodes[0]
    j = (idx // nodes[0]) % nodes[1]
    k = idx // (nodes[0] * nodes[1]) if d == 3 else None
    sel = {"left": i == 0, "right": i == nodes[0] - 1, "bottom": j == 0, "top": j == nodes[1] - 1}
    if d == 3:
        sel["front"] = k == 0
        sel["back"] = k == nodes[2] - 1
    for f in faces:
        fx = dirs[f] if dirs else (1,) * d
        for a in range(d):
            if fx[a]:
                m[sel[f], a] = 1
    return m


def elastic(E, nu, rho):
    return make_material(_ffi.MAT_ELASTIC, E=E, nu=nu, density=rho)


def neohookean(E, nu, rho):
    return make_material(_ffi.MAT_NEOHOOKEAN, E=E, nu=nu, density=rho)


def j2(E, nu, rho, fy, k1):
    return make_material(_ffi.MAT_ELASTOPLASTIC, E=E, nu=nu, density=rho, fy=fy, k1=k1)


def johnson_cook(E, nu, rho, A, B, C, n, epsdot0, cellsize):
    return make_material(_ffi.MAT_JOHNSONCOOK, E=E, nu=nu, density=rho, A=A, B=B, C=C, n=n, epsdot0=epsdot0,
                         m=1.0, chi=0.9, Cp=385.0, cellsize=cellsize)


def rigid(vx, vy, rho, fixed=False):
    return make_material(_ffi.MAT_RIGID, density=rho, vx=vx, vy=vy, fixed=fixed)


# ---------------------------------------------------------------- 2D
def two_disks(cells=20, ppc=2, steps=3500, basis=_ffi.BASIS_LINEAR, material="elastic"):
    """Main_TwoDisks.jl:40-116: unit square, disks r=0.2 at (0.2,0.2) and
    (0.8,0.8) moving at +-(0.1,0.1), E=1000, nu=0.3, rho=1000, dt=1e-3."""
    n = cells + 1
    dx = 1.0 / cells
    fOffset = dx / ppc
    c1 = buildParticleForCircle([0.2, 0.2], 0.2, fOffset)
    c2 = buildParticleForCircle([0.8, 0.8], 0.2, fOffset)
    mat = {"elastic": elastic, "neohookean": neohookean}[material](1000.0, 0.3, 1000.0) \
        if material != "j2" else j2(1000.0, 0.3, 1000.0, 5.0, 100.0)
    vol = fOffset * fOffset
    lp = 0.5 * fOffset if basis == _ffi.BASIS_GIMP else None
    s1 = _solid(mat, c1, (0.1, 0.1), vol, 1000.0, lp=lp)
    s2 = _solid(mat, c2, (-0.1, -0.1), vol, 1000.0, lp=lp)
    p = make_params(2, basis, _ffi.ALG_USL, 1e-3, (0.0, 0.0), (1.0, 1.0), (n, n), tolerance=1e-9)
    return Scene("two_disks", p, [s1, s2], steps=steps, meta=dict(interval=10, total_time=3.5))


def disks_contact(cells=32, ppc=2, r=0.15, gap_cells=2, v=0.1, offset=0.0, basis=_ffi.BASIS_LINEAR,
                  material="elastic", contact=True, friction=0.0, steps=1000, dt=1e-3):
    """Two Two-Disks disks (Main_TwoDisks.jl:40-116 material: E=1000,
    nu=0.3, rho=1000) on a horizontal line of the unit square, `gap_cells`
    cells apart, moving towards each other at +-v; `offset` shifts the left
    disk down and the right disk up (an oblique impact).  Centres sit on grid
    nodes so that no particle lies on a grid line.  With contact on, the
    disks interact only through the contact correction
    (SolveExplicitFEMP2DContact.jl:311-387)."""
    n = cells + 1
    dx = 1.0 / cells
    fOffset = dx / ppc
    rr = math.floor(r / fOffset) * fOffset
    half = round((rr + 0.5 * gap_cells * dx) / dx) * dx
    dyc = round(0.5 * offset / dx) * dx
    c1 = buildParticleForCircle([0.5 - half, 0.5 - dyc], r, fOffset)
    c2 = buildParticleForCircle([0.5 + half, 0.5 + dyc], r, fOffset)
    mat = {"elastic": elastic, "neohookean": neohookean}[material](1000.0, 0.3, 1000.0) \
        if material != "j2" else j2(1000.0, 0.3, 1000.0, 5.0, 100.0)
    vol = fOffset * fOffset
    lp = 0.5 * fOffset if basis == _ffi.BASIS_GIMP else None
    s1 = _solid(mat, c1, (v, 0.0), vol, 1000.0, lp=lp)
    s2 = _solid(mat, c2, (-v, 0.0), vol, 1000.0, lp=lp)
    p = make_params(2, basis, _ffi.ALG_USL, dt, (0.0, 0.0), (1.0, 1.0), (n, n), tolerance=1e-9, contact=contact,
                    friction=friction)
    return Scene("disks_contact", p, [s1, s2], steps=steps)


def disk_wall(cells=32, ppc=2, r=0.15, gap_cells=2, v=(0.0, -0.1), wall_v=(0.0, 0.0), fixed=False,
              basis=_ffi.BASIS_LINEAR, material="elastic", friction=0.0, steps=600, dt=1e-3, wall=True):
    """A Two-Disks disk (E=1000, nu=0.3, rho=1000) moving at v onto a rigid
    wall (RigidMaterial(vx, vy, rho), Main_LateralCompression_Cellular.jl:
    58-59) two particles thick below it, `gap_cells` cells apart, with
    multi-body contact (SolveExplicitFEMP2DContact.jl:272-358: the rigid
    body's normal, the rigid velocity as the centre-of-mass velocity, only
    the disk corrected).  The disk centre sits on a grid node, the wall's
    particles at quarter cells, so no particle lies on a grid line."""
    n = cells + 1
    dx = 1.0 / cells
    fOffset = dx / ppc
    rr = math.floor(r / fOffset) * fOffset
    cy = round((0.5 + 0.5 * gap_cells * dx) / dx) * dx
    c = buildParticleForCircle([0.5, cy], r, fOffset)
    mat = {"elastic": elastic, "neohookean": neohookean}[material](1000.0, 0.3, 1000.0) \
        if material != "j2" else j2(1000.0, 0.3, 1000.0, 5.0, 100.0)
    vol = fOffset * fOffset
    lp = 0.5 * fOffset if basis == _ffi.BASIS_GIMP else None
    solids = [_solid(mat, c, v, vol, 1000.0, lp=lp)]
    if wall:
        top = cy - rr - gap_cells * dx  # the wall's upper surface
        w = buildParticleForRectangle([0.5, top - fOffset], 24 * dx, 2 * fOffset, fOffset)
        solids.append(_solid(rigid(wall_v[0], wall_v[1], 7800.0, fixed=fixed), w, (0.0, 0.0), vol, 7800.0))
    p = make_params(2, basis, _ffi.ALG_USL, dt, (0.0, 0.0), (1.0, 1.0), (n, n), tolerance=1e-9, contact=True,
                    friction=friction)
    return Scene("disk_wall", p, solids, steps=steps)


def spheres_contact(cells=24, r=0.15, gap_cells=2, v=0.1, offset=0.0, basis=_ffi.BASIS_QUAD_BSPLINE,
                    material="elastic", contact=True, steps=600, dt=1e-3, alpha=1.0):
    """Two spheres (2 particles per cell and direction, Two-Disks material
    E=1000, nu=0.3, rho=1000) on the x axis of the unit cube, `gap_cells`
    cells apart, moving towards each other at +-v, 3D MUSL; `offset` shifts
    the left sphere down in y and the right one up (an oblique impact).
    Centres sit on grid nodes.  With contact on, the spheres interact only
    through the normal contact correction
    (SolveExplicitFEMP3DContact.jl:398-511)."""
    n = cells + 1
    dx = 1.0 / cells
    f = dx / 2
    rr = math.floor(r / f) * f
    half = round((rr + 0.5 * gap_cells * dx) / dx) * dx
    dyc = round(0.5 * offset / dx) * dx
    c1 = buildParticleForSphere([0.5 - half, 0.5 - dyc, 0.5], r, f)
    c2 = buildParticleForSphere([0.5 + half, 0.5 + dyc, 0.5], r, f)
    mat = {"elastic": elastic, "neohookean": neohookean}[material](1000.0, 0.3, 1000.0) \
        if material != "j2" else j2(1000.0, 0.3, 1000.0, 5.0, 100.0)
    vol = f ** 3
    lp = 0.5 * f if basis == _ffi.BASIS_GIMP else None
    s1 = _solid(mat, c1, (v, 0.0, 0.0), vol, 1000.0, lp=lp)
    s2 = _solid(mat, c2, (-v, 0.0, 0.0), vol, 1000.0, lp=lp)
    p = make_params(3, basis, _ffi.ALG_MUSL, dt, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (n, n, n), alpha=alpha,
                    contact=contact)
    return Scene("spheres_contact", p, [s1, s2], steps=steps)


def rings2d(cells=40, basis=_ffi.BASIS_LINEAR, material="neohookean", platen=True, steps=50, v=2.0, seed=0):
    """Two rings pressed by a moving rigid platen (Main_TwoRings.jl,
    Main_LateralCompression_Cellular.jl:58-59), unit-free small geometry."""
    L = 1.0
    dx = L / cells
    fOffset = dx / 2
    rho, E, nu = 1000.0, 1.0e5, 0.3
    mk = {"elastic": lambda: elastic(E, nu, rho), "neohookean": lambda: neohookean(E, nu, rho),
          "j2": lambda: j2(E, nu, rho, 500.0, 1.0e3)}[material]
    lp = 0.5 * fOffset if basis == _ffi.BASIS_GIMP else None
    solids = []
    for cx, cy, vx in ((0.3, 0.4, 0.0), (0.6, 0.4, 0.0)):
        pts = buildParticleForRing([cx, cy], 0.08, 0.13, fOffset)
        solids.append(_solid(mk(), pts, (vx, 0.0), fOffset * fOffset, rho, lp=lp))
    if platen:
        top = buildParticleForRectangle([0.45, 0.6], 0.5, 2 * fOffset, fOffset)
        solids.append(_solid(rigid(0.0, -v, 7800.0), top, (0.0, 0.0), fOffset * fOffset, 7800.0))
    c = math.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / rho)
    dt = 0.2 * dx / c
    p = make_params(2, basis, _ffi.ALG_USL, dt, (0.0, 0.0), (L, L), (cells + 1, cells + 1), tolerance=1e-12,
                    body=(0.0, -9.81))
    fixed = _fixed_faces((cells + 1, cells + 1), ["bottom"], {"bottom": (0, 1)})
    return Scene("rings2d", p, solids, fixed=fixed, steps=steps)


def two_rings(side, k, basis, E, vel, steps=1000, nu=0.3, meta=None):
    """The two-rings impact of Main_TwoRings.jl:40-130 (domain 200 x 100 mm,
    rings r 30..40 mm at x = 50 and 150, Neo-Hookean, rho 1010e-9,
    USL(1e-18)) on `side` x (side+1)/2 nodes (side odd, so dy = dx) with k
    particles per cell and direction.  The ring centres sit on the nearest
    grid nodes: the builder offsets particles by half-integer multiples of
    dx/k from the centre, so none lies on a grid line, where the hat
    gradient's sign(0) = 0 (Basis.jl:219-251) leaves sum(grad N) != 0 and
    strains a translating body.  dt = 0.3 dx / c where the reference takes
    0.6, at which hat-function runs go unstable at the free surface within
    50 steps."""
    l, w, rin, rout, rho = 200.0, 100.0, 30.0, 40.0, 1010e-9
    nx, ny = side, (side + 1) // 2
    dx = l / (nx - 1)
    dy = w / (ny - 1)
    fOffset = dx / k
    lp = 0.5 * fOffset if basis == _ffi.BASIS_GIMP else None
    cy = round(0.5 * w / dy) * dy
    solids = []
    for cx, vx in ((round(0.25 * l / dx) * dx, vel), (round(0.75 * l / dx) * dx, -vel)):
        pts = buildParticleForRing([cx, cy], rin, rout, fOffset)
        solids.append(_solid(neohookean(E, nu, rho), pts, (vx, 0.0), fOffset * fOffset, rho, lp=lp))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    c = math.sqrt((lam + 2 * mu) / rho)
    dt = 0.3 * dx / c
    p = make_params(2, basis, _ffi.ALG_USL, dt, (0.0, 0.0), (l, w), (nx, ny), tolerance=1e-18)
    return Scene("two_rings", p, solids, steps=steps, meta=meta or {"side": side, "E": E, "vel": vel, "ppc": k})


def two_rings_instance(seed, steps=1000, draw="mixed"):
    """One instance of the batched workload: two_rings() with drawn sizes.
    rng = default_rng(seed):
      grid side S ~ log-uniform [64, 512] nodes along x (made odd);
      particle count ~ log-uniform [5k, 80k], through k particles per cell
        and direction, clamped to 1..4 (denser packings are unstable at this
        dt, so small grids get fewer particles than drawn);
      E ~ U[0.5, 5] MPa (nu 0.3); impact speed ~ U[10, 100] (mm/s, the
      reference's units); basis linear or GIMP with probability 1/2.
    draw="stable" keeps every drawn number but uses GIMP for every instance
    and a tenth of the drawn speed (1..10 mm/s): over 1000 steps the mixed
    draw loses 92 of its 256 instances (seeds 1000..1255) to NaN, the linear
    hat ones to the small-mass cell-crossing instability of the reference's
    USL/FLIP step (a node a crossing particle barely reaches gets a tiny mass
    but a full hat gradient) and fast GIMP ones to collisions that invert
    particles in the single shared velocity field; the restatement on the CPU
    fails on the same instances.  The stable draw loses none."""
    rng = np.random.default_rng(seed)
    side = int(round(math.exp(rng.uniform(math.log(64), math.log(512))))) | 1
    np_target = math.exp(rng.uniform(math.log(5000.0), math.log(80000.0)))
    E = rng.uniform(0.5, 5.0)
    vel = rng.uniform(10.0, 100.0)
    basis = _ffi.BASIS_GIMP if rng.uniform() < 0.5 else _ffi.BASIS_LINEAR
    if draw == "stable":
        basis = _ffi.BASIS_GIMP
        vel = 0.1 * vel
    elif draw != "mixed":
        raise ValueError(f"unknown draw {draw!r}")
    dx = 200.0 / (side - 1)
    area = 2.0 * math.pi * (40.0 * 40.0 - 30.0 * 30.0)
    k = min(4, max(1, int(round(math.sqrt(np_target / area) * dx))))
    return two_rings(side, k, basis, E, vel, steps=steps,
                     meta={"seed": seed, "side": side, "E": E, "vel": vel, "ppc": k, "draw": draw})


def _ring_q4(cx, cy, ri, ro, nr, nt, jitter, rng):
    """Polar Q4 mesh of a ring: (nt*(nr+1), 2) corners, (nt*nr, 4) elements
    (counter-clockwise)."""
    r = np.linspace(ri, ro, nr + 1)
    t = np.linspace(0.0, 2 * math.pi, nt, endpoint=False)
    tt, rr = np.meshgrid(t, r, indexing="ij")
    x = cx + rr * np.cos(tt)
    y = cy + rr * np.sin(tt)
    if jitter > 0.0:
        x = x + rng.uniform(-jitter, jitter, x.shape)
        y = y + rng.uniform(-jitter, jitter, y.shape)
    corners = np.stack([x.ravel(), y.ravel()], axis=1)
    elems = []
    for a in range(nt):
        b = (a + 1) % nt
        for q in range(nr):
            n00 = a * (nr + 1) + q
            n01 = a * (nr + 1) + q + 1
            n10 = b * (nr + 1) + q
            n11 = b * (nr + 1) + q + 1
            elems.append((n00, n01, n11, n10))
    return corners, np.asarray(elems, dtype=np.int64)


def cpdi_tubes(rings=(2, 2), cells=64, nr=4, nt=96, jitter_cells=0.1, steps=20, seed=20261202, platens=True,
               contact=False, friction=0.0):
    """Lateral compression of a tube array with CPDI-Q4 domains
    (Main_LateralCompression_Cellular.jl:26-86, Main_LateralCompressionTube_CPDI.jl):
    rings of outer radius 5 mm and wall 0.5 mm side by side, each its own
    solid, Neo-Hookean E=1 GPa, nu=0.3, rho=1000, between rigid platens
    closing at 1 m/s each; ring corners jittered by +-0.1 cell (seeded); nr x
    nt Q4 domains per ring.  With contact on, rings and platens interact
    through the multi-body contact correction (rigid platens by their own
    normal); otherwise every solid shares one velocity field.  The platens'
    particles sit at quarter cells, off the grid lines."""
    rng = np.random.default_rng(seed)
    ro, wall = 5e-3, 0.5e-3
    nxr, nyr = rings
    width = 2 * ro * nxr
    height = 2 * ro * nyr
    margin = 4 * ro
    L = max(width, height) + 2 * margin
    dx = L / cells
    rho, E, nu = 1000.0, 1.0e9, 0.3
    solids = []
    jit = jitter_cells * dx
    for j in range(nyr):
        for i in range(nxr):
            cx = margin + ro + 2 * ro * i
            cy = margin + ro + 2 * ro * j
            corners, elems = _ring_q4(cx, cy, ro - wall, ro, nr, nt, jit, rng)
            cc = corners[elems]
            x, y = cc[:, :, 0], cc[:, :, 1]
            area = 0.5 * (x[:, 0] * y[:, 1] - x[:, 1] * y[:, 0] + x[:, 1] * y[:, 2] - x[:, 2] * y[:, 1]
                          + x[:, 2] * y[:, 3] - x[:, 3] * y[:, 2] + x[:, 3] * y[:, 0] - x[:, 0] * y[:, 3])
            centre = cc.mean(axis=1)
            solids.append(dict(material=neohookean(E, nu, rho), pos=np.ascontiguousarray(centre),
                               vel=np.zeros_like(centre), mass=rho * area, volume=area.copy(), volume0=area.copy(),
                               corners=corners, elems=elems))
    if platens:
        f = dx / 2
        # two rows of particles at quarter cells just outside the top and
        # bottom ring rows, 4 cells wider than the array on each side
        i0 = int(math.floor((margin - 4 * dx) / dx))
        i1 = int(math.ceil((margin + width + 4 * dx) / dx))
        xs = np.concatenate([(np.arange(i0, i1) + q) * dx for q in (0.25, 0.75)])
        xs.sort()
        jt = int(math.ceil((margin + height) / dx))
        jb = int(math.floor(margin / dx)) - 1
        for rows, vy in (((jt + 0.25, jt + 0.75), -1.0), ((jb + 0.25, jb + 0.75), 1.0)):
            X, Y = np.meshgrid(xs, np.asarray(rows) * dx, indexing="xy")
            pts = np.stack([X.ravel(), Y.ravel()], axis=1)
            # rigid platens use the hat-cell stencil only; give them unit
            # CPDI domains so the solid is well formed
            nc = pts.shape[0]
            pc = np.concatenate([pts + [-f / 2, -f / 2], pts + [f / 2, -f / 2], pts + [f / 2, f / 2],
                                 pts + [-f / 2, f / 2]])
            pe = np.stack([np.arange(nc), np.arange(nc) + nc, np.arange(nc) + 2 * nc, np.arange(nc) + 3 * nc], axis=1)
            solids.append(dict(material=rigid(0.0, vy, 7800.0), pos=pts, vel=np.zeros_like(pts),
                               mass=np.full(nc, 7800.0 * f * f), volume=np.full(nc, f * f), volume0=np.full(nc, f * f),
                               corners=pc, elems=pe))
    c = math.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / rho)
    dt = 0.2 * dx / c
    p = make_params(2, _ffi.BASIS_CPDI_Q4, _ffi.ALG_USL, dt, (0.0, 0.0), (L, L), (cells + 1, cells + 1),
                    tolerance=1e-14, contact=contact, friction=friction)
    return Scene("cpdi_tubes", p, solids, steps=steps, meta={"rings": rings, "dx": dx})


def musl_block2d(cells=256, basis=_ffi.BASIS_QUAD_BSPLINE, steps=20, seed=0, alpha=0.99):
    """2D MUSL with the blended mass matrix at size: a stress-free elastic
    block (0.6 x 0.6 of the unit square, 2x2 particles per cell, E=1 MPa,
    nu=0.3, rho=1000) translating at (0.3, -0.2) m/s with seeded noise (0.05)
    under gravity; dt = 0.02 cell crossings at the dilatational wave speed, a
    step this scheme is stable at (it diverges at 0.05 on a stressed block)."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / cells
    f = dx / 2
    xs = np.arange(0.2 + f / 2, 0.8, f)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    rho, E, nu = 1000.0, 1.0e6, 0.3
    s = _solid(elastic(E, nu, rho), pts, (0.3, -0.2), f * f, rho)
    s["vel"] = s["vel"] + rng.normal(scale=0.05, size=s["vel"].shape)
    c = math.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / rho)
    p = make_params(2, basis, _ffi.ALG_MUSL, 0.02 * dx / c, (0.0, 0.0), (1.0, 1.0), (cells + 1, cells + 1),
                    alpha=alpha, body=(0.0, -9.81))
    return Scene("musl_block2d", p, [s], steps=steps)


# ---------------------------------------------------------------- 3D
def impact3d(nodes=(256, 256, 64), plate_mm=(100.0, 100.0, 20.0), sphere_r_mm=5.0, steps=2000, cfl=0.2,
             basis=_ffi.BASIS_QUAD_BSPLINE, plate_material="j2", sphere_material="elastic", v0=300.0):
    """3D high-velocity impact: sphere (r=5 mm, 300 m/s, stiff elastic: 3D
    solids cannot be rigid) on a J2 plate 100x100x20 mm (E=200 GPa, nu=0.3,
    fy=300 MPa, linear hardening 1 GPa, rho=7800); quadratic B-splines, MUSL(1),
    2x2x2 particles per cell, dx=0.72 mm, all grid faces fixed."""
    dx = 0.72e-3
    nx, ny, nz = nodes
    f = dx / 2
    E, nu, rho = 200e9, 0.3, 7800.0
    W, H, T = (v * 1e-3 for v in plate_mm)
    # particle lattice at (k+1/4) and (k+3/4) of a cell
    cx = 0.5 * (nx - 1) * dx
    wfl = math.floor(W / f) * f
    x0 = cx - 0.5 * wfl + 0.5 * f
    shift = (math.floor(x0 / dx) + 0.25) * dx - x0 if (x0 / dx) % 0.5 != 0.25 else 0.0
    cxs = cx + shift
    hfl = math.floor(H / f) * f
    y0 = cx - 0.5 * hfl + 0.5 * f
    shy = (math.floor(y0 / dx) + 0.25) * dx - y0
    cys = cx + shy
    zbot = 2.25 * dx
    czp = zbot + 0.5 * T - 0.5 * f
    plate = buildParticleForBlock([cxs, cys, czp], W, H, T, f)
    zt = plate[:, 2].max()
    sr = math.floor(sphere_r_mm * 1e-3 / f) * f
    czs = zt + f + sr - 0.5 * f
    sph = buildParticleForSphere([cxs, cys, czs], sphere_r_mm * 1e-3, f)
    vol = f ** 3
    pm = {"j2": j2(E, nu, rho, 300e6, 1e9), "elastic": elastic(E, nu, rho),
          "neohookean": neohookean(E, nu, rho),
          "jc": johnson_cook(E, nu, rho, 300e6, 1e9, 0.0, 1.0, 1.0, dx)}[plate_material]
    sm = {"elastic": elastic(E, nu, rho), "neohookean": neohookean(E, nu, rho)}[sphere_material]
    lp = 0.5 * f if basis == _ffi.BASIS_GIMP else None
    solids = [_solid(sm, sph, (0.0, 0.0, -v0), vol, rho, lp=lp), _solid(pm, plate, (0.0, 0.0, 0.0), vol, rho, lp=lp)]
    lam = E * nu / (1 + nu) / (1 - 2 * nu)
    mu = E / 2 / (1 + nu)
    c = math.sqrt((lam + 2 * mu) / rho)
    dt = cfl * dx / c
    hi = ((nx - 1) * dx, (ny - 1) * dx, (nz - 1) * dx)
    p = make_params(3, basis, _ffi.ALG_MUSL, dt, (0.0, 0.0, 0.0), hi, nodes, alpha=1.0)
    fixed = _fixed_faces(nodes, ["left", "right", "bottom", "top", "front", "back"])
    return Scene("impact3d", p, solids, fixed=fixed, steps=steps)


def vortex_annulus(cells=40, k=2, basis=_ffi.BASIS_LINEAR, G=0.5, T=1.0, t_end=0.1, cfl=0.1, E=1.0e6):
    """The vortex manufactured solution (Main_Vortex_FEMP.jl:41-109 set-up,
    BodyForce.jl:94-121): an annulus 0.75 < r < 1.25 of Neo-Hookean solid
    (E = 1 MPa, nu = 0.3, rho = 1050) centred in the grid [0, 2.8]^2 (the
    driver's B-spline grid: the reference's B-spline support assumes a grid
    starting at 0),
    here as particles (k per cell and direction, buildParticleForRing) on the
    2D USL step with VortexBodyForce2D(G, T, Ri, Ro, rho, mu, centre).  The
    exact solution rotates each point by alpha = G sin(pi t/T) h(r); the
    particles start with its velocity at t = 0, alpha'(0) (-y, x) with
    alpha'(0) = G pi/T h(r).  The driver steps its FEMP solver at 0.02
    dx/c; the MPM step here at cfl dx/c, to t_end (meta: steps, the vortex
    force)."""
    from .bodyforce import VortexBodyForce2D
    nu, rho = 0.3, 1050.0
    Ri, Ro, l = 0.75, 1.25, 1.4
    mu = E / (2 * (1 + nu))
    dx = 2 * l / cells
    f = dx / k
    pts = buildParticleForRing([l, l], Ri, Ro, f)
    vf = VortexBodyForce2D(G, T, Ri, Ro, rho, mu, l, l)
    rel = pts - l
    r = np.sqrt((rel ** 2).sum(axis=1))
    s_ = (r - vf.R) / (Ri - Ro)
    ad = G * math.pi / T * (1 - 8 * s_ ** 2 + 16 * s_ ** 4)
    lp = 0.5 * f if basis == _ffi.BASIS_GIMP else None
    sol = _solid(neohookean(E, nu, rho), pts, (0.0, 0.0), f * f, rho, lp=lp)
    sol["vel"] = np.ascontiguousarray(np.stack([-ad * rel[:, 1], ad * rel[:, 0]], axis=1))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    c = math.sqrt((lam + 2 * mu) / rho)
    dt = cfl * dx / c
    steps = int(math.ceil(t_end / dt))
    p = make_params(2, basis, _ffi.ALG_USL, dt, (0.0, 0.0), (2 * l, 2 * l), (cells + 1, cells + 1))
    return Scene("vortex_annulus", p, [sol], steps=steps, meta={"vortex": vf, "steps": steps, "dt": dt})


def taylor_bar(scale=1.0, steps=400, basis=_ffi.BASIS_GIMP, material="j2", E=117e9, fy=400e6, H=100e6, v0=190.0,
               alpha=0.95):
    """Taylor bar (Main_TaylorBar.jl:43-130): copper cylinder r=3.2 mm,
    L=32.4 mm along y, v0=190 m/s onto a rigid wall, J2, GIMP.  `scale`
    multiplies the particle spacing (1.0: ~0.128 mm, 2 particles per cell and
    axis).  The wall is the reference's Dirichlet bottom (nodes fixed in y),
    taken as the node planes j <= 2 so that particles landing on it keep
    their GIMP window inside the grid (a particle needs its half-width of
    room above the lowest node); the bar starts half a cell above it and
    lands within the first tens of steps.  MUSL(alpha) with alpha = 0.95:
    the reference's 3D step zeroes momentum0 and momentum on fixed nodes but
    not momentum2 (SolveExplicitSolidMechanics3D.jl:89-105, 142-156), and
    G2P-2 takes the velocity gradient from momentum2, so with pure FLIP
    (alpha = 1) a wall changes no particle velocity and the bar develops no
    strain against it; with the PIC share the wall's zero momentum damps
    the velocities next to it."""
    r, L = 3.2e-3, 32.4e-3
    f = 0.128e-3 * scale
    dx = 2 * f
    rho, nu = 8930.0, 0.35
    pts = buildParticleForCylinder([0.0, 0.0, 0.0], r, 0.5 * L, f, f)
    margin = 4 * dx
    jw = 2  # the wall: node planes 0..jw fixed in y
    ext = (2 * r + 2 * margin, L * 1.2 + 2 * margin, 2 * r + 2 * margin)
    nodes = tuple(int(math.ceil(e / dx)) + 1 for e in ext)
    shift_y = (jw + 0.5) * dx - pts[:, 1].min() + 0.5 * f
    pts = pts + np.array([0.5 * (nodes[0] - 1) * dx + 0.25 * dx, shift_y, 0.5 * (nodes[2] - 1) * dx + 0.25 * dx])
    mat = j2(E, nu, rho, fy, H) if material == "j2" else johnson_cook(E, nu, rho, fy, H, 0.0, 1.0, 1.0, dx)
    lp = 0.5 * f if basis == _ffi.BASIS_GIMP else None
    solids = [_solid(mat, pts, (0.0, -v0, 0.0), f ** 3, rho, lp=lp)]
    lam = E * nu / (1 + nu) / (1 - 2 * nu)
    mu = E / 2 / (1 + nu)
    c = math.sqrt((lam + 2 * mu) / rho)
    dt = 0.2 * dx / c
    hi = tuple((n - 1) * dx for n in nodes)
    p = make_params(3, basis, _ffi.ALG_MUSL, dt, (0.0, 0.0, 0.0), hi, nodes, alpha=alpha)
    fixed = np.zeros((int(np.prod(nodes)), 3), dtype=np.uint8)
    j = (np.arange(fixed.shape[0]) // nodes[0]) % nodes[1]
    fixed[j <= jw, 1] = 1
    return Scene("taylor_bar", p, solids, fixed=fixed, steps=steps)


def block3d(nodes=(12, 12, 12), basis=_ffi.BASIS_QUAD_BSPLINE, material="elastic", steps=10, v=(1.0, -2.0, 0.5),
            seed=0, ppc=2):
    """Small 3D block with a random velocity field and random initial stress,
    for parity tests of one material/basis pair; ppc particles per cell and
    direction."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / (nodes[0] - 1)
    f = dx / ppc
    ctr = [0.5 + 0.13 * dx, 0.5 + 0.07 * dx, 0.5 + 0.11 * dx]
    pts = buildParticleForBlock(ctr, 0.45, 0.45, 0.45, f)
    rho, E, nu = 1000.0, 1.0e6, 0.3
    mat = {"elastic": elastic(E, nu, rho), "neohookean": neohookean(E, nu, rho),
           "j2": j2(E, nu, rho, 2.0e3, 1.0e4), "jc": johnson_cook(E, nu, rho, 2.0e3, 1.0e4, 0.0, 1.0, 1.0, dx),
           # rate-dependent Johnson-Cook: C != 0 and a small reference rate, so
           # the (1 + rate ratio)^C factor and the alphadot filter take part
           "jc_rate": johnson_cook(E, nu, rho, 2.0e3, 1.0e4, 0.1, 0.5, 1.0e-5, dx)}[material]
    lp = 0.5 * f if basis == _ffi.BASIS_GIMP else None
    s = _solid(mat, pts, v, f ** 3, rho, lp=lp)
    s["vel"] = s["vel"] + rng.normal(scale=0.5, size=s["vel"].shape)
    n = pts.shape[0]
    sig = rng.normal(scale=1.0e3, size=(n, 3, 3))
    sig = 0.5 * (sig + sig.transpose(0, 2, 1))
    s["stress"] = np.ascontiguousarray(sig.transpose(0, 2, 1).reshape(n, 9))
    c = math.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / rho)
    dt = 0.2 * dx / c
    hi = tuple((k - 1) * dx for k in nodes)
    p = make_params(3, basis, _ffi.ALG_MUSL, dt, (0.0, 0.0, 0.0), hi, nodes, alpha=0.95)
    fixed = _fixed_faces(nodes, ["bottom"], {"bottom": (0, 1, 0)})
    return Scene("block3d", p, [s], fixed=fixed, steps=steps)


def block2d(cells=16, basis=_ffi.BASIS_LINEAR, material="elastic", steps=10, seed=0, algorithm=_ffi.ALG_USL,
            alpha=1.0, ppc=2):
    """Small 2D block with random velocity and stress (parity tests); USL, or
    MUSL with the blended mass matrix and FLIP fraction `alpha`; ppc
    particles per cell and direction."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / cells
    f = dx / ppc
    pts = buildParticleForRectangle([0.5 + 0.13 * dx, 0.5 + 0.07 * dx], 0.45, 0.45, f)
    rho, E, nu = 1000.0, 1.0e6, 0.3
    mat = {"elastic": elastic(E, nu, rho), "neohookean": neohookean(E, nu, rho),
           "j2": j2(E, nu, rho, 2.0e3, 1.0e4)}[material]
    lp = 0.5 * f if basis == _ffi.BASIS_GIMP else None
    s = _solid(mat, pts, (1.0, -0.5), f * f, rho, lp=lp)
    s["vel"] = s["vel"] + rng.normal(scale=0.5, size=s["vel"].shape)
    n = pts.shape[0]
    sig = rng.normal(scale=1.0e3, size=(n, 2, 2))
    sig = 0.5 * (sig + sig.transpose(0, 2, 1))
    s["stress"] = np.ascontiguousarray(sig.transpose(0, 2, 1).reshape(n, 4))
    c = math.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / rho)
    dt = 0.2 * dx / c
    p = make_params(2, basis, algorithm, dt, (0.0, 0.0), (1.0, 1.0), (cells + 1, cells + 1), tolerance=1e-12,
                    alpha=alpha, body=(0.0, -9.81))
    fixed = _fixed_faces((cells + 1, cells + 1), ["bottom"], {"bottom": (0, 1)})
    return Scene("block2d", p, [s], fixed=fixed, steps=steps)
