/*
 * jump_mpm.h -- C ABI of the MI355X explicit MPM step (libjumphip.so).
 *
 * This is the drop-in boundary for jump's explicit updated-Lagrangian MPM
 * solvers:
 *   solve_explicit_dynamics_2D(grid,solids,basis,alg::USL,output,fixes,data)
 *       SolveExplicitSolidMechanics2D.jl:335-585
 *   solve_explicit_dynamics_2D(grid,solids,basis,alg::MUSL,output,fixes,data)
 *       SolveExplicitSolidMechanics2D.jl:25-329 (blended consistent mass)
 *   solve_explicit_dynamics_3D(grid,solids,basis,alg::MUSL,output,fixes,data)
 *       SolveExplicitSolidMechanics3D.jl:14-223
 * both exported by module Problem (Problem.jl:163-166).  A host binding keeps
 * the reference's Grid/Solid/Material/Basis objects and calls these entry
 * points instead of running the Julia time loop.
 *
 * Conventions
 *  - Every function returns 0 (JUMP_OK) or a negative jump_status; no C++
 *    exception crosses the ABI.  jump_last_error() returns the message of the
 *    last failure on a context.
 *  - All floating point data are IEEE binary64.
 *  - Per-particle vectors are passed in the reference's AoS layout, i.e. the
 *    memory of Vector{SVector{d,Float64}} (pos[d*i+k]) and of
 *    Vector{SMatrix{d,d,Float64}} (column-major: M[d*d*i + a + d*b] = M[a,b]).
 *    Host buffers are owned by the caller and copied; the context owns all
 *    device memory.
 *  - Node and particle indices are 0-based (the Julia binding subtracts 1).
 *  - One HIP stream per context; a context is not thread-safe.  Calls that
 *    return host data block until the stream has drained.
 */
#ifndef JUMP_MPM_H
#define JUMP_MPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JUMP_ABI_VERSION 1

typedef enum {
  JUMP_OK = 0,
  JUMP_ERR_ARG = -1,         /* invalid argument                           */
  JUMP_ERR_STATE = -2,       /* call not valid in the current state        */
  JUMP_ERR_DEVICE = -3,      /* HIP runtime failure                        */
  JUMP_ERR_NOMEM = -4,       /* allocation failure                         */
  JUMP_ERR_UNSUPPORTED = -5, /* combination the reference does not define  */
  JUMP_ERR_NUMERIC = -6,     /* step finished with a particle error flag   */
  JUMP_ERR_IO = -7           /* a dump file could not be written           */
} jump_status;

/* Particle error flags raised during a step (OR-ed; jump_flags()).
 * The reference only logs these (SolveExplicitSolidMechanics3D.jl:201-206,
 * Basis.jl:71-74,113-126) and continues; the step here also continues, skips
 * the offending particle's grid access, and reports the flag. */
#define JUMP_FLAG_OUT_OF_GRID 1u  /* stencil leaves the grid: particle skipped */
#define JUMP_FLAG_NEGATIVE_J 2u   /* det(F) < 0 after the update              */
#define JUMP_FLAG_NONFINITE 4u    /* NaN/Inf in updated particle state        */
#define JUMP_FLAG_CONTACT_OVERFLOW 8u /* contact: a node touched by > 2 solids  */
#define JUMP_FLAG_SOLVER 16u        /* reserved: the 2D MUSL solve has a fixed
                                     * iteration count and never raises it     */

/* Basis types: Basis.jl:37-48 (LinearBasis, QuadBsplineBasis,
 * CubicBsplineBasis, CPDIQ4Basis) plus uGIMP, which the reference lacks. */
typedef enum {
  JUMP_BASIS_LINEAR = 0,
  JUMP_BASIS_QUAD_BSPLINE = 1,
  JUMP_BASIS_CUBIC_BSPLINE = 2,
  JUMP_BASIS_CPDI_Q4 = 3, /* 2D only */
  JUMP_BASIS_GIMP = 4
} jump_basis;

/* Algorithm tags: Algorithm.jl:16-18 (USL(tolerance)), :37-39 (MUSL(alpha)).
 * The reference implements USL and MUSL in 2D and MUSL in 3D on this path.
 * The 2D MUSL step solves with the blended mass matrix
 * B = 0.1 diag(m) + 0.9 M (M_IJ = sum m N_I N_J); the reference assembles it
 * densely and factors it, this library keeps it as a (2W-1)^2-point stencil
 * (W the stencil width) and solves it by the Chebyshev iteration
 * preconditioned with diag(m), whose spectrum bound [0.1, 1] holds for any
 * particle layout: 51 iterations bound the error's B-norm by 1e-14 of the
 * solution's.  No contact and no rigid solids with it.
 * JUMP_ALG_APIC (Algorithm.jl:41, declared by the reference with the
 * per-particle affine matrix Cmat of Solid.jl:82, but never stepped): the 2D
 * USL step with the affine particle-in-cell transfers of Jiang et al., "The
 * affine particle-in-cell method" (ACM TOG 34(4), 2015): P2G carries
 * m N (v_p + C_p (x_I - x_p)), G2P takes v_p = sum N v_I and
 * C_p = B_p D_p^-1 (B_p = sum N v_I (x_I - x_p)^T, D_p = sum N (x_I - x_p)
 * (x_I - x_p)^T; for the linear hat C_p = sum v_I grad N^T).  2D only; the
 * linear, B-spline and GIMP bases; no CPDI, contact, batches or reverse mode.
 * C_p starts at zero, as Cmat does (Solid.jl:111); jump_get_affine and
 * jump_set_affine read and write it. */
typedef enum { JUMP_ALG_USL = 0, JUMP_ALG_MUSL = 1, JUMP_ALG_APIC = 2 } jump_algorithm;

/* Material kinds: Material.jl:29-31 (Rigid), :37-81 (Elastic/Hooke),
 * :100-122 (NeoHookean), :128-275 (ElastoPlastic J2), :280-441 (JohnsonCook). */
typedef enum {
  JUMP_MAT_RIGID = 0,
  JUMP_MAT_ELASTIC = 1,
  JUMP_MAT_NEOHOOKEAN = 2,
  JUMP_MAT_ELASTOPLASTIC = 3,
  JUMP_MAT_JOHNSONCOOK = 4
} jump_material_kind;

/* P2G summation order.  The reference folds each node's contributions in
 * particle order (SolveExplicitSolidMechanics2D.jl:391-423).
 * JUMP_ORDER_REFERENCE reproduces that fold exactly (grid and particle state
 * bit-identical to the reference's arithmetic); JUMP_ORDER_BINNED sums bin by
 * bin in a fixed order: deterministic run to run, equal to the reference up
 * to floating-point reassociation of the node sums. */
typedef enum { JUMP_ORDER_REFERENCE = 0, JUMP_ORDER_BINNED = 1 } jump_order;

/* Step parameters: the Grid2D/Grid3D constructor arguments (Grid.jl:104,198),
 * the basis and algorithm objects, and the keys of the reference's `data`
 * Dict (SolveExplicitSolidMechanics2D.jl:337-344: "dt", "bodyforce",
 * "ghostcell"). */
typedef struct {
  int32_t dim;       /* 2 or 3                                            */
  int32_t basis;     /* jump_basis                                        */
  int32_t algorithm; /* jump_algorithm                                    */
  int32_t ghostcell; /* data["ghostcell"] present                         */
  double tolerance;  /* USL(tolerance): nodes with m <= tolerance skipped */
  double alpha;      /* MUSL(alpha): FLIP/PIC blend                       */
  double dt;         /* data["dt"]                                        */
  double xmin, xmax, ymin, ymax, zmin, zmax;
  int64_t nx, ny, nz; /* node counts iN_x, iN_y, iN_z (nz ignored in 2D)   */
  double body[3];     /* ConstantBodyForce2D/3D g (BodyForce.jl:56-91)     */
  int32_t order;      /* jump_order (0: reference summation order)          */
  int32_t contact;    /* 1 = multi-body contact between solids             */
  double friction;    /* Coulomb coefficient of the 2D contact             */
} jump_params;

/* Multi-body contact (contact = 1).  The reference's ULMPM step has none:
 * all solids share one velocity field.  This is the algorithm of its FEMP
 * contact solvers (SolveExplicitFEMP2DContact.jl:311-387 on the 2D USL
 * step; SolveExplicitFEMP3DContact.jl:398-511 on the 3D MUSL step, where
 * the correction is normal-only, v -= D n, and friction is not used; the
 * re-projected momentum is not corrected again) on the MPM grid:
 * each node keeps the mass, momenta, force and normal sum(m_p grad N) of at
 * most two solids (in particle order; a third raises
 * JUMP_FLAG_CONTACT_OVERFLOW); where two solids meet, the common normal is
 * n = (n1 - n2)/|n1 - n2| (-n for the second), and a solid whose velocity
 * relative to the centre-of-mass velocity approaches along its normal
 * (D = (v - v_cm).n > 0) gets v -= D (n + mu'/|C| (n_y C, -n_x C)),
 * C = (v - v_cm) x n, mu' = min(friction, |C|/D) (the tangential term is 0
 * when C = 0, where the reference divides 0 by 0).  G2P then reads each
 * particle's own solid's nodal velocities.  A 2D rigid solid
 * (SolveExplicitFEMP2DContact.jl:272-304, :330-358) reaches the four nodes
 * of each of its particles' hat cells and brings only its normal
 * sum(m_p grad N); where it meets a deformable solid, n is the rigid body's
 * normal (the deformable solid uses -n), v_cm is the rigid velocity (0 if
 * fixed) and only the deformable solid is corrected; rigid solids move with
 * their prescribed velocity.  Batches and the reverse mode do not combine
 * with contact. */

/* Material parameters.  Lame constants are derived exactly as the reference
 * constructors do (Material.jl:54-57,112-113,159-161,313-325). */
typedef struct {
  int32_t kind;  /* jump_material_kind                                  */
  int32_t fixed; /* rigid only: zero the grid momentum under the body   */
  double E, nu, density;
  double fy, k1; /* ElastoPlastic: yield stress, linear hardening          */
  double A, B, C, n, epsdot0, m, chi, Cp, cellsize; /* JohnsonCook          */
  double vx, vy; /* rigid only: prescribed velocity (RigidMaterial(vx,vy,rho)) */
} jump_material;

/* One solid (Solid2D/Solid3D, Solid.jl:67-297).  Arrays are AoS as above;
 * pos, vel: d*n; mass, volume, volume0: n; F, strain, stress: d*d*n.
 * F/strain/stress may be NULL (identity / zero).  For CPDI-Q4 (2D), corners
 * (2*ncorners, the solid's `nodes`) and elems (4*n, 0-based, the solid's
 * `elems` rows) describe the particle domains.  For GIMP, lp[k] is the
 * particle half-width along axis k (must be <= dx_k/2). */
typedef struct {
  int64_t n;
  const double* pos;
  const double* vel;
  const double* mass;
  const double* volume;
  const double* volume0;
  const double* F;
  const double* strain;
  const double* stress;
  int64_t ncorners;
  const double* corners;
  const int64_t* elems;
  double lp[3];
} jump_solid_desc;

typedef struct jump_ctx jump_ctx;

/* Create a context for one grid/basis/algorithm.  Replaces the set-up part of
 * solve_explicit_dynamics_2D/3D (SolveExplicitSolidMechanics2D.jl:335-371,
 * SolveExplicitSolidMechanics3D.jl:14-39) and Grid2D/Grid3D allocation
 * (Grid.jl:104-158, 198-267; the 2D dense mass matrix M is not allocated:
 * USL never reads it and 2D MUSL keeps its nonzero band only).  Selects
 * device `device` (HIP ordinal).  A grid (or, in jump_batch_create, the
 * grids of a batch together) has at most 2^30 - 2 nodes and at most 65535
 * per axis; larger ones are JUMP_ERR_UNSUPPORTED. */
int jump_ctx_create(const jump_params* params, int device, jump_ctx** out);
int jump_ctx_destroy(jump_ctx* ctx);
const char* jump_last_error(const jump_ctx* ctx);

/* Append one solid (Solid2D(coords,mat) Solid.jl:104-125, Solid2D(file,mat)
 * for CPDI Solid.jl:162-221, Solid3D Solid.jl:283-296) with its material.
 * Solids are processed in the order added, as the reference's `solids`
 * vector.  Writes the solid index to *solid_index. */
int jump_add_solid(jump_ctx* ctx, const jump_solid_desc* desc,
                   const jump_material* mat, int32_t* solid_index);

/* Dirichlet mask, the reference's grid.fixedNodes (dim x nodeCount, Int64,
 * Grid.jl:130-131): mask[dim*node + k] != 0 fixes direction k of node.
 * Set by fix_Dirichlet_grid (Dirichlet.jl:28-82) on the host side. */
int jump_set_fixed(jump_ctx* ctx, const uint8_t* mask);

/* Run nsteps iterations of the reference's `while t < Tf` body
 * (SolveExplicitSolidMechanics2D.jl:373-581 for USL, :86-327 for 2D MUSL,
 *  SolveExplicitSolidMechanics3D.jl:41-220 for MUSL) without output, each
 * ending with t += dt.  *t_inout holds t (data["time"]) and is advanced.
 * Returns JUMP_ERR_NUMERIC (after finishing all steps) if a particle error
 * flag was raised; the flags are then available from jump_flags(). */
int jump_step(jump_ctx* ctx, int64_t nsteps, double* t_inout);

/* OR of JUMP_FLAG_* raised since the last call; clears them. */
int jump_flags(jump_ctx* ctx, uint32_t* flags);

/* Copy a solid's particle state back (any pointer may be NULL), in the
 * order the particles were added, AoS as for jump_add_solid. */
int jump_get_solid(jump_ctx* ctx, int32_t solid, double* pos, double* vel,
                   double* mass, double* volume, double* F, double* strain,
                   double* stress);

/* Material history held per particle (the reference keeps it in the
 * material, indexed by particle: Material.jl:140-155, 301-303).
 * alpha: equivalent plastic strain (J2, JC); alphadot (JC); vm: von Mises
 * stress (3D elastic, J2, JC); pstrain: plastic strain, 3 Voigt components
 * [e11, e22, 2*e12] per particle in 2D (Material.jl:140,191) or a full 3x3
 * column-major matrix in 3D (Material.jl:151).  Any pointer may be NULL. */
int jump_get_solid_state(jump_ctx* ctx, int32_t solid, double* alpha,
                         double* alphadot, double* vm, double* pstrain);

/* Body forces.  The 2D steps evaluate the body force at each particle's
 * position at the start of the step, bodyforce(body, xx[ip], t)
 * (SolveExplicitSolidMechanics2D.jl:408 USL, :125 MUSL), and P2G adds
 * m_p b N to the nodal force; the 3D step applies none.
 * JUMP_BODY_CONSTANT: params.body (ConstantBodyForce2D, BodyForce.jl:58-60,
 * :86-88), the default.  JUMP_BODY_VORTEX2D: VortexBodyForce2D(G, T, Ri, Ro,
 * rho0, mu, xcenter, ycenter) (BodyForce.jl:34-53, :94-121), the
 * manufactured body force under which an annulus Ri < r < Ro of a
 * Neo-Hookean solid (shear modulus mu, density rho0) rotates by
 * alpha(r, t) = G sin(pi t / T) h(r), h = 1 - 8 s^2 + 16 s^4,
 * s = (r - (Ri + Ro)/2) / (Ri - Ro) (the exact map of Main_Vortex_FEMP.jl);
 * prm = those 8 values.  JUMP_BODY_VORTEX2D evaluates it at each particle's
 * current position, as the ULMPM step calls bodyforce; the manufactured
 * solution is written in material coordinates (the total-Lagrangian FEMP
 * solver evaluates it at the reference positions), and
 * JUMP_BODY_VORTEX2D_X0 evaluates it at each particle's position when it
 * was added, under which the step converges to the exact map.  The time of
 * a step is the t of jump_step, accumulated by dt per step as the
 * reference's loop does.  2D only; the reverse mode and batches refuse a
 * vortex force.  kind JUMP_BODY_CONSTANT takes prm = the 2 components and
 * replaces params.body. */
typedef enum { JUMP_BODY_CONSTANT = 0, JUMP_BODY_VORTEX2D = 1, JUMP_BODY_VORTEX2D_X0 = 2 } jump_body_kind;
int jump_set_body_force(jump_ctx* ctx, int32_t kind, const double* prm, int32_t nprm);

/* APIC (JUMP_ALG_APIC): the affine matrix C_p of each particle of a solid,
 * 4 doubles per particle, column-major, as Solid.Cmat holds it (Solid.jl:82,
 * "APIC BpDp"; MMatrix{2,2}).  jump_set_affine writes it, for a start from
 * an affine velocity field (C_p = grad v).  Other algorithms:
 * JUMP_ERR_UNSUPPORTED. */
int jump_get_affine(jump_ctx* ctx, int32_t solid, double* cmat);
int jump_set_affine(jump_ctx* ctx, int32_t solid, const double* cmat);

/* Copy the grid arrays of the last step (Grid.jl:78-90, 181-186), AoS
 * d*nodeCount for the vector fields.  Any pointer may be NULL. */
int jump_get_grid(jump_ctx* ctx, double* mass, double* momentum0,
                  double* momentum, double* momentum2, double* force);

/* Kinetic and strain energy over all solids, compute(::EnergiesFix)
 * (Fix.jl:248-270). */
int jump_energies(jump_ctx* ctx, double* ke, double* se);

/* LAMMPS particle dump of all solids, the reference's OvitoOutput
 * (plotParticles_2D/3D, Output.jl:146-260): header, then one row per
 * particle "id type x y z <outs>" (type = 1-based solid index; 2D rows
 * write z as 0) with numbers in Julia's shortest Float64 text, appended to
 * `path` as the reference's open(file, "a") does.  `outs`: space-separated
 * column names from pstrain, vonMises, pressure, vx, vy, and sigmaxx,
 * sigmayy (2D) or vz (3D); others are JUMP_ERR_UNSUPPORTED.  Box bounds are
 * the grid extents (grid.lx, ly, lz).  The state is copied to pinned host
 * memory on the context's stream, ahead of any later step, and a worker
 * thread formats and writes the text while the device goes on stepping.
 * jump_dump_wait blocks until every queued dump is written and returns
 * JUMP_ERR_IO if any dump since the previous jump_dump_wait failed. */
int jump_dump_lammps(jump_ctx* ctx, const char* path, const char* outs);
int jump_dump_wait(jump_ctx* ctx);

/* Checkpoint / restart of the particle state (the reference has none; its
 * dumps, Output.jl:146-290, are write-only).  A checkpoint holds every
 * per-particle array: positions, velocities, mass, volumes, F, strain,
 * stress and the material history; the grid is rebuilt by the next step.
 * The caller keeps t.  jump_checkpoint_bytes gives the buffer size;
 * jump_checkpoint_load accepts only a checkpoint of a context with the same
 * dimension and particle count (the same solids) and rejects anything else
 * with JUMP_ERR_ARG. */
int jump_checkpoint_bytes(jump_ctx* ctx, int64_t* bytes);
int jump_checkpoint_save(jump_ctx* ctx, void* buf, int64_t bytes);
int jump_checkpoint_load(jump_ctx* ctx, const void* buf, int64_t bytes);

/* Batches: many independent 2D USL problems stepped with shared launches
 * (the reference runs one problem per solve_explicit_dynamics_2D call,
 * SolveExplicitSolidMechanics2D.jl:335).  jump_batch_create takes n
 * configured contexts (2D USL, solids added, Dirichlet mask set; any basis
 * but CPDI, any grid size, dt, tolerance and materials) on one device and
 * copies their state; jump_batch_step advances every member nsteps steps
 * (t_inout: n times, each advanced by its member's dt); jump_batch_sync
 * copies the state back into the member contexts, whose getters then read
 * it; jump_batch_energies writes n kinetic and strain energies.  Each member
 * steps bit for bit as it would alone.  The member contexts are not stepped
 * while in a batch; destroy the batch before them. */
typedef struct jump_batch jump_batch;
int jump_batch_create(jump_ctx* const* ctxs, int32_t n, int device, jump_batch** out);
int jump_batch_step(jump_batch* batch, int64_t nsteps, double* t_inout);
int jump_batch_sync(jump_batch* batch);
int jump_batch_energies(jump_batch* batch, double* ke, double* se);
int jump_batch_destroy(jump_batch* batch);
const char* jump_batch_last_error(const jump_batch* batch);

/* Launches of a batch (no counterpart in the reference, which runs one
 * problem per call).  graph != 0 (the default): each group's step sequence
 * is captured once into a HIP graph (one step and a chunk of 8) and replayed.
 * autotune != 0 (the default): at the first jump_batch_step the batched P2G
 * gather is timed in each of the JUMP_P2G_SHAPES launch shapes on the
 * batch's own particles and the fastest is kept; autotune == 0: every group
 * uses `shape` (JUMP_ERR_ARG outside 0..JUMP_P2G_SHAPES-1).  Shapes: 0 active nodes
 * compacted, 128 threads per workgroup; 1 compacted, 64; 2 compacted, 256;
 * 3 every node in grid-row order, 128; 4 rows, 64; 5 rows, 256; 6, 7, 8 a
 * persistent kernel of 128, 64 or 256 threads per workgroup, as many
 * workgroups as stay resident, each taking (instance, tile) items of the
 * compacted order from a work queue.  Every shape sums each node in the
 * same order, so neither option changes any result.
 * jump_batch_tuning reports group g's chosen shape and the measured ms per
 * launch of each shape (ms: JUMP_P2G_SHAPES doubles, may be NULL) and the
 * number of groups (members with the same basis and summation order, in
 * order of first appearance); JUMP_ERR_STATE before the group is tuned.  A
 * group of one member steps as a single context and has no batched gather:
 * it is not tuned, and its shape is reported as -1. */
#define JUMP_P2G_SHAPES 9
int jump_batch_options(jump_batch* batch, int32_t graph, int32_t autotune, int32_t shape);
int jump_batch_tuning(jump_batch* batch, int32_t group, int32_t* ngroups, int32_t* shape, double* ms);

/* Reverse mode of the 2D USL step and of the 3D MUSL step (the reference has
 * no derivatives; this is the pullback a ChainRulesCore rrule around
 * solve_explicit_dynamics_2D/3D, SolveExplicitSolidMechanics2D.jl:335 and
 * SolveExplicitSolidMechanics3D.jl:14, would call).  Covers the linear and
 * GIMP bases and Hooke / Neo-Hookean solids, and in 3D J2 solids
 * (Material.jl:228-275: the derivative of the return-map branch each
 * particle takes), else JUMP_ERR_UNSUPPORTED.  jump_adjoint_forward runs
 * nsteps steps like jump_step and records the particle state every K steps.
 * jump_adjoint_backward takes the adjoint of the final positions (pos_bar,
 * AoS n x dim, required) and velocities (vel_bar, may be NULL) and writes,
 * per solid, dL/dE, dL/dnu and dL/dv0 (the derivative with respect to a
 * uniform change of that solid's initial velocity, dim per solid); outputs
 * may be NULL.  Each step of the sweep recomputes its states from the
 * nearest snapshot (checkpoint and recompute); the context is left in the
 * recorded final state.  The record belongs to the particle state the run
 * ended in: jump_step (nsteps > 0), jump_checkpoint_load, jump_add_solid or
 * jump_batch_sync on the context discard it, and jump_adjoint_backward then
 * returns JUMP_ERR_STATE instead of rewinding the particles.
 * checkpoint_every = 0 picks the interval: the whole run when its per-step
 * states, grids and bins fit in three quarters of the free device memory
 * (no recomputation), else the largest interval that fits.
 * jump_adjoint_interval reports the interval of the recorded run.
 * jump_adjoint_plastic gives, per solid, dL/dfy (yield stress) and dL/dk1
 * (linear hardening) from the last jump_adjoint_backward of the current
 * record (0 for solids without J2 and in 2D; JUMP_ERR_STATE before it). */
int jump_adjoint_forward(jump_ctx* ctx, int64_t nsteps, int32_t checkpoint_every, double* t_inout);
int jump_adjoint_interval(jump_ctx* ctx, int32_t* checkpoint_every);
int jump_adjoint_backward(jump_ctx* ctx, const double* pos_bar, const double* vel_bar, double* dE,
                          double* dnu, double* dvel0);
int jump_adjoint_plastic(jump_ctx* ctx, double* dfy, double* dk1);

/* Kinetic and strain energy of particles held in the caller's device memory,
 * allocating nothing.  No counterpart in the reference: this is the reduction
 * of jump_energies (Fix.jl:248-270, the same sums in the same order, so the
 * same bits as jump_energies on the same particles) for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  All pointers are device pointers on the
 * current device: mass, volume (n), vel (dim*n), strain, stress (dim*dim*n,
 * column-major, symmetric), in the AoS layout of jump_add_solid; ke_se
 * receives KE then SE.  The work is queued on `stream` (a hipStream_t; NULL:
 * the null stream) and is complete when the stream reaches it.  Partial sums
 * live in a fixed device array of the library (1 MiB per device, held until
 * the process ends): n is at most 16,777,216 (else JUMP_ERR_UNSUPPORTED),
 * and calls that overlap on one device race on that array. */
int jump_energies_noalloc(int32_t dim, int64_t n, const double* mass, const double* vel, const double* volume,
                          const double* strain, const double* stress, double* ke_se, void* stream);

/* Diagnostics (no counterpart in the reference, which recomputes the support
 * inside every shape-function call, Basis.jl:63-78,390-425).
 * jump_get_bins: runs the binning of the current particle positions and
 * returns, per particle in insertion order, the 0-based linear index of the
 * first node of its stencil (-1: outside the grid, or rigid in 2D outside
 * a contact context; a contact context bins a rigid particle by its hat
 * cell).  For
 * tests of the bit-exact support rule. */
int jump_get_bins(jump_ctx* ctx, int32_t* keys);
int jump_abi_version(void);

/* Per-phase GPU time (diagnostic, no counterpart in the reference, whose only
 * timing hook is `@time main()`, Main_TwoDisks.jl:142).  While enabled, every
 * phase of jump_step is bracketed by HIP events recorded on the context's
 * stream; jump_profile(ctx, on) enables/disables and resets the totals,
 * jump_profile_read returns the summed milliseconds and the number of
 * intervals of one phase. */
typedef enum {
  JUMP_PHASE_SORT = 0, /* binning + counting sort                    */
  JUMP_PHASE_P2G = 1,  /* P2G fused with the grid update             */
  JUMP_PHASE_RIGID = 2,/* 2D rigid-body grid override and motion     */
  JUMP_PHASE_G2P = 3,  /* USL G2P + stress                           */
  JUMP_PHASE_G2P1 = 4, /* MUSL particle velocity                     */
  JUMP_PHASE_P2G2 = 5, /* MUSL momentum re-projection                */
  JUMP_PHASE_G2P2 = 6, /* MUSL position, deformation, stress         */
  JUMP_PHASE_SOLVE = 7,/* 2D MUSL blended mass solve                 */
  JUMP_PHASE_COUNT = 8
} jump_phase;
int jump_profile(jump_ctx* ctx, int32_t enable);
int jump_profile_read(jump_ctx* ctx, int32_t phase, double* total_ms, int64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* JUMP_MPM_H */
