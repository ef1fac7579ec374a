// oracle_capi.cpp -- extern "C" wrapper of the CPU restatement (T = double).
//
// TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py
// cpu_baseline).  The entry points mirror include/jump_mpm.h with an `orc_`
// prefix so that a test drives the oracle and the HIP library through the same
// call sequence; a few extra hooks expose single functions (B-spline kernels,
// supports, stress updates) for closed-form checks.
#include <cstdio>
#include <exception>
#include <new>

#include "mpm_oracle.hpp"
#include "omp_step.hpp"
#include "dual.hpp"

using orc::i64;
typedef orc::Ctx<double> OCtx;

struct orc_ctx {
  OCtx c;
};

static int fail(orc_ctx* h, int code, const char* msg) {
  if (h) h->c.err = msg;
  return code;
}

extern "C" {

int orc_ctx_create(const jump_params* p, orc_ctx** out) {
  if (!p || !out) return JUMP_ERR_ARG;
  *out = nullptr;
  if (p->dim != 2 && p->dim != 3) return JUMP_ERR_ARG;
  if (p->nx < 2 || p->ny < 2 || (p->dim == 3 && p->nz < 2)) return JUMP_ERR_ARG;
  if (p->basis < 0 || p->basis > JUMP_BASIS_GIMP) return JUMP_ERR_ARG;
  if (p->dim == 3 && p->basis == JUMP_BASIS_CPDI_Q4) return JUMP_ERR_UNSUPPORTED;
  if (p->dim == 2 && p->algorithm == JUMP_ALG_MUSL && (p->contact || p->nx * p->ny > 4096))
    return JUMP_ERR_UNSUPPORTED;  // the dense 2D MUSL restatement: no contact, up to 4096 nodes
  if (p->dim == 3 && p->algorithm != JUMP_ALG_MUSL) return JUMP_ERR_UNSUPPORTED;
  if (p->algorithm == JUMP_ALG_APIC && (p->contact || p->basis == JUMP_BASIS_CPDI_Q4))
    return JUMP_ERR_UNSUPPORTED;  // APIC: the 2D step without contact, point bases
  if (p->algorithm < JUMP_ALG_USL || p->algorithm > JUMP_ALG_APIC) return JUMP_ERR_ARG;
  orc_ctx* h = new (std::nothrow) orc_ctx();
  if (!h) return JUMP_ERR_NOMEM;
  OCtx& g = h->c;
  g.dim = p->dim;
  g.basis = p->basis;
  g.alg = p->algorithm;
  g.ghostcell = p->ghostcell != 0;
  g.tol = p->tolerance;
  g.alpha = p->alpha;
  g.dt = p->dt;
  for (int k = 0; k < 3; k++) g.body[k] = p->body[k];
  g.contact = p->contact != 0;
  g.fric = p->friction;
  g.nx = p->nx;
  g.ny = p->ny;
  g.nz = p->dim == 3 ? p->nz : 1;
  g.nxy = g.nx * g.ny;
  g.nn = g.nxy * g.nz;
  // Grid.jl:105-108, 199-204
  g.xmin = p->xmin; g.ymin = p->ymin; g.zmin = p->dim == 3 ? p->zmin : 0.0;
  g.dx = (p->xmax - p->xmin) / double(double(p->nx) - 1.0);
  g.dy = (p->ymax - p->ymin) / double(double(p->ny) - 1.0);
  g.dz = p->dim == 3 ? (p->zmax - p->zmin) / double(double(p->nz) - 1.0) : 1.0;
  g.dxI = 1.0 / g.dx; g.dyI = 1.0 / g.dy; g.dzI = 1.0 / g.dz;
  try {
    size_t nn = size_t(g.nn), d = size_t(g.dim);
    g.gmass.assign(nn, 0.0);
    g.gp0.assign(d * nn, 0.0);
    g.gp.assign(d * nn, 0.0);
    g.gp2.assign(d * nn, 0.0);
    g.gf.assign(d * nn, 0.0);
    g.fixed.assign(d * nn, 0);
    g.near.assign(64, 0);
    g.funcs.assign(64, 0.0);
    g.ders.assign(3 * 64, 0.0);
  } catch (const std::exception&) {
    delete h;
    return JUMP_ERR_NOMEM;
  }
  *out = h;
  return JUMP_OK;
}

int orc_ctx_destroy(orc_ctx* h) {
  delete h;
  return JUMP_OK;
}

const char* orc_last_error(const orc_ctx* h) { return h ? h->c.err.c_str() : "null context"; }

int orc_add_solid(orc_ctx* h, const jump_solid_desc* ds, const jump_material* mat, int32_t* idx) {
  if (!h || !ds || !mat) return JUMP_ERR_ARG;
  OCtx& g = h->c;
  const int d = g.dim, dd = d * d;
  if (ds->n < 0 || (ds->n > 0 && (!ds->pos || !ds->vel || !ds->mass || !ds->volume || !ds->volume0)))
    return fail(h, JUMP_ERR_ARG, "missing particle arrays");
  if (mat->kind < JUMP_MAT_RIGID || mat->kind > JUMP_MAT_JOHNSONCOOK) return fail(h, JUMP_ERR_ARG, "bad material");
  if (d == 3 && mat->kind == JUMP_MAT_RIGID) return fail(h, JUMP_ERR_UNSUPPORTED, "3D solids cannot be rigid");
  if (d == 2 && mat->kind == JUMP_MAT_JOHNSONCOOK) return fail(h, JUMP_ERR_UNSUPPORTED, "Johnson-Cook is 3D only");
  if (d == 2 && g.alg == JUMP_ALG_MUSL && mat->kind == JUMP_MAT_RIGID)
    return fail(h, JUMP_ERR_UNSUPPORTED, "the 2D MUSL step has no rigid-body treatment");
  if (g.basis == JUMP_BASIS_CPDI_Q4 && (!ds->corners || !ds->elems || ds->ncorners <= 0))
    return fail(h, JUMP_ERR_ARG, "CPDI solid needs corners and elems");
  orc::Solid<double> s;
  size_t n = size_t(ds->n);
  s.n = ds->n;
  s.mat = orc::make_material<double>(*mat);
  s.rigid = mat->kind == JUMP_MAT_RIGID;
  s.mass.assign(ds->mass, ds->mass + n);
  s.vol.assign(ds->volume, ds->volume + n);
  s.vol0.assign(ds->volume0, ds->volume0 + n);
  s.pos.assign(ds->pos, ds->pos + d * n);
  s.vel.assign(ds->vel, ds->vel + d * n);
  s.pos0 = s.pos;
  s.F.assign(dd * n, 0.0);
  s.strain.assign(dd * n, 0.0);
  s.stress.assign(dd * n, 0.0);
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < d; k++) s.F[dd * i + k + d * k] = 1.0;
  if (ds->F) s.F.assign(ds->F, ds->F + dd * n);
  if (ds->strain) s.strain.assign(ds->strain, ds->strain + dd * n);
  if (ds->stress) s.stress.assign(ds->stress, ds->stress + dd * n);
  s.alpha.assign(n, 0.0);
  s.alphadot.assign(n, 0.0);
  s.vm.assign(n, 0.0);
  s.pstrain.assign((d == 2 ? 3 : 9) * n, 0.0);
  if (g.alg == JUMP_ALG_APIC) s.cmat.assign(4 * n, 0.0);  // Solid.jl:111
  if (g.basis == JUMP_BASIS_CPDI_Q4) {
    s.corners.assign(ds->corners, ds->corners + 2 * size_t(ds->ncorners));
    s.elems.assign(ds->elems, ds->elems + 4 * n);
    for (size_t i = 0; i < 4 * n; i++)
      if (s.elems[i] < 0 || s.elems[i] >= ds->ncorners) return fail(h, JUMP_ERR_ARG, "CPDI elems out of range");
  }
  if (g.basis == JUMP_BASIS_GIMP && !s.rigid) {
    const double hs[3] = {g.dx, g.dy, g.dz};
    for (int k = 0; k < d; k++) {
      if (!(ds->lp[k] > 0.0) || ds->lp[k] > 0.5 * hs[k]) return fail(h, JUMP_ERR_ARG, "GIMP lp must be in (0, dx/2]");
      s.lp[k] = ds->lp[k];
    }
  }
  g.solids.push_back(std::move(s));
  if (idx) *idx = int32_t(g.solids.size() - 1);
  return JUMP_OK;
}

int orc_set_fixed(orc_ctx* h, const uint8_t* mask) {
  if (!h || !mask) return JUMP_ERR_ARG;
  OCtx& g = h->c;
  for (size_t i = 0; i < g.fixed.size(); i++) g.fixed[i] = mask[i] ? 1 : 0;
  return JUMP_OK;
}

int orc_step(orc_ctx* h, int64_t nsteps, double* t) {
  if (!h || !t || nsteps < 0) return JUMP_ERR_ARG;
  OCtx& g = h->c;
  for (int64_t s = 0; s < nsteps; s++) {
    if (g.dim == 2 && g.contact) orc::step_usl_2d_contact(g, *t);
    else if (g.dim == 2 && g.alg == JUMP_ALG_MUSL) orc::step_musl_2d(g, *t);
    else if (g.dim == 2 && g.alg == JUMP_ALG_APIC) orc::step_apic_2d(g, *t);
    else if (g.dim == 2) orc::step_usl_2d(g, *t);
    else if (g.contact) orc::step_musl_3d_contact(g);
    else orc::step_musl_3d(g);
    *t += g.dt;
  }
  return g.flags ? JUMP_ERR_NUMERIC : JUMP_OK;
}

// The 3D MUSL step on nthreads OpenMP threads (omp_step.hpp: the CPU
// baseline at many cores; equal to orc_step up to the rounding of the node
// sums, which it forms block by block).  Other steps: JUMP_ERR_UNSUPPORTED.
int orc_step_parallel(orc_ctx* h, int64_t nsteps, double* t, int nthreads) {
  if (!h || !t || nsteps < 0 || nthreads < 1) return JUMP_ERR_ARG;
  OCtx& g = h->c;
  if (g.dim != 3 || g.contact) return JUMP_ERR_UNSUPPORTED;
  for (int64_t s = 0; s < nsteps; s++) {
    orc::step_musl_3d_omp(g, nthreads);
    *t += g.dt;
  }
  return g.flags ? JUMP_ERR_NUMERIC : JUMP_OK;
}

// ForwardDiff-style gradient of the CPU path.  For the linear functional
// L = sum_p pos_bar_p . x_p + vel_bar_p . v_p of the state after nsteps steps
// (all solids' particles in order, AoS), the derivatives of L with respect
// to solid `solid`'s E, nu, yield stress fy, hardening k1 and the components
// of a uniform change of its initial velocity (grad: 4 + dim values), from a
// run of the restatement on dual numbers seeded on those parameters.  The
// context itself is not stepped.
}  // extern "C"
typedef orc::Dual<7> D7;
static void to_dual(const orc::MatConst<double>& a, orc::MatConst<D7>& b) {
  b.kind = a.kind; b.fixed = a.fixed;
  b.E = a.E; b.nu = a.nu; b.density = a.density; b.lambda = a.lambda; b.mu = a.mu; b.kappa = a.kappa;
  b.fy = a.fy; b.k1 = a.k1; b.A = a.A; b.B = a.B; b.C = a.C; b.n = a.n; b.epsdot0 = a.epsdot0; b.m = a.m;
  b.chi = a.chi; b.Cp = a.Cp; b.cellsize = a.cellsize; b.velo = a.velo; b.delta = a.delta; b.vx = a.vx; b.vy = a.vy;
}
template <class V> static std::vector<D7> dvec(const std::vector<V>& a) { return std::vector<D7>(a.begin(), a.end()); }
extern "C" {

int orc_dual_gradient(orc_ctx* h, int32_t solid, int64_t nsteps, const double* pos_bar, const double* vel_bar,
                      double* grad) {
  if (!h || !pos_bar || !grad || nsteps < 0) return JUMP_ERR_ARG;
  const OCtx& g = h->c;
  if (solid < 0 || size_t(solid) >= g.solids.size()) return fail(h, JUMP_ERR_ARG, "bad solid index");
  if (g.contact || (g.dim == 2 && g.alg != JUMP_ALG_USL)) return fail(h, JUMP_ERR_UNSUPPORTED, "dual run: USL 2D, MUSL 3D");
  if (g.body_kind != JUMP_BODY_CONSTANT) return fail(h, JUMP_ERR_UNSUPPORTED, "dual run: constant body force only");
  orc::Ctx<D7> q;
  q.dim = g.dim; q.basis = g.basis; q.alg = g.alg; q.ghostcell = g.ghostcell;
  q.tol = g.tol; q.alpha = g.alpha; q.dt = g.dt;
  for (int k = 0; k < 3; k++) q.body[k] = g.body[k];
  q.nx = g.nx; q.ny = g.ny; q.nz = g.nz; q.nxy = g.nxy; q.nn = g.nn;
  q.xmin = g.xmin; q.ymin = g.ymin; q.zmin = g.zmin;
  q.dx = g.dx; q.dy = g.dy; q.dz = g.dz; q.dxI = g.dxI; q.dyI = g.dyI; q.dzI = g.dzI;
  q.gmass = dvec(g.gmass); q.gp0 = dvec(g.gp0); q.gp = dvec(g.gp); q.gp2 = dvec(g.gp2); q.gf = dvec(g.gf);
  q.fixed = g.fixed;
  q.near = g.near;
  q.funcs = dvec(g.funcs);
  q.ders = dvec(g.ders);
  const int d = g.dim;
  for (size_t si = 0; si < g.solids.size(); si++) {
    const auto& a = g.solids[si];
    orc::Solid<D7> b;
    b.n = a.n; b.rigid = a.rigid;
    to_dual(a.mat, b.mat);
    b.mass = dvec(a.mass); b.vol0 = dvec(a.vol0); b.vol = dvec(a.vol); b.pos = dvec(a.pos); b.vel = dvec(a.vel);
    b.F = dvec(a.F); b.strain = dvec(a.strain); b.stress = dvec(a.stress);
    b.alpha = dvec(a.alpha); b.alphadot = dvec(a.alphadot); b.vm = dvec(a.vm); b.pstrain = dvec(a.pstrain);
    b.corners = dvec(a.corners); b.elems = a.elems;
    for (int k = 0; k < 3; k++) b.lp[k] = a.lp[k];
    if (int(si) == solid) {  // seeds: E, nu, fy, k1, v0
      b.mat.E.d[0] = 1.0;
      b.mat.nu.d[1] = 1.0;
      orc::derive_constants(b.mat);
      b.mat.fy.d[2] = 1.0;
      b.mat.k1.d[3] = 1.0;
      for (orc::i64 ip = 0; ip < b.n; ip++)
        for (int k = 0; k < d; k++) b.vel[size_t(d * ip + k)].d[4 + k] = 1.0;
    }
    q.solids.push_back(std::move(b));
  }
  for (int64_t s = 0; s < nsteps; s++) {
    if (d == 2) orc::step_usl_2d(q, D7(0.0));
    else orc::step_musl_3d(q);
  }
  for (int k = 0; k < 4 + d; k++) grad[k] = 0.0;
  size_t off = 0;
  for (const auto& b : q.solids) {
    for (orc::i64 ip = 0; ip < b.n; ip++)
      for (int c = 0; c < d; c++) {
        const D7& x = b.pos[size_t(d * ip + c)];
        const D7& v = b.vel[size_t(d * ip + c)];
        const double pb = pos_bar[d * (off + size_t(ip)) + size_t(c)];
        const double vb = vel_bar ? vel_bar[d * (off + size_t(ip)) + size_t(c)] : 0.0;
        for (int k = 0; k < 4 + d; k++) grad[k] += pb * x.d[k] + vb * v.d[k];
      }
    off += size_t(b.n);
  }
  return q.flags ? JUMP_ERR_NUMERIC : JUMP_OK;
}

int orc_flags(orc_ctx* h, uint32_t* flags) {
  if (!h || !flags) return JUMP_ERR_ARG;
  *flags = h->c.flags;
  h->c.flags = 0;
  return JUMP_OK;
}

static void copy_out(const std::vector<double>& v, double* dst) {
  if (dst) std::copy(v.begin(), v.end(), dst);
}

int orc_get_solid(orc_ctx* h, int32_t si, double* pos, double* vel, double* mass, double* vol, double* F,
                  double* strain, double* stress) {
  if (!h || si < 0 || size_t(si) >= h->c.solids.size()) return JUMP_ERR_ARG;
  const auto& s = h->c.solids[size_t(si)];
  copy_out(s.pos, pos); copy_out(s.vel, vel); copy_out(s.mass, mass); copy_out(s.vol, vol);
  copy_out(s.F, F); copy_out(s.strain, strain); copy_out(s.stress, stress);
  return JUMP_OK;
}

int orc_get_solid_state(orc_ctx* h, int32_t si, double* alpha, double* alphadot, double* vm, double* pstrain) {
  if (!h || si < 0 || size_t(si) >= h->c.solids.size()) return JUMP_ERR_ARG;
  const auto& s = h->c.solids[size_t(si)];
  copy_out(s.alpha, alpha); copy_out(s.alphadot, alphadot); copy_out(s.vm, vm); copy_out(s.pstrain, pstrain);
  return JUMP_OK;
}

int orc_set_body_force(orc_ctx* h, int32_t kind, const double* prm, int32_t nprm) {
  if (!h || (!prm && nprm > 0)) return JUMP_ERR_ARG;
  OCtx& g = h->c;
  if (kind == JUMP_BODY_CONSTANT) {
    if (nprm != g.dim) return fail(h, JUMP_ERR_ARG, "a constant body force takes dim values");
    for (int k = 0; k < g.dim; k++) g.body[k] = prm[k];
    g.body_kind = JUMP_BODY_CONSTANT;
    return JUMP_OK;
  }
  if ((kind != JUMP_BODY_VORTEX2D && kind != JUMP_BODY_VORTEX2D_X0) || nprm != 8)
    return fail(h, JUMP_ERR_ARG, "bad body force");
  if (g.dim != 2) return fail(h, JUMP_ERR_UNSUPPORTED, "the vortex body force is 2D");
  if (!(prm[1] != 0.0 && prm[2] != prm[3] && prm[4] > 0.0)) return fail(h, JUMP_ERR_ARG, "bad vortex parameters");
  for (int k = 0; k < 8; k++) g.vortex[k] = prm[k];
  g.body_kind = kind;
  return JUMP_OK;
}

int orc_get_affine(orc_ctx* h, int32_t si, double* cmat) {
  if (!h || si < 0 || size_t(si) >= h->c.solids.size() || !cmat) return JUMP_ERR_ARG;
  if (h->c.alg != JUMP_ALG_APIC) return JUMP_ERR_UNSUPPORTED;
  copy_out(h->c.solids[size_t(si)].cmat, cmat);
  return JUMP_OK;
}

int orc_set_affine(orc_ctx* h, int32_t si, const double* cmat) {
  if (!h || si < 0 || size_t(si) >= h->c.solids.size() || !cmat) return JUMP_ERR_ARG;
  if (h->c.alg != JUMP_ALG_APIC) return JUMP_ERR_UNSUPPORTED;
  auto& s = h->c.solids[size_t(si)];
  s.cmat.assign(cmat, cmat + 4 * size_t(s.n));
  return JUMP_OK;
}

int orc_get_grid(orc_ctx* h, double* mass, double* p0, double* p, double* p2, double* f) {
  if (!h) return JUMP_ERR_ARG;
  const OCtx& g = h->c;
  copy_out(g.gmass, mass); copy_out(g.gp0, p0); copy_out(g.gp, p); copy_out(g.gp2, p2); copy_out(g.gf, f);
  return JUMP_OK;
}

int orc_energies(orc_ctx* h, double* ke, double* se) {
  if (!h || !ke || !se) return JUMP_ERR_ARG;
  orc::energies(h->c, *ke, *se);
  return JUMP_OK;
}

// ------------------------------------------------------------ test hooks
// Quadratic (family 2) and cubic (family 3) B-spline kernels of type 1..4.
int orc_bspline(int family, int type, double r, double* N, double* dN) {
  void (*fq[4])(const double&, double&, double&) = {orc::qbs1<double>, orc::qbs2<double>, orc::qbs3<double>,
                                                     orc::qbs4<double>};
  void (*fc[4])(const double&, double&, double&) = {orc::cbs1<double>, orc::cbs2<double>, orc::cbs3<double>,
                                                     orc::cbs4<double>};
  if (type < 1 || type > 4) return JUMP_ERR_ARG;
  if (family == 2) fq[type - 1](r, *N, *dN);
  else if (family == 3) fc[type - 1](r, *N, *dN);
  else return JUMP_ERR_ARG;
  return JUMP_OK;
}

int orc_gimp1d(double xi, double h, double lp, double* N, double* dN) {
  orc::gimp_axis<double>(xi, h, 1.0 / h, lp, *N, *dN);
  return JUMP_OK;
}

// Support of particle p of solid s: returns the count (or -1 outside the grid),
// writes 0-based nodes, N and gradients (dim per node).  grad=0 selects the
// getShapeFunctions variant.
int orc_support(orc_ctx* h, int32_t si, int64_t p, int grad, int64_t* near, double* funcs, double* ders) {
  if (!h || si < 0 || size_t(si) >= h->c.solids.size()) return JUMP_ERR_ARG;
  OCtx& g = h->c;
  const auto& s = g.solids[size_t(si)];
  if (p < 0 || p >= s.n) return JUMP_ERR_ARG;
  int cnt = orc::support(g, s, p, grad != 0);
  for (int i = 0; i < cnt; i++) {
    near[i] = g.near[i];
    funcs[i] = g.funcs[i];
    if (grad)
      for (int k = 0; k < g.dim; k++) ders[g.dim * i + k] = g.ders[g.dim * i + k];
  }
  return cnt;
}

// One constitutive update on full d x d column-major tensors.  state layout:
// [alpha, alphadot, vm, pstrain(3 or 9)].
int orc_stress(int dim, const jump_material* m, const double* eps, const double* deps, const double* F, double J,
               double dt, double* sig, double* state) {
  if (!m || (dim != 2 && dim != 3)) return JUMP_ERR_ARG;
  orc::MatConst<double> c = orc::make_material<double>(*m);
  double& alpha = state[0];
  double& alphadot = state[1];
  double& vm = state[2];
  double* ep = state + 3;
  if (dim == 2) {
    switch (m->kind) {
      case JUMP_MAT_ELASTIC: orc::stress_hooke_2d(c, eps, sig); break;
      case JUMP_MAT_NEOHOOKEAN: orc::stress_nh(2, c, F, J, sig); break;
      case JUMP_MAT_ELASTOPLASTIC: orc::stress_j2_2d(c, eps, sig, ep, alpha, vm); break;
      default: return JUMP_ERR_UNSUPPORTED;
    }
  } else {
    switch (m->kind) {
      case JUMP_MAT_ELASTIC: orc::stress_hooke_3d(c, eps, sig, vm); break;
      case JUMP_MAT_NEOHOOKEAN: orc::stress_nh(3, c, F, J, sig); break;
      case JUMP_MAT_ELASTOPLASTIC: orc::stress_j2_3d(c, eps, sig, ep, alpha, vm); break;
      case JUMP_MAT_JOHNSONCOOK: orc::stress_jc(c, eps, deps, dt, sig, alpha, alphadot, vm); break;
      default: return JUMP_ERR_UNSUPPORTED;
    }
  }
  return JUMP_OK;
}

double orc_det(int dim, const double* A) { return dim == 2 ? orc::det2(A) : orc::det3(A); }

}  // extern "C"
