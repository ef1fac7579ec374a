// mpm_oracle.hpp -- CPU restatement of jump's explicit ULMPM step.
//
// TEST INFRASTRUCTURE ONLY.  This is the parity oracle for the HIP path: it is
// linked by tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline
// leg, never by the product library.  It restates the reference loops in the
// reference's own operation order (left-to-right sums, no FMA contraction:
// build with -ffp-contract=off), templated on the scalar type T so that a
// dual-number instantiation can differentiate it.
//
// Reference files followed (file:line in vinhphunguyen/jump):
//   USL 2D step        SolveExplicitSolidMechanics2D.jl:335-585
//   MUSL 3D step       SolveExplicitSolidMechanics3D.jl:14-223
//   support lookup     Basis.jl:63-78, 81-98, 104-137
//   shape functions    Basis.jl:219-251, 292-332, 374-589, 636-820,
//                      852-983, 1023-1231, 1271-1450, 1454-1490, 1496-1652
//   B-spline kernels   Bsplines.jl:18-171
//   materials          Material.jl:53-81, 110-122, 157-275, 311-441
//   grid               Grid.jl:104-158, 198-267, 275-291 (M omitted)
//   Dirichlet          Dirichlet.jl:28-82
//   energies           Fix.jl:248-270
// Operation order inherited from StaticArrays.jl (2x2/3x3 products, det,
// dot, sum) is assumed textbook: parity unpinned at the last ulp there.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../include/jump_mpm.h"

namespace orc {

using i64 = long long;

// ---------------------------------------------------------------- scalars
inline double val(double x) { return x; }

// Julia sign(x): +-1 for nonzero, x itself for +-0.
template <class T> inline T jsign(const T& x) {
  if (val(x) > 0) return T(1.0);
  if (val(x) < 0) return T(-1.0);
  return x * 0.0;
}
template <class T> inline T jabs(const T& x) { return val(x) < 0 ? -x : x; }
// floor(Int64, x)
template <class T> inline i64 ifloor(const T& x) { return (i64)std::floor(val(x)); }
template <class T> inline T tfloor(const T& x) { return T(std::floor(val(x))); }

// ---------------------------------------------------------------- material
template <class T> struct MatConst {
  int kind = JUMP_MAT_ELASTIC;
  bool fixed = false;
  T E, nu, density, lambda, mu, kappa;
  T fy, k1;
  T A, B, C, n, epsdot0, m, chi, Cp, cellsize, velo, delta;
  T vx, vy;
};

// Material.jl:54-57, 112-113, 159-161, 313-325: the constants derived from
// E, nu and the density (separate so that a dual-number instantiation can
// seed E and nu and derive the rest)
template <class T> void derive_constants(MatConst<T>& c) {
  c.lambda = c.E * c.nu / (1.0 + c.nu) / (1.0 - 2.0 * c.nu);
  c.mu = c.E / 2.0 / (1.0 + c.nu);
  c.kappa = c.lambda + c.mu;
  if (c.kind == JUMP_MAT_JOHNSONCOOK) {
    c.velo = sqrt((c.lambda + 2.0 * c.mu) / c.density);
    c.delta = c.chi / (c.density * c.Cp);
  } else {
    c.velo = 0.0; c.delta = 0.0;
  }
}
template <class T> MatConst<T> make_material(const jump_material& in) {
  MatConst<T> c;
  c.kind = in.kind;
  c.fixed = in.fixed != 0;
  c.E = in.E; c.nu = in.nu; c.density = in.density;
  c.fy = in.fy; c.k1 = in.k1;
  c.A = in.A; c.B = in.B; c.C = in.C; c.n = in.n; c.epsdot0 = in.epsdot0;
  c.m = in.m; c.chi = in.chi; c.Cp = in.Cp; c.cellsize = in.cellsize;
  derive_constants(c);
  c.vx = in.vx; c.vy = in.vy;
  return c;
}

// ---------------------------------------------------------------- containers
template <class T> struct Solid {
  i64 n = 0;
  bool rigid = false;
  MatConst<T> mat;
  std::vector<T> mass, vol0, vol, pos, vel, F, strain, stress;  // AoS
  std::vector<T> pos0;  // positions when added (the material coordinates of a body force)
  // material history, per particle (reference keeps it in the material)
  std::vector<T> alpha, alphadot, vm, pstrain;  // pstrain: 3 (2D) or 9 (3D)
  std::vector<T> cmat;  // APIC: the affine matrix C_p = B_p D_p^-1, 4 per particle column-major (Solid.jl:82)
  // CPDI-Q4 particle domains: static corners and 0-based connectivity
  std::vector<T> corners;
  std::vector<i64> elems;
  T lp[3] = {0.0, 0.0, 0.0};  // GIMP half-widths
};

template <class T> struct Ctx {
  int dim = 2, basis = 0, alg = 0;
  bool ghostcell = false;
  T tol, alpha, dt;
  T body[3];
  // grid (Grid.jl:104-108, 198-204)
  i64 nx = 0, ny = 0, nz = 1, nxy = 0, nn = 0;
  T xmin, ymin, zmin, dx, dy, dz, dxI, dyI, dzI;
  std::vector<T> gmass, gp0, gp, gp2, gf;  // AoS dim*nn
  std::vector<uint8_t> fixed;              // dim*nn
  std::vector<Solid<T>> solids;
  uint32_t flags = 0;
  bool contact = false;  // multi-body contact (2D USL, 3D MUSL)
  int body_kind = 0;     // JUMP_BODY_CONSTANT (body), JUMP_BODY_VORTEX2D(_X0) (vortex)
  double vortex[8] = {};  // G, T, Ri, Ro, rho0, mu, xcenter, ycenter
  T fric;
  // contact: per node two solid slots; per-slot vectors are indexed
  // (2 node + slot) dim + component
  std::vector<int> cbody;                    // 2*nn, -1 empty
  std::vector<T> cm, cp0, cf, cn, cv0, cv;   // 2*nn, 2*dim*nn
  std::vector<T> cp, cp2;                    // 3D: momentum, momentum2
  std::vector<T> mv;                         // 2D MUSL: vx0, vy0, vx, vy per node
  std::string err;
  // scratch for the support (Basis.jl initialise, :1658-1727)
  std::vector<i64> near;
  std::vector<T> funcs, ders;
};

// node coordinate along an axis: xmin + (i-1)*dx with 1-based i (Grid.jl:149-152)
template <class T> inline T node_x(const Ctx<T>& g, i64 i1) { return g.xmin + double(i1 - 1) * g.dx; }
template <class T> inline T node_y(const Ctx<T>& g, i64 j1) { return g.ymin + double(j1 - 1) * g.dy; }

// The body force at a particle, bodyforce(body, xx[ip], t)
// (SolveExplicitSolidMechanics2D.jl:408, :125): the constant g
// (BodyForce.jl:86-88), or VortexBodyForce2D (BodyForce.jl:94-121) at the
// particle's current position (the ULMPM call) or at its position when it
// was added (_X0: material coordinates, as the manufactured solution and the
// FEMP solver take them), written
// in the reference's operation order; Julia's x^2 and x^3 are the products
// x*x and x*x*x, x^4 is taken as (x*x)*(x*x) (parity unpinned for that
// rounding: the reference's ^ is Julia's power by squaring).  The vortex is
// evaluated on the values (the dual-number run refuses it).
inline double p2(double x) { return x * x; }
inline double p3(double x) { return x * x * x; }
inline double p4(double x) { return (x * x) * (x * x); }
template <class T> void body_at(const Ctx<T>& g, const Solid<T>& s, i64 ip, const T& t, T* b) {
  if (g.body_kind != JUMP_BODY_VORTEX2D && g.body_kind != JUMP_BODY_VORTEX2D_X0) {
    b[0] = g.body[0];
    b[1] = g.body[1];
    return;
  }
  const double G = g.vortex[0], Tp = g.vortex[1], Ri = g.vortex[2], Ro = g.vortex[3];
  const double rho0 = g.vortex[4], mu = g.vortex[5], xc = g.vortex[6], yc = g.vortex[7];
  const double tt = val(t);
  const T* xp = g.body_kind == JUMP_BODY_VORTEX2D_X0 ? &s.pos0[2 * ip] : &s.pos[2 * ip];
  const double x = val(xp[0]) - xc, y = val(xp[1]) - yc;
  const double r = std::sqrt(x * x + y * y);
  const double theta = std::atan2(y, x);
  const double R = (Ri + Ro) / 2;
  const double h = 1 - 8 * p2((r - R) / (Ri - Ro)) + 16 * p4((r - R) / (Ri - Ro));
  const double hp = (-16) * (r - R) / p2(Ri - Ro) + 16 * 4 * p3(r - R) / p4(Ri - Ro);
  const double hpp = -16 / p2(Ri - Ro) + 16 * 4 * 3 * p2(r - R) / p4(Ri - Ro);
  const double pi = M_PI;
  const double gg = G * std::sin(pi * tt / Tp);
  const double gp = G * pi / Tp * std::cos(pi * tt / Tp);
  const double gpp = -p2(pi) / p2(Tp) * gg;
  const double alpha = gg * h;
  const double mdr = mu / rho0;
  const double br = (mdr * (3 * gg * hp + r * gg * hpp) - r * gpp * h) * std::sin(alpha) +
                    (mdr * r * p2(gg * hp) - r * p2(gp * h)) * std::cos(alpha);
  const double bt = (-mdr * (3 * gg * hp + r * gg * hpp) + r * gpp * h) * std::cos(alpha) +
                    (mdr * r * p2(gg * hp) + r * p2(gp * h)) * std::sin(alpha);
  b[0] = br * std::cos(theta) - bt * std::sin(theta);
  b[1] = br * std::sin(theta) + bt * std::cos(theta);
}
template <class T> inline T node_z(const Ctx<T>& g, i64 k1) { return g.zmin + double(k1 - 1) * g.dz; }

// ---------------------------------------------------------------- B-splines
// Bsplines.jl:18-83 (quadratic, boundary types 1-4)
template <class T> inline void qbs1(const T& r, T& N, T& dN) {
  if (val(r) < -1.5 || val(r) > 1.5) { N = 0.0; dN = 0.0; return; }
  if (val(r) >= -1.5 && val(r) <= -0.5) { N = 0.5 * r * r + 1.5 * r + 1.125; dN = r + 1.5; }
  else if (val(r) <= 0.0) { N = 1.0 + r; dN = T(1.0); }
  else if (val(r) <= 0.5) { N = 1.0 - r; dN = T(-1.0); }
  else { N = 0.5 * r * r - 1.5 * r + 1.125; dN = r - 1.5; }
}
template <class T> inline void qbs2(const T& r, T& N, T& dN) {
  if (val(r) < -1.0 || val(r) > 1.5) { N = 0.0; dN = 0.0; return; }
  if (val(r) >= -1.0 && val(r) <= -0.5) { N = 1.0 + r; dN = T(1.0); }
  else if (val(r) <= 0.5) { N = (-r) * r + 0.75; dN = -2.0 * r; }
  else { N = 0.5 * r * r - 1.5 * r + 1.125; dN = r - 1.5; }
}
template <class T> inline void qbs3(const T& r, T& N, T& dN) {
  if (val(r) < -1.5 || val(r) > 1.5) { N = 0.0; dN = 0.0; return; }
  if (val(r) >= -1.5 && val(r) <= -0.5) { N = 0.5 * r * r + 1.5 * r + 1.125; dN = r + 1.5; }
  else if (val(r) <= 0.5) { N = (-r) * r + 0.75; dN = -2.0 * r; }
  else { N = 0.5 * r * r - 1.5 * r + 1.125; dN = r - 1.5; }
}
template <class T> inline void qbs4(const T& r, T& N, T& dN) {
  if (val(r) < -1.5 || val(r) > 1.0) { N = 0.0; dN = 0.0; return; }
  if (val(r) >= -1.5 && val(r) <= -0.5) { N = 0.5 * r * r + 1.5 * r + 1.125; dN = r + 1.5; }
  else if (val(r) <= 0.5) { N = (-r) * r + 0.75; dN = -2.0 * r; }
  else { N = 1.0 - r; dN = T(-1.0); }
}
// Bsplines.jl:89-171 (cubic, boundary types 1-4); decimal literals as written
template <class T> inline void cbs1(const T& r, T& N, T& dN) {
  if (val(r) < -2.0 || val(r) > 2.0) { N = 0.0; dN = 0.0; return; }
  T r2 = r * r;
  if (val(r) >= -2.0 && val(r) <= -1.0) {
    N = 0.16666666666666666 * r * r2 + r2 + 2.0 * r + 1.3333333333333333; dN = 0.5 * r2 + 2.0 * r + 2.0;
  } else if (-1.0 <= val(r) && val(r) <= 0.0) {
    N = -0.16666666666666666 * r * r2 + r + 1.0; dN = -0.5 * r2 + 1.0;
  } else if (0.0 <= val(r) && val(r) <= 1.0) {
    N = 0.16666666666666666 * r * r2 - r + 1.0; dN = 0.5 * r2 - 1.0;
  } else {
    N = -0.16666666666666666 * r * r2 + r2 - 2.0 * r + 1.3333333333333333; dN = -0.5 * r2 + 2.0 * r - 2.0;
  }
}
template <class T> inline void cbs2(const T& r, T& N, T& dN) {
  if (val(r) < -1.0 || val(r) > 2.0) { N = 0.0; dN = 0.0; return; }
  T r2 = r * r;
  if (val(r) >= -1.0 && val(r) <= 0.0) {
    N = -0.3333333333333333 * r * r2 - r2 + 0.6666666666666666; dN = -r2 - 2.0 * r;
  } else if (0.0 <= val(r) && val(r) <= 1.0) {
    N = 0.5 * r * r2 - r2 + 0.6666666666666666; dN = 1.5 * r2 - 2.0 * r;
  } else {
    N = -0.16666666666666666 * r * r2 + r2 - 2.0 * r + 1.3333333333333333; dN = -0.5 * r2 + 2.0 * r - 2.0;
  }
}
template <class T> inline void cbs3(const T& r, T& N, T& dN) {
  if (val(r) < -2.0 || val(r) > 2.0) { N = 0.0; dN = 0.0; return; }
  T r2 = r * r;
  if (val(r) >= -2.0 && val(r) <= -1.0) {
    N = 0.16666666666666666 * r * r2 + r2 + 2.0 * r + 1.3333333333333333; dN = 0.5 * r2 + 2.0 * r + 2.0;
  } else if (-1.0 <= val(r) && val(r) <= 0.0) {
    N = -0.5 * r * r2 - r2 + 0.6666666666666666; dN = -1.5 * r2 - 2.0 * r;
  } else if (0.0 <= val(r) && val(r) <= 1.0) {
    N = 0.5 * r * r2 - r2 + 0.6666666666666666; dN = 1.5 * r2 - 2.0 * r;
  } else {
    N = -0.16666666666666666 * r * r2 + r2 - 2.0 * r + 1.3333333333333333; dN = -0.5 * r2 + 2.0 * r - 2.0;
  }
}
template <class T> inline void cbs4(const T& r, T& N, T& dN) {
  if (val(r) < -2.0 || val(r) > 1.0) { N = 0.0; dN = 0.0; return; }
  T r2 = r * r;
  if (val(r) >= -2.0 && val(r) <= -1.0) {
    N = 0.16666666666666666 * r * r2 + r2 + 2.0 * r + 1.3333333333333333; dN = 0.5 * r2 + 2.0 * r + 2.0;
  } else if (-1.0 <= val(r) && val(r) <= 0.0) {
    N = -0.5 * r * r2 - r2 + 0.6666666666666666; dN = -1.5 * r2 - 2.0 * r;
  } else {
    N = 0.3333333333333333 * r * r2 - r2 + 0.6666666667; dN = r2 - 2.0 * r;
  }
}

// Boundary-type selection for one axis (Basis.jl:431-439).  `quirk` selects
// the getShapeFunctions variant whose test `id == 1 | id == N` Julia parses as
// `id == (1|id) == N`: true only when id == N and N is odd (Basis.jl:1079,
// 1091, 1188, 1200, 1212).
template <class T> inline void qbs_axis(i64 id, i64 n, const T& r, bool quirk, T& N, T& dN) {
  bool t1 = quirk ? (id == (1 | id) && (1 | id) == n) : (id == 1 || id == n);
  if (t1) qbs1(r, N, dN);
  else if (id == 2) qbs2(r, N, dN);
  else if (id == n - 1) qbs4(r, N, dN);
  else qbs3(r, N, dN);
}
// Basis.jl:680-687; `sel4` is the node index tested for type 4, which the 3D
// getShapeFunctions takes from the y index on the z axis (Basis.jl:1434).
template <class T> inline void cbs_axis(i64 id, i64 n, i64 sel4, const T& r, T& N, T& dN) {
  if (id == 1 || id == n) cbs1(r, N, dN);
  else if (id == 2) cbs2(r, N, dN);
  else if (sel4 == n - 1) cbs4(r, N, dN);
  else cbs3(r, N, dN);
}

// uGIMP 1D weight (Bardenhagen & Kober 2004, eq. for the uniform GIMP with
// particle half-width lp <= h/2).  Not in the reference (its basis set is
// Basis.jl:37-48); parity for this function is pinned by closed forms.
template <class T> inline void gimp_axis(const T& xi, const T& h, const T& hI, const T& lp, T& N, T& dN) {
  T a = jabs(xi);
  if (val(a) < val(lp)) {
    N = 1.0 - (xi * xi + lp * lp) / (2.0 * h * lp);
    dN = -xi / (h * lp);
  } else if (val(a) <= val(h - lp)) {
    N = 1.0 - a * hI;
    dN = -jsign(xi) * hI;
  } else if (val(a) < val(h + lp)) {
    T q = h + lp - a;
    N = q * q / (4.0 * h * lp);
    dN = -jsign(xi) * q / (2.0 * h * lp);
  } else {
    N = 0.0;
    dN = 0.0;
  }
}

// QBS / cubic / GIMP stencil on one axis, 1-based first node; false when the
// stencil leaves [1, n].
// QBS: Basis.jl:390-425 (2D; the 3D version :487-538 passes xmin = 0).
inline bool qbs_first(double num, i64 n, i64& first) {
  double whole = std::floor(num);
  double rem = num - whole;
  i64 i = (i64)std::floor(num + 1.0);
  if (i == 1) first = i;
  else if (i == n - 1) first = i - 1;
  else first = rem < 0.5 ? i - 1 : i;
  return first >= 1 && first + 2 <= n;
}
// cubic: Basis.jl:652-674
inline bool cbs_first(double num, i64 n, i64& first) {
  i64 i = (i64)std::floor(num + 1.0);
  if (i == 1) first = 1;
  else if (i == n - 1) first = i - 2;
  else first = i - 1;
  return first >= 1 && first + 3 <= n;
}
// GIMP: nodes whose support (|x - xI| < h + lp) can contain x, window of 3
inline bool gimp_first(double x, double xmin, double dxI, double lp, double xlast, i64 n, i64& first) {
  if (x - lp < xmin || x + lp > xlast) return false;
  i64 i = (i64)std::floor((x - lp - xmin) * dxI + 1.0);
  first = std::min(i, n - 2);
  return first >= 1;
}

// ---------------------------------------------------------------- support
// Returns the support size, or -1 if the stencil leaves the grid.
// Writes 0-based node indices to g.near, N to g.funcs, grad to g.ders[dim*i+k].
// `grad` false selects the getShapeFunctions variants (funcs only).

// Linear, 2D: Basis.jl:63-78 + 219-251 (funcs only: 852-878, identical N)
template <class T> int sup_linear_2d(Ctx<T>& g, const T* xp) {
  i64 i = ifloor((xp[0] - g.xmin) * g.dxI + 1.0);
  i64 j = ifloor((xp[1] - g.ymin) * g.dyI + 1.0);
  if (i < 1 || i + 1 > g.nx || j < 1 || j + 1 > g.ny) return -1;
  i64 base = (i - 1) + g.nx * (j - 1);
  i64 idx[4] = {base, base + 1, base + g.nx, base + 1 + g.nx};
  for (int a = 0; a < 4; a++) {
    i64 ii = idx[a] % g.nx + 1, jj = idx[a] / g.nx + 1;
    T xip = xp[0] - node_x(g, ii);
    T yip = xp[1] - node_y(g, jj);
    T Nx = 1.0 - jabs(xip) * g.dxI;
    T Ny = 1.0 - jabs(yip) * g.dyI;
    g.near[a] = idx[a];
    g.funcs[a] = Nx * Ny;
    g.ders[2 * a + 0] = (-Ny) * jsign(xip) * g.dxI;
    g.ders[2 * a + 1] = (-Nx) * jsign(yip) * g.dyI;
  }
  return 4;
}

// Linear, 3D: Basis.jl:104-137 + 292-332 (funcs only: 910-946)
template <class T> int sup_linear_3d(Ctx<T>& g, const T* xp) {
  i64 i = ifloor((xp[0] - g.xmin) * g.dxI + 1.0);
  i64 j = ifloor((xp[1] - g.ymin) * g.dyI + 1.0);
  i64 k = ifloor((xp[2] - g.zmin) * g.dzI + 1.0);
  if (i < 1 || i + 1 > g.nx || j < 1 || j + 1 > g.ny || k < 1 || k + 1 > g.nz) return -1;
  i64 b = (i - 1) + g.nx * (j - 1) + g.nxy * (k - 1);
  i64 idx[8] = {b, b + 1, b + g.nx, b + g.nx + 1, b + g.nxy, b + 1 + g.nxy, b + g.nx + g.nxy, b + g.nx + 1 + g.nxy};
  for (int a = 0; a < 8; a++) {
    i64 ii = idx[a] % g.nx + 1, jj = (idx[a] / g.nx) % g.ny + 1, kk = idx[a] / g.nxy + 1;
    T xip = xp[0] - node_x(g, ii);
    T yip = xp[1] - node_y(g, jj);
    T zip = xp[2] - node_z(g, kk);
    T Nx = 1.0 - jabs(xip) * g.dxI;
    T Ny = 1.0 - jabs(yip) * g.dyI;
    T Nz = 1.0 - jabs(zip) * g.dzI;
    g.near[a] = idx[a];
    g.funcs[a] = Nx * Ny * Nz;
    g.ders[3 * a + 0] = (-Ny) * Nz * jsign(xip) * g.dxI;
    g.ders[3 * a + 1] = (-Nx) * Nz * jsign(yip) * g.dyI;
    g.ders[3 * a + 2] = (-Nx) * Ny * jsign(zip) * g.dzI;
  }
  return 8;
}

// Quadratic B-spline, 2D: Basis.jl:374-463 (grad), 1023-1108 (funcs, quirk)
template <class T> int sup_qbs_2d(Ctx<T>& g, const T* xp, bool grad) {
  i64 fx, fy;
  if (!qbs_first(val((xp[0] - g.xmin) * g.dxI), g.nx, fx)) return -1;
  if (!qbs_first(val((xp[1] - g.ymin) * g.dyI), g.ny, fy)) return -1;
  int idx = 0;
  for (int a = 0; a < 3; a++) {
    i64 id = fx + a;
    T r = (xp[0] - node_x(g, id)) * g.dxI;
    T Nx, dNx;
    qbs_axis(id, g.nx, r, !grad, Nx, dNx);
    for (int b = 0; b < 3; b++) {
      i64 jd = fy + b;
      T ry = (xp[1] - double(jd - 1) * g.dy) * g.dyI;  // ymin ignored (Basis.jl:442)
      T Ny, dNy;
      qbs_axis(jd, g.ny, ry, !grad, Ny, dNy);
      g.near[idx] = g.nx * (jd - 1) + id - 1;
      g.funcs[idx] = Nx * Ny;
      if (grad) {
        g.ders[2 * idx + 0] = dNx * g.dxI * Ny;
        g.ders[2 * idx + 1] = dNy * g.dyI * Nx;
      }
      idx++;
    }
  }
  return 9;
}

// Quadratic B-spline, 3D: Basis.jl:469-589 (grad), 1114-1231 (funcs, quirk).
// The stencil ignores xmin/ymin/zmin (numx = xp*dxI, :487-497) and y/z use
// (jd-1)*dy, (kd-1)*dz (:555, :567).
template <class T> int sup_qbs_3d(Ctx<T>& g, const T* xp, bool grad) {
  i64 fx, fy, fz;
  if (!qbs_first(val(xp[0] * g.dxI), g.nx, fx)) return -1;
  if (!qbs_first(val(xp[1] * g.dyI), g.ny, fy)) return -1;
  if (!qbs_first(val(xp[2] * g.dzI), g.nz, fz)) return -1;
  int idx = 0;
  for (int a = 0; a < 3; a++) {
    i64 id = fx + a;
    T r = (xp[0] - node_x(g, id)) * g.dxI;
    T Nx, dNx;
    qbs_axis(id, g.nx, r, !grad, Nx, dNx);
    for (int b = 0; b < 3; b++) {
      i64 jd = fy + b;
      T ry = (xp[1] - double(jd - 1) * g.dy) * g.dyI;
      T Ny, dNy;
      qbs_axis(jd, g.ny, ry, !grad, Ny, dNy);
      for (int c = 0; c < 3; c++) {
        i64 kd = fz + c;
        T rz = (xp[2] - double(kd - 1) * g.dz) * g.dzI;
        T Nz, dNz;
        qbs_axis(kd, g.nz, rz, !grad, Nz, dNz);
        g.near[idx] = g.nxy * (kd - 1) + g.nx * (jd - 1) + id - 1;
        g.funcs[idx] = Nx * Ny * Nz;
        if (grad) {
          g.ders[3 * idx + 0] = dNx * g.dxI * Ny * Nz;
          g.ders[3 * idx + 1] = dNy * g.dyI * Nx * Nz;
          g.ders[3 * idx + 2] = dNz * g.dzI * Nx * Ny;
        }
        idx++;
      }
    }
  }
  return 27;
}

// Cubic B-spline, 2D: Basis.jl:636-712 (grad), 1271-1345 (funcs; no quirk)
template <class T> int sup_cbs_2d(Ctx<T>& g, const T* xp, bool grad) {
  i64 fx, fy;
  if (!cbs_first(val((xp[0] - g.xmin) * g.dxI), g.nx, fx)) return -1;
  if (!cbs_first(val((xp[1] - g.ymin) * g.dyI), g.ny, fy)) return -1;
  int idx = 0;
  for (int a = 0; a < 4; a++) {
    i64 id = fx + a;
    T r = (xp[0] - node_x(g, id)) * g.dxI;
    T Nx, dNx;
    cbs_axis(id, g.nx, id, r, Nx, dNx);
    for (int b = 0; b < 4; b++) {
      i64 jd = fy + b;
      T ry = (xp[1] - double(jd - 1) * g.dy) * g.dyI;
      T Ny, dNy;
      cbs_axis(jd, g.ny, jd, ry, Ny, dNy);
      g.near[idx] = g.nx * (jd - 1) + id - 1;
      g.funcs[idx] = Nx * Ny;
      if (grad) {
        g.ders[2 * idx + 0] = dNx * g.dxI * Ny;
        g.ders[2 * idx + 1] = dNy * g.dyI * Nx;
      }
      idx++;
    }
  }
  return 16;
}

// Cubic B-spline, 3D: Basis.jl:717-820 (grad), 1350-1450 (funcs; the z axis
// selects type 4 from jd, :1434)
template <class T> int sup_cbs_3d(Ctx<T>& g, const T* xp, bool grad) {
  i64 fx, fy, fz;
  if (!cbs_first(val((xp[0] - g.xmin) * g.dxI), g.nx, fx)) return -1;
  if (!cbs_first(val((xp[1] - g.ymin) * g.dyI), g.ny, fy)) return -1;
  if (!cbs_first(val((xp[2] - g.zmin) * g.dzI), g.nz, fz)) return -1;
  int idx = 0;
  for (int a = 0; a < 4; a++) {
    i64 id = fx + a;
    T r = (xp[0] - node_x(g, id)) * g.dxI;
    T Nx, dNx;
    cbs_axis(id, g.nx, id, r, Nx, dNx);
    for (int b = 0; b < 4; b++) {
      i64 jd = fy + b;
      T ry = (xp[1] - double(jd - 1) * g.dy) * g.dyI;
      T Ny, dNy;
      cbs_axis(jd, g.ny, jd, ry, Ny, dNy);
      for (int c = 0; c < 4; c++) {
        i64 kd = fz + c;
        T rz = (xp[2] - double(kd - 1) * g.dz) * g.dzI;
        T Nz, dNz;
        cbs_axis(kd, g.nz, grad ? kd : jd, rz, Nz, dNz);
        g.near[idx] = g.nxy * (kd - 1) + g.nx * (jd - 1) + id - 1;
        g.funcs[idx] = Nx * Ny * Nz;
        if (grad) {
          g.ders[3 * idx + 0] = dNx * g.dxI * Ny * Nz;
          g.ders[3 * idx + 1] = dNy * g.dyI * Nx * Nz;
          g.ders[3 * idx + 2] = dNz * g.dzI * Nx * Ny;
        }
        idx++;
      }
    }
  }
  return 64;
}

// uGIMP, 2D/3D tensor product over a 3-node window per axis (x outermost).
template <class T> int sup_gimp(Ctx<T>& g, const Solid<T>& s, const T* xp, bool grad) {
  const int d = g.dim;
  i64 first[3] = {1, 1, 1};
  const T mins[3] = {g.xmin, g.ymin, g.zmin};
  const T hs[3] = {g.dx, g.dy, g.dz};
  const T his[3] = {g.dxI, g.dyI, g.dzI};
  const i64 ns[3] = {g.nx, g.ny, g.nz};
  T w[3][3], dw[3][3];
  for (int k = 0; k < d; k++) {
    double xlast = val(mins[k] + double(ns[k] - 1) * hs[k]);
    if (!gimp_first(val(xp[k]), val(mins[k]), val(his[k]), val(s.lp[k]), xlast, ns[k], first[k])) return -1;
    for (int a = 0; a < 3; a++) {
      T xI = mins[k] + double(first[k] + a - 1) * hs[k];
      gimp_axis(xp[k] - xI, hs[k], his[k], s.lp[k], w[k][a], dw[k][a]);
    }
  }
  int idx = 0;
  if (d == 2) {
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        g.near[idx] = g.nx * (first[1] + b - 1) + first[0] + a - 1;
        g.funcs[idx] = w[0][a] * w[1][b];
        if (grad) {
          g.ders[2 * idx + 0] = dw[0][a] * w[1][b];
          g.ders[2 * idx + 1] = w[0][a] * dw[1][b];
        }
        idx++;
      }
    return 9;
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int c = 0; c < 3; c++) {
        g.near[idx] = g.nxy * (first[2] + c - 1) + g.nx * (first[1] + b - 1) + first[0] + a - 1;
        g.funcs[idx] = w[0][a] * w[1][b] * w[2][c];
        if (grad) {
          g.ders[3 * idx + 0] = dw[0][a] * w[1][b] * w[2][c];
          g.ders[3 * idx + 1] = w[0][a] * dw[1][b] * w[2][c];
          g.ders[3 * idx + 2] = w[0][a] * w[1][b] * dw[2][c];
        }
        idx++;
      }
  return 27;
}

// CPDI-Q4: Basis.jl:1496-1581.  The support is the union of the four corner
// hat stencils (getNearestGridPoints, :81-98); the reference iterates it in
// Julia Set order, this restatement in ascending node index (the order only
// affects the G2P summation order: parity unpinned there).
template <class T> int sup_cpdi(Ctx<T>& g, const Solid<T>& s, i64 p, bool grad) {
  const i64* e = &s.elems[4 * p];
  const T* c = s.corners.data();
  T x11 = c[2 * e[0]], x12 = c[2 * e[0] + 1];
  T x21 = c[2 * e[1]], x22 = c[2 * e[1] + 1];
  T x31 = c[2 * e[2]], x32 = c[2 * e[2] + 1];
  T x41 = c[2 * e[3]], x42 = c[2 * e[3] + 1];
  T Vp = 0.5 * (x11 * x22 - x21 * x12 + x21 * x32 - x31 * x22 + x31 * x42 - x41 * x32 + x41 * x12 - x11 * x42);
  T c1 = (x21 - x11) * (x42 - x12) - (x22 - x12) * (x41 - x11);
  T c2 = (x21 - x11) * (x32 - x22) - (x22 - x12) * (x31 - x21);
  T c3 = (x31 - x41) * (x42 - x12) - (x32 - x42) * (x41 - x11);
  T c4 = (x31 - x41) * (x32 - x22) - (x32 - x42) * (x31 - x21);
  T sf = 1.0 / (36.0 * Vp);
  T wf[4] = {sf * (4.0 * c1 + 2.0 * c2 + 2.0 * c3 + c4), sf * (2.0 * c1 + 4.0 * c2 + c3 + 2.0 * c4),
             sf * (c1 + 2.0 * c2 + 2.0 * c3 + 4.0 * c4), sf * (2.0 * c1 + c2 + 4.0 * c3 + 2.0 * c4)};
  T sg = 1.0 / (2.0 * Vp);
  T wg[2][4] = {{sg * (x22 - x42), sg * (x32 - x12), sg * (-x22 + x42), sg * (-x32 + x12)},
                {sg * (x41 - x21), sg * (x11 - x31), sg * (-x41 + x21), sg * (-x11 + x31)}};
  i64 pts[16];
  const T xs[4] = {x11, x21, x31, x41}, ys[4] = {x12, x22, x32, x42};
  for (int q = 0; q < 4; q++) {
    i64 i = ifloor((xs[q] - g.xmin) * g.dxI + 1.0);
    i64 j = ifloor((ys[q] - g.ymin) * g.dyI + 1.0);
    if (i < 1 || i + 1 > g.nx || j < 1 || j + 1 > g.ny) return -1;
    i64 b = g.nx * (j - 1) + i - 1;
    pts[4 * q + 0] = b; pts[4 * q + 1] = b + 1; pts[4 * q + 2] = b + g.nx; pts[4 * q + 3] = b + 1 + g.nx;
  }
  std::sort(pts, pts + 16);
  int cnt = int(std::unique(pts, pts + 16) - pts);
  for (int a = 0; a < cnt; a++) {
    i64 id = pts[a];
    T xI = node_x(g, id % g.nx + 1), yI = node_y(g, id / g.nx + 1);
    T Ns = 0.0, dN1 = 0.0, dN2 = 0.0;
    for (int q = 0; q < 4; q++) {
      // getShapeIp, Basis.jl:1479-1490 (clamped hat)
      T Nx = 1.0 - jabs(xs[q] - xI) * g.dxI;
      T Ny = 1.0 - jabs(ys[q] - yI) * g.dyI;
      if (val(Nx) < 0.0) Nx = 0.0;
      if (val(Ny) < 0.0) Ny = 0.0;
      T N = Nx * Ny;
      Ns += wf[q] * N;
      dN1 += wg[0][q] * N;
      dN2 += wg[1][q] * N;
    }
    g.near[a] = id;
    g.funcs[a] = Ns;
    if (grad) { g.ders[2 * a] = dN1; g.ders[2 * a + 1] = dN2; }
  }
  return cnt;
}

template <class T> int support(Ctx<T>& g, const Solid<T>& s, i64 p, bool grad) {
  const T* xp = &s.pos[g.dim * p];
  if (g.dim == 2) {
    switch (g.basis) {
      case JUMP_BASIS_LINEAR: return sup_linear_2d(g, xp);
      case JUMP_BASIS_QUAD_BSPLINE: return sup_qbs_2d(g, xp, grad);
      case JUMP_BASIS_CUBIC_BSPLINE: return sup_cbs_2d(g, xp, grad);
      case JUMP_BASIS_CPDI_Q4: return sup_cpdi(g, s, p, grad);
      case JUMP_BASIS_GIMP: return sup_gimp(g, s, xp, grad);
    }
  } else {
    switch (g.basis) {
      case JUMP_BASIS_LINEAR: return sup_linear_3d(g, xp);
      case JUMP_BASIS_QUAD_BSPLINE: return sup_qbs_3d(g, xp, grad);
      case JUMP_BASIS_CUBIC_BSPLINE: return sup_cbs_3d(g, xp, grad);
      case JUMP_BASIS_GIMP: return sup_gimp(g, s, xp, grad);
    }
  }
  return -1;
}

// ---------------------------------------------------------------- small matrices
// column-major d x d: M[a + d*b] = M[a,b]
template <class T> inline void matmul2(const T* A, const T* B, T* C) {
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) C[i + 2 * j] = A[i + 0] * B[0 + 2 * j] + A[i + 2] * B[1 + 2 * j];
}
template <class T> inline void matmul3(const T* A, const T* B, T* C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[i + 3 * j] = A[i + 0] * B[0 + 3 * j] + A[i + 3] * B[1 + 3 * j] + A[i + 6] * B[2 + 3 * j];
}
template <class T> inline T det2(const T* A) { return A[0] * A[3] - A[2] * A[1]; }
template <class T> inline T det3(const T* A) {
  // dot(col1, cross(col2, col3))
  T c0 = A[4] * A[8] - A[5] * A[7];
  T c1 = A[5] * A[6] - A[3] * A[8];
  T c2 = A[3] * A[7] - A[4] * A[6];
  return A[0] * c0 + A[1] * c1 + A[2] * c2;
}

// ---------------------------------------------------------------- stress updates
// Hooke 2D, Material.jl:68-72
template <class T> void stress_hooke_2d(const MatConst<T>& m, const T* eps, T* sig) {
  T tr = m.lambda * (eps[0] + eps[3]);
  T tm = 2.0 * m.mu;
  sig[0] = tr + tm * eps[0];
  sig[1] = tm * eps[1];
  sig[2] = tm * eps[2];
  sig[3] = tr + tm * eps[3];
}
// Hooke 3D, Material.jl:74-81 (vmStr from the incoming sigma)
template <class T> void stress_hooke_3d(const MatConst<T>& m, const T* eps, T* sig, T& vm) {
  const T* s = sig;
  vm = sqrt(s[0] * s[0] + s[4] * s[4] + s[8] * s[8] + 2.0 * (s[3] * s[3] + s[6] * s[6] + s[7] * s[7]));
  T tr = m.lambda * (eps[0] + eps[4] + eps[8]);
  T tm = 2.0 * m.mu;
  for (int i = 0; i < 9; i++) sig[i] = tm * eps[i];
  sig[0] = tr + sig[0];
  sig[4] = tr + sig[4];
  sig[8] = tr + sig[8];
}
// Neo-Hookean, Material.jl:119-122: (1/J)(mu (F F' - I) + lambda log(J) I)
template <class T> void stress_nh(int d, const MatConst<T>& m, const T* F, const T& J, T* sig) {
  T FFt[9];
  for (int i = 0; i < d; i++)
    for (int j = 0; j < d; j++) {
      T acc = F[i] * F[j];
      for (int k = 1; k < d; k++) acc = acc + F[i + d * k] * F[j + d * k];
      FFt[i + d * j] = acc;
    }
  T lnJ = m.lambda * log(J);
  T iJ = 1.0 / J;
  for (int i = 0; i < d; i++)
    for (int j = 0; j < d; j++) {
      T a = (i == j) ? FFt[i + d * j] - 1.0 : FFt[i + d * j];
      T b = m.mu * a;
      if (i == j) b = b + lnJ;
      sig[i + d * j] = iJ * b;
    }
}
// J2 plane (Voigt) update, Material.jl:174-225.  The reference method is
// typed on MMatrix and unreachable from the SMatrix solid (a MethodError); the
// stress it computes is returned here.
template <class T> void stress_j2_2d(const MatConst<T>& m, const T* eps0, T* sig, T* ep, T& alpha, T& vm) {
  T e[3] = {eps0[0], eps0[3], 2.0 * eps0[2]};
  // epsilon_dev = I_dev * e, I_dev = [.5 -.5 0; -.5 .5 0; 0 0 1]
  T ed[3] = {0.5 * e[0] + (-0.5) * e[1] + 0.0 * e[2], (-0.5) * e[0] + 0.5 * e[1] + 0.0 * e[2],
             0.0 * e[0] + 0.0 * e[1] + 1.0 * e[2]};
  T tm = 2.0 * m.mu;
  T u[3] = {tm * (ed[0] - ep[0]), tm * (ed[1] - ep[1]), tm * (ed[2] - ep[2])};
  // s_trial = I * u, I = diag(1, 1, 0.5)
  T st[3] = {1.0 * u[0] + 0.0 * u[1] + 0.0 * u[2], 0.0 * u[0] + 1.0 * u[1] + 0.0 * u[2],
             0.0 * u[0] + 0.0 * u[1] + 0.5 * u[2]};
  T nrm = sqrt(st[0] * st[0] + st[1] * st[1] + 2.0 * (st[2] * st[2]));
  T kt = m.kappa * (e[0] + e[1]);
  T strial[3] = {kt * 1.0 + st[0], kt * 1.0 + st[1], kt * 0.0 + st[2]};
  T f = nrm - (m.k1 * alpha + m.fy);
  T s[3];
  if (val(f) <= 0.0) {
    s[0] = strial[0]; s[1] = strial[1]; s[2] = strial[2];
  } else {
    T inv = 1.0 / nrm;
    T nv[3] = {inv * st[0], inv * st[1], inv * st[2]};
    T lam = (nrm - m.k1 * alpha - m.fy) / (2.0 * m.mu + m.k1);
    alpha = alpha + lam;
    // Iinv = diag(1, 1, 2)
    T in0 = 1.0 * nv[0] + 0.0 * nv[1] + 0.0 * nv[2];
    T in1 = 0.0 * nv[0] + 1.0 * nv[1] + 0.0 * nv[2];
    T in2 = 0.0 * nv[0] + 0.0 * nv[1] + 2.0 * nv[2];
    ep[0] = ep[0] + lam * in0; ep[1] = ep[1] + lam * in1; ep[2] = ep[2] + lam * in2;
    T c = tm * lam;
    s[0] = kt * 1.0 + st[0] - c * nv[0];
    s[1] = kt * 1.0 + st[1] - c * nv[1];
    s[2] = kt * 0.0 + st[2] - c * nv[2];
  }
  sig[0] = s[0]; sig[3] = s[1]; sig[2] = s[2]; sig[1] = s[2];
  T dv[3] = {0.5 * s[0] + (-0.5) * s[1] + 0.0 * s[2], (-0.5) * s[0] + 0.5 * s[1] + 0.0 * s[2],
             0.0 * s[0] + 0.0 * s[1] + 1.0 * s[2]};
  vm = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + 2.0 * (dv[2] * dv[2]));
}
// J2 3D, Material.jl:228-275 (returns sigma; see the 2D note)
template <class T> void stress_j2_3d(const MatConst<T>& m, const T* eps0, T* sig, T* ep, T& alpha, T& vm) {
  T tr = eps0[0] + eps0[4] + eps0[8];
  T c = 0.333333333 * tr;
  T ed[9];
  for (int i = 0; i < 9; i++) ed[i] = eps0[i];
  ed[0] = eps0[0] - c; ed[4] = eps0[4] - c; ed[8] = eps0[8] - c;
  T tm = 2.0 * m.mu;
  T st[9];
  for (int i = 0; i < 9; i++) st[i] = tm * (ed[i] - ep[i]);
  T nrm = sqrt(st[0] * st[0] + st[4] * st[4] + st[8] * st[8] + 2.0 * (st[3] * st[3] + st[7] * st[7] + st[6] * st[6]));
  T kt = m.kappa * tr;
  T strial[9];
  for (int i = 0; i < 9; i++) strial[i] = st[i];
  strial[0] = kt + st[0]; strial[4] = kt + st[4]; strial[8] = kt + st[8];
  T f = nrm - (m.k1 * alpha + m.fy);
  T s[9];
  if (val(f) <= 0.0) {
    for (int i = 0; i < 9; i++) s[i] = strial[i];
  } else {
    T inv = 1.0 / nrm;
    T nv[9];
    for (int i = 0; i < 9; i++) nv[i] = inv * st[i];
    T lam = (nrm - m.k1 * alpha - m.fy) / (2.0 * m.mu + m.k1);
    alpha = alpha + lam;
    for (int i = 0; i < 9; i++) ep[i] = ep[i] + lam * nv[i];
    T cl = tm * lam;
    for (int i = 0; i < 9; i++) s[i] = strial[i] - cl * nv[i];
  }
  for (int i = 0; i < 9; i++) sig[i] = s[i];
  T ts = s[0] + s[4] + s[8];
  T cs = 0.333333333 * ts;
  T d0 = s[0] - cs, d4 = s[4] - cs, d8 = s[8] - cs;
  vm = sqrt(d0 * d0 + d4 * d4 + d8 * d8 + 2.0 * (s[3] * s[3] + s[6] * s[6] + s[7] * s[7]));
}
// Johnson-Cook 3D, Material.jl:386-441
template <class T> void stress_jc(const MatConst<T>& m, const T* eps0, const T* deps, const T& dtime, T* sig,
                                  T& alpha, T& alphadot, T& vm) {
  T eff = alpha;
  T ratio = alphadot / m.epsdot0;
  if (!(val(ratio) >= 1.0)) ratio = 1.0;  // max(ratio, 1.0)
  T ys;
  if (val(eff) < 1.0e-10) ys = m.A;
  else if (val(m.C) != 0.0) ys = (m.A + m.B * pow(eff, m.n)) * pow(1.0 + ratio, m.C);
  else ys = m.A + m.B * pow(eff, m.n);
  T Gd = m.mu;
  T g2 = 2.0 * Gd;
  T trial[9];
  for (int i = 0; i < 9; i++) trial[i] = sig[i] + g2 * deps[i];
  T c = 0.333333333 * (trial[0] + trial[4] + trial[8]);
  T dv[9];
  for (int i = 0; i < 9; i++) dv[i] = trial[i];
  dv[0] = trial[0] - c; dv[4] = trial[4] - c; dv[8] = trial[8] - c;
  T J2 = 1.224744871391589 *
         sqrt(dv[0] * dv[0] + dv[4] * dv[4] + dv[8] * dv[8] + 2.0 * (dv[3] * dv[3] + dv[6] * dv[6] + dv[7] * dv[7]));
  T dp;
  if (val(J2) < val(ys)) {
    dp = 0.0;
  } else {
    dp = (J2 - ys) / (3.0 * Gd);
    T sc = ys / J2;
    for (int i = 0; i < 9; i++) dv[i] = dv[i] * sc;
  }
  alpha = eff + dp;
  T tav = 1000.0 * m.cellsize / m.velo;
  alphadot = alphadot - alphadot * dtime / tav;
  alphadot = alphadot + dp / tav;
  if (!(val(alphadot) >= 0.0)) alphadot = 0.0;  // max(0.0, alphadot)
  vm = sqrt(dv[0] * dv[0] + dv[4] * dv[4] + dv[8] * dv[8] + 2.0 * (dv[3] * dv[3] + dv[6] * dv[6] + dv[7] * dv[7]));
  T kt = m.kappa * (eps0[0] + eps0[4] + eps0[8]);
  for (int i = 0; i < 9; i++) sig[i] = dv[i];
  sig[0] = dv[0] + kt; sig[4] = dv[4] + kt; sig[8] = dv[8] + kt;
}

// ---------------------------------------------------------------- steps
template <class T> void grid_reset(Ctx<T>& g, bool musl) {
  std::fill(g.gmass.begin(), g.gmass.end(), T(0.0));
  std::fill(g.gp0.begin(), g.gp0.end(), T(0.0));
  std::fill(g.gf.begin(), g.gf.end(), T(0.0));
  if (musl) std::fill(g.gp2.begin(), g.gp2.end(), T(0.0));
}

// The particle update after a USL 2D gather (SolveExplicitSolidMechanics2D.jl:
// 506-532): new v and x, strain += dt D, F = F (I + L dt), V = det(F) V0,
// stress; L = (dN1 v1, dN1 v2, dN2 v1, dN2 v2) column-major.
template <class T> void usl2d_particle_update(Ctx<T>& g, Solid<T>& s, i64 ip, const T& v0, const T& v1, const T& x0,
                                              const T& x1, const T* L) {
  const T dt = g.dt;
  s.vel[2 * ip] = v0; s.vel[2 * ip + 1] = v1;
  s.pos[2 * ip] = x0; s.pos[2 * ip + 1] = x1;
  T D[4] = {0.5 * (L[0] + L[0]), 0.5 * (L[1] + L[2]), 0.5 * (L[2] + L[1]), 0.5 * (L[3] + L[3])};
  T* eps = &s.strain[4 * ip];
  for (int k = 0; k < 4; k++) eps[k] = eps[k] + dt * D[k];
  T A[4] = {1.0 + L[0] * dt, L[1] * dt, L[2] * dt, 1.0 + L[3] * dt};
  T* F = &s.F[4 * ip];
  T Fn[4];
  matmul2(F, A, Fn);
  for (int k = 0; k < 4; k++) F[k] = Fn[k];
  T J = det2(F);
  s.vol[ip] = J * s.vol0[ip];
  if (val(J) < 0.0) g.flags |= JUMP_FLAG_NEGATIVE_J;
  T* sg = &s.stress[4 * ip];
  switch (s.mat.kind) {
    case JUMP_MAT_ELASTIC: stress_hooke_2d(s.mat, eps, sg); break;
    case JUMP_MAT_NEOHOOKEAN: stress_nh(2, s.mat, F, J, sg); break;
    case JUMP_MAT_ELASTOPLASTIC: stress_j2_2d(s.mat, eps, sg, &s.pstrain[3 * ip], s.alpha[ip], s.vm[ip]); break;
    default: break;
  }
  if (!std::isfinite(val(x0)) || !std::isfinite(val(x1)) || !std::isfinite(val(v0)) || !std::isfinite(val(v1)))
    g.flags |= JUMP_FLAG_NONFINITE;
}

// SolveExplicitSolidMechanics2D.jl:381-572 (one iteration, no output)
template <class T> void step_usl_2d(Ctx<T>& g, const T& t) {
  const T dt = g.dt;
  grid_reset(g, false);  // :381-385
  // P2G, deformable solids (:391-423)
  for (auto& s : g.solids) {
    if (s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T fV = s.vol[ip], fM = s.mass[ip];
      const T* vp = &s.vel[2 * ip];
      const T* sg = &s.stress[4 * ip];
      T bb[2];
      body_at(g, s, ip, t, bb);
      T b0 = fM * bb[0], b1 = fM * bb[1];
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        T Ni = g.funcs[i];
        T d1 = g.ders[2 * i], d2 = g.ders[2 * i + 1];
        T Nim = Ni * fM;
        g.gmass[in] += Nim;
        g.gp0[2 * in] += Nim * vp[0];
        g.gp0[2 * in + 1] += Nim * vp[1];
        T fx = (-fV) * (sg[0] * d1 + sg[2] * d2) + b0 * Ni;
        T fy = (-fV) * (sg[1] * d1 + sg[3] * d2) + b1 * Ni;
        g.gf[2 * in] += fx;
        g.gf[2 * in + 1] += fy;
      }
    }
  }
  // grid update + Dirichlet (:429-442)
  for (i64 i = 0; i < g.nn; i++) {
    for (int k = 0; k < 2; k++) g.gp[2 * i + k] = g.gp0[2 * i + k] + g.gf[2 * i + k] * dt;
    for (int k = 0; k < 2; k++)
      if (g.fixed[2 * i + k]) { g.gp0[2 * i + k] = 0.0; g.gp[2 * i + k] = 0.0; }
  }
  // rigid solids override the grid momentum (:448-479)
  for (auto& s : g.solids) {
    if (!s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = sup_linear_2d(g, &s.pos[2 * ip]);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      for (int i = 0; i < 4; i++) {
        i64 id = g.near[i];
        T mi = g.gmass[id];
        if (s.mat.fixed) {
          g.gp[2 * id] = 0.0; g.gp[2 * id + 1] = 0.0;
          g.gp0[2 * id] = 0.0; g.gp0[2 * id + 1] = 0.0;
        } else {
          if (val(s.mat.vx) != 0.0) { g.gp[2 * id] = mi * s.mat.vx; g.gp0[2 * id] = mi * s.mat.vx; }
          if (val(s.mat.vy) != 0.0) { g.gp[2 * id + 1] = mi * s.mat.vy; g.gp0[2 * id + 1] = mi * s.mat.vy; }
        }
      }
    }
  }
  // G2P, deformable solids (:485-534)
  for (auto& s : g.solids) {
    if (s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T L[4] = {0.0, 0.0, 0.0, 0.0};
      T v0 = s.vel[2 * ip], v1 = s.vel[2 * ip + 1];
      T x0 = s.pos[2 * ip], x1 = s.pos[2 * ip + 1];
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        T Ni = g.funcs[i];
        T d1 = g.ders[2 * i], d2 = g.ders[2 * i + 1];
        T mI = g.gmass[in];
        if (val(mI) > val(g.tol)) {
          T invM = 1.0 / mI;
          T vI0 = g.gp[2 * in] * invM, vI1 = g.gp[2 * in + 1] * invM;
          v0 += Ni * (g.gp[2 * in] - g.gp0[2 * in]) * invM;
          v1 += Ni * (g.gp[2 * in + 1] - g.gp0[2 * in + 1]) * invM;
          x0 += Ni * vI0 * dt;
          x1 += Ni * vI1 * dt;
          // SMatrix{2,2}(dN1 v1, dN1 v2, dN2 v1, dN2 v2): L[a,b] = v_a dN_b
          L[0] += d1 * vI0; L[1] += d1 * vI1; L[2] += d2 * vI0; L[3] += d2 * vI1;
        }
      }
      usl2d_particle_update(g, s, ip, v0, v1, x0, x1, L);
    }
  }
  // rigid solids move (:539-551)
  for (auto& s : g.solids) {
    if (!s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      s.pos[2 * ip] = s.pos[2 * ip] + dt * s.mat.vx;
      s.pos[2 * ip + 1] = s.pos[2 * ip + 1] + dt * s.mat.vy;
    }
  }
  // CPDI corner update (:556-572) only moves a local copy: no effect.
  (void)t;
}

// APIC in the 2D USL step: the affine particle-in-cell transfers of Jiang,
// Schroeder, Selle, Teran and Stomakhin, "The affine particle-in-cell
// method", ACM TOG 34(4) 2015, on the step above.  The reference declares
// the algorithm (Algorithm.jl:41) and the per-particle affine matrix
// (Solid.jl:82, "APIC BpDp", zero-initialised at :111) but has no step for
// it.  Differences from the USL step:
//   P2G momentum0 += N m (v_p + C_p (x_I - x_p))
//   G2P over the nodes with m_I > tol, vI = p_I / m_I:
//     v_p = sum N vI (PIC: replaces the FLIP increment)
//     B = sum (N vI) (x_I - x_p)^T,  D = sum (N (x_I - x_p)) (x_I - x_p)^T over
//     the whole stencil,  C_p = B D^-1 (2x2 inverse by cofactors);
//     for the linear hat D is singular on grid lines and C_p = sum vI grad N^T
//     (= L), the multilinear identity B D^-1 = sum vI grad N^T of the paper
//   x, L, F, strain, stress: as USL, from the position before the step.
// Node coordinates x_I = min + (i-1) h (Grid.jl:149-152).  Floating-point
// order (the GPU follows it term by term):
//   d = (xI - xp, yI - yp)
//   P2G: a = (C0 d0 + C2 d1, C1 d0 + C3 d1); p0 += Nim (v + a)
//   G2P: w = N vI; B0 += w0 d0; B1 += w1 d0; B2 += w0 d1; B3 += w1 d1
//        e = N d; D0 += e0 d0; D1 += e0 d1; D3 += e1 d1
//        det = D0 D3 - D1 D1; (i0, i1, i3) = (D3, -D1, D0) / det
//        C = (B0 i0 + B2 i1, B1 i0 + B3 i1, B0 i1 + B2 i3, B1 i1 + B3 i3)
template <class T> void step_apic_2d(Ctx<T>& g, const T& t) {
  const T dt = g.dt;
  grid_reset(g, false);
  for (auto& s : g.solids) {
    if (s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T fV = s.vol[ip], fM = s.mass[ip];
      const T* vp = &s.vel[2 * ip];
      const T* sg = &s.stress[4 * ip];
      const T* C = &s.cmat[4 * ip];
      const T xp0 = s.pos[2 * ip], xp1 = s.pos[2 * ip + 1];
      T bb[2];
      body_at(g, s, ip, t, bb);
      T b0 = fM * bb[0], b1 = fM * bb[1];
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        T Ni = g.funcs[i];
        T d1 = g.ders[2 * i], d2 = g.ders[2 * i + 1];
        T e0 = node_x(g, in % g.nx + 1) - xp0;
        T e1 = node_y(g, in / g.nx + 1) - xp1;
        T a0 = C[0] * e0 + C[2] * e1, a1 = C[1] * e0 + C[3] * e1;
        T Nim = Ni * fM;
        g.gmass[in] += Nim;
        g.gp0[2 * in] += Nim * (vp[0] + a0);
        g.gp0[2 * in + 1] += Nim * (vp[1] + a1);
        T fx = (-fV) * (sg[0] * d1 + sg[2] * d2) + b0 * Ni;
        T fy = (-fV) * (sg[1] * d1 + sg[3] * d2) + b1 * Ni;
        g.gf[2 * in] += fx;
        g.gf[2 * in + 1] += fy;
      }
    }
  }
  for (i64 i = 0; i < g.nn; i++) {
    for (int k = 0; k < 2; k++) g.gp[2 * i + k] = g.gp0[2 * i + k] + g.gf[2 * i + k] * dt;
    for (int k = 0; k < 2; k++)
      if (g.fixed[2 * i + k]) { g.gp0[2 * i + k] = 0.0; g.gp[2 * i + k] = 0.0; }
  }
  for (auto& s : g.solids) {  // rigid override, as the USL step
    if (!s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = sup_linear_2d(g, &s.pos[2 * ip]);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      for (int i = 0; i < 4; i++) {
        i64 id = g.near[i];
        T mi = g.gmass[id];
        if (s.mat.fixed) {
          g.gp[2 * id] = 0.0; g.gp[2 * id + 1] = 0.0;
          g.gp0[2 * id] = 0.0; g.gp0[2 * id + 1] = 0.0;
        } else {
          if (val(s.mat.vx) != 0.0) { g.gp[2 * id] = mi * s.mat.vx; g.gp0[2 * id] = mi * s.mat.vx; }
          if (val(s.mat.vy) != 0.0) { g.gp[2 * id + 1] = mi * s.mat.vy; g.gp0[2 * id + 1] = mi * s.mat.vy; }
        }
      }
    }
  }
  for (auto& s : g.solids) {
    if (s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T L[4] = {0.0, 0.0, 0.0, 0.0}, Bm[4] = {0.0, 0.0, 0.0, 0.0}, Dm[4] = {0.0, 0.0, 0.0, 0.0};
      T v0 = 0.0, v1 = 0.0;
      const T xp0 = s.pos[2 * ip], xp1 = s.pos[2 * ip + 1];
      T x0 = xp0, x1 = xp1;
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        T Ni = g.funcs[i];
        T d1 = g.ders[2 * i], d2 = g.ders[2 * i + 1];
        T e0 = node_x(g, in % g.nx + 1) - xp0;
        T e1 = node_y(g, in / g.nx + 1) - xp1;
        T n0 = Ni * e0, n1 = Ni * e1;
        Dm[0] += n0 * e0; Dm[1] += n0 * e1; Dm[3] += n1 * e1;
        T mI = g.gmass[in];
        if (val(mI) > val(g.tol)) {
          T invM = 1.0 / mI;
          T vI0 = g.gp[2 * in] * invM, vI1 = g.gp[2 * in + 1] * invM;
          T w0 = Ni * vI0, w1 = Ni * vI1;
          v0 += w0;
          v1 += w1;
          x0 += Ni * vI0 * dt;
          x1 += Ni * vI1 * dt;
          L[0] += d1 * vI0; L[1] += d1 * vI1; L[2] += d2 * vI0; L[3] += d2 * vI1;
          Bm[0] += w0 * e0; Bm[1] += w1 * e0; Bm[2] += w0 * e1; Bm[3] += w1 * e1;
        }
      }
      T* C = &s.cmat[4 * ip];
      if (g.basis == JUMP_BASIS_LINEAR) {
        for (int k = 0; k < 4; k++) C[k] = L[k];
      } else {
        T det = Dm[0] * Dm[3] - Dm[1] * Dm[1];
        T i0 = Dm[3] / det, i1 = (-Dm[1]) / det, i3 = Dm[0] / det;
        C[0] = Bm[0] * i0 + Bm[2] * i1;
        C[1] = Bm[1] * i0 + Bm[3] * i1;
        C[2] = Bm[0] * i1 + Bm[2] * i3;
        C[3] = Bm[1] * i1 + Bm[3] * i3;
      }
      usl2d_particle_update(g, s, ip, v0, v1, x0, x1, L);
    }
  }
  for (auto& s : g.solids) {
    if (!s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      s.pos[2 * ip] = s.pos[2 * ip] + dt * s.mat.vx;
      s.pos[2 * ip + 1] = s.pos[2 * ip + 1] + dt * s.mat.vy;
    }
  }
  (void)t;
}

// Multi-body contact in the 2D USL step (optional; the reference's ULMPM
// step has none).  The algorithm of SolveExplicitFEMP2DContact.jl:311-387 on
// the MPM grid: per node the fields of at most two solids, taken in particle
// order; where both are present the common normal and the approach
// correction; G2P reads each particle's own solid's velocities.
// Floating-point order (the GPU follows it term by term):
//   slot sums as the P2G sums above, plus normal += (fM * dN)
//   p = p0 + f dt; Dirichlet zeroes p0, p; invM = 1/m; v0 = p0 invM; v = p invM
//   mS = m_a + m_b; pS = p_a + p_b; vcm = pS / mS
//   n = n_a - n_b; n /= sqrt(n.n)   (slot b uses -n)
//   dv = v - vcm; D = dv.n; C = dv_x n_y - dv_y n_x; mu' = min(fric, |C|/D)
//   v = v - D (n + (mu'/|C|) (n_y C, -(n_x C)))   (tangential 0 if C == 0)
// Rigid solids (SolveExplicitFEMP2DContact.jl:272-304, :330-358): a rigid
// particle reaches the four nodes of its hat cell (Basis.jl:63-78, the nodes
// of the ULMPM rigid override) and adds only normal += (fM * dN_hat) to its
// slot.  Where a rigid and a deformable solid meet:
//   vcm = (mS * v_r) / mS (mS: the deformable mass; v_r = 0 if fixed)
//   n = n_r / sqrt(n_r.n_r) for the rigid slot, -n for the deformable one
//   and only the deformable slot is corrected.  G2P skips rigid solids; they
//   move by x += dt v_r as in the ULMPM step.
template <class T> void step_usl_2d_contact(Ctx<T>& g, const T& t) {
  const T dt = g.dt;
  const i64 nn = g.nn;
  g.cbody.assign(size_t(2 * nn), -1);
  g.cm.assign(size_t(2 * nn), T(0.0));
  g.cp0.assign(size_t(4 * nn), T(0.0));
  g.cf.assign(size_t(4 * nn), T(0.0));
  g.cn.assign(size_t(4 * nn), T(0.0));
  g.cv0.assign(size_t(4 * nn), T(0.0));
  g.cv.assign(size_t(4 * nn), T(0.0));
  grid_reset(g, false);
  for (size_t si = 0; si < g.solids.size(); si++) {
    auto& s = g.solids[si];
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = s.rigid ? sup_linear_2d(g, &s.pos[2 * ip]) : support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      if (s.rigid) {
        const T fM = s.mass[ip];
        for (int i = 0; i < sup; i++) {
          i64 in = g.near[i];
          int k;
          if (g.cbody[2 * in] < 0 || g.cbody[2 * in] == int(si)) k = 0;
          else if (g.cbody[2 * in + 1] < 0 || g.cbody[2 * in + 1] == int(si)) k = 1;
          else { g.flags |= JUMP_FLAG_CONTACT_OVERFLOW; continue; }
          g.cbody[2 * in + k] = int(si);
          const i64 c2 = 4 * in + 2 * k;
          g.cn[c2] += fM * g.ders[2 * i];
          g.cn[c2 + 1] += fM * g.ders[2 * i + 1];
        }
        continue;
      }
      T fV = s.vol[ip], fM = s.mass[ip];
      const T* vp = &s.vel[2 * ip];
      const T* sg = &s.stress[4 * ip];
      T bb[2];
      body_at(g, s, ip, t, bb);
      T b0 = fM * bb[0], b1 = fM * bb[1];
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        int k;
        if (g.cbody[2 * in] < 0 || g.cbody[2 * in] == int(si)) k = 0;
        else if (g.cbody[2 * in + 1] < 0 || g.cbody[2 * in + 1] == int(si)) k = 1;
        else { g.flags |= JUMP_FLAG_CONTACT_OVERFLOW; continue; }
        g.cbody[2 * in + k] = int(si);
        T Ni = g.funcs[i];
        T d1 = g.ders[2 * i], d2 = g.ders[2 * i + 1];
        T Nim = Ni * fM;
        const i64 c = 2 * in + k, c2 = 4 * in + 2 * k;
        g.cm[c] += Nim;
        g.cp0[c2] += Nim * vp[0];
        g.cp0[c2 + 1] += Nim * vp[1];
        g.cf[c2] += (-fV) * (sg[0] * d1 + sg[2] * d2) + b0 * Ni;
        g.cf[c2 + 1] += (-fV) * (sg[1] * d1 + sg[3] * d2) + b1 * Ni;
        g.cn[c2] += fM * d1;
        g.cn[c2 + 1] += fM * d2;
      }
    }
  }
  for (i64 i = 0; i < nn; i++) {
    T p[2][2];
    for (int k = 0; k < 2; k++) {
      const i64 c2 = 4 * i + 2 * k;
      for (int a = 0; a < 2; a++) {
        p[k][a] = g.cp0[c2 + a] + g.cf[c2 + a] * dt;
        if (g.fixed[2 * i + a]) { g.cp0[c2 + a] = 0.0; p[k][a] = 0.0; }
      }
      const int b = g.cbody[2 * i + k];
      if (b >= 0 && g.solids[size_t(b)].rigid) {
        const auto& rm = g.solids[size_t(b)].mat;
        g.cv0[c2] = g.cv[c2] = rm.fixed ? T(0.0) : rm.vx;
        g.cv0[c2 + 1] = g.cv[c2 + 1] = rm.fixed ? T(0.0) : rm.vy;
      } else if (b >= 0) {
        T invM = 1.0 / g.cm[2 * i + k];
        for (int a = 0; a < 2; a++) {
          g.cv0[c2 + a] = g.cp0[c2 + a] * invM;
          g.cv[c2 + a] = p[k][a] * invM;
        }
      }
    }
    // totals (the grid fields a caller reads back)
    g.gmass[i] = g.cm[2 * i] + g.cm[2 * i + 1];
    for (int a = 0; a < 2; a++) {
      g.gp0[2 * i + a] = g.cp0[4 * i + a] + g.cp0[4 * i + 2 + a];
      g.gp[2 * i + a] = p[0][a] + p[1][a];
      g.gf[2 * i + a] = g.cf[4 * i + a] + g.cf[4 * i + 2 + a];
    }
    if (g.cbody[2 * i] < 0 || g.cbody[2 * i + 1] < 0) continue;
    const bool r0 = g.solids[size_t(g.cbody[2 * i])].rigid, r1 = g.solids[size_t(g.cbody[2 * i + 1])].rigid;
    if (r0 && r1) continue;  // rigid bodies are never corrected
    // contact node
    const T mS = g.cm[2 * i] + g.cm[2 * i + 1];
    T vcm[2], nc[2];
    if (r0 || r1) {
      const i64 rr = r0 ? 0 : 1;
      const T* vr = &g.cv[4 * i + 2 * rr];
      vcm[0] = (mS * vr[0]) / mS;
      vcm[1] = (mS * vr[1]) / mS;
      const T* nr = &g.cn[4 * i + 2 * rr];
      const T nrm = sqrt(nr[0] * nr[0] + nr[1] * nr[1]);
      nc[0] = r0 ? nr[0] / nrm : -(nr[0] / nrm);
      nc[1] = r0 ? nr[1] / nrm : -(nr[1] / nrm);
    } else {
      vcm[0] = (p[0][0] + p[1][0]) / mS;
      vcm[1] = (p[0][1] + p[1][1]) / mS;
      nc[0] = g.cn[4 * i] - g.cn[4 * i + 2];
      nc[1] = g.cn[4 * i + 1] - g.cn[4 * i + 3];
      const T nrm = sqrt(nc[0] * nc[0] + nc[1] * nc[1]);
      nc[0] = nc[0] / nrm;
      nc[1] = nc[1] / nrm;
    }
    for (int k = 0; k < 2; k++) {
      if ((k == 0 && r0) || (k == 1 && r1)) continue;
      const T nk[2] = {k == 0 ? nc[0] : -nc[0], k == 0 ? nc[1] : -nc[1]};
      T* v = &g.cv[4 * i + 2 * k];
      const T dv[2] = {v[0] - vcm[0], v[1] - vcm[1]};
      const T D = dv[0] * nk[0] + dv[1] * nk[1];
      if (!(val(D) > 0.0)) continue;
      const T C = dv[0] * nk[1] - dv[1] * nk[0];
      const T absC = C < 0.0 ? -C : C;
      const T mup = val(g.fric) < val(absC / D) ? g.fric : absC / D;
      T tx = 0.0, ty = 0.0;
      if (val(absC) > 0.0) {
        const T cc = mup / absC;
        tx = cc * (nk[1] * C);
        ty = cc * (-(nk[0] * C));
      }
      v[0] = v[0] - D * (nk[0] + tx);
      v[1] = v[1] - D * (nk[1] + ty);
    }
  }
  // G2P with each particle's own solid's velocities
  for (size_t si = 0; si < g.solids.size(); si++) {
    auto& s = g.solids[si];
    if (s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T L[4] = {0.0, 0.0, 0.0, 0.0};
      T v0 = s.vel[2 * ip], v1 = s.vel[2 * ip + 1];
      T x0 = s.pos[2 * ip], x1 = s.pos[2 * ip + 1];
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        int k = g.cbody[2 * in] == int(si) ? 0 : (g.cbody[2 * in + 1] == int(si) ? 1 : -1);
        if (k < 0) continue;
        T Ni = g.funcs[i];
        T d1 = g.ders[2 * i], d2 = g.ders[2 * i + 1];
        if (val(g.cm[2 * in + k]) > val(g.tol)) {
          const T* vb = &g.cv[4 * in + 2 * k];
          const T* vb0 = &g.cv0[4 * in + 2 * k];
          v0 += Ni * (vb[0] - vb0[0]);
          v1 += Ni * (vb[1] - vb0[1]);
          x0 += Ni * vb[0] * dt;
          x1 += Ni * vb[1] * dt;
          L[0] += d1 * vb[0]; L[1] += d1 * vb[1]; L[2] += d2 * vb[0]; L[3] += d2 * vb[1];
        }
      }
      usl2d_particle_update(g, s, ip, v0, v1, x0, x1, L);
    }
  }
  // rigid solids move (SolveExplicitSolidMechanics2D.jl:539-551)
  for (auto& s : g.solids) {
    if (!s.rigid) continue;
    for (i64 ip = 0; ip < s.n; ip++) {
      s.pos[2 * ip] = s.pos[2 * ip] + dt * s.mat.vx;
      s.pos[2 * ip + 1] = s.pos[2 * ip + 1] + dt * s.mat.vy;
    }
  }
  (void)t;
}

// The particle update of G2P-2 (SolveExplicitSolidMechanics3D.jl:186-211)
// from the gathered position xx and velocity gradient L[a,b] = dN_a v_b.
template <class T> void musl3d_particle_update(Ctx<T>& g, Solid<T>& s, i64 ip, const T* xx, const T* L) {
  const T dt = g.dt;
  for (int k = 0; k < 3; k++) s.pos[3 * ip + k] = xx[k];
  T D[9], dD[9];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) D[a + 3 * b] = 0.5 * (L[a + 3 * b] + L[b + 3 * a]);
  T* eps = &s.strain[9 * ip];
  for (int k = 0; k < 9; k++) { dD[k] = dt * D[k]; eps[k] = eps[k] + dD[k]; }
  T A[9];
  for (int k = 0; k < 9; k++) A[k] = L[k] * dt;
  A[0] = 1.0 + A[0]; A[4] = 1.0 + A[4]; A[8] = 1.0 + A[8];
  T* F = &s.F[9 * ip];
  T Fn[9];
  matmul3(F, A, Fn);
  for (int k = 0; k < 9; k++) F[k] = Fn[k];
  T J = det3(F);
  s.vol[ip] = J * s.vol0[ip];
  if (val(J) < 0.0) g.flags |= JUMP_FLAG_NEGATIVE_J;
  T* sg = &s.stress[9 * ip];
  switch (s.mat.kind) {
    case JUMP_MAT_ELASTIC: stress_hooke_3d(s.mat, eps, sg, s.vm[ip]); break;
    case JUMP_MAT_NEOHOOKEAN: stress_nh(3, s.mat, F, J, sg); break;
    case JUMP_MAT_ELASTOPLASTIC: stress_j2_3d(s.mat, eps, sg, &s.pstrain[9 * ip], s.alpha[ip], s.vm[ip]); break;
    case JUMP_MAT_JOHNSONCOOK: stress_jc(s.mat, eps, dD, dt, sg, s.alpha[ip], s.alphadot[ip], s.vm[ip]); break;
    default: break;
  }
  if (!std::isfinite(val(xx[0])) || !std::isfinite(val(xx[1])) || !std::isfinite(val(xx[2])) ||
      !std::isfinite(val(s.vel[3 * ip])) || !std::isfinite(val(s.vel[3 * ip + 1])) ||
      !std::isfinite(val(s.vel[3 * ip + 2])))
    g.flags |= JUMP_FLAG_NONFINITE;
}

// SolveExplicitSolidMechanics3D.jl:46-211 (one iteration, no output)
template <class T> void step_musl_3d(Ctx<T>& g) {
  const T dt = g.dt, alpha = g.alpha;
  grid_reset(g, true);  // :46-51
  // P2G (:55-84)
  for (auto& s : g.solids) {
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T fV = s.vol[ip], fM = s.mass[ip];
      const T* vp = &s.vel[3 * ip];
      const T* sg = &s.stress[9 * ip];
      for (int i = 0; i < sup; i++) {
        i64 id = g.near[i];
        T Ni = g.funcs[i];
        T dx = g.ders[3 * i], dy = g.ders[3 * i + 1], dz = g.ders[3 * i + 2];
        T Nim = Ni * fM;
        g.gmass[id] += Nim;
        for (int k = 0; k < 3; k++) g.gp0[3 * id + k] += Nim * vp[k];
        // sigma[1,1]dx + sigma[1,2]dy + sigma[1,3]dz, ... (symmetric indices)
        T f0 = sg[0] * dx + sg[3] * dy + sg[6] * dz;
        T f1 = sg[3] * dx + sg[4] * dy + sg[7] * dz;
        T f2 = sg[6] * dx + sg[7] * dy + sg[8] * dz;
        g.gf[3 * id] -= fV * f0;
        g.gf[3 * id + 1] -= fV * f1;
        g.gf[3 * id + 2] -= fV * f2;
      }
    }
  }
  // grid update + BC (:89-105)
  for (i64 i = 0; i < g.nn; i++) {
    for (int k = 0; k < 3; k++) g.gp[3 * i + k] = g.gp0[3 * i + k] + g.gf[3 * i + k] * dt;
    for (int k = 0; k < 3; k++)
      if (g.fixed[3 * i + k]) { g.gp0[3 * i + k] = 0.0; g.gp[3 * i + k] = 0.0; }
  }
  // G2P-1: particle velocity + re-projection into momentum2 (:109-140)
  for (auto& s : g.solids) {
    for (i64 ip = 0; ip < s.n; ip++) {
      T* vv = &s.vel[3 * ip];
      T mp = s.mass[ip];
      T vx[3] = {0.0, 0.0, 0.0};
      int sup = support(g, s, ip, false);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      for (int i = 0; i < sup; i++) {
        i64 id = g.near[i];
        T Ni = g.funcs[i];
        T mI = g.gmass[id];
        if (val(mI) > 0.0) {
          T mii = (1.0 / mI) * Ni;
          for (int k = 0; k < 3; k++) vx[k] += (g.gp[3 * id + k] - alpha * g.gp0[3 * id + k]) * mii;
        }
      }
      for (int k = 0; k < 3; k++) vv[k] = alpha * vv[k] + vx[k];
      for (int i = 0; i < sup; i++) {
        i64 id = g.near[i];
        T Nm = g.funcs[i] * mp;
        for (int k = 0; k < 3; k++) g.gp2[3 * id + k] += Nm * vv[k];
      }
    }
  }
  // BC again (:142-156): momentum0/momentum unchanged since the grid update
  for (i64 i = 0; i < g.nn; i++)
    for (int k = 0; k < 3; k++)
      if (g.fixed[3 * i + k]) { g.gp0[3 * i + k] = 0.0; g.gp[3 * i + k] = 0.0; }
  // G2P-2: position, deformation, stress (:161-211)
  for (auto& s : g.solids) {
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T L[9];
      for (int k = 0; k < 9; k++) L[k] = 0.0;
      T xx[3] = {s.pos[3 * ip], s.pos[3 * ip + 1], s.pos[3 * ip + 2]};
      for (int i = 0; i < sup; i++) {
        i64 id = g.near[i];
        T Ni = g.funcs[i];
        T dNx = g.ders[3 * i], dNy = g.ders[3 * i + 1], dNz = g.ders[3 * i + 2];
        T m = g.gmass[id];
        if (val(m) > 0.0) {
          T mii = 1.0 / m;
          T vIx = mii * g.gp2[3 * id], vIy = mii * g.gp2[3 * id + 1], vIz = mii * g.gp2[3 * id + 2];
          T c = Ni * mii * dt;
          for (int k = 0; k < 3; k++) xx[k] += c * g.gp[3 * id + k];
          // SMatrix{3,3}(dNx vIx, dNy vIx, dNz vIx, ...): L[a,b] = dN_a v_b
          L[0] += dNx * vIx; L[1] += dNy * vIx; L[2] += dNz * vIx;
          L[3] += dNx * vIy; L[4] += dNy * vIy; L[5] += dNz * vIy;
          L[6] += dNx * vIz; L[7] += dNy * vIz; L[8] += dNz * vIz;
        }
      }
      musl3d_particle_update(g, s, ip, xx, L);
    }
  }
}

// Dense LU with partial pivoting (the factorisation Julia's `\` applies to
// a general square matrix), in place on the n x n row-major A; then solves
// A x = b for each right-hand side column of B (n x m, row-major), in place.
// Returns false for a singular pivot.
template <class T> bool lu_solve(std::vector<T>& A, i64 n, std::vector<T>& B, int m) {
  std::vector<i64> piv(static_cast<size_t>(n));
  for (i64 k = 0; k < n; k++) {
    i64 p = k;
    for (i64 i = k + 1; i < n; i++)
      if (std::fabs(val(A[size_t(i * n + k)])) > std::fabs(val(A[size_t(p * n + k)]))) p = i;
    piv[size_t(k)] = p;
    if (val(A[size_t(p * n + k)]) == 0.0) return false;
    if (p != k)
      for (i64 j = 0; j < n; j++) std::swap(A[size_t(k * n + j)], A[size_t(p * n + j)]);
    const T inv = 1.0 / A[size_t(k * n + k)];
    for (i64 i = k + 1; i < n; i++) {
      T& l = A[size_t(i * n + k)];
      if (val(l) == 0.0) continue;
      l = l * inv;
      for (i64 j = k + 1; j < n; j++) A[size_t(i * n + j)] = A[size_t(i * n + j)] - l * A[size_t(k * n + j)];
    }
  }
  for (int c = 0; c < m; c++) {
    for (i64 k = 0; k < n; k++)
      if (piv[size_t(k)] != k) std::swap(B[size_t(k * m + c)], B[size_t(piv[size_t(k)] * m + c)]);
    for (i64 i = 0; i < n; i++)  // L y = b (unit diagonal)
      for (i64 j = 0; j < i; j++) B[size_t(i * m + c)] = B[size_t(i * m + c)] - A[size_t(i * n + j)] * B[size_t(j * m + c)];
    for (i64 i = n - 1; i >= 0; i--) {  // U x = y
      for (i64 j = i + 1; j < n; j++) B[size_t(i * m + c)] = B[size_t(i * m + c)] - A[size_t(i * n + j)] * B[size_t(j * m + c)];
      B[size_t(i * m + c)] = B[size_t(i * m + c)] / A[size_t(i * n + i)];
    }
  }
  return true;
}

// SolveExplicitSolidMechanics2D.jl:25-329 (2D MUSL with the blended mass
// matrix; one iteration, no output).  P2G adds the consistent mass matrix
// M_IJ = sum m N_I N_J (:120-125); the blended B = 0.1 diag(m) + 0.9 M, with
// 1 on the diagonal of massless nodes (:150-156), is solved for the nodal
// velocities v0 = B \ p0 and v = v0 + B \ (dt f) (:160-178).  The Dirichlet
// mask zeroes only the nodal momenta (:184-199), which the G2P passes do not
// read.  G2P-1 (getShapeFunctions, :203-226): v = alpha v + sum N (v_I -
// alpha v0_I) over nodes with mass; G2P-2 (:246-286): x += N v_I dt, L from
// v_I, then the particle update.  The CPDI corner update reads the
// re-projected momentum, which this step never fills, so corners stay put
// (:289-309).  Dense, as the reference: up to 4096 nodes.
template <class T> bool step_musl_2d(Ctx<T>& g, const T& t) {
  const T dt = g.dt, alpha = g.alpha;
  const i64 nn = g.nn;
  if (nn > 4096) return false;
  grid_reset(g, true);
  std::vector<T> M(size_t(nn * nn), T(0.0));
  for (auto& s : g.solids) {
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T fV = s.vol[ip], fM = s.mass[ip];
      const T* vp = &s.vel[2 * ip];
      const T* sg = &s.stress[4 * ip];
      T bb[2];
      body_at(g, s, ip, t, bb);
      T b0 = fM * bb[0], b1 = fM * bb[1];
      for (int i = 0; i < sup; i++) {
        i64 id = g.near[i];
        T Ni = g.funcs[i];
        T d1 = g.ders[2 * i], d2 = g.ders[2 * i + 1];
        T Nim = Ni * fM;
        g.gmass[id] += Nim;
        for (int j = 0; j < sup; j++) M[size_t(id * nn + g.near[j])] += Nim * g.funcs[j];
        g.gp0[2 * id] += Nim * vp[0];
        g.gp0[2 * id + 1] += Nim * vp[1];
        g.gf[2 * id] += (-fV) * (sg[0] * d1 + sg[2] * d2) + b0 * Ni;
        g.gf[2 * id + 1] += (-fV) * (sg[1] * d1 + sg[3] * d2) + b1 * Ni;
      }
    }
  }
  const T eps = 0.1, oneme = 1.0 - eps;
  for (i64 i = 0; i < nn; i++)
    for (i64 j = 0; j < nn; j++) {
      T& b = M[size_t(i * nn + j)];
      b = i == j ? eps * g.gmass[i] + oneme * b : oneme * b;
    }
  for (i64 i = 0; i < nn; i++)
    if (val(g.gmass[i]) == 0.0) M[size_t(i * nn + i)] = 1.0;
  // right-hand sides p0x, p0y, dt fx, dt fy (before the Dirichlet mask)
  std::vector<T> X(size_t(4 * nn));
  for (i64 i = 0; i < nn; i++) {
    X[size_t(4 * i)] = g.gp0[2 * i];
    X[size_t(4 * i + 1)] = g.gp0[2 * i + 1];
    X[size_t(4 * i + 2)] = dt * g.gf[2 * i];
    X[size_t(4 * i + 3)] = dt * g.gf[2 * i + 1];
  }
  if (!lu_solve(M, nn, X, 4)) g.flags |= JUMP_FLAG_NONFINITE;
  g.mv.assign(size_t(4 * nn), T(0.0));  // vx0, vy0, vx, vy
  for (i64 i = 0; i < nn; i++) {
    g.mv[size_t(4 * i)] = X[size_t(4 * i)];
    g.mv[size_t(4 * i + 1)] = X[size_t(4 * i + 1)];
    g.mv[size_t(4 * i + 2)] = X[size_t(4 * i)] + X[size_t(4 * i + 2)];
    g.mv[size_t(4 * i + 3)] = X[size_t(4 * i + 1)] + X[size_t(4 * i + 3)];
  }
  for (i64 i = 0; i < nn; i++) {
    for (int k = 0; k < 2; k++) g.gp[2 * i + k] = g.gp0[2 * i + k] + g.gf[2 * i + k] * dt;
    for (int k = 0; k < 2; k++)
      if (g.fixed[2 * i + k]) { g.gp0[2 * i + k] = 0.0; g.gp[2 * i + k] = 0.0; }
  }
  // G2P-1: particle velocity only
  for (auto& s : g.solids) {
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, false);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T vx = 0.0, vy = 0.0;
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        T Ni = g.funcs[i];
        if (val(g.gmass[in]) > 0.0) {
          vx += Ni * (g.mv[size_t(4 * in + 2)] - alpha * g.mv[size_t(4 * in)]);
          vy += Ni * (g.mv[size_t(4 * in + 3)] - alpha * g.mv[size_t(4 * in + 1)]);
        }
      }
      s.vel[2 * ip] = alpha * s.vel[2 * ip] + vx;
      s.vel[2 * ip + 1] = alpha * s.vel[2 * ip + 1] + vy;
    }
  }
  // G2P-2: position, deformation, stress
  for (auto& s : g.solids) {
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T L[4] = {0.0, 0.0, 0.0, 0.0};
      T x0 = s.pos[2 * ip], x1 = s.pos[2 * ip + 1];
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        T Ni = g.funcs[i];
        T d1 = g.ders[2 * i], d2 = g.ders[2 * i + 1];
        if (val(g.gmass[in]) > 0.0) {
          T vI0 = g.mv[size_t(4 * in + 2)], vI1 = g.mv[size_t(4 * in + 3)];
          x0 += Ni * vI0 * dt;
          x1 += Ni * vI1 * dt;
          L[0] += d1 * vI0; L[1] += d1 * vI1; L[2] += d2 * vI0; L[3] += d2 * vI1;
        }
      }
      usl2d_particle_update(g, s, ip, s.vel[2 * ip], s.vel[2 * ip + 1], x0, x1, L);
    }
  }
  return true;
}

// The 3D MUSL step with multi-body contact: the normal-only correction of
// the FEMP 3D contact solver for two deformable solids
// (SolveExplicitFEMP3DContact.jl:398-511: n = (n1 - n2)/|n1 - n2|, -n for
// the second; a solid whose velocity approaches the centre-of-mass velocity
// along its normal, D = (v - v_cm).n > 0, gets v -= D n), at the grid update
// of the MUSL step.  Per-solid slots as in 2D (first two solids of a node in
// particle order); normals sum m grad N over the solid's particles.  The
// FEMP solver also requires |(x1 - x2).n| < dy/40 for nodal positions
// summed over its mesh's surface particles; an MPM solid has no surface
// particles, and over all particles that test lets two spheres interpenetrate
// by more than two cells before it passes, so it is not applied.  Dirichlet
// is applied to the slot momenta before v_cm.  Then G2P-1, re-projection and
// G2P-2 of SolveExplicitSolidMechanics3D.jl:109-211, each particle reading
// its own solid's slot; the re-projected momentum is not corrected again.
template <class T> void step_musl_3d_contact(Ctx<T>& g) {
  const T dt = g.dt, alpha = g.alpha;
  const i64 nn = g.nn;
  g.cbody.assign(size_t(2 * nn), -1);
  for (auto* v : {&g.cm}) v->assign(size_t(2 * nn), T(0.0));
  for (auto* v : {&g.cp0, &g.cp, &g.cf, &g.cn, &g.cp2}) v->assign(size_t(6 * nn), T(0.0));
  grid_reset(g, true);
  auto slot_of = [&](i64 in, int si) {
    return g.cbody[2 * in] == si ? 0 : (g.cbody[2 * in + 1] == si ? 1 : -1);
  };
  // P2G per solid slot (:55-84 per slot, normals and positions added)
  for (size_t si = 0; si < g.solids.size(); si++) {
    auto& s = g.solids[si];
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T fV = s.vol[ip], fM = s.mass[ip];
      const T* vp = &s.vel[3 * ip];
      const T* sg = &s.stress[9 * ip];
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        int k;
        if (g.cbody[2 * in] < 0 || g.cbody[2 * in] == int(si)) k = 0;
        else if (g.cbody[2 * in + 1] < 0 || g.cbody[2 * in + 1] == int(si)) k = 1;
        else { g.flags |= JUMP_FLAG_CONTACT_OVERFLOW; continue; }
        g.cbody[2 * in + k] = int(si);
        const i64 c = 2 * in + k, c3 = 3 * c;
        T Ni = g.funcs[i];
        T dx = g.ders[3 * i], dy = g.ders[3 * i + 1], dz = g.ders[3 * i + 2];
        T Nim = Ni * fM;
        g.cm[c] += Nim;
        for (int a = 0; a < 3; a++) g.cp0[c3 + a] += Nim * vp[a];
        T f0 = sg[0] * dx + sg[3] * dy + sg[6] * dz;
        T f1 = sg[3] * dx + sg[4] * dy + sg[7] * dz;
        T f2 = sg[6] * dx + sg[7] * dy + sg[8] * dz;
        g.cf[c3] -= fV * f0;
        g.cf[c3 + 1] -= fV * f1;
        g.cf[c3 + 2] -= fV * f2;
        g.cn[c3] += fM * dx;
        g.cn[c3 + 1] += fM * dy;
        g.cn[c3 + 2] += fM * dz;
      }
    }
  }
  // grid update + BC per slot (:89-105), totals, contact correction
  for (i64 i = 0; i < nn; i++) {
    for (int k = 0; k < 2; k++) {
      const i64 c3 = 3 * (2 * i + k);
      for (int a = 0; a < 3; a++) {
        g.cp[c3 + a] = g.cp0[c3 + a] + g.cf[c3 + a] * dt;
        if (g.fixed[3 * i + a]) { g.cp0[c3 + a] = 0.0; g.cp[c3 + a] = 0.0; }
      }
    }
    g.gmass[i] = g.cm[2 * i] + g.cm[2 * i + 1];
    for (int a = 0; a < 3; a++) {
      g.gp0[3 * i + a] = g.cp0[6 * i + a] + g.cp0[6 * i + 3 + a];
      g.gp[3 * i + a] = g.cp[6 * i + a] + g.cp[6 * i + 3 + a];
      g.gf[3 * i + a] = g.cf[6 * i + a] + g.cf[6 * i + 3 + a];
    }
    if (g.cbody[2 * i] < 0 || g.cbody[2 * i + 1] < 0) continue;
    const T mS = g.cm[2 * i] + g.cm[2 * i + 1];
    T vcm[3], nc[3];
    for (int a = 0; a < 3; a++) vcm[a] = (g.cp[6 * i + a] + g.cp[6 * i + 3 + a]) / mS;
    for (int a = 0; a < 3; a++) nc[a] = g.cn[6 * i + a] - g.cn[6 * i + 3 + a];
    const T nrm = sqrt(nc[0] * nc[0] + nc[1] * nc[1] + nc[2] * nc[2]);
    for (int a = 0; a < 3; a++) nc[a] = nc[a] / nrm;
    for (int k = 0; k < 2; k++) {
      const i64 c3 = 3 * (2 * i + k);
      const T m = g.cm[2 * i + k];
      T v[3], nk[3];
      for (int a = 0; a < 3; a++) {
        v[a] = g.cp[c3 + a] / m;
        nk[a] = k == 0 ? nc[a] : -nc[a];
      }
      const T Dn = (v[0] - vcm[0]) * nk[0] + (v[1] - vcm[1]) * nk[1] + (v[2] - vcm[2]) * nk[2];
      if (!(val(Dn) > 0.0)) continue;
      for (int a = 0; a < 3; a++) g.cp[c3 + a] = (v[a] - Dn * nk[a]) * m;
    }
  }
  // G2P-1 and re-projection (:109-140), each particle on its own slot
  for (size_t si = 0; si < g.solids.size(); si++) {
    auto& s = g.solids[si];
    for (i64 ip = 0; ip < s.n; ip++) {
      T* vv = &s.vel[3 * ip];
      T mp = s.mass[ip];
      T vx[3] = {0.0, 0.0, 0.0};
      int sup = support(g, s, ip, false);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        int k = slot_of(in, int(si));
        if (k < 0) continue;
        const i64 c = 2 * in + k;
        T mI = g.cm[c];
        if (val(mI) > 0.0) {
          T mii = (1.0 / mI) * g.funcs[i];
          for (int a = 0; a < 3; a++) vx[a] += (g.cp[3 * c + a] - alpha * g.cp0[3 * c + a]) * mii;
        }
      }
      for (int a = 0; a < 3; a++) vv[a] = alpha * vv[a] + vx[a];
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        int k = slot_of(in, int(si));
        if (k < 0) continue;
        T Nm = g.funcs[i] * mp;
        for (int a = 0; a < 3; a++) g.cp2[3 * (2 * in + k) + a] += Nm * vv[a];
      }
    }
  }
  for (i64 i = 0; i < nn; i++)
    for (int a = 0; a < 3; a++) g.gp2[3 * i + a] = g.cp2[6 * i + a] + g.cp2[6 * i + 3 + a];
  // G2P-2 (:161-211) on the own slot
  for (size_t si = 0; si < g.solids.size(); si++) {
    auto& s = g.solids[si];
    for (i64 ip = 0; ip < s.n; ip++) {
      int sup = support(g, s, ip, true);
      if (sup < 0) { g.flags |= JUMP_FLAG_OUT_OF_GRID; continue; }
      T L[9];
      for (int k = 0; k < 9; k++) L[k] = 0.0;
      T xx[3] = {s.pos[3 * ip], s.pos[3 * ip + 1], s.pos[3 * ip + 2]};
      for (int i = 0; i < sup; i++) {
        i64 in = g.near[i];
        int k = slot_of(in, int(si));
        if (k < 0) continue;
        const i64 c = 2 * in + k;
        T Ni = g.funcs[i];
        T dNx = g.ders[3 * i], dNy = g.ders[3 * i + 1], dNz = g.ders[3 * i + 2];
        T m = g.cm[c];
        if (val(m) > 0.0) {
          T mii = 1.0 / m;
          T vIx = mii * g.cp2[3 * c], vIy = mii * g.cp2[3 * c + 1], vIz = mii * g.cp2[3 * c + 2];
          T cc = Ni * mii * dt;
          for (int a = 0; a < 3; a++) xx[a] += cc * g.cp[3 * c + a];
          L[0] += dNx * vIx; L[1] += dNy * vIx; L[2] += dNz * vIx;
          L[3] += dNx * vIy; L[4] += dNy * vIy; L[5] += dNz * vIy;
          L[6] += dNx * vIz; L[7] += dNy * vIz; L[8] += dNz * vIz;
        }
      }
      musl3d_particle_update(g, s, ip, xx, L);
    }
  }
}

// compute(::EnergiesFix), Fix.jl:248-270
template <class T> void energies(const Ctx<T>& g, T& KE, T& SE) {
  KE = 0.0; SE = 0.0;
  const int d = g.dim, dd = d * d;
  for (auto& s : g.solids) {
    for (i64 ip = 0; ip < s.n; ip++) {
      const T* v = &s.vel[d * ip];
      T vv = v[0] * v[0];
      for (int k = 1; k < d; k++) vv = vv + v[k] * v[k];
      KE += 0.5 * s.mass[ip] * vv;
      const T* sg = &s.stress[dd * ip];
      const T* ep = &s.strain[dd * ip];
      T acc = sg[0] * ep[0];
      for (int k = 1; k < dd; k++) acc = acc + sg[k] * ep[k];
      SE += 0.5 * s.vol[ip] * acc;
    }
  }
}

}  // namespace orc
