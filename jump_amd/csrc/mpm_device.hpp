// mpm_device.hpp -- device-side data layout, shape functions and constitutive
// updates of the explicit MPM step on gfx950.
//
// Particle state is structure-of-arrays in HBM (one FP64 array per component)
// so that per-particle kernels read and write with unit stride.  Symmetric
// tensors (strain, stress, plastic strain) keep 3 (2D) or 6 (3D) components:
// the reference's D = (L+L')/2, eps, sigma and F*F' are bitwise symmetric, so
// the compression is exact.  Component order:
//   2D: 0=11 1=22 2=12          3D: 0=11 1=22 2=33 3=12 4=13 5=23
// Deformation gradient F keeps d*d components, index a + d*b = F[a,b].
//
// Arithmetic follows the reference's operation order term by term (Julia
// evaluates a+b+c as (a+b)+c and never contracts to FMA; the library is built
// with -ffp-contract=off), so per-particle results equal the CPU restatement
// bit for bit given the same grid values.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/jump_mpm.h"
#include "mpm_basis.hpp"

namespace jh {

struct GridDesc {
  int dim;
  int apic;  // the step's transfers are APIC (2D): P2G adds m N C_p (x_I - x_p), G2P forms C_p
  int bodyfield;  // the body force is per particle (Particles::bf), not the constant body
  int nx, ny, nz, nxy, nn;
  double xmin, ymin, zmin;
  double dx, dy, dz, dxI, dyI, dzI;
  double body[3];
};

// Per-solid constants (material, rigid prescription, GIMP half-widths).
// inst: the batch instance the solid belongs to (0 outside batches);
// rigid_rank: order of a rigid solid among the rigid solids of its instance.
struct SolidDev {
  int kind, rigid, fixed, inst;
  int rigid_rank, pad;
  double lambda, mu, kappa, fy, k1;
  double A, B, C, n, epsdot0, cellsize, velo;
  double vx, vy;
  double lp[3];
};

// One field of an interleaved pair array: element i is p[2*i].  The fields
// the P2G gather reads (x, m, V, v, sigma, in that order) are stored two per
// 16-byte slot, so a lane fetches two of them with one load instruction.
// On gfx950 a wave's load instruction costs the vector memory pipeline about
// the same whether each lane reads 8 or 16 bytes, and the gather is bound by
// that instruction count.  Byte offsets are 32-bit: i < 2^28.
struct SPtr {
  double* p;
  __device__ __forceinline__ double& operator[](int i) const {
    return *reinterpret_cast<double*>(reinterpret_cast<char*>(p) + unsigned(i) * 16u);
  }
};
// Slot of gather field f (0..D-1 x, D m, D+1 V, D+2..2D+1 v, then sigma):
// pair f/2, half f%2.  3D: (x,y)(z,m)(V,vx)(vy,vz)(s11,s22)(s33,s12)(s13,s23);
// 2D: (x,y)(m,V)(vx,vy)(s11,s22)(s12,-).
template <int D> struct PairLayout {
  static constexpr int NF = D == 3 ? 14 : 9;
  static constexpr int NP = (NF + 1) / 2;
};

struct Particles {
  SPtr x[3];
  SPtr v[3];
  SPtr m;
  SPtr V;
  double* V0;
  double* F[9];
  double* eps[6];
  SPtr sig[6];
  double* alpha;
  double* alphadot;
  double* vm;
  double* ep[6];
  int* sid;      // solid index of each particle
  int* elem[4];  // CPDI: global corner index of each domain corner
  double2* pr[7];  // the pair arrays x, m, V, v and sigma live in
  double* C[4];    // APIC: affine matrix C_p, column-major (2D)
  double* bf[2];   // per-particle body force of the step (2D, a position-dependent body force)
  double* X0[2];   // positions when added (a body force in material coordinates)
  int n;
};

struct GridArrays {
  double* m;
  double* p0[3];
  double* p[3];
  double* p2[3];
  double* f[3];
  const uint8_t* fixed;  // bit k set: direction k fixed
  // 3D MUSL, for the G2P passes: the node factors the reference forms
  // without the particle's weight, minv = 1/m (0 where m <= 0) and
  // q = p - alpha p0 (SolveExplicitSolidMechanics3D.jl:124-126, 176-180),
  // and the momenta, in pairs of their own arrays (16 bytes per node, each
  // array contiguous): G2P-1 loads two pairs per node, G2P-2 four.
  //   mq = (minv, q0), qq = (q1, q2), pxy = (px, py)    by the force pass
  //   p2xy = (p2x, p2y), pz2 = (pz, p2z)                 by the re-projection
  double2* mq;
  double2* qq;
  double2* pxy;
  double2* p2xy;
  double2* pz2;
};

// Multi-body contact: each node keeps up to two solids ("slots", filled in
// particle order).  body[k][n]: solid index in slot k (-1: empty); m: the
// slot's mass; v0/v: its velocity before / after the force update, v after
// the contact correction.  [slot][component][node].
struct ContactArrays {
  int* body[2];
  double* m[2];
  double* v0[2][2];
  double* v[2][2];
};
// 3D (MUSL): per slot the momenta p0, p (after the force update, Dirichlet
// and the contact correction) and the re-projected p2.  [slot][comp][node].
struct ContactArrays3 {
  int* body[2];
  double* m[2];
  double* p0[2][3];
  double* p[2][3];
  double* p2[2][3];
};

// A batch steps several independent 2D problems with the same launches.
// Instance i has its own grid, whose nodes (and bins) occupy
// [node_off[i], node_off[i+1]) of the concatenated node arrays, and its own
// dt and USL tolerance.  Particles find their instance through their solid,
// nodes through node_off.
struct InstDev {
  GridDesc G;
  double dt, tol;
};
struct BatchDev {
  const InstDev* inst;
  const int* node_off;  // ninst + 1 entries
  int ninst;
  int NN;  // total nodes
};
__device__ __forceinline__ int inst_of_node(const BatchDev& Bt, int n) {
  int lo = 0, hi = Bt.ninst - 1;  // largest i with node_off[i] <= n
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (Bt.node_off[mid] <= n) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
__device__ __forceinline__ GridArrays grid_view(GridArrays g, int off) {
  g.m += off;
  for (int k = 0; k < 3; k++) {
    if (g.p0[k]) g.p0[k] += off;
    if (g.p[k]) g.p[k] += off;
    if (g.p2[k]) g.p2[k] += off;
    if (g.f[k]) g.f[k] += off;
  }
  g.fixed += off;
  return g;
}

// ------------------------------------------------------------------ helpers
// Element i of an array whose base is wave-uniform (a kernel argument): the
// address is the base plus a 32-bit byte offset, which gfx950 global loads
// take as SGPR base + VGPR offset, so a gather of many fields at one index
// costs one shift instead of a 64-bit add per field.  Requires i * sizeof(T)
// < 2^32 (the library caps a context at kMaxParticles particles, so that
// also 16-byte elements stay in range).
constexpr long long kMaxParticles = 1ll << 28;
// Node-indexed int arrays (bin starts, work order, active-node scan) hold
// NN + 1 entries read through ldu, so (NN + 1) * 4 < 2^32: a context or batch
// has at most kMaxNodes nodes.
constexpr long long kMaxNodes = (1ll << 30) - 2;
template <class T> __device__ __forceinline__ T ldu(const T* __restrict__ base, unsigned i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * unsigned(sizeof(T)));
}

// Index of symmetric component (a,b) in the compressed storage.
template <int D> __device__ __forceinline__ constexpr int sym(int a, int b) {
  if (D == 2) return a == b ? a : 2;
  return a == b ? a : (a + b == 1 ? 3 : (a + b == 2 ? 4 : 5));
}

// ------------------------------------------------------------------ stencils
// Width of the tensor-product stencil per axis.
template <int B> struct Width { static constexpr int W = 2; };
template <> struct Width<JUMP_BASIS_QUAD_BSPLINE> { static constexpr int W = 3; };
template <> struct Width<JUMP_BASIS_CUBIC_BSPLINE> { static constexpr int W = 4; };
template <> struct Width<JUMP_BASIS_GIMP> { static constexpr int W = 3; };

// Axis description passed to the per-axis routines.
struct Axis {
  double mn, h, hI;
  int n;
};
__device__ __forceinline__ Axis axis_of(const GridDesc& G, int k) {
  Axis a;
  if (k == 0) { a.mn = G.xmin; a.h = G.dx; a.hI = G.dxI; a.n = G.nx; }
  else if (k == 1) { a.mn = G.ymin; a.h = G.dy; a.hI = G.dyI; a.n = G.ny; }
  else { a.mn = G.zmin; a.h = G.dz; a.hI = G.dzI; a.n = G.nz; }
  return a;
}

// First node (1-based) of the stencil on one axis; false when any stencil
// node leaves [1, n].  Computed in FP64 before any integer conversion so
// that out-of-range or NaN coordinates cannot overflow.
//   linear : floor((x-min)/h + 1)                       Basis.jl:67-68
//   QBS    : Basis.jl:390-425 (2D), 487-538 (3D: min taken as 0)
//   cubic  : Basis.jl:652-674
//   GIMP   : nodes within h+lp of x (3-node window)
template <int B, int D>
__device__ __forceinline__ bool first_node(double x, int k, const Axis& A, double lp, int& first) {
  if (B == JUMP_BASIS_LINEAR) {
    double i = floor((x - A.mn) * A.hI + 1.0);
    if (!(i >= 1.0 && i + 1.0 <= double(A.n))) return false;
    first = int(i);
    return true;
  } else if (B == JUMP_BASIS_QUAD_BSPLINE) {
    double num = (D == 3) ? x * A.hI : (x - A.mn) * A.hI;
    double whole = floor(num);
    double rem = num - whole;
    double i = floor(num + 1.0);
    double f;
    if (i == 1.0) f = i;
    else if (i == double(A.n - 1)) f = i - 1.0;
    else f = rem < 0.5 ? i - 1.0 : i;
    if (!(f >= 1.0 && f + 2.0 <= double(A.n))) return false;
    first = int(f);
    return true;
  } else if (B == JUMP_BASIS_CUBIC_BSPLINE) {
    double num = (x - A.mn) * A.hI;
    double i = floor(num + 1.0);
    double f;
    if (i == 1.0) f = 1.0;
    else if (i == double(A.n - 1)) f = i - 2.0;
    else f = i - 1.0;
    if (!(f >= 1.0 && f + 3.0 <= double(A.n))) return false;
    first = int(f);
    return true;
  } else {  // GIMP
    double xlast = A.mn + double(A.n - 1) * A.h;
    if (!(x - lp >= A.mn && x + lp <= xlast)) return false;
    double i = floor((x - lp - A.mn) * A.hI + 1.0);
    double f = i < double(A.n - 2) ? i : double(A.n - 2);
    if (!(f >= 1.0)) return false;
    first = int(f);
    return true;
  }
}

// Weight and derivative factor of node `id` (1-based) on axis k.
// For the linear basis `dw` is sign(x - xI) (the reference's gradient is
// -N_other*sign*hI); for the B-splines it is dN/dr; for GIMP dN/dx.
//   r on the x axis uses the node position (with xmin); on the y/z axes the
//   B-spline bases use (id-1)*h without the origin (Basis.jl:442,555,567).
// INTERIOR: the caller knows the node is interior on this axis (type 3).
template <int B, bool QUIRK, bool SPLIT = false, bool INTERIOR = false>
__device__ __forceinline__ void axis_weight(double x, int k, const Axis& A, double lp, int id, int sel4, double& w,
                                            double& dw) {
  if (B == JUMP_BASIS_LINEAR) {
    double d = x - (A.mn + double(id - 1) * A.h);
    w = 1.0 - fabs(d) * A.hI;
    dw = jsign(d);
  } else if (B == JUMP_BASIS_QUAD_BSPLINE) {
    double r = (k == 0) ? (x - (A.mn + double(id - 1) * A.h)) * A.hI : (x - double(id - 1) * A.h) * A.hI;
    if (INTERIOR) qbs_eval3(r, w, dw);
    else if (SPLIT) qbs_eval_split(qbs_type(id, A.n, QUIRK), r, w, dw);
    else qbs_eval(qbs_type(id, A.n, QUIRK), r, w, dw);
  } else if (B == JUMP_BASIS_CUBIC_BSPLINE) {
    double r = (k == 0) ? (x - (A.mn + double(id - 1) * A.h)) * A.hI : (x - double(id - 1) * A.h) * A.hI;
    cbs_eval(cbs_type(id, A.n, sel4), r, w, dw);
  } else {
    double xi = x - (A.mn + double(id - 1) * A.h);
    gimp_eval(xi, A.h, A.hI, lp, w, dw);
  }
}

// Shape function and gradient of a tensor-product node from the per-axis
// factors, in each basis's own operation order.
//   linear 2D  Basis.jl:246-248     linear 3D  Basis.jl:322-326
//   B-spline   Basis.jl:454-456, 578-581
template <int B>
__device__ __forceinline__ void combine2(const Axis& X, const Axis& Y, double wx, double dwx, double wy, double dwy,
                                         double& N, double* dN) {
  N = wx * wy;
  if (B == JUMP_BASIS_LINEAR) {
    dN[0] = (-wy) * dwx * X.hI;
    dN[1] = (-wx) * dwy * Y.hI;
  } else if (B == JUMP_BASIS_GIMP) {
    dN[0] = dwx * wy;
    dN[1] = wx * dwy;
  } else {
    dN[0] = dwx * X.hI * wy;
    dN[1] = dwy * Y.hI * wx;
  }
}
template <int B>
__device__ __forceinline__ void combine3(const Axis& X, const Axis& Y, const Axis& Z, double wx, double dwx, double wy,
                                         double dwy, double wz, double dwz, double& N, double* dN) {
  N = wx * wy * wz;
  if (B == JUMP_BASIS_LINEAR) {
    dN[0] = (-wy) * wz * dwx * X.hI;
    dN[1] = (-wx) * wz * dwy * Y.hI;
    dN[2] = (-wx) * wy * dwz * Z.hI;
  } else if (B == JUMP_BASIS_GIMP) {
    dN[0] = dwx * wy * wz;
    dN[1] = wx * dwy * wz;
    dN[2] = wx * wy * dwz;
  } else {
    dN[0] = dwx * X.hI * wy * wz;
    dN[1] = dwy * Y.hI * wx * wz;
    dN[2] = dwz * Z.hI * wx * wy;
  }
}

// ------------------------------------------------------------------ CPDI-Q4
// Basis.jl:1496-1581: weights of the four (static) domain corners and the
// 1-based hat cell of each corner.
struct CpdiW {
  double wf[4], wg0[4], wg1[4];
  double cx[4], cy[4];
  int ci[4], cj[4];
};
__device__ __forceinline__ bool cpdi_setup(const double* corners, int e0, int e1, int e2, int e3, const GridDesc& G,
                                           CpdiW& c) {
  double x11 = corners[2 * e0], x12 = corners[2 * e0 + 1];
  double x21 = corners[2 * e1], x22 = corners[2 * e1 + 1];
  double x31 = corners[2 * e2], x32 = corners[2 * e2 + 1];
  double x41 = corners[2 * e3], x42 = corners[2 * e3 + 1];
  double Vp = 0.5 * (x11 * x22 - x21 * x12 + x21 * x32 - x31 * x22 + x31 * x42 - x41 * x32 + x41 * x12 - x11 * x42);
  double c1 = (x21 - x11) * (x42 - x12) - (x22 - x12) * (x41 - x11);
  double c2 = (x21 - x11) * (x32 - x22) - (x22 - x12) * (x31 - x21);
  double c3 = (x31 - x41) * (x42 - x12) - (x32 - x42) * (x41 - x11);
  double c4 = (x31 - x41) * (x32 - x22) - (x32 - x42) * (x31 - x21);
  double sf = 1.0 / (36.0 * Vp);
  c.wf[0] = sf * (4.0 * c1 + 2.0 * c2 + 2.0 * c3 + c4);
  c.wf[1] = sf * (2.0 * c1 + 4.0 * c2 + c3 + 2.0 * c4);
  c.wf[2] = sf * (c1 + 2.0 * c2 + 2.0 * c3 + 4.0 * c4);
  c.wf[3] = sf * (2.0 * c1 + c2 + 4.0 * c3 + 2.0 * c4);
  double sg = 1.0 / (2.0 * Vp);
  c.wg0[0] = sg * (x22 - x42); c.wg0[1] = sg * (x32 - x12); c.wg0[2] = sg * (-x22 + x42); c.wg0[3] = sg * (-x32 + x12);
  c.wg1[0] = sg * (x41 - x21); c.wg1[1] = sg * (x11 - x31); c.wg1[2] = sg * (-x41 + x21); c.wg1[3] = sg * (-x11 + x31);
  c.cx[0] = x11; c.cx[1] = x21; c.cx[2] = x31; c.cx[3] = x41;
  c.cy[0] = x12; c.cy[1] = x22; c.cy[2] = x32; c.cy[3] = x42;
  for (int q = 0; q < 4; q++) {
    double i = floor((c.cx[q] - G.xmin) * G.dxI + 1.0);
    double j = floor((c.cy[q] - G.ymin) * G.dyI + 1.0);
    if (!(i >= 1.0 && i + 1.0 <= double(G.nx) && j >= 1.0 && j + 1.0 <= double(G.ny))) return false;
    c.ci[q] = int(i);
    c.cj[q] = int(j);
  }
  return true;
}
// Is node (ii, jj) (1-based) in the union of the corner hat stencils?
__device__ __forceinline__ bool cpdi_member(const CpdiW& c, int ii, int jj) {
  bool in = false;
  for (int q = 0; q < 4; q++) in = in || (ii >= c.ci[q] && ii <= c.ci[q] + 1 && jj >= c.cj[q] && jj <= c.cj[q] + 1);
  return in;
}
// phi_I and gradient at node (ii, jj): sum over corners of the clamped hat
// (getShapeIp, Basis.jl:1479-1490; accumulation :1556-1579).
__device__ __forceinline__ void cpdi_node(const CpdiW& c, const GridDesc& G, int ii, int jj, double& N, double* dN) {
  double xI = G.xmin + double(ii - 1) * G.dx;
  double yI = G.ymin + double(jj - 1) * G.dy;
  double Ns = 0.0, d1 = 0.0, d2 = 0.0;
  for (int q = 0; q < 4; q++) {
    double Nx = 1.0 - fabs(c.cx[q] - xI) * G.dxI;
    double Ny = 1.0 - fabs(c.cy[q] - yI) * G.dyI;
    if (Nx < 0.0) Nx = 0.0;
    if (Ny < 0.0) Ny = 0.0;
    double Nq = Nx * Ny;
    Ns += c.wf[q] * Nq;
    d1 += c.wg0[q] * Nq;
    d2 += c.wg1[q] * Nq;
  }
  N = Ns;
  dN[0] = d1;
  dN[1] = d2;
}

// ------------------------------------------------------------------ materials
// Hooke 2D, Material.jl:68-72
__device__ __forceinline__ void hooke2(const SolidDev& M, const double* e, double* s) {
  double tr = M.lambda * (e[0] + e[1]);
  double tm = 2.0 * M.mu;
  s[0] = tr + tm * e[0];
  s[1] = tr + tm * e[1];
  s[2] = tm * e[2];
}
// Hooke 3D, Material.jl:74-81; vm from the incoming stress
__device__ __forceinline__ void hooke3(const SolidDev& M, const double* e, double* s, double& vm) {
  vm = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
  double tr = M.lambda * (e[0] + e[1] + e[2]);
  double tm = 2.0 * M.mu;
  s[0] = tr + tm * e[0];
  s[1] = tr + tm * e[1];
  s[2] = tr + tm * e[2];
  s[3] = tm * e[3];
  s[4] = tm * e[4];
  s[5] = tm * e[5];
}
// Neo-Hookean, Material.jl:119-122
template <int D>
__device__ __forceinline__ void neohooke(const SolidDev& M, const double* F, double J, double* s) {
  double lnJ = M.lambda * log(J);
  double iJ = 1.0 / J;
  for (int a = 0; a < D; a++)
    for (int b = a; b < D; b++) {
      double acc = F[a] * F[b];
      for (int k = 1; k < D; k++) acc = acc + F[a + D * k] * F[b + D * k];
      double t = (a == b) ? M.mu * (acc - 1.0) + lnJ : M.mu * acc;
      s[sym<D>(a, b)] = iJ * t;
    }
}
// J2 plane Voigt update, Material.jl:174-225 (returns the computed stress;
// ep holds [ep11, ep22, engineering ep12]).
__device__ __forceinline__ void j2_2d(const SolidDev& M, const double* eps, double* s, double* ep, double& alpha,
                                      double& vm) {
  double e0 = eps[0], e1 = eps[1], e2 = 2.0 * eps[2];
  double ed0 = 0.5 * e0 + (-0.5) * e1 + 0.0 * e2;
  double ed1 = (-0.5) * e0 + 0.5 * e1 + 0.0 * e2;
  double ed2 = 0.0 * e0 + 0.0 * e1 + 1.0 * e2;
  double tm = 2.0 * M.mu;
  double u0 = tm * (ed0 - ep[0]), u1 = tm * (ed1 - ep[1]), u2 = tm * (ed2 - ep[2]);
  double st0 = 1.0 * u0 + 0.0 * u1 + 0.0 * u2;
  double st1 = 0.0 * u0 + 1.0 * u1 + 0.0 * u2;
  double st2 = 0.0 * u0 + 0.0 * u1 + 0.5 * u2;
  double nrm = sqrt(st0 * st0 + st1 * st1 + 2.0 * (st2 * st2));
  double kt = M.kappa * (e0 + e1);
  double f = nrm - (M.k1 * alpha + M.fy);
  double r0, r1, r2;
  if (f <= 0.0) {
    r0 = kt * 1.0 + st0; r1 = kt * 1.0 + st1; r2 = kt * 0.0 + st2;
  } else {
    double inv = 1.0 / nrm;
    double n0 = inv * st0, n1 = inv * st1, n2 = inv * st2;
    double lam = (nrm - M.k1 * alpha - M.fy) / (2.0 * M.mu + M.k1);
    alpha = alpha + lam;
    double i0 = 1.0 * n0 + 0.0 * n1 + 0.0 * n2;
    double i1 = 0.0 * n0 + 1.0 * n1 + 0.0 * n2;
    double i2 = 0.0 * n0 + 0.0 * n1 + 2.0 * n2;
    ep[0] = ep[0] + lam * i0; ep[1] = ep[1] + lam * i1; ep[2] = ep[2] + lam * i2;
    double c = tm * lam;
    r0 = kt * 1.0 + st0 - c * n0;
    r1 = kt * 1.0 + st1 - c * n1;
    r2 = kt * 0.0 + st2 - c * n2;
  }
  s[0] = r0; s[1] = r1; s[2] = r2;
  double d0 = 0.5 * r0 + (-0.5) * r1 + 0.0 * r2;
  double d1 = (-0.5) * r0 + 0.5 * r1 + 0.0 * r2;
  double d2 = 0.0 * r0 + 0.0 * r1 + 1.0 * r2;
  vm = sqrt(d0 * d0 + d1 * d1 + 2.0 * (d2 * d2));
}
// J2 3D, Material.jl:228-275 (returns sigma)
// Returns whether the return map changed ep and alpha (the plastic branch).
__device__ __forceinline__ bool j2_3d(const SolidDev& M, const double* e, double* s, double* ep, double& alpha,
                                      double& vm) {
  double tr = e[0] + e[1] + e[2];
  double c = 0.333333333 * tr;
  double ed[6] = {e[0] - c, e[1] - c, e[2] - c, e[3], e[4], e[5]};
  double tm = 2.0 * M.mu;
  double st[6];
  for (int i = 0; i < 6; i++) st[i] = tm * (ed[i] - ep[i]);
  // s12, s23, s13 in the reference's order
  double nrm = sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2] + 2.0 * (st[3] * st[3] + st[5] * st[5] + st[4] * st[4]));
  double kt = M.kappa * tr;
  double f = nrm - (M.k1 * alpha + M.fy);
  double r[6];
  if (f <= 0.0) {
    r[0] = kt + st[0]; r[1] = kt + st[1]; r[2] = kt + st[2];
    r[3] = st[3]; r[4] = st[4]; r[5] = st[5];
  } else {
    double inv = 1.0 / nrm;
    double nv[6];
    for (int i = 0; i < 6; i++) nv[i] = inv * st[i];
    double lam = (nrm - M.k1 * alpha - M.fy) / (2.0 * M.mu + M.k1);
    alpha = alpha + lam;
    for (int i = 0; i < 6; i++) ep[i] = ep[i] + lam * nv[i];
    double cl = tm * lam;
    r[0] = (kt + st[0]) - cl * nv[0];
    r[1] = (kt + st[1]) - cl * nv[1];
    r[2] = (kt + st[2]) - cl * nv[2];
    r[3] = st[3] - cl * nv[3];
    r[4] = st[4] - cl * nv[4];
    r[5] = st[5] - cl * nv[5];
  }
  for (int i = 0; i < 6; i++) s[i] = r[i];
  double cs = 0.333333333 * (r[0] + r[1] + r[2]);
  double d0 = r[0] - cs, d1 = r[1] - cs, d2 = r[2] - cs;
  vm = sqrt(d0 * d0 + d1 * d1 + d2 * d2 + 2.0 * (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]));
  return !(f <= 0.0);
}
// Johnson-Cook, Material.jl:386-441
__device__ __forceinline__ void johnson_cook(const SolidDev& M, const double* e, const double* de, double dtime,
                                             double* s, double& alpha, double& alphadot, double& vm) {
  double eff = alpha;
  double ratio = alphadot / M.epsdot0;
  if (!(ratio >= 1.0)) ratio = 1.0;
  double ys;
  if (eff < 1.0e-10) ys = M.A;
  else if (M.C != 0.0) ys = (M.A + M.B * pow(eff, M.n)) * pow(1.0 + ratio, M.C);
  else ys = M.A + M.B * pow(eff, M.n);
  double g2 = 2.0 * M.mu;
  double t[6];
  for (int i = 0; i < 6; i++) t[i] = s[i] + g2 * de[i];
  double c = 0.333333333 * (t[0] + t[1] + t[2]);
  double d[6] = {t[0] - c, t[1] - c, t[2] - c, t[3], t[4], t[5]};
  double J2 = 1.224744871391589 *
              sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + 2.0 * (d[3] * d[3] + d[4] * d[4] + d[5] * d[5]));
  double dp;
  if (J2 < ys) {
    dp = 0.0;
  } else {
    dp = (J2 - ys) / (3.0 * M.mu);
    double sc = ys / J2;
    for (int i = 0; i < 6; i++) d[i] = d[i] * sc;
  }
  alpha = eff + dp;
  double tav = 1000.0 * M.cellsize / M.velo;
  alphadot = alphadot - alphadot * dtime / tav;
  alphadot = alphadot + dp / tav;
  if (!(alphadot >= 0.0)) alphadot = 0.0;
  vm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + 2.0 * (d[3] * d[3] + d[4] * d[4] + d[5] * d[5]));
  double kt = M.kappa * (e[0] + e[1] + e[2]);
  s[0] = d[0] + kt; s[1] = d[1] + kt; s[2] = d[2] + kt;
  s[3] = d[3]; s[4] = d[4]; s[5] = d[5];
}

}  // namespace jh
