// mpm_basis.hpp -- one-dimensional shape-function kernels of the reference
// (quadratic and cubic B-splines with their boundary types, uGIMP) and the
// sign function the linear basis uses.  Pure FP64 arithmetic with no device
// state: the kernels call these, and a host build of the same header lets
// the CPU test suite check the branch-free forms against the piecewise
// definitions bit for bit.  Build with -ffp-contract=off.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define JH_HD __host__ __device__ __forceinline__
#else
#define JH_HD inline
#endif

namespace jh {

JH_HD double jsign(double d) { return d > 0.0 ? 1.0 : (d < 0.0 ? -1.0 : d); }

// ------------------------------------------------------------------ B-splines
// Bsplines.jl:18-83
JH_HD void qbs_t1(double r, double& N, double& dN) {
  if (r < -1.5 || r > 1.5) { N = 0.0; dN = 0.0; return; }
  if (r >= -1.5 && r <= -0.5) { N = 0.5 * r * r + 1.5 * r + 1.125; dN = r + 1.5; }
  else if (r <= 0.0) { N = 1.0 + r; dN = 1.0; }
  else if (r <= 0.5) { N = 1.0 - r; dN = -1.0; }
  else { N = 0.5 * r * r - 1.5 * r + 1.125; dN = r - 1.5; }
}
JH_HD void qbs_t2(double r, double& N, double& dN) {
  if (r < -1.0 || r > 1.5) { N = 0.0; dN = 0.0; return; }
  if (r >= -1.0 && r <= -0.5) { N = 1.0 + r; dN = 1.0; }
  else if (r <= 0.5) { N = (-r) * r + 0.75; dN = -2.0 * r; }
  else { N = 0.5 * r * r - 1.5 * r + 1.125; dN = r - 1.5; }
}
JH_HD void qbs_t3(double r, double& N, double& dN) {
  if (r < -1.5 || r > 1.5) { N = 0.0; dN = 0.0; return; }
  if (r >= -1.5 && r <= -0.5) { N = 0.5 * r * r + 1.5 * r + 1.125; dN = r + 1.5; }
  else if (r <= 0.5) { N = (-r) * r + 0.75; dN = -2.0 * r; }
  else { N = 0.5 * r * r - 1.5 * r + 1.125; dN = r - 1.5; }
}
JH_HD void qbs_t4(double r, double& N, double& dN) {
  if (r < -1.5 || r > 1.0) { N = 0.0; dN = 0.0; return; }
  if (r >= -1.5 && r <= -0.5) { N = 0.5 * r * r + 1.5 * r + 1.125; dN = r + 1.5; }
  else if (r <= 0.5) { N = (-r) * r + 0.75; dN = -2.0 * r; }
  else { N = 1.0 - r; dN = -1.0; }
}
// Bsplines.jl:89-171
JH_HD void cbs_t1(double r, double& N, double& dN) {
  if (r < -2.0 || r > 2.0) { N = 0.0; dN = 0.0; return; }
  double r2 = r * r;
  if (r >= -2.0 && r <= -1.0) { N = 0.16666666666666666 * r * r2 + r2 + 2.0 * r + 1.3333333333333333; dN = 0.5 * r2 + 2.0 * r + 2.0; }
  else if (-1.0 <= r && r <= 0.0) { N = -0.16666666666666666 * r * r2 + r + 1.0; dN = -0.5 * r2 + 1.0; }
  else if (0.0 <= r && r <= 1.0) { N = 0.16666666666666666 * r * r2 - r + 1.0; dN = 0.5 * r2 - 1.0; }
  else { N = -0.16666666666666666 * r * r2 + r2 - 2.0 * r + 1.3333333333333333; dN = -0.5 * r2 + 2.0 * r - 2.0; }
}
JH_HD void cbs_t2(double r, double& N, double& dN) {
  if (r < -1.0 || r > 2.0) { N = 0.0; dN = 0.0; return; }
  double r2 = r * r;
  if (r >= -1.0 && r <= 0.0) { N = -0.3333333333333333 * r * r2 - r2 + 0.6666666666666666; dN = -r2 - 2.0 * r; }
  else if (0.0 <= r && r <= 1.0) { N = 0.5 * r * r2 - r2 + 0.6666666666666666; dN = 1.5 * r2 - 2.0 * r; }
  else { N = -0.16666666666666666 * r * r2 + r2 - 2.0 * r + 1.3333333333333333; dN = -0.5 * r2 + 2.0 * r - 2.0; }
}
JH_HD void cbs_t3(double r, double& N, double& dN) {
  if (r < -2.0 || r > 2.0) { N = 0.0; dN = 0.0; return; }
  double r2 = r * r;
  if (r >= -2.0 && r <= -1.0) { N = 0.16666666666666666 * r * r2 + r2 + 2.0 * r + 1.3333333333333333; dN = 0.5 * r2 + 2.0 * r + 2.0; }
  else if (-1.0 <= r && r <= 0.0) { N = -0.5 * r * r2 - r2 + 0.6666666666666666; dN = -1.5 * r2 - 2.0 * r; }
  else if (0.0 <= r && r <= 1.0) { N = 0.5 * r * r2 - r2 + 0.6666666666666666; dN = 1.5 * r2 - 2.0 * r; }
  else { N = -0.16666666666666666 * r * r2 + r2 - 2.0 * r + 1.3333333333333333; dN = -0.5 * r2 + 2.0 * r - 2.0; }
}
JH_HD void cbs_t4(double r, double& N, double& dN) {
  if (r < -2.0 || r > 1.0) { N = 0.0; dN = 0.0; return; }
  double r2 = r * r;
  if (r >= -2.0 && r <= -1.0) { N = 0.16666666666666666 * r * r2 + r2 + 2.0 * r + 1.3333333333333333; dN = 0.5 * r2 + 2.0 * r + 2.0; }
  else if (-1.0 <= r && r <= 0.0) { N = -0.5 * r * r2 - r2 + 0.6666666666666666; dN = -1.5 * r2 - 2.0 * r; }
  else { N = 0.3333333333333333 * r * r2 - r2 + 0.6666666667; dN = r2 - 2.0 * r; }
}

// Boundary type of a node on one axis (1-based id, n nodes).  `quirk`: the
// getShapeFunctions test `id == 1 | id == n`, which Julia evaluates as
// `id == (1|id) == n` (Basis.jl:1079,1188).
JH_HD int qbs_type(int id, int n, bool quirk) {
  bool t1 = quirk ? (id == (1 | id) && (1 | id) == n) : (id == 1 || id == n);
  return t1 ? 1 : (id == 2 ? 2 : (id == n - 1 ? 4 : 3));
}
JH_HD void qbs_eval_branchy(int type, double r, double& N, double& dN) {
  switch (type) {
    case 1: qbs_t1(r, N, dN); break;
    case 2: qbs_t2(r, N, dN); break;
    case 4: qbs_t4(r, N, dN); break;
    default: qbs_t3(r, N, dN); break;
  }
}
// The same four functions without branches: every piece is formed with the
// expression (and rounding) of qbs_t1..qbs_t4 and the comparison chain picks
// one, so the result is bitwise that of qbs_eval_branchy, NaN included.  The
// pieces:  A = 0.5r^2 + 1.5r + 1.125   B = -r^2 + 0.75   C = 0.5r^2 - 1.5r + 1.125
//          L = 1 + r                   R = 1 - r
// type 1: [-1.5,-0.5] A, (-0.5,0] L, (0,0.5] R, else C     support |r| <= 1.5
// type 2: [-1,-0.5] L, (-0.5,0.5] B, else C                support [-1, 1.5]
// type 3: [-1.5,-0.5] A, (-0.5,0.5] B, else C              support |r| <= 1.5
// type 4: [-1.5,-0.5] A, (-0.5,0.5] B, else R              support [-1.5, 1]
// Type 3 alone (interior nodes: every node but the first two and last two
// of an axis): the pieces A, B, C with the expressions above, so bitwise the
// general form for type 3.  A and C share one form, (q + s) + 1.125 with
// s = 1.5r for A and -(1.5r) for C (q - 1.5r is q + (-(1.5r)) exactly), and
// r + 1.5 or r + (-1.5); the support test is |r| > 1.5, and below -1.5 the
// zero overrides A, so A's own test is r <= -0.5.  NaN fails every
// comparison and takes C, as the piecewise form does.  23 instead of 29
// instructions per axis on gfx950.
JH_HD void qbs_eval3(double r, double& N, double& dN) {
  const double q = 0.5 * r * r;
  const bool zero = fabs(r) > 1.5;
  const bool left = r <= -0.5;
  const bool midB = !left && r <= 0.5;
  const double s15 = 1.5 * r;
  const double s = left ? s15 : -s15;
  const double pAC = q + s + 1.125, dAC = r + (left ? 1.5 : -1.5);
  const double pB = (-r) * r + 0.75, dB = -2.0 * r;
  const double n = midB ? pB : pAC, d = midB ? dB : dAC;
  N = zero ? 0.0 : n;
  dN = zero ? 0.0 : d;
}
JH_HD void qbs_eval(int type, double r, double& N, double& dN) {
  const double q = 0.5 * r * r;
  const double s = 1.5 * r;
  const double pA = q + s + 1.125, dA = r + 1.5;
  const double pC = q - s + 1.125, dC = r - 1.5;
  const double pB = (-r) * r + 0.75, dB = -2.0 * r;
  const double pL = 1.0 + r, pR = 1.0 - r;
  const double lo = type == 2 ? -1.0 : -1.5;
  const double hi = type == 4 ? 1.0 : 1.5;
  const bool zero = r < lo || r > hi;
  const bool left = r >= lo && r <= -0.5;
  const bool mid = r <= 0.5;
  const bool midL = r <= 0.0;
  // right piece, middle piece, left piece for this type
  double nr = type == 4 ? pR : pC, dr = type == 4 ? -1.0 : dC;
  double nm = type == 1 ? (midL ? pL : pR) : pB, dm = type == 1 ? (midL ? 1.0 : -1.0) : dB;
  double nl = type == 2 ? pL : pA, dl = type == 2 ? 1.0 : dA;
  double n = mid ? nm : nr, d = mid ? dm : dr;
  n = left ? nl : n;
  d = left ? dl : d;
  N = zero ? 0.0 : n;
  dN = zero ? 0.0 : d;
}
// qbs_eval with the interior pieces alone where the node is interior.  The
// P2G gathers use it: a wave's lanes are consecutive nodes of one grid row,
// so the test on the type is uniform except next to the grid's ends (impact
// force pass 2.12 -> 1.98 ms).  The G2P stencil walks keep qbs_eval: there a
// lane's nine nodes per axis mix types and the branch cost 3-4%.
JH_HD void qbs_eval_split(int type, double r, double& N, double& dN) {
  if (type == 3) qbs_eval3(r, N, dN);
  else qbs_eval(type, r, N, dN);
}
// `sel4`: index compared with n-1 for type 4 (the 3D getShapeFunctions z axis
// compares the y index, Basis.jl:1434).
JH_HD int cbs_type(int id, int n, int sel4) {
  return (id == 1 || id == n) ? 1 : (id == 2 ? 2 : (sel4 == n - 1 ? 4 : 3));
}
JH_HD void cbs_eval(int type, double r, double& N, double& dN) {
  switch (type) {
    case 1: cbs_t1(r, N, dN); break;
    case 2: cbs_t2(r, N, dN); break;
    case 4: cbs_t4(r, N, dN); break;
    default: cbs_t3(r, N, dN); break;
  }
}
// uGIMP 1D weight and derivative (Bardenhagen & Kober 2004), lp <= h/2.
JH_HD void gimp_eval(double xi, double h, double hI, double lp, double& N, double& dN) {
  double a = fabs(xi);
  if (a < lp) {
    N = 1.0 - (xi * xi + lp * lp) / (2.0 * h * lp);
    dN = -xi / (h * lp);
  } else if (a <= h - lp) {
    N = 1.0 - a * hI;
    dN = -jsign(xi) * hI;
  } else if (a < h + lp) {
    double q = h + lp - a;
    N = q * q / (4.0 * h * lp);
    dN = -jsign(xi) * q / (2.0 * h * lp);
  } else {
    N = 0.0;
    dN = 0.0;
  }
}

}  // namespace jh
