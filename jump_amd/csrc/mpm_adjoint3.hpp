// mpm_adjoint3.hpp -- reverse mode of the 3D MUSL step (linear and GIMP
// bases; Hooke, Neo-Hookean and J2 solids): given the adjoints of the
// particle state after a step, the adjoints of the state before it and of
// the material constants lambda, mu, fy (yield stress) and k1 (hardening).
//
// Forward step (SolveExplicitSolidMechanics3D.jl:55-211), particle p, node
// I, N = N_I(x_p), g = grad N_I(x_p) (for these bases getShapeFunctions
// and getShapeAndGradient give the same N):
//   m_I = sum N m_p     P0_I = sum N m_p v_p     f_I = -sum V_p sigma_p g
//   P_I = P0_I + dt f_I                                  (Dirichlet: 0)
//   G2P-1, nodes with m_I > 0:  v'_p = alpha v_p + sum N (P_I - alpha P0_I)/m_I
//   re-projection:              P2_I = sum N m_p v'_p
//   G2P-2, nodes with m_I > 0:  x'_p = x_p + sum N dt P_I/m_I
//                               L[a,b] = sum g_a P2_Ib/m_I (the reference's
//                               transposed gradient, column-major a + 3b)
//   eps' = eps + dt sym(L)   F' = F (I + dt L)   J = det F'   V' = J V0
//   sigma' = Hooke(eps') | (mu (F'F'^T - I) + lambda ln J I)/J
//          | J2 return map of (eps', plastic strain, alpha)
// The adjoint runs the parts backwards, one kernel each:
//   k_adj3_g2p2   per particle: the stress update, V', F', eps' and the
//                 G2P-2 sums (x picks up dN/dx and the Hessian of N); stores
//                 the particle's sources of the node adjoints (x'_b, L_b)
//   k_adj3_nodes2 per node: P_b, P2_b, m_b of the G2P-2 sums
//   k_adj3_g2p1   per particle: the re-projection and G2P-1 sums (v'_b
//                 total, v_b = alpha v'_b); stores v'_b as the source
//   k_adj3_nodes1 per node: P_b, P0_b, m_b of the G2P-1 sums
//   k_adj3_grid   per node: P = P0 + dt f with the Dirichlet mask
//   k_adj3_p2g    per particle: the P2G sums gathered from the nodes
// Nodes without mass and the Dirichlet components are constants of the
// step, as in the forward.  The J2 branch (elastic or plastic) is taken as
// in the forward; the derivative is that of the branch taken.  Every sum has
// a fixed order, so gradients are bitwise reproducible.
#pragma once

#include "mpm_adjoint.hpp"

namespace jh {

struct AdjParticles3 {
  double* xb[3];
  double* vb[3];
  double* Fb[9];
  double* eb[6];
  double* sb[6];
  double* Vb;
  double* ab;      // alpha (equivalent plastic strain)
  double* epb[6];  // plastic strain
  double* lam;     // accumulated adjoints of lambda, mu, fy, k1, per particle
  double* mu;
  double* fy;
  double* k1;
};
struct AdjNodes3 {
  double* mb;
  double* Pb[3];
  double* P0b[3];
  double* P2b[3];
};
// Per-particle sources of the node adjoints, gathered per node, in 16-byte
// pairs that the node gathers load whole: xL holds x'_b then L_b (G2P-2),
// (x0, x1) (x2, L0) (L1, L2) (L3, L4) (L5, L6) (L7, L8); vb holds v'_b, total
// (G2P-1), as (v0, v1) (v2, 0).
struct AdjSources3 {
  double2* xL[6];
  double2* vb[2];
  __device__ __forceinline__ void put_xL(int i, const double* xb, const double* Lb) const {
    const double v[12] = {xb[0], xb[1], xb[2], Lb[0], Lb[1], Lb[2], Lb[3], Lb[4], Lb[5], Lb[6], Lb[7], Lb[8]};
#pragma unroll
    for (int j = 0; j < 6; j++) xL[j][i] = make_double2(v[2 * j], v[2 * j + 1]);
  }
  __device__ __forceinline__ void get_xL(int i, double* xb, double* Lb) const {
    double v[12];
#pragma unroll
    for (int j = 0; j < 6; j++) {
      const double2 w = xL[j][i];
      v[2 * j] = w.x;
      v[2 * j + 1] = w.y;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) xb[k] = v[k];
#pragma unroll
    for (int k = 0; k < 9; k++) Lb[k] = v[3 + k];
  }
  __device__ __forceinline__ void put_vb(int i, const double* v) const {
    vb[0][i] = make_double2(v[0], v[1]);
    vb[1][i] = make_double2(v[2], 0.0);
  }
};

// A particle's 3D stencil formed once (as Stencil2): each() calls
// fn(node, N, g[3], H[6] = {xx, xy, xz, yy, yz, zz}) in walk_stencil_h3's
// order.
template <int B> struct Stencil3 {
  static constexpr int W = Width<B>::W;
  int f[3];
  double w[3][W], d1[3][W], d2[3][W];
  __device__ __forceinline__ bool init(const Particles& P, const GridDesc& G, const SolidDev& s, int p) {
    const Axis A[3] = {axis_of(G, 0), axis_of(G, 1), axis_of(G, 2)};
    for (int a = 0; a < 3; a++)
      if (!first_node<B, 3>(P.x[a][p], a, A[a], s.lp[a], f[a])) return false;
    for (int a = 0; a < 3; a++)
      for (int o = 0; o < W; o++) axis_w2<B>(P.x[a][p], A[a], s.lp[a], f[a] + o, w[a][o], d1[a][o], d2[a][o]);
    return true;
  }
  template <class Fn> __device__ __forceinline__ void each(const GridDesc& G, Fn&& fn) const {
    for (int c = 0; c < W; c++)
      for (int b = 0; b < W; b++)
        for (int a = 0; a < W; a++) {
          const double N = w[0][a] * w[1][b] * w[2][c];
          const double g[3] = {d1[0][a] * w[1][b] * w[2][c], w[0][a] * d1[1][b] * w[2][c],
                               w[0][a] * w[1][b] * d1[2][c]};
          const double H[6] = {d2[0][a] * w[1][b] * w[2][c], d1[0][a] * d1[1][b] * w[2][c],
                               d1[0][a] * w[1][b] * d1[2][c], w[0][a] * d2[1][b] * w[2][c],
                               w[0][a] * d1[1][b] * d1[2][c], w[0][a] * w[1][b] * d2[2][c]};
          fn((f[0] + a - 1) + G.nx * ((f[1] + b - 1) + G.ny * (f[2] + c - 1)), N, g, H);
        }
  }
};

// Per node, the factors the 3D G2P adjoints read, formed as the forward's
// G2P passes form them (k_g2p1_musl, k_g2p2_musl: 1/m, then products with
// it), so that the velocity gradient k_adj3_g2p2 recomputes is the
// forward's and the J2 branch it differentiates is the forward's:
// mii = 1/m (0 without mass), q = p - alpha p0, qm = q mii, v = p mii,
// v2 = p2 mii (k_adj3_nodevel).
struct NodeVel3 {
  double* mii;
  double* q[3];
  double* qm[3];
  double* v[3];
  double* v2[3];
};
__global__ void __launch_bounds__(kBlock) k_adj3_nodevel(int nn, GridArrays g, double alpha, NodeVel3 nv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const double m = g.m[n];
  const double mii = m > 0.0 ? 1.0 / m : 0.0;
  nv.mii[n] = mii;
  if (!(m > 0.0)) return;
  for (int k = 0; k < 3; k++) {
    const double q = g.p[k][n] - alpha * g.p0[k][n];
    nv.q[k][n] = q;
    nv.qm[k][n] = q * mii;
    nv.v[k][n] = g.p[k][n] * mii;
    nv.v2[k][n] = g.p2[k][n] * mii;
  }
}

// Calls fn(node, N, g[3], H[6] = {xx, xy, xz, yy, yz, zz}) for the support of p.
template <int B, class Fn>
__device__ __forceinline__ bool walk_stencil_h3(const Particles& P, const GridDesc& G, const SolidDev& s, int p,
                                                Fn&& fn) {
  constexpr int W = Width<B>::W;
  const Axis A[3] = {axis_of(G, 0), axis_of(G, 1), axis_of(G, 2)};
  int f[3];
  for (int a = 0; a < 3; a++)
    if (!first_node<B, 3>(P.x[a][p], a, A[a], s.lp[a], f[a])) return false;
  double w[3][W], d1[3][W], d2[3][W];
  for (int a = 0; a < 3; a++)
    for (int o = 0; o < W; o++) axis_w2<B>(P.x[a][p], A[a], s.lp[a], f[a] + o, w[a][o], d1[a][o], d2[a][o]);
  for (int c = 0; c < W; c++)
    for (int b = 0; b < W; b++)
      for (int a = 0; a < W; a++) {
        const double N = w[0][a] * w[1][b] * w[2][c];
        const double g[3] = {d1[0][a] * w[1][b] * w[2][c], w[0][a] * d1[1][b] * w[2][c], w[0][a] * w[1][b] * d1[2][c]};
        const double H[6] = {d2[0][a] * w[1][b] * w[2][c], d1[0][a] * d1[1][b] * w[2][c], d1[0][a] * w[1][b] * d1[2][c],
                             w[0][a] * d2[1][b] * w[2][c], w[0][a] * d1[1][b] * d1[2][c], w[0][a] * w[1][b] * d2[2][c]};
        fn((f[0] + a - 1) + G.nx * ((f[1] + b - 1) + G.ny * (f[2] + c - 1)), N, g, H);
      }
  return true;
}
// H[6] as a 3x3 lookup
__device__ __forceinline__ double hess(const double* H, int a, int b) {
  const int idx[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
  return H[idx[3 * a + b]];
}

// Reverse of the J2 3D return map (mpm_device.hpp j2_3d; Material.jl:228-275)
// at the new strain e (6), plastic strain ep (6) and alpha before the update.
// rb: adjoint of the returned stress; ab_out, epb_out: adjoints of alpha'
// and ep'.  Adds the strain adjoint to eb and returns the adjoints of alpha
// and ep before the update and of lambda, mu, fy, k1.
__device__ __forceinline__ void j2_3d_reverse(const SolidDev& M, const double* e, const double* ep, double alpha,
                                              const double* rb, double ab_out, const double* epb_out, double* eb,
                                              double& ab_in, double* epb_in, double& lamb, double& mub, double& fyb,
                                              double& k1b) {
  const double tr = e[0] + e[1] + e[2];
  const double c = 0.333333333 * tr;
  const double ed[6] = {e[0] - c, e[1] - c, e[2] - c, e[3], e[4], e[5]};
  const double tm = 2.0 * M.mu;
  double st[6];
  for (int i = 0; i < 6; i++) st[i] = tm * (ed[i] - ep[i]);
  const double nrm = sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2] + 2.0 * (st[3] * st[3] + st[5] * st[5] + st[4] * st[4]));
  const double f = nrm - (M.k1 * alpha + M.fy);
  double stb[6], tmb = 0.0, nrmb = 0.0;
  const double ktb = rb[0] + rb[1] + rb[2];
  for (int i = 0; i < 6; i++) {
    stb[i] = rb[i];
    epb_in[i] = epb_out[i];
  }
  ab_in = ab_out;
  if (f > 0.0) {
    const double inv = 1.0 / nrm;
    const double den = 2.0 * M.mu + M.k1;
    const double lam = (nrm - M.k1 * alpha - M.fy) / den;
    double nv[6], nvb[6], lb = ab_out, invb = 0.0;
    for (int i = 0; i < 6; i++) {
      nv[i] = inv * st[i];
      // r = kt I + st - tm lam nv;  ep' = ep + lam nv;  alpha' = alpha + lam
      nvb[i] = -tm * lam * rb[i] + lam * epb_out[i];
      lb += -tm * nv[i] * rb[i] + nv[i] * epb_out[i];
      tmb += -lam * nv[i] * rb[i];
      stb[i] += inv * nvb[i];
      invb += st[i] * nvb[i];
    }
    nrmb += -invb * inv * inv;
    // lam = (nrm - k1 alpha - fy) / (2 mu + k1)
    nrmb += lb / den;
    ab_in += -M.k1 * lb / den;
    fyb += -lb / den;
    k1b += (-alpha - lam) * lb / den;
    mub += -2.0 * lam * lb / den;
  }
  if (nrmb != 0.0)
    for (int i = 0; i < 6; i++) stb[i] += nrmb * (i < 3 ? 1.0 : 2.0) * st[i] / nrm;
  double edb[6];
  for (int i = 0; i < 6; i++) {
    tmb += stb[i] * (ed[i] - ep[i]);
    edb[i] = tm * stb[i];
    epb_in[i] -= tm * stb[i];
  }
  mub += 2.0 * tmb;
  // kt = (lambda + mu) tr
  const double kap = M.lambda + M.mu;
  double trb = kap * ktb;
  lamb += tr * ktb;
  mub += tr * ktb;
  // ed = e - c I, c = 0.333333333 tr
  for (int i = 0; i < 6; i++) eb[i] += edb[i];
  trb += 0.333333333 * -(edb[0] + edb[1] + edb[2]);
  for (int i = 0; i < 3; i++) eb[i] += trb;
}

template <int B>
__global__ void __launch_bounds__(kBlock) k_adj3_g2p2(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                      GridArrays g, double dt, AdjParticles3 ad, AdjSources3 src,
                                                      NodeVel3 nv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  // forward recomputation of L (column-major a + 3b), eps', F', J: the
  // forward's L[a,b] = g_a v2_b with v2 = p2 (1/m)
  double L[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  Stencil3<B> stc;
  const bool ok = stc.init(P, G, s, i);
  if (ok)
    stc.each(G, [&](int n, double, const double* gr, const double*) {
      if (nv.mii[n] > 0.0) {
        const double u[3] = {nv.v2[0][n], nv.v2[1][n], nv.v2[2][n]};
        for (int b = 0; b < 3; b++)
          for (int a = 0; a < 3; a++) L[a + 3 * b] += gr[a] * u[b];
      }
    });
  if (!ok) {  // the forward left this particle unchanged: identity adjoint, no sources
    const double z[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    src.put_xL(i, z, z);
    src.put_vb(i, z);
    return;
  }
  const int sa[6] = {0, 1, 2, 0, 0, 1}, sbb[6] = {0, 1, 2, 1, 2, 2};
  double e[6];
  for (int c = 0; c < 6; c++) e[c] = P.eps[c][i] + dt * (0.5 * (L[sa[c] + 3 * sbb[c]] + L[sbb[c] + 3 * sa[c]]));
  double A[9], F[9], Fn[9];
  for (int k = 0; k < 9; k++) {
    A[k] = dt * L[k] + ((k % 4) == 0 ? 1.0 : 0.0);
    F[k] = P.F[k][i];
  }
  for (int b = 0; b < 3; b++)
    for (int a = 0; a < 3; a++) {
      double acc = 0.0;
      for (int k = 0; k < 3; k++) acc += F[a + 3 * k] * A[k + 3 * b];
      Fn[a + 3 * b] = acc;
    }
  // cofactors of F' (dJ/dF'[a,b]) and J
  double cof[9];
  cof[0] = Fn[4] * Fn[8] - Fn[7] * Fn[5];
  cof[3] = -(Fn[1] * Fn[8] - Fn[7] * Fn[2]);
  cof[6] = Fn[1] * Fn[5] - Fn[4] * Fn[2];
  cof[1] = -(Fn[3] * Fn[8] - Fn[6] * Fn[5]);
  cof[4] = Fn[0] * Fn[8] - Fn[6] * Fn[2];
  cof[7] = -(Fn[0] * Fn[5] - Fn[3] * Fn[2]);
  cof[2] = Fn[3] * Fn[7] - Fn[6] * Fn[4];
  cof[5] = -(Fn[0] * Fn[7] - Fn[6] * Fn[1]);
  cof[8] = Fn[0] * Fn[4] - Fn[3] * Fn[1];
  const double J = Fn[0] * cof[0] + Fn[3] * cof[3] + Fn[6] * cof[6];
  // incoming adjoints of the state after the step
  double sgb[6], eb[6], Fnb[9];
  for (int c = 0; c < 6; c++) {
    sgb[c] = ad.sb[c][i];
    eb[c] = ad.eb[c][i];
  }
  for (int k = 0; k < 9; k++) Fnb[k] = ad.Fb[k][i];
  double Jb = ad.Vb[i] * P.V0[i];
  double lamb = 0.0, mub = 0.0, fyb = 0.0, k1b = 0.0;
  // alpha and the plastic strain pass through unchanged unless J2 updates them
  double ab_in = ad.ab[i], epb_in[6];
  for (int c = 0; c < 6; c++) epb_in[c] = ad.epb[c][i];
  if (s.kind == JUMP_MAT_ELASTIC) {
    const double tr = e[0] + e[1] + e[2], trb = sgb[0] + sgb[1] + sgb[2];
    for (int c = 0; c < 6; c++) eb[c] += (c < 3 ? s.lambda * trb : 0.0) + 2.0 * s.mu * sgb[c];
    lamb += tr * trb;
    for (int c = 0; c < 6; c++) mub += 2.0 * e[c] * sgb[c];
  } else if (s.kind == JUMP_MAT_NEOHOOKEAN) {
    double Bm[6];
    for (int c = 0; c < 6; c++) {
      double acc = 0.0;
      for (int k = 0; k < 3; k++) acc += Fn[sa[c] + 3 * k] * Fn[sbb[c] + 3 * k];
      Bm[c] = acc;
    }
    const double lnJ = log(J), iJ = 1.0 / J;
    double Tb[6], Tsum = 0.0;
    for (int c = 0; c < 6; c++) {
      const double T = c < 3 ? s.mu * (Bm[c] - 1.0) + s.lambda * lnJ : s.mu * Bm[c];
      Tsum += T * sgb[c];
      Tb[c] = iJ * sgb[c];
    }
    const double Tbd = Tb[0] + Tb[1] + Tb[2];
    Jb += -Tsum * iJ * iJ + s.lambda * Tbd * iJ;
    lamb += lnJ * Tbd;
    for (int c = 0; c < 6; c++) {
      mub += (c < 3 ? Bm[c] - 1.0 : Bm[c]) * Tb[c];
      const double Bb = s.mu * Tb[c];
      for (int k = 0; k < 3; k++) {
        if (c < 3) {
          Fnb[sa[c] + 3 * k] += 2.0 * Bb * Fn[sa[c] + 3 * k];
        } else {
          Fnb[sa[c] + 3 * k] += Bb * Fn[sbb[c] + 3 * k];
          Fnb[sbb[c] + 3 * k] += Bb * Fn[sa[c] + 3 * k];
        }
      }
    }
  } else if (s.kind == JUMP_MAT_ELASTOPLASTIC) {
    double ep[6], epbo[6];
    for (int c = 0; c < 6; c++) {
      ep[c] = P.ep[c][i];
      epbo[c] = ad.epb[c][i];
    }
    j2_3d_reverse(s, e, ep, P.alpha[i], sgb, ad.ab[i], epbo, eb, ab_in, epb_in, lamb, mub, fyb, k1b);
  }
  for (int k = 0; k < 9; k++) Fnb[k] += Jb * cof[k];
  // F' = F A
  double Fb[9], Ab[9];
  for (int b = 0; b < 3; b++)
    for (int a = 0; a < 3; a++) {
      double f1 = 0.0, f2 = 0.0;
      for (int k = 0; k < 3; k++) {
        f1 += Fnb[a + 3 * k] * A[b + 3 * k];  // F_b[a,b] = sum_k F'_b[a,k] A[b,k]
        f2 += F[k + 3 * a] * Fnb[k + 3 * b];  // A_b[a,b] = sum_k F[k,a] F'_b[k,b]
      }
      Fb[a + 3 * b] = f1;
      Ab[a + 3 * b] = f2;
    }
  // A = I + dt L;  eps' = eps + dt sym(L)
  double Lb[9];
  for (int k = 0; k < 9; k++) Lb[k] = dt * Ab[k];
  for (int c = 0; c < 6; c++) {
    if (c < 3) {
      Lb[sa[c] + 3 * sa[c]] += dt * eb[c];
    } else {
      Lb[sa[c] + 3 * sbb[c]] += 0.5 * dt * eb[c];
      Lb[sbb[c] + 3 * sa[c]] += 0.5 * dt * eb[c];
    }
  }
  const double xb1[3] = {ad.xb[0][i], ad.xb[1][i], ad.xb[2][i]};
  src.put_xL(i, xb1, Lb);
  // x' = x + sum N dt P/m and L through N(x), g(x)
  double xb[3] = {xb1[0], xb1[1], xb1[2]};
  stc.each(G, [&](int n, double, const double* gr, const double* H) {
    if (!(nv.mii[n] > 0.0)) return;
    const double u[3] = {nv.v[0][n], nv.v[1][n], nv.v[2][n]};
    const double u2[3] = {nv.v2[0][n], nv.v2[1][n], nv.v2[2][n]};
    const double Nb = dt * (u[0] * xb1[0] + u[1] * xb1[1] + u[2] * xb1[2]);
    double gb[3];  // adjoint of g_a: sum_b Lb[a,b] u2_b
    for (int a = 0; a < 3; a++) gb[a] = Lb[a] * u2[0] + Lb[a + 3] * u2[1] + Lb[a + 6] * u2[2];
    for (int c = 0; c < 3; c++) xb[c] += Nb * gr[c] + hess(H, 0, c) * gb[0] + hess(H, 1, c) * gb[1] + hess(H, 2, c) * gb[2];
  });
  for (int k = 0; k < 3; k++) ad.xb[k][i] = xb[k];
  for (int k = 0; k < 9; k++) ad.Fb[k][i] = Fb[k];
  // sigma and V are overwritten by the update: their adjoints before the
  // step are their P2G uses alone, which k_adj3_p2g stores
  for (int c = 0; c < 6; c++) {
    ad.eb[c][i] = eb[c];
    ad.epb[c][i] = epb_in[c];
  }
  ad.ab[i] = ab_in;
  ad.lam[i] += lamb;
  ad.mu[i] += mub;
  ad.fy[i] += fyb;
  ad.k1[i] += k1b;
}

// Node gather over the particles of the node's window (the step's bins).
// fn(particle, N, g[3]) for every particle whose stencil covers node n that
// lane sl of the node's S lanes takes: entries sl, sl+S, ... of each bin
// (lanes_sum).
template <int B, int S, class Fn>
__device__ __forceinline__ void window_gather3(const Particles& P, const GridDesc& G, const SolidDev* SD,
                                               const int* start, const unsigned long long* sorted, int n, int sl,
                                               Fn&& fn) {
  constexpr int W = Width<B>::W;
  const int ni = n % G.nx, nj = (n / G.nx) % G.ny, nk = n / G.nxy;
  const Axis X = axis_of(G, 0), Y = axis_of(G, 1), Z = axis_of(G, 2);
  XWeightCache xc;
  for (int bk = max(nk - W + 1, 0); bk <= nk; bk++)
    for (int bj = max(nj - W + 1, 0); bj <= nj; bj++)
      for (int bi = max(ni - W + 1, 0); bi <= ni; bi++) {
        const int bin = bi + G.nx * (bj + G.ny * bk);
        const int e = start[bin + 1];
        for (int c = start[bin] + sl; c < e; c += S) {
          const int q = int(unsigned(sorted[c]));
          const SolidDev& s = SD[P.sid[q]];
          const double2 xy = ldu(P.pr[0], unsigned(q));
          double d2x, wy, d1y, d2y, wz, d1z, d2z;
          // a lane's particles of one bin share x on a lattice of S per cell
          // and direction: its x weight is formed again only when some
          // active lane of the wave has another x (or GIMP half-width)
          if (!__all(xy.x == xc.x && s.lp[0] == xc.lp)) {
            axis_w2<B>(xy.x, X, s.lp[0], ni + 1, xc.w, xc.d1, d2x);
            xc.x = xy.x;
            xc.lp = s.lp[0];
          }
          const double wx = xc.w, d1x = xc.d1;
          axis_w2<B>(xy.y, Y, s.lp[1], nj + 1, wy, d1y, d2y);
          axis_w2<B>(P.x[2][q], Z, s.lp[2], nk + 1, wz, d1z, d2z);
          const double gr[3] = {d1x * wy * wz, wx * d1y * wz, wx * wy * d1z};
          fn(q, wx * wy * wz, gr);
        }
      }
}

//   P_b  += N dt x'_b / m      P2_b[b] += sum_a g_a L_b[a,b] / m
//   m_b  -= (N dt P.x'_b + sum_ab g_a P2_b L_b[a,b]) / m^2
// S lanes per node (lanes_sum); the nodes in the step's order, those with a
// particle in their window first (order, nodepos).
template <int B, int S>
__global__ void __launch_bounds__(kBlock) k_adj3_nodes2(Particles P, GridDesc G, const SolidDev* __restrict__ SD,
                                                        const int* __restrict__ start,
                                                        const unsigned long long* __restrict__ sorted, GridArrays g,
                                                        double dt, AdjSources3 src, AdjNodes3 an,
                                                        const int* __restrict__ order,
                                                        const int* __restrict__ nodepos) {
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int q = int(tid / S), sl = int(tid % S);
  const bool live = q < G.nn;  // every lane reaches the sums below
  const int n = live ? order[q] : 0;
  const double m = live && q < nodepos[G.nn] ? g.m[n] : 0.0;
  double Pb[3] = {0.0, 0.0, 0.0}, P2b[3] = {0.0, 0.0, 0.0}, mb = 0.0;
  if (m > 0.0) {
    const double iM = 1.0 / m;
    const double u[3] = {g.p[0][n] * iM, g.p[1][n] * iM, g.p[2][n] * iM};
    const double u2[3] = {g.p2[0][n] * iM, g.p2[1][n] * iM, g.p2[2][n] * iM};
    window_gather3<B, S>(P, G, SD, start, sorted, n, sl, [&](int q, double N, const double* gr) {
      double xb[3], Lq[9];
      src.get_xL(q, xb, Lq);
      double ub[3], u2b[3];
      for (int b = 0; b < 3; b++) {
        ub[b] = N * dt * xb[b];
        u2b[b] = gr[0] * Lq[3 * b] + gr[1] * Lq[1 + 3 * b] + gr[2] * Lq[2 + 3 * b];
        Pb[b] += ub[b] * iM;
        P2b[b] += u2b[b] * iM;
      }
      mb -= (u[0] * ub[0] + u[1] * ub[1] + u[2] * ub[2] + u2[0] * u2b[0] + u2[1] * u2b[1] + u2[2] * u2b[2]) * iM;
    });
  }
  for (int k = 0; k < 3; k++) {
    Pb[k] = lanes_sum<S>(Pb[k]);
    P2b[k] = lanes_sum<S>(P2b[k]);
  }
  mb = lanes_sum<S>(mb);
  if (!live || sl != 0) return;
  an.mb[n] = mb;
  for (int k = 0; k < 3; k++) {
    an.Pb[k][n] = Pb[k];
    an.P2b[k][n] = P2b[k];
    an.P0b[k][n] = 0.0;
  }
}

// Re-projection P2 = sum N m_p v' and G2P-1 v' = alpha v + sum N (P - alpha P0)/m.
template <int B>
__global__ void __launch_bounds__(kBlock) k_adj3_g2p1(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                      GridArrays g, double alpha, AdjNodes3 an, AdjParticles3 ad,
                                                      AdjSources3 src, NodeVel3 nv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  const double mp = P.m[i];
  // forward recomputation of v', and the re-projection's part of v'_b
  double vn[3] = {alpha * P.v[0][i], alpha * P.v[1][i], alpha * P.v[2][i]};
  double vx[3] = {0.0, 0.0, 0.0}, vb[3] = {ad.vb[0][i], ad.vb[1][i], ad.vb[2][i]};
  Stencil3<B> stc;
  const bool ok = stc.init(P, G, s, i);
  if (ok)
    stc.each(G, [&](int n, double N, const double*, const double*) {
      const double mii = nv.mii[n];
      if (mii > 0.0)
        for (int k = 0; k < 3; k++) vx[k] += nv.q[k][n] * (mii * N);
      for (int k = 0; k < 3; k++) vb[k] += N * mp * an.P2b[k][n];
    });
  if (!ok) {
    const double z[3] = {0.0, 0.0, 0.0};
    src.put_vb(i, z);
    return;
  }
  for (int k = 0; k < 3; k++) vn[k] += vx[k];
  double xb[3] = {0.0, 0.0, 0.0};
  stc.each(G, [&](int n, double, const double* gr, const double*) {
    double Nb = mp * (vn[0] * an.P2b[0][n] + vn[1] * an.P2b[1][n] + vn[2] * an.P2b[2][n]);
    if (nv.mii[n] > 0.0)
      for (int k = 0; k < 3; k++) Nb += nv.qm[k][n] * vb[k];
    for (int c = 0; c < 3; c++) xb[c] += Nb * gr[c];
  });
  src.put_vb(i, vb);
  for (int k = 0; k < 3; k++) {
    ad.xb[k][i] += xb[k];
    ad.vb[k][i] = alpha * vb[k];
  }
}

//   P_b += N v'_b / m    P0_b -= alpha N v'_b / m    m_b -= N (P - alpha P0).v'_b / m^2
// S lanes per node (lanes_sum); the nodes in the step's order (order,
// nodepos).
template <int B, int S>
__global__ void __launch_bounds__(kBlock) k_adj3_nodes1(Particles P, GridDesc G, const SolidDev* __restrict__ SD,
                                                        const int* __restrict__ start,
                                                        const unsigned long long* __restrict__ sorted, GridArrays g,
                                                        double alpha, AdjSources3 src, AdjNodes3 an,
                                                        const int* __restrict__ order,
                                                        const int* __restrict__ nodepos) {
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int q = int(tid / S), sl = int(tid % S);
  const bool live = q < G.nn;  // every lane reaches the sums below
  const int n = live ? order[q] : 0;
  const double m = live && q < nodepos[G.nn] ? g.m[n] : 0.0;
  double Pb[3] = {0.0, 0.0, 0.0}, mb = 0.0;
  if (m > 0.0) {
    const double iM = 1.0 / m;
    const double q[3] = {(g.p[0][n] - alpha * g.p0[0][n]) * iM, (g.p[1][n] - alpha * g.p0[1][n]) * iM,
                         (g.p[2][n] - alpha * g.p0[2][n]) * iM};
    window_gather3<B, S>(P, G, SD, start, sorted, n, sl, [&](int p, double N, const double*) {
      const double2 v01 = src.vb[0][p], v2 = src.vb[1][p];
      const double vp[3] = {v01.x, v01.y, v2.x};
      for (int k = 0; k < 3; k++) {
        const double b = N * vp[k];
        Pb[k] += b * iM;
        mb -= q[k] * b * iM;
      }
    });
  }
  for (int k = 0; k < 3; k++) Pb[k] = lanes_sum<S>(Pb[k]);
  mb = lanes_sum<S>(mb);
  if (!(m > 0.0) || sl != 0) return;
  an.mb[n] += mb;
  for (int k = 0; k < 3; k++) {
    an.Pb[k][n] += Pb[k];
    an.P0b[k][n] -= alpha * Pb[k];
  }
}

// P = P0 + dt f; Dirichlet components of P0 and P are constants.  Pb holds
// the force adjoint afterwards.
__global__ void __launch_bounds__(kBlock) k_adj3_grid(int nn, const uint8_t* __restrict__ fixed, double dt,
                                                      AdjNodes3 an) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const uint8_t fx = fixed[n];
  for (int k = 0; k < 3; k++) {
    if (fx & (1u << k)) {
      an.P0b[k][n] = 0.0;
      an.Pb[k][n] = 0.0;
    } else {
      const double pb = an.Pb[k][n];
      an.P0b[k][n] += pb;
      an.Pb[k][n] = dt * pb;
    }
  }
}

template <int B>
__global__ void __launch_bounds__(kBlock) k_adj3_p2g(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                     AdjNodes3 an, AdjParticles3 ad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  const double m = P.m[i], V = P.V[i];
  const double v[3] = {P.v[0][i], P.v[1][i], P.v[2][i]};
  double sg[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) sg[a][b] = P.sig[sym<3>(a, b)][i];
  double xb[3] = {0.0, 0.0, 0.0}, vb[3] = {0.0, 0.0, 0.0}, sb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, Vb = 0.0;
  const bool ok = walk_stencil_h3<B>(P, G, s, i, [&](int n, double N, const double* gr, const double* H) {
    const double mb = an.mb[n];
    const double p0b[3] = {an.P0b[0][n], an.P0b[1][n], an.P0b[2][n]};
    const double fb[3] = {an.Pb[0][n], an.Pb[1][n], an.Pb[2][n]};
    const double Nb = mb * m + m * (v[0] * p0b[0] + v[1] * p0b[1] + v[2] * p0b[2]);
    double sgf[3], gb[3];  // (sigma g)_a and the adjoint of g
    for (int a = 0; a < 3; a++) {
      vb[a] += N * m * p0b[a];
      sgf[a] = sg[a][0] * gr[0] + sg[a][1] * gr[1] + sg[a][2] * gr[2];
      gb[a] = -V * (sg[0][a] * fb[0] + sg[1][a] * fb[1] + sg[2][a] * fb[2]);
    }
    Vb -= sgf[0] * fb[0] + sgf[1] * fb[1] + sgf[2] * fb[2];
    // f_a = -V sum_b sigma_ab g_b, sigma stored once per symmetric pair
    sb[0] -= V * gr[0] * fb[0];
    sb[1] -= V * gr[1] * fb[1];
    sb[2] -= V * gr[2] * fb[2];
    sb[3] -= V * (gr[1] * fb[0] + gr[0] * fb[1]);
    sb[4] -= V * (gr[2] * fb[0] + gr[0] * fb[2]);
    sb[5] -= V * (gr[2] * fb[1] + gr[1] * fb[2]);
    for (int c = 0; c < 3; c++) xb[c] += Nb * gr[c] + hess(H, 0, c) * gb[0] + hess(H, 1, c) * gb[1] + hess(H, 2, c) * gb[2];
  });
  for (int k = 0; k < 3; k++) {
    ad.xb[k][i] += xb[k];
    ad.vb[k][i] += vb[k];
  }
  // stored where the step updated the particle (k_adj3_g2p2 consumed the
  // adjoints of the updated sigma and V; the sums start at +0 and are never
  // -0, so they equal 0 + the sum bit for bit), passed through elsewhere
  if (ok) {
    for (int c = 0; c < 6; c++) ad.sb[c][i] = sb[c];
    ad.Vb[i] = Vb;
  } else {
    for (int c = 0; c < 6; c++) ad.sb[c][i] += sb[c];
    ad.Vb[i] += Vb;
  }
}

// Per-solid sums of per-particle adjoint arrays (nsum of them, array k at
// arr[k]), in two passes of fixed order: block (c, solid) sums chunk c of the
// solid's particle range (kSumChunk particles) into part[solid][c][k]
// (solid = si0 + blockIdx.y, the grid's y extent being limited); then
// one block per solid sums its nchunk partials.
constexpr int kAdjSums = 7;
constexpr int kSumChunk = 8192;
struct AdjSumArrays {
  const double* a[kAdjSums];
};
__device__ __forceinline__ void block_sum(double (*sh)[kBlock], double* acc, int nsum) {
  for (int k = 0; k < kAdjSums; k++) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int k = 0; k < nsum; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
    __syncthreads();
  }
}
__global__ void __launch_bounds__(kBlock) k_adj_sums_part(const long long* __restrict__ range, AdjSumArrays arr,
                                                          int nsum, int si0, double* __restrict__ part) {
  __shared__ double sh[kAdjSums][kBlock];
  const int c = blockIdx.x, si = si0 + blockIdx.y;
  const long long e = range[2 * si + 1];
  const long long b = range[2 * si] + (long long)c * kSumChunk, ce = b + kSumChunk < e ? b + kSumChunk : e;
  double acc[kAdjSums];
  for (int k = 0; k < kAdjSums; k++) acc[k] = 0.0;
  for (long long i = b + threadIdx.x; i < ce; i += kBlock)
    for (int k = 0; k < nsum; k++) acc[k] += arr.a[k][i];
  block_sum(sh, acc, nsum);
  if (threadIdx.x < nsum) part[(size_t(si) * gridDim.x + c) * nsum + threadIdx.x] = sh[threadIdx.x][0];
}
__global__ void __launch_bounds__(kBlock) k_adj_sums_final(const double* __restrict__ part, int nchunk, int nsum,
                                                           double* __restrict__ out) {
  __shared__ double sh[kAdjSums][kBlock];
  const int si = blockIdx.x;
  double acc[kAdjSums];
  for (int k = 0; k < kAdjSums; k++) acc[k] = 0.0;
  for (int c = threadIdx.x; c < nchunk; c += kBlock)
    for (int k = 0; k < nsum; k++) acc[k] += part[(size_t(si) * nchunk + c) * nsum + k];
  block_sum(sh, acc, nsum);
  if (threadIdx.x < nsum) out[nsum * si + threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace jh
