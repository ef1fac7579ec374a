// mpm_adjoint.hpp -- reverse mode of the 2D USL step (linear and GIMP bases,
// Hooke and Neo-Hookean solids): given the adjoints of the particle state
// after a step, the adjoints of the state before it and of the material
// constants lambda, mu.
//
// Forward step (SolveExplicitSolidMechanics2D.jl:391-534), particle p, node I,
// N = N_I(x_p), g = grad N_I(x_p):
//   m_I = sum N m_p          P0_I = sum N m_p v_p
//   f_I = sum -V_p sigma_p g + m_p b N          P_I = P0_I + dt f_I  (Dirichlet: 0)
//   for nodes with m_I > tol:  u = P/m,  a = (P - P0)/m
//   v' = v + sum N a     x' = x + dt sum N u     L = sum u (x) g   (L[a][b] = u_a g_b)
//   eps' = eps + dt sym(L)   F' = F (I + dt L)   J = det F'   V' = J V0
//   sigma' = Hooke(eps') | (mu (F'F'^T - I) + lambda ln J I) / J
// The adjoint runs the three parts backwards:
//   k_adj_g2p    per particle: stress, V, F, eps, then the G2P sums; x picks
//                up dN/dx and d(grad N)/dx terms; the particle's sources of
//                the node adjoints are stored
//   k_adj_nodes  per node: the node adjoints of m, P, P0, gathered from the
//                particles of its window (no atomics)
//   k_adj_grid   per node: P = P0 + dt f with the Dirichlet mask
//   k_adj_p2g    per particle: the P2G sums gathered from the nodes
// Nodes below the tolerance and the Dirichlet components are constants of
// the step, as in the forward.  Every sum has a fixed order, so gradients
// are bitwise reproducible run to run.
#pragma once

#include "mpm_kernels.hpp"

namespace jh {

struct AdjParticles {
  double* xb[2];
  double* vb[2];
  double* Fb[4];
  double* eb[3];
  double* sb[3];
  double* Vb;
  double* lam;  // accumulated adjoint of lambda, per particle
  double* mu;   // accumulated adjoint of mu, per particle
};
struct AdjNodes {
  double* mb;
  double* Pb[2];
  double* P0b[2];
};
// Per-particle sources of the G2P node adjoints: the adjoints of x' and v'
// and of L, written by k_adj_g2p and gathered per node by k_adj_nodes.
// In pairs, one 16-byte load each where k_adj_nodes gathers them.
struct AdjSources {
  double2* xb;     // (x, y)
  double2* vb;     // (x, y)
  double2* Lb[2];  // (Lb0, Lb1), (Lb2, Lb3)
};

// Weight, first and second derivative of node id (1-based) on one axis.
template <int B>
__device__ __forceinline__ void axis_w2(double x, const Axis& A, double lp, int id, double& w, double& d1, double& d2) {
  const double xi = x - (A.mn + double(id - 1) * A.h);
  if (B == JUMP_BASIS_LINEAR) {
    w = 1.0 - fabs(xi) * A.hI;
    d1 = -jsign(xi) * A.hI;
    d2 = 0.0;
  } else {  // uGIMP
    double dn;
    gimp_eval(xi, A.h, A.hI, lp, w, dn);
    d1 = dn;
    const double a = fabs(xi);
    if (a < lp) d2 = -1.0 / (A.h * lp);
    else if (a <= A.h - lp) d2 = 0.0;
    else if (a < A.h + lp) d2 = 1.0 / (2.0 * A.h * lp);
    else d2 = 0.0;
  }
}

// A particle's stencil: first node and the per-axis weights with their
// first and second derivatives, formed once; each() calls
// fn(node, N, g[2], H[3] = {Nxx, Nxy, Nyy}, k) for the support, k the node's
// place in the walk (y outer, x inner).
template <int B> struct Stencil2 {
  static constexpr int W = Width<B>::W;
  int f[2];
  double w[2][W], d1[2][W], d2[2][W];
  __device__ __forceinline__ bool init(const Particles& P, const GridDesc& G, const SolidDev& s, int p) {
    const Axis A[2] = {axis_of(G, 0), axis_of(G, 1)};
    for (int a = 0; a < 2; a++)
      if (!first_node<B, 2>(P.x[a][p], a, A[a], s.lp[a], f[a])) return false;
    for (int a = 0; a < 2; a++)
      for (int o = 0; o < W; o++) axis_w2<B>(P.x[a][p], A[a], s.lp[a], f[a] + o, w[a][o], d1[a][o], d2[a][o]);
    return true;
  }
  template <class Fn> __device__ __forceinline__ void each(const GridDesc& G, Fn&& fn) const {
#pragma unroll
    for (int b = 0; b < W; b++)
#pragma unroll
      for (int a = 0; a < W; a++) {
        const double N = w[0][a] * w[1][b];
        const double g[2] = {d1[0][a] * w[1][b], w[0][a] * d1[1][b]};
        const double H[3] = {d2[0][a] * w[1][b], d1[0][a] * d1[1][b], w[0][a] * d2[1][b]};
        fn((f[0] + a - 1) + G.nx * (f[1] + b - 1), N, g, H, a + W * b);
      }
  }
};
// Calls fn(node, N, g[2], H[3] = {Nxx, Nxy, Nyy}) for the support of p.
template <int B, class Fn>
__device__ __forceinline__ bool walk_stencil_h(const Particles& P, const GridDesc& G, const SolidDev& s, int p,
                                               Fn&& fn) {
  Stencil2<B> st;
  if (!st.init(P, G, s, p)) return false;
  st.each(G, [&](int n, double N, const double* g, const double* H, int) { fn(n, N, g, H); });
  return true;
}

// Per node, the velocities k_adj_g2p reads (nodes with mass above the
// tolerance only): uL = p/m as the forward's G2P forms it (so the velocity
// gradient it recomputes is the forward's), u = p (1/m) and a = (p - p0) (1/m)
// (k_adj_nodevel).
struct NodeVel2 {
  double2* uL;
  double2* u;
  double2* a;
};
__global__ void __launch_bounds__(kBlock) k_adj_nodevel(int nn, GridArrays g, double tol, NodeVel2 nv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const double mI = g.m[n];
  if (!(mI > tol)) return;
  const double p0 = g.p[0][n], p1 = g.p[1][n];
  nv.uL[n] = make_double2(p0 / mI, p1 / mI);
  const double iM = 1.0 / mI;
  nv.u[n] = make_double2(p0 * iM, p1 * iM);
  nv.a[n] = make_double2((p0 - g.p0[0][n]) * iM, (p1 - g.p0[1][n]) * iM);
}

template <int B>
__global__ void __launch_bounds__(kBlock) k_adj_g2p(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                    GridArrays g, double tol, double dt, AdjParticles ad,
                                                    AdjSources src, NodeVel2 nv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  if (s.rigid) return;
  // forward recomputation of L, eps', F', J
  double L0 = 0.0, L1 = 0.0, L2 = 0.0, L3 = 0.0;
  Stencil2<B> stc;
  unsigned mask = 0;  // the stencil's nodes with mass above the tolerance
  const bool ok = stc.init(P, G, s, i);
  if (ok)
    stc.each(G, [&](int n, double, const double* gr, const double*, int k) {
      if (g.m[n] > tol) {
        mask |= 1u << k;
        const double2 uL = nv.uL[n];
        L0 += gr[0] * uL.x;
        L1 += gr[0] * uL.y;
        L2 += gr[1] * uL.x;
        L3 += gr[1] * uL.y;
      }
    });
  if (!ok) {  // the forward left this particle unchanged: identity adjoint, no node sources
    src.xb[i] = make_double2(0.0, 0.0);
    src.vb[i] = make_double2(0.0, 0.0);
    src.Lb[0][i] = make_double2(0.0, 0.0);
    src.Lb[1][i] = make_double2(0.0, 0.0);
    return;
  }
  const double e[3] = {P.eps[0][i] + dt * L0, P.eps[1][i] + dt * L3, P.eps[2][i] + dt * (0.5 * (L2 + L1))};
  const double A[4] = {1.0 + L0 * dt, L1 * dt, L2 * dt, 1.0 + L3 * dt};
  const double F[4] = {P.F[0][i], P.F[1][i], P.F[2][i], P.F[3][i]};
  const double Fn[4] = {F[0] * A[0] + F[2] * A[1], F[1] * A[0] + F[3] * A[1], F[0] * A[2] + F[2] * A[3],
                        F[1] * A[2] + F[3] * A[3]};
  const double J = Fn[0] * Fn[3] - Fn[2] * Fn[1];
  // incoming adjoints of the state after the step
  const double sb[3] = {ad.sb[0][i], ad.sb[1][i], ad.sb[2][i]};
  double eb[3] = {ad.eb[0][i], ad.eb[1][i], ad.eb[2][i]};
  double Fnb[4] = {ad.Fb[0][i], ad.Fb[1][i], ad.Fb[2][i], ad.Fb[3][i]};
  double Jb = ad.Vb[i] * P.V0[i];
  double lamb = 0.0, mub = 0.0;
  if (s.kind == JUMP_MAT_ELASTIC) {
    const double tr = e[0] + e[1], trb = sb[0] + sb[1];
    eb[0] += s.lambda * trb + 2.0 * s.mu * sb[0];
    eb[1] += s.lambda * trb + 2.0 * s.mu * sb[1];
    eb[2] += 2.0 * s.mu * sb[2];
    lamb += tr * trb;
    mub += 2.0 * (e[0] * sb[0] + e[1] * sb[1] + e[2] * sb[2]);
  } else if (s.kind == JUMP_MAT_NEOHOOKEAN) {
    const double B11 = Fn[0] * Fn[0] + Fn[2] * Fn[2], B22 = Fn[1] * Fn[1] + Fn[3] * Fn[3];
    const double B12 = Fn[0] * Fn[1] + Fn[2] * Fn[3];
    const double lnJ = log(J), iJ = 1.0 / J;
    const double T11 = s.mu * (B11 - 1.0) + s.lambda * lnJ, T22 = s.mu * (B22 - 1.0) + s.lambda * lnJ,
                 T12 = s.mu * B12;
    const double Tb11 = iJ * sb[0], Tb22 = iJ * sb[1], Tb12 = iJ * sb[2];
    Jb += -(T11 * sb[0] + T22 * sb[1] + T12 * sb[2]) * iJ * iJ;
    Jb += s.lambda * (Tb11 + Tb22) * iJ;
    lamb += lnJ * (Tb11 + Tb22);
    mub += (B11 - 1.0) * Tb11 + (B22 - 1.0) * Tb22 + B12 * Tb12;
    const double Bb11 = s.mu * Tb11, Bb22 = s.mu * Tb22, Bb12 = s.mu * Tb12;
    Fnb[0] += 2.0 * Bb11 * Fn[0] + Bb12 * Fn[1];
    Fnb[2] += 2.0 * Bb11 * Fn[2] + Bb12 * Fn[3];
    Fnb[1] += 2.0 * Bb22 * Fn[1] + Bb12 * Fn[0];
    Fnb[3] += 2.0 * Bb22 * Fn[3] + Bb12 * Fn[2];
  }
  // J = Fn0 Fn3 - Fn2 Fn1
  Fnb[0] += Jb * Fn[3];
  Fnb[3] += Jb * Fn[0];
  Fnb[2] -= Jb * Fn[1];
  Fnb[1] -= Jb * Fn[2];
  // F' = F A
  double Fb[4], Ab[4];
  Fb[0] = Fnb[0] * A[0] + Fnb[2] * A[2];
  Fb[1] = Fnb[1] * A[0] + Fnb[3] * A[2];
  Fb[2] = Fnb[0] * A[1] + Fnb[2] * A[3];
  Fb[3] = Fnb[1] * A[1] + Fnb[3] * A[3];
  Ab[0] = Fnb[0] * F[0] + Fnb[1] * F[1];
  Ab[1] = Fnb[0] * F[2] + Fnb[1] * F[3];
  Ab[2] = Fnb[2] * F[0] + Fnb[3] * F[1];
  Ab[3] = Fnb[2] * F[2] + Fnb[3] * F[3];
  // A = I + dt L;  eps' = eps + dt sym(L)
  const double Lb0 = dt * Ab[0] + dt * eb[0];
  const double Lb3 = dt * Ab[3] + dt * eb[1];
  const double Lb1 = dt * Ab[1] + 0.5 * dt * eb[2];
  const double Lb2 = dt * Ab[2] + 0.5 * dt * eb[2];
  const double xb1[2] = {ad.xb[0][i], ad.xb[1][i]};
  const double vb1[2] = {ad.vb[0][i], ad.vb[1][i]};
  src.xb[i] = make_double2(xb1[0], xb1[1]);
  src.vb[i] = make_double2(vb1[0], vb1[1]);
  src.Lb[0][i] = make_double2(Lb0, Lb1);
  src.Lb[1][i] = make_double2(Lb2, Lb3);
  double xb[2] = {xb1[0], xb1[1]};
  stc.each(G, [&](int n, double N, const double* gr, const double* H, int k) {
    if (!((mask >> k) & 1u)) return;
    const double2 u2 = nv.u[n], a2 = nv.a[n];
    const double u[2] = {u2.x, u2.y};
    const double a[2] = {a2.x, a2.y};
    const double Nb = dt * (u[0] * xb1[0] + u[1] * xb1[1]) + a[0] * vb1[0] + a[1] * vb1[1];
    const double gb[2] = {Lb0 * u[0] + Lb1 * u[1], Lb2 * u[0] + Lb3 * u[1]};
    xb[0] += Nb * gr[0] + H[0] * gb[0] + H[1] * gb[1];
    xb[1] += Nb * gr[1] + H[1] * gb[0] + H[2] * gb[1];
  });
  ad.xb[0][i] = xb[0];
  ad.xb[1][i] = xb[1];
  // v' = v + ...: identity part of v stays in ad.vb
  for (int k = 0; k < 4; k++) ad.Fb[k][i] = Fb[k];
  // sigma and V are overwritten by the update: their adjoints before the
  // step are their P2G uses alone, which k_adj_p2g stores
  for (int k = 0; k < 3; k++) ad.eb[k][i] = eb[k];
  ad.lam[i] += lamb;
  ad.mu[i] += mub;
}

// Lanes per node of the node gathers below.  With one lane per node, the
// lanes of a wave sit on neighbouring nodes of a grid row, whose windows are
// one cell apart: at each visit lane i reads a particle one cell's worth of
// particles after lane i-1's (4 apart at 4 particles per cell and
// direction), so every 16-byte wave-load touches 4x the cache lines it
// uses.  With S lanes per node, lane s of a node takes entries s, s+S,
// s+2S, ... of each window bin; on a lattice of S particles per cell and
// direction (x fastest) the S lanes read S consecutive particles, and the
// wave's loads are contiguous.  The partial sums are combined in a fixed
// order, so the gathers stay deterministic.
// The x weight and its derivative of the last particle a lane weighed
// against its node in a node gather (k_adj_nodes, window_gather3).
struct XWeightCache {
  double x = __builtin_nan(""), lp = __builtin_nan("");  // NaN: nothing cached
  double w = 0.0, d1 = 0.0;
};
template <int S> __device__ __forceinline__ double lanes_sum(double v) {
  // ((l0 + l1) + (l2 + l3)) in every lane of the group: addition commutes
  // exactly, so all S lanes hold the same bits
  if (S >= 2) v += __shfl_xor(v, 1);
  if (S >= 4) v += __shfl_xor(v, 2);
  return v;
}

// Node adjoints of the G2P sums, gathered per node over the particles of
// its window (the step's bins, any fixed order: the sum is not bit-matched
// to a reference).  u = P/m, a = (P - P0)/m:
//   Pb  += (ub + ab)/m    P0b -= ab/m    mb -= (u.ub + a.ab)/m
//   ub = N dt xb' + Lb g (L[a][b] = u_a g_b),   ab = N vb'
// S lanes per node (lanes_sum).  The nodes are taken in the step's order,
// those with a particle in their window first (order, nodepos); the rest
// have no mass and get zeros.
template <int B, int S>
__global__ void __launch_bounds__(kBlock) k_adj_nodes(Particles P, GridDesc G, const SolidDev* __restrict__ S_,
                                                      const int* __restrict__ start,
                                                      const unsigned long long* __restrict__ sorted, GridArrays g,
                                                      double tol, double dt, AdjSources src, AdjNodes an,
                                                      const int* __restrict__ order,
                                                      const int* __restrict__ nodepos) {
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int q = int(tid / S), sl = int(tid % S);
  const bool live = q < G.nn;  // every lane reaches the sums below
  const int n = live ? order[q] : 0;
  const double mI = live && q < nodepos[G.nn] ? g.m[n] : 0.0;
  double Pb0 = 0.0, Pb1 = 0.0, P0b0 = 0.0, P0b1 = 0.0, mb = 0.0;
  if (mI > tol) {
    constexpr int W = Width<B>::W;
    const int ni = n % G.nx, nj = n / G.nx;
    const double iM = 1.0 / mI;
    const double u[2] = {g.p[0][n] * iM, g.p[1][n] * iM};
    const double a[2] = {(g.p[0][n] - g.p0[0][n]) * iM, (g.p[1][n] - g.p0[1][n]) * iM};
    const Axis X = axis_of(G, 0), Y = axis_of(G, 1);
    XWeightCache xc;
    for (int bj = max(nj - W + 1, 0); bj <= nj; bj++)
      for (int bi = max(ni - W + 1, 0); bi <= ni; bi++) {
        const int bin = bi + G.nx * bj;
        const int e = start[bin + 1];
        int c = start[bin] + sl;
        if (c >= e) continue;
        // the next particle's index, solid and position are loaded while
        // the current one is summed (the order of the sums is unchanged)
        int q = int(unsigned(sorted[c]));
        int sq = P.sid[q];
        double2 xy = ldu(P.pr[0], unsigned(q));
        for (; c < e; c += S) {
          int qn = q, sn = sq;
          double2 xyn = xy;
          if (c + S < e) {
            qn = int(unsigned(sorted[c + S]));
            sn = P.sid[qn];
            xyn = ldu(P.pr[0], unsigned(qn));
          }
          const SolidDev& s = S_[sq];
          double d2x, wy, d1y, d2y;
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
          const double N = wx * wy, gx = d1x * wy, gy = wx * d1y;
          const double2 xb2 = src.xb[q], vb2 = src.vb[q], L01 = src.Lb[0][q], L23 = src.Lb[1][q];
          const double xb[2] = {xb2.x, xb2.y}, vb[2] = {vb2.x, vb2.y};
          const double Lb[4] = {L01.x, L01.y, L23.x, L23.y};
          const double ub0 = N * dt * xb[0] + Lb[0] * gx + Lb[2] * gy;
          const double ub1 = N * dt * xb[1] + Lb[1] * gx + Lb[3] * gy;
          const double ab0 = N * vb[0], ab1 = N * vb[1];
          Pb0 += (ub0 + ab0) * iM;
          Pb1 += (ub1 + ab1) * iM;
          P0b0 -= ab0 * iM;
          P0b1 -= ab1 * iM;
          mb -= (u[0] * ub0 + u[1] * ub1 + a[0] * ab0 + a[1] * ab1) * iM;
          q = qn;
          sq = sn;
          xy = xyn;
        }
      }
  }
  Pb0 = lanes_sum<S>(Pb0);
  Pb1 = lanes_sum<S>(Pb1);
  P0b0 = lanes_sum<S>(P0b0);
  P0b1 = lanes_sum<S>(P0b1);
  mb = lanes_sum<S>(mb);
  if (!live || sl != 0) return;
  an.mb[n] = mb;
  an.Pb[0][n] = Pb0;
  an.Pb[1][n] = Pb1;
  an.P0b[0][n] = P0b0;
  an.P0b[1][n] = P0b1;
}

// P = P0 + dt f; Dirichlet components of P0 and P are constants.
__global__ void __launch_bounds__(kBlock) k_adj_grid(int nn, const uint8_t* __restrict__ fixed, double dt,
                                                     AdjNodes an) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const uint8_t fx = fixed[n];
  for (int k = 0; k < 2; k++) {
    if (fx & (1u << k)) {
      an.P0b[k][n] = 0.0;
      an.Pb[k][n] = 0.0;  // holds the force adjoint from here on
    } else {
      const double pb = an.Pb[k][n];
      an.P0b[k][n] += pb;
      an.Pb[k][n] = dt * pb;
    }
  }
}

template <int B>
__global__ void __launch_bounds__(kBlock) k_adj_p2g(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                    AdjNodes an, AdjParticles ad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  if (s.rigid) return;
  const double m = P.m[i], V = P.V[i];
  const double v[2] = {P.v[0][i], P.v[1][i]};
  const double sg[3] = {P.sig[0][i], P.sig[1][i], P.sig[2][i]};
  double xb[2] = {0.0, 0.0}, vb[2] = {0.0, 0.0}, sb[3] = {0.0, 0.0, 0.0}, Vb = 0.0;
  const bool ok = walk_stencil_h<B>(P, G, s, i, [&](int n, double N, const double* gr, const double* H) {
    const double mb = an.mb[n];
    const double p0b[2] = {an.P0b[0][n], an.P0b[1][n]};
    const double fb[2] = {an.Pb[0][n], an.Pb[1][n]};
    const double Nb = mb * m + m * (v[0] * p0b[0] + v[1] * p0b[1]) + m * (G.body[0] * fb[0] + G.body[1] * fb[1]);
    vb[0] += N * m * p0b[0];
    vb[1] += N * m * p0b[1];
    sb[0] -= V * gr[0] * fb[0];
    sb[1] -= V * gr[1] * fb[1];
    sb[2] -= V * (gr[1] * fb[0] + gr[0] * fb[1]);
    Vb -= (sg[0] * gr[0] + sg[2] * gr[1]) * fb[0] + (sg[2] * gr[0] + sg[1] * gr[1]) * fb[1];
    const double gb[2] = {-V * (sg[0] * fb[0] + sg[2] * fb[1]), -V * (sg[2] * fb[0] + sg[1] * fb[1])};
    xb[0] += Nb * gr[0] + H[0] * gb[0] + H[1] * gb[1];
    xb[1] += Nb * gr[1] + H[1] * gb[0] + H[2] * gb[1];
  });
  ad.xb[0][i] += xb[0];
  ad.xb[1][i] += xb[1];
  ad.vb[0][i] += vb[0];
  ad.vb[1][i] += vb[1];
  // Where the step updated the particle (ok), k_adj_g2p consumed the
  // adjoints of the updated sigma and V, and these sums are stored; they
  // start at +0 and are never -0, so they equal 0 + the sum bit for bit.
  // Elsewhere the adjoints pass through, plus the (zero) sums.
  if (ok) {
    for (int k = 0; k < 3; k++) ad.sb[k][i] = sb[k];
    ad.Vb[i] = Vb;
  } else {
    for (int k = 0; k < 3; k++) ad.sb[k][i] += sb[k];
    ad.Vb[i] += Vb;
  }
}

// Many device-to-device copies in one launch (the state, grid and bins a
// recorded step keeps): segment blockIdx.y copies words[y] 4-byte words from
// src[y] to dst[y], 16 bytes per thread, grid-stride over blockIdx.x.  Every
// array the library allocates starts 256-byte aligned and is a whole number
// of 4-byte words long.
constexpr int kCopySegs = 48;
struct CopyTable {
  const unsigned* src[kCopySegs];
  unsigned* dst[kCopySegs];
  long long words[kCopySegs];
};
// SAVE: the copy goes into a record, which is read back only in a later
// backward sweep, so it is written with non-temporal stores and leaves the
// caches to the particle state the next step reads.
constexpr bool kCopyNT = true;
template <bool SAVE>
__global__ void __launch_bounds__(kBlock) k_copy_table(CopyTable t) {
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  const int sg = blockIdx.y;
  const unsigned* __restrict__ src = t.src[sg];
  unsigned* __restrict__ dst = t.dst[sg];
  const long long n = t.words[sg], n4 = n >> 2;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const u32x4* __restrict__ s4 = reinterpret_cast<const u32x4*>(src);
  u32x4* __restrict__ d4 = reinterpret_cast<u32x4*>(dst);
  for (long long i = i0; i < n4; i += stride) {
    if (SAVE && kCopyNT) __builtin_nontemporal_store(s4[i], &d4[i]);
    else d4[i] = s4[i];
  }
  const long long tail = 4 * n4 + i0;  // the last n % 4 words
  if (tail < n) dst[tail] = src[tail];
}

__global__ void __launch_bounds__(kBlock) k_fill_f64(double* __restrict__ a, double v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Per-solid sums of the accumulated lambda and mu adjoints and of the
// initial-velocity adjoints: one block per solid over its particle range.
__global__ void __launch_bounds__(kBlock) k_adj_solid_sums(const long long* __restrict__ range, AdjParticles ad,
                                                           double* __restrict__ out) {
  __shared__ double sh[4][kBlock];
  const long long b = range[2 * blockIdx.x], e = range[2 * blockIdx.x + 1];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long long i = b + threadIdx.x; i < e; i += kBlock) {
    acc[0] += ad.lam[i];
    acc[1] += ad.mu[i];
    acc[2] += ad.vb[0][i];
    acc[3] += ad.vb[1][i];
  }
  for (int k = 0; k < 4; k++) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int k = 0; k < 4; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace jh
