// mpm_kernels.hpp -- HIP kernels of one explicit MPM step (gfx950, wave64).
//
// Step structure (USL 2D, SolveExplicitSolidMechanics2D.jl:373-581;
// MUSL 3D, SolveExplicitSolidMechanics3D.jl:41-220):
//   k_bin        per particle: stencil origin node (the reference's integer
//                support rule, bit-exact) -> bin key, per-bin count
//   k_scan_*     exclusive scan of the bin counts -> bin_start
//   k_scatter    particle index into its bin segment
//   k_binsort    each bin segment sorted by particle index (deterministic)
//   k_p2g        per grid node: gather of every particle whose stencil covers
//                the node, in bin order then particle order -- a segmented
//                reduction with a fixed order, no floating-point atomics,
//                bitwise reproducible run to run; fused with the grid update
//                p = p0 + f dt and the Dirichlet mask
//   k_rigid_*    2D rigid bodies: grid override and motion
//   k_g2p_*      per particle: gather of grid velocity, position, F, J, V and
//                the constitutive update, in the reference's node order
//   k_energy_*   kinetic/strain energy, fixed-order tree reduction
#pragma once

#include "mpm_device.hpp"

namespace jh {

constexpr int kBlock = 256;

// The 8 XCDs of an MI355X each own an L2; workgroups are dispatched to them
// round-robin.  Kernels whose neighbouring workgroups share data remap the
// hardware block index so that each XCD receives runs of kXcdChunk
// consecutive logical blocks (spatial locality in its L2), the runs dealt to
// the XCDs in turn (balance).  The grid is rounded up to a multiple of
// kXcds * kXcdChunk; surplus blocks exit.
constexpr int kXcds = 8;
constexpr int kXcdChunk = 16;
__device__ __forceinline__ int xcd_block(int nlogical) {
  const int b = int(blockIdx.x);
  const int group = b / (kXcds * kXcdChunk);
  const int r = b % (kXcds * kXcdChunk);
  const int logical = group * (kXcds * kXcdChunk) + (r % kXcds) * kXcdChunk + r / kXcds;
  return logical < nlogical ? logical : -1;
}
__host__ __device__ constexpr int xcd_grid(int nlogical) {
  return (nlogical + kXcds * kXcdChunk - 1) / (kXcds * kXcdChunk) * (kXcds * kXcdChunk);
}

__global__ void k_fill_i32(int* __restrict__ a, int v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ------------------------------------------------------------------ binning
// key[i] = linear index of the particle's first stencil node (0-based), or -1
// for particles that take no part in P2G (2D rigid bodies, out of grid).
// bxyz[i] packs the origin's 0-based coordinates, 21 bits per axis (all
// ones for excluded particles).
// A slot in bin kb for each active lane: consecutive lanes with the same bin
// (neighbouring particles usually share one) take theirs with one atomic,
// the first lane of the run adding the run's length.  Active lanes of a
// wave are contiguous from lane 0 (the grid is cut at the particle count).
__device__ __forceinline__ int wave_slot(int* cnt, int kb) {
  const int lane = int(__lane_id());
  const int kprev = __shfl_up(kb, 1u);
  const unsigned long long heads = __ballot(lane == 0 || kprev != kb);
  const unsigned long long brk = heads | ~__ballot(1);  // a run ends at the next head or inactive lane
  const int leader = 63 - __clzll((long long)(heads & ((2ull << lane) - 1ull)));
  const unsigned long long after = brk & ~((2ull << leader) - 1ull);
  const int end = after ? __ffsll((long long)after) - 1 : 64;
  int base = 0;
  if (lane == leader) base = atomicAdd(&cnt[kb], end - leader);
  base = __shfl(base, leader);
  return base + (lane - leader);
}

template <int D, int B, bool MULTI>
__global__ void __launch_bounds__(kBlock) k_bin(Particles P, GridDesc G0, const SolidDev* __restrict__ S,
                                                const double* __restrict__ corners, int* __restrict__ key,
                                                int* __restrict__ rank, unsigned long long* __restrict__ bxyz,
                                                int* __restrict__ cnt, unsigned* __restrict__ flags, BatchDev Bt,
                                                int* __restrict__ dirty, int gen, int rigid_hat, int W) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  // Run replay (single problems): a particle whose bin changed since the
  // last binning changes the windows of the nodes that contain its old or
  // new bin (W per axis); stamp them
  const int old = (!MULTI && dirty) ? key[i] : -1;
  auto mark = [&](int b) {
    if (MULTI || !dirty || b < 0) return;
    const int bx = b % G0.nx, by = (b / G0.nx) % G0.ny, bz = D == 3 ? b / G0.nxy : 0;
    for (int c = 0; c < (D == 3 ? W : 1); c++)
      for (int y = 0; y < W; y++)
        for (int a = 0; a < W; a++)
          if (bx + a < G0.nx && by + y < G0.ny && (D == 2 || bz + c < G0.nz))
            dirty[(bx + a) + G0.nx * (by + y) + (D == 3 ? G0.nxy * (bz + c) : 0)] = gen;
  };
  const SolidDev& s = S[P.sid[i]];
  const GridDesc& G = MULTI ? Bt.inst[s.inst].G : G0;
  const int off = MULTI ? Bt.node_off[s.inst] : 0;
  const int NN = MULTI ? Bt.NN : G0.nn;
  int k = -1;
  bool ok = false;
  int f[3] = {1, 1, 1};
  // 2D rigid solids take no part in P2G, except in a contact step
  // (rigid_hat), where they are binned by their hat cell: the nodes their
  // velocity is prescribed at
  const bool hat = D == 2 && s.rigid;
  if (!hat || rigid_hat) {
    ok = true;
    if (hat) {
      for (int a = 0; a < 2; a++) ok = ok && first_node<JUMP_BASIS_LINEAR, 2>(P.x[a][i], a, axis_of(G, a), 0.0, f[a]);
    } else if (B == JUMP_BASIS_CPDI_Q4) {
      CpdiW c;
      ok = cpdi_setup(corners, P.elem[0][i], P.elem[1][i], P.elem[2][i], P.elem[3][i], G, c);
      if (ok) {
        f[0] = min(min(c.ci[0], c.ci[1]), min(c.ci[2], c.ci[3]));
        f[1] = min(min(c.cj[0], c.cj[1]), min(c.cj[2], c.cj[3]));
      }
    } else {
      for (int a = 0; a < D; a++) ok = ok && first_node<B, D>(P.x[a][i], a, axis_of(G, a), s.lp[a], f[a]);
    }
    if (ok)
      k = off + (f[0] - 1) + G.nx * ((f[1] - 1) + (D == 3 ? G.ny * (f[2] - 1) : 0));
    else
      atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
  }
  // its slot in the bin, any order (k_binsort orders each bin); unbinned
  // particles count in the bin past the last node
  rank[i] = wave_slot(cnt, ok ? k : NN);
  if (ok) {
    if (k != old) {
      mark(old);
      mark(k);
    }
    key[i] = k;
    // bin coordinates 21 bits per axis; a 2D batch keeps the instance in
    // the third field, so a window of one instance never admits another's
    bxyz[i] = (unsigned long long)(f[0] - 1) | ((unsigned long long)(f[1] - 1) << 21) |
              ((unsigned long long)(D == 3 ? f[2] - 1 : (MULTI ? s.inst : 0)) << 42);
    return;
  }
  if (old >= 0) mark(old);
  key[i] = -1;
  bxyz[i] = ~0ull;
}

// CPDI stencil extent (static corners): max bounding-box width in nodes.
__global__ void __launch_bounds__(kBlock) k_cpdi_width(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                       const double* __restrict__ corners, int* __restrict__ wmax) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  if (S[P.sid[i]].rigid) return;
  CpdiW c;
  if (!cpdi_setup(corners, P.elem[0][i], P.elem[1][i], P.elem[2][i], P.elem[3][i], G, c)) return;
  int w = max(max(max(c.ci[0], c.ci[1]), max(c.ci[2], c.ci[3])) - min(min(c.ci[0], c.ci[1]), min(c.ci[2], c.ci[3])),
              max(max(c.cj[0], c.cj[1]), max(c.cj[2], c.cj[3])) - min(min(c.cj[0], c.cj[1]), min(c.cj[2], c.cj[3]))) +
          2;
  atomicMax(wmax, w);
}

// Exclusive scan, 1024 items per block (256 threads x 4), three passes.
constexpr int kScanItems = 1024;

// Exclusive scan over the block (kBlock threads): a shuffle scan per wave,
// then the wave totals through LDS.  lds: at least kBlock / 64 ints.
__device__ __forceinline__ int block_excl_scan(int v, int* lds, int& total) {
  constexpr int NW = kBlock / 64;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  int pre = 0;
  total = 0;
#pragma unroll
  for (int j = 0; j < NW; j++) {
    const int s = lds[j];
    if (j < w) pre += s;
    total += s;
  }
  __syncthreads();  // lds is free again
  return pre + x - v;
}

__global__ void __launch_bounds__(kBlock) k_scan_block(const int* __restrict__ in, int* __restrict__ out,
                                                       int* __restrict__ sums, int n) {
  __shared__ int lds[kBlock / 64];
  int base = blockIdx.x * kScanItems + threadIdx.x * 4;
  int v[4];
  for (int j = 0; j < 4; j++) v[j] = (base + j < n) ? in[base + j] : 0;
  int tsum = v[0] + v[1] + v[2] + v[3];
  int total;
  int off = block_excl_scan(tsum, lds, total);
  for (int j = 0; j < 4; j++) {
    if (base + j < n) out[base + j] = off;
    off += v[j];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the block sums in place.  Each thread
// scans a contiguous run of ceil(nb / kBlock) sums, then the runs' totals
// are scanned across the block.
__global__ void __launch_bounds__(kBlock) k_scan_sums(int* __restrict__ sums, int nb) {
  __shared__ int lds[kBlock / 64];
  const int per = (nb + kBlock - 1) / kBlock;
  const int lo = min(nb, int(threadIdx.x) * per), hi = min(nb, lo + per);
  int run = 0;
  for (int i = lo; i < hi; i++) run += sums[i];
  int total;
  int acc = block_excl_scan(run, lds, total);
  for (int i = lo; i < hi; i++) {
    const int v = sums[i];
    sums[i] = acc;
    acc += v;
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_add(int* __restrict__ out, const int* __restrict__ sums, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / kScanItems];
}

// Counting-sort scatter.  sorted[pos] = (bin << 32) | particle: carrying the
// bin lets a reader that advances through a bin's segment see where it ends
// from the entry itself.  Particles outside the grid go to the extra bin nn,
// so every entry of sorted[0, np) is written.
__global__ void __launch_bounds__(kBlock) k_scatter(const int* __restrict__ key, const int* __restrict__ start,
                                                    const int* __restrict__ rank,
                                                    unsigned long long* __restrict__ sorted, int np, int nn) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int k = key[i];
  if (k < 0) k = nn;
  int pos = start[k] + rank[i];
  sorted[pos] = ((unsigned long long)unsigned(k) << 32) | unsigned(i);
}

// Order each bin segment by particle index so the P2G sum order is fixed
// (the bin half of every entry in a segment is equal).
constexpr int kSortReg = 16;  // bins up to this size sort in registers
__global__ void __launch_bounds__(kBlock) k_binsort(const int* __restrict__ start,
                                                    unsigned long long* __restrict__ sorted, int nbins) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  int s = start[b], e = start[b + 1];
  if (e - s <= 1) return;
  if (e - s <= kSortReg) {
    // in registers: odd-even transposition sort, padded with the largest key
    // (about 8 particles per bin at 2 per cell and axis)
    const int n = e - s;
    unsigned long long r[kSortReg];
#pragma unroll
    for (int q = 0; q < kSortReg; q++) {
      r[q] = ~0ull;
      if (q < n) r[q] = sorted[s + q];
    }
#pragma unroll
    for (int round = 0; round < kSortReg; round++)
#pragma unroll
      for (int q = round & 1; q + 1 < kSortReg; q += 2) {
        const unsigned long long lo = r[q] < r[q + 1] ? r[q] : r[q + 1];
        const unsigned long long hi = r[q] < r[q + 1] ? r[q + 1] : r[q];
        r[q] = lo;
        r[q + 1] = hi;
      }
#pragma unroll
    for (int q = 0; q < kSortReg; q++)
      if (q < n) sorted[s + q] = r[q];
    return;
  }
  for (int i = s + 1; i < e; i++) {
    unsigned long long v = sorted[i];
    int j = i - 1;
    while (j >= s && sorted[j] > v) {
      sorted[j + 1] = sorted[j];
      j--;
    }
    sorted[j + 1] = v;
  }
}

// ------------------------------------------------------------------ P2G
// Weight N and gradient of node (ii,jj,kk) (1-based) for a particle at xp.
// QUIRK selects the getShapeFunctions variant (used by the MUSL
// re-projection): QBS boundary typing, cubic z-axis type-4 test on jd.
// INTERIOR: quadratic B-splines at a node interior on every axis (type 3).
template <int D, int B, bool QUIRK, bool GRAD, bool INTERIOR = false>
__device__ __forceinline__ void node_weight(const Particles& P, const GridDesc& G, const double* lp,
                                            const double* corners, int p, const double* xp, int ii, int jj, int kk,
                                            double& N, double* dN, bool& member) {
  member = true;
  if (B == JUMP_BASIS_CPDI_Q4) {
    CpdiW c;
    cpdi_setup(corners, P.elem[0][p], P.elem[1][p], P.elem[2][p], P.elem[3][p], G, c);
    member = cpdi_member(c, ii, jj);
    if (member) cpdi_node(c, G, ii, jj, N, dN);
    return;
  }
  Axis X = axis_of(G, 0), Y = axis_of(G, 1);
  double wx, dwx, wy, dwy;
  axis_weight<B, QUIRK, true, INTERIOR>(xp[0], 0, X, lp[0], ii, ii, wx, dwx);
  axis_weight<B, QUIRK, true, INTERIOR>(xp[1], 1, Y, lp[1], jj, jj, wy, dwy);
  if (D == 2) {
    if (GRAD) combine2<B>(X, Y, wx, dwx, wy, dwy, N, dN);
    else N = wx * wy;
  } else {
    Axis Z = axis_of(G, 2);
    double wz, dwz;
    int sel4 = (QUIRK && B == JUMP_BASIS_CUBIC_BSPLINE) ? jj : kk;
    axis_weight<B, QUIRK, true, INTERIOR>(xp[2], 2, Z, lp[2], kk, sel4, wz, dwz);
    if (GRAD) combine3<B>(X, Y, Z, wx, dwx, wy, dwy, wz, dwz, N, dN);
    else N = wx * wy * wz;
  }
}

// The y and z axis weights of the last particle a lane weighed against its
// node (3D force pass replay).  A replayed run is consecutive particles of
// one lattice row, whose y and z are equal, and the lanes of a wave (nodes
// of a grid row) walk their runs in step.
struct YZCache {
  double y = __builtin_nan(""), z = __builtin_nan("");  // NaN: nothing cached
  double wy = 0.0, dwy = 0.0, wz = 0.0, dwz = 0.0;
};
// node_weight for D = 3 (not CPDI, not GIMP: lp is not compared), reusing
// the cached y and z weights when every active lane of the wave has the
// particle y and z it cached; otherwise every lane forms them again.
// Either way each value is the one node_weight forms from the same inputs.
template <int B, bool QUIRK, bool INTERIOR>
__device__ __forceinline__ void node_weight_yz(const GridDesc& G, const double* lp, const double* xp, int ii, int jj,
                                               int kk, double& N, double* dN, YZCache& c) {
  static_assert(B != JUMP_BASIS_CPDI_Q4 && B != JUMP_BASIS_GIMP, "weights that depend only on the position");
  const Axis X = axis_of(G, 0), Y = axis_of(G, 1), Z = axis_of(G, 2);
  double wx, dwx;
  axis_weight<B, QUIRK, true, INTERIOR>(xp[0], 0, X, lp[0], ii, ii, wx, dwx);
  if (!__all(xp[1] == c.y && xp[2] == c.z)) {
    const int sel4 = (QUIRK && B == JUMP_BASIS_CUBIC_BSPLINE) ? jj : kk;
    axis_weight<B, QUIRK, true, INTERIOR>(xp[1], 1, Y, lp[1], jj, jj, c.wy, c.dwy);
    axis_weight<B, QUIRK, true, INTERIOR>(xp[2], 2, Z, lp[2], kk, sel4, c.wz, c.dwz);
    c.y = xp[1];
    c.z = xp[2];
  }
  combine3<B>(X, Y, Z, wx, dwx, c.wy, c.dwy, c.wz, c.dwz, N, dN);
}

// Largest stencil window per axis: the basis width, or for CPDI the widest
// corner bounding box (checked on the host).
template <int B> struct MaxW { static constexpr int W = Width<B>::W; };
template <> struct MaxW<JUMP_BASIS_CPDI_Q4> { static constexpr int W = 5; };
constexpr int kNoParticle = 0x7fffffff;
constexpr int kStaleHead = -1;  // a consumed bin whose next head is not loaded yet

// Particle fields one P2G pass reads, fetched as whole 16-byte slots.
// PASS 0: x, m, V, v, sigma (every slot); PASS 1 (re-projection): x, m, v
// (the first four slots in 3D, of the third only vx).
template <int D, int PASS> struct P2GFields {
  static constexpr int NF = PairLayout<D>::NF;
  static constexpr int NPAIR = PASS == 0 ? PairLayout<D>::NP : (D == 3 ? 4 : 3);
  __device__ static __forceinline__ void load(const Particles& P, unsigned q, double* v, int j0 = 0) {
#pragma unroll
    for (int j = j0; j < NPAIR; j++) {
      if (PASS == 1 && D == 3 && j == 2) {  // only vx of (V, vx)
        v[5] = ldu(reinterpret_cast<const double*>(P.pr[2]), 2 * q + 1);
        continue;
      }
      const double2 w = ldu(P.pr[j], q);
      v[2 * j] = w.x;
      if (2 * j + 1 < NF) v[2 * j + 1] = w.y;
    }
  }
};

// One particle's contribution to the node accumulators, in the reference's
// operation order (SolveExplicitSolidMechanics2D.jl:411-421;
// SolveExplicitSolidMechanics3D.jl:69-82 and :135-138).
// v: the particle's fields as P2GFields loads them.
// AFF: the node's grid may be an APIC step's (single 2D problems; batches
// refuse APIC), read from G.apic.
// yz: the lane's cached y and z weights (node_weight_yz), or null.
template <int D, int B, int PASS, bool AFF = false, bool INTERIOR = false, bool YZ = false>
__device__ __forceinline__ void p2g_add_v(const Particles& P, const GridDesc& G, const SolidDev* S,
                                          const double* corners, int p, const double* v, int ni, int nj, int nk,
                                          double& am, double* ap, double* af, YZCache* yz = nullptr) {
  double lp[3] = {0.0, 0.0, 0.0};
  if (B == JUMP_BASIS_GIMP) {
    const SolidDev& s = S[P.sid[p]];
    for (int k = 0; k < D; k++) lp[k] = s.lp[k];
  }
  double N, dN[3];
  bool member = true;
  if constexpr (YZ) node_weight_yz<B, PASS == 1, INTERIOR>(G, lp, v, ni + 1, nj + 1, nk + 1, N, dN, *yz);
  else node_weight<D, B, PASS == 1, PASS == 0, INTERIOR>(P, G, lp, corners, p, v, ni + 1, nj + 1, nk + 1, N, dN,
                                                         member);
  if (!member) return;
  const double fM = v[D];
  if (PASS == 0) {
    const double fV = v[D + 1];
    const double* vp = v + D + 2;
    const double* sg = v + 2 * D + 2;
    const double Nim = N * fM;
    am += Nim;
    if (AFF && D == 2 && G.apic) {
      // APIC: momentum of the affine field v_p + C_p (x_I - x_p) at the node
      const double e0 = (G.xmin + double(ni) * G.dx) - v[0], e1 = (G.ymin + double(nj) * G.dy) - v[1];
      const double C0 = P.C[0][p], C1 = P.C[1][p], C2 = P.C[2][p], C3 = P.C[3][p];
      const double a0 = C0 * e0 + C2 * e1, a1 = C1 * e0 + C3 * e1;
      ap[0] += Nim * (vp[0] + a0);
      ap[1] += Nim * (vp[1] + a1);
    } else {
      for (int k = 0; k < D; k++) ap[k] += Nim * vp[k];
    }
    if (D == 2) {
      const double b0 = AFF && G.bodyfield ? P.bf[0][p] : G.body[0];
      const double b1 = AFF && G.bodyfield ? P.bf[1][p] : G.body[1];
      af[0] += (-fV) * (sg[0] * dN[0] + sg[2] * dN[1]) + (fM * b0) * N;
      af[1] += (-fV) * (sg[2] * dN[0] + sg[1] * dN[1]) + (fM * b1) * N;
    } else {
      af[0] -= fV * (sg[0] * dN[0] + sg[3] * dN[1] + sg[4] * dN[2]);
      af[1] -= fV * (sg[3] * dN[0] + sg[1] * dN[1] + sg[5] * dN[2]);
      af[2] -= fV * (sg[4] * dN[0] + sg[5] * dN[1] + sg[2] * dN[2]);
    }
  } else {
    const double Nm = N * fM;
    for (int k = 0; k < D; k++) ap[k] += Nm * v[D + 2 + k];
  }
}

template <int D, int B, int PASS, bool AFF = false>
__device__ __forceinline__ void p2g_add(const Particles& P, const GridDesc& G, const SolidDev* S,
                                        const double* corners, int p, int ni, int nj, int nk, double& am,
                                        double* ap, double* af) {
  double v[P2GFields<D, PASS>::NF + 1];
  P2GFields<D, PASS>::load(P, unsigned(p), v);
  p2g_add_v<D, B, PASS, AFF>(P, G, S, corners, p, v, ni, nj, nk, am, ap, af);
}

// PASS 0: mass, momentum0, force, then p = p0 + f dt and the Dirichlet mask
//         (SolveExplicitSolidMechanics2D.jl:391-442; 3D.jl:55-105).
// PASS 1: MUSL re-projection momentum2 = sum N m v (3D.jl:135-138).
// One thread per grid node gathers every particle whose stencil covers the
// node: the particles of the W^d bins of its window (bin = stencil origin
// node; each bin's list is sorted by particle index).  No floating-point
// atomics: every node's sum has a fixed order, so runs are bitwise
// reproducible.
//
// ORDER = JUMP_ORDER_REFERENCE.  The reference adds each particle's
// contribution to a node in particle order (solids in order, then particle
// index: SolveExplicitSolidMechanics2D.jl:391-401).  The kernel enumerates
// the union of the window's bins in ascending index, so every node folds its
// contributions in exactly the reference's order and the grid equals the
// reference's bit for bit.  Enumeration: every window bin keeps, in LDS, its
// next unconsumed particle (its head) and that particle's position in
// `sorted`.  The smallest head is the next particle; consuming it advances
// only its own bin (one load of the following entry, whose bin half says
// whether the segment goes on).  After particle p, if p+1 lies in the window
// (its packed bin coordinates) it is next and its bin is known, so the scan
// over all heads runs once per run of consecutive indices, which
// lattice-built solids make long.
// ORDER = JUMP_ORDER_BINNED.  The window's bins in a fixed order, each bin's
// particles in index order: deterministic, no merge, a different rounding
// than the reference's fold.
constexpr int kP2GBlock = 128;
// Lanes per node of the 3D single-problem gathers in reference order
// (p2g_block's LPN): the force pass and the re-projection.
constexpr int kP2GLanes0 = 1;
constexpr int kP2GLanes1 = 2;
// The 3D replays (the force pass and the re-projection) reuse a lane's y
// and z weights across a run (YZCache, node_weight_yz).
constexpr bool kYZCache = true;
// Pairs the 3D force pass's replay loads a particle ahead: (x, y) (z, m)
// (V, vx) ...  Measured (ms per launch on the impact): 2 1.334, 3 1.327,
// 4 1.333, 5 1.345; all seven (before the weight cache) 1.46 -> 1.53.
constexpr int kPrefetchPairs = 3;
// Node order of the single 2D gathers: the active nodes compacted (true) or
// grid rows, for the contact step and the plain USL step.
// Measured: the two-rings USL gather (1025 x 513 nodes, thick rings) 0.609
// ms in rows, 0.415 compacted; the CPDI tube array's contact gather (thin
// rings on 2048 x 2048 nodes) 1.878 ms in rows, 2.150 compacted (a wave of
// compacted nodes spans several ring walls and touches more cache lines).
constexpr bool kContactCompact = false;
constexpr bool kUsl2dCompact = true;
// Work order of the P2G threads.  Only nodes whose window holds a particle
// ("active") have work.  With waves on 64-node grid rows, lanes past the
// solid's edge idle: about half the lane slots on the 4.2M-particle impact.
// COMPACT runs the gather over the active nodes compacted in row-major
// order (then the inactive ones, which only get zeros): nearly every lane
// works, but a wave can straddle two rows and then gathers from two places.
// The re-projection pass (few loads per pair, issue-bound) gains from it;
// the force pass (sixteen loads per pair) is bound by the cache lines its
// loads touch, where idle lanes cost nothing and straddling costs, so it
// keeps grid rows (measured: 2.6 ms rows, 3.9 ms compacted).
//   k_node_flags: flag[n] = 1 when the W^d window of node n has a particle
//                 (each window row of bins is one contiguous range of start)
//   k_node_order: order[pos] = n, actives first, both parts ascending
template <int D, bool MULTI>
__global__ void __launch_bounds__(kBlock) k_node_flags(GridDesc G0, const int* __restrict__ start0, int W,
                                                       int* __restrict__ flag, BatchDev Bt) {
  const int n0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int NN = MULTI ? Bt.NN : G0.nn;
  if (n0 > NN) return;
  if (n0 == NN) { flag[n0] = 0; return; }
  const int inst = MULTI ? inst_of_node(Bt, n0) : 0;
  const GridDesc& G = MULTI ? Bt.inst[inst].G : G0;
  const int off = MULTI ? Bt.node_off[inst] : 0;
  const int* start = start0 + off;
  const int n = n0 - off;
  const int ni = n % G.nx, nj = (n / G.nx) % G.ny, nk = (D == 3) ? n / G.nxy : 0;
  const int lo = max(ni - W + 1, 0);
  int any = 0;
  for (int bk = (D == 3 ? max(nk - W + 1, 0) : 0); bk <= nk; bk++)
    for (int bj = max(nj - W + 1, 0); bj <= nj; bj++) {
      const int row = G.nx * (bj + G.ny * bk);
      any |= ldu(start, unsigned(row + ni + 1)) - ldu(start, unsigned(row + lo));
    }
  flag[n0] = any != 0;
}
__global__ void __launch_bounds__(kBlock) k_node_order(const int* __restrict__ pos, int nn, int* __restrict__ order) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const int p = pos[n];
  const int nact = pos[nn];
  order[(pos[n + 1] - p) ? p : nact + n - p] = n;
}

// Calls visit(p) for every particle whose stencil covers node (ni,nj,nk):
// the particles of the window's W^d bins, in ascending index for
// JUMP_ORDER_REFERENCE (k-way merge of the bins, heads and cursors in the
// LDS arrays s_cur/s_head of this thread's column), bin after bin for
// JUMP_ORDER_BINNED.  start: the node's own instance's bins; off: that
// instance's first node; sorted/bxyz: all particles.
// DEFER: a consumed bin's new head is loaded when the run of consecutive
// indices ends (one load per bin the run touched, all issued together)
// instead of after every particle.  Pays where the visit is light: the 3D
// re-projection pass (1.62 -> 1.19 ms on the 4.2M-particle impact) and the
// 2D force pass (two rings, 1.8M particles: step 1.52 -> 1.26 ms).  The 3D
// force pass hides the per-particle load behind its work (2.41 vs 2.47 ms).
struct NoRuns {
  __device__ void operator()(int, int) const {}
};
// The 3D re-projection pass visits exactly the particles of the force
// pass's merge, in the same order (same bins, same windows), and a node's
// window holds the same particles from one step to the next unless one of
// them changed bin.  The force pass records each node's runs of consecutive
// particle indices; the re-projection replays them, and so does the next
// step's force pass at every node whose window no particle entered or left
// (k_bin stamps the others).  The merge is about 0.84 of the force pass's
// 1.97 ms on the 4.2M-particle impact.  A run is one 32-bit entry, first
// particle << 3 | (length - 1), longer runs taking several; entry r of node
// n is runs[r NN + n].  A node with more than kRunCap entries keeps the
// merge in both passes.  On the impact the plate's interior nodes hold 36
// entries (6 x 6 rows of 6 particles); nodes where the sphere has entered
// the crater reached 57 by step 400.  At 48 such nodes merged again in the
// re-projection and lengthened its tail (0.70 -> 0.93 ms after step 160).
constexpr int kRunCap = 64;
// 2D problems keep run lists up to this many nodes (8.6 GB of lists).
constexpr long long kRunNodes2D = 1ll << 25;
struct RunList {
  unsigned* runs;    // kRunCap planes of NN entries, by node
  int* nruns;        // per node: entries recorded (> kRunCap: overflow)
  const int* dirty;  // per node: the step stamp of its last window change
  int gen;           // this step's stamp
  int all;           // every node merges (the first step after an upload)
};
template <int D, int B, int ORDER, bool DEFER, int BLK, class Visit, class OnRun = NoRuns>
__device__ __forceinline__ void for_window_particles(const Particles& P, const GridDesc& G, const int* start,
                                                     const unsigned long long* sorted,
                                                     const unsigned long long* bxyz, int off, int inst, int W,
                                                     int ni, int nj, int nk, int (*s_cur)[BLK],
                                                     int (*s_head)[BLK], Visit&& visit,
                                                     OnRun&& on_run = OnRun{}, int slot = -1) {
  constexpr int MW = MaxW<B>::W;
  constexpr int NB = (D == 3) ? MW * MW * MW : MW * MW;
  // the thread's column of s_cur/s_head (lanes that share a node share it)
  const int t = slot < 0 ? int(threadIdx.x) : slot;
  if constexpr (ORDER == JUMP_ORDER_BINNED) {
    const int k0 = (D == 3) ? max(nk - W + 1, 0) : 0;
    for (int bk = k0; bk <= nk; bk++)
      for (int bj = max(nj - W + 1, 0); bj <= nj; bj++)
        for (int bi = max(ni - W + 1, 0); bi <= ni; bi++) {
          const int bin = bi + G.nx * (bj + G.ny * bk);
          const int e = ldu(start, unsigned(bin + 1));
          for (int q = ldu(start, unsigned(bin)); q < e; q++)
            visit(int(unsigned(ldu(sorted, unsigned(q)))));
        }
  } else {
    const int wx0 = ni - W + 1, wy0 = nj - W + 1, wz0 = (D == 3) ? nk - W + 1 : 0;  // window origin bin
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int a0 = b % MW, a1 = (b / MW) % MW, a2 = (D == 3) ? b / (MW * MW) : 0;
      int c = 0, h = kNoParticle;
      if (a0 < W && a1 < W && a2 < W && wx0 + a0 >= 0 && wy0 + a1 >= 0 && wz0 + a2 >= 0) {
        const int bin = (wx0 + a0) + G.nx * ((wy0 + a1) + G.ny * (wz0 + a2));
        c = ldu(start, unsigned(bin));
        if (c < ldu(start, unsigned(bin + 1))) h = int(unsigned(ldu(sorted, unsigned(c))));
      }
      s_cur[b][t] = c;
      s_head[b][t] = h;
    }
    while (true) {
      // the smallest head is the window's next particle in index order
      int best = kNoParticle;
#pragma unroll
      for (int b = 0; b < NB; b++) best = min(best, s_head[b][t]);
      if (best == kNoParticle) break;
      int b = 0;
#pragma unroll
      for (int k = NB - 1; k >= 0; k--)
        if (s_head[k][t] == best) b = k;
      int p = best;
      int len = 0;  // particles of this run visited so far
      while (true) {
        // p is the head of window bin b: consume it and advance the bin
        const int a0 = b % MW, a1 = (b / MW) % MW, a2 = (D == 3) ? b / (MW * MW) : 0;
        const unsigned bin = unsigned(off + (wx0 + a0) + G.nx * ((wy0 + a1) + G.ny * (wz0 + a2)));
        const int c = s_cur[b][t] + 1;
        s_cur[b][t] = c;
        // is p+1 in the window?  (bin coordinates packed 21 bits per axis;
        // excluded particles hold all ones)  The load that decides it is
        // issued before the visit, so its latency runs under the visit's
        // work: a run's critical path is this chain of loads (4.99 vs 5.33
        // ms per step on the 4.2M-particle impact with it after the visit).
        const int q = p + 1;
        const unsigned long long cc = q < P.n ? ldu(bxyz, unsigned(q)) : ~0ull;
        if (DEFER) {
          s_head[b][t] = kStaleHead;
          visit(p);
        } else {
          const unsigned long long e = ldu(sorted, unsigned(c));  // sorted[np] is a sentinel
          visit(p);
          s_head[b][t] = unsigned(e >> 32) == bin ? int(unsigned(e)) : kNoParticle;
        }
        const unsigned ux = unsigned(int(cc & 0x1fffffu) - wx0);
        const unsigned uy = unsigned(int((cc >> 21) & 0x1fffffu) - wy0);
        const unsigned uz = unsigned(int((cc >> 42) & 0x1fffffu) - wz0);
        // 2D: the third field is the batch instance and must be this node's
        len++;
        if (!(ux < unsigned(W) && uy < unsigned(W) && (D == 2 ? uz == unsigned(inst) : uz < unsigned(W))) ||
            cc == ~0ull)
          break;
        b = int(ux) + MW * (int(uy) + (D == 3 ? MW * int(uz) : 0));
        p = q;
      }
      on_run(best, len);  // the run: particles best .. best + len - 1
      if (DEFER) {
#pragma unroll
        for (int k = 0; k < NB; k++) {
          if (s_head[k][t] != kStaleHead) continue;
          const int a0 = k % MW, a1 = (k / MW) % MW, a2 = (D == 3) ? k / (MW * MW) : 0;
          const unsigned bin = unsigned(off + (wx0 + a0) + G.nx * ((wy0 + a1) + G.ny * (wz0 + a2)));
          const unsigned long long e = ldu(sorted, unsigned(s_cur[k][t]));
          s_head[k][t] = unsigned(e >> 32) == bin ? int(unsigned(e)) : kNoParticle;
        }
      }
    }
  }
}

template <bool V> struct BoolTag { static constexpr bool value = V; };

// The terms one particle adds to a 3D node's accumulators, each formed by
// the expression p2g_add_v adds, so that folding the terms in particle order
// gives the same bits: PASS 0 t = (N m, N m v, force term (subtracted)),
// PASS 1 t = N m v.
template <int B, int PASS> struct Terms3 { static constexpr int NT = PASS == 0 ? 7 : 3; };
template <int B, int PASS, bool INTERIOR = false, bool YZ = false>
__device__ __forceinline__ void p2g_terms3(const Particles& P, const GridDesc& G, const SolidDev* S,
                                           const double* corners, int p, const double* v, int ni, int nj, int nk,
                                           double* t, YZCache* yz = nullptr) {
  double lp[3] = {0.0, 0.0, 0.0};
  if (B == JUMP_BASIS_GIMP) {
    const SolidDev& s = S[P.sid[p]];
    for (int k = 0; k < 3; k++) lp[k] = s.lp[k];
  }
  double N, dN[3];
  bool member;
  if constexpr (YZ) node_weight_yz<B, PASS == 1, INTERIOR>(G, lp, v, ni + 1, nj + 1, nk + 1, N, dN, *yz);
  else node_weight<3, B, PASS == 1, PASS == 0, INTERIOR>(P, G, lp, corners, p, v, ni + 1, nj + 1, nk + 1, N, dN,
                                                         member);
  const double fM = v[3];
  if (PASS == 0) {
    const double fV = v[4];
    const double* vp = v + 5;
    const double* sg = v + 8;
    const double Nim = N * fM;
    t[0] = Nim;
    for (int k = 0; k < 3; k++) t[1 + k] = Nim * vp[k];
    t[4] = fV * (sg[0] * dN[0] + sg[3] * dN[1] + sg[4] * dN[2]);
    t[5] = fV * (sg[3] * dN[0] + sg[1] * dN[1] + sg[5] * dN[2]);
    t[6] = fV * (sg[4] * dN[0] + sg[5] * dN[1] + sg[2] * dN[2]);
  } else {
    const double Nm = N * fM;
    for (int k = 0; k < 3; k++) t[k] = Nm * v[5 + k];
  }
}
template <int PASS>
__device__ __forceinline__ void p2g_fold3(const double* t, double& am, double* ap, double* af) {
  if (PASS == 0) {
    am += t[0];
    for (int k = 0; k < 3; k++) ap[k] += t[1 + k];
    for (int k = 0; k < 3; k++) af[k] -= t[4 + k];
  } else {
    for (int k = 0; k < 3; k++) ap[k] += t[k];
  }
}
// A double from lane 2j (SEL 0) or 2j+1 (SEL 1) of this lane's pair, by DPP
// within the quad (quad_perm [0,0,2,2] or [1,1,3,3]).
template <int SEL> __device__ __forceinline__ double pair_lane(double x) {
  constexpr int ctrl = SEL == 0 ? 0xA0 : 0xF5;
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
  const int lo = __builtin_amdgcn_update_dpp(0, int(unsigned(b)), ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(unsigned(b >> 32)), ctrl, 0xF, 0xF, false);
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(unsigned(hi)) << 32) |
                                                     unsigned(lo)));
}
// Two lanes per node: lane 0 of the pair holds the terms of the earlier
// particle, lane 1 those of the next (if `two`); both lanes fold both, in
// that order, so both hold the node's sums.
template <int PASS>
__device__ __forceinline__ void p2g_fold_pair(const double* t, bool two, double& am, double* ap, double* af) {
  constexpr int NT = PASS == 0 ? 7 : 3;
  double a[NT], b[NT];
#pragma unroll
  for (int j = 0; j < NT; j++) {
    a[j] = pair_lane<0>(t[j]);
    b[j] = pair_lane<1>(t[j]);
  }
  p2g_fold3<PASS>(a, am, ap, af);
  if (two) p2g_fold3<PASS>(b, am, ap, af);
}

// The nodes of logical block L (BLK threads, one node each, or LPN = 2: two
// lanes per node): the body of the gather kernels below.
//
// LPN = 2 (3D, reference order).  A wave whose lanes sit on neighbouring
// nodes of a row, one lane each, reads at every visit particles a cell's
// worth of particles apart (2 on the impact's lattice), so each 16-byte
// field load touches twice the cache lines it uses.  With two lanes per
// node, lane h of the pair computes the terms of the node's particles
// h, h+2, h+4, ... of its sequence; the pair swaps them (DPP) and both lanes
// fold them in the sequence's order: every sum keeps the reference's order
// and bits, and a wave's loads read consecutive particles.
template <int D, int B, int PASS, int ORDER, bool COMPACT, bool MULTI, int BLK, int LPN = 1>
__device__ __forceinline__ void p2g_block(const Particles& P, const GridDesc& G0, const SolidDev* __restrict__ S,
                                          const double* __restrict__ corners, const int* __restrict__ start0,
                                          const unsigned long long* __restrict__ sorted,
                                          const unsigned long long* __restrict__ bxyz,
                                          const int* __restrict__ order, const int* __restrict__ nodepos,
                                          const GridArrays& g0, int Wrt, double dt0, const BatchDev& Bt,
                                          double alpha, const RunList& rl, int L, int (*s_cur)[BLK / LPN],
                                          int (*s_head)[BLK / LPN]) {
  static_assert(LPN == 1 || (LPN == 2 && D == 3 && !MULTI && ORDER == JUMP_ORDER_REFERENCE),
                "two lanes per node: the 3D single-problem gathers in reference order");
  const int t = threadIdx.x;
  const int slot = t / LPN, h = t % LPN;  // node of the block, lane within the node
  const int NN = MULTI ? Bt.NN : G0.nn;
  const int i = L * (BLK / LPN) + slot;
  if (i >= NN) return;
  const int ng = COMPACT ? ldu(order, unsigned(i)) : i;  // node in the (concatenated) node arrays
  // run replay: single problems in reference order, where the lists exist
  constexpr bool RUNS = !MULTI && ORDER == JUMP_ORDER_REFERENCE;
  const bool runs_on = RUNS && rl.runs != nullptr;
  // A node whose window held no particle when its runs were recorded, and
  // which no particle has entered since, still holds the zeros that both
  // passes wrote in the step that emptied it (or in the first step after an
  // upload, when every node merges and writes): skip the writes.  3.6M of
  // the impact's 4.2M nodes, 0.46 GB of writes per force pass.
  // (3D only: in 2D the rigid override writes grid nodes after the gather)
  const bool still_empty = D == 3 && runs_on && !rl.all && rl.nruns[ng] == 0 && rl.dirty[ng] != rl.gen;
  if (COMPACT && i >= ldu(nodepos, unsigned(NN))) {  // inactive node: no particle reaches it
    if (PASS == 1 && still_empty) return;
    if (PASS == 0) {
      g0.m[ng] = 0.0;
      for (int k = 0; k < D; k++) { g0.p0[k][ng] = 0.0; g0.p[k][ng] = 0.0; g0.f[k][ng] = 0.0; }
    } else {
      // (p2xy, pz2 are left alone: G2P reads them only where 1/m > 0, and
      // a node with mass has a particle in its window)
      for (int k = 0; k < D; k++) g0.p2[k][ng] = 0.0;
    }
    return;
  }
  // the node's own grid: the single grid, or its batch instance's
  const int inst = MULTI ? inst_of_node(Bt, ng) : 0;
  const GridDesc& G = MULTI ? Bt.inst[inst].G : G0;
  const int off = MULTI ? Bt.node_off[inst] : 0;
  const double dt = MULTI ? Bt.inst[inst].dt : dt0;
  const int* start = start0 + off;
  const GridArrays g = MULTI ? grid_view(g0, off) : g0;
  const int n = ng - off;
  const int ni = n % G.nx;
  const int nj = (n / G.nx) % G.ny;
  const int nk = (D == 3) ? n / G.nxy : 0;
  const int W = (B == JUMP_BASIS_CPDI_Q4) ? Wrt : Width<B>::W;
  double am = 0.0, ap[D], af[D];
  for (int k = 0; k < D; k++) { ap[k] = 0.0; af[k] = 0.0; }
  constexpr bool DEFER = PASS == 1 || D == 2;
  if (PASS == 0 && still_empty) return;
  constexpr bool AFF = D == 2 && PASS == 0 && !MULTI;
  constexpr int NF = P2GFields<D, PASS>::NF;
  int pend = -1;  // LPN 2: the first particle of a pair not summed yet
  auto visit = [&](int p) {
    if constexpr (LPN == 1) {
      p2g_add<D, B, PASS, AFF>(P, G, S, corners, p, ni, nj, nk, am, ap, af);
    } else {
      if (pend < 0) {
        pend = p;
        return;
      }
      const int mine = h ? p : pend;
      double v[NF + 1], tt[Terms3<B, PASS>::NT];
      P2GFields<D, PASS>::load(P, unsigned(mine), v);
      p2g_terms3<B, PASS>(P, G, S, corners, mine, v, ni, nj, nk, tt);
      p2g_fold_pair<PASS>(tt, true, am, ap, af);
      pend = -1;
    }
  };
  // the pair's last particle when the sequence has odd length (both lanes)
  auto flush = [&]() {
    if constexpr (LPN == 2) {
      if (pend >= 0) {
        double v[NF + 1], tt[Terms3<B, PASS>::NT];
        P2GFields<D, PASS>::load(P, unsigned(pend), v);
        p2g_terms3<B, PASS>(P, G, S, corners, pend, v, ni, nj, nk, tt);
        p2g_fold3<PASS>(tt, am, ap, af);
        pend = -1;
      }
    }
  };
  // Replay.  The re-projection loads the next particle's fields (and the
  // next run's entry) while the current particle is summed; the sums keep
  // their order (0.683 -> 0.651 ms on the impact).  The force pass loads
  // the next particle's first kPrefetchPairs pairs ahead, (x, y) (z, m),
  // which its weights wait on, and (V, vx) (1.463 -> 1.378 ms for the first
  // two; all 14 fields ahead: 1.46 -> 1.53 ms, 117 -> 140 VGPRs).
  auto replay_as = [&](auto inner, int nr) {
    constexpr bool INT = decltype(inner)::value;
    if constexpr (LPN == 2 && PASS == 1) {
      // this lane's particles of each run, the next one's fields (and the
      // next run's entry) loaded while the current one is summed; the y and
      // z weights of a run's row formed once (kYZCache)
      constexpr bool YZ = kYZCache && B != JUMP_BASIS_GIMP && B != JUMP_BASIS_CPDI_Q4;
      YZCache yz;
      unsigned e = nr > 0 ? rl.runs[n] : 0u;
      for (int r = 0; r < nr; r++) {
        const int p0 = int(e >> 3), len = int(e & 7u) + 1;
        if (r + 1 < nr) e = rl.runs[size_t(r + 1) * NN + n];
        double v[NF + 1], w[NF + 1];
        if (h < len) P2GFields<D, PASS>::load(P, unsigned(p0 + h), v);
        for (int k = 0; k < len; k += 2) {
          const int mine = k + h;
          double tt[Terms3<B, PASS>::NT];
          if (mine < len) {
            if (mine + 2 < len) P2GFields<D, PASS>::load(P, unsigned(p0 + mine + 2), w);
            p2g_terms3<B, PASS, INT, YZ>(P, G, S, corners, p0 + mine, v, ni, nj, nk, tt, &yz);
            if (mine + 2 < len)
#pragma unroll
              for (int q = 0; q < NF + 1; q++) v[q] = w[q];
          } else {
#pragma unroll
            for (int j = 0; j < Terms3<B, PASS>::NT; j++) tt[j] = 0.0;
          }
          p2g_fold_pair<PASS>(tt, k + 1 < len, am, ap, af);
        }
      }
      return;
    } else if constexpr (LPN == 2) {
      constexpr bool YZ = kYZCache && B != JUMP_BASIS_GIMP && B != JUMP_BASIS_CPDI_Q4;
      YZCache yz;
      for (int r = 0; r < nr; r++) {
        const unsigned e = rl.runs[size_t(r) * NN + n];
        const int p0 = int(e >> 3), len = int(e & 7u) + 1;
        double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);  // (x, y) (z, m) of this lane's next
        if (PASS == 0 && h < len) {
          a = ldu(P.pr[0], unsigned(p0 + h));
          b = ldu(P.pr[1], unsigned(p0 + h));
        }
        for (int k = 0; k < len; k += 2) {
          const int mine = k + h;
          double tt[Terms3<B, PASS>::NT];
          if (mine < len) {
            double v[NF + 1];
            if (PASS == 0) {
              v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
              P2GFields<D, PASS>::load(P, unsigned(p0 + mine), v, 2);
              if (mine + 2 < len) {
                a = ldu(P.pr[0], unsigned(p0 + mine + 2));
                b = ldu(P.pr[1], unsigned(p0 + mine + 2));
              }
            } else {
              P2GFields<D, PASS>::load(P, unsigned(p0 + mine), v);
            }
            p2g_terms3<B, PASS, INT, YZ>(P, G, S, corners, p0 + mine, v, ni, nj, nk, tt, &yz);
          } else {
#pragma unroll
            for (int j = 0; j < Terms3<B, PASS>::NT; j++) tt[j] = 0.0;
          }
          p2g_fold_pair<PASS>(tt, k + 1 < len, am, ap, af);
        }
      }
      return;
    } else {
      if constexpr (PASS == 0) {
        // 3D, weights of the position alone: the y and z weights of a run's
        // row are formed once (kYZCache)
        constexpr bool YZ = kYZCache && D == 3 && B != JUMP_BASIS_GIMP && B != JUMP_BASIS_CPDI_Q4;
        YZCache yz;
        constexpr int NPF = D == 3 ? kPrefetchPairs : 2;  // pairs loaded a particle ahead
        for (int r = 0; r < nr; r++) {
          const unsigned e = rl.runs[size_t(r) * NN + n];
          const int p0 = int(e >> 3), len = int(e & 7u) + 1;
          double2 pf[NPF];  // (x, y) (z, m) ...
#pragma unroll
          for (int j = 0; j < NPF; j++) pf[j] = ldu(P.pr[j], unsigned(p0));
          for (int k = 0; k < len; k++) {
            double v[NF + 1];
#pragma unroll
            for (int j = 0; j < NPF; j++) {
              v[2 * j] = pf[j].x;
              if (2 * j + 1 < NF) v[2 * j + 1] = pf[j].y;
            }
            P2GFields<D, PASS>::load(P, unsigned(p0 + k), v, NPF);
            if (k + 1 < len) {
#pragma unroll
              for (int j = 0; j < NPF; j++) pf[j] = ldu(P.pr[j], unsigned(p0 + k + 1));
            }
            p2g_add_v<D, B, PASS, AFF, INT, YZ>(P, G, S, corners, p0 + k, v, ni, nj, nk, am, ap, af, &yz);
          }
        }
        return;
      }
      unsigned e = nr > 0 ? rl.runs[n] : 0u;
      for (int r = 0; r < nr; r++) {
        const int p0 = int(e >> 3), len = int(e & 7u) + 1;
        if (r + 1 < nr) e = rl.runs[size_t(r + 1) * NN + n];
        double v[NF + 1], w[NF + 1];
        P2GFields<D, PASS>::load(P, unsigned(p0), v);
        for (int k = 0; k < len; k++) {
          if (k + 1 < len) P2GFields<D, PASS>::load(P, unsigned(p0 + k + 1), w);
          p2g_add_v<D, B, PASS, AFF, INT>(P, G, S, corners, p0 + k, v, ni, nj, nk, am, ap, af);
          if (k + 1 < len)
#pragma unroll
            for (int q = 0; q < NF + 1; q++) v[q] = w[q];
        }
      }
    }
  };
  // Nodes interior on every axis (quadratic B-splines, 3D: all but the two
  // outermost node layers of each face) replay with the interior weights
  // alone, with no test of the node type in the loop.
  auto replay = [&](int nr) {
    if constexpr (D == 3 && B == JUMP_BASIS_QUAD_BSPLINE) {
      if (qbs_type(ni + 1, G.nx, PASS == 1) == 3 && qbs_type(nj + 1, G.ny, PASS == 1) == 3 &&
          qbs_type(nk + 1, G.nz, PASS == 1) == 3) {
        replay_as(BoolTag<true>{}, nr);
        return;
      }
    }
    replay_as(BoolTag<false>{}, nr);
  };
  if (runs_on && PASS == 1 && rl.nruns[n] <= kRunCap) {
    replay(rl.nruns[n]);  // the runs this step's force pass left
  } else if (runs_on && PASS == 0 && !rl.all && rl.dirty[n] != rl.gen && rl.nruns[n] <= kRunCap) {
    replay(rl.nruns[n]);  // no particle entered or left the window since they were recorded
  } else if (runs_on && PASS == 0) {
    // merge, recording the runs (none for a node whose window is empty)
    // (the two lanes of a node record the same entries; lane 0 stores them)
    int nr = 0;
    for_window_particles<D, B, ORDER, DEFER>(P, G, start, sorted, bxyz, off, inst, W, ni, nj, nk, s_cur, s_head, visit,
                                             [&](int p0, int len) {
                                               while (len > 0) {
                                                 const int c = min(len, 8);
                                                 if (nr < kRunCap && h == 0)
                                                   rl.runs[size_t(nr) * NN + n] = (unsigned(p0) << 3) | unsigned(c - 1);
                                                 nr++;
                                                 p0 += c;
                                                 len -= c;
                                               }
                                             },
                                             slot);
    flush();
    if (h == 0) rl.nruns[n] = nr;
  } else {
    for_window_particles<D, B, ORDER, DEFER>(P, G, start, sorted, bxyz, off, inst, W, ni, nj, nk, s_cur, s_head,
                                             visit, NoRuns{}, slot);
    flush();
  }
  if (LPN == 2 && h != 0) return;  // both lanes hold the node's sums
  if (PASS == 0) {
    const uint8_t fx = g.fixed[n];
    g.m[n] = am;
    double q[D], pn[D];
    for (int k = 0; k < D; k++) {
      double p0 = ap[k];
      double pk = p0 + af[k] * dt;
      if (fx & (1u << k)) { p0 = 0.0; pk = 0.0; }
      g.p0[k][n] = p0;
      g.p[k][n] = pk;
      g.f[k][n] = af[k];
      q[k] = pk - alpha * p0;
      pn[k] = pk;
    }
    if (D == 3 && !MULTI) {
      g.mq[n] = make_double2(am > 0.0 ? 1.0 / am : 0.0, q[0]);
      g.qq[n] = make_double2(q[1], q[D - 1]);
      g.pxy[n] = make_double2(pn[0], pn[1]);
    }
  } else {
    for (int k = 0; k < D; k++) g.p2[k][n] = ap[k];
    if (D == 3 && !MULTI) {
      g.p2xy[n] = make_double2(ap[0], ap[1]);
      g.pz2[n] = make_double2(g.p[D - 1][n], ap[D - 1]);
    }
  }
}

// BLK: threads (nodes) per workgroup; the batched step lets the autotuner
// choose it (the single-problem steps use kP2GBlock).
template <int D, int B, int PASS, int ORDER, bool COMPACT, bool MULTI, int BLK = kP2GBlock, int LPN = 1>
__global__ void __launch_bounds__(BLK) k_p2g(Particles P, GridDesc G0, const SolidDev* __restrict__ S,
                                                   const double* __restrict__ corners, const int* __restrict__ start0,
                                                   const unsigned long long* __restrict__ sorted,
                                                   const unsigned long long* __restrict__ bxyz,
                                                   const int* __restrict__ order, const int* __restrict__ nodepos,
                                                   GridArrays g0, int Wrt, double dt0, BatchDev Bt,
                                                   double alpha, RunList rl) {
  constexpr int MW = MaxW<B>::W;
  constexpr int NB = (D == 3) ? MW * MW * MW : MW * MW;
  // Reference order: per window bin, the position in `sorted` of its next
  // unconsumed particle and that particle (kNoParticle when exhausted).
  __shared__ int s_cur[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][BLK / LPN];
  __shared__ int s_head[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][BLK / LPN];
  const int NN = MULTI ? Bt.NN : G0.nn;
  // XCD: the 3D re-projection pass deals its (compacted) blocks to the XCDs
  // in runs of kXcdChunk (1.115 -> 1.066 ms on the 4.2M-particle impact;
  // the force pass keeps the hardware order, it was slower with any run
  // length tried, 2.12 -> 2.19-3.47 ms, as with 8x4x4 node tiles, 2.9-3.8,
  // and again once it replayed runs, 1.47 -> 1.92 ms with runs of 16)
  constexpr bool XCD = D == 3 && !MULTI && COMPACT;
  // 3D: blocks in descending node order.  The costliest nodes (the crater,
  // where windows change and merge again, and the sphere's mixed windows)
  // sit at the top of the plate, the highest node ids; started last they
  // set the kernel's tail.  Both grids have exactly nlog (logical) blocks.
  constexpr bool REV = D == 3 && !MULTI;
  const int nlog = (NN + BLK / LPN - 1) / (BLK / LPN);
  const int L0 = XCD ? xcd_block(nlog) : int(blockIdx.x);
  if (L0 < 0) return;
  const int L = REV ? nlog - 1 - L0 : L0;
  p2g_block<D, B, PASS, ORDER, COMPACT, MULTI, BLK, LPN>(P, G0, S, corners, start0, sorted, bxyz, order, nodepos,
                                                          g0, Wrt, dt0, Bt, alpha, rl, L, s_cur, s_head);
}

// The batched force pass as a persistent kernel over a work queue: a grid
// sized to the resident workgroups, each taking the next item (BLK
// consecutive nodes of the compacted active-node order, all instances'
// nodes in instance order: an (instance, tile) item) from the counter
// `queue` (zeroed before the launch) until the items run out; every
// workgroup leaves the loop together when the counter passes the last item.
// Each node is summed as in k_p2g, so the results are the same bits.
template <int B, int ORDER, int BLK>
__global__ void __launch_bounds__(BLK) k_p2g_queue(Particles P, GridDesc G0, const SolidDev* __restrict__ S,
                                                   const double* __restrict__ corners, const int* __restrict__ start0,
                                                   const unsigned long long* __restrict__ sorted,
                                                   const unsigned long long* __restrict__ bxyz,
                                                   const int* __restrict__ order, const int* __restrict__ nodepos,
                                                   GridArrays g0, int Wrt, double dt0, BatchDev Bt,
                                                   int* __restrict__ queue) {
  constexpr int MW = MaxW<B>::W;
  constexpr int NB = MW * MW;
  __shared__ int s_cur[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][BLK];
  __shared__ int s_head[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][BLK];
  __shared__ int s_item;
  const int nlog = (Bt.NN + BLK - 1) / BLK;
  while (true) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int L = s_item;
    __syncthreads();  // every thread has read the item before the next is taken
    if (L >= nlog) break;
    p2g_block<2, B, 0, ORDER, true, true, BLK>(P, G0, S, corners, start0, sorted, bxyz, order, nodepos, g0, Wrt,
                                               dt0, Bt, 0.0, RunList{}, L, s_cur, s_head);
  }
}

// ------------------------------------------------------------------ contact (2D)
// P2G and grid update with multi-body contact, the algorithm of the FEMP
// contact solver (SolveExplicitFEMP2DContact.jl:189-268 per-solid P2G with
// normals sum(m_p grad N) at :237; :311-387 the correction) on the USL grid.
// One thread per node gathers its window's particles in index order (the
// solids' order); the first solid reaching the node takes slot 0, the next
// slot 1, a third is dropped and flagged.  Per slot: mass, momentum0, force
// (as the single-field pass) and normal; then p = p0 + f dt, the Dirichlet
// mask, v0 = p0/m, v = p/m.  At a node with two solids the common normal is
// n = (n0 - n1)/|n0 - n1| (slot 1 uses -n) and each slot whose velocity
// approaches the centre-of-mass velocity v_cm = (p0 + p1)/(m0 + m1) along
// its normal (D = (v - v_cm).n > 0) gets v -= D (n + mu'/|C| (n_y C, -n_x C)),
// C = (v - v_cm) x n, mu' = min(mu, |C|/D) (:366-381; no tangential term
// when C = 0).  The single-field arrays receive the slot totals.
//
// Rigid solids (SolveExplicitFEMP2DContact.jl:272-304, :330-358): a rigid
// particle reaches the four nodes of its hat cell, as in the rigid override
// of the ULMPM step, and adds only its normal sum(m_p grad N_hat) to its
// slot (no mass, momentum or force).  Where a rigid and a deformable solid
// meet, the common normal is the rigid body's own, normalised (the
// deformable slot uses -n), the centre-of-mass velocity is the rigid
// velocity, formed as (m_S v_r)/m_S from the deformable mass m_S as the
// reference overrides the system momentum, and only the deformable slot is
// corrected.  A rigid slot holds the rigid velocity (0 if fixed) as v and v0.
struct ContactSlot {
  int body = -1;
  bool rigid = false;
  double m = 0.0, p0[2] = {0.0, 0.0}, f[2] = {0.0, 0.0}, nr[2] = {0.0, 0.0};
};

// Hat weight and gradient of node (ii, jj) (1-based) for a particle at x;
// false when the node is not a corner of the particle's hat cell
// (Basis.jl:63-78, 219-251).
__device__ __forceinline__ bool hat_node(const GridDesc& G, const double* x, int ii, int jj, double& N, double* dN) {
  const Axis X = axis_of(G, 0), Y = axis_of(G, 1);
  int ci, cj;
  if (!first_node<JUMP_BASIS_LINEAR, 2>(x[0], 0, X, 0.0, ci) || !first_node<JUMP_BASIS_LINEAR, 2>(x[1], 1, Y, 0.0, cj))
    return false;
  if (ii - ci < 0 || ii - ci > 1 || jj - cj < 0 || jj - cj > 1) return false;
  double wx, dwx, wy, dwy;
  axis_weight<JUMP_BASIS_LINEAR, false>(x[0], 0, X, 0.0, ii, ii, wx, dwx);
  axis_weight<JUMP_BASIS_LINEAR, false>(x[1], 1, Y, 0.0, jj, jj, wy, dwy);
  combine2<JUMP_BASIS_LINEAR>(X, Y, wx, dwx, wy, dwy, N, dN);
  return true;
}

// BF: the body force is per particle (GridDesc::bodyfield), not the constant.
// COMPACT: threads take the nodes in the order active nodes first (order,
// nodepos: k_node_order); a node with an empty window skips the walk and
// writes what an empty walk leaves.
template <int B, int ORDER, bool BF = false, bool COMPACT = false>
__global__ void __launch_bounds__(kP2GBlock) k_p2g_contact(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                           const double* __restrict__ corners,
                                                           const int* __restrict__ start,
                                                           const unsigned long long* __restrict__ sorted,
                                                           const unsigned long long* __restrict__ bxyz, GridArrays g,
                                                           ContactArrays c, int Wrt, double dt, double fric,
                                                           unsigned* __restrict__ flags, RunList rl,
                                                           const int* __restrict__ order,
                                                           const int* __restrict__ nodepos) {
  constexpr int MW = MaxW<B>::W;
  constexpr int NB = MW * MW;
  __shared__ int s_cur[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][kP2GBlock];
  __shared__ int s_head[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][kP2GBlock];
  const int i = blockIdx.x * kP2GBlock + threadIdx.x;
  if (i >= G.nn) return;
  const int n = COMPACT ? ldu(order, unsigned(i)) : i;
  const bool active = !COMPACT || i < ldu(nodepos, unsigned(G.nn));
  const int ni = n % G.nx, nj = n / G.nx;
  const int W = (B == JUMP_BASIS_CPDI_Q4) ? Wrt : Width<B>::W;
  ContactSlot s0, s1;
  bool overflow = false;
  auto add = [&](ContactSlot& s, int sid, double N, const double* dN, const double* v, const double* bp) {
    const double fM = v[2], fV = v[3];
    const double* vp = v + 4;
    const double* sg = v + 6;
    const double Nim = N * fM;
    s.body = sid;
    s.m += Nim;
    s.p0[0] += Nim * vp[0];
    s.p0[1] += Nim * vp[1];
    s.f[0] += (-fV) * (sg[0] * dN[0] + sg[2] * dN[1]) + (fM * bp[0]) * N;
    s.f[1] += (-fV) * (sg[2] * dN[0] + sg[1] * dN[1]) + (fM * bp[1]) * N;
    s.nr[0] += fM * dN[0];
    s.nr[1] += fM * dN[1];
  };
  auto visit = [&](int p) {
    double v[PairLayout<2>::NF + 1];
    P2GFields<2, 0>::load(P, unsigned(p), v);
    double lp[3] = {0.0, 0.0, 0.0};
    const int sid = ldu(P.sid, unsigned(p));
    if (B == JUMP_BASIS_GIMP) {
      lp[0] = S[sid].lp[0];
      lp[1] = S[sid].lp[1];
    }
    double N, dN[3];
    if (S[sid].rigid) {
      if (!hat_node(G, v, ni + 1, nj + 1, N, dN)) return;
      ContactSlot* s = (s0.body < 0 || s0.body == sid) ? &s0 : ((s1.body < 0 || s1.body == sid) ? &s1 : nullptr);
      if (!s) {
        overflow = true;
        return;
      }
      s->body = sid;
      s->rigid = true;
      s->nr[0] += v[2] * dN[0];
      s->nr[1] += v[2] * dN[1];
      return;
    }
    bool member;
    node_weight<2, B, false, true>(P, G, lp, corners, p, v, ni + 1, nj + 1, 1, N, dN, member);
    if (!member) return;
    const double bp[2] = {BF ? P.bf[0][p] : G.body[0], BF ? P.bf[1][p] : G.body[1]};
    if (s0.body < 0 || s0.body == sid) add(s0, sid, N, dN, v, bp);
    else if (s1.body < 0 || s1.body == sid) add(s1, sid, N, dN, v, bp);
    else overflow = true;
  };
  // Run replay as in p2g_block: a node whose window no particle entered or
  // left since the last step visits the runs of particle indices it
  // recorded then, in the same order (CPDI bins follow the static corners,
  // so only the platens' particles ever change them).
  const bool runs_on = ORDER == JUMP_ORDER_REFERENCE && rl.runs != nullptr;
  if (!active) {
    // empty window: no particle to visit (its run list, if stale, is
    // recorded again when a particle enters and stamps the node)
  } else if (runs_on && !rl.all && rl.dirty[n] != rl.gen && rl.nruns[n] <= kRunCap) {
    const int nr = rl.nruns[n];
    for (int r = 0; r < nr; r++) {
      const unsigned e = rl.runs[size_t(r) * G.nn + n];
      const int p0 = int(e >> 3), len = int(e & 7u) + 1;
      for (int k = 0; k < len; k++) visit(p0 + k);
    }
  } else if (runs_on) {
    int nr = 0;
    for_window_particles<2, B, ORDER, true>(P, G, start, sorted, bxyz, 0, 0, W, ni, nj, 0, s_cur, s_head, visit,
                                            [&](int p0, int len) {
                                              while (len > 0) {
                                                const int cc = min(len, 8);
                                                if (nr < kRunCap)
                                                  rl.runs[size_t(nr) * G.nn + n] = (unsigned(p0) << 3) | unsigned(cc - 1);
                                                nr++;
                                                p0 += cc;
                                                len -= cc;
                                              }
                                            });
    rl.nruns[n] = nr;
  } else {
    for_window_particles<2, B, ORDER, true>(P, G, start, sorted, bxyz, 0, 0, W, ni, nj, 0, s_cur, s_head, visit);
  }
  if (overflow) atomicOr(flags, JUMP_FLAG_CONTACT_OVERFLOW);
  const uint8_t fx = g.fixed[n];
  double pk[2][2], v[2][2];
  ContactSlot* sl[2] = {&s0, &s1};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    ContactSlot& s = *sl[k];
#pragma unroll
    for (int a = 0; a < 2; a++) {
      pk[k][a] = s.p0[a] + s.f[a] * dt;
      if (fx & (1u << a)) { s.p0[a] = 0.0; pk[k][a] = 0.0; }
    }
    double v0[2] = {0.0, 0.0};
    v[k][0] = 0.0;
    v[k][1] = 0.0;
    if (s.rigid) {
      const SolidDev& r = S[s.body];
      v0[0] = v[k][0] = r.fixed ? 0.0 : r.vx;
      v0[1] = v[k][1] = r.fixed ? 0.0 : r.vy;
    } else if (s.body >= 0) {
      const double invM = 1.0 / s.m;
#pragma unroll
      for (int a = 0; a < 2; a++) {
        v0[a] = s.p0[a] * invM;
        v[k][a] = pk[k][a] * invM;
      }
    }
    c.body[k][n] = s.body;
    c.m[k][n] = s.m;
    c.v0[k][0][n] = v0[0];
    c.v0[k][1][n] = v0[1];
  }
  g.m[n] = s0.m + s1.m;
#pragma unroll
  for (int a = 0; a < 2; a++) {
    g.p0[a][n] = s0.p0[a] + s1.p0[a];
    g.p[a][n] = pk[0][a] + pk[1][a];
    g.f[a][n] = s0.f[a] + s1.f[a];
  }
  if (s0.body >= 0 && s1.body >= 0 && !(s0.rigid && s1.rigid)) {
    const double mS = s0.m + s1.m;
    double vcm[2], nc[2];
    if (s0.rigid || s1.rigid) {
      // the rigid body's velocity and normal; slot 1 rigid: slot 0 takes -n
      const ContactSlot& r = s0.rigid ? s0 : s1;
      const double vr[2] = {v[s0.rigid ? 0 : 1][0], v[s0.rigid ? 0 : 1][1]};
      vcm[0] = (mS * vr[0]) / mS;
      vcm[1] = (mS * vr[1]) / mS;
      const double nrm = sqrt(r.nr[0] * r.nr[0] + r.nr[1] * r.nr[1]);
      nc[0] = s0.rigid ? r.nr[0] / nrm : -(r.nr[0] / nrm);
      nc[1] = s0.rigid ? r.nr[1] / nrm : -(r.nr[1] / nrm);
    } else {
      vcm[0] = (pk[0][0] + pk[1][0]) / mS;
      vcm[1] = (pk[0][1] + pk[1][1]) / mS;
      nc[0] = s0.nr[0] - s1.nr[0];
      nc[1] = s0.nr[1] - s1.nr[1];
      const double nrm = sqrt(nc[0] * nc[0] + nc[1] * nc[1]);
      nc[0] = nc[0] / nrm;
      nc[1] = nc[1] / nrm;
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
      if (sl[k]->rigid) continue;  // a rigid body is never corrected
      const double nk[2] = {k == 0 ? nc[0] : -nc[0], k == 0 ? nc[1] : -nc[1]};
      const double dv[2] = {v[k][0] - vcm[0], v[k][1] - vcm[1]};
      const double Dn = dv[0] * nk[0] + dv[1] * nk[1];
      if (!(Dn > 0.0)) continue;
      const double C = dv[0] * nk[1] - dv[1] * nk[0];
      const double absC = C < 0.0 ? -C : C;
      const double mup = fric < absC / Dn ? fric : absC / Dn;
      double tx = 0.0, ty = 0.0;
      if (absC > 0.0) {
        const double cc = mup / absC;
        tx = cc * (nk[1] * C);
        ty = cc * (-(nk[0] * C));
      }
      v[k][0] = v[k][0] - Dn * (nk[0] + tx);
      v[k][1] = v[k][1] - Dn * (nk[1] + ty);
    }
  }
#pragma unroll
  for (int k = 0; k < 2; k++) {
    c.v[k][0][n] = v[k][0];
    c.v[k][1][n] = v[k][1];
  }
}

// ------------------------------------------------------------------ rigid (2D)
// SolveExplicitSolidMechanics2D.jl:448-479.  All particles of one rigid solid
// write the same function of the node mass, so concurrent writes agree.  The
// reference overrides solid after solid; a launch handles the rigid solids of
// one rank (the r-th rigid solid of every instance), ranks in order.
template <bool MULTI>
__global__ void __launch_bounds__(kBlock) k_rigid_override(Particles P, GridDesc G0, const SolidDev* __restrict__ S,
                                                           int rank, GridArrays g0, unsigned* __restrict__ flags,
                                                           BatchDev Bt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  if (!s.rigid || s.rigid_rank != rank) return;
  const GridDesc& G = MULTI ? Bt.inst[s.inst].G : G0;
  const GridArrays g = MULTI ? grid_view(g0, Bt.node_off[s.inst]) : g0;
  int f[2];
  bool ok = first_node<JUMP_BASIS_LINEAR, 2>(P.x[0][i], 0, axis_of(G, 0), 0.0, f[0]) &&
            first_node<JUMP_BASIS_LINEAR, 2>(P.x[1][i], 1, axis_of(G, 1), 0.0, f[1]);
  if (!ok) {
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  int b = (f[0] - 1) + G.nx * (f[1] - 1);
  int ids[4] = {b, b + 1, b + G.nx, b + 1 + G.nx};
  for (int a = 0; a < 4; a++) {
    int id = ids[a];
    double mi = g.m[id];
    if (s.fixed) {
      g.p[0][id] = 0.0; g.p[1][id] = 0.0;
      g.p0[0][id] = 0.0; g.p0[1][id] = 0.0;
    } else {
      if (s.vx != 0.0) { g.p[0][id] = mi * s.vx; g.p0[0][id] = mi * s.vx; }
      if (s.vy != 0.0) { g.p[1][id] = mi * s.vy; g.p0[1][id] = mi * s.vy; }
    }
  }
}

// SolveExplicitSolidMechanics2D.jl:539-551
template <bool MULTI>
__global__ void __launch_bounds__(kBlock) k_rigid_move(Particles P, const SolidDev* __restrict__ S, double dt0,
                                                       BatchDev Bt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  if (!s.rigid) return;
  const double dt = MULTI ? Bt.inst[s.inst].dt : dt0;
  P.x[0][i] = P.x[0][i] + dt * s.vx;
  P.x[1][i] = P.x[1][i] + dt * s.vy;
}

// ------------------------------------------------------------------ stencil walk
// Calls fn(node, N, dN) for every support node of particle p in the
// reference's node order: linear -- y (z) outer, x inner (Basis.jl:77,132);
// B-splines/GIMP -- x outer, z inner (Basis.jl:428-459); CPDI -- ascending
// node index.  Returns false if the stencil leaves the grid.
template <int D, int B, bool QUIRK, bool GRAD, class Fn>
__device__ __forceinline__ bool walk_stencil(const Particles& P, const GridDesc& G, const SolidDev& s,
                                             const double* corners, int p, Fn&& fn) {
  if (B == JUMP_BASIS_CPDI_Q4) {
    CpdiW c;
    if (!cpdi_setup(corners, P.elem[0][p], P.elem[1][p], P.elem[2][p], P.elem[3][p], G, c)) return false;
    int i0 = min(min(c.ci[0], c.ci[1]), min(c.ci[2], c.ci[3]));
    int i1 = max(max(c.ci[0], c.ci[1]), max(c.ci[2], c.ci[3])) + 1;
    int j0 = min(min(c.cj[0], c.cj[1]), min(c.cj[2], c.cj[3]));
    int j1 = max(max(c.cj[0], c.cj[1]), max(c.cj[2], c.cj[3])) + 1;
    for (int jj = j0; jj <= j1; jj++)
      for (int ii = i0; ii <= i1; ii++) {
        if (!cpdi_member(c, ii, jj)) continue;
        double N, dN[3];
        cpdi_node(c, G, ii, jj, N, dN);
        fn((ii - 1) + G.nx * (jj - 1), N, dN);
      }
    return true;
  }
  constexpr int W = Width<B>::W;
  Axis A[3] = {axis_of(G, 0), axis_of(G, 1), axis_of(G, 2)};
  int f[3] = {1, 1, 1};
  double w[3][W], dw[3][W];
  for (int a = 0; a < D; a++)
    if (!first_node<B, D>(P.x[a][p], a, A[a], s.lp[a], f[a])) return false;
  for (int a = 0; a < D; a++) {
    if (D == 3 && a == 2 && QUIRK && B == JUMP_BASIS_CUBIC_BSPLINE) break;  // z depends on jd, below
#pragma unroll
    for (int o = 0; o < W; o++) axis_weight<B, QUIRK>(P.x[a][p], a, A[a], s.lp[a], f[a] + o, f[a] + o, w[a][o], dw[a][o]);
  }
  if (D == 2) {
    if (B == JUMP_BASIS_LINEAR) {
#pragma unroll
      for (int b = 0; b < W; b++)
#pragma unroll
        for (int a = 0; a < W; a++) {
          double N, dN[3];
          if (GRAD) combine2<B>(A[0], A[1], w[0][a], dw[0][a], w[1][b], dw[1][b], N, dN);
          else N = w[0][a] * w[1][b];
          fn((f[0] + a - 1) + G.nx * (f[1] + b - 1), N, dN);
        }
    } else {
#pragma unroll
      for (int a = 0; a < W; a++)
#pragma unroll
        for (int b = 0; b < W; b++) {
          double N, dN[3];
          if (GRAD) combine2<B>(A[0], A[1], w[0][a], dw[0][a], w[1][b], dw[1][b], N, dN);
          else N = w[0][a] * w[1][b];
          fn((f[0] + a - 1) + G.nx * (f[1] + b - 1), N, dN);
        }
    }
  } else {
    if (B == JUMP_BASIS_LINEAR) {
#pragma unroll
      for (int c = 0; c < W; c++)
#pragma unroll
        for (int b = 0; b < W; b++)
#pragma unroll
          for (int a = 0; a < W; a++) {
            double N, dN[3];
            if (GRAD) combine3<B>(A[0], A[1], A[2], w[0][a], dw[0][a], w[1][b], dw[1][b], w[2][c], dw[2][c], N, dN);
            else N = w[0][a] * w[1][b] * w[2][c];
            fn((f[0] + a - 1) + G.nx * (f[1] + b - 1) + G.nxy * (f[2] + c - 1), N, dN);
          }
    } else {
#pragma unroll
      for (int a = 0; a < W; a++)
#pragma unroll
        for (int b = 0; b < W; b++) {
          if (QUIRK && B == JUMP_BASIS_CUBIC_BSPLINE) {
#pragma unroll
            for (int c = 0; c < W; c++)
              axis_weight<B, QUIRK>(P.x[2][p], 2, A[2], s.lp[2], f[2] + c, f[1] + b, w[2][c], dw[2][c]);
          }
#pragma unroll
          for (int c = 0; c < W; c++) {
            double N, dN[3];
            if (GRAD) combine3<B>(A[0], A[1], A[2], w[0][a], dw[0][a], w[1][b], dw[1][b], w[2][c], dw[2][c], N, dN);
            else N = w[0][a] * w[1][b] * w[2][c];
            fn((f[0] + a - 1) + G.nx * (f[1] + b - 1) + G.nxy * (f[2] + c - 1), N, dN);
          }
        }
    }
  }
  return true;
}

__device__ __forceinline__ bool finite3(double a, double b, double c) { return isfinite(a) && isfinite(b) && isfinite(c); }

// ------------------------------------------------------------------ G2P, USL 2D
// The particle update after the USL 2D gather (SolveExplicitSolidMechanics2D.jl:
// 506-532): new v and x, strain += dt D, F = F (I + L dt), V = det(F) V0,
// stress.  L = (L0, L1, L2, L3) = (dN_x v_x, dN_x v_y, dN_y v_x, dN_y v_y).
// The state after the step goes to Po: P itself, or with OOP another set of
// state arrays (the reverse mode's records), in which case the mass, which
// shares its pair with V, is carried over too.
template <bool OOP>
__device__ __forceinline__ void usl2d_particle_update_to(const Particles& P, const Particles& Po, const SolidDev& s,
                                                         int i, double dt, double v0, double v1, double x0, double x1,
                                                         double L0, double L1, double L2, double L3,
                                                         unsigned* flags) {
  Po.v[0][i] = v0; Po.v[1][i] = v1;
  Po.x[0][i] = x0; Po.x[1][i] = x1;
  if (OOP) Po.m[i] = P.m[i];
  // D = (L + L')/2, strain += dt D, F *= (I + L dt), J = det F, V = J V0
  const double D11 = 0.5 * (L0 + L0), D22 = 0.5 * (L3 + L3), D12 = 0.5 * (L2 + L1);
  double e[3] = {P.eps[0][i] + dt * D11, P.eps[1][i] + dt * D22, P.eps[2][i] + dt * D12};
  Po.eps[0][i] = e[0]; Po.eps[1][i] = e[1]; Po.eps[2][i] = e[2];
  const double A0 = 1.0 + L0 * dt, A1 = L1 * dt, A2 = L2 * dt, A3 = 1.0 + L3 * dt;
  const double F0 = P.F[0][i], F1 = P.F[1][i], F2 = P.F[2][i], F3 = P.F[3][i];
  double Fn[4];
  Fn[0] = F0 * A0 + F2 * A1;
  Fn[1] = F1 * A0 + F3 * A1;
  Fn[2] = F0 * A2 + F2 * A3;
  Fn[3] = F1 * A2 + F3 * A3;
  for (int k = 0; k < 4; k++) Po.F[k][i] = Fn[k];
  const double J = Fn[0] * Fn[3] - Fn[2] * Fn[1];
  Po.V[i] = J * P.V0[i];
  unsigned fl = 0;
  if (J < 0.0) fl |= JUMP_FLAG_NEGATIVE_J;
  double sg[3] = {P.sig[0][i], P.sig[1][i], P.sig[2][i]};
  switch (s.kind) {
    case JUMP_MAT_ELASTIC: hooke2(s, e, sg); break;
    case JUMP_MAT_NEOHOOKEAN: neohooke<2>(s, Fn, J, sg); break;
    case JUMP_MAT_ELASTOPLASTIC: {
      double ep[3] = {P.ep[0][i], P.ep[1][i], P.ep[2][i]};
      double al = P.alpha[i], vm;
      j2_2d(s, e, sg, ep, al, vm);
      Po.ep[0][i] = ep[0]; Po.ep[1][i] = ep[1]; Po.ep[2][i] = ep[2];
      Po.alpha[i] = al;
      Po.vm[i] = vm;
      break;
    }
    default: break;
  }
  Po.sig[0][i] = sg[0]; Po.sig[1][i] = sg[1]; Po.sig[2][i] = sg[2];
  if (!(isfinite(x0) && isfinite(x1) && isfinite(v0) && isfinite(v1))) fl |= JUMP_FLAG_NONFINITE;
  if (fl) atomicOr(flags, fl);
}

__device__ __forceinline__ void usl2d_particle_update(const Particles& P, const SolidDev& s, int i, double dt,
                                                      double v0, double v1, double x0, double x1, double L0,
                                                      double L1, double L2, double L3, unsigned* flags) {
  usl2d_particle_update_to<false>(P, P, s, i, dt, v0, v1, x0, x1, L0, L1, L2, L3, flags);
}
// The recorded state of a 2D particle the step leaves as it is (a rigid
// particle before k_rigid_move, one whose stencil leaves the grid), carried
// from P to Po.
__device__ __forceinline__ void carry_state2d(const Particles& P, const Particles& Po, int i) {
#pragma unroll
  for (int j = 0; j < PairLayout<2>::NP; j++) Po.pr[j][i] = P.pr[j][i];
#pragma unroll
  for (int k = 0; k < 4; k++) Po.F[k][i] = P.F[k][i];
#pragma unroll
  for (int k = 0; k < 3; k++) Po.eps[k][i] = P.eps[k][i];
}

// VortexBodyForce2D (BodyForce.jl:94-121) at every particle's position at
// the start of the step, time t: bodyforce(body, xx[ip], t)
// (SolveExplicitSolidMechanics2D.jl:408, :125), in the reference's
// operation order (x^2 = x*x, x^3 = x*x*x, x^4 = (x*x)*(x*x)).  X0: at the
// particle's position when it was added instead (material coordinates, in
// which the manufactured solution is written).  v: G, T, Ri, Ro, rho0, mu,
// xcenter, ycenter.
struct VortexDev {
  double v[8];
};
template <bool X0>
__global__ void __launch_bounds__(kBlock) k_body_vortex2d(Particles P, VortexDev vd, double t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  auto p2 = [](double x) { return x * x; };
  auto p3 = [](double x) { return x * x * x; };
  auto p4 = [](double x) { return (x * x) * (x * x); };
  const double G = vd.v[0], Tp = vd.v[1], Ri = vd.v[2], Ro = vd.v[3];
  const double rho0 = vd.v[4], mu = vd.v[5], xc = vd.v[6], yc = vd.v[7];
  const double x = (X0 ? P.X0[0][i] : P.x[0][i]) - xc, y = (X0 ? P.X0[1][i] : P.x[1][i]) - yc;
  const double r = sqrt(x * x + y * y);
  const double theta = atan2(y, x);
  const double R = (Ri + Ro) / 2;
  const double h = 1 - 8 * p2((r - R) / (Ri - Ro)) + 16 * p4((r - R) / (Ri - Ro));
  const double hp = (-16) * (r - R) / p2(Ri - Ro) + 16 * 4 * p3(r - R) / p4(Ri - Ro);
  const double hpp = -16 / p2(Ri - Ro) + 16 * 4 * 3 * p2(r - R) / p4(Ri - Ro);
  const double pi = M_PI;
  const double gg = G * sin(pi * t / Tp);
  const double gp = G * pi / Tp * cos(pi * t / Tp);
  const double gpp = -p2(pi) / p2(Tp) * gg;
  const double alpha = gg * h;
  const double mdr = mu / rho0;
  const double br = (mdr * (3 * gg * hp + r * gg * hpp) - r * gpp * h) * sin(alpha) +
                    (mdr * r * p2(gg * hp) - r * p2(gp * h)) * cos(alpha);
  const double bt = (-mdr * (3 * gg * hp + r * gg * hpp) + r * gpp * h) * cos(alpha) +
                    (mdr * r * p2(gg * hp) + r * p2(gp * h)) * sin(alpha);
  P.bf[0][i] = br * cos(theta) - bt * sin(theta);
  P.bf[1][i] = br * sin(theta) + bt * cos(theta);
}

// APIC G2P (Jiang et al., "The affine particle-in-cell method", ACM TOG
// 34(4) 2015) in the 2D USL step; the reference declares APIC and the
// affine matrix Cmat (Algorithm.jl:41, Solid.jl:82) without a step.  Over
// the nodes with m_I > tol, vI = p_I / m_I: v_p = sum N vI (PIC),
// B = sum (N vI) (x_I - x_p)^T, x and L as the USL gather; over the whole
// stencil D = sum (N (x_I - x_p)) (x_I - x_p)^T; C_p = B D^-1 by cofactors,
// or C_p = L for the linear hat (D singular on grid lines; B D^-1 equals
// sum vI grad N^T for multilinear weights).
template <int B>
__device__ __forceinline__ void g2p_apic2d(const Particles& P, const GridDesc& G, const SolidDev& s,
                                           const double* corners, const GridArrays& g, double tol, double dt, int i,
                                           unsigned* flags) {
  const double xp0 = P.x[0][i], xp1 = P.x[1][i];
  double v0 = 0.0, v1 = 0.0, x0 = xp0, x1 = xp1;
  double L0 = 0.0, L1 = 0.0, L2 = 0.0, L3 = 0.0;
  double B0 = 0.0, B1 = 0.0, B2 = 0.0, B3 = 0.0, D0 = 0.0, D1 = 0.0, D3 = 0.0;
  bool ok = walk_stencil<2, B, false, true>(P, G, s, corners, i, [&](int n, double Ni, const double* dN) {
    const double e0 = (G.xmin + double(n % G.nx) * G.dx) - xp0, e1 = (G.ymin + double(n / G.nx) * G.dy) - xp1;
    const double n0 = Ni * e0, n1 = Ni * e1;
    D0 += n0 * e0;
    D1 += n0 * e1;
    D3 += n1 * e1;
    const double mI = g.m[n];
    if (mI > tol) {
      const double invM = 1.0 / mI;
      const double vI0 = g.p[0][n] * invM, vI1 = g.p[1][n] * invM;
      const double w0 = Ni * vI0, w1 = Ni * vI1;
      v0 += w0;
      v1 += w1;
      x0 += Ni * vI0 * dt;
      x1 += Ni * vI1 * dt;
      L0 += dN[0] * vI0;
      L1 += dN[0] * vI1;
      L2 += dN[1] * vI0;
      L3 += dN[1] * vI1;
      B0 += w0 * e0;
      B1 += w1 * e0;
      B2 += w0 * e1;
      B3 += w1 * e1;
    }
  });
  if (!ok) {
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  double C[4];
  if (B == JUMP_BASIS_LINEAR) {
    C[0] = L0; C[1] = L1; C[2] = L2; C[3] = L3;
  } else {
    const double det = D0 * D3 - D1 * D1;
    const double i0 = D3 / det, i1 = (-D1) / det, i3 = D0 / det;
    C[0] = B0 * i0 + B2 * i1;
    C[1] = B1 * i0 + B3 * i1;
    C[2] = B0 * i1 + B2 * i3;
    C[3] = B1 * i1 + B3 * i3;
  }
  for (int k = 0; k < 4; k++) P.C[k][i] = C[k];
  usl2d_particle_update(P, s, i, dt, v0, v1, x0, x1, L0, L1, L2, L3, flags);
}

// SolveExplicitSolidMechanics2D.jl:485-534.  OOP (single problems, not
// APIC): the state after the step goes to Po instead of P (the reverse
// mode's rebuilt records), every particle's recorded state included.
template <int B, bool MULTI, bool OOP = false>
__global__ void __launch_bounds__(kBlock) k_g2p_usl2d(Particles P, GridDesc G0, const SolidDev* __restrict__ S,
                                                      const double* __restrict__ corners, GridArrays g0, double tol0,
                                                      double dt0, unsigned* __restrict__ flags, BatchDev Bt,
                                                      Particles Po) {
  static_assert(!(OOP && MULTI), "single problems only");
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  if (s.rigid) {
    if (OOP) carry_state2d(P, Po, i);  // then moved by k_rigid_move on Po
    return;
  }
  const GridDesc& G = MULTI ? Bt.inst[s.inst].G : G0;
  const GridArrays g = MULTI ? grid_view(g0, Bt.node_off[s.inst]) : g0;
  const double tol = MULTI ? Bt.inst[s.inst].tol : tol0;
  const double dt = MULTI ? Bt.inst[s.inst].dt : dt0;
  if constexpr (!MULTI && !OOP) {
    if (G.apic) {  // (batches refuse APIC)
      g2p_apic2d<B>(P, G, s, corners, g, tol, dt, i, flags);
      return;
    }
  }
  double v0 = P.v[0][i], v1 = P.v[1][i];
  double x0 = P.x[0][i], x1 = P.x[1][i];
  double L0 = 0.0, L1 = 0.0, L2 = 0.0, L3 = 0.0;
  bool ok = walk_stencil<2, B, false, true>(P, G, s, corners, i, [&](int n, double Ni, const double* dN) {
    const double mI = g.m[n];
    if (mI > tol) {
      const double invM = 1.0 / mI;
      const double pn0 = g.p[0][n], pn1 = g.p[1][n];
      const double vI0 = pn0 * invM, vI1 = pn1 * invM;
      v0 += Ni * (pn0 - g.p0[0][n]) * invM;
      v1 += Ni * (pn1 - g.p0[1][n]) * invM;
      x0 += Ni * vI0 * dt;
      x1 += Ni * vI1 * dt;
      L0 += dN[0] * vI0;
      L1 += dN[0] * vI1;
      L2 += dN[1] * vI0;
      L3 += dN[1] * vI1;
    }
  });
  if (!ok) {
    if (OOP) carry_state2d(P, Po, i);
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  usl2d_particle_update_to<OOP>(P, OOP ? Po : P, s, i, dt, v0, v1, x0, x1, L0, L1, L2, L3, flags);
}

// G2P with contact: each particle reads the slot of its own solid at every
// support node (SolveExplicitFEMP2DContact.jl:393-426 form: v += N (v_I -
// v0_I), x += N v_I dt), then the same particle update.
template <int B>
__global__ void __launch_bounds__(kBlock) k_g2p_usl2d_contact(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                              const double* __restrict__ corners, ContactArrays c,
                                                              double tol, double dt, unsigned* __restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int sid = P.sid[i];
  const SolidDev& s = S[sid];
  if (s.rigid) return;  // moved by k_rigid_move
  double v0 = P.v[0][i], v1 = P.v[1][i];
  double x0 = P.x[0][i], x1 = P.x[1][i];
  double L0 = 0.0, L1 = 0.0, L2 = 0.0, L3 = 0.0;
  bool ok = walk_stencil<2, B, false, true>(P, G, s, corners, i, [&](int n, double Ni, const double* dN) {
    const int k = c.body[0][n] == sid ? 0 : (c.body[1][n] == sid ? 1 : -1);
    if (k < 0) return;
    if (c.m[k][n] > tol) {
      const double vI0 = c.v[k][0][n], vI1 = c.v[k][1][n];
      v0 += Ni * (vI0 - c.v0[k][0][n]);
      v1 += Ni * (vI1 - c.v0[k][1][n]);
      x0 += Ni * vI0 * dt;
      x1 += Ni * vI1 * dt;
      L0 += dN[0] * vI0;
      L1 += dN[0] * vI1;
      L2 += dN[1] * vI0;
      L3 += dN[1] * vI1;
    }
  });
  if (!ok) {
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  usl2d_particle_update(P, s, i, dt, v0, v1, x0, x1, L0, L1, L2, L3, flags);
}

// ------------------------------------------------------------------ G2P, MUSL 3D
// G2P-1, SolveExplicitSolidMechanics3D.jl:109-134 (uses getShapeFunctions).
// OOP: the new velocity goes to Po (the reverse mode's records); a particle
// whose stencil leaves the grid is carried over by k_g2p2_musl.
template <int B, bool OOP>
__global__ void __launch_bounds__(kBlock) k_g2p1_musl(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                      const double* __restrict__ corners, GridArrays g, double alpha,
                                                      unsigned* __restrict__ flags, Particles Po) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  double vx[3] = {0.0, 0.0, 0.0};
  bool ok = walk_stencil<3, B, true, false>(P, G, s, corners, i, [&](int n, double Ni, const double*) {
    const double2 a = g.mq[n];  // (1/m, q0), 1/m = 0 without mass; q = p - alpha p0
    if (a.x > 0.0) {
      const double mii = a.x * Ni;
      const double2 b = g.qq[n];
      vx[0] += a.y * mii;
      vx[1] += b.x * mii;
      vx[2] += b.y * mii;
    }
  });
  if (!ok) {
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  double vn[3];
  for (int k = 0; k < 3; k++) {
    vn[k] = alpha * P.v[k][i] + vx[k];
    (OOP ? Po : P).v[k][i] = vn[k];
  }
  if (!finite3(vn[0], vn[1], vn[2])) atomicOr(flags, JUMP_FLAG_NONFINITE);  // final for the step
}

// The particle update of G2P-2 (SolveExplicitSolidMechanics3D.jl:186-211)
// from the gathered position xx and velocity gradient L[a,b] = dN_a v_b.
// The state after the step goes to Po: P itself, or with OOP another set of
// state arrays (the reverse mode's records), which then also take the mass
// (it shares its pair with z) and alpha and the plastic strain after every
// step, not only after a plastic one.
template <bool OOP>
__device__ __forceinline__ void musl3d_particle_update_to(const Particles& P, const Particles& Po, const SolidDev& s,
                                                          int i, double dt, const double* xx, const double* L,
                                                          unsigned* flags) {
  for (int k = 0; k < 3; k++) Po.x[k][i] = xx[k];
  if (OOP) Po.m[i] = P.m[i];
  // symmetric part: D[a,b] = 0.5 (L[a,b] + L[b,a])
  const double D[6] = {0.5 * (L[0] + L[0]), 0.5 * (L[4] + L[4]), 0.5 * (L[8] + L[8]),
                       0.5 * (L[3] + L[1]), 0.5 * (L[6] + L[2]), 0.5 * (L[7] + L[5])};
  double dD[6], e[6];
  for (int k = 0; k < 6; k++) {
    dD[k] = dt * D[k];
    e[k] = P.eps[k][i] + dD[k];
    Po.eps[k][i] = e[k];
  }
  double A[9];
  for (int k = 0; k < 9; k++) A[k] = L[k] * dt;
  A[0] = 1.0 + A[0]; A[4] = 1.0 + A[4]; A[8] = 1.0 + A[8];
  double F[9], Fn[9];
  for (int k = 0; k < 9; k++) F[k] = P.F[k][i];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) Fn[a + 3 * b] = F[a] * A[3 * b] + F[a + 3] * A[1 + 3 * b] + F[a + 6] * A[2 + 3 * b];
  for (int k = 0; k < 9; k++) Po.F[k][i] = Fn[k];
  // det = dot(col1, cross(col2, col3))
  const double c0 = Fn[4] * Fn[8] - Fn[5] * Fn[7];
  const double c1 = Fn[5] * Fn[6] - Fn[3] * Fn[8];
  const double c2 = Fn[3] * Fn[7] - Fn[4] * Fn[6];
  const double J = Fn[0] * c0 + Fn[1] * c1 + Fn[2] * c2;
  Po.V[i] = J * P.V0[i];
  unsigned fl = 0;
  if (J < 0.0) fl |= JUMP_FLAG_NEGATIVE_J;
  double sg[6];
  for (int k = 0; k < 6; k++) sg[k] = P.sig[k][i];
  switch (s.kind) {
    case JUMP_MAT_ELASTIC: {
      double vm;
      hooke3(s, e, sg, vm);
      Po.vm[i] = vm;
      break;
    }
    case JUMP_MAT_NEOHOOKEAN: neohooke<3>(s, Fn, J, sg); break;
    case JUMP_MAT_ELASTOPLASTIC: {
      double ep[6];
      for (int k = 0; k < 6; k++) ep[k] = P.ep[k][i];
      double al = P.alpha[i], vm;
      // an elastic step leaves ep and alpha as they are: no write-back in place
      const bool plastic = j2_3d(s, e, sg, ep, al, vm);
      if (OOP || plastic) {
        for (int k = 0; k < 6; k++) Po.ep[k][i] = ep[k];
        Po.alpha[i] = al;
      }
      Po.vm[i] = vm;
      break;
    }
    case JUMP_MAT_JOHNSONCOOK: {
      double al = P.alpha[i], ad = P.alphadot[i], vm;
      johnson_cook(s, e, dD, dt, sg, al, ad, vm);
      Po.alpha[i] = al;
      Po.alphadot[i] = ad;
      Po.vm[i] = vm;
      break;
    }
    default: break;
  }
  for (int k = 0; k < 6; k++) Po.sig[k][i] = sg[k];
  // (the velocity, final after G2P-1, is checked there)
  if (!finite3(xx[0], xx[1], xx[2])) fl |= JUMP_FLAG_NONFINITE;
  if (fl) atomicOr(flags, fl);
}

__device__ __forceinline__ void musl3d_particle_update(const Particles& P, const SolidDev& s, int i, double dt,
                                                       const double* xx, const double* L, unsigned* flags) {
  musl3d_particle_update_to<false>(P, P, s, i, dt, xx, L, flags);
}
// The recorded state of a 3D particle the step leaves as it is (its stencil
// leaves the grid), carried from P to Po.
__device__ __forceinline__ void carry_state3d(const Particles& P, const Particles& Po, int i) {
#pragma unroll
  for (int j = 0; j < PairLayout<3>::NP; j++) Po.pr[j][i] = P.pr[j][i];
#pragma unroll
  for (int k = 0; k < 9; k++) Po.F[k][i] = P.F[k][i];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    Po.eps[k][i] = P.eps[k][i];
    Po.ep[k][i] = P.ep[k][i];
  }
  Po.alpha[i] = P.alpha[i];
}

// G2P-2, SolveExplicitSolidMechanics3D.jl:161-211.  OOP: P holds the state
// before the step and Po the state after it (Po's velocity from
// k_g2p1_musl); a particle whose stencil leaves the grid is carried over.
template <int B, bool OOP>
__global__ void __launch_bounds__(kBlock) k_g2p2_musl(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                      const double* __restrict__ corners, GridArrays g, double dt,
                                                      unsigned* __restrict__ flags, Particles Po) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  double xx[3] = {P.x[0][i], P.x[1][i], P.x[2][i]};
  double L[9];
  for (int k = 0; k < 9; k++) L[k] = 0.0;
  bool ok = walk_stencil<3, B, false, true>(P, G, s, corners, i, [&](int n, double Ni, const double* dN) {
    const double mii = g.mq[n].x;  // 1/m, 0 without mass
    if (mii > 0.0) {
      const double2 a = g.p2xy[n], b = g.pz2[n], c2 = g.pxy[n];
      const double vIx = mii * a.x, vIy = mii * a.y, vIz = mii * b.y;
      const double c = Ni * mii * dt;
      xx[0] += c * c2.x;
      xx[1] += c * c2.y;
      xx[2] += c * b.x;
      // L[a,b] = dN_a v_b (Basis gradient index first, as SMatrix{3,3}(...) at :189-191)
      L[0] += dN[0] * vIx; L[1] += dN[1] * vIx; L[2] += dN[2] * vIx;
      L[3] += dN[0] * vIy; L[4] += dN[1] * vIy; L[5] += dN[2] * vIy;
      L[6] += dN[0] * vIz; L[7] += dN[1] * vIz; L[8] += dN[2] * vIz;
    }
  });
  if (!ok) {
    if (OOP) carry_state3d(P, Po, i);
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  musl3d_particle_update_to<OOP>(P, OOP ? Po : P, s, i, dt, xx, L, flags);
}

// ------------------------------------------------------------------ 2D MUSL
// SolveExplicitSolidMechanics2D.jl:25-329: MUSL in 2D with the blended
// (consistent) mass matrix B = 0.1 diag(m) + 0.9 M, M_IJ = sum m N_I N_J.
// The reference assembles M densely and solves with a dense LU, which
// bounds it to a few thousand nodes; here M is a (2R+1)^2-point stencil per
// node (R = stencil width - 1: a node couples to the nodes of the particles
// it shares) and B is solved by preconditioned conjugate gradients on the
// device.  The nodal velocities v0 = B \ p0 and v = v0 + B \ (dt f) feed
// both G2P passes; the Dirichlet mask zeroes only the nodal momenta (:184-
// 199), which the G2P passes do not read.
//   brow[k nn + n]: row n of B at offset k = (a+R) + (2R+1)(b+R), the
//                   coefficient of node (i+a, j+b)
//   rhs[c nn + n]:  p0x, p0y, dt fx, dt fy
//   mv[c nn + n]:   vx0, vy0, vx, vy
struct Musl2Arrays {
  double* brow;
  double* rhs;
  double* mv;
};
constexpr double kMuslEps = 0.1;  // the lumped share of the blend (:150)

// P2G: mass, momentum0, force as the USL pass, and the row of M, each node
// in the reference's particle order (the row entries accumulate
// (N_I m) N_J as massMatrix[id,idj] += Nim * Nj, :120-125).
// BF: the body force is per particle (GridDesc::bodyfield), not the constant.
template <int B, int ORDER, int R, bool BF = false>
__global__ void __launch_bounds__(kP2GBlock) k_p2g_musl2d(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                          const double* __restrict__ corners,
                                                          const int* __restrict__ start,
                                                          const unsigned long long* __restrict__ sorted,
                                                          const unsigned long long* __restrict__ bxyz, GridArrays g,
                                                          Musl2Arrays mu, int Wrt, double dt) {
  constexpr int MW = MaxW<B>::W;
  constexpr int NB = MW * MW;
  constexpr int K1 = 2 * R + 1;
  __shared__ int s_cur[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][kP2GBlock];
  __shared__ int s_head[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][kP2GBlock];
  const int n = blockIdx.x * kP2GBlock + threadIdx.x;
  if (n >= G.nn) return;
  const int ni = n % G.nx, nj = n / G.nx;
  const int W = (B == JUMP_BASIS_CPDI_Q4) ? Wrt : Width<B>::W;
  double am = 0.0, ap[2] = {0.0, 0.0}, af[2] = {0.0, 0.0};
  double mrow[K1 * K1];
#pragma unroll
  for (int k = 0; k < K1 * K1; k++) mrow[k] = 0.0;
  const Axis X = axis_of(G, 0), Y = axis_of(G, 1);
  for_window_particles<2, B, ORDER, true>(P, G, start, sorted, bxyz, 0, 0, W, ni, nj, 0, s_cur, s_head, [&](int p) {
    double v[PairLayout<2>::NF + 1];
    P2GFields<2, 0>::load(P, unsigned(p), v);
    double lp[3] = {0.0, 0.0, 0.0};
    if (B == JUMP_BASIS_GIMP) {
      const SolidDev& s = S[P.sid[p]];
      lp[0] = s.lp[0];
      lp[1] = s.lp[1];
    }
    double N, dN[3];
    bool member;
    node_weight<2, B, false, true>(P, G, lp, corners, p, v, ni + 1, nj + 1, 1, N, dN, member);
    if (!member) return;
    const double fM = v[2], fV = v[3];
    const double Nim = N * fM;
    am += Nim;
    ap[0] += Nim * v[4];
    ap[1] += Nim * v[5];
    const double b0 = BF ? P.bf[0][p] : G.body[0], b1 = BF ? P.bf[1][p] : G.body[1];
    af[0] += (-fV) * (v[6] * dN[0] + v[8] * dN[1]) + (fM * b0) * N;
    af[1] += (-fV) * (v[8] * dN[0] + v[7] * dN[1]) + (fM * b1) * N;
    if (B == JUMP_BASIS_CPDI_Q4) {
      CpdiW c;
      cpdi_setup(corners, P.elem[0][p], P.elem[1][p], P.elem[2][p], P.elem[3][p], G, c);
#pragma unroll
      for (int b = -R; b <= R; b++)
#pragma unroll
        for (int a = -R; a <= R; a++) {
          if (!cpdi_member(c, ni + a + 1, nj + b + 1)) continue;
          double NJ, dNJ[3];
          cpdi_node(c, G, ni + a + 1, nj + b + 1, NJ, dNJ);
          mrow[(a + R) + K1 * (b + R)] += Nim * NJ;
        }
    } else {
      // the particle's stencil: nodes fx..fx+W-1, fy..fy+W-1 (0-based)
      const unsigned long long cc = ldu(bxyz, unsigned(p));
      const int fx = int(cc & 0x1fffffu), fy = int((cc >> 21) & 0x1fffffu);
      double wx[K1], wy[K1], dw;
#pragma unroll
      for (int a = -R; a <= R; a++) {
        wx[a + R] = 0.0;
        wy[a + R] = 0.0;
        if (ni + a >= fx && ni + a < fx + W) axis_weight<B, false>(v[0], 0, X, lp[0], ni + a + 1, ni + a + 1, wx[a + R], dw);
        if (nj + a >= fy && nj + a < fy + W) axis_weight<B, false>(v[1], 1, Y, lp[1], nj + a + 1, nj + a + 1, wy[a + R], dw);
      }
#pragma unroll
      for (int b = -R; b <= R; b++)
#pragma unroll
        for (int a = -R; a <= R; a++) {
          if (ni + a < fx || ni + a >= fx + W || nj + b < fy || nj + b >= fy + W) continue;
          mrow[(a + R) + K1 * (b + R)] += Nim * (wx[a + R] * wy[b + R]);
        }
    }
  });
  const int nn = G.nn;
  // right-hand sides before the Dirichlet mask (:159-171)
  mu.rhs[n] = ap[0];
  mu.rhs[nn + n] = ap[1];
  mu.rhs[2 * nn + n] = dt * af[0];
  mu.rhs[3 * nn + n] = dt * af[1];
  // the blended row (:150-156)
  const double oneme = 1.0 - kMuslEps;
#pragma unroll
  for (int k = 0; k < K1 * K1; k++) {
    double b = oneme * mrow[k];
    if (k == R + K1 * R) b = am == 0.0 ? 1.0 : kMuslEps * am + b;
    mu.brow[size_t(k) * nn + n] = b;
  }
  // grid fields as the reference leaves them (:184-199)
  const uint8_t fx = g.fixed[n];
  g.m[n] = am;
  for (int k = 0; k < 2; k++) {
    double p0 = ap[k];
    double pk = p0 + af[k] * dt;
    if (fx & (1u << k)) { p0 = 0.0; pk = 0.0; }
    g.p0[k][n] = p0;
    g.p[k][n] = pk;
    g.f[k][n] = af[k];
    g.p2[k][n] = 0.0;
  }
}

// The blended mass matrix B X = RHS, four right-hand sides at once, is
// solved preconditioned with the lumped mass D = diag(m) (1 at massless
// nodes, whose rows are the identity).  With non-negative shape functions
// that sum to at most one, 0.1 D <= B <= D: x.Mx = sum_p m_p (N_p.x)^2
// <= sum_p m_p sum_I N_I x_I^2 = sum_I m_I x_I^2 (Cauchy-Schwarz) and M is
// positive semi-definite, so the spectrum of D^-1 B lies in [0.1, 1]
// whatever the particle layout.
constexpr int kChebBlock = 256;
// The iteration works on TX x TY node tiles (one node per thread), staging d
// over the tile and its halo in LDS (each d then feeds up to (2R+1)^2 rows):
// 32 x 8, or 64 x 4 above kChebWideNodes nodes, whose rows read whole
// wave-wide segments.  Measured (solve ms at 4k / 66k / 263k nodes): 32 x 8
// 0.41 / 0.54 / 1.26, 64 x 4 0.44 / 0.55 / 1.10.
constexpr int kChebWideNodes = 150000;
template <int TX, int TY> __host__ __device__ inline int cheb_tiles(const GridDesc& G) {
  static_assert(TX * TY == kChebBlock, "one node per thread");
  return ((G.nx + TX - 1) / TX) * ((G.ny + TY - 1) / TY);
}

// Chebyshev iteration (Saad, Iterative Methods for
// Sparse Linear Systems, 2nd ed., Algorithm 12.1, with the preconditioner
// D = diag(m)).  The spectrum of D^-1 B lies in [0.1, 1] (above), so with
// theta = 0.55, delta = 0.45, sigma = theta / delta:
//   x0 = 0, r0 = b, d0 = D^-1 r0 / theta, rho0 = 1 / sigma
//   x_k+1 = x_k + d_k,  r_k+1 = r_k - B d_k,  rho_k+1 = 1 / (2 sigma - rho_k)
//   d_k+1 = rho_k+1 rho_k d_k + (2 rho_k+1 / delta) D^-1 r_k+1
// and the B-norm of the error after k iterations is at most
// 2 q^k / (1 + q^2k) of the first, q = (sqrt 10 - 1) / (sqrt 10 + 1), for
// any particle layout.  The iteration count follows from that bound and
// the coefficients are formed on the host, so an iteration needs no inner
// product and no read-back: it is one launch with no reduction.  r and d
// alternate between two slots per launch, so no workgroup reads what
// another writes in the same launch; x is read and written at its owner's
// nodes only.
struct ChebArrays {
  double* dinv;  // 1/m, 1 at massless nodes (nn)
  double* x;     // four columns of nn each
  double* r;     // slot k at r + 4 nn k, likewise d
  double* d;
};
__global__ void __launch_bounds__(kBlock) k_cheb_init(int nn, Musl2Arrays mu, const double* __restrict__ gm,
                                                      ChebArrays ch, double itheta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const double mi = gm[i];
  const double di = mi > 0.0 ? 1.0 / mi : 1.0;
  ch.dinv[i] = di;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const size_t k = size_t(c) * nn + i;
    const double b = mu.rhs[k];
    ch.x[k] = 0.0;
    ch.r[k] = b;  // slot 0
    ch.d[k] = b * di * itheta;
  }
}

// (B v)_i for the four columns of v at node (ix, iy), v held in LDS with
// the node at (lx, ly).  The row's (2R+1)^2 coefficients are all loaded
// first (every row is stored whole), so their loads are in flight together;
// the sum then takes them in a fixed order (b outer, a inner), skipping
// neighbours outside the grid.
template <int R, int LX, int LY>
__device__ __forceinline__ void cheb_spmv(const GridDesc& G, const Musl2Arrays& mu, const double (*v)[LY][LX], int ix,
                                          int iy, int lx, int ly, double* y) {
  constexpr int K1 = 2 * R + 1;
  const int nn = G.nn, i = ix + G.nx * iy;
  double cf[K1 * K1];
#pragma unroll
  for (int k = 0; k < K1 * K1; k++) cf[k] = mu.brow[size_t(k) * nn + i];
#pragma unroll
  for (int c = 0; c < 4; c++) y[c] = 0.0;
#pragma unroll
  for (int b = -R; b <= R; b++) {
    if (iy + b < 0 || iy + b >= G.ny) continue;
#pragma unroll
    for (int a = -R; a <= R; a++) {
      if (ix + a < 0 || ix + a >= G.nx) continue;
      const double w = cf[(a + R) + K1 * (b + R)];
#pragma unroll
      for (int c = 0; c < 4; c++) y[c] += w * v[c][ly + b][lx + a];
    }
  }
}

// One iteration on the workgroup's tile: B d_k with d_k staged over the tile
// and its R-wide halo, then x, r_k+1 and d_k+1 = c1 d_k + c2 D^-1 r_k+1.
// Reads slot `slot` of r and d, writes the other.  The last iteration writes
// the nodal velocities instead: mv = (x0, x1, x0 + x2, x1 + x3).
template <int R, int TX, int TY>
__global__ void __launch_bounds__(kChebBlock) k_cheb_iter(GridDesc G, Musl2Arrays mu, ChebArrays ch, int slot,
                                                        double c1, double c2, int last) {
  constexpr int LX = TX + 2 * R, LY = TY + 2 * R;
  __shared__ double ds[4][LY][LX];
  const int t = threadIdx.x;
  const int nn = G.nn;
  const size_t n4 = 4 * size_t(nn);
  const double* d_o = ch.d + size_t(slot) * n4;
  const double* r_o = ch.r + size_t(slot) * n4;
  double* d_n = ch.d + size_t(slot ^ 1) * n4;
  double* r_n = ch.r + size_t(slot ^ 1) * n4;
  const int ntx = (G.nx + TX - 1) / TX;
  const int x0 = int(blockIdx.x % unsigned(ntx)) * TX, y0 = int(blockIdx.x / unsigned(ntx)) * TY;
  for (int e = t; e < LX * LY; e += kChebBlock) {
    const int lx = e % LX, ly = e / LX;
    const int gx = x0 + lx - R, gy = y0 + ly - R;
    const bool in = gx >= 0 && gx < G.nx && gy >= 0 && gy < G.ny;
    const int j = in ? gx + G.nx * gy : 0;
#pragma unroll
    for (int c = 0; c < 4; c++) ds[c][ly][lx] = in ? d_o[size_t(c) * nn + j] : 0.0;
  }
  __syncthreads();
  const int tx = t % TX, ty = t / TX;
  const int ix = x0 + tx, iy = y0 + ty;
  if (ix >= G.nx || iy >= G.ny) return;
  const int i = ix + G.nx * iy;
  double y[4];
  cheb_spmv<R, LX, LY>(G, mu, ds, ix, iy, tx + R, ty + R, y);
  const double di = ch.dinv[i];
  double xn[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const size_t q = size_t(c) * nn + i;
    const double dk = ds[c][ty + R][tx + R];
    xn[c] = ch.x[q] + dk;
    const double rn = r_o[q] - y[c];
    if (!last) {
      ch.x[q] = xn[c];
      r_n[q] = rn;
      d_n[q] = c1 * dk + c2 * (rn * di);
    }
  }
  if (last) {
    mu.mv[i] = xn[0];
    mu.mv[size_t(nn) + i] = xn[1];
    mu.mv[2 * size_t(nn) + i] = xn[0] + xn[2];
    mu.mv[3 * size_t(nn) + i] = xn[1] + xn[3];
  }
}

// G2P-1 (getShapeFunctions, :203-226): v = alpha v + sum N (v_I - alpha v0_I)
// over nodes with mass.
template <int B>
__global__ void __launch_bounds__(kBlock) k_g2p1_musl2d(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                        const double* __restrict__ corners, GridArrays g,
                                                        Musl2Arrays mu, double alpha, unsigned* __restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  const int nn = G.nn;
  double vx = 0.0, vy = 0.0;
  bool ok = walk_stencil<2, B, true, false>(P, G, s, corners, i, [&](int n, double Ni, const double*) {
    if (g.m[n] > 0.0) {
      vx += Ni * (mu.mv[2 * nn + n] - alpha * mu.mv[n]);
      vy += Ni * (mu.mv[3 * nn + n] - alpha * mu.mv[nn + n]);
    }
  });
  if (!ok) {
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  P.v[0][i] = alpha * P.v[0][i] + vx;
  P.v[1][i] = alpha * P.v[1][i] + vy;
}

// G2P-2 (getShapeAndGradient, :246-286): x += N v_I dt, L from v_I over
// nodes with mass, then the USL particle update.
template <int B>
__global__ void __launch_bounds__(kBlock) k_g2p2_musl2d(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                        const double* __restrict__ corners, GridArrays g,
                                                        Musl2Arrays mu, double dt, unsigned* __restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SolidDev& s = S[P.sid[i]];
  const int nn = G.nn;
  double x0 = P.x[0][i], x1 = P.x[1][i];
  double L0 = 0.0, L1 = 0.0, L2 = 0.0, L3 = 0.0;
  bool ok = walk_stencil<2, B, false, true>(P, G, s, corners, i, [&](int n, double Ni, const double* dN) {
    if (g.m[n] > 0.0) {
      const double vI0 = mu.mv[2 * nn + n], vI1 = mu.mv[3 * nn + n];
      x0 += Ni * vI0 * dt;
      x1 += Ni * vI1 * dt;
      L0 += dN[0] * vI0;
      L1 += dN[0] * vI1;
      L2 += dN[1] * vI0;
      L3 += dN[1] * vI1;
    }
  });
  if (!ok) {
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  usl2d_particle_update(P, s, i, dt, P.v[0][i], P.v[1][i], x0, x1, L0, L1, L2, L3, flags);
}

// ------------------------------------------------------------------ contact (3D)
// The MUSL step with multi-body contact: per-solid node slots as in 2D; the
// normal-only correction of SolveExplicitFEMP3DContact.jl:398-511 (v -= D n
// for a slot approaching the centre-of-mass velocity along its normal, n =
// (n0 - n1)/|n0 - n1| from the slots' sums of m grad N, -n for slot 1) at
// the grid update; G2P-1, the re-projection and G2P-2 each on the particle's
// own solid's slot.  The FEMP solver's gap test on its mesh's surface
// positions has no counterpart on MPM solids and is not applied.
struct ContactSlot3 {
  int body = -1;
  double m = 0.0, p0[3] = {0.0, 0.0, 0.0}, f[3] = {0.0, 0.0, 0.0}, nr[3] = {0.0, 0.0, 0.0};
};

template <int B, int ORDER>
__global__ void __launch_bounds__(kP2GBlock) k_p2g_contact3d(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                             const int* __restrict__ start,
                                                             const unsigned long long* __restrict__ sorted,
                                                             const unsigned long long* __restrict__ bxyz,
                                                             GridArrays g, ContactArrays3 c, double dt,
                                                             unsigned* __restrict__ flags) {
  constexpr int MW = MaxW<B>::W;
  constexpr int NB = MW * MW * MW;
  __shared__ int s_cur[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][kP2GBlock];
  __shared__ int s_head[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][kP2GBlock];
  const int n = blockIdx.x * kP2GBlock + threadIdx.x;
  if (n >= G.nn) return;
  const int ni = n % G.nx, nj = (n / G.nx) % G.ny, nk = n / G.nxy;
  ContactSlot3 s0, s1;
  bool overflow = false;
  auto add = [&](ContactSlot3& s, int sid, double N, const double* dN, const double* v) {
    const double fM = v[3], fV = v[4];
    const double* vp = v + 5;
    const double* sg = v + 8;
    const double Nim = N * fM;
    s.body = sid;
    s.m += Nim;
    for (int a = 0; a < 3; a++) s.p0[a] += Nim * vp[a];
    s.f[0] -= fV * (sg[0] * dN[0] + sg[3] * dN[1] + sg[4] * dN[2]);
    s.f[1] -= fV * (sg[3] * dN[0] + sg[1] * dN[1] + sg[5] * dN[2]);
    s.f[2] -= fV * (sg[4] * dN[0] + sg[5] * dN[1] + sg[2] * dN[2]);
    for (int a = 0; a < 3; a++) s.nr[a] += fM * dN[a];
  };
  for_window_particles<3, B, ORDER, false>(P, G, start, sorted, bxyz, 0, 0, Width<B>::W, ni, nj, nk, s_cur, s_head,
                                           [&](int p) {
    double v[PairLayout<3>::NF + 1];
    P2GFields<3, 0>::load(P, unsigned(p), v);
    double lp[3] = {0.0, 0.0, 0.0};
    const int sid = ldu(P.sid, unsigned(p));
    if (B == JUMP_BASIS_GIMP)
      for (int a = 0; a < 3; a++) lp[a] = S[sid].lp[a];
    double N, dN[3];
    bool member;
    node_weight<3, B, false, true>(P, G, lp, nullptr, p, v, ni + 1, nj + 1, nk + 1, N, dN, member);
    if (s0.body < 0 || s0.body == sid) add(s0, sid, N, dN, v);
    else if (s1.body < 0 || s1.body == sid) add(s1, sid, N, dN, v);
    else overflow = true;
  });
  if (overflow) atomicOr(flags, JUMP_FLAG_CONTACT_OVERFLOW);
  const uint8_t fx = g.fixed[n];
  double pk[2][3];
  ContactSlot3* sl[2] = {&s0, &s1};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    ContactSlot3& s = *sl[k];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      pk[k][a] = s.p0[a] + s.f[a] * dt;
      if (fx & (1u << a)) { s.p0[a] = 0.0; pk[k][a] = 0.0; }
    }
  }
  g.m[n] = s0.m + s1.m;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    g.p0[a][n] = s0.p0[a] + s1.p0[a];
    g.p[a][n] = pk[0][a] + pk[1][a];
    g.f[a][n] = s0.f[a] + s1.f[a];
  }
  if (s0.body >= 0 && s1.body >= 0) {
    const double mS = s0.m + s1.m;
    double vcm[3], nc[3];
#pragma unroll
    for (int a = 0; a < 3; a++) vcm[a] = (pk[0][a] + pk[1][a]) / mS;
#pragma unroll
    for (int a = 0; a < 3; a++) nc[a] = s0.nr[a] - s1.nr[a];
    const double nrm = sqrt(nc[0] * nc[0] + nc[1] * nc[1] + nc[2] * nc[2]);
#pragma unroll
    for (int a = 0; a < 3; a++) nc[a] = nc[a] / nrm;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const double m = sl[k]->m;
      double v[3], nk[3];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        v[a] = pk[k][a] / m;
        nk[a] = k == 0 ? nc[a] : -nc[a];
      }
      const double Dn = (v[0] - vcm[0]) * nk[0] + (v[1] - vcm[1]) * nk[1] + (v[2] - vcm[2]) * nk[2];
      if (!(Dn > 0.0)) continue;
#pragma unroll
      for (int a = 0; a < 3; a++) pk[k][a] = (v[a] - Dn * nk[a]) * m;
    }
  }
#pragma unroll
  for (int k = 0; k < 2; k++) {
    c.body[k][n] = sl[k]->body;
    c.m[k][n] = sl[k]->m;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      c.p0[k][a][n] = sl[k]->p0[a];
      c.p[k][a][n] = pk[k][a];
    }
  }
}

// Slot of solid sid at node n (-1: the solid was dropped there).
__device__ __forceinline__ int contact_slot(const ContactArrays3& c, int n, int sid) {
  return c.body[0][n] == sid ? 0 : (c.body[1][n] == sid ? 1 : -1);
}

// G2P-1 on the own slot (SolveExplicitSolidMechanics3D.jl:109-134).
template <int B>
__global__ void __launch_bounds__(kBlock) k_g2p1_musl_contact(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                              ContactArrays3 c, double alpha,
                                                              unsigned* __restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int sid = P.sid[i];
  const SolidDev& s = S[sid];
  double vx[3] = {0.0, 0.0, 0.0};
  bool ok = walk_stencil<3, B, true, false>(P, G, s, nullptr, i, [&](int n, double Ni, const double*) {
    const int k = contact_slot(c, n, sid);
    if (k < 0) return;
    const double mI = (k ? c.m[1] : c.m[0])[n];
    if (mI > 0.0) {
      const double mii = (1.0 / mI) * Ni;
#pragma unroll
      for (int a = 0; a < 3; a++)
        vx[a] += ((k ? c.p[1][a] : c.p[0][a])[n] - alpha * (k ? c.p0[1][a] : c.p0[0][a])[n]) * mii;
    }
  });
  if (!ok) {
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  double vn[3];
  for (int a = 0; a < 3; a++) {
    vn[a] = alpha * P.v[a][i] + vx[a];
    P.v[a][i] = vn[a];
  }
  if (!finite3(vn[0], vn[1], vn[2])) atomicOr(flags, JUMP_FLAG_NONFINITE);  // final for the step
}

// Re-projection per slot (SolveExplicitSolidMechanics3D.jl:135-138): the
// node gathers its window's particles in index order, each into its own
// solid's slot (as PASS 1 of k_p2g); the slot totals go to g.p2.
template <int B, int ORDER>
__global__ void __launch_bounds__(kP2GBlock) k_p2g2_contact3d(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                              const int* __restrict__ start,
                                                              const unsigned long long* __restrict__ sorted,
                                                              const unsigned long long* __restrict__ bxyz,
                                                              GridArrays g, ContactArrays3 c) {
  constexpr int MW = MaxW<B>::W;
  constexpr int NB = MW * MW * MW;
  __shared__ int s_cur[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][kP2GBlock];
  __shared__ int s_head[ORDER == JUMP_ORDER_REFERENCE ? NB : 1][kP2GBlock];
  const int n = blockIdx.x * kP2GBlock + threadIdx.x;
  if (n >= G.nn) return;
  const int ni = n % G.nx, nj = (n / G.nx) % G.ny, nk = n / G.nxy;
  const int b0 = c.body[0][n], b1 = c.body[1][n];
  double a0x = 0.0, a0y = 0.0, a0z = 0.0, a1x = 0.0, a1y = 0.0, a1z = 0.0;
  for_window_particles<3, B, ORDER, true>(P, G, start, sorted, bxyz, 0, 0, Width<B>::W, ni, nj, nk, s_cur, s_head,
                                          [&](int p) {
    const int sid = ldu(P.sid, unsigned(p));
    if (sid != b0 && sid != b1) return;  // dropped at this node (overflow)
    double q[3] = {0.0, 0.0, 0.0};
    double am = 0.0, af[3];
    p2g_add<3, B, 1>(P, G, S, nullptr, p, ni, nj, nk, am, q, af);
    // both sums formed, one kept: selects keep the six sums in registers
    const bool first = sid == b0;
    a0x = first ? a0x + q[0] : a0x;
    a0y = first ? a0y + q[1] : a0y;
    a0z = first ? a0z + q[2] : a0z;
    a1x = first ? a1x : a1x + q[0];
    a1y = first ? a1y : a1y + q[1];
    a1z = first ? a1z : a1z + q[2];
  });
  c.p2[0][0][n] = a0x; c.p2[0][1][n] = a0y; c.p2[0][2][n] = a0z;
  c.p2[1][0][n] = a1x; c.p2[1][1][n] = a1y; c.p2[1][2][n] = a1z;
  g.p2[0][n] = a0x + a1x;
  g.p2[1][n] = a0y + a1y;
  g.p2[2][n] = a0z + a1z;
}

// G2P-2 on the own slot (SolveExplicitSolidMechanics3D.jl:161-211).
template <int B>
__global__ void __launch_bounds__(kBlock) k_g2p2_musl_contact(Particles P, GridDesc G, const SolidDev* __restrict__ S,
                                                              ContactArrays3 c, double dt,
                                                              unsigned* __restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int sid = P.sid[i];
  const SolidDev& s = S[sid];
  double xx[3] = {P.x[0][i], P.x[1][i], P.x[2][i]};
  double L[9];
  for (int k = 0; k < 9; k++) L[k] = 0.0;
  bool ok = walk_stencil<3, B, false, true>(P, G, s, nullptr, i, [&](int n, double Ni, const double* dN) {
    const int k = contact_slot(c, n, sid);
    if (k < 0) return;
    // the slot's arrays by selects (an index into the argument struct
    // would copy it to scratch)
    const double m = (k ? c.m[1] : c.m[0])[n];
    if (m > 0.0) {
      const double mii = 1.0 / m;
      const double vIx = mii * (k ? c.p2[1][0] : c.p2[0][0])[n], vIy = mii * (k ? c.p2[1][1] : c.p2[0][1])[n],
                   vIz = mii * (k ? c.p2[1][2] : c.p2[0][2])[n];
      const double cc = Ni * mii * dt;
      xx[0] += cc * (k ? c.p[1][0] : c.p[0][0])[n];
      xx[1] += cc * (k ? c.p[1][1] : c.p[0][1])[n];
      xx[2] += cc * (k ? c.p[1][2] : c.p[0][2])[n];
      L[0] += dN[0] * vIx; L[1] += dN[1] * vIx; L[2] += dN[2] * vIx;
      L[3] += dN[0] * vIy; L[4] += dN[1] * vIy; L[5] += dN[2] * vIy;
      L[6] += dN[0] * vIz; L[7] += dN[1] * vIz; L[8] += dN[2] * vIz;
    }
  });
  if (!ok) {
    atomicOr(flags, JUMP_FLAG_OUT_OF_GRID);
    return;
  }
  musl3d_particle_update(P, s, i, dt, xx, L, flags);
}

// ------------------------------------------------------------------ energies
// Fix.jl:248-270.  Per-block sums in a fixed tree order, then one block sums
// the partials in block order: deterministic.
//
// Particle fields of the sums: the context's layout (SoA, symmetric tensors
// compressed) or a caller's arrays in the reference's AoS layout (vel d*n,
// strain and stress d*d*n column-major).  The strain-energy term runs over
// the full tensor in column-major order, so both layouts give the same bits
// for the symmetric tensors the reference keeps.
template <int D> struct EnergySoA {
  Particles P;
  __device__ double m(int i) const { return P.m[i]; }
  __device__ double V(int i) const { return P.V[i]; }
  __device__ double v(int i, int k) const { return P.v[k][i]; }
  __device__ double se(int i, int r, int c) const { return P.sig[sym<D>(r, c)][i] * P.eps[sym<D>(r, c)][i]; }
};
template <int D> struct EnergyAoS {
  const double *mass, *vel, *vol, *strain, *stress;
  __device__ double m(int i) const { return mass[i]; }
  __device__ double V(int i) const { return vol[i]; }
  __device__ double v(int i, int k) const { return vel[size_t(D) * i + k]; }
  __device__ double se(int i, int r, int c) const {
    const size_t q = size_t(D * D) * i + r + D * c;
    return stress[q] * strain[q];
  }
};
// Partial sums over blocks of kBlock particles of [begin, end), the blocks
// aligned at begin (a batch instance reduces exactly as a context of its own).
template <int D, class Src>
__global__ void __launch_bounds__(kBlock) k_energy_partial(Src src, int begin, int end, double* __restrict__ part) {
  __shared__ double ske[kBlock], sse[kBlock];
  int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  double ke = 0.0, se = 0.0;
  if (i < end) {
    double vv = src.v(i, 0) * src.v(i, 0);
    for (int k = 1; k < D; k++) vv = vv + src.v(i, k) * src.v(i, k);
    ke = 0.5 * src.m(i) * vv;
    // sum(sigma .* eps), column-major: 11 21 (31) 12 22 (32) (13 23 33)
    double acc = src.se(i, 0, 0);
    for (int c = 0; c < D; c++)
      for (int r = 0; r < D; r++)
        if (r + c > 0) acc = acc + src.se(i, r, c);
    se = 0.5 * src.V(i) * acc;
  }
  ske[threadIdx.x] = ke;
  sse[threadIdx.x] = se;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ske[threadIdx.x] += ske[threadIdx.x + off];
      sse[threadIdx.x] += sse[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ske[0];
    part[2 * blockIdx.x + 1] = sse[0];
  }
}

__global__ void __launch_bounds__(kBlock) k_energy_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double ske[kBlock], sse[kBlock];
  double ke = 0.0, se = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) {
    ke += part[2 * b];
    se += part[2 * b + 1];
  }
  ske[threadIdx.x] = ke;
  sse[threadIdx.x] = se;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ske[threadIdx.x] += ske[threadIdx.x + off];
      sse[threadIdx.x] += sse[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = ske[0];
    out[1] = sse[0];
  }
}

}  // namespace jh
