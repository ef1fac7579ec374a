// omp_step.hpp -- the 3D MUSL step of the CPU restatement, parallel over
// particles with OpenMP (the CPU baseline at many cores).
//
// TEST AND BENCHMARK INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and
// tests/): the reference's loops are single-threaded
// (SolveExplicitSolidMechanics3D.jl:46-211); this is the same arithmetic per
// particle and per node with the particle loops shared among threads.
//
// The scatters (P2G, SolveExplicitSolidMechanics3D.jl:55-84, and the
// re-projection, :135-138) are made race-free by colouring: particles are
// grouped by the block of their stencil origin, blocks of W x W x W cells (W
// the stencil width), and the blocks coloured by the parity of their block
// coordinates (8 colours).  Two blocks of one colour are a block apart, so
// their particles' stencils (W nodes per axis from the origin) share no node;
// the blocks of one colour run in parallel, the colours one after another.
// Each node therefore sums its contributions block by block (in colour
// order, particle order within a block) instead of in global particle order:
// results agree with the serial restatement to rounding, not bit for bit.
// The gathers (G2P-1, G2P-2) and the grid update are plain parallel loops.
#pragma once

#include <omp.h>

#include <vector>

#include "mpm_oracle.hpp"

namespace orc {

// A context sharing g's grid description, with its own support scratch
// (the support functions write only near/funcs/ders).
template <class T> Ctx<T> scratch_ctx(const Ctx<T>& g) {
  Ctx<T> c;
  c.dim = g.dim; c.basis = g.basis; c.alg = g.alg; c.ghostcell = g.ghostcell;
  c.tol = g.tol; c.alpha = g.alpha; c.dt = g.dt;
  for (int k = 0; k < 3; k++) c.body[k] = g.body[k];
  c.nx = g.nx; c.ny = g.ny; c.nz = g.nz; c.nxy = g.nxy; c.nn = g.nn;
  c.xmin = g.xmin; c.ymin = g.ymin; c.zmin = g.zmin;
  c.dx = g.dx; c.dy = g.dy; c.dz = g.dz; c.dxI = g.dxI; c.dyI = g.dyI; c.dzI = g.dzI;
  c.near.assign(64, 0);
  c.funcs.assign(64, T(0.0));
  c.ders.assign(3 * 64, T(0.0));
  return c;
}

inline int stencil_width(int basis) {
  return basis == JUMP_BASIS_LINEAR ? 2 : (basis == JUMP_BASIS_CUBIC_BSPLINE ? 4 : 3);
}

struct Colouring {
  // per colour, its blocks; per block, its particles as (solid, index)
  std::vector<std::vector<std::vector<std::pair<int, i64>>>> blocks;
};

// Particles by colour and block of their stencil origin (near[0] of the
// support: every basis lists its first node first).
template <class T> void colour_particles(Ctx<T>& g, std::vector<Ctx<T>>& tc, Colouring& col, uint32_t& flags) {
  const int W = stencil_width(g.basis);
  const i64 bx = (g.nx + W - 1) / W, by = (g.ny + W - 1) / W, bz = (g.nz + W - 1) / W;
  std::vector<i64> ofs;  // running particle offsets per solid
  i64 total = 0;
  for (auto& s : g.solids) { ofs.push_back(total); total += s.n; }
  std::vector<i64> blk(size_t(total), -1);
  uint32_t fl = 0;
#pragma omp parallel for schedule(static) reduction(| : fl)
  for (i64 q = 0; q < total; q++) {
    size_t si = 0;
    while (si + 1 < ofs.size() && ofs[si + 1] <= q) si++;
    Ctx<T>& c = tc[size_t(omp_get_thread_num())];
    const i64 ip = q - ofs[si];
    if (support(c, g.solids[si], ip, false) < 0) { fl |= JUMP_FLAG_OUT_OF_GRID; continue; }
    const i64 o = c.near[0];
    const i64 i = o % g.nx, j = (o / g.nx) % g.ny, k = o / g.nxy;
    blk[size_t(q)] = (i / W) + bx * ((j / W) + by * (k / W));
  }
  flags |= fl;
  col.blocks.assign(8, {});
  std::vector<std::vector<std::pair<int, i64>>> all(size_t(bx * by * bz));
  for (size_t si = 0, q = 0; si < g.solids.size(); si++)
    for (i64 ip = 0; ip < g.solids[si].n; ip++, q++)
      if (blk[q] >= 0) all[size_t(blk[q])].push_back({int(si), ip});
  for (i64 b = 0; b < bx * by * bz; b++) {
    if (all[size_t(b)].empty()) continue;
    const i64 i = b % bx, j = (b / bx) % by, k = b / (bx * by);
    col.blocks[size_t((i & 1) + 2 * (j & 1) + 4 * (k & 1))].push_back(std::move(all[size_t(b)]));
  }
}

// One 3D MUSL step (SolveExplicitSolidMechanics3D.jl:46-211) on nthreads.
template <class T> void step_musl_3d_omp(Ctx<T>& g, int nthreads) {
  omp_set_num_threads(nthreads);
  std::vector<Ctx<T>> tc;
  for (int k = 0; k < nthreads; k++) tc.push_back(scratch_ctx(g));
  const T dt = g.dt, alpha = g.alpha;
  const i64 nn = g.nn;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < nn; i++) {
    g.gmass[size_t(i)] = 0.0;
    for (int k = 0; k < 3; k++) { g.gp0[3 * i + k] = 0.0; g.gf[3 * i + k] = 0.0; g.gp2[3 * i + k] = 0.0; }
  }
  uint32_t flags = 0;
  Colouring col;
  colour_particles(g, tc, col, flags);
  // P2G (:55-84), colour by colour
  for (auto& blocks : col.blocks) {
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t b = 0; b < blocks.size(); b++) {
      Ctx<T>& c = tc[size_t(omp_get_thread_num())];
      for (auto& pr : blocks[b]) {
        const Solid<T>& s = g.solids[size_t(pr.first)];
        const i64 ip = pr.second;
        const int sup = support(c, s, ip, true);
        T fV = s.vol[ip], fM = s.mass[ip];
        const T* vp = &s.vel[3 * ip];
        const T* sg = &s.stress[9 * ip];
        for (int i = 0; i < sup; i++) {
          i64 id = c.near[i];
          T Ni = c.funcs[i];
          T dx = c.ders[3 * i], dy = c.ders[3 * i + 1], dz = c.ders[3 * i + 2];
          T Nim = Ni * fM;
          g.gmass[id] += Nim;
          for (int k = 0; k < 3; k++) g.gp0[3 * id + k] += Nim * vp[k];
          T f0 = sg[0] * dx + sg[3] * dy + sg[6] * dz;
          T f1 = sg[3] * dx + sg[4] * dy + sg[7] * dz;
          T f2 = sg[6] * dx + sg[7] * dy + sg[8] * dz;
          g.gf[3 * id] -= fV * f0;
          g.gf[3 * id + 1] -= fV * f1;
          g.gf[3 * id + 2] -= fV * f2;
        }
      }
    }
  }
  // grid update + BC (:89-105)
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < nn; i++) {
    for (int k = 0; k < 3; k++) g.gp[3 * i + k] = g.gp0[3 * i + k] + g.gf[3 * i + k] * dt;
    for (int k = 0; k < 3; k++)
      if (g.fixed[3 * i + k]) { g.gp0[3 * i + k] = 0.0; g.gp[3 * i + k] = 0.0; }
  }
  // G2P-1 (:109-134): particle velocities, every particle in parallel
  for (auto& s : g.solids) {
#pragma omp parallel for schedule(static)
    for (i64 ip = 0; ip < s.n; ip++) {
      Ctx<T>& c = tc[size_t(omp_get_thread_num())];
      const int sup = support(c, s, ip, false);
      if (sup < 0) continue;
      T* vv = &s.vel[3 * ip];
      T vx[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < sup; i++) {
        i64 id = c.near[i];
        T mI = g.gmass[id];
        if (val(mI) > 0.0) {
          T mii = (1.0 / mI) * c.funcs[i];
          for (int k = 0; k < 3; k++) vx[k] += (g.gp[3 * id + k] - alpha * g.gp0[3 * id + k]) * mii;
        }
      }
      for (int k = 0; k < 3; k++) vv[k] = alpha * vv[k] + vx[k];
    }
  }
  // the re-projection (:135-138), colour by colour
  for (auto& blocks : col.blocks) {
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t b = 0; b < blocks.size(); b++) {
      Ctx<T>& c = tc[size_t(omp_get_thread_num())];
      for (auto& pr : blocks[b]) {
        const Solid<T>& s = g.solids[size_t(pr.first)];
        const i64 ip = pr.second;
        const int sup = support(c, s, ip, false);
        const T mp = s.mass[ip];
        const T* vv = &s.vel[3 * ip];
        for (int i = 0; i < sup; i++) {
          i64 id = c.near[i];
          T Nm = c.funcs[i] * mp;
          for (int k = 0; k < 3; k++) g.gp2[3 * id + k] += Nm * vv[k];
        }
      }
    }
  }
  // G2P-2 (:161-211)
  for (auto& s : g.solids) {
#pragma omp parallel for schedule(static)
    for (i64 ip = 0; ip < s.n; ip++) {
      Ctx<T>& c = tc[size_t(omp_get_thread_num())];
      const int sup = support(c, s, ip, true);
      if (sup < 0) continue;
      T L[9];
      for (int k = 0; k < 9; k++) L[k] = 0.0;
      T xx[3] = {s.pos[3 * ip], s.pos[3 * ip + 1], s.pos[3 * ip + 2]};
      for (int i = 0; i < sup; i++) {
        i64 id = c.near[i];
        T Ni = c.funcs[i];
        T dNx = c.ders[3 * i], dNy = c.ders[3 * i + 1], dNz = c.ders[3 * i + 2];
        T m = g.gmass[id];
        if (val(m) > 0.0) {
          T mii = 1.0 / m;
          T vIx = mii * g.gp2[3 * id], vIy = mii * g.gp2[3 * id + 1], vIz = mii * g.gp2[3 * id + 2];
          T cc = Ni * mii * dt;
          for (int k = 0; k < 3; k++) xx[k] += cc * g.gp[3 * id + k];
          L[0] += dNx * vIx; L[1] += dNy * vIx; L[2] += dNz * vIx;
          L[3] += dNx * vIy; L[4] += dNy * vIy; L[5] += dNz * vIy;
          L[6] += dNx * vIz; L[7] += dNy * vIz; L[8] += dNz * vIz;
        }
      }
      musl3d_particle_update(c, s, ip, xx, L);
    }
  }
  for (auto& c : tc) flags |= c.flags;
  g.flags |= flags;
}

}  // namespace orc
