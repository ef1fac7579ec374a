// jump_api.hip -- context, memory management and the exported C ABI of
// libjumphip.so (declared in include/jump_mpm.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <climits>
#include <cstddef>
#include <memory>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mpm_adjoint3.hpp"
#include "mpm_output.hpp"

using namespace jh;

namespace {

struct HostSolid {
  int64_t n = 0;
  jump_material mat{};
  SolidDev dev{};
  int64_t offset = 0;  // first particle in the concatenated arrays
};

inline int blocks_for(int64_t n) { return int((n + kBlock - 1) / kBlock); }

}  // namespace

struct jump_ctx {
  jump_params prm{};
  GridDesc G{};
  int D = 2;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  std::vector<HostSolid> solids;
  std::vector<double> corners;  // CPDI corner coordinates of all solids

  // Host staging in device layout (SoA, symmetric tensors compressed);
  // authoritative while !on_device.
  int64_t np = 0;
  std::vector<double> hx[3], hv[3], hm, hV, hV0, hF[9], heps[6], hsig[6], halpha, halphadot, hvm, hep[6];
  std::vector<double> hC[4];  // APIC: affine matrices, column-major
  std::vector<double> hX0[2]; // positions when added (a body force in material coordinates)
  int body_kind = JUMP_BODY_CONSTANT;  // the body force (jump_set_body_force)
  double vortex[8] = {};               // JUMP_BODY_VORTEX2D: G, T, Ri, Ro, rho0, mu, xcenter, ycenter
  std::vector<std::vector<double2>> hpairs;  // transfer buffers of the interleaved gather fields
  std::vector<int> hsid, helem[4];
  std::vector<uint8_t> hfixed;
  bool on_device = false;  // device holds the current particle state
  int64_t dev_capacity = -1;

  // device memory
  Particles P{};
  GridArrays g{};
  ContactArrays cg{};    // contact mode only, 2D
  ContactArrays3 cg3{};  // contact mode only, 3D
  Musl2Arrays mu{};      // 2D MUSL only: matrix rows, right-hand sides, velocities
  ChebArrays ch{};       // 2D MUSL only: the Chebyshev iteration's vectors
  RunList runs{};        // 3D: the force pass's particle runs per node, replayed by the re-projection
  const Particles* Pout = nullptr;  // 2D USL, 3D MUSL: where the step writes the state after it (null: in place)
  int* d_dirty = nullptr;  // 3D: per node, the stamp of the step its window last changed
  int runs_gen = 0;        // stamp of the last step
  bool runs_valid = false; // the recorded runs match the particles' current bins
  std::vector<void*> part_allocs;
  std::vector<void*> grid_allocs;
  double* d_corners = nullptr;
  SolidDev* d_solids = nullptr;
  int *d_key = nullptr, *d_rank = nullptr, *d_cnt = nullptr, *d_start = nullptr,
      *d_sums = nullptr;
  unsigned long long* d_bxyz = nullptr;
  int *d_nodepos = nullptr, *d_order = nullptr;  // P2G work order: scan of active flags, node per slot
  unsigned long long* d_sorted = nullptr;  // (bin << 32) | particle, by bin then index; [np] = sentinel
  unsigned* d_flags = nullptr;
  double* d_red = nullptr;
  int* d_wmax = nullptr;
  int* d_queue = nullptr;       // work-queue counter of the persistent batched gather
  int queue_grid[3] = {0, 0, 0};  // its resident workgroups per block size (64, 128, 256), 0: not yet known
  uint8_t* d_fixed = nullptr;
  int cpdi_w = 0;
  bool grid_valid = false;
  int alloc_gen = 0;    // bumped whenever the particle arrays are reallocated (captured graphs hold their addresses)
  int p2g_shape = 0;    // batched gather launch shape (0..JUMP_P2G_SHAPES-1), chosen by the autotuner

  // Batch instances (an internal context of jump_batch holds several 2D
  // problems; an ordinary context is one instance with grid G).  NN: nodes
  // of all instances together.
  int ninst = 1;
  int NN = 0;
  std::vector<InstDev> hinst;
  std::vector<int> hnode_off;
  InstDev* d_inst = nullptr;
  int* d_node_off = nullptr;
  int max_rigid_rank = -1;  // highest rigid_rank of any solid (-1: none)
  BatchDev bt() const { return BatchDev{d_inst, d_node_off, ninst, NN}; }

  // reverse mode (jump_adjoint_*): state snapshots of the recorded run and
  // the adjoint arrays
  struct Adjoint {
    int64_t nsteps = 0;
    int K = 0;
    bool ready = false;
    int64_t cached_seg = -1;   // segment whose states and grids segbuf/gridbuf hold
    int64_t snap_np = -1, adj_np = -1;  // particle counts the buffers were sized for
    std::vector<void*> allocs;      // snapshots and grids of the recorded run
    std::vector<void*> adj_allocs;  // adjoint arrays
    std::vector<double*> ckpt;    // state at steps 0, K, 2K, ...
    std::vector<double*> segbuf;  // the states of one segment after its first (that one is its checkpoint; [0] unused)
    std::vector<double*> gridbuf; // and their grids (m, p0, p)
    std::vector<void*> binbuf;    // and their bins (start, sorted)
    double* final_state = nullptr;
    AdjParticles ap{};
    AdjNodes an{};
    AdjSources src{};
    NodeVel2 nv{};        // 2D: the nodal velocities of the recorded step
    AdjParticles3 ap3{};  // 3D
    AdjNodes3 an3{};
    AdjSources3 src3{};
    NodeVel3 nv3{};  // 3D: the nodal factors of the recorded step
    std::vector<double> dfy, dk1;  // per solid, from the last backward sweep
    int lanes = 1;  // lanes per node of the backward sweep's node gathers (1, 2, 4)
  } adj;
  uint32_t flags = 0;

  // LAMMPS dumps (jump_dump_*): pinned host buffer sets, reused while free,
  // and the writer thread.  A set is busy from its device copy until the
  // writer has formatted it.
  struct DumpBuf {
    char* host = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    std::atomic<bool> busy{false};
  };
  std::vector<std::unique_ptr<DumpBuf>> dump_bufs;
  DumpWriter dump_writer;
  bool dump_failed = false;  // a dump failed before jump_dump_wait was called

  // per-phase timing with HIP events on the context stream (jump_profile)
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_recs;
  double prof_ms[JUMP_PHASE_COUNT] = {};
  int64_t prof_cnt[JUMP_PHASE_COUNT] = {};
  hipEvent_t prof_open = nullptr;
};

// ------------------------------------------------------------------ utilities
static int set_err(jump_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return set_err(ctx, JUMP_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)
#define HIP_TRY_B(expr)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return batch_err(b, JUMP_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_));   \
  } while (0)

#define HIP_TRY_CTX_LAST(c)                                                                        \
  do {                                                                                             \
    hipError_t e_ = hipGetLastError();                                                             \
    if (e_ != hipSuccess) return set_err((c), JUMP_ERR_DEVICE, std::string("launch: ") + hipGetErrorString(e_)); \
  } while (0)

static int nsym(int D) { return D == 2 ? 3 : 6; }

template <class T> static int dev_alloc(jump_ctx* ctx, std::vector<void*>& list, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return set_err(ctx, JUMP_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  list.push_back(q);
  *p = static_cast<T*>(q);
  return JUMP_OK;
}

static void free_list(std::vector<void*>& list) {
  for (void* q : list) (void)hipFree(q);
  list.clear();
}

// Lame constants as the reference constructors compute them.
static SolidDev make_solid_dev(const jump_material& m, const double lp[3]) {
  SolidDev s{};
  s.kind = m.kind;
  s.rigid = m.kind == JUMP_MAT_RIGID;
  s.fixed = m.fixed != 0;
  s.lambda = m.E * m.nu / (1.0 + m.nu) / (1.0 - 2.0 * m.nu);
  s.mu = m.E / 2.0 / (1.0 + m.nu);
  s.kappa = s.lambda + s.mu;
  s.fy = m.fy;
  s.k1 = m.k1;
  s.A = m.A; s.B = m.B; s.C = m.C; s.n = m.n; s.epsdot0 = m.epsdot0; s.cellsize = m.cellsize;
  s.velo = m.kind == JUMP_MAT_JOHNSONCOOK ? sqrt((s.lambda + 2.0 * s.mu) / m.density) : 0.0;
  s.vx = m.vx;
  s.vy = m.vy;
  for (int k = 0; k < 3; k++) s.lp[k] = lp[k];
  return s;
}

// ------------------------------------------------------------------ transfers
// The gather fields (x, m, V, v, sigma) live two per 16-byte slot: field f
// of that order is half f%2 of pair array f/2.
static SPtr pair_field(const Particles& P, int f) { return SPtr{reinterpret_cast<double*>(P.pr[f / 2]) + (f % 2)}; }
static SPtr* gather_field(Particles& P, int D, int f) {
  if (f < D) return &P.x[f];
  if (f == D) return &P.m;
  if (f == D + 1) return &P.V;
  if (f < 2 * D + 2) return &P.v[f - D - 2];
  return &P.sig[f - 2 * D - 2];
}
static std::vector<double>* gather_host(jump_ctx* ctx, int f) {
  const int D = ctx->D;
  if (f < D) return &ctx->hx[f];
  if (f == D) return &ctx->hm;
  if (f == D + 1) return &ctx->hV;
  if (f < 2 * D + 2) return &ctx->hv[f - D - 2];
  return &ctx->hsig[f - 2 * D - 2];
}
static int nfields(int D) { return D == 3 ? PairLayout<3>::NF : PairLayout<2>::NF; }

static int alloc_particles(jump_ctx* ctx) {
  free_list(ctx->part_allocs);
  const int D = ctx->D, S = nsym(D);
  const size_t n = size_t(ctx->np);
  Particles& P = ctx->P;
  P = Particles{};
  int r = JUMP_OK;
  const int npair = D == 3 ? PairLayout<3>::NP : PairLayout<2>::NP;
  for (int j = 0; j < npair && !r; j++) r = dev_alloc(ctx, ctx->part_allocs, &P.pr[j], n);
  if (r) return r;
  for (int f = 0; f < (D == 3 ? PairLayout<3>::NF : PairLayout<2>::NF); f++) *gather_field(P, D, f) = pair_field(P, f);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &P.V0, n);
  for (int k = 0; k < D * D && !r; k++) r = dev_alloc(ctx, ctx->part_allocs, &P.F[k], n);
  for (int k = 0; k < S && !r; k++) r = dev_alloc(ctx, ctx->part_allocs, &P.eps[k], n);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &P.alpha, n);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &P.alphadot, n);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &P.vm, n);
  for (int k = 0; k < S && !r; k++) r = dev_alloc(ctx, ctx->part_allocs, &P.ep[k], n);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &P.sid, n);
  if (ctx->prm.basis == JUMP_BASIS_CPDI_Q4)
    for (int k = 0; k < 4 && !r; k++) r = dev_alloc(ctx, ctx->part_allocs, &P.elem[k], n);
  if (ctx->G.apic)
    for (int k = 0; k < 4 && !r; k++) r = dev_alloc(ctx, ctx->part_allocs, &P.C[k], n);
  if (ctx->G.bodyfield)
    for (int k = 0; k < 2 && !r; k++) r = dev_alloc(ctx, ctx->part_allocs, &P.bf[k], n);
  if (ctx->body_kind == JUMP_BODY_VORTEX2D_X0)
    for (int k = 0; k < 2 && !r; k++) r = dev_alloc(ctx, ctx->part_allocs, &P.X0[k], n);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &ctx->d_key, n);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &ctx->d_rank, n);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &ctx->d_sorted, n + 1);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &ctx->d_bxyz, n);
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &ctx->d_red, 2 * size_t(blocks_for(ctx->np) + 1));
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &ctx->d_solids, ctx->solids.size());
  if (!r) r = dev_alloc(ctx, ctx->part_allocs, &ctx->d_corners, ctx->corners.size());
  if (r) return r;
  // sentinel after the last entry: its bin half (all ones) matches no bin
  HIP_TRY(hipMemsetAsync(ctx->d_sorted + n, 0xff, sizeof(unsigned long long), ctx->stream));
  P.n = int(ctx->np);
  ctx->dev_capacity = ctx->np;
  ctx->alloc_gen++;
  return r;
}

template <class T> static int h2d(jump_ctx* ctx, T* dst, const std::vector<T>& src) {
  if (src.empty()) return JUMP_OK;
  HIP_TRY(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return JUMP_OK;
}
template <class T> static int d2h(jump_ctx* ctx, std::vector<T>& dst, const T* src) {
  if (dst.empty()) return JUMP_OK;
  HIP_TRY(hipMemcpyAsync(dst.data(), src, dst.size() * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  return JUMP_OK;
}

// Host staging -> device (all particle arrays).
static int upload(jump_ctx* ctx) {
  const int D = ctx->D, S = nsym(D);
  int r = alloc_particles(ctx);
  if (r) return r;
  Particles& P = ctx->P;
  {  // interleave the gather fields into their pair slots
    const size_t n = size_t(ctx->np);
    const int nf = nfields(D);
    ctx->hpairs.assign((nf + 1) / 2, std::vector<double2>(n, double2{0.0, 0.0}));
    for (int f = 0; f < nf; f++) {
      const std::vector<double>& src = *gather_host(ctx, f);
      std::vector<double2>& dst = ctx->hpairs[f / 2];
      for (size_t i = 0; i < n; i++) (f % 2 ? dst[i].y : dst[i].x) = src[i];
    }
    for (int j = 0; j < (nf + 1) / 2 && !r; j++) r = h2d(ctx, P.pr[j], ctx->hpairs[j]);
  }
  if (!r) r = h2d(ctx, P.V0, ctx->hV0);
  for (int k = 0; k < D * D && !r; k++) r = h2d(ctx, P.F[k], ctx->hF[k]);
  for (int k = 0; k < S && !r; k++) r = h2d(ctx, P.eps[k], ctx->heps[k]);
  if (!r) r = h2d(ctx, P.alpha, ctx->halpha);
  if (!r) r = h2d(ctx, P.alphadot, ctx->halphadot);
  if (!r) r = h2d(ctx, P.vm, ctx->hvm);
  for (int k = 0; k < S && !r; k++) r = h2d(ctx, P.ep[k], ctx->hep[k]);
  if (!r) r = h2d(ctx, P.sid, ctx->hsid);
  if (ctx->prm.basis == JUMP_BASIS_CPDI_Q4)
    for (int k = 0; k < 4 && !r; k++) r = h2d(ctx, P.elem[k], ctx->helem[k]);
  if (ctx->G.apic)
    for (int k = 0; k < 4 && !r; k++) r = h2d(ctx, P.C[k], ctx->hC[k]);
  if (ctx->body_kind == JUMP_BODY_VORTEX2D_X0)
    for (int k = 0; k < 2 && !r; k++) r = h2d(ctx, P.X0[k], ctx->hX0[k]);
  std::vector<SolidDev> sd;
  for (auto& s : ctx->solids) sd.push_back(s.dev);
  if (!r) r = h2d(ctx, ctx->d_solids, sd);
  if (!r) r = h2d(ctx, ctx->d_corners, ctx->corners);
  if (r) return r;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->on_device = true;
  ctx->cpdi_w = 0;
  ctx->runs_valid = false;  // new particle state: every node merges again
  return JUMP_OK;
}

// Device -> host staging (all particle arrays).
static int download(jump_ctx* ctx) {
  if (!ctx->on_device) return JUMP_OK;
  const int D = ctx->D, S = nsym(D);
  Particles& P = ctx->P;
  int r = JUMP_OK;
  const int nf = nfields(D);
  ctx->hpairs.resize((nf + 1) / 2);
  for (int j = 0; j < (nf + 1) / 2 && !r; j++) {
    ctx->hpairs[j].resize(size_t(ctx->np));
    r = d2h(ctx, ctx->hpairs[j], P.pr[j]);
  }
  if (!r) r = d2h(ctx, ctx->hV0, P.V0);
  for (int k = 0; k < D * D && !r; k++) r = d2h(ctx, ctx->hF[k], P.F[k]);
  for (int k = 0; k < S && !r; k++) r = d2h(ctx, ctx->heps[k], P.eps[k]);
  if (!r) r = d2h(ctx, ctx->halpha, P.alpha);
  if (!r) r = d2h(ctx, ctx->halphadot, P.alphadot);
  if (!r) r = d2h(ctx, ctx->hvm, P.vm);
  for (int k = 0; k < S && !r; k++) r = d2h(ctx, ctx->hep[k], P.ep[k]);
  if (ctx->G.apic)
    for (int k = 0; k < 4 && !r; k++) r = d2h(ctx, ctx->hC[k], P.C[k]);
  if (r) return r;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int f = 0; f < nf; f++) {
    std::vector<double>& dst = *gather_host(ctx, f);
    const std::vector<double2>& src = ctx->hpairs[f / 2];
    for (size_t i = 0; i < dst.size(); i++) dst[i] = f % 2 ? src[i].y : src[i].x;
  }
  return JUMP_OK;
}

static int ensure_device(jump_ctx* ctx) {
  if (ctx->on_device) return JUMP_OK;
  return upload(ctx);
}

static int alloc_grid(jump_ctx* ctx) {
  const size_t nn = size_t(ctx->NN);
  const int D = ctx->D;
  GridArrays& g = ctx->g;
  int r = dev_alloc(ctx, ctx->grid_allocs, &g.m, nn);
  for (int k = 0; k < D && !r; k++) r = dev_alloc(ctx, ctx->grid_allocs, &g.p0[k], nn);
  for (int k = 0; k < D && !r; k++) r = dev_alloc(ctx, ctx->grid_allocs, &g.p[k], nn);
  for (int k = 0; k < D && !r; k++) r = dev_alloc(ctx, ctx->grid_allocs, &g.p2[k], nn);
  for (int k = 0; k < D && !r; k++) r = dev_alloc(ctx, ctx->grid_allocs, &g.f[k], nn);
  if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_fixed, nn);
  if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_cnt, nn + 1);
  if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_start, nn + 1);
  if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_nodepos, nn + 1);
  if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_order, nn);
  if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_sums, (nn + 1) / kScanItems + 2);
  if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_flags, 1);
  if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_wmax, 1);
  if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_queue, 1);
  if (!r && D == 3) {
    for (double2** v : {&g.mq, &g.qq, &g.pxy, &g.p2xy, &g.pz2})
      if (!r) r = dev_alloc(ctx, ctx->grid_allocs, v, nn);
  }
  // run lists of the reference-order gathers: 3D, and single 2D problems of
  // up to kRunNodes2D nodes (kRunCap entries of 4 bytes per node)
  if (!r && (D == 3 || (ctx->ninst == 1 && nn <= size_t(kRunNodes2D)))) {
    if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->runs.runs, size_t(kRunCap) * nn);
    if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->runs.nruns, nn);
    if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_dirty, nn);
    if (!r) HIP_TRY(hipMemsetAsync(ctx->d_dirty, 0, nn * sizeof(int), ctx->stream));
  }
  if (ctx->prm.contact && D == 3) {
    ContactArrays3& c = ctx->cg3;
    for (int k = 0; k < 2 && !r; k++) {
      r = dev_alloc(ctx, ctx->grid_allocs, &c.body[k], nn);
      if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &c.m[k], nn);
      for (int a = 0; a < 3 && !r; a++) r = dev_alloc(ctx, ctx->grid_allocs, &c.p0[k][a], nn);
      for (int a = 0; a < 3 && !r; a++) r = dev_alloc(ctx, ctx->grid_allocs, &c.p[k][a], nn);
      for (int a = 0; a < 3 && !r; a++) r = dev_alloc(ctx, ctx->grid_allocs, &c.p2[k][a], nn);
    }
  } else if (D == 2 && ctx->prm.algorithm == JUMP_ALG_MUSL) {
    // rows of the blended mass matrix: (2R+1)^2 coefficients per node, R
    // the stencil width less one (CPDI: the widest domain, up to 5 nodes)
    const int W = ctx->prm.basis == JUMP_BASIS_CPDI_Q4 ? MaxW<JUMP_BASIS_CPDI_Q4>::W
                  : ctx->prm.basis == JUMP_BASIS_LINEAR ? 2
                  : ctx->prm.basis == JUMP_BASIS_CUBIC_BSPLINE ? 4
                                                               : 3;
    const size_t K1 = size_t(2 * W - 1);
    Musl2Arrays& m = ctx->mu;
    r = dev_alloc(ctx, ctx->grid_allocs, &m.brow, K1 * K1 * nn);
    if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &m.rhs, 4 * nn);
    if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &m.mv, 4 * nn);
    ChebArrays& c = ctx->ch;
    if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &c.x, 4 * nn);
    for (double** v : {&c.r, &c.d})  // two slots each
      if (!r) r = dev_alloc(ctx, ctx->grid_allocs, v, 8 * nn);
    if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &c.dinv, nn);
  } else if (ctx->prm.contact) {
    ContactArrays& c = ctx->cg;
    for (int k = 0; k < 2 && !r; k++) {
      r = dev_alloc(ctx, ctx->grid_allocs, &c.body[k], nn);
      if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &c.m[k], nn);
      for (int a = 0; a < 2 && !r; a++) r = dev_alloc(ctx, ctx->grid_allocs, &c.v0[k][a], nn);
      for (int a = 0; a < 2 && !r; a++) r = dev_alloc(ctx, ctx->grid_allocs, &c.v[k][a], nn);
    }
  }
  if (!r && ctx->ninst > 1) {
    r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_inst, ctx->hinst.size());
    if (!r) r = dev_alloc(ctx, ctx->grid_allocs, &ctx->d_node_off, ctx->hnode_off.size());
    if (!r) {
      HIP_TRY(hipMemcpyAsync(ctx->d_inst, ctx->hinst.data(), ctx->hinst.size() * sizeof(InstDev),
                             hipMemcpyHostToDevice, ctx->stream));
      HIP_TRY(hipMemcpyAsync(ctx->d_node_off, ctx->hnode_off.data(), ctx->hnode_off.size() * sizeof(int),
                             hipMemcpyHostToDevice, ctx->stream));
    }
  }
  if (r) return r;
  g.fixed = ctx->d_fixed;
  HIP_TRY(hipMemsetAsync(ctx->d_fixed, 0, nn, ctx->stream));
  HIP_TRY(hipMemsetAsync(ctx->d_flags, 0, sizeof(unsigned), ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return JUMP_OK;
}

// ------------------------------------------------------------------ profiling
static hipEvent_t prof_event(jump_ctx* ctx) {
  if (ctx->ev_used == ctx->ev_pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->ev_pool.push_back(e);
  }
  return ctx->ev_pool[ctx->ev_used++];
}
static void prof_begin(jump_ctx* ctx) {
  if (!ctx->prof) return;
  ctx->prof_open = prof_event(ctx);
  if (ctx->prof_open) (void)hipEventRecord(ctx->prof_open, ctx->stream);
}
static void prof_end(jump_ctx* ctx, int phase) {
  if (!ctx->prof || !ctx->prof_open) return;
  hipEvent_t e = prof_event(ctx);
  if (!e) return;
  (void)hipEventRecord(e, ctx->stream);
  ctx->prof_recs.push_back({phase, {ctx->prof_open, e}});
  ctx->prof_open = nullptr;
}
// After the stream has drained: fold the recorded intervals into the totals.
static void prof_collect(jump_ctx* ctx) {
  for (auto& r : ctx->prof_recs) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r.second.first, r.second.second) == hipSuccess) {
      ctx->prof_ms[r.first] += ms;
      ctx->prof_cnt[r.first] += 1;
    }
  }
  ctx->prof_recs.clear();
  ctx->ev_used = 0;
}

// ------------------------------------------------------------------ the step
#define LAUNCH(kernel, n, ...)                                                                \
  do {                                                                                        \
    if ((n) > 0) hipLaunchKernelGGL((kernel), dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, __VA_ARGS__); \
  } while (0)
#define LAUNCH_P2G(kernel, ...)                                                                      \
  do {                                                                                               \
    hipLaunchKernelGGL((kernel), dim3((ctx->NN + kP2GBlock - 1) / kP2GBlock), dim3(kP2GBlock), 0,     \
                       ctx->stream, __VA_ARGS__);                                                    \
  } while (0)

// Exclusive prefix sum of in[0, nn+1) into out (three passes).
static void scan_nodes(jump_ctx* ctx, const int* in, int* out) {
  const int n = ctx->NN + 1;
  const int nb = (n + kScanItems - 1) / kScanItems;
  hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(kBlock), 0, ctx->stream, in, out, ctx->d_sums, n);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->d_sums, nb);
  LAUNCH(k_scan_add, n, out, ctx->d_sums, n);
}

// dirty, gen: run replay (k_bin stamps nodes whose window changed)
template <int D, int B, bool MULTI> static int bin_and_sort(jump_ctx* ctx, int* dirty = nullptr, int gen = 0) {
  const int nn = ctx->NN;
  const int np = int(ctx->np);
  const int W = (B == JUMP_BASIS_CPDI_Q4) ? ctx->cpdi_w : Width<B>::W;
  LAUNCH(k_fill_i32, nn + 1, ctx->d_cnt, 0, nn + 1);
  LAUNCH((k_bin<D, B, MULTI>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->d_key, ctx->d_rank,
         ctx->d_bxyz, ctx->d_cnt, ctx->d_flags, ctx->bt(), dirty, gen, ctx->prm.contact, W);
  scan_nodes(ctx, ctx->d_cnt, ctx->d_start);
  LAUNCH(k_scatter, np, ctx->d_key, ctx->d_start, ctx->d_rank, ctx->d_sorted, np, nn);
  LAUNCH(k_binsort, nn, ctx->d_start, ctx->d_sorted, nn);
  // P2G work order: active nodes first (d_cnt is free again as the flags)
  LAUNCH((k_node_flags<D, MULTI>), nn + 1, ctx->G, ctx->d_start, W, ctx->d_cnt, ctx->bt());
  scan_nodes(ctx, ctx->d_cnt, ctx->d_nodepos);
  LAUNCH(k_node_order, nn, ctx->d_nodepos, nn, ctx->d_order);
  return JUMP_OK;
}

// Launch shapes of the batched gather (P2G over the nodes of all instances):
// the node order (active nodes compacted, or every node in grid-row order)
// and the workgroup size, as one workgroup per tile of nodes (shapes 0-5), or
// as a persistent kernel whose resident workgroups take (instance, tile)
// items of the compacted order from a work queue (6-8).  Every shape sums
// each node in the same order, so the choice changes the time only; the
// autotuner times them on the batch's own particles (tune_batched_p2g).

template <int B, int ORDER, bool COMPACT, int BLK> static void p2g_batched_as(jump_ctx* ctx) {
  hipLaunchKernelGGL((k_p2g<2, B, 0, ORDER, COMPACT, true, BLK>), dim3((ctx->NN + BLK - 1) / BLK), dim3(BLK), 0,
                     ctx->stream, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->d_start, ctx->d_sorted,
                     ctx->d_bxyz, ctx->d_order, ctx->d_nodepos, ctx->g, ctx->cpdi_w, ctx->prm.dt, ctx->bt(), 0.0,
                     RunList{});
}
// The persistent gather: as many workgroups as stay resident on the device
// (occupancy x CUs, looked up once per context and block size), at most one
// per item; the queue counter is zeroed by a kernel (graph capture records
// no memset).
template <int B, int ORDER, int BLK> static void p2g_queue_as(jump_ctx* ctx, int slot) {
  if (ctx->queue_grid[slot] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_p2g_queue<B, ORDER, BLK>, BLK, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess)
      per_cu = cus = 0;
    ctx->queue_grid[slot] = std::max(1, per_cu * cus);
  }
  const int nlog = (ctx->NN + BLK - 1) / BLK;
  const int grid = std::max(1, std::min(nlog, ctx->queue_grid[slot]));
  LAUNCH(k_fill_i32, 1, ctx->d_queue, 0, 1);
  hipLaunchKernelGGL((k_p2g_queue<B, ORDER, BLK>), dim3(grid), dim3(BLK), 0, ctx->stream, ctx->P, ctx->G,
                     ctx->d_solids, ctx->d_corners, ctx->d_start, ctx->d_sorted, ctx->d_bxyz, ctx->d_order,
                     ctx->d_nodepos, ctx->g, ctx->cpdi_w, ctx->prm.dt, ctx->bt(), ctx->d_queue);
}
template <int B, int ORDER> static void p2g_batched(jump_ctx* ctx, int shape) {
  static_assert(JUMP_P2G_SHAPES == 9, "one case per shape");
  switch (shape) {
    case 1: p2g_batched_as<B, ORDER, true, 64>(ctx); break;
    case 2: p2g_batched_as<B, ORDER, true, 256>(ctx); break;
    case 3: p2g_batched_as<B, ORDER, false, 128>(ctx); break;
    case 4: p2g_batched_as<B, ORDER, false, 64>(ctx); break;
    case 5: p2g_batched_as<B, ORDER, false, 256>(ctx); break;
    case 6: p2g_queue_as<B, ORDER, 128>(ctx, 1); break;
    case 7: p2g_queue_as<B, ORDER, 64>(ctx, 0); break;
    case 8: p2g_queue_as<B, ORDER, 256>(ctx, 2); break;
    default: p2g_batched_as<B, ORDER, true, 128>(ctx); break;
  }
}

// The run lists for this step's reference-order gather: this step's stamp
// (after an upload every node merges and records).  None (runs null) in the
// binned order or where the context has no lists.
template <int ORDER> static RunList begin_runs(jump_ctx* ctx) {
  RunList rl = ctx->runs;
  if (ORDER != JUMP_ORDER_REFERENCE || !rl.runs) return RunList{};
  if (ctx->runs_gen == INT_MAX) {  // stamps wrapped: start over
    ctx->runs_gen = 0;
    ctx->runs_valid = false;
    LAUNCH(k_fill_i32, ctx->NN, ctx->d_dirty, 0, ctx->NN);
  }
  rl.dirty = ctx->d_dirty;
  rl.gen = ++ctx->runs_gen;
  rl.all = ctx->runs_valid ? 0 : 1;
  return rl;
}

// MULTI (a batch of 2D instances): the gather runs over all instances' nodes
// in the launch shape the autotuner chose (by default the compacted active
// nodes, since the instances' grids are mostly larger than their solids).
template <int B, int ORDER, bool MULTI> static void step_usl2d(jump_ctx* ctx) {
  const int np = int(ctx->np);
  const double dt = ctx->prm.dt;
  const RunList rl = MULTI ? RunList{} : begin_runs<ORDER>(ctx);
  prof_begin(ctx);
  bin_and_sort<2, B, MULTI>(ctx, rl.runs ? ctx->d_dirty : nullptr, rl.gen);
  prof_end(ctx, JUMP_PHASE_SORT);
  prof_begin(ctx);
  if constexpr (MULTI && B != JUMP_BASIS_CPDI_Q4) {  // (batches refuse CPDI)
    p2g_batched<B, ORDER>(ctx, ctx->p2g_shape);
  } else {
    LAUNCH_P2G((k_p2g<2, B, 0, ORDER, MULTI || kUsl2dCompact, MULTI>), ctx->P, ctx->G, ctx->d_solids, ctx->d_corners,
               ctx->d_start,
               ctx->d_sorted, ctx->d_bxyz, ctx->d_order, ctx->d_nodepos, ctx->g, ctx->cpdi_w, dt, ctx->bt(), 0.0,
               rl);
  }
  if (rl.runs) ctx->runs_valid = true;
  prof_end(ctx, JUMP_PHASE_P2G);
  prof_begin(ctx);
  for (int r = 0; r <= ctx->max_rigid_rank; r++)
    LAUNCH((k_rigid_override<MULTI>), np, ctx->P, ctx->G, ctx->d_solids, r, ctx->g, ctx->d_flags, ctx->bt());
  prof_end(ctx, JUMP_PHASE_RIGID);
  prof_begin(ctx);
  const bool oop = !MULTI && ctx->Pout;  // the reverse mode's records (single problems)
  const Particles& Po = oop ? *ctx->Pout : ctx->P;
  if (oop)
    LAUNCH((k_g2p_usl2d<B, false, true>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->g,
           ctx->prm.tolerance, dt, ctx->d_flags, ctx->bt(), Po);
  else
    LAUNCH((k_g2p_usl2d<B, MULTI>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->g, ctx->prm.tolerance,
           dt, ctx->d_flags, ctx->bt(), Po);
  prof_end(ctx, JUMP_PHASE_G2P);
  prof_begin(ctx);
  if (ctx->max_rigid_rank >= 0) LAUNCH((k_rigid_move<MULTI>), np, Po, ctx->d_solids, dt, ctx->bt());
  prof_end(ctx, JUMP_PHASE_RIGID);
}

// 2D USL with multi-body contact: P2G per solid slot, grid update and
// contact correction in one gather; G2P from each particle's own slot.
template <int B, int ORDER> static void step_usl2d_contact(jump_ctx* ctx) {
  const int np = int(ctx->np);
  const double dt = ctx->prm.dt;
  const RunList rl = begin_runs<ORDER>(ctx);
  prof_begin(ctx);
  bin_and_sort<2, B, false>(ctx, rl.runs ? ctx->d_dirty : nullptr, rl.gen);
  prof_end(ctx, JUMP_PHASE_SORT);
  prof_begin(ctx);
  if (ctx->G.bodyfield)
    LAUNCH_P2G((k_p2g_contact<B, ORDER, true, kContactCompact>), ctx->P, ctx->G, ctx->d_solids, ctx->d_corners,
               ctx->d_start, ctx->d_sorted, ctx->d_bxyz, ctx->g, ctx->cg, ctx->cpdi_w, dt, ctx->prm.friction,
               ctx->d_flags, rl, ctx->d_order, ctx->d_nodepos);
  else
    LAUNCH_P2G((k_p2g_contact<B, ORDER, false, kContactCompact>), ctx->P, ctx->G, ctx->d_solids, ctx->d_corners,
               ctx->d_start, ctx->d_sorted, ctx->d_bxyz, ctx->g, ctx->cg, ctx->cpdi_w, dt, ctx->prm.friction,
               ctx->d_flags, rl, ctx->d_order, ctx->d_nodepos);
  if (rl.runs) ctx->runs_valid = true;
  prof_end(ctx, JUMP_PHASE_P2G);
  prof_begin(ctx);
  LAUNCH((k_g2p_usl2d_contact<B>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->cg, ctx->prm.tolerance, dt,
         ctx->d_flags);
  prof_end(ctx, JUMP_PHASE_G2P);
  prof_begin(ctx);
  if (ctx->max_rigid_rank >= 0) LAUNCH((k_rigid_move<false>), np, ctx->P, ctx->d_solids, dt, ctx->bt());
  prof_end(ctx, JUMP_PHASE_RIGID);
}

// 3D MUSL with multi-body contact: slot P2G with the grid update and the
// normal correction, G2P-1, slot re-projection and G2P-2 on the own slot.
template <int B, int ORDER> static void step_musl3d_contact(jump_ctx* ctx) {
  const int np = int(ctx->np);
  const double dt = ctx->prm.dt;
  prof_begin(ctx);
  bin_and_sort<3, B, false>(ctx);
  prof_end(ctx, JUMP_PHASE_SORT);
  prof_begin(ctx);
  LAUNCH_P2G((k_p2g_contact3d<B, ORDER>), ctx->P, ctx->G, ctx->d_solids, ctx->d_start, ctx->d_sorted, ctx->d_bxyz,
             ctx->g, ctx->cg3, dt, ctx->d_flags);
  prof_end(ctx, JUMP_PHASE_P2G);
  prof_begin(ctx);
  LAUNCH((k_g2p1_musl_contact<B>), np, ctx->P, ctx->G, ctx->d_solids, ctx->cg3, ctx->prm.alpha, ctx->d_flags);
  prof_end(ctx, JUMP_PHASE_G2P1);
  prof_begin(ctx);
  LAUNCH_P2G((k_p2g2_contact3d<B, ORDER>), ctx->P, ctx->G, ctx->d_solids, ctx->d_start, ctx->d_sorted, ctx->d_bxyz,
             ctx->g, ctx->cg3);
  prof_end(ctx, JUMP_PHASE_P2G2);
  prof_begin(ctx);
  LAUNCH((k_g2p2_musl_contact<B>), np, ctx->P, ctx->G, ctx->d_solids, ctx->cg3, dt, ctx->d_flags);
  prof_end(ctx, JUMP_PHASE_G2P2);
}

// 3D: the particle arrays of P with the pairs that hold the velocity, (V, vx)
// and (vy, vz), taken from Q (the re-projection reads only vx of the first).
static Particles with_velocity_of(const Particles& P, const Particles& Q) {
  Particles R = P;
  R.pr[2] = Q.pr[2];
  R.pr[3] = Q.pr[3];
  for (int f = 0; f < PairLayout<3>::NF; f++) *gather_field(R, 3, f) = pair_field(R, f);
  return R;
}
template <int B, int ORDER> static void step_musl3d(jump_ctx* ctx) {
  const int np = int(ctx->np);
  const double dt = ctx->prm.dt;
  // run replay (reference order)
  const RunList rl = begin_runs<ORDER>(ctx);
  prof_begin(ctx);
  bin_and_sort<3, B, false>(ctx, rl.runs ? ctx->d_dirty : nullptr, rl.gen);
  prof_end(ctx, JUMP_PHASE_SORT);
  prof_begin(ctx);
  constexpr int LN0 = ORDER == JUMP_ORDER_REFERENCE ? kP2GLanes0 : 1;  // lanes per node
  constexpr int LN1 = ORDER == JUMP_ORDER_REFERENCE ? kP2GLanes1 : 1;
  hipLaunchKernelGGL((k_p2g<3, B, 0, ORDER, false, false, kP2GBlock, LN0>),
                     dim3((ctx->NN + kP2GBlock / LN0 - 1) / (kP2GBlock / LN0)), dim3(kP2GBlock), 0, ctx->stream,
                     ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->d_start, ctx->d_sorted, ctx->d_bxyz,
                     ctx->d_order, ctx->d_nodepos, ctx->g, 0, dt, ctx->bt(), ctx->prm.alpha, rl);
  if (rl.runs) ctx->runs_valid = true;
  prof_end(ctx, JUMP_PHASE_P2G);
  // the reverse mode's records: the state before the step in ctx->P, after
  // it in *ctx->Pout; the re-projection reads the new velocities there
  const bool oop = ctx->Pout != nullptr;
  const Particles& Po = oop ? *ctx->Pout : ctx->P;
  const Particles Pv = oop ? with_velocity_of(ctx->P, Po) : ctx->P;
  prof_begin(ctx);
  if (oop)
    LAUNCH((k_g2p1_musl<B, true>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->g, ctx->prm.alpha,
           ctx->d_flags, Po);
  else
    LAUNCH((k_g2p1_musl<B, false>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->g, ctx->prm.alpha,
           ctx->d_flags, Po);
  prof_end(ctx, JUMP_PHASE_G2P1);
  prof_begin(ctx);
  {
    const int rows = (ctx->NN + kP2GBlock / LN1 - 1) / (kP2GBlock / LN1);
    hipLaunchKernelGGL((k_p2g<3, B, 1, ORDER, true, false, kP2GBlock, LN1>), dim3(xcd_grid(rows)), dim3(kP2GBlock),
                       0, ctx->stream, Pv, ctx->G, ctx->d_solids, ctx->d_corners, ctx->d_start, ctx->d_sorted,
                       ctx->d_bxyz, ctx->d_order, ctx->d_nodepos, ctx->g, 0, dt, ctx->bt(), ctx->prm.alpha,
                       rl);
  }
  prof_end(ctx, JUMP_PHASE_P2G2);
  // The second Dirichlet pass (SolveExplicitSolidMechanics3D.jl:142-156)
  // zeroes components of momentum0/momentum already zeroed by the grid update.
  prof_begin(ctx);
  if (oop)
    LAUNCH((k_g2p2_musl<B, true>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->g, dt, ctx->d_flags, Po);
  else
    LAUNCH((k_g2p2_musl<B, false>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->g, dt, ctx->d_flags, Po);
  prof_end(ctx, JUMP_PHASE_G2P2);
}

// 2D MUSL with the blended mass matrix: P2G with the matrix rows and the
// grid fields, the solve for the nodal velocities, G2P-1, G2P-2.  R: the
// matrix stencil radius (width - 1).
constexpr double kMuslTol = 1e-14;  // bound on the error's B-norm, relative to the solution's
constexpr int kMuslMaxIt = 2000;
// The Chebyshev iteration count for kMuslTol: the least k with
// 2 q^k / (1 + q^2k) <= kMuslTol, q = (sqrt 10 - 1) / (sqrt 10 + 1).
static int cheb_iterations() {
  const double q = (std::sqrt(10.0) - 1.0) / (std::sqrt(10.0) + 1.0);
  double qk = 1.0;
  for (int k = 1; k < kMuslMaxIt; k++) {
    qk *= q;
    if (2.0 * qk / (1.0 + qk * qk) <= kMuslTol) return k;
  }
  return kMuslMaxIt;
}
// The blended mass solve: K Chebyshev iterations of one launch each; the
// last writes the nodal velocities mv.  The launches do not depend on the
// data, so the host never waits.
template <int R> static void solve_musl2d(jump_ctx* ctx) {
  const int nn = ctx->NN;
  const bool wide = nn > kChebWideNodes;
  const int ntile = wide ? cheb_tiles<64, 4>(ctx->G) : cheb_tiles<32, 8>(ctx->G);
  const ChebArrays& ch = ctx->ch;
  const double theta = 0.55, delta = 0.45, sigma = theta / delta;
  LAUNCH(k_cheb_init, nn, nn, ctx->mu, ctx->g.m, ch, 1.0 / theta);
  static const int K = cheb_iterations();
  // iteration k forms d_k+1 = c1 d_k + c2 D^-1 r_k+1
  double rho = 1.0 / sigma;
  for (int k = 0; k < K; k++) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    if (wide)
      hipLaunchKernelGGL((k_cheb_iter<R, 64, 4>), dim3(ntile), dim3(kChebBlock), 0, ctx->stream, ctx->G, ctx->mu, ch,
                         k & 1, rn * rho, 2.0 * rn / delta, k + 1 == K ? 1 : 0);
    else
      hipLaunchKernelGGL((k_cheb_iter<R, 32, 8>), dim3(ntile), dim3(kChebBlock), 0, ctx->stream, ctx->G, ctx->mu, ch,
                         k & 1, rn * rho, 2.0 * rn / delta, k + 1 == K ? 1 : 0);
    rho = rn;
  }
}
template <int B, int ORDER, int R> static void step_musl2d(jump_ctx* ctx) {
  const int np = int(ctx->np);
  const double dt = ctx->prm.dt;
  prof_begin(ctx);
  bin_and_sort<2, B, false>(ctx);
  prof_end(ctx, JUMP_PHASE_SORT);
  prof_begin(ctx);
  if (ctx->G.bodyfield)
    LAUNCH_P2G((k_p2g_musl2d<B, ORDER, R, true>), ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->d_start,
               ctx->d_sorted, ctx->d_bxyz, ctx->g, ctx->mu, ctx->cpdi_w, dt);
  else
    LAUNCH_P2G((k_p2g_musl2d<B, ORDER, R, false>), ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->d_start,
               ctx->d_sorted, ctx->d_bxyz, ctx->g, ctx->mu, ctx->cpdi_w, dt);
  prof_end(ctx, JUMP_PHASE_P2G);
  prof_begin(ctx);
  solve_musl2d<R>(ctx);
  prof_end(ctx, JUMP_PHASE_SOLVE);
  prof_begin(ctx);
  LAUNCH((k_g2p1_musl2d<B>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->g, ctx->mu, ctx->prm.alpha,
         ctx->d_flags);
  prof_end(ctx, JUMP_PHASE_G2P1);
  prof_begin(ctx);
  LAUNCH((k_g2p2_musl2d<B>), np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->g, ctx->mu, dt, ctx->d_flags);
  prof_end(ctx, JUMP_PHASE_G2P2);
}
template <int ORDER> static void step_musl2d_cpdi(jump_ctx* ctx) {
  switch (ctx->cpdi_w) {
    case 2: step_musl2d<JUMP_BASIS_CPDI_Q4, ORDER, 1>(ctx); break;
    case 3: step_musl2d<JUMP_BASIS_CPDI_Q4, ORDER, 2>(ctx); break;
    case 4: step_musl2d<JUMP_BASIS_CPDI_Q4, ORDER, 3>(ctx); break;
    default: step_musl2d<JUMP_BASIS_CPDI_Q4, ORDER, 4>(ctx); break;
  }
}

static int prepare_cpdi(jump_ctx* ctx) {
  if (ctx->prm.basis != JUMP_BASIS_CPDI_Q4 || ctx->cpdi_w > 0) return JUMP_OK;
  int zero = 0;
  HIP_TRY(hipMemcpyAsync(ctx->d_wmax, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  LAUNCH(k_cpdi_width, ctx->np, ctx->P, ctx->G, ctx->d_solids, ctx->d_corners, ctx->d_wmax);
  int w = 0;
  HIP_TRY(hipMemcpyAsync(&w, ctx->d_wmax, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (w > MaxW<JUMP_BASIS_CPDI_Q4>::W)
    return set_err(ctx, JUMP_ERR_UNSUPPORTED, "CPDI particle domain spans more than 4 cells");
  ctx->cpdi_w = std::max(w, 2);
  return JUMP_OK;
}

// t0: the time of the first step (the reference's accumulated t), which a
// position-dependent body force reads.
static int run_steps(jump_ctx* ctx, int64_t nsteps, double t0 = 0.0) {
  const bool ref = ctx->prm.order == JUMP_ORDER_REFERENCE;
  double t = t0;
  for (int64_t s = 0; s < nsteps; s++) {
    if (ctx->G.bodyfield) {  // bodyforce(body, xx[ip], t) at the start of the step
      VortexDev vd;
      for (int k = 0; k < 8; k++) vd.v[k] = ctx->vortex[k];
      if (ctx->body_kind == JUMP_BODY_VORTEX2D_X0) LAUNCH(k_body_vortex2d<true>, int(ctx->np), ctx->P, vd, t);
      else LAUNCH(k_body_vortex2d<false>, int(ctx->np), ctx->P, vd, t);
    }
    t += ctx->prm.dt;
    if (ctx->D == 2 && ctx->prm.algorithm == JUMP_ALG_MUSL) {
      switch (ctx->prm.basis) {
        case JUMP_BASIS_LINEAR:
          if (ref) step_musl2d<JUMP_BASIS_LINEAR, JUMP_ORDER_REFERENCE, 1>(ctx);
          else step_musl2d<JUMP_BASIS_LINEAR, JUMP_ORDER_BINNED, 1>(ctx);
          break;
        case JUMP_BASIS_QUAD_BSPLINE:
          if (ref) step_musl2d<JUMP_BASIS_QUAD_BSPLINE, JUMP_ORDER_REFERENCE, 2>(ctx);
          else step_musl2d<JUMP_BASIS_QUAD_BSPLINE, JUMP_ORDER_BINNED, 2>(ctx);
          break;
        case JUMP_BASIS_CUBIC_BSPLINE:
          if (ref) step_musl2d<JUMP_BASIS_CUBIC_BSPLINE, JUMP_ORDER_REFERENCE, 3>(ctx);
          else step_musl2d<JUMP_BASIS_CUBIC_BSPLINE, JUMP_ORDER_BINNED, 3>(ctx);
          break;
        case JUMP_BASIS_GIMP:
          if (ref) step_musl2d<JUMP_BASIS_GIMP, JUMP_ORDER_REFERENCE, 2>(ctx);
          else step_musl2d<JUMP_BASIS_GIMP, JUMP_ORDER_BINNED, 2>(ctx);
          break;
        case JUMP_BASIS_CPDI_Q4:
          if (ref) step_musl2d_cpdi<JUMP_ORDER_REFERENCE>(ctx);
          else step_musl2d_cpdi<JUMP_ORDER_BINNED>(ctx);
          break;
      }
    } else if (ctx->D == 2) {
      switch (ctx->prm.basis) {
#define USL_CASE(BB)                                                     \
  case BB:                                                               \
    if (ctx->prm.contact) {                                              \
      if (ref) step_usl2d_contact<BB, JUMP_ORDER_REFERENCE>(ctx);        \
      else step_usl2d_contact<BB, JUMP_ORDER_BINNED>(ctx);               \
    } else if (ctx->ninst > 1) {                                         \
      if (ref) step_usl2d<BB, JUMP_ORDER_REFERENCE, true>(ctx);          \
      else step_usl2d<BB, JUMP_ORDER_BINNED, true>(ctx);                 \
    } else {                                                             \
      if (ref) step_usl2d<BB, JUMP_ORDER_REFERENCE, false>(ctx);         \
      else step_usl2d<BB, JUMP_ORDER_BINNED, false>(ctx);                \
    }                                                                    \
    break;
        USL_CASE(JUMP_BASIS_LINEAR)
        USL_CASE(JUMP_BASIS_QUAD_BSPLINE)
        USL_CASE(JUMP_BASIS_CUBIC_BSPLINE)
        USL_CASE(JUMP_BASIS_CPDI_Q4)
        USL_CASE(JUMP_BASIS_GIMP)
#undef USL_CASE
      }
    } else {
      switch (ctx->prm.basis) {
#define MUSL_CASE(BB)                                                    \
  case BB:                                                               \
    if (ctx->prm.contact) {                                              \
      if (ref) step_musl3d_contact<BB, JUMP_ORDER_REFERENCE>(ctx);       \
      else step_musl3d_contact<BB, JUMP_ORDER_BINNED>(ctx);              \
    } else {                                                             \
      if (ref) step_musl3d<BB, JUMP_ORDER_REFERENCE>(ctx);               \
      else step_musl3d<BB, JUMP_ORDER_BINNED>(ctx);                      \
    }                                                                    \
    break;
        MUSL_CASE(JUMP_BASIS_LINEAR)
        MUSL_CASE(JUMP_BASIS_QUAD_BSPLINE)
        MUSL_CASE(JUMP_BASIS_CUBIC_BSPLINE)
        MUSL_CASE(JUMP_BASIS_GIMP)
#undef MUSL_CASE
      }
    }
  }
  HIP_TRY(hipGetLastError());
  return JUMP_OK;
}

// ------------------------------------------------------------------ C ABI
extern "C" {

int jump_ctx_create(const jump_params* p, int device, jump_ctx** out) {
  if (!p || !out) return JUMP_ERR_ARG;
  *out = nullptr;
  if (p->dim != 2 && p->dim != 3) return JUMP_ERR_ARG;
  if (p->basis < JUMP_BASIS_LINEAR || p->basis > JUMP_BASIS_GIMP) return JUMP_ERR_ARG;
  if (p->dim == 3 && p->basis == JUMP_BASIS_CPDI_Q4) return JUMP_ERR_UNSUPPORTED;
  // The reference implements USL and MUSL (blended mass matrix) in 2D and
  // MUSL in 3D on this path (3D USL is marked unfinished and broken).
  if (p->algorithm != JUMP_ALG_USL && p->algorithm != JUMP_ALG_MUSL && p->algorithm != JUMP_ALG_APIC)
    return JUMP_ERR_ARG;
  if (p->dim == 3 && p->algorithm != JUMP_ALG_MUSL) return JUMP_ERR_UNSUPPORTED;
  if (p->dim == 2 && p->algorithm == JUMP_ALG_MUSL && p->contact) return JUMP_ERR_UNSUPPORTED;
  // APIC: the 2D USL step with affine transfers, for the point bases
  if (p->algorithm == JUMP_ALG_APIC && (p->contact || p->basis == JUMP_BASIS_CPDI_Q4)) return JUMP_ERR_UNSUPPORTED;
  if (p->nx < 2 || p->ny < 2 || (p->dim == 3 && p->nz < 2)) return JUMP_ERR_ARG;
  const int64_t nz = p->dim == 3 ? p->nz : 1;
  if (p->nx >= 65536 || p->ny >= 65536 || nz >= 65536) return JUMP_ERR_UNSUPPORTED;
  if (p->nx * p->ny * nz > kMaxNodes) return JUMP_ERR_UNSUPPORTED;
  if (!(p->dt > 0.0)) return JUMP_ERR_ARG;
  if (p->order != JUMP_ORDER_REFERENCE && p->order != JUMP_ORDER_BINNED) return JUMP_ERR_ARG;
  if (p->contact != 0 && p->contact != 1) return JUMP_ERR_ARG;
  if (p->contact && !(p->friction >= 0.0 && std::isfinite(p->friction))) return JUMP_ERR_ARG;
  jump_ctx* ctx = new (std::nothrow) jump_ctx();
  if (!ctx) return JUMP_ERR_NOMEM;
  ctx->prm = *p;
  ctx->D = p->dim;
  ctx->device = device;
  GridDesc& G = ctx->G;
  G.dim = p->dim;
  G.apic = p->algorithm == JUMP_ALG_APIC ? 1 : 0;
  G.nx = int(p->nx);
  G.ny = int(p->ny);
  G.nz = int(nz);
  G.nxy = G.nx * G.ny;
  G.nn = G.nxy * G.nz;
  // Grid.jl:105-108, 199-204
  G.xmin = p->xmin;
  G.ymin = p->ymin;
  G.zmin = p->dim == 3 ? p->zmin : 0.0;
  G.dx = (p->xmax - p->xmin) / double(double(p->nx) - 1.0);
  G.dy = (p->ymax - p->ymin) / double(double(p->ny) - 1.0);
  G.dz = p->dim == 3 ? (p->zmax - p->zmin) / double(double(p->nz) - 1.0) : 1.0;
  G.dxI = 1.0 / G.dx;
  G.dyI = 1.0 / G.dy;
  G.dzI = 1.0 / G.dz;
  for (int k = 0; k < 3; k++) G.body[k] = p->body[k];
  ctx->NN = G.nn;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return JUMP_ERR_DEVICE;
  }
  int r = alloc_grid(ctx);
  if (r) {
    free_list(ctx->grid_allocs);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return r;
  }
  ctx->hfixed.assign(size_t(G.nn), 0);
  *out = ctx;
  return JUMP_OK;
}

int jump_ctx_destroy(jump_ctx* ctx) {
  if (!ctx) return JUMP_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->dump_writer.stop();  // writes what is queued
  for (auto& b : ctx->dump_bufs) {
    if (b->host) (void)hipHostFree(b->host);
    if (b->ev) (void)hipEventDestroy(b->ev);
  }
  free_list(ctx->adj.allocs);
  free_list(ctx->adj.adj_allocs);
  free_list(ctx->part_allocs);
  free_list(ctx->grid_allocs);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return JUMP_OK;
}

const char* jump_last_error(const jump_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int jump_add_solid(jump_ctx* ctx, const jump_solid_desc* ds, const jump_material* mat, int32_t* idx) {
  if (!ctx) return JUMP_ERR_ARG;
  if (!ds || !mat) return set_err(ctx, JUMP_ERR_ARG, "null solid or material");
  const int D = ctx->D, S = nsym(D);
  if (ds->n < 0 || (ds->n > 0 && (!ds->pos || !ds->vel || !ds->mass || !ds->volume || !ds->volume0)))
    return set_err(ctx, JUMP_ERR_ARG, "missing particle arrays");
  if (mat->kind < JUMP_MAT_RIGID || mat->kind > JUMP_MAT_JOHNSONCOOK) return set_err(ctx, JUMP_ERR_ARG, "bad material");
  if (D == 3 && mat->kind == JUMP_MAT_RIGID)
    return set_err(ctx, JUMP_ERR_UNSUPPORTED, "3D solids cannot be rigid (Solid3D has no rigid flag)");
  if (ctx->D == 2 && ctx->prm.algorithm == JUMP_ALG_MUSL && mat->kind == JUMP_MAT_RIGID)
    return set_err(ctx, JUMP_ERR_UNSUPPORTED, "the 2D MUSL step has no rigid bodies");
  if (D == 2 && mat->kind == JUMP_MAT_JOHNSONCOOK) return set_err(ctx, JUMP_ERR_UNSUPPORTED, "Johnson-Cook is 3D only");
  if (ctx->np + ds->n > kMaxParticles) return set_err(ctx, JUMP_ERR_UNSUPPORTED, "too many particles (at most 2^28 per context)");
  const bool cpdi = ctx->prm.basis == JUMP_BASIS_CPDI_Q4;
  if (cpdi && (!ds->corners || !ds->elems || ds->ncorners <= 0))
    return set_err(ctx, JUMP_ERR_ARG, "CPDI solid needs corners and elems");
  double lp[3] = {0.0, 0.0, 0.0};
  if (ctx->prm.basis == JUMP_BASIS_GIMP && mat->kind != JUMP_MAT_RIGID) {
    const double hs[3] = {ctx->G.dx, ctx->G.dy, ctx->G.dz};
    for (int k = 0; k < D; k++) {
      if (!(ds->lp[k] > 0.0) || ds->lp[k] > 0.5 * hs[k])
        return set_err(ctx, JUMP_ERR_ARG, "GIMP lp must be in (0, dx/2]");
      lp[k] = ds->lp[k];
    }
  }
  const size_t n = size_t(ds->n);
  // symmetric inputs only (the reference's strain/stress are symmetric)
  for (size_t i = 0; i < n; i++)
    for (int a = 0; a < D; a++)
      for (int b = a + 1; b < D; b++) {
        if (ds->strain && ds->strain[D * D * i + a + D * b] != ds->strain[D * D * i + b + D * a])
          return set_err(ctx, JUMP_ERR_ARG, "strain must be symmetric");
        if (ds->stress && ds->stress[D * D * i + a + D * b] != ds->stress[D * D * i + b + D * a])
          return set_err(ctx, JUMP_ERR_ARG, "stress must be symmetric");
      }
  if (cpdi)
    for (size_t i = 0; i < 4 * n; i++)
      if (ds->elems[i] < 0 || ds->elems[i] >= ds->ncorners) return set_err(ctx, JUMP_ERR_ARG, "CPDI elems out of range");
  // bring the current device state back before growing the arrays
  int r = download(ctx);
  if (r) return r;
  ctx->on_device = false;
  HostSolid hs;
  hs.n = ds->n;
  hs.mat = *mat;
  hs.dev = make_solid_dev(*mat, lp);
  if (hs.dev.rigid) {
    hs.dev.rigid_rank = ctx->max_rigid_rank + 1;  // rigid solids are overridden in solid order
    ctx->max_rigid_rank = hs.dev.rigid_rank;
  }
  hs.offset = ctx->np;
  const int sid = int(ctx->solids.size());
  const int coff = int(ctx->corners.size() / 2);
  if (cpdi) ctx->corners.insert(ctx->corners.end(), ds->corners, ds->corners + 2 * ds->ncorners);
  for (size_t i = 0; i < n; i++) {
    for (int k = 0; k < D; k++) {
      ctx->hx[k].push_back(ds->pos[D * i + k]);
      ctx->hv[k].push_back(ds->vel[D * i + k]);
    }
    ctx->hm.push_back(ds->mass[i]);
    ctx->hV.push_back(ds->volume[i]);
    ctx->hV0.push_back(ds->volume0[i]);
    for (int k = 0; k < D * D; k++)
      ctx->hF[k].push_back(ds->F ? ds->F[D * D * i + k] : ((k % (D + 1)) == 0 ? 1.0 : 0.0));
    // compressed symmetric components in storage order
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    const int ia2[3] = {0, 1, 0}, ib2[3] = {0, 1, 1};
    for (int c = 0; c < S; c++) {
      int a = D == 2 ? ia2[c] : ia[c], b = D == 2 ? ib2[c] : ib[c];
      ctx->heps[c].push_back(ds->strain ? ds->strain[D * D * i + a + D * b] : 0.0);
      ctx->hsig[c].push_back(ds->stress ? ds->stress[D * D * i + a + D * b] : 0.0);
      ctx->hep[c].push_back(0.0);
    }
    ctx->halpha.push_back(0.0);
    ctx->halphadot.push_back(0.0);
    ctx->hvm.push_back(0.0);
    ctx->hsid.push_back(sid);
    if (D == 2)
      for (int k = 0; k < 2; k++) ctx->hX0[k].push_back(ds->pos[2 * i + k]);
    if (ctx->G.apic)
      for (int k = 0; k < 4; k++) ctx->hC[k].push_back(0.0);  // Solid.jl:111
    if (cpdi)
      for (int q = 0; q < 4; q++) ctx->helem[q].push_back(coff + int(ds->elems[4 * i + q]));
  }
  ctx->np += ds->n;
  ctx->solids.push_back(hs);
  ctx->adj.ready = false;  // a recorded run no longer matches the particles
  if (idx) *idx = sid;
  return JUMP_OK;
}

int jump_set_fixed(jump_ctx* ctx, const uint8_t* mask) {
  if (!ctx) return JUMP_ERR_ARG;
  if (!mask) return set_err(ctx, JUMP_ERR_ARG, "null mask");
  const int D = ctx->D;
  for (int64_t i = 0; i < ctx->G.nn; i++) {
    uint8_t b = 0;
    for (int k = 0; k < D; k++)
      if (mask[D * i + k]) b |= uint8_t(1u << k);
    ctx->hfixed[size_t(i)] = b;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(ctx->d_fixed, ctx->hfixed.data(), ctx->hfixed.size(), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return JUMP_OK;
}

int jump_step(jump_ctx* ctx, int64_t nsteps, double* t_inout) {
  if (!ctx) return JUMP_ERR_ARG;
  if (!t_inout || nsteps < 0) return set_err(ctx, JUMP_ERR_ARG, "bad step arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  int r = ensure_device(ctx);
  if (r) return r;
  // the particles move on from a recorded run: its pullback would rewind them
  if (nsteps > 0) ctx->adj.ready = false;
  if (ctx->np == 0) {
    for (int64_t s = 0; s < nsteps; s++) *t_inout += ctx->prm.dt;
    return JUMP_OK;
  }
  r = prepare_cpdi(ctx);
  if (r) return r;
  r = run_steps(ctx, nsteps, *t_inout);
  if (r) return r;
  unsigned fl = 0;
  HIP_TRY(hipMemcpyAsync(&fl, ctx->d_flags, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemsetAsync(ctx->d_flags, 0, sizeof(unsigned), ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  prof_collect(ctx);
  for (int64_t s = 0; s < nsteps; s++) *t_inout += ctx->prm.dt;
  if (nsteps > 0) ctx->grid_valid = true;
  ctx->flags |= fl;
  if (fl) return set_err(ctx, JUMP_ERR_NUMERIC, "particle error flags raised during the step");
  return JUMP_OK;
}

// ------------------------------------------------------------------ checkpoints
// A checkpoint is every per-particle array of the state (positions,
// velocities, mass, volumes, F, strain, stress, material history) in the
// staging layout, after a fixed header.  The grid is not part of it: the
// step rebuilds every grid field from the particles.
namespace {
struct CkptHeader {
  char magic[8];
  int32_t version, dim;
  int64_t np, narrays;
};
constexpr char kCkptMagic[8] = {'J', 'U', 'M', 'P', 'C', 'K', 'P', 'T'};
}  // namespace

static std::vector<std::vector<double>*> state_arrays(jump_ctx* ctx) {
  const int D = ctx->D, S = nsym(D);
  std::vector<std::vector<double>*> a;
  for (int k = 0; k < D; k++) a.push_back(&ctx->hx[k]);
  for (int k = 0; k < D; k++) a.push_back(&ctx->hv[k]);
  a.push_back(&ctx->hm);
  a.push_back(&ctx->hV);
  a.push_back(&ctx->hV0);
  for (int k = 0; k < D * D; k++) a.push_back(&ctx->hF[k]);
  for (int k = 0; k < S; k++) a.push_back(&ctx->heps[k]);
  for (int k = 0; k < S; k++) a.push_back(&ctx->hsig[k]);
  a.push_back(&ctx->halpha);
  a.push_back(&ctx->halphadot);
  a.push_back(&ctx->hvm);
  for (int k = 0; k < S; k++) a.push_back(&ctx->hep[k]);
  if (ctx->G.apic)
    for (int k = 0; k < 4; k++) a.push_back(&ctx->hC[k]);
  return a;
}

int jump_checkpoint_bytes(jump_ctx* ctx, int64_t* bytes) {
  if (!ctx || !bytes) return JUMP_ERR_ARG;
  *bytes = int64_t(sizeof(CkptHeader)) + int64_t(state_arrays(ctx).size()) * ctx->np * int64_t(sizeof(double));
  return JUMP_OK;
}

int jump_checkpoint_save(jump_ctx* ctx, void* buf, int64_t bytes) {
  if (!ctx || !buf) return JUMP_ERR_ARG;
  int64_t need = 0;
  jump_checkpoint_bytes(ctx, &need);
  if (bytes < need) return set_err(ctx, JUMP_ERR_ARG, "checkpoint buffer too small");
  if (ctx->on_device) {
    HIP_TRY(hipSetDevice(ctx->device));
    int r = download(ctx);
    if (r) return r;
  }
  auto arrays = state_arrays(ctx);
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = 1;
  h.dim = ctx->D;
  h.np = ctx->np;
  h.narrays = int64_t(arrays.size());
  char* out = static_cast<char*>(buf);
  std::memcpy(out, &h, sizeof h);
  out += sizeof h;
  for (auto* a : arrays) {
    std::memcpy(out, a->data(), size_t(ctx->np) * sizeof(double));
    out += size_t(ctx->np) * sizeof(double);
  }
  return JUMP_OK;
}

int jump_checkpoint_load(jump_ctx* ctx, const void* buf, int64_t bytes) {
  if (!ctx || !buf) return JUMP_ERR_ARG;
  if (bytes < int64_t(sizeof(CkptHeader))) return set_err(ctx, JUMP_ERR_ARG, "checkpoint too short");
  CkptHeader h;
  std::memcpy(&h, buf, sizeof h);
  auto arrays = state_arrays(ctx);
  if (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != 1)
    return set_err(ctx, JUMP_ERR_ARG, "not a checkpoint of this library");
  if (h.dim != ctx->D || h.np != ctx->np || h.narrays != int64_t(arrays.size()))
    return set_err(ctx, JUMP_ERR_ARG, "checkpoint does not match this context's solids");
  int64_t need = 0;
  jump_checkpoint_bytes(ctx, &need);
  if (bytes < need) return set_err(ctx, JUMP_ERR_ARG, "checkpoint truncated");
  const char* in = static_cast<const char*>(buf) + sizeof h;
  for (auto* a : arrays) {
    a->resize(size_t(ctx->np));
    std::memcpy(a->data(), in, size_t(ctx->np) * sizeof(double));
    in += size_t(ctx->np) * sizeof(double);
  }
  ctx->on_device = false;  // the next call that needs the device uploads the restored state
  ctx->grid_valid = false;
  ctx->flags = 0;
  ctx->adj.ready = false;  // a recorded run no longer matches the particles
  return JUMP_OK;
}

int jump_flags(jump_ctx* ctx, uint32_t* flags) {
  if (!ctx || !flags) return JUMP_ERR_ARG;
  *flags = ctx->flags;
  ctx->flags = 0;
  return JUMP_OK;
}

// Copy the particle range of one solid, converting SoA -> AoS.
int jump_get_solid(jump_ctx* ctx, int32_t si, double* pos, double* vel, double* mass, double* vol, double* F,
                   double* strain, double* stress) {
  if (!ctx) return JUMP_ERR_ARG;
  if (si < 0 || size_t(si) >= ctx->solids.size()) return set_err(ctx, JUMP_ERR_ARG, "bad solid index");
  HIP_TRY(hipSetDevice(ctx->device));
  int r = download(ctx);
  if (r) return r;
  const HostSolid& s = ctx->solids[size_t(si)];
  const int D = ctx->D;
  const size_t o = size_t(s.offset);
  for (size_t i = 0; i < size_t(s.n); i++) {
    for (int k = 0; k < D; k++) {
      if (pos) pos[D * i + k] = ctx->hx[k][o + i];
      if (vel) vel[D * i + k] = ctx->hv[k][o + i];
    }
    if (mass) mass[i] = ctx->hm[o + i];
    if (vol) vol[i] = ctx->hV[o + i];
    for (int k = 0; k < D * D; k++)
      if (F) F[D * D * i + k] = ctx->hF[k][o + i];
    for (int a = 0; a < D; a++)
      for (int b = 0; b < D; b++) {
        int c = a == b ? a : (D == 2 ? 2 : (a + b == 1 ? 3 : (a + b == 2 ? 4 : 5)));
        if (strain) strain[D * D * i + a + D * b] = ctx->heps[c][o + i];
        if (stress) stress[D * D * i + a + D * b] = ctx->hsig[c][o + i];
      }
  }
  return JUMP_OK;
}

int jump_get_solid_state(jump_ctx* ctx, int32_t si, double* alpha, double* alphadot, double* vm, double* pstrain) {
  if (!ctx) return JUMP_ERR_ARG;
  if (si < 0 || size_t(si) >= ctx->solids.size()) return set_err(ctx, JUMP_ERR_ARG, "bad solid index");
  HIP_TRY(hipSetDevice(ctx->device));
  int r = download(ctx);
  if (r) return r;
  const HostSolid& s = ctx->solids[size_t(si)];
  const int D = ctx->D;
  const size_t o = size_t(s.offset);
  for (size_t i = 0; i < size_t(s.n); i++) {
    if (alpha) alpha[i] = ctx->halpha[o + i];
    if (alphadot) alphadot[i] = ctx->halphadot[o + i];
    if (vm) vm[i] = ctx->hvm[o + i];
    if (pstrain) {
      if (D == 2) {
        for (int c = 0; c < 3; c++) pstrain[3 * i + c] = ctx->hep[c][o + i];
      } else {
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            int c = a == b ? a : (a + b == 1 ? 3 : (a + b == 2 ? 4 : 5));
            pstrain[9 * i + a + 3 * b] = ctx->hep[c][o + i];
          }
      }
    }
  }
  return JUMP_OK;
}

int jump_get_affine(jump_ctx* ctx, int32_t si, double* cmat) {
  if (!ctx) return JUMP_ERR_ARG;
  if (si < 0 || size_t(si) >= ctx->solids.size() || !cmat) return set_err(ctx, JUMP_ERR_ARG, "bad solid or null array");
  if (!ctx->G.apic) return set_err(ctx, JUMP_ERR_UNSUPPORTED, "affine matrices exist in APIC contexts only");
  HIP_TRY(hipSetDevice(ctx->device));
  int r = download(ctx);
  if (r) return r;
  const HostSolid& s = ctx->solids[size_t(si)];
  const size_t o = size_t(s.offset);
  for (size_t i = 0; i < size_t(s.n); i++)
    for (int k = 0; k < 4; k++) cmat[4 * i + size_t(k)] = ctx->hC[k][o + i];
  return JUMP_OK;
}

int jump_set_affine(jump_ctx* ctx, int32_t si, const double* cmat) {
  if (!ctx) return JUMP_ERR_ARG;
  if (si < 0 || size_t(si) >= ctx->solids.size() || !cmat) return set_err(ctx, JUMP_ERR_ARG, "bad solid or null array");
  if (!ctx->G.apic) return set_err(ctx, JUMP_ERR_UNSUPPORTED, "affine matrices exist in APIC contexts only");
  HIP_TRY(hipSetDevice(ctx->device));
  int r = download(ctx);  // the other solids' current state
  if (r) return r;
  const HostSolid& s = ctx->solids[size_t(si)];
  const size_t o = size_t(s.offset);
  for (size_t i = 0; i < size_t(s.n); i++)
    for (int k = 0; k < 4; k++) ctx->hC[k][o + i] = cmat[4 * i + size_t(k)];
  ctx->on_device = false;  // the next call that needs the device uploads it
  ctx->adj.ready = false;
  return JUMP_OK;
}

int jump_set_body_force(jump_ctx* ctx, int32_t kind, const double* prm, int32_t nprm) {
  if (!ctx) return JUMP_ERR_ARG;
  if (!prm && nprm > 0) return set_err(ctx, JUMP_ERR_ARG, "null parameters");
  if (kind == JUMP_BODY_CONSTANT) {
    if (nprm != ctx->D) return set_err(ctx, JUMP_ERR_ARG, "a constant body force takes dim values");
    for (int k = 0; k < ctx->D; k++) ctx->G.body[k] = ctx->prm.body[k] = prm[k];
  } else if (kind == JUMP_BODY_VORTEX2D || kind == JUMP_BODY_VORTEX2D_X0) {
    if (nprm != 8) return set_err(ctx, JUMP_ERR_ARG, "VortexBodyForce2D takes 8 values");
    if (ctx->D != 2) return set_err(ctx, JUMP_ERR_UNSUPPORTED, "the vortex body force is 2D");
    if (!(prm[1] != 0.0 && prm[2] != prm[3] && prm[4] > 0.0))
      return set_err(ctx, JUMP_ERR_ARG, "bad vortex parameters (T != 0, Ri != Ro, rho0 > 0)");
    for (int k = 0; k < 8; k++) ctx->vortex[k] = prm[k];
  } else {
    return set_err(ctx, JUMP_ERR_ARG, "bad body force kind");
  }
  const int field = kind == JUMP_BODY_CONSTANT ? 0 : 1;
  if (field != ctx->G.bodyfield || kind != ctx->body_kind) {  // the particle arrays change
    HIP_TRY(hipSetDevice(ctx->device));
    int r = download(ctx);
    if (r) return r;
    ctx->on_device = false;
    ctx->G.bodyfield = field;
  }
  ctx->body_kind = kind;
  ctx->adj.ready = false;
  return JUMP_OK;
}

int jump_get_grid(jump_ctx* ctx, double* mass, double* p0, double* p, double* p2, double* f) {
  if (!ctx) return JUMP_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t nn = size_t(ctx->G.nn);
  const int D = ctx->D;
  if (!ctx->grid_valid) {
    if (mass) std::fill(mass, mass + nn, 0.0);
    double* vs[4] = {p0, p, p2, f};
    for (double* v : vs)
      if (v) std::fill(v, v + D * nn, 0.0);
    return JUMP_OK;
  }
  std::vector<double> tmp(nn);
  if (mass) HIP_TRY(hipMemcpyAsync(mass, ctx->g.m, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  double* outs[4] = {p0, p, p2, f};
  double* const* srcs[4] = {ctx->g.p0, ctx->g.p, ctx->g.p2, ctx->g.f};
  for (int a = 0; a < 4; a++) {
    if (!outs[a]) continue;
    for (int k = 0; k < D; k++) {
      HIP_TRY(hipMemcpyAsync(tmp.data(), srcs[a][k], nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      for (size_t i = 0; i < nn; i++) outs[a][D * i + k] = tmp[i];
    }
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return JUMP_OK;
}

// KE and SE of the particles [begin, end), reduced over blocks of kBlock
// particles aligned at begin, then over the block sums.
static int range_energies(jump_ctx* ctx, int64_t begin, int64_t end, double* ke, double* se) {
  if (end <= begin) {
    *ke = 0.0;
    *se = 0.0;
    return JUMP_OK;
  }
  const int nb = blocks_for(end - begin);
  double* out = ctx->d_red + 2 * size_t(nb);
  if (ctx->D == 2) LAUNCH((k_energy_partial<2, EnergySoA<2>>), end - begin, EnergySoA<2>{ctx->P}, int(begin), int(end), ctx->d_red);
  else LAUNCH((k_energy_partial<3, EnergySoA<3>>), end - begin, EnergySoA<3>{ctx->P}, int(begin), int(end), ctx->d_red);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->d_red, nb, out);
  double h[2];
  HIP_TRY(hipMemcpyAsync(h, out, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  *ke = h[0];
  *se = h[1];
  return JUMP_OK;
}

int jump_energies(jump_ctx* ctx, double* ke, double* se) {
  if (!ctx) return JUMP_ERR_ARG;
  if (!ke || !se) return set_err(ctx, JUMP_ERR_ARG, "null output");
  HIP_TRY(hipSetDevice(ctx->device));
  int r = ensure_device(ctx);
  if (r) return r;
  if (ctx->np == 0) {
    *ke = 0.0;
    *se = 0.0;
    return JUMP_OK;
  }
  r = range_energies(ctx, 0, ctx->np, ke, se);
  return r;
}

}  // extern "C"

// ------------------------------------------------------------------ energies without allocation
// jump_energies_noalloc: the reduction of jump_energies over a caller's device
// arrays, with its per-block partial sums in `scratch`, a fixed device array
// of this library (2 doubles per block of kBlock particles, 1 MiB: at most
// kNoAllocMaxBlocks * kBlock = 16,777,216 particles per call).  Each device
// holds its own copy of the array; its address is looked up once per device
// (hipGetSymbolAddress, with that device current) and kept.  Every call on a
// device shares that device's array, so calls that overlap on one device
// race; the partials are fully rewritten by each call, nothing has to start
// at zero.
namespace {
constexpr int kNoAllocMaxBlocks = 1 << 16;
static __device__ __align__(256) double scratch[2 * kNoAllocMaxBlocks];
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of this device's `scratch` (the array's name in host
// code is a host-side stand-in, not a device address).
int scratch_ptr(int dev, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return JUMP_ERR_DEVICE;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return JUMP_OK;
}
}  // namespace

extern "C" int jump_energies_noalloc(int32_t dim, int64_t n, const double* mass, const double* vel,
                                     const double* volume, const double* strain, const double* stress,
                                     double* ke_se, void* stream) {
  if (dim != 2 && dim != 3) return JUMP_ERR_ARG;
  if (n < 0 || !ke_se) return JUMP_ERR_ARG;
  if (n > 0 && (!mass || !vel || !volume || !strain || !stress)) return JUMP_ERR_ARG;
  if (n > int64_t(kNoAllocMaxBlocks) * kBlock) return JUMP_ERR_UNSUPPORTED;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return JUMP_ERR_DEVICE;
  if (dev >= kNoAllocMaxDevices) return JUMP_ERR_UNSUPPORTED;
  double* part = nullptr;
  int r = scratch_ptr(dev, &part);
  if (r) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = blocks_for(n);
  if (nb > 0) {
    if (dim == 2)
      hipLaunchKernelGGL((k_energy_partial<2, EnergyAoS<2>>), dim3(nb), dim3(kBlock), 0, st,
                         EnergyAoS<2>{mass, vel, volume, strain, stress}, 0, int(n), part);
    else
      hipLaunchKernelGGL((k_energy_partial<3, EnergyAoS<3>>), dim3(nb), dim3(kBlock), 0, st,
                         EnergyAoS<3>{mass, vel, volume, strain, stress}, 0, int(n), part);
  }
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, st, part, nb, ke_se);  // nb = 0 writes zeros
  return hipGetLastError() == hipSuccess ? JUMP_OK : JUMP_ERR_DEVICE;
}

extern "C" {

int jump_dump_lammps(jump_ctx* ctx, const char* path, const char* outs) {
  if (!ctx) return JUMP_ERR_ARG;
  if (!path || !*path) return set_err(ctx, JUMP_ERR_ARG, "empty dump path");
  const int D = ctx->D;
  DumpJob job;
  if (!parse_dump_cols(outs, D, job.names, job.cols))
    return set_err(ctx, JUMP_ERR_UNSUPPORTED, std::string("unknown dump column in \"") + outs + "\"");
  HIP_TRY(hipSetDevice(ctx->device));
  int r = ensure_device(ctx);
  if (r) return r;
  // the pair arrays and history fields the columns read (positions always)
  bool want[7] = {}, walpha = false, wvm = false;
  const int fv = D + 2, fs = 2 * D + 2;
  for (int a = 0; a < D; a++) want[a / 2] = true;
  for (int c : job.cols) {
    if (c == kColPstrain) walpha = true;
    if (c == kColVonMises) wvm = true;
    if (c == kColVx || c == kColVy || c == kColVz)
      for (int a = 0; a < D; a++) want[(fv + a) / 2] = true;
    if (c == kColSxx || c == kColSyy || c == kColPressure)
      for (int k = 0; k < D; k++) want[(fs + k) / 2] = true;
  }
  const size_t n = size_t(ctx->np);
  size_t bytes = (walpha ? n * 8 : 0) + (wvm ? n * 8 : 0);
  for (bool w : want) bytes += w ? n * 16 : 0;
  bytes = std::max<size_t>(bytes, 16);
  // a free buffer set (at most two: a third dump waits for the writer)
  jump_ctx::DumpBuf* buf = nullptr;
  for (auto& b : ctx->dump_bufs)
    if (!b->busy.load()) { buf = b.get(); break; }
  if (!buf && ctx->dump_bufs.size() < 2) {
    ctx->dump_bufs.push_back(std::make_unique<jump_ctx::DumpBuf>());
    buf = ctx->dump_bufs.back().get();
  }
  if (!buf) {
    // wait() reports (and clears) a failed earlier dump: keep it for jump_dump_wait
    if (!ctx->dump_writer.wait()) ctx->dump_failed = true;
    buf = ctx->dump_bufs[0].get();
  }
  if (buf->bytes < bytes) {
    if (buf->host) HIP_TRY(hipHostFree(buf->host));
    buf->host = nullptr;
    buf->bytes = 0;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&buf->host), bytes, hipHostMallocDefault));
    buf->bytes = bytes;
  }
  if (!buf->ev) HIP_TRY(hipEventCreateWithFlags(&buf->ev, hipEventDisableTiming));
  buf->busy = true;
  char* q = buf->host;
  for (int k = 0; k < 7; k++) {
    if (!want[k]) continue;
    HIP_TRY(hipMemcpyAsync(q, ctx->P.pr[k], n * 16, hipMemcpyDeviceToHost, ctx->stream));
    job.pr[k] = reinterpret_cast<const double*>(q);
    q += n * 16;
  }
  if (walpha) {
    HIP_TRY(hipMemcpyAsync(q, ctx->P.alpha, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    job.alpha = reinterpret_cast<const double*>(q);
    q += n * 8;
  }
  if (wvm) {
    HIP_TRY(hipMemcpyAsync(q, ctx->P.vm, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    job.vm = reinterpret_cast<const double*>(q);
  }
  HIP_TRY(hipEventRecord(buf->ev, ctx->stream));
  job.path = path;
  job.dim = D;
  job.np = ctx->np;
  job.lims[0] = ctx->prm.xmax - ctx->prm.xmin;  // Grid.jl: lx = xmax - xmin
  job.lims[1] = ctx->prm.ymax - ctx->prm.ymin;
  job.lims[2] = D == 3 ? ctx->prm.zmax - ctx->prm.zmin : 0.0;
  for (const HostSolid& hs : ctx->solids) {
    job.solid_end.push_back(hs.offset + hs.n);
    // Output.jl:176-177 (2D: elastic or rigid), :243 (3D: elastic)
    job.elastic.push_back(hs.mat.kind == JUMP_MAT_ELASTIC || (D == 2 && hs.mat.kind == JUMP_MAT_RIGID));
  }
  hipEvent_t ev = buf->ev;
  job.ready = [ev] { return hipEventSynchronize(ev) == hipSuccess; };
  job.release = [buf] { buf->busy = false; };
  ctx->dump_writer.push(std::move(job));
  return JUMP_OK;
}

int jump_dump_wait(jump_ctx* ctx) {
  if (!ctx) return JUMP_ERR_ARG;
  const bool ok = ctx->dump_writer.wait() && !ctx->dump_failed;
  ctx->dump_failed = false;
  if (!ok) return set_err(ctx, JUMP_ERR_IO, "a LAMMPS dump could not be written");
  return JUMP_OK;
}

int jump_get_bins(jump_ctx* ctx, int32_t* keys) {
  if (!ctx || !keys) return JUMP_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  int r = ensure_device(ctx);
  if (r) return r;
  if (ctx->np == 0) return JUMP_OK;
  r = prepare_cpdi(ctx);
  if (r) return r;
  if (ctx->D == 2) {
    switch (ctx->prm.basis) {
      case JUMP_BASIS_LINEAR: bin_and_sort<2, JUMP_BASIS_LINEAR, false>(ctx); break;
      case JUMP_BASIS_QUAD_BSPLINE: bin_and_sort<2, JUMP_BASIS_QUAD_BSPLINE, false>(ctx); break;
      case JUMP_BASIS_CUBIC_BSPLINE: bin_and_sort<2, JUMP_BASIS_CUBIC_BSPLINE, false>(ctx); break;
      case JUMP_BASIS_CPDI_Q4: bin_and_sort<2, JUMP_BASIS_CPDI_Q4, false>(ctx); break;
      case JUMP_BASIS_GIMP: bin_and_sort<2, JUMP_BASIS_GIMP, false>(ctx); break;
    }
  } else {
    switch (ctx->prm.basis) {
      case JUMP_BASIS_LINEAR: bin_and_sort<3, JUMP_BASIS_LINEAR, false>(ctx); break;
      case JUMP_BASIS_QUAD_BSPLINE: bin_and_sort<3, JUMP_BASIS_QUAD_BSPLINE, false>(ctx); break;
      case JUMP_BASIS_CUBIC_BSPLINE: bin_and_sort<3, JUMP_BASIS_CUBIC_BSPLINE, false>(ctx); break;
      case JUMP_BASIS_GIMP: bin_and_sort<3, JUMP_BASIS_GIMP, false>(ctx); break;
    }
  }
  ctx->runs_valid = false;  // the bins now differ from those the runs were recorded on
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(keys, ctx->d_key, size_t(ctx->np) * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemsetAsync(ctx->d_flags, 0, sizeof(unsigned), ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return JUMP_OK;
}

int jump_profile(jump_ctx* ctx, int32_t enable) {
  if (!ctx) return JUMP_ERR_ARG;
  ctx->prof = enable != 0;
  for (int k = 0; k < JUMP_PHASE_COUNT; k++) {
    ctx->prof_ms[k] = 0.0;
    ctx->prof_cnt[k] = 0;
  }
  ctx->prof_recs.clear();
  ctx->ev_used = 0;
  return JUMP_OK;
}

int jump_profile_read(jump_ctx* ctx, int32_t phase, double* total_ms, int64_t* count) {
  if (!ctx || phase < 0 || phase >= JUMP_PHASE_COUNT || !total_ms || !count) return JUMP_ERR_ARG;
  *total_ms = ctx->prof_ms[phase];
  *count = ctx->prof_cnt[phase];
  return JUMP_OK;
}

int jump_abi_version(void) { return JUMP_ABI_VERSION; }

}  // extern "C"

// ------------------------------------------------------------------ batches
// A batch steps several independent 2D USL contexts with shared launches.
// Members with the same basis and summation order form one group, held by an
// internal context whose particles are the members' particles back to back
// and whose nodes are the members' grids back to back (instance i at
// node_off[i]); every kernel finds the instance of its particle or node and
// works in that instance's grid.  Each member's particles keep their order
// and each member's reductions their tree, so a member steps bit for bit as
// it would alone.
//
// Launches.  Each group's step is captured once into a HIP graph (one step,
// and a chunk of kGraphChunk steps) and replayed, so a step costs one graph
// launch per group instead of about fifteen kernel launches; the graphs are
// captured again if the group's particle arrays move.  Before the first step
// the autotuner times the batched gather's launch shapes on the group's own
// particles and keeps the fastest (kernel times: about 0.1 to 1 ms per
// launch on 256 instances, where shapes differ by up to tens of percent).
constexpr int kGraphChunk = 8;
struct jump_batch {
  int device = 0;
  std::string err;
  std::vector<jump_ctx*> members;
  bool use_graph = true, autotune = true;
  struct Group {
    jump_ctx* inner = nullptr;
    std::vector<int> members;        // member index of each instance
    std::vector<int64_t> part_off;   // first particle of each instance in the group
    bool tuned = false;
    double tune_ms[JUMP_P2G_SHAPES] = {};  // the autotuner's ms per launch of each shape
    hipGraphExec_t exec1 = nullptr, execN = nullptr;  // one step, kGraphChunk steps
    int graph_gen = -1, graph_shape = -1;              // what the graphs were captured with
  };
  std::vector<Group> groups;
};

static void drop_graphs(jump_batch::Group& g) {
  if (g.exec1) (void)hipGraphExecDestroy(g.exec1);
  if (g.execN) (void)hipGraphExecDestroy(g.execN);
  g.exec1 = g.execN = nullptr;
  g.graph_gen = g.graph_shape = -1;
}

// Capture nsteps steps of the group's launch sequence (no host
// synchronisation, memset or allocation happens inside a 2D USL step).
static int capture_steps(jump_ctx* in, int64_t nsteps, hipGraphExec_t* out) {
  hipGraph_t graph = nullptr;
  if (hipStreamBeginCapture(in->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return set_err(in, JUMP_ERR_DEVICE, "hipStreamBeginCapture failed");
  const int r = run_steps(in, nsteps);
  const hipError_t e = hipStreamEndCapture(in->stream, &graph);
  if (r || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return r ? r : set_err(in, JUMP_ERR_DEVICE, "hipStreamEndCapture failed");
  }
  const hipError_t e2 = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e2 != hipSuccess) return set_err(in, JUMP_ERR_DEVICE, "hipGraphInstantiate failed");
  return JUMP_OK;
}

// The autotuner: bin once, then time each batched gather shape (one warm-up
// launch and kTuneReps timed ones, HIP events on the group's stream) and
// keep the fastest.  The gather only writes the grid, which the next step
// rebuilds, so the particle state is untouched.
constexpr int kTuneReps = 3;
template <int B, int ORDER> static int tune_batched_p2g(jump_ctx* in, double* ms) {
  // the batched kernels read the instance table, which only a group of two
  // or more members has (a group of one steps as a single context)
  if (in->ninst < 2 || !in->d_inst || !in->d_node_off)
    return set_err(in, JUMP_ERR_STATE, "only a group of several instances has a batched gather to tune");
  bin_and_sort<2, B, true>(in);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return set_err(in, JUMP_ERR_DEVICE, "hipEventCreate failed");
  }
  int best = 0, r = JUMP_OK;
  for (int sh = 0; sh < JUMP_P2G_SHAPES && !r; sh++) {
    p2g_batched<B, ORDER>(in, sh);
    float t = 0.0f;
    if (hipEventRecord(e0, in->stream) != hipSuccess) r = JUMP_ERR_DEVICE;
    for (int k = 0; k < kTuneReps && !r; k++) p2g_batched<B, ORDER>(in, sh);
    if (!r && (hipEventRecord(e1, in->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
               hipEventElapsedTime(&t, e0, e1) != hipSuccess))
      r = JUMP_ERR_DEVICE;
    ms[sh] = double(t) / kTuneReps;
    if (ms[sh] < ms[best]) best = sh;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (r) return set_err(in, r, "autotuner timing failed");
  in->p2g_shape = best;
  HIP_TRY_CTX_LAST(in);
  return JUMP_OK;
}
static int tune_group(jump_ctx* in, double* ms) {
  const bool ref = in->prm.order == JUMP_ORDER_REFERENCE;
  switch (in->prm.basis) {
    case JUMP_BASIS_LINEAR:
      return ref ? tune_batched_p2g<JUMP_BASIS_LINEAR, JUMP_ORDER_REFERENCE>(in, ms)
                 : tune_batched_p2g<JUMP_BASIS_LINEAR, JUMP_ORDER_BINNED>(in, ms);
    case JUMP_BASIS_QUAD_BSPLINE:
      return ref ? tune_batched_p2g<JUMP_BASIS_QUAD_BSPLINE, JUMP_ORDER_REFERENCE>(in, ms)
                 : tune_batched_p2g<JUMP_BASIS_QUAD_BSPLINE, JUMP_ORDER_BINNED>(in, ms);
    case JUMP_BASIS_CUBIC_BSPLINE:
      return ref ? tune_batched_p2g<JUMP_BASIS_CUBIC_BSPLINE, JUMP_ORDER_REFERENCE>(in, ms)
                 : tune_batched_p2g<JUMP_BASIS_CUBIC_BSPLINE, JUMP_ORDER_BINNED>(in, ms);
    case JUMP_BASIS_GIMP:
      return ref ? tune_batched_p2g<JUMP_BASIS_GIMP, JUMP_ORDER_REFERENCE>(in, ms)
                 : tune_batched_p2g<JUMP_BASIS_GIMP, JUMP_ORDER_BINNED>(in, ms);
    default:
      return set_err(in, JUMP_ERR_UNSUPPORTED, "no batched gather for this basis");
  }
}

static int batch_err(jump_batch* b, int code, const std::string& m) {
  b->err = m;
  return code;
}

const char* jump_batch_last_error(const jump_batch* b) { return b ? b->err.c_str() : "null batch"; }

int jump_batch_destroy(jump_batch* b) {
  if (!b) return JUMP_OK;
  (void)hipSetDevice(b->device);
  for (auto& g : b->groups) {
    if (g.inner && g.inner->stream) (void)hipStreamSynchronize(g.inner->stream);
    drop_graphs(g);
    jump_ctx_destroy(g.inner);
  }
  delete b;
  return JUMP_OK;
}

int jump_batch_create(jump_ctx* const* ctxs, int32_t n, int device, jump_batch** out) {
  if (!ctxs || n <= 0 || !out) return JUMP_ERR_ARG;
  *out = nullptr;
  for (int i = 0; i < n; i++) {
    const jump_ctx* c = ctxs[i];
    if (!c) return JUMP_ERR_ARG;
    if (c->D != 2 || c->prm.algorithm != JUMP_ALG_USL) return JUMP_ERR_UNSUPPORTED;  // batches are 2D USL
    if (c->G.bodyfield) return JUMP_ERR_UNSUPPORTED;  // the graphs replay fixed times
    if (c->prm.basis == JUMP_BASIS_CPDI_Q4 || c->prm.contact) return JUMP_ERR_UNSUPPORTED;
    if (c->device != device) return JUMP_ERR_ARG;
  }
  jump_batch* b = new (std::nothrow) jump_batch();
  if (!b) return JUMP_ERR_NOMEM;
  b->device = device;
  b->members.assign(ctxs, ctxs + n);
  // group by (basis, order), members in the given order
  std::vector<std::pair<int, int>> keys;
  std::vector<int> gof(n);
  for (int i = 0; i < n; i++) {
    const std::pair<int, int> k{ctxs[i]->prm.basis, ctxs[i]->prm.order};
    auto it = std::find(keys.begin(), keys.end(), k);
    gof[i] = int(it - keys.begin());
    if (it == keys.end()) keys.push_back(k);
  }
  b->groups.resize(keys.size());
  for (size_t gi = 0; gi < keys.size(); gi++) {
    jump_batch::Group& grp = b->groups[gi];
    for (int i = 0; i < n; i++)
      if (gof[i] == int(gi)) grp.members.push_back(i);
    jump_ctx* first = ctxs[grp.members[0]];
    jump_ctx* in = nullptr;
    int r = jump_ctx_create(&first->prm, device, &in);
    if (r) {
      jump_batch_destroy(b);
      return r;
    }
    grp.inner = in;
    // instances: grids back to back
    in->ninst = int(grp.members.size());
    in->hinst.clear();
    in->hnode_off.assign(1, 0);
    int64_t nn = 0;
    for (int m : grp.members) {
      const jump_ctx* c = ctxs[m];
      in->hinst.push_back(InstDev{c->G, c->prm.dt, c->prm.tolerance});
      nn += c->G.nn;
      if (nn > kMaxNodes) {
        jump_batch_destroy(b);
        return JUMP_ERR_UNSUPPORTED;
      }
      in->hnode_off.push_back(int(nn));
    }
    in->NN = int(nn);
    free_list(in->grid_allocs);
    r = alloc_grid(in);
    if (r) {
      jump_batch_destroy(b);
      return r;
    }
    in->hfixed.clear();
    // particles, solids and Dirichlet masks, member after member
    for (size_t ii = 0; ii < grp.members.size(); ii++) {
      jump_ctx* c = ctxs[grp.members[ii]];
      if (c->on_device) {
        r = hipSetDevice(c->device) == hipSuccess ? download(c) : JUMP_ERR_DEVICE;
        if (r) {
          jump_batch_destroy(b);
          return r;
        }
      }
      grp.part_off.push_back(in->np);
      const int soff = int(in->solids.size());
      for (const HostSolid& hs0 : c->solids) {
        HostSolid hs = hs0;
        hs.offset += in->np;
        hs.dev.inst = int(ii);
        in->solids.push_back(hs);
        in->max_rigid_rank = std::max(in->max_rigid_rank, hs.dev.rigid ? hs.dev.rigid_rank : -1);
      }
      auto cat = [](std::vector<double>& d, const std::vector<double>& s2) { d.insert(d.end(), s2.begin(), s2.end()); };
      for (int k = 0; k < 2; k++) { cat(in->hx[k], c->hx[k]); cat(in->hv[k], c->hv[k]); }
      cat(in->hm, c->hm);
      cat(in->hV, c->hV);
      cat(in->hV0, c->hV0);
      for (int k = 0; k < 4; k++) cat(in->hF[k], c->hF[k]);
      for (int k = 0; k < 3; k++) { cat(in->heps[k], c->heps[k]); cat(in->hsig[k], c->hsig[k]); cat(in->hep[k], c->hep[k]); }
      cat(in->halpha, c->halpha);
      cat(in->halphadot, c->halphadot);
      cat(in->hvm, c->hvm);
      for (int sid : c->hsid) in->hsid.push_back(sid + soff);
      in->np += c->np;
      in->hfixed.insert(in->hfixed.end(), c->hfixed.begin(), c->hfixed.end());
    }
    if (in->np > kMaxParticles) {
      jump_batch_destroy(b);
      return JUMP_ERR_UNSUPPORTED;
    }
    if (hipSetDevice(device) != hipSuccess ||
        hipMemcpyAsync(in->d_fixed, in->hfixed.data(), in->hfixed.size(), hipMemcpyHostToDevice, in->stream) !=
            hipSuccess ||
        hipStreamSynchronize(in->stream) != hipSuccess) {
      jump_batch_destroy(b);
      return JUMP_ERR_DEVICE;
    }
    in->on_device = false;
  }
  *out = b;
  return JUMP_OK;
}

int jump_batch_step(jump_batch* b, int64_t nsteps, double* t_inout) {
  if (!b) return JUMP_ERR_ARG;
  if (!t_inout || nsteps < 0) return batch_err(b, JUMP_ERR_ARG, "bad step arguments");
  HIP_TRY_B(hipSetDevice(b->device));
  unsigned fl = 0;
  for (auto& grp : b->groups) {  // every group's launches are queued before any wait
    jump_ctx* in = grp.inner;
    int r = ensure_device(in);
    if (!r && b->autotune && !grp.tuned && nsteps > 0 && in->ninst > 1) {
      r = tune_group(in, grp.tune_ms);
      grp.tuned = r == JUMP_OK;
    }
    if (!r && b->use_graph && nsteps > 0) {
      if (grp.graph_gen != in->alloc_gen || grp.graph_shape != in->p2g_shape) {
        drop_graphs(grp);
        r = capture_steps(in, 1, &grp.exec1);
        if (!r) r = capture_steps(in, kGraphChunk, &grp.execN);
        if (!r) {
          grp.graph_gen = in->alloc_gen;
          grp.graph_shape = in->p2g_shape;
        }
      }
      int64_t k = 0;
      for (; !r && k + kGraphChunk <= nsteps; k += kGraphChunk)
        if (hipGraphLaunch(grp.execN, in->stream) != hipSuccess) r = set_err(in, JUMP_ERR_DEVICE, "hipGraphLaunch");
      for (; !r && k < nsteps; k++)
        if (hipGraphLaunch(grp.exec1, in->stream) != hipSuccess) r = set_err(in, JUMP_ERR_DEVICE, "hipGraphLaunch");
    } else if (!r) {
      r = run_steps(in, nsteps);
    }
    if (r) return batch_err(b, r, in->err);
  }
  for (auto& grp : b->groups) {
    jump_ctx* in = grp.inner;
    unsigned f = 0;
    HIP_TRY_B(hipMemcpyAsync(&f, in->d_flags, sizeof(unsigned), hipMemcpyDeviceToHost, in->stream));
    HIP_TRY_B(hipMemsetAsync(in->d_flags, 0, sizeof(unsigned), in->stream));
    HIP_TRY_B(hipStreamSynchronize(in->stream));
    if (nsteps > 0) in->grid_valid = true;
    in->flags |= f;
    fl |= f;
  }
  for (size_t m = 0; m < b->members.size(); m++)
    for (int64_t s = 0; s < nsteps; s++) t_inout[m] += b->members[m]->prm.dt;
  if (fl) return batch_err(b, JUMP_ERR_NUMERIC, "particle error flags raised during the step");
  return JUMP_OK;
}

int jump_batch_sync(jump_batch* b) {
  if (!b) return JUMP_ERR_ARG;
  HIP_TRY_B(hipSetDevice(b->device));
  for (auto& grp : b->groups) {
    jump_ctx* in = grp.inner;
    int r = download(in);
    if (r) return batch_err(b, r, in->err);
    for (size_t ii = 0; ii < grp.members.size(); ii++) {
      jump_ctx* c = b->members[grp.members[ii]];
      const int64_t o = grp.part_off[ii], n = c->np;
      auto put = [&](std::vector<double>& d, const std::vector<double>& s2) {
        std::copy(s2.begin() + o, s2.begin() + o + n, d.begin());
      };
      for (int k = 0; k < 2; k++) { put(c->hx[k], in->hx[k]); put(c->hv[k], in->hv[k]); }
      put(c->hm, in->hm);
      put(c->hV, in->hV);
      put(c->hV0, in->hV0);
      for (int k = 0; k < 4; k++) put(c->hF[k], in->hF[k]);
      for (int k = 0; k < 3; k++) { put(c->heps[k], in->heps[k]); put(c->hsig[k], in->hsig[k]); put(c->hep[k], in->hep[k]); }
      put(c->halpha, in->halpha);
      put(c->halphadot, in->halphadot);
      put(c->hvm, in->hvm);
      c->on_device = false;  // the member's own device copy (if any) is stale now
      c->grid_valid = false;
      c->adj.ready = false;
    }
  }
  return JUMP_OK;
}

int jump_batch_energies(jump_batch* b, double* ke, double* se) {
  if (!b || !ke || !se) return JUMP_ERR_ARG;
  HIP_TRY_B(hipSetDevice(b->device));
  for (auto& grp : b->groups) {
    jump_ctx* in = grp.inner;
    int r = ensure_device(in);
    if (r) return batch_err(b, r, in->err);
    for (size_t ii = 0; ii < grp.members.size(); ii++) {
      const int m = grp.members[ii];
      const int64_t o = grp.part_off[ii];
      r = range_energies(in, o, o + b->members[m]->np, &ke[m], &se[m]);
      if (r) return batch_err(b, r, in->err);
    }
  }
  return JUMP_OK;
}

int jump_batch_options(jump_batch* b, int32_t graph, int32_t autotune, int32_t shape) {
  if (!b) return JUMP_ERR_ARG;
  if (!autotune && (shape < 0 || shape >= JUMP_P2G_SHAPES)) return batch_err(b, JUMP_ERR_ARG, "bad gather shape");
  b->use_graph = graph != 0;
  b->autotune = autotune != 0;
  for (auto& grp : b->groups) {
    if (!b->autotune) {
      grp.inner->p2g_shape = shape;
      grp.tuned = false;
    }
  }
  return JUMP_OK;
}

int jump_batch_tuning(jump_batch* b, int32_t g, int32_t* ngroups, int32_t* shape, double* ms) {
  if (!b) return JUMP_ERR_ARG;
  if (ngroups) *ngroups = int32_t(b->groups.size());
  if (g < 0 || size_t(g) >= b->groups.size()) return batch_err(b, JUMP_ERR_ARG, "bad group index");
  const auto& grp = b->groups[size_t(g)];
  if (grp.inner->ninst < 2) {  // one member: stepped as a single context, no batched gather
    if (shape) *shape = -1;
    if (ms)
      for (int k = 0; k < JUMP_P2G_SHAPES; k++) ms[k] = 0.0;
    return JUMP_OK;
  }
  if (!grp.tuned) return batch_err(b, JUMP_ERR_STATE, "the group has not been tuned (no step with autotune on)");
  if (shape) *shape = grp.inner->p2g_shape;
  if (ms)
    for (int k = 0; k < JUMP_P2G_SHAPES; k++) ms[k] = grp.tune_ms[k];
  return JUMP_OK;
}

// ------------------------------------------------------------------ reverse mode
// The 2D USL and 3D MUSL steps differentiated by hand (mpm_adjoint.hpp,
// mpm_adjoint3.hpp).  The recorded run keeps the particle state every K
// steps; the backward sweep rebuilds each segment's states from its snapshot
// (keeping the grid and bins each step used), then walks the segment's steps
// backwards.  Memory: (nsteps/K + K) snapshots of 136 (2D) or 288 (3D)
// bytes per particle, K grids of 40 or 80 bytes per node and K bins.

// Copies the arrays of `segs` (device pointer, bytes) to or from the buffer
// `buf`, laid out back to back with each array starting 256-byte aligned, in
// one launch.
static size_t seg_offset(size_t bytes) { return (bytes + 255) & ~size_t(255); }
static int copy_segments(jump_ctx* ctx, char* buf, const std::vector<std::pair<void*, size_t>>& segs, bool save) {
  CopyTable t{};
  size_t off = 0;
  long long most = 0;
  if (segs.size() > size_t(kCopySegs)) return set_err(ctx, JUMP_ERR_STATE, "too many arrays for one copy launch");
  for (size_t k = 0; k < segs.size(); k++) {
    unsigned* b = reinterpret_cast<unsigned*>(buf + off);
    unsigned* a = static_cast<unsigned*>(segs[k].first);
    t.src[k] = save ? a : b;
    t.dst[k] = save ? b : a;
    t.words[k] = (long long)(segs[k].second / 4);
    most = std::max(most, t.words[k]);
    off += seg_offset(segs[k].second);
  }
  if (segs.empty() || most == 0) return JUMP_OK;
  const unsigned bx = unsigned(std::min<long long>((most / 4 + kBlock) / kBlock, 1024));
  if (save) hipLaunchKernelGGL(k_copy_table<true>, dim3(bx, unsigned(segs.size())), dim3(kBlock), 0, ctx->stream, t);
  else hipLaunchKernelGGL(k_copy_table<false>, dim3(bx, unsigned(segs.size())), dim3(kBlock), 0, ctx->stream, t);
  return JUMP_OK;
}

// A snapshot is the particle state a step reads: the pair arrays (x, m, V,
// v, sigma), F and the strain, and in 3D the J2 history (alpha and the
// plastic strain): 17 arrays of n doubles in 2D, 36 in 3D (pairs counted
// twice), each 256-byte aligned in the buffer.
static std::vector<std::pair<void*, size_t>> snap_segments(const jump_ctx* ctx) {
  const size_t n = size_t(ctx->np);
  const int D = ctx->D;
  std::vector<std::pair<void*, size_t>> arrays;
  const int npair = D == 2 ? PairLayout<2>::NP : PairLayout<3>::NP;
  for (int j = 0; j < npair; j++) arrays.push_back({ctx->P.pr[j], n * sizeof(double2)});
  for (int k = 0; k < D * D; k++) arrays.push_back({ctx->P.F[k], n * sizeof(double)});
  for (int k = 0; k < nsym(D); k++) arrays.push_back({ctx->P.eps[k], n * sizeof(double)});
  if (D == 3) {
    arrays.push_back({ctx->P.alpha, n * sizeof(double)});
    for (int k = 0; k < 6; k++) arrays.push_back({ctx->P.ep[k], n * sizeof(double)});
  }
  return arrays;
}
static size_t snap_bytes(const jump_ctx* ctx) {
  size_t b = 0;
  for (auto& a : snap_segments(ctx)) b += seg_offset(a.second);
  return b;
}

static int snap_copy(jump_ctx* ctx, double* buf, bool save) {
  int r = copy_segments(ctx, reinterpret_cast<char*>(buf), snap_segments(ctx), save);
  if (!save) ctx->runs_valid = false;  // a restored state: the next step merges at every node
  return r;
}

static int adjoint_supported(jump_ctx* ctx) {
  const bool usl2d = ctx->D == 2 && ctx->prm.algorithm == JUMP_ALG_USL;
  const bool musl3d = ctx->D == 3 && ctx->prm.algorithm == JUMP_ALG_MUSL;
  if (!(usl2d || musl3d) || ctx->ninst != 1)
    return set_err(ctx, JUMP_ERR_UNSUPPORTED, "the adjoint covers the 2D USL and 3D MUSL steps");
  if (ctx->prm.contact) return set_err(ctx, JUMP_ERR_UNSUPPORTED, "the adjoint does not cover contact");
  if (ctx->G.bodyfield)
    return set_err(ctx, JUMP_ERR_UNSUPPORTED, "the adjoint covers a constant body force only");
  if (ctx->prm.basis != JUMP_BASIS_LINEAR && ctx->prm.basis != JUMP_BASIS_GIMP)
    return set_err(ctx, JUMP_ERR_UNSUPPORTED, "the adjoint covers the linear and GIMP bases");
  for (auto& s : ctx->solids) {
    const int k = s.dev.kind;
    if (k != JUMP_MAT_ELASTIC && k != JUMP_MAT_NEOHOOKEAN && !(musl3d && k == JUMP_MAT_ELASTOPLASTIC))
      return set_err(ctx, JUMP_ERR_UNSUPPORTED,
                     "the adjoint covers Hooke and Neo-Hookean solids, and J2 solids in 3D");
  }
  return JUMP_OK;
}

// The bins a step sorted its particles into: start (NN+1) and sorted (np+1).
// The bins record also keeps the step's active-node order (d_order,
// d_nodepos), which the backward sweep's node gathers walk.
static std::vector<std::pair<void*, size_t>> bins_segments(const jump_ctx* ctx) {
  return {{ctx->d_start, size_t(ctx->NN + 1) * sizeof(int)},
          {ctx->d_sorted, size_t(ctx->np + 1) * sizeof(unsigned long long)},
          {ctx->d_order, size_t(ctx->NN) * sizeof(int)},
          {ctx->d_nodepos, size_t(ctx->NN + 1) * sizeof(int)}};
}
static size_t bins_bytes(const jump_ctx* ctx) {
  size_t b = 0;
  for (auto& a : bins_segments(ctx)) b += seg_offset(a.second);
  return b;
}
static int bins_copy(jump_ctx* ctx, void* buf, bool save) {
  return copy_segments(ctx, static_cast<char*>(buf), bins_segments(ctx), save);
}

// The grid a step's G2P used: nodal mass and momenta (2D: m, p0, p, 5
// arrays of NN; 3D: m, p0, p and the re-projected p2, 10 arrays).
static int grid_arrays(const jump_ctx* ctx) { return ctx->D == 2 ? 5 : 10; }
static std::vector<std::pair<void*, size_t>> grid_segments(const jump_ctx* ctx) {
  const size_t nn = size_t(ctx->NN);
  double* arrs[10] = {ctx->g.m, ctx->g.p0[0], ctx->g.p0[1], ctx->g.p[0], ctx->g.p[1]};
  if (ctx->D == 3) {
    double* a3[10] = {ctx->g.m, ctx->g.p0[0], ctx->g.p0[1], ctx->g.p0[2], ctx->g.p[0],
                      ctx->g.p[1], ctx->g.p[2], ctx->g.p2[0], ctx->g.p2[1], ctx->g.p2[2]};
    for (int a = 0; a < 10; a++) arrs[a] = a3[a];
  }
  std::vector<std::pair<void*, size_t>> segs;
  for (int a = 0; a < grid_arrays(ctx); a++) segs.push_back({arrs[a], nn * sizeof(double)});
  return segs;
}
static int grid_copy(jump_ctx* ctx, double* buf, bool save) {
  return copy_segments(ctx, reinterpret_cast<char*>(buf), grid_segments(ctx), save);
}
static size_t grid_bytes(const jump_ctx* ctx) {
  return size_t(grid_arrays(ctx)) * seg_offset(size_t(ctx->NN) * sizeof(double));
}

// A recorded step as the backward sweep's kernels read it: the particle
// state before the step, the grid and the bins the step used.  The arrays a
// snapshot, grid or bin record holds point into the record (the layouts of
// snap_segments, grid_segments and bins_segments) instead of being copied
// back (about 0.1 ms per step on the 1.8M-particle two-rings problem); the
// rest (V0, sid, ...) are the context's own.  The adjoint kernels only read
// them.
struct AdjStep {
  Particles P;
  GridArrays g;
  const int* start;
  const unsigned long long* sorted;
  const int* order;    // the step's nodes, those with a particle in their window first
  const int* nodepos;  // nodepos[NN]: the number of those
};
template <class T>
static void seg_redirect(const std::vector<std::pair<void*, size_t>>& segs, void* buf, T*& p) {
  size_t off = 0;
  for (auto& sg : segs) {
    if (p && sg.first == static_cast<const void*>(p)) {
      p = reinterpret_cast<T*>(static_cast<char*>(buf) + off);
      return;
    }
    off += seg_offset(sg.second);
  }
}
// The particle arrays with those a snapshot holds pointing into it.
static Particles snap_view(const jump_ctx* ctx, double* snap) {
  Particles P = ctx->P;
  const auto ss = snap_segments(ctx);
  for (auto& q : P.pr) seg_redirect(ss, snap, q);
  for (auto& q : P.F) seg_redirect(ss, snap, q);
  for (auto& q : P.eps) seg_redirect(ss, snap, q);
  seg_redirect(ss, snap, P.alpha);
  for (auto& q : P.ep) seg_redirect(ss, snap, q);
  for (int f = 0; f < (ctx->D == 3 ? PairLayout<3>::NF : PairLayout<2>::NF); f++)
    *gather_field(P, ctx->D, f) = pair_field(P, f);
  return P;
}
static AdjStep adj_step_view(const jump_ctx* ctx, double* snap, double* grid, void* bins) {
  AdjStep st{snap_view(ctx, snap), ctx->g, ctx->d_start, ctx->d_sorted, ctx->d_order, ctx->d_nodepos};
  const auto gs = grid_segments(ctx);
  seg_redirect(gs, grid, st.g.m);
  for (int k = 0; k < 3; k++) {
    seg_redirect(gs, grid, st.g.p0[k]);
    seg_redirect(gs, grid, st.g.p[k]);
    seg_redirect(gs, grid, st.g.p2[k]);
  }
  const auto bs = bins_segments(ctx);
  seg_redirect(bs, bins, st.start);
  seg_redirect(bs, bins, st.sorted);
  seg_redirect(bs, bins, st.order);
  seg_redirect(bs, bins, st.nodepos);
  return st;
}

// One step from the state in `in` to the state in `out` (null `in`: the
// live arrays, read and written in place unless `out` is given), for the
// steps the reverse mode records (2D USL, 3D MUSL): it steps from one
// record into the next without copying.
static int step_between(jump_ctx* ctx, const Particles* in, const Particles* out) {
  const Particles live = ctx->P;
  if (in) ctx->P = *in;
  ctx->Pout = out;
  const int r = run_steps(ctx, 1);
  ctx->P = live;
  ctx->Pout = nullptr;
  return r;
}

// Lanes per node of the backward sweep's node gathers (k_adj_nodes,
// k_adj3_nodes1/2): the particles per cell and direction of the run's
// particles, (cell volume / mean initial particle volume)^(1/D), as 1, 2 or
// 4.  It is fixed when the run is recorded, so every backward sweep of the
// run sums in the same order whatever the checkpoint interval.  Measured on
// the 2D two-rings adjoint (4 per cell and direction) and the Taylor bar (2).
static int adjoint_lanes(const jump_ctx* ctx) {
  double v = 0.0;
  for (double x : ctx->hV0) v += x;
  if (ctx->hV0.empty() || !(v > 0.0)) return 1;
  v /= double(ctx->hV0.size());
  const double cell = ctx->G.dx * ctx->G.dy * (ctx->D == 3 ? ctx->G.dz : 1.0);
  const double r = std::pow(cell / v, 1.0 / double(ctx->D));
  return r >= 3.5 ? 4 : (r >= 1.5 ? 2 : 1);
}

template <int B, int S> static void adj_step_back3_s(jump_ctx* ctx, const AdjStep& st) {
  const int np = int(ctx->np), nn = ctx->NN;
  const int64_t nl = int64_t(nn) * S;
  const double dt = ctx->prm.dt, alpha = ctx->prm.alpha;
  auto& A = ctx->adj;
  LAUNCH(k_adj3_nodevel, nn, nn, st.g, alpha, A.nv3);
  LAUNCH((k_adj3_g2p2<B>), np, st.P, ctx->G, ctx->d_solids, st.g, dt, A.ap3, A.src3, A.nv3);
  LAUNCH((k_adj3_nodes2<B, S>), nl, st.P, ctx->G, ctx->d_solids, st.start, st.sorted, st.g, dt, A.src3, A.an3,
         st.order, st.nodepos);
  LAUNCH((k_adj3_g2p1<B>), np, st.P, ctx->G, ctx->d_solids, st.g, alpha, A.an3, A.ap3, A.src3, A.nv3);
  LAUNCH((k_adj3_nodes1<B, S>), nl, st.P, ctx->G, ctx->d_solids, st.start, st.sorted, st.g, alpha, A.src3, A.an3,
         st.order, st.nodepos);
  LAUNCH(k_adj3_grid, nn, nn, ctx->d_fixed, dt, A.an3);
  LAUNCH((k_adj3_p2g<B>), np, st.P, ctx->G, ctx->d_solids, A.an3, A.ap3);
}
template <int B> static void adj_step_back3(jump_ctx* ctx, const AdjStep& st) {
  switch (ctx->adj.lanes) {
    case 4: adj_step_back3_s<B, 4>(ctx, st); break;
    case 2: adj_step_back3_s<B, 2>(ctx, st); break;
    default: adj_step_back3_s<B, 1>(ctx, st); break;
  }
}

template <int B, int S> static void adj_step_back_s(jump_ctx* ctx, const AdjStep& st) {
  const int np = int(ctx->np), nn = ctx->NN;
  const double dt = ctx->prm.dt;
  AdjNodes& an = ctx->adj.an;
  LAUNCH(k_adj_nodevel, nn, nn, st.g, ctx->prm.tolerance, ctx->adj.nv);
  LAUNCH((k_adj_g2p<B>), np, st.P, ctx->G, ctx->d_solids, st.g, ctx->prm.tolerance, dt, ctx->adj.ap, ctx->adj.src,
         ctx->adj.nv);
  LAUNCH((k_adj_nodes<B, S>), int64_t(nn) * S, st.P, ctx->G, ctx->d_solids, st.start, st.sorted, st.g,
         ctx->prm.tolerance, dt, ctx->adj.src, an, st.order, st.nodepos);
  LAUNCH(k_adj_grid, nn, nn, ctx->d_fixed, dt, an);
  LAUNCH((k_adj_p2g<B>), np, st.P, ctx->G, ctx->d_solids, an, ctx->adj.ap);
}
template <int B> static void adj_step_back(jump_ctx* ctx, const AdjStep& st) {
  switch (ctx->adj.lanes) {
    case 4: adj_step_back_s<B, 4>(ctx, st); break;
    case 2: adj_step_back_s<B, 2>(ctx, st); break;
    default: adj_step_back_s<B, 1>(ctx, st); break;
  }
}

// Device bytes a recorded run of nsteps steps at interval K holds: nck
// snapshots, K grids and bins, K - 1 per-step states (a segment's first
// state is its snapshot), the final state.
static size_t adjoint_record_bytes(const jump_ctx* ctx, int64_t nsteps, int64_t K) {
  const size_t snap = snap_bytes(ctx);
  const size_t kk = size_t(std::min(K, nsteps));
  const size_t nck = size_t((nsteps + K - 1) / K);
  return (nck + kk) * snap + kk * (grid_bytes(ctx) + bins_bytes(ctx));
}

// The largest interval whose record fits in 3/4 of the device memory that is
// free once this context's previous record is released.  The adjoint arrays
// of the backward sweep (2D 25, 3D 55 doubles per particle) are set aside first.
static int adjoint_auto_interval(jump_ctx* ctx, int64_t nsteps, int64_t* K) {
  size_t freeb = 0, total = 0;
  HIP_TRY(hipMemGetInfo(&freeb, &total));
  const auto& A = ctx->adj;
  if (A.snap_np == ctx->np) {  // the buffers held now are freed or reused
    const size_t states = A.ckpt.size() + (A.segbuf.empty() ? 0 : A.segbuf.size() - 1) + (A.final_state ? 1 : 0);
    freeb += states * snap_bytes(ctx) + A.gridbuf.size() * grid_bytes(ctx) + A.binbuf.size() * bins_bytes(ctx);
  }
  // the adjoint arrays of the backward sweep: 2D 25 doubles per particle and
  // 6 per node; 3D 55 and 10
  const size_t reserve = (ctx->D == 2 ? 25 : 55) * size_t(ctx->np) * sizeof(double) +
                         (ctx->D == 2 ? 6 : 10) * size_t(ctx->NN) * sizeof(double);
  const size_t budget = freeb > reserve ? (freeb - reserve) / 4 * 3 : 0;
  int64_t k = std::min<int64_t>(nsteps, INT32_MAX);
  while (k > 1 && adjoint_record_bytes(ctx, nsteps, k) > budget) k = k * 7 / 8;
  *K = std::max<int64_t>(k, 1);
  return JUMP_OK;
}

int jump_adjoint_forward(jump_ctx* ctx, int64_t nsteps, int32_t K_arg, double* t_inout) {
  if (!ctx) return JUMP_ERR_ARG;
  if (!t_inout || nsteps <= 0 || K_arg < 0) return set_err(ctx, JUMP_ERR_ARG, "bad adjoint run arguments");
  int r = adjoint_supported(ctx);
  if (r) return r;
  HIP_TRY(hipSetDevice(ctx->device));
  r = ensure_device(ctx);
  if (r) return r;
  auto& A = ctx->adj;
  int64_t K = K_arg;
  if (K == 0 && (r = adjoint_auto_interval(ctx, nsteps, &K))) return r;
  const size_t sd = snap_bytes(ctx) / sizeof(double);
  const int64_t nck = (nsteps + K - 1) / K;
  const size_t nseg = size_t(std::min<int64_t>(K, nsteps));
  // buffers are kept from run to run while they are large enough
  if (A.ckpt.size() < size_t(nck) || A.segbuf.size() < nseg || A.snap_np != ctx->np) {
    free_list(A.allocs);
    A.ckpt.clear();
    A.segbuf.clear();
    A.gridbuf.clear();
    A.final_state = nullptr;
    auto alloc = [&](double** p, size_t count) { return dev_alloc(ctx, A.allocs, p, count); };
    A.ckpt.resize(size_t(nck));
    for (auto& c : A.ckpt)
      if ((r = alloc(&c, sd))) return r;
    A.segbuf.assign(nseg, nullptr);
    for (size_t q = 1; q < nseg; q++)
      if ((r = alloc(&A.segbuf[q], sd))) return r;
    A.gridbuf.resize(nseg);
    for (auto& c : A.gridbuf)
      if ((r = alloc(&c, grid_bytes(ctx) / sizeof(double)))) return r;
    A.binbuf.resize(nseg);
    for (auto& c : A.binbuf) {
      char* q = nullptr;
      if ((r = dev_alloc(ctx, A.allocs, &q, bins_bytes(ctx)))) return r;
      c = q;
    }
    if ((r = alloc(&A.final_state, sd))) return r;
    A.snap_np = ctx->np;
  }
  A.ready = false;
  A.dfy.clear();
  A.dk1.clear();
  A.lanes = adjoint_lanes(ctx);
  unsigned fl = 0;
  const int64_t last0 = (nck - 1) * K;  // the last segment's states and grids are kept as they go by
  const int64_t nlast = nsteps - last0;
  for (int64_t s = 0; s < nsteps; s++) {
    if (s % K == 0 && (r = snap_copy(ctx, A.ckpt[size_t(s / K)], true))) return r;
    if (s >= last0 && nlast > 1) {
      // the step reads its state where it is kept and writes the next one
      // into its record (the last into the live arrays)
      const size_t q = size_t(s - last0);
      const Particles in = q == 0 ? ctx->P : snap_view(ctx, A.segbuf[q]);
      const Particles out = int64_t(q) + 1 < nlast ? snap_view(ctx, A.segbuf[q + 1]) : ctx->P;
      r = step_between(ctx, q == 0 ? nullptr : &in, &out);
    } else {  // before the last segment, or a last segment of one step: in place
      r = run_steps(ctx, 1);
    }
    if (r) return r;
    if (s >= last0 && (r = grid_copy(ctx, A.gridbuf[size_t(s - last0)], true))) return r;
    if (s >= last0 && (r = bins_copy(ctx, A.binbuf[size_t(s - last0)], true))) return r;
  }
  A.cached_seg = nck - 1;
  if ((r = snap_copy(ctx, A.final_state, true))) return r;
  HIP_TRY(hipMemcpyAsync(&fl, ctx->d_flags, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemsetAsync(ctx->d_flags, 0, sizeof(unsigned), ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int64_t s = 0; s < nsteps; s++) *t_inout += ctx->prm.dt;
  ctx->grid_valid = true;
  A.nsteps = nsteps;
  A.K = int(K);
  A.ready = true;
  ctx->flags |= fl;
  if (fl) return set_err(ctx, JUMP_ERR_NUMERIC, "particle error flags raised during the recorded run");
  return JUMP_OK;
}

int jump_adjoint_interval(jump_ctx* ctx, int32_t* K) {
  if (!ctx || !K) return JUMP_ERR_ARG;
  if (!ctx->adj.ready) return set_err(ctx, JUMP_ERR_STATE, "no recorded run (call jump_adjoint_forward first)");
  *K = int32_t(ctx->adj.K);
  return JUMP_OK;
}

int jump_adjoint_backward(jump_ctx* ctx, const double* pos_bar, const double* vel_bar, double* dE, double* dnu,
                          double* dvel0) {
  if (!ctx) return JUMP_ERR_ARG;
  auto& A = ctx->adj;
  if (!A.ready) return set_err(ctx, JUMP_ERR_STATE, "no recorded run (call jump_adjoint_forward first)");
  if (!pos_bar) return set_err(ctx, JUMP_ERR_ARG, "pos_bar is required");
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t n = size_t(ctx->np), nn = size_t(ctx->NN);
  const int D = ctx->D;
  int r = JUMP_OK;
  if (A.adj_np != ctx->np) {
    free_list(A.adj_allocs);
    A.ap = AdjParticles{};
    A.an = AdjNodes{};
    A.src = AdjSources{};
    A.nv = NodeVel2{};
    A.ap3 = AdjParticles3{};
    A.an3 = AdjNodes3{};
    A.src3 = AdjSources3{};
    A.nv3 = NodeVel3{};
    A.adj_np = ctx->np;
    auto alloc = [&](double** p, size_t count) { return dev_alloc(ctx, A.adj_allocs, p, count); };
    if (D == 2) {
      AdjParticles& ap = A.ap;
      AdjNodes& an = A.an;
      for (int k = 0; k < 2 && !r; k++) r = alloc(&ap.xb[k], n);
      for (int k = 0; k < 2 && !r; k++) r = alloc(&ap.vb[k], n);
      for (int k = 0; k < 4 && !r; k++) r = alloc(&ap.Fb[k], n);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&ap.eb[k], n);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&ap.sb[k], n);
      if (!r) r = alloc(&ap.Vb, n);
      if (!r) r = alloc(&ap.lam, n);
      if (!r) r = alloc(&ap.mu, n);
      AdjSources& src = A.src;
      for (double2** v : {&src.xb, &src.vb, &src.Lb[0], &src.Lb[1]})
        if (!r) r = dev_alloc(ctx, A.adj_allocs, v, n);
      for (double2** v : {&A.nv.uL, &A.nv.u, &A.nv.a})
        if (!r) r = dev_alloc(ctx, A.adj_allocs, v, nn);
      if (!r) r = alloc(&an.mb, nn);
      for (int k = 0; k < 2 && !r; k++) r = alloc(&an.Pb[k], nn);
      for (int k = 0; k < 2 && !r; k++) r = alloc(&an.P0b[k], nn);
    } else {
      AdjParticles3& ap = A.ap3;
      for (int k = 0; k < 3 && !r; k++) r = alloc(&ap.xb[k], n);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&ap.vb[k], n);
      for (int k = 0; k < 9 && !r; k++) r = alloc(&ap.Fb[k], n);
      for (int k = 0; k < 6 && !r; k++) r = alloc(&ap.eb[k], n);
      for (int k = 0; k < 6 && !r; k++) r = alloc(&ap.sb[k], n);
      for (int k = 0; k < 6 && !r; k++) r = alloc(&ap.epb[k], n);
      for (double** v : {&ap.Vb, &ap.ab, &ap.lam, &ap.mu, &ap.fy, &ap.k1})
        if (!r) r = alloc(v, n);
      AdjSources3& src = A.src3;
      for (int k = 0; k < 6 && !r; k++) r = dev_alloc(ctx, A.adj_allocs, &src.xL[k], n);
      for (int k = 0; k < 2 && !r; k++) r = dev_alloc(ctx, A.adj_allocs, &src.vb[k], n);
      NodeVel3& nv = A.nv3;
      if (!r) r = alloc(&nv.mii, nn);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&nv.q[k], nn);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&nv.qm[k], nn);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&nv.v[k], nn);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&nv.v2[k], nn);
      AdjNodes3& an = A.an3;
      if (!r) r = alloc(&an.mb, nn);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&an.Pb[k], nn);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&an.P0b[k], nn);
      for (int k = 0; k < 3 && !r; k++) r = alloc(&an.P2b[k], nn);
    }
    if (r) return r;
  }
  // adjoints of the final state: positions and velocities as given, the rest 0
  {
    std::vector<double> h(n);
    for (int k = 0; k < D; k++) {
      double* xbk = D == 2 ? A.ap.xb[k] : A.ap3.xb[k];
      double* vbk = D == 2 ? A.ap.vb[k] : A.ap3.vb[k];
      for (size_t i = 0; i < n; i++) h[i] = pos_bar[D * i + k];
      HIP_TRY(hipMemcpyAsync(xbk, h.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      for (size_t i = 0; i < n; i++) h[i] = vel_bar ? vel_bar[D * i + k] : 0.0;
      HIP_TRY(hipMemcpyAsync(vbk, h.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    std::vector<double*> zs;
    if (D == 2) {
      const AdjParticles& ap = A.ap;
      zs = {ap.Fb[0], ap.Fb[1], ap.Fb[2], ap.Fb[3], ap.eb[0], ap.eb[1], ap.eb[2],
            ap.sb[0], ap.sb[1], ap.sb[2], ap.Vb, ap.lam, ap.mu};
    } else {
      const AdjParticles3& ap = A.ap3;
      for (int k = 0; k < 9; k++) zs.push_back(ap.Fb[k]);
      for (int k = 0; k < 6; k++) { zs.push_back(ap.eb[k]); zs.push_back(ap.sb[k]); zs.push_back(ap.epb[k]); }
      for (double* z : {ap.Vb, ap.ab, ap.lam, ap.mu, ap.fy, ap.k1}) zs.push_back(z);
    }
    for (double* z : zs) LAUNCH(k_fill_f64, n, z, 0.0, int(n));
  }
  const int64_t nseg = (A.nsteps + A.K - 1) / A.K;
  for (int64_t seg = nseg - 1; seg >= 0 && !r; seg--) {
    const int64_t s0 = seg * A.K, s1 = std::min<int64_t>(s0 + A.K, A.nsteps);
    // the state step s0 + q read: the segment's checkpoint for its first step
    auto seg_state = [&](size_t q) { return q == 0 ? A.ckpt[size_t(seg)] : A.segbuf[q]; };
    if (seg != A.cached_seg) {
      // rebuild the segment: states and the grids and bins their steps used,
      // from record to record: each step reads its state from the snapshot
      // or the record the previous step wrote, and writes the next record
      // (the segment's last state into the live arrays)
      ctx->runs_valid = false;  // the steps start from the snapshot: merge at every node
      const int64_t n = s1 - s0;
      for (int64_t q = 0; q < n && !r; q++) {
        const Particles in = snap_view(ctx, seg_state(size_t(q)));
        const Particles out = q + 1 < n ? snap_view(ctx, A.segbuf[size_t(q + 1)]) : ctx->P;
        if ((r = step_between(ctx, &in, &out))) break;
        if ((r = grid_copy(ctx, A.gridbuf[size_t(q)], true))) break;
        r = bins_copy(ctx, A.binbuf[size_t(q)], true);
      }
      A.cached_seg = seg;
    }
    for (int64_t j = s1 - 1; j >= s0 && !r; j--) {
      const size_t q = size_t(j - s0);
      const AdjStep st = adj_step_view(ctx, seg_state(q), A.gridbuf[q], A.binbuf[q]);
      if (D == 2) {
        if (ctx->prm.basis == JUMP_BASIS_LINEAR) adj_step_back<JUMP_BASIS_LINEAR>(ctx, st);
        else adj_step_back<JUMP_BASIS_GIMP>(ctx, st);
      } else {
        if (ctx->prm.basis == JUMP_BASIS_LINEAR) adj_step_back3<JUMP_BASIS_LINEAR>(ctx, st);
        else adj_step_back3<JUMP_BASIS_GIMP>(ctx, st);
      }
    }
  }
  if (r) return r;
  HIP_TRY(hipGetLastError());
  // leave the recorded final state in place
  if ((r = snap_copy(ctx, A.final_state, false))) return r;
  HIP_TRY(hipMemsetAsync(ctx->d_flags, 0, sizeof(unsigned), ctx->stream));
  // per-solid sums on the device: lambda, mu, (fy, k1,) v0
  const size_t ns = ctx->solids.size();
  AdjSumArrays arr{};
  int nsum = 0;
  if (D == 2) {
    for (const double* a : {(const double*)A.ap.lam, (const double*)A.ap.mu, (const double*)A.ap.vb[0],
                            (const double*)A.ap.vb[1]})
      arr.a[nsum++] = a;
  } else {
    for (const double* a : {(const double*)A.ap3.lam, (const double*)A.ap3.mu, (const double*)A.ap3.fy,
                            (const double*)A.ap3.k1, (const double*)A.ap3.vb[0], (const double*)A.ap3.vb[1],
                            (const double*)A.ap3.vb[2]})
      arr.a[nsum++] = a;
  }
  std::vector<long long> range(2 * ns);
  long long most = 1;
  for (size_t si = 0; si < ns; si++) {
    range[2 * si] = ctx->solids[si].offset;
    range[2 * si + 1] = ctx->solids[si].offset + ctx->solids[si].n;
    most = std::max<long long>(most, ctx->solids[si].n);
  }
  const int nchunk = int((most + kSumChunk - 1) / kSumChunk);
  long long* d_range = nullptr;
  double *d_sums = nullptr, *d_part = nullptr;
  std::vector<void*> tmp;
  if ((r = dev_alloc(ctx, tmp, &d_range, range.size())) || (r = dev_alloc(ctx, tmp, &d_sums, size_t(nsum) * ns)) ||
      (r = dev_alloc(ctx, tmp, &d_part, size_t(nsum) * ns * size_t(nchunk)))) {
    free_list(tmp);
    return r;
  }
  std::vector<double> sums(size_t(nsum) * ns);
  hipError_t he = hipMemcpyAsync(d_range, range.data(), range.size() * sizeof(long long), hipMemcpyHostToDevice,
                                 ctx->stream);
  if (he == hipSuccess) {
    for (size_t s0 = 0; s0 < ns; s0 += 65535)
      hipLaunchKernelGGL(k_adj_sums_part, dim3(unsigned(nchunk), unsigned(std::min<size_t>(ns - s0, 65535))),
                         dim3(kBlock), 0, ctx->stream, d_range, arr, nsum, int(s0), d_part);
    hipLaunchKernelGGL(k_adj_sums_final, dim3(unsigned(ns)), dim3(kBlock), 0, ctx->stream, d_part, nchunk, nsum,
                       d_sums);
    he = hipMemcpyAsync(sums.data(), d_sums, sums.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  }
  if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
  free_list(tmp);
  if (he != hipSuccess) return set_err(ctx, JUMP_ERR_DEVICE, "adjoint sums");
  A.dfy.assign(ns, 0.0);
  A.dk1.assign(ns, 0.0);
  for (size_t si = 0; si < ns; si++) {
    const HostSolid& hs = ctx->solids[si];
    const double* sm = &sums[size_t(nsum) * si];
    const double lb = sm[0], mb = sm[1];
    // lambda = E nu / ((1+nu)(1-2nu)),  mu = E / (2(1+nu))
    const double E = hs.mat.E, nu = hs.mat.nu, den = (1.0 + nu) * (1.0 - 2.0 * nu);
    const double dl_dE = nu / den, dm_dE = 1.0 / (2.0 * (1.0 + nu));
    const double dl_dnu = E * (1.0 + 2.0 * nu * nu) / (den * den), dm_dnu = -E / (2.0 * (1.0 + nu) * (1.0 + nu));
    if (dE) dE[si] = lb * dl_dE + mb * dm_dE;
    if (dnu) dnu[si] = lb * dl_dnu + mb * dm_dnu;
    const int vo = D == 2 ? 2 : 4;
    if (D == 3) {
      A.dfy[si] = sm[2];
      A.dk1[si] = sm[3];
    }
    if (dvel0)
      for (int k = 0; k < D; k++) dvel0[D * si + size_t(k)] = sm[vo + k];
  }
  ctx->grid_valid = false;
  return JUMP_OK;
}

int jump_adjoint_plastic(jump_ctx* ctx, double* dfy, double* dk1) {
  if (!ctx) return JUMP_ERR_ARG;
  const auto& A = ctx->adj;
  if (!A.ready || A.dfy.size() != ctx->solids.size())
    return set_err(ctx, JUMP_ERR_STATE, "no backward sweep of the current recorded run");
  for (size_t si = 0; si < ctx->solids.size(); si++) {
    if (dfy) dfy[si] = A.dfy[si];
    if (dk1) dk1[si] = A.dk1[si];
  }
  return JUMP_OK;
}
