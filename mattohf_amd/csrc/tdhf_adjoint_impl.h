// tdhf_adjoint_impl.h -- host side of the forward/backward pair (included by
// tdhf.hip after the single-trajectory helpers).
//
// A forward call propagates ntraj trajectories of the context's cell (same
// NK, tables and ion potential; own initial orbitals and field each) for nt
// RK4 steps in one launch per stage (grid y = trajectory) and keeps what the
// backward sweep needs:
//   - when 4*nt stage states per trajectory fit the memory budget, every
//     stage state y1..y4 of every step (the forward kernels write them
//     straight into the tape);
//   - otherwise the state at every S-th step, S = ceil(sqrt(nt)) (or the
//     least-memory spacing ~sqrt(nt/4) if that does not fit), and the
//     backward call recomputes each segment once, writing its stage states
//     into a segment tape.
// The backward call starts from the cotangent of the final state (dE/du for
// the energy objective, 2 dt J dJ/du for the current-squared one) and runs
// the reverse RK4 stages (adjoint_kernel, two exchange transforms per pair),
// adding 2 dt J_n dJ_n/du at every step boundary for the current-squared
// objective and collecting dL/dA per step; A enters step it as A(it-1)
// (stage 1), Aave(it) (stages 2, 3) and A(it) (stage 4), with
// Aave(it) = (A(it) + A(it-1))/2.
//
// Tape layout: [step][stage][trajectory][state], so the stage-s states of one
// step form one batched state array.
#pragma once

struct AdjState {
  bool have_fwd = false;
  int nt = 0, ntraj = 1, objective = TDHF_OBJ_ENERGY;
  int seg = 0;             // steps per segment (== nt when the whole tape is kept)
  bool full = false;
  std::vector<double> A, Aave;   // ntraj * nt, trajectory-major
  std::vector<double> jt, nel;   // J and N after every step, ntraj * nt
  double *dA = nullptr, *dAave = nullptr, *dJ = nullptr;
  zd* st = nullptr;        // u0 | ksum | sa | sb, ntraj states each
  zd* tape = nullptr;      // 4 * seg stage blocks of ntraj states
  zd* ckpt = nullptr;      // ceil(nt/seg) blocks of ntraj states (checkpoint mode)
  zd* buf = nullptr;       // UA | UB | GA | GB, ntraj states each
  double* obs = nullptr;   // ntraj * (nt+1) observables slots
  double* res = nullptr;   // ntraj * (nt+1) * 7 reduced observables
  double* abar = nullptr;  // ntraj * seg * 4 * rows per-row field derivatives
  double* stsum = nullptr; // ntraj * nt * 4 stage field derivatives
  // reverse stages by pair tiles (NG = 128, NB == NBocc; adjtile_kernel):
  // per-row partials, density and s partials, the tile list, a zero row
  zd* xpart = nullptr;
  double* dpart = nullptr;
  int4* tiles = nullptr;
  double* zrow = nullptr;
  int ntiles = 0, nblk = 0;
  // captured step blocks, keyed by the buffers they use
  std::vector<std::pair<std::vector<const void*>, hipGraphExec_t>> graphs;
};

namespace {

// The forward record of each context.  Contexts may be used from different
// host threads (one thread per context), so the map itself is guarded; a
// context's record is only touched by its own thread (references into a
// std::map stay valid while other keys are added or removed).
std::map<const tdhf_ctx*, AdjState>& adj_states() {
  static std::map<const tdhf_ctx*, AdjState> m;
  return m;
}
std::mutex& adj_mutex() {
  static std::mutex m;
  return m;
}
AdjState& adj_state_of(const tdhf_ctx* c) {
  std::lock_guard<std::mutex> lk(adj_mutex());
  return adj_states()[c];
}
AdjState* adj_find(const tdhf_ctx* c) {
  std::lock_guard<std::mutex> lk(adj_mutex());
  auto it = adj_states().find(c);
  return it == adj_states().end() ? nullptr : &it->second;
}

void adj_free(AdjState& s) {
  dfree(s.dA);
  dfree(s.dAave);
  dfree(s.dJ);
  dfree(s.st);
  dfree(s.tape);
  dfree(s.ckpt);
  dfree(s.buf);
  dfree(s.obs);
  dfree(s.res);
  dfree(s.abar);
  dfree(s.stsum);
  dfree(s.xpart);
  dfree(s.dpart);
  dfree(s.tiles);
  dfree(s.zrow);
  for (auto& g : s.graphs) (void)hipGraphExecDestroy(g.second);
  s = AdjState();
}

// Steps per captured block of the taped forward and of the reverse sweep
// (even: the reverse sweep's buffer rotation has period 2).
constexpr int kAdjGraphSteps = 16;

// Replays the captured block for key, capturing it with enqueue() first.
template <class Enq>
int adj_graph_launch(tdhf_ctx* c, AdjState& s, const std::vector<const void*>& key, Enq enqueue) {
  hipGraphExec_t ex = nullptr;
  for (auto& g : s.graphs)
    if (g.first == key) ex = g.second;
  if (!ex) {
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != 0) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(c, TDHF_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
      return fail(c, TDHF_EDEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
    s.graphs.emplace_back(key, ex);
  }
  HIPCHK(c, hipGraphLaunch(ex, c->stream));
  return 0;
}

size_t adj_lds_for(int ng, int nw) {
  switch (ng) {
    case 16: return tdhf::adjoint_lds_bytes<16>(nw);
    case 32: return tdhf::adjoint_lds_bytes<32>(nw);
    case 64: return tdhf::adjoint_lds_bytes<64>(nw);
    case 128: return tdhf::adjoint_lds_bytes<128>(nw);
  }
  return 0;
}

// Waves per workgroup of the reverse stage's second launch after the pair
// tiles (no pair loop): two threads per grid point for the circulant products.
constexpr int kAdjTileEpiWaves = 4;
constexpr int kAdjTileWaves = 8;  // adjtile_kernel<8>: 32-row tiles

// One reverse stage.  With ts holding pair tiles (adj_tiles_setup) and a
// stage mode, the exchange part comes from adjtile_kernel and
// adjoint_kernel<128, true> finishes the stage; otherwise adjoint_kernel
// evaluates every ordered pair itself.
int launch_adj(tdhf_ctx* c, const tdhf::AdjArgs& a, int ntraj, const AdjState* ts = nullptr) {
  const int rows = c->nb * c->nk;
  if (ts && ts->tiles && a.mode >= tdhf::A_S1 && a.mode <= tdhf::A_S4) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
      if (c->ev_used + 2 > c->ev_pool.size())
        for (int i = 0; i < 64; ++i) {
          hipEvent_t e;
          HIPCHK(c, hipEventCreate(&e));
          c->ev_pool.push_back(e);
        }
      e0 = c->ev_pool[c->ev_used++];
      e1 = c->ev_pool[c->ev_used++];
      HIPCHK(c, hipEventRecord(e0, c->stream));
    }
    tdhf::AdjTileArgs t{};
    t.nk = c->nk;
    t.nb = c->nb;
    t.vtl = c->d_vtl;
    t.tw = c->d_tw;
    t.zrow = ts->zrow;
    t.tiles = ts->tiles;
    t.y = a.y;
    t.gb = a.gb;
    t.step_base = a.step_base;
    t.jstep = a.jstep;
    t.tape = a.tape;
    t.tape_seg = a.tape_seg;
    t.blk = a.blk;
    t.mode = a.mode;
    t.tstride = a.tstride;
    t.xpart = ts->xpart;
    t.dpart = ts->dpart;
    t.xstride = (size_t)rows * ts->nblk * c->ng;
    t.dstride = (size_t)2 * ts->nblk * c->ng;
    t.nblk = ts->nblk;
    hipLaunchKernelGGL(tdhf::adjtile_kernel<kAdjTileWaves>, dim3(ts->ntiles),
                       dim3(64 * kAdjTileWaves), tdhf::adjtile_lds_bytes(kAdjTileWaves), c->stream,
                       t);
    HIPCHK(c, hipGetLastError());
    tdhf::AdjArgs b = a;
    b.xpart = t.xpart;
    b.dpart = t.dpart;
    b.nblk = t.nblk;
    b.xstride = t.xstride;
    b.dstride = t.dstride;
    hipLaunchKernelGGL((tdhf::adjoint_kernel<128, true>), dim3(rows, ntraj),
                       dim3(64 * kAdjTileEpiWaves),
                       tdhf::adjoint_lds_bytes<128>(kAdjTileEpiWaves, true), c->stream, b);
    HIPCHK(c, hipGetLastError());
    if (c->timing) HIPCHK(c, hipEventRecord(e1, c->stream));
    return 0;
  }
  // the context's launch shape also for large uniform batches: at NG=128
  // with 8 pulses, 2 or 4 waves per workgroup measured 15 % and 37 % slower
  // than 8 (unlike the forward kernel, see launch_apply)
  int nw = choose_nw(c->ng, c->nk, c->nb);
  // fewer waves where the tiles would exceed the 160 KiB of LDS (NG = 128)
  while (nw > (c->ng + 63) / 64 && adj_lds_for(c->ng, nw) > 160 * 1024) --nw;
  const size_t lds = adj_lds_for(c->ng, nw);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    if (c->ev_used + 2 > c->ev_pool.size())
      for (int i = 0; i < 64; ++i) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->ev_pool.push_back(e);
      }
    e0 = c->ev_pool[c->ev_used++];
    e1 = c->ev_pool[c->ev_used++];
    HIPCHK(c, hipEventRecord(e0, c->stream));
  }
  const dim3 grid(rows, ntraj), block(64 * nw);
  switch (c->ng) {
    case 16: hipLaunchKernelGGL(tdhf::adjoint_kernel<16>, grid, block, lds, c->stream, a); break;
    case 32: hipLaunchKernelGGL(tdhf::adjoint_kernel<32>, grid, block, lds, c->stream, a); break;
    case 64: hipLaunchKernelGGL(tdhf::adjoint_kernel<64>, grid, block, lds, c->stream, a); break;
    case 128: hipLaunchKernelGGL(tdhf::adjoint_kernel<128>, grid, block, lds, c->stream, a); break;
  }
  HIPCHK(c, hipGetLastError());
  if (c->timing) HIPCHK(c, hipEventRecord(e1, c->stream));
  return 0;
}

// Raises the LDS limit of the adjoint kernel (outside any stream capture).
// The limit is shared by every context of the process, so it is set to the
// largest launch any context makes (the most waves that fit 160 KiB).
int adj_prepare(tdhf_ctx* c) {
  int nw = 8;
  while (nw > (c->ng + 63) / 64 && adj_lds_for(c->ng, nw) > 160 * 1024) --nw;
  const size_t lds = adj_lds_for(c->ng, nw);
  const void* fn = nullptr;
  switch (c->ng) {
    case 16: fn = (const void*)tdhf::adjoint_kernel<16>; break;
    case 32: fn = (const void*)tdhf::adjoint_kernel<32>; break;
    case 64: fn = (const void*)tdhf::adjoint_kernel<64>; break;
    case 128: fn = (const void*)tdhf::adjoint_kernel<128>; break;
  }
  if (lds > 64 * 1024)
    HIPCHK(c, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return 0;
}

// Pair tiles for the reverse stages (NG = 128, every band occupied): the
// partial buffers and the tile list of the record's trajectories, set up once
// before the first reverse sweep.  Without device memory for them the
// reverse stages stay row-wise (the same results to rounding).
int adj_tiles_setup(tdhf_ctx* c, AdjState& s) {
  if (c->ng != 128 || c->nb != c->nbocc || s.tiles) return 0;
  const int rows = c->nb * c->nk;
  if ((long)rows >= (1L << 22)) return 0;  // the kernel's float k-point division
  constexpr int TR = tdhf::adjtile_rows(kAdjTileWaves);
  const int nblk = (rows + TR - 1) / TR;
  std::vector<int4> map;
  for (int pass = 0; pass < 2; ++pass)  // off-diagonal tiles (the costlier) first
    for (int t = 0; t < s.ntraj; ++t)
      for (int bi = 0; bi < nblk; ++bi)
        for (int bj = bi; bj < nblk; ++bj)
          if ((bi == bj) == (pass == 1)) map.push_back(make_int4(t, bi, bj, 0));
  const size_t xs = (size_t)s.ntraj * rows * nblk * c->ng;
  const size_t ds = (size_t)s.ntraj * 2 * nblk * c->ng;
  auto get = [](void** p, size_t bytes) { return hipMalloc(p, bytes + 16) == hipSuccess; };
  zd* xp = nullptr;
  double *dp = nullptr, *zr = nullptr;
  int4* tl = nullptr;
  if (!get((void**)&xp, xs * sizeof(zd)) || !get((void**)&dp, ds * sizeof(double)) ||
      !get((void**)&tl, map.size() * sizeof(int4)) || !get((void**)&zr, c->ng * sizeof(double))) {
    dfree(xp);
    dfree(dp);
    dfree(tl);
    dfree(zr);
    (void)hipGetLastError();  // the failed allocation is not an error of the call
    return 0;
  }
  s.xpart = xp;
  s.dpart = dp;
  s.tiles = tl;
  s.zrow = zr;
  s.ntiles = (int)map.size();
  s.nblk = nblk;
  HIPCHK(c, hipMemcpyAsync(tl, map.data(), map.size() * sizeof(int4), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemsetAsync(zr, 0, c->ng * sizeof(double), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipFuncSetAttribute((const void*)tdhf::adjtile_kernel<kAdjTileWaves>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tdhf::adjtile_lds_bytes(kAdjTileWaves)));
  HIPCHK(c, hipFuncSetAttribute((const void*)tdhf::adjoint_kernel<128, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tdhf::adjoint_lds_bytes<128>(kAdjTileEpiWaves, true)));
  return 0;
}

int launch_jgrad(tdhf_ctx* c, const AdjState& s, int step, const zd* y, zd* ua, zd* gb,
                 int init, int jstep = -1) {
  tdhf::JgradArgs j{};
  j.nb = c->nb;
  j.nbocc = c->nbocc;
  j.dt = c->dt;
  j.dv = c->dv;
  j.vcell = c->vcell;
  j.occ = c->occ;
  j.kx = c->d_kx;
  j.kc = c->d_kc;
  j.Aarr = s.dA;
  j.J = s.dJ;
  j.step = step;
  j.y = y;
  j.ua = ua;
  j.gb = gb;
  j.init = init;
  j.tstride = nstate(c);
  j.fstride = s.nt;
  if (jstep >= 0) {  // step and state from the device counter and the tape
    j.step_base = c->d_step;
    j.jstep = jstep;
    j.tape = s.tape;
    j.tape_seg = s.seg;
    j.blk = (size_t)s.ntraj * nstate(c);
  }
  const dim3 grid(c->nb * c->nk, s.ntraj);
  switch (c->ng) {
    case 16: hipLaunchKernelGGL(tdhf::jgrad_kernel<16>, grid, dim3(16), 0, c->stream, j); break;
    case 32: hipLaunchKernelGGL(tdhf::jgrad_kernel<32>, grid, dim3(32), 0, c->stream, j); break;
    case 64: hipLaunchKernelGGL(tdhf::jgrad_kernel<64>, grid, dim3(64), 0, c->stream, j); break;
    case 128: hipLaunchKernelGGL(tdhf::jgrad_kernel<128>, grid, dim3(128), 0, c->stream, j); break;
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

tdhf::AdjArgs adj_base(const tdhf_ctx* c, const AdjState& s) {
  tdhf::AdjArgs a{};
  a.nk = c->nk;
  a.nb = c->nb;
  a.nbocc = c->nbocc;
  a.bx = c->bx;
  a.dv = c->dv;
  a.occ = c->occ;
  a.dt = c->dt;
  a.kx = c->d_kx;
  a.vion = c->d_vion;
  a.vtl = c->d_vtl;
  a.kc = c->d_kc;
  a.tw = c->d_tw;
  a.Aarr = s.dA;
  a.Aavearr = s.dAave;
  a.tstride = nstate(c);
  a.fstride = s.nt;
  return a;
}

// Forward steps [s0, s0+n) of every trajectory on the batch state u0 (s.st),
// with the stage states written to tape (step it in slot it % seg) or, when
// tape == nullptr, untaped.  Observables of step it go to slot it.  Blocks of
// kAdjGraphSteps steps are replayed from a captured graph.
int adj_forward_steps(tdhf_ctx* c, AdjState& s, int s0, int n, zd* tape) {
  const int rows = c->nb * c->nk;
  const size_t ns = nstate(c), blk = (size_t)s.ntraj * ns;
  zd *u0 = s.st, *ks = s.st + blk, *sa = s.st + 2 * blk, *sb = s.st + 3 * blk;
  tdhf::ApplyArgs a = base_args(c);
  a.Aarr = s.dA;
  a.Aavearr = s.dAave;
  a.step_base = c->d_step;
  a.obs = s.obs;
  a.u0 = u0;
  a.ksum = ks;
  a.tape = tape;
  a.tape_seg = s.seg;
  a.nstate = blk;
  a.tstride = ns;
  a.fstride = s.nt;
  a.ostride = (size_t)(s.nt + 1) * c->obs_stride;
  auto enq = [&](int m) -> int {
    for (int j = 0; j < m; ++j) {
      a.jstep = j;
      a.mode = tdhf::M_S1; a.u = a.w = u0; a.unext = sa;
      CHKRET(launch_apply(c, a, rows, s.ntraj));
      a.mode = tdhf::M_S2; a.u = a.w = sa; a.unext = sb;
      CHKRET(launch_apply(c, a, rows, s.ntraj));
      a.mode = tdhf::M_S3; a.u = a.w = sb; a.unext = sa;
      CHKRET(launch_apply(c, a, rows, s.ntraj));
      a.mode = tdhf::M_S4; a.u = a.w = sa; a.unext = nullptr;
      CHKRET(launch_apply(c, a, rows, s.ntraj));
    }
    hipLaunchKernelGGL(tdhf::step_advance_kernel, dim3(1), dim3(1), 0, c->stream, c->d_step, m);
    HIPCHK(c, hipGetLastError());
    return 0;
  };
  hipLaunchKernelGGL(tdhf::step_set_kernel, dim3(1), dim3(1), 0, c->stream, c->d_step, s0);
  HIPCHK(c, hipGetLastError());
  int done = 0;
  if (!c->timing) {
    const std::vector<const void*> key = {(const void*)1, tape, s.st, s.obs};
    while (n - done >= kAdjGraphSteps) {
      CHKRET(adj_graph_launch(c, s, key, [&] { return enq(kAdjGraphSteps); }));
      done += kAdjGraphSteps;
    }
  }
  if (done < n) CHKRET(enq(n - done));
  return 0;
}

// Reverse sweep over steps [s0, s0+n) whose stage states are in s.tape
// (step it in slot it % seg).  ub/ua/ga/gb rotate as described in
// tdhf_adjoint.h.  The per-row field derivatives go to s.abar (same slots)
// and are summed per stage into s.stsum.  Blocks of kAdjGraphSteps steps are
// replayed from a captured graph (their buffer rotation returns to the start).
int adj_backward_steps(tdhf_ctx* c, AdjState& s, int s0, int n, zd** ub, zd** ua, zd** ga,
                       zd** gb) {
  const size_t ns = nstate(c), blk = (size_t)s.ntraj * ns;
  const int rows = c->nb * c->nk;
  tdhf::AdjArgs a = adj_base(c, s);
  a.abar_tstride = (size_t)s.seg * 4 * rows;
  a.step_base = c->d_step;
  a.tape = s.tape;
  a.tape_seg = s.seg;
  a.blk = blk;
  a.abar = s.abar;
  // m steps from the counter down; rotates the four buffers in place
  auto enq = [&](int m, zd** pub, zd** pua, zd** pga, zd** pgb) -> int {
    for (int j = 0; j < m; ++j) {
      a.jstep = j;
      a.ub = *pub;
      a.ua = *pua;
      for (int st = 4; st >= 1; --st) {
        a.mode = st;
        a.gb = *pga;
        a.gout = *pgb;
        CHKRET(launch_adj(c, a, s.ntraj, &s));
        std::swap(*pga, *pgb);
      }
      // current-squared objective: J after step it-1 is evaluated on y1 of step it
      if (s.objective == TDHF_OBJ_JSQ) CHKRET(launch_jgrad(c, s, 0, nullptr, *pua, *pga, 0, j));
      std::swap(*pub, *pua);  // the accumulated cotangent becomes ubar of the step before
    }
    hipLaunchKernelGGL(tdhf::step_advance_kernel, dim3(1), dim3(1), 0, c->stream, c->d_step, -m);
    HIPCHK(c, hipGetLastError());
    return 0;
  };
  hipLaunchKernelGGL(tdhf::step_set_kernel, dim3(1), dim3(1), 0, c->stream, c->d_step, s0 + n - 1);
  HIPCHK(c, hipGetLastError());
  const bool graphs = !c->timing && n >= kAdjGraphSteps;
  const int r = graphs ? n % kAdjGraphSteps : n;
  if (r > 0) CHKRET(enq(r, ub, ua, ga, gb));
  for (int done = r; done < n; done += kAdjGraphSteps) {
    const std::vector<const void*> key = {(const void*)2, *ub, *ua, *ga, *gb};
    zd *qb = *ub, *qa = *ua, *qga = *ga, *qgb = *gb;  // rotated copies during capture
    CHKRET(adj_graph_launch(c, s, key, [&] { return enq(kAdjGraphSteps, &qb, &qa, &qga, &qgb); }));
  }
  // stage sums of the segment, per trajectory: stsum[(t*nt + it)*4 + st]
  for (int t = 0; t < s.ntraj; ++t) {
    const long m = (long)n * 4;
    hipLaunchKernelGGL(tdhf::row_sum_kernel, dim3((unsigned)((m + 63) / 64)), dim3(64), 0,
                       c->stream, s.abar + (size_t)t * s.seg * 4 * rows + (size_t)(s0 % s.seg) * 4 * rows,
                       rows, m, s.stsum + ((size_t)t * s.nt + s0) * 4);
    HIPCHK(c, hipGetLastError());
  }
  return 0;
}

int adj_fail_free(AdjState& s, int rc) {
  adj_free(s);
  return rc;
}

}  // namespace

// One progress report (tdhf_set_progress): after the segment has finished.
int adj_report(tdhf_ctx* c, int phase, long done, long total) {
  if (!c->progress) return 0;
  CHKRET(sync(c));
  c->progress(c->progress_user, phase, done, total);
  return 0;
}

extern "C" int tdhf_set_progress(tdhf_ctx* c, tdhf_progress_fn fn, void* user) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  c->progress = fn;
  c->progress_user = user;
  return 0;
}

extern "C" int tdhf_fwd_batch(tdhf_ctx* c, int ntraj, double* u, int nt, const double* A,
                              int objective, const double* eref, double mem_budget, double* L,
                              double* etot, double* nele, double* jtot) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u || nt < 1 || !A || ntraj < 1) return fail(c, TDHF_EINVAL, "invalid forward arguments");
  if (objective != TDHF_OBJ_ENERGY && objective != TDHF_OBJ_JSQ)
    return fail(c, TDHF_EINVAL, "unknown objective %d", objective);
  if (c->nb != c->nbocc)
    return fail(c, TDHF_EINVAL, "the adjoint path propagates occupied bands only (NB must equal NBocc)");
  AdjState& s = adj_state_of(c);
  adj_free(s);
  s.nt = nt;
  s.ntraj = ntraj;
  s.objective = objective;
  s.A.assign(A, A + (size_t)ntraj * nt);
  s.Aave.resize((size_t)ntraj * nt);
  for (int t = 0; t < ntraj; ++t)
    for (int it = 0; it < nt; ++it) {
      const double* At = A + (size_t)t * nt;
      s.Aave[(size_t)t * nt + it] = 0.5 * (At[it] + (it > 0 ? At[it - 1] : 0.0));
    }
  const size_t ns = nstate(c), blk = (size_t)ntraj * ns, bb = blk * sizeof(zd);
  const int rows = c->nb * c->nk;
  double budget = mem_budget;
  if (!(budget > 0.0)) {
    size_t fr = 0, tot = 0;
    HIPCHK(c, hipMemGetInfo(&fr, &tot));
    budget = 0.25 * (double)fr;
  }
  if (4.0 * nt * (double)bb <= budget) {
    s.full = true;
    s.seg = nt;
  } else {
    // sqrt(nt) spacing; if that does not fit, the spacing of least memory,
    // ceil(nt/S) + 4 S states with S ~ sqrt(nt/4)
    auto need = [&](int S) { return ((double)((nt + S - 1) / S) + 4.0 * S) * (double)bb; };
    int S = (int)std::ceil(std::sqrt((double)nt));
    if (need(S) > budget) S = std::max(1, (int)std::ceil(std::sqrt(nt / 4.0)));
    if (need(S) > budget)
      return adj_fail_free(s, fail(c, TDHF_ENOMEM, "memory budget %.3g B too small for %d steps",
                                   budget, nt));
    s.full = false;
    s.seg = S;
  }
  int rc = 0;
  const long nsl = (long)nt + 1;
  if ((rc = dalloc(c, &s.dA, s.A.size())) || (rc = dalloc(c, &s.dAave, s.A.size())) ||
      (rc = dalloc(c, &s.st, 4 * blk)) || (rc = dalloc(c, &s.tape, (size_t)4 * s.seg * blk)) ||
      (rc = dalloc(c, &s.obs, (size_t)ntraj * nsl * c->obs_stride)) ||
      (rc = dalloc(c, &s.res, (size_t)ntraj * nsl * 7)))
    return adj_fail_free(s, rc);
  if (!s.full && (rc = dalloc(c, &s.ckpt, (size_t)((nt + s.seg - 1) / s.seg) * blk)))
    return adj_fail_free(s, rc);
  HIPCHK(c, hipMemcpyAsync(s.dA, s.A.data(), s.A.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(s.dAave, s.Aave.data(), s.Aave.size() * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.st, u, bb, hipMemcpyHostToDevice, c->stream));
  CHKRET(adj_prepare(c));
  if (s.full) {
    CHKRET(adj_forward_steps(c, s, 0, nt, s.tape));
  } else {
    const long nseg = (nt + s.seg - 1) / s.seg;
    for (int s0 = 0; s0 < nt; s0 += s.seg) {
      HIPCHK(c, hipMemcpyAsync(s.ckpt + (size_t)(s0 / s.seg) * blk, s.st, bb,
                               hipMemcpyDeviceToDevice, c->stream));
      CHKRET(adj_forward_steps(c, s, s0, std::min(s.seg, nt - s0), nullptr));
      CHKRET(adj_report(c, 0, s0 / s.seg + 1, nseg));
    }
  }
  // observables after the last step (slot nt) and the per-step reduction
  hipLaunchKernelGGL(tdhf::step_set_kernel, dim3(1), dim3(1), 0, c->stream, c->d_step, nt);
  HIPCHK(c, hipGetLastError());
  tdhf::ApplyArgs a = base_args(c);
  a.mode = tdhf::M_OBS;
  a.u = a.w = s.st;
  a.Aarr = s.dA;
  a.Aavearr = s.dAave;
  a.step_base = c->d_step;
  a.jstep = -1;
  a.obs = s.obs;
  a.tstride = ns;
  a.fstride = nt;
  a.ostride = (size_t)nsl * c->obs_stride;
  CHKRET(launch_apply(c, a, rows, ntraj));
  for (int t = 0; t < ntraj; ++t) {
    hipLaunchKernelGGL(tdhf::obs_reduce_kernel, dim3((unsigned)((nsl + 63) / 64)), dim3(64), 0,
                       c->stream, s.obs + (size_t)t * nsl * c->obs_stride, c->obs_stride, rows,
                       (int)nsl, c->dv, c->vcell, s.res + (size_t)t * nsl * 7);
    HIPCHK(c, hipGetLastError());
  }
  std::vector<double> res((size_t)ntraj * nsl * 7);
  HIPCHK(c, hipMemcpyAsync(res.data(), s.res, res.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(u, s.st, bb, hipMemcpyDeviceToHost, c->stream));
  CHKRET(sync(c));
  s.jt.resize((size_t)ntraj * nt);
  s.nel.resize((size_t)ntraj * nt);
  for (int t = 0; t < ntraj; ++t) {
    double lsum = 0.0;
    for (int it = 0; it < nt; ++it) {
      const double* r = &res[((size_t)t * nsl + it + 1) * 7];
      const size_t k = (size_t)t * nt + it;
      if (etot) etot[k] = r[0];
      if (nele) nele[k] = r[1];
      if (jtot) jtot[k] = r[2];
      s.jt[k] = r[2];
      s.nel[k] = r[1];
      lsum += c->dt * r[2] * r[2];
    }
    if (L) {
      if (objective == TDHF_OBJ_ENERGY)
        L[t] = res[((size_t)t * nsl + nt) * 7] - (eref ? eref[t] : 0.0);
      else
        L[t] = lsum;
    }
  }
  if (objective == TDHF_OBJ_JSQ) {
    CHKRET(dalloc(c, &s.dJ, s.jt.size()));
    HIPCHK(c, hipMemcpyAsync(s.dJ, s.jt.data(), s.jt.size() * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    CHKRET(sync(c));
  }
  // the observables slots are not needed by the backward sweep
  dfree(s.obs);
  s.obs = nullptr;
  s.have_fwd = true;
  return 0;
}

extern "C" int tdhf_bwd_batch(tdhf_ctx* c, double* dLdA, double* dLdu0) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  AdjState* ps = adj_find(c);
  if (!ps || !ps->have_fwd)
    return fail(c, TDHF_EINVAL, "the backward call needs a preceding forward call on this context");
  AdjState& s = *ps;
  if (!dLdA) return fail(c, TDHF_EINVAL, "dLdA is NULL");
  if (!s.full && !s.ckpt)
    return fail(c, TDHF_EINVAL, "the checkpointed forward state was consumed by an earlier backward call");
  const size_t ns = nstate(c), blk = (size_t)s.ntraj * ns;
  const int rows = c->nb * c->nk;
  const int nt = s.nt;
  if (!s.buf) CHKRET(dalloc(c, &s.buf, 4 * blk));
  if (!s.abar) CHKRET(dalloc(c, &s.abar, (size_t)s.ntraj * s.seg * 4 * rows));
  if (!s.stsum) CHKRET(dalloc(c, &s.stsum, (size_t)s.ntraj * nt * 4));
  CHKRET(adj_tiles_setup(c, s));
  zd *ua = s.buf, *ub = s.buf + blk, *ga = s.buf + 2 * blk, *gb = s.buf + 3 * blk;
  // cotangent of the final state (s.st holds it after the forward run)
  if (s.objective == TDHF_OBJ_ENERGY) {
    tdhf::AdjArgs e = adj_base(c, s);
    e.mode = tdhf::A_EGRAD;
    e.step = nt - 1;
    e.y = s.st;
    e.cout = ub;
    e.gout = ga;
    CHKRET(launch_adj(c, e, s.ntraj));
  } else {
    CHKRET(launch_jgrad(c, s, nt - 1, s.st, ub, ga, 1));
  }
  if (s.full) {
    CHKRET(adj_backward_steps(c, s, 0, nt, &ub, &ua, &ga, &gb));
  } else {
    const long nseg = (nt + s.seg - 1) / s.seg;
    for (int s0 = ((nt - 1) / s.seg) * s.seg; s0 >= 0; s0 -= s.seg) {
      const int n = std::min(s.seg, nt - s0);
      HIPCHK(c, hipMemcpyAsync(s.st, s.ckpt + (size_t)(s0 / s.seg) * blk, blk * sizeof(zd),
                               hipMemcpyDeviceToDevice, c->stream));
      CHKRET(adj_forward_steps(c, s, s0, n, s.tape));
      CHKRET(adj_backward_steps(c, s, s0, n, &ub, &ua, &ga, &gb));
      CHKRET(adj_report(c, 1, nseg - s0 / s.seg, nseg));
    }
    // s.st now holds the initial state: a second backward call would need the
    // final one, so the forward record is spent
    dfree(s.ckpt);
    s.ckpt = nullptr;
  }
  std::vector<double> sts((size_t)s.ntraj * nt * 4);
  HIPCHK(c, hipMemcpyAsync(sts.data(), s.stsum, sts.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  if (dLdu0)
    HIPCHK(c, hipMemcpyAsync(dLdu0, ub, blk * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
  CHKRET(sync(c));
  for (int t = 0; t < s.ntraj; ++t) {
    double* g = dLdA + (size_t)t * nt;
    const double* st4 = sts.data() + (size_t)t * nt * 4;
    for (int it = 0; it < nt; ++it) g[it] = 0.0;
    for (int it = 0; it < nt; ++it) {
      const double* st = st4 + (size_t)it * 4;
      if (it > 0) g[it - 1] += st[0] + 0.5 * (st[1] + st[2]);
      g[it] += st[3] + 0.5 * (st[1] + st[2]);
    }
    if (s.objective == TDHF_OBJ_ENERGY) {
      g[nt - 1] += -s.jt[(size_t)t * nt + nt - 1] * c->vcell;  // dEtot/dA at the final state
    } else {
      for (int it = 0; it < nt; ++it) {  // dJ/dA = -N/vcell after every step
        const size_t k = (size_t)t * nt + it;
        g[it] += 2.0 * c->dt * s.jt[k] * (-s.nel[k] / c->vcell);
      }
    }
  }
  return 0;
}

extern "C" int tdhf_fwd(tdhf_ctx* c, double* u, int nt, const double* A, double eref,
                        double mem_budget, double* L, double* etot, double* nele, double* jtot) {
  return tdhf_fwd_batch(c, 1, u, nt, A, TDHF_OBJ_ENERGY, &eref, mem_budget, L, etot, nele, jtot);
}

extern "C" int tdhf_bwd(tdhf_ctx* c, double* dLdA, double* dLdu0) {
  return tdhf_bwd_batch(c, dLdA, dLdu0);
}

extern "C" int tdhf_adjoint_info(tdhf_ctx* c, int* full_tape, int* segment) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  const AdjState* ps = adj_find(c);
  if (!ps || !ps->have_fwd) return fail(c, TDHF_EINVAL, "no forward run recorded");
  if (full_tape) *full_tape = ps->full ? 1 : 0;
  if (segment) *segment = ps->seg;
  return 0;
}

namespace {
void adj_release(const tdhf_ctx* c) {
  AdjState* ps = adj_find(c);
  if (!ps) return;
  adj_free(*ps);
  std::lock_guard<std::mutex> lk(adj_mutex());
  adj_states().erase(c);
}
}  // namespace
