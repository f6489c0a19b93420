// tdhf_batch_impl.h -- batched propagation of independent trajectories
// (included by tdhf.hip after the single-trajectory helpers).
//
// Every trajectory has its own k-mesh (NK may differ), ion potential, field
// A(t) and RK4 buffers; NG, NB, NBocc, acell and dt are shared.  One launch of
// apply_kernel per RK4 stage covers the rows (k,i) of all trajectories: the
// schedule (rowmap) lists rows longest-first within each XCD, and pins each
// trajectory to one XCD (workgroups are dispatched round-robin over the 8
// XCDs, so entry b runs on XCD b % 8) so that its orbitals stay in that XCD's
// L2.  The exchange is evaluated row-wise by that kernel or, for NG = 64 with
// every band occupied, by pair tiles (tdhf_pairtile.h), whose per-row partial
// sums pair_epilogue_kernel then adds before the rest of H and the stage
// update (two launches per stage).  The RK4 steps are captured as a hipGraph and
// replayed; the exchange evaluation and tile size, the row schedule and the
// launch shape (waves per workgroup) are picked by tdhf_batch_autotune.
#pragma once

#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <numeric>

struct tdhf_batch {
  tdhf_params p{};
  int ntraj = 0;
  std::vector<int> nk;
  std::string err = "no error";
  hipStream_t stream = nullptr;
  std::vector<tdhf_ctx*> tabs;  // one context (tables) per distinct NK
  std::vector<int> tab_of;
  std::vector<size_t> off;      // state offset (complex elements) per trajectory
  size_t total = 0;             // complex elements of all trajectories
  double* d_vion = nullptr;
  zd* d_state = nullptr;        // 4 * total: u0 | ksum | sa | sb
  tdhf::TrajDesc* d_td = nullptr;
  std::vector<tdhf::TrajDesc> h_td;
  int2* d_rowmap = nullptr;
  int ngrid = 0;
  double *d_A = nullptr, *d_Aave = nullptr;
  long A_cap = 0;
  double* d_obs = nullptr;
  long obs_cap = 0;
  std::vector<size_t> obs_off;  // doubles per trajectory per slot: stride_t
  double* d_res = nullptr;
  long res_cap = 0;
  int* d_step = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0;
  int graph_nw = 0;
  int graph_sched = 0;
  int nw = 4;
  int sched = 0;  // row schedule (build_schedule)
  // exchange: 0 = row-wise in the fused kernel; 1, 2, 3 = pair tiles of 64,
  // 32, 16 rows (pairtile_kernel<8, 4, 2>) + pair_epilogue_kernel
  int xmode = 0;
  int graph_xmode = 0;
  int xrows = 0;      // tile rows the partial buffers and the tile list are laid out for
  zd* d_xpart = nullptr;
  double* d_dpart = nullptr;
  int4* d_tiles = nullptr;
  int ntiles = 0;
  int2* d_erows = nullptr;  // (trajectory, row) of every row, for pair_epilogue_kernel
  int nerows = 0;
  size_t tiles_cap = 0;  // entries d_tiles holds
  std::vector<size_t> xoff, doff;  // per trajectory (complex / double elements)
  size_t zoff = 0;                 // NG zeros in d_dpart after the partials
  bool timing = false;
  long launches = 0;
  double total_ms = 0.0;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
};

namespace {

int bfail(tdhf_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  g_err = buf;
  return code;
}

#define BHIP(b, expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return bfail((b), TDHF_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                   __FILE__, __LINE__);                                                       \
  } while (0)

template <typename T>
int balloc(tdhf_batch* b, T** p, size_t n) {
  if (*p) return 0;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T) + 16);
  if (e != hipSuccess) {
    *p = nullptr;
    return bfail(b, TDHF_ENOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                 hipGetErrorString(e));
  }
  return 0;
}

size_t bstate(const tdhf_batch* b, int t) { return (size_t)b->p.ng * b->p.nb * b->nk[t]; }

int batch_valid_nw(const tdhf_batch* b, int nw) {
  return nw >= 1 && nw <= 8 && 64 * nw >= b->p.ng;
}

constexpr int kNumSchedules = 3;

// Row schedules (workgroup b of a launch is dispatched to XCD b % 8):
//  0: trajectories bin-packed over the 8 XCDs by cost (longest-processing-
//     time first), each XCD's queue longest trajectory first with its rows
//     contiguous, queues interleaved so that entry b is on XCD b % 8: a
//     trajectory's orbitals stay in one XCD's L2;
//  1: as 0, but each XCD's queue takes its trajectories' rows round-robin
//     (row r of every trajectory before row r+1), so the rows running at once
//     on an XCD belong to different trajectories;
//  2: no XCD pinning: all trajectories longest first, rows contiguous, in
//     dispatch order (a trajectory's rows spread over the XCDs).
void build_schedule(const tdhf_batch* b, std::vector<int2>& map, int sched) {
  constexpr int kXcd = 8;
  const int nb = b->p.nb;
  std::vector<int> order(b->ntraj);
  std::iota(order.begin(), order.end(), 0);
  auto cost = [&](int t) { return (double)b->nk[t] * b->nk[t] * nb * b->p.nbocc + b->nk[t] * nb; };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
  map.clear();
  if (sched == 2) {
    for (int t : order)
      for (int r = 0; r < nb * b->nk[t]; ++r) map.push_back(make_int2(t, r));
    return;
  }
  std::vector<double> load(kXcd, 0.0);
  std::vector<std::vector<int>> bins(kXcd);
  for (int t : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += cost(t);
    bins[x].push_back(t);  // already in decreasing cost order
  }
  std::vector<std::vector<int2>> q(kXcd);
  size_t longest = 0;
  for (int x = 0; x < kXcd; ++x) {
    if (sched == 1) {
      int rmax = 0;
      for (int t : bins[x]) rmax = std::max(rmax, nb * b->nk[t]);
      for (int r = 0; r < rmax; ++r)
        for (int t : bins[x])
          if (r < nb * b->nk[t]) q[x].push_back(make_int2(t, r));
    } else {
      for (int t : bins[x])
        for (int r = 0; r < nb * b->nk[t]; ++r) q[x].push_back(make_int2(t, r));
    }
    longest = std::max(longest, q[x].size());
  }
  for (size_t i = 0; i < longest; ++i)
    for (int x = 0; x < kXcd; ++x) map.push_back(i < q[x].size() ? q[x][i] : make_int2(-1, 0));
  while (!map.empty() && map.back().x < 0) map.pop_back();
}

int batch_build_tiles(tdhf_batch* b);

// Installs row schedule sched (device row map and, with pair tiles, the tile
// list; the step graph is recaptured).
int batch_set_schedule(tdhf_batch* b, int sched) {
  std::vector<int2> map;
  build_schedule(b, map, sched);
  if ((int)map.size() > b->ngrid) {
    dfree(b->d_rowmap);
    CHKRET(balloc(b, &b->d_rowmap, map.size()));
  }
  BHIP(b, hipMemcpyAsync(b->d_rowmap, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice,
                         b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  b->ngrid = (int)map.size();
  b->sched = sched;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  // the pair-tile list follows the row schedule's XCD pinning
  if (b->xmode) CHKRET(batch_build_tiles(b));
  return 0;
}

// Pair tiles: NG = 64, every band occupied, and fewer than 2^22 rows per
// trajectory (the kernel's float k-point division is exact below that).
bool batch_pairtile_ok(const tdhf_batch* b) {
  if (b->p.ng != 64 || b->p.nb != b->p.nbocc) return false;
  for (int k : b->nk)
    if ((long)k * b->p.nb >= (1L << 22)) return false;
  return true;
}

// Tile waves and rows of a pair-tile exchange mode (1, 2, 3: 8, 4, 2 waves).
int xmode_waves(int xmode) { return 16 >> xmode; }
int xmode_rows(int xmode) { return tdhf::tile_rows(xmode_waves(xmode)); }

// Pair-tile schedule: the tiles (bi <= bj) of every trajectory, bin-packed
// over the 8 XCDs by cost (an off-diagonal tile of W waves is 8 W passes, a
// diagonal one 4 W + 2) -- with pin, whole trajectories (a trajectory's tiles
// share one XCD's L2; row schedules 0 and 1), else tile by tile (row
// schedule 2: small batches use the whole chip) -- each XCD's tiles costliest
// first, queues interleaved so that entry b runs on XCD b % 8.
void build_tile_schedule(const tdhf_batch* b, int xmode, bool pin, std::vector<int4>& map) {
  constexpr int kXcd = 8;
  const int W = xmode_waves(xmode), TR = xmode_rows(xmode);
  const double coff = 8.0 * W, cdiag = tdhf::diag_passes(W);
  auto nblk = [&](int t) { return (b->p.nb * b->nk[t] + TR - 1) / TR; };
  auto cost = [&](int t) {
    const int n = nblk(t);
    return coff * n * (n - 1) / 2 + cdiag * n;
  };
  std::vector<int> order(b->ntraj);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
  std::vector<double> load(kXcd, 0.0);
  std::vector<std::vector<int4>> q(kXcd);
  auto lightest = [&] { return (int)(std::min_element(load.begin(), load.end()) - load.begin()); };
  if (pin) {
    for (int t : order) {
      const int x = lightest();
      load[x] += cost(t);
      const int n = nblk(t);
      for (int bi = 0; bi < n; ++bi)
        for (int bj = bi; bj < n; ++bj) q[x].push_back(make_int4(t, bi, bj, 0));
    }
  } else {
    // off-diagonal tiles first (the costlier), each to the lightest XCD
    for (int pass = 0; pass < 2; ++pass)
      for (int t : order) {
        const int n = nblk(t);
        for (int bi = 0; bi < n; ++bi)
          for (int bj = bi; bj < n; ++bj) {
            if ((bi == bj) != (pass == 1)) continue;
            const int x = lightest();
            load[x] += bi == bj ? cdiag : coff;
            q[x].push_back(make_int4(t, bi, bj, 0));
          }
      }
  }
  size_t longest = 0;
  for (auto& v : q) {
    std::stable_sort(v.begin(), v.end(),
                     [](const int4& x, const int4& y) { return (x.y != x.z) > (y.y != y.z); });
    longest = std::max(longest, v.size());
  }
  map.clear();
  for (size_t i = 0; i < longest; ++i)
    for (int x = 0; x < kXcd; ++x) map.push_back(i < q[x].size() ? q[x][i] : make_int4(-1, 0, 0, 0));
  while (!map.empty() && map.back().x < 0) map.pop_back();
}

int batch_update_desc(tdhf_batch* b);

// Switches the exchange between the row-wise kernel (0) and the pair tiles
// of 64, 32 or 16 rows (1, 2, 3): partial buffers and the tile schedule are
// set up when the tile size changes.
int batch_set_xmode(tdhf_batch* b, int xmode) {
  if (xmode && !batch_pairtile_ok(b))
    return bfail(b, TDHF_EINVAL, "pair tiles need NG = 64, NB == NBocc and NK*NB < 2^22");
  if (xmode && b->xrows != xmode_rows(xmode)) {
    const int TR = xmode_rows(xmode);
    BHIP(b, hipStreamSynchronize(b->stream));  // no launch in flight uses the old buffers
    dfree(b->d_xpart);
    dfree(b->d_dpart);
    dfree(b->d_tiles);
    b->d_xpart = nullptr;
    b->d_dpart = nullptr;
    b->d_tiles = nullptr;
    b->xrows = 0;
    b->xoff.resize(b->ntraj);
    b->doff.resize(b->ntraj);
    size_t xs = 0, ds = 0;
    for (int t = 0; t < b->ntraj; ++t) {
      const int rows = b->p.nb * b->nk[t];
      const int nblk = (rows + TR - 1) / TR;
      b->xoff[t] = xs;
      b->doff[t] = ds;
      xs += (size_t)rows * nblk * b->p.ng;
      ds += (size_t)nblk * b->p.ng;
    }
    CHKRET(balloc(b, &b->d_xpart, xs));
    // the density partials, then NG zeros (the table row of absent partners)
    CHKRET(balloc(b, &b->d_dpart, ds + b->p.ng));
    b->zoff = ds;
    BHIP(b, hipMemsetAsync(b->d_dpart + ds, 0, b->p.ng * sizeof(double), b->stream));
    b->tiles_cap = 0;
    BHIP(b, hipFuncSetAttribute((const void*)tdhf::pairtile_kernel<8>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tdhf::pairtile_lds_bytes(8)));
    BHIP(b, hipFuncSetAttribute((const void*)tdhf::pairtile_kernel<4>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tdhf::pairtile_lds_bytes(4)));
    BHIP(b, hipFuncSetAttribute((const void*)tdhf::pairtile_kernel<2>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tdhf::pairtile_lds_bytes(2)));
    BHIP(b, hipFuncSetAttribute((const void*)tdhf::pair_epilogue_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tdhf::pair_epilogue_lds_bytes(8)));
    b->xrows = TR;
  }
  if (xmode && !b->d_erows) {
    // the epilogue's rows, trajectory by trajectory (a wave's 8 rows mostly
    // share a trajectory's tables and density partials)
    std::vector<int2> er;
    for (int t = 0; t < b->ntraj; ++t)
      for (int r = 0; r < b->p.nb * b->nk[t]; ++r) er.push_back(make_int2(t, r));
    CHKRET(balloc(b, &b->d_erows, er.size()));
    BHIP(b, hipMemcpyAsync(b->d_erows, er.data(), er.size() * sizeof(int2), hipMemcpyHostToDevice,
                           b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    b->nerows = (int)er.size();
  }
  b->xmode = xmode;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  if (xmode) CHKRET(batch_build_tiles(b));
  return batch_update_desc(b);
}

// (Re)builds the pair-tile list for the current tile size and row schedule.
int batch_build_tiles(tdhf_batch* b) {
  std::vector<int4> map;
  build_tile_schedule(b, b->xmode, b->sched != 2, map);
  if (map.size() > b->tiles_cap) {
    BHIP(b, hipStreamSynchronize(b->stream));
    dfree(b->d_tiles);
    CHKRET(balloc(b, &b->d_tiles, map.size()));
    b->tiles_cap = map.size();
  }
  BHIP(b, hipMemcpyAsync(b->d_tiles, map.data(), map.size() * sizeof(int4), hipMemcpyHostToDevice,
                         b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  b->ntiles = (int)map.size();
  return 0;
}

int batch_update_desc(tdhf_batch* b) {
  for (int t = 0; t < b->ntraj; ++t) {
    tdhf::TrajDesc& d = b->h_td[t];
    tdhf_ctx* c = b->tabs[b->tab_of[t]];
    d.nk = b->nk[t];
    d.obs_stride = 3 * b->p.nb * b->nk[t] + 3;
    d.occ = 2.0 / (double)b->nk[t];
    d.kx = c->d_kx;
    d.vion = b->d_vion + (size_t)b->p.ng * t;
    d.vtl = c->d_vtl;
    d.kc = c->d_kc;
    d.u0 = b->d_state + b->off[t];
    d.ksum = b->d_state + b->total + b->off[t];
    d.sa = b->d_state + 2 * b->total + b->off[t];
    d.sb = b->d_state + 3 * b->total + b->off[t];
    d.A = b->d_A ? b->d_A + (size_t)b->A_cap * t : nullptr;
    d.Aave = b->d_Aave ? b->d_Aave + (size_t)b->A_cap * t : nullptr;
    d.obs = b->d_obs ? b->d_obs + b->obs_off[t] * (size_t)b->obs_cap : nullptr;
    const int rows = b->p.nb * b->nk[t];
    d.nblk = b->xmode ? (rows + b->xrows - 1) / b->xrows : 0;
    d.xpart = b->xmode ? b->d_xpart + b->xoff[t] : nullptr;
    d.dpart = b->xmode ? b->d_dpart + b->doff[t] : nullptr;
    d.vhg = c->d_vhg;
  }
  BHIP(b, hipMemcpyAsync(b->d_td, b->h_td.data(), b->h_td.size() * sizeof(tdhf::TrajDesc),
                         hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return 0;
}

int batch_launch(tdhf_batch* b, const tdhf::ApplyArgs& a, int nw) {
  // with timing on, one event pair brackets the stage (pair tiles + epilogue)
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (b->timing) {
    if (b->ev_used + 2 > b->ev.size())
      for (int i = 0; i < 64; ++i) {
        hipEvent_t e;
        BHIP(b, hipEventCreate(&e));
        b->ev.push_back(e);
      }
    e0 = b->ev[b->ev_used++];
    e1 = b->ev[b->ev_used++];
    BHIP(b, hipEventRecord(e0, b->stream));
  }
  if (b->xmode) {
    tdhf::PairTileArgs pa{};
    pa.td = b->d_td;
    pa.tiles = b->d_tiles;
    pa.tw = a.tw;
    pa.zrow = b->d_dpart + b->zoff;
    pa.nb = b->p.nb;
    pa.mode = a.mode;
    switch (xmode_waves(b->xmode)) {
      case 8:
        hipLaunchKernelGGL(tdhf::pairtile_kernel<8>, dim3(b->ntiles), dim3(512),
                           tdhf::pairtile_lds_bytes(8), b->stream, pa);
        break;
      case 4:
        hipLaunchKernelGGL(tdhf::pairtile_kernel<4>, dim3(b->ntiles), dim3(256),
                           tdhf::pairtile_lds_bytes(4), b->stream, pa);
        break;
      default:
        hipLaunchKernelGGL(tdhf::pairtile_kernel<2>, dim3(b->ntiles), dim3(128),
                           tdhf::pairtile_lds_bytes(2), b->stream, pa);
        break;
    }
    BHIP(b, hipGetLastError());
  }
  if (b->xmode) {  // pair tiles: NG = 64 (batch_pairtile_ok); 8 rows per wave
    const int rpb = 8 * nw;
    hipLaunchKernelGGL(tdhf::pair_epilogue_kernel, dim3((b->nerows + rpb - 1) / rpb),
                       dim3(64 * nw), tdhf::pair_epilogue_lds_bytes(nw), b->stream, a, b->d_erows,
                       b->nerows);
  } else {
    const size_t lds = lds_for(b->p.ng, nw);
    switch (b->p.ng) {
      case 16: launch_apply_t<16>(a, b->ngrid, 1, nw, lds, b->stream); break;
      case 32: launch_apply_t<32>(a, b->ngrid, 1, nw, lds, b->stream); break;
      case 64: launch_apply_t<64>(a, b->ngrid, 1, nw, lds, b->stream); break;
      case 128: launch_apply_t<128>(a, b->ngrid, 1, nw, lds, b->stream); break;
    }
  }
  BHIP(b, hipGetLastError());
  if (b->timing) BHIP(b, hipEventRecord(e1, b->stream));
  return 0;
}

int batch_harvest(tdhf_batch* b) {
  BHIP(b, hipStreamSynchronize(b->stream));
  if (!b->timing) return 0;
  for (size_t i = 0; i + 1 < b->ev_used; i += 2) {
    float ms = 0.f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev[i], b->ev[i + 1]));
    b->total_ms += ms;
    b->launches += 1;
  }
  b->ev_used = 0;
  return 0;
}

tdhf::ApplyArgs batch_args(const tdhf_batch* b) {
  tdhf::ApplyArgs a{};
  a.nb = b->p.nb;
  a.nbocc = b->p.nbocc;
  a.nbw = b->p.nb;
  a.bx = 2.0 * (4.0 * std::atan(1.0)) / b->p.acell;
  a.dv = std::fabs(b->p.acell) / (double)b->p.ng;
  a.dt = b->p.dt;
  a.tw = b->tabs[0]->d_tw;
  a.td = b->d_td;
  a.rowmap = b->d_rowmap;
  a.step_base = b->d_step;
  return a;
}

int batch_enqueue_steps(tdhf_batch* b, int n, int nw) {
  tdhf::ApplyArgs a = batch_args(b);
  for (int j = 0; j < n; ++j) {
    a.jstep = j;
    for (int m = tdhf::M_S1; m <= tdhf::M_S4; ++m) {
      a.mode = m;
      CHKRET(batch_launch(b, a, nw));
    }
  }
  return 0;
}

int batch_run_steps(tdhf_batch* b, int nt) {
  int done = 0;
  if (!b->timing && nt >= kGraphSteps) {
    if (b->gexec && (b->graph_nw != b->nw || b->graph_sched != b->sched ||
                     b->graph_xmode != b->xmode)) {
      (void)hipGraphExecDestroy(b->gexec);
      b->gexec = nullptr;
    }
    if (!b->gexec) {
      hipGraph_t g = nullptr;
      BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = batch_enqueue_steps(b, kGraphSteps, b->nw);
      if (rc == 0)
        hipLaunchKernelGGL(tdhf::step_advance_kernel, dim3(1), dim3(1), 0, b->stream, b->d_step,
                           kGraphSteps);
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return bfail(b, TDHF_EDEVICE, "graph capture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return bfail(b, TDHF_EDEVICE, "graph instantiate: %s", hipGetErrorString(e));
      b->graph_steps = kGraphSteps;
      b->graph_nw = b->nw;
      b->graph_sched = b->sched;
      b->graph_xmode = b->xmode;
    }
    while (nt - done >= b->graph_steps) {
      BHIP(b, hipGraphLaunch(b->gexec, b->stream));
      done += b->graph_steps;
    }
  }
  if (done < nt) {
    CHKRET(batch_enqueue_steps(b, nt - done, b->nw));
    hipLaunchKernelGGL(tdhf::step_advance_kernel, dim3(1), dim3(1), 0, b->stream, b->d_step,
                       nt - done);
    BHIP(b, hipGetLastError());
  }
  return 0;
}

int batch_ensure(tdhf_batch* b, long nt) {
  if (nt > b->A_cap || nt + 1 > b->obs_cap) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    dfree(b->d_A);
    dfree(b->d_Aave);
    dfree(b->d_obs);
    dfree(b->d_res);
    const long cap = std::max(nt, 1L);
    CHKRET(balloc(b, &b->d_A, (size_t)cap * b->ntraj));
    CHKRET(balloc(b, &b->d_Aave, (size_t)cap * b->ntraj));
    size_t per_slot = 0;
    for (int t = 0; t < b->ntraj; ++t) per_slot += 3 * (size_t)b->p.nb * b->nk[t] + 3;
    CHKRET(balloc(b, &b->d_obs, per_slot * (size_t)(cap + 1)));
    CHKRET(balloc(b, &b->d_res, (size_t)7 * b->ntraj * (cap + 1)));
    b->A_cap = cap;
    b->obs_cap = cap + 1;
    b->res_cap = cap + 1;
    CHKRET(batch_update_desc(b));
  }
  return 0;
}

void batch_release(tdhf_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  for (auto e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
  dfree(b->d_vion); dfree(b->d_state); dfree(b->d_td); dfree(b->d_rowmap); dfree(b->d_erows);
  dfree(b->d_A); dfree(b->d_Aave); dfree(b->d_obs); dfree(b->d_res); dfree(b->d_step);
  dfree(b->d_xpart); dfree(b->d_dpart); dfree(b->d_tiles);
  for (auto c : b->tabs) tdhf_destroy(c);
  b->tabs.clear();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->stream = nullptr;
}

// autotuner cache: shape -> (waves per workgroup, row schedule, exchange
// mode), shared by the batches of the process (guarded: batches may be
// created and tuned on different host threads)
std::map<std::vector<int>, std::array<int, 3>>& tune_cache() {
  static std::map<std::vector<int>, std::array<int, 3>> m;
  return m;
}
std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}
bool tune_lookup(const std::vector<int>& key, std::array<int, 3>* out) {
  std::lock_guard<std::mutex> lk(tune_mutex());
  auto it = tune_cache().find(key);
  if (it == tune_cache().end()) return false;
  *out = it->second;
  return true;
}
void tune_store(const std::vector<int>& key, const std::array<int, 3>& v) {
  std::lock_guard<std::mutex> lk(tune_mutex());
  tune_cache()[key] = v;
}

std::vector<int> tune_key(const tdhf_batch* b) {
  std::vector<int> k = {b->p.ng, b->p.nb, b->p.nbocc};
  std::vector<int> n = b->nk;
  std::sort(n.begin(), n.end());
  k.insert(k.end(), n.begin(), n.end());
  return k;
}

}  // namespace

extern "C" {

const char* tdhf_batch_last_error(const tdhf_batch* b) { return b ? b->err.c_str() : g_err.c_str(); }

int tdhf_batch_create(const tdhf_params* p, int ntraj, const int* nk, const double* vion,
                      tdhf_batch** out) {
  if (!out) return bfail(nullptr, TDHF_EINVAL, "out is NULL");
  *out = nullptr;
  if (!p || ntraj < 1 || !nk || !vion) return bfail(nullptr, TDHF_EINVAL, "invalid batch arguments");
  tdhf_params q = *p;
  for (int t = 0; t < ntraj; ++t) {
    q.nk = nk[t];
    CHKRET(valid_params(&q, nullptr));
  }
  tdhf_batch* b = new tdhf_batch();
  b->p = *p;
  b->ntraj = ntraj;
  b->nk.assign(nk, nk + ntraj);
  auto bail = [&](int rc) {
    g_err = b->err;
    batch_release(b);
    delete b;
    return rc;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    bfail(b, TDHF_EDEVICE, "hipStreamCreate failed");
    return bail(TDHF_EDEVICE);
  }
  // tables: one context per distinct NK (its ion potential is not used)
  std::map<int, int> idx;
  b->tab_of.resize(ntraj);
  for (int t = 0; t < ntraj; ++t) {
    auto it = idx.find(nk[t]);
    if (it == idx.end()) {
      q.nk = nk[t];
      tdhf_ctx* c = nullptr;
      int rc = tdhf_create(&q, vion + (size_t)p->ng * t, nullptr, nullptr, &c);
      if (rc) {
        b->err = g_err;
        return bail(rc);
      }
      it = idx.emplace(nk[t], (int)b->tabs.size()).first;
      b->tabs.push_back(c);
    }
    b->tab_of[t] = it->second;
  }
  b->off.resize(ntraj);
  b->obs_off.resize(ntraj);
  size_t acc = 0, oacc = 0;
  for (int t = 0; t < ntraj; ++t) {
    b->off[t] = acc;
    acc += bstate(b, t);
    b->obs_off[t] = oacc;
    oacc += 3 * (size_t)p->nb * nk[t] + 3;
  }
  b->total = acc;
  int rc;
  if ((rc = balloc(b, &b->d_vion, (size_t)p->ng * ntraj)) || (rc = balloc(b, &b->d_state, 4 * acc)) ||
      (rc = balloc(b, &b->d_td, (size_t)ntraj)) || (rc = balloc(b, &b->d_step, 1)))
    return bail(rc);
  if (hipMemcpy(b->d_vion, vion, sizeof(double) * p->ng * ntraj, hipMemcpyHostToDevice) != hipSuccess) {
    bfail(b, TDHF_EDEVICE, "vion upload failed");
    return bail(TDHF_EDEVICE);
  }
  // launch shape and row schedule: the cached autotuner choice, else the
  // fewest waves per workgroup that cover the grid (one at NG <= 64: the
  // groups of a single wave need no workgroup barrier, and eight such
  // workgroups share a CU) and the XCD-pinned schedule
  std::array<int, 3> ct{};
  const bool cached = tune_lookup(tune_key(b), &ct);
  b->nw = cached ? ct[0] : (p->ng + 63) / 64;
  b->ngrid = 0;
  if ((rc = batch_set_schedule(b, cached ? ct[1] : 0))) return bail(rc);
  b->h_td.resize(ntraj);
  if ((rc = batch_update_desc(b))) return bail(rc);
  if (cached && ct[2] && (rc = batch_set_xmode(b, ct[2]))) return bail(rc);
  (void)set_apply_lds_limit(p->ng);
  *out = b;
  return 0;
}

int tdhf_batch_destroy(tdhf_batch* b) {
  if (!b) return 0;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_release(b);
  delete b;
  return 0;
}

int tdhf_batch_upload(tdhf_batch* b, int traj, const double* u) {
  if (!b || !u || traj >= b->ntraj) return bfail(b, TDHF_EINVAL, "invalid upload arguments");
  if (traj < 0)
    BHIP(b, hipMemcpyAsync(b->d_state, u, b->total * sizeof(zd), hipMemcpyHostToDevice, b->stream));
  else
    BHIP(b, hipMemcpyAsync(b->d_state + b->off[traj], u, bstate(b, traj) * sizeof(zd),
                           hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return 0;
}

int tdhf_batch_download(tdhf_batch* b, int traj, double* u) {
  if (!b || !u || traj >= b->ntraj) return bfail(b, TDHF_EINVAL, "invalid download arguments");
  if (traj < 0)
    BHIP(b, hipMemcpyAsync(u, b->d_state, b->total * sizeof(zd), hipMemcpyDeviceToHost, b->stream));
  else
    BHIP(b, hipMemcpyAsync(u, b->d_state + b->off[traj], bstate(b, traj) * sizeof(zd),
                           hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return 0;
}

int tdhf_batch_propagate(tdhf_batch* b, int nt, const double* A, const double* Aave, double* etot,
                         double* nele, double* jtot, double* init3) {
  if (!b) return bfail(nullptr, TDHF_EINVAL, "batch is NULL");
  if (nt < 1 || !A || !Aave) return bfail(b, TDHF_EINVAL, "invalid field arrays");
  CHKRET(batch_ensure(b, nt));
  for (int t = 0; t < b->ntraj; ++t) {
    BHIP(b, hipMemcpyAsync(b->d_A + (size_t)b->A_cap * t, A + (size_t)nt * t, nt * sizeof(double),
                           hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_Aave + (size_t)b->A_cap * t, Aave + (size_t)nt * t,
                           nt * sizeof(double), hipMemcpyHostToDevice, b->stream));
  }
  const int zero = 0;
  BHIP(b, hipMemcpyAsync(b->d_step, &zero, sizeof zero, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  CHKRET(batch_run_steps(b, nt));
  // observables of the final states (slot nt)
  tdhf::ApplyArgs a = batch_args(b);
  a.mode = tdhf::M_OBS;
  a.jstep = -1;
  CHKRET(batch_launch(b, a, b->nw));
  const int ns = nt + 1;
  const int n = b->ntraj * ns;
  hipLaunchKernelGGL(tdhf::obs_reduce_batch_kernel, dim3((n + 63) / 64), dim3(64), 0, b->stream,
                     b->d_td, b->ntraj, b->p.nb, ns, std::fabs(b->p.acell) / (double)b->p.ng,
                     std::fabs(b->p.acell), b->d_res);
  BHIP(b, hipGetLastError());
  std::vector<double> res((size_t)7 * n);
  BHIP(b, hipMemcpyAsync(res.data(), b->d_res, res.size() * sizeof(double), hipMemcpyDeviceToHost,
                         b->stream));
  CHKRET(batch_harvest(b));
  for (int t = 0; t < b->ntraj; ++t) {
    const double* r0 = &res[(size_t)7 * ns * t];
    if (init3) {
      init3[3 * t] = r0[0];
      init3[3 * t + 1] = r0[1];
      init3[3 * t + 2] = r0[2];
    }
    for (int it = 0; it < nt; ++it) {
      const double* r = r0 + (size_t)7 * (it + 1);
      if (etot) etot[(size_t)nt * t + it] = r[0];
      if (nele) nele[(size_t)nt * t + it] = r[1];
      if (jtot) jtot[(size_t)nt * t + it] = r[2];
    }
  }
  return 0;
}

int tdhf_batch_run(const tdhf_params* p, int ntraj, const int* nk, const double* vion, double* u,
                   int nt, const double* A, const double* Aave, double* etot, double* nele,
                   double* jtot, double* init3) {
  tdhf_batch* b = nullptr;
  CHKRET(tdhf_batch_create(p, ntraj, nk, vion, &b));
  int rc = tdhf_batch_upload(b, -1, u);
  if (!rc) rc = tdhf_batch_propagate(b, nt, A, Aave, etot, nele, jtot, init3);
  if (!rc) rc = tdhf_batch_download(b, -1, u);
  if (rc) g_err = b->err;
  tdhf_batch_destroy(b);
  return rc;
}

int tdhf_batch_set_waves(tdhf_batch* b, int nw) {
  if (!b || !batch_valid_nw(b, nw)) return bfail(b, TDHF_EINVAL, "invalid waves per workgroup %d", nw);
  b->nw = nw;
  return 0;
}

int tdhf_batch_get_waves(const tdhf_batch* b) { return b ? b->nw : 0; }

int tdhf_batch_set_schedule(tdhf_batch* b, int sched) {
  if (!b || sched < 0 || sched >= kNumSchedules)
    return bfail(b, TDHF_EINVAL, "invalid row schedule %d", sched);
  return batch_set_schedule(b, sched);
}

int tdhf_batch_get_schedule(const tdhf_batch* b) { return b ? b->sched : -1; }

int tdhf_batch_set_exchange(tdhf_batch* b, int mode) {
  if (!b || mode < 0 || mode > 3) return bfail(b, TDHF_EINVAL, "invalid exchange mode %d", mode);
  return batch_set_xmode(b, mode);
}

int tdhf_batch_get_exchange(const tdhf_batch* b) { return b ? b->xmode : -1; }

int tdhf_batch_autotune(tdhf_batch* b, int nt_trial, const double* A, const double* Aave,
                        int* best_nw) {
  if (!b || nt_trial < 1 || !A || !Aave) return bfail(b, TDHF_EINVAL, "invalid autotune arguments");
  // keep the state: the trial steps run on a copy, freed on every return
  zd* save = nullptr;
  CHKRET(balloc(b, &save, b->total));
  struct FreeOnReturn {
    zd*& p;
    ~FreeOnReturn() { dfree(p); }
  } save_guard{save};
  BHIP(b, hipMemcpyAsync(save, b->d_state, b->total * sizeof(zd), hipMemcpyDeviceToDevice, b->stream));
  int best = b->nw, best_s = b->sched, best_x = b->xmode;
  double best_t = 1e300;
  const int cand[4] = {1, 2, 4, 8};
  for (int xm = 0; xm <= (batch_pairtile_ok(b) ? 3 : 0); ++xm)
  for (int sc = 0; sc < kNumSchedules; ++sc) {
    CHKRET(batch_set_xmode(b, xm));
    CHKRET(batch_set_schedule(b, sc));
    for (int nw : cand) {
      if (!batch_valid_nw(b, nw)) continue;
      b->nw = nw;
      int rc = 0;
      double dt_ms = 1e300;
      // the first repetition warms up (graph capture); the faster of the
      // next two counts
      for (int rep = 0; rep < 3 && !rc; ++rep) {
        BHIP(b, hipMemcpyAsync(b->d_state, save, b->total * sizeof(zd), hipMemcpyDeviceToDevice,
                               b->stream));
        BHIP(b, hipStreamSynchronize(b->stream));
        auto t0 = std::chrono::steady_clock::now();
        rc = tdhf_batch_propagate(b, nt_trial, A, Aave, nullptr, nullptr, nullptr, nullptr);
        auto t1 = std::chrono::steady_clock::now();
        if (rep > 0) dt_ms = std::min(dt_ms, std::chrono::duration<double, std::milli>(t1 - t0).count());
      }
      if (rc) return rc;
      if (dt_ms < best_t) {
        best_t = dt_ms;
        best = nw;
        best_s = sc;
        best_x = xm;
      }
    }
  }
  BHIP(b, hipMemcpyAsync(b->d_state, save, b->total * sizeof(zd), hipMemcpyDeviceToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  b->nw = best;
  CHKRET(batch_set_xmode(b, best_x));
  CHKRET(batch_set_schedule(b, best_s));
  tune_store(tune_key(b), {best, best_s, best_x});
  if (best_nw) *best_nw = best;
  return 0;
}

int tdhf_batch_set_kernel_timing(tdhf_batch* b, int enable) {
  if (!b) return bfail(nullptr, TDHF_EINVAL, "batch is NULL");
  CHKRET(batch_harvest(b));
  b->timing = enable != 0;
  b->launches = 0;
  b->total_ms = 0.0;
  return 0;
}

int tdhf_batch_kernel_stats(tdhf_batch* b, long* launches, double* total_ms) {
  if (!b) return bfail(nullptr, TDHF_EINVAL, "batch is NULL");
  CHKRET(batch_harvest(b));
  if (launches) *launches = b->launches;
  if (total_ms) *total_ms = b->total_ms;
  return 0;
}

}  // extern "C"
