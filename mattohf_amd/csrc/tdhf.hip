// tdhf.hip -- C-ABI host layer of libtdhf.so (gfx950).
//
// Owns the device state of one system (tables, orbital buffers, observables
// slots, the RK4 step graph) and implements the entry points of
// include/tdhf.h.  All arithmetic of the path runs in the kernels of
// tdhf_kernels.h / tdhf_gs.h; this file only sets up tables, launches and
// copies.  There is no CPU fallback: a missing or non-gfx950 device is an
// error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <map>
#include <mutex>
#include <numeric>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tdhf.h"
#include "tdhf_kernels.h"
#include "tdhf_pairtile.h"
#include "tdhf_gs.h"
#include "tdhf_adjoint.h"
#include "tdhf_adjtile.h"
#include "tdhf_post.h"

using tdhf::zd;

namespace {

thread_local std::string g_err = "no error";

// Stage buffers of the allocation-free RK4 step (tdhf_rk4_step_noalloc):
// ksum | sa | sb, each of at most kNoAllocMaxState complex elements
// (NG*NB*NK <= 131072, e.g. NG=128, NB=8, NK=128), 6 MiB in all.  A
// file-scope device array: every GPU the library runs on holds its own copy
// until the process ends.  Host code never names the array in a kernel
// argument or a copy; it uses the device address taken with
// hipGetSymbolAddress for each device (noalloc_scratch).
constexpr size_t kNoAllocMaxState = (size_t)1 << 17;
__device__ __align__(256) double scratch[6 * kNoAllocMaxState];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // per device id

}  // namespace

struct tdhf_ctx {
  tdhf_params p{};
  int ng = 0, nk = 0, nb = 0, nbocc = 0;
  double acell = 0, dt = 0, dv = 0, vcell = 0, bx = 0, occ = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  tdhf_progress_fn progress = nullptr;  // checkpointed adjoint calls (tdhf_set_progress)
  void* progress_user = nullptr;
  std::string err = "no error";
  std::vector<double> kx;
  // device tables
  double *d_kx = nullptr, *d_vion = nullptr, *d_vcg = nullptr;
  double *d_vq = nullptr, *d_vtl = nullptr;
  double* d_kc = nullptr;  // circulant kernels K_H, K_2, K_1 of the apply kernel
  double* d_vhg = nullptr; // vCG/NG in the transforms' G layout (pair-tile epilogue)
  zd* d_tw = nullptr;
  // orbital buffers (NG*NB*NK complex each)
  zd *d_u0 = nullptr, *d_ksum = nullptr, *d_sa = nullptr, *d_sb = nullptr;
  zd *d_w = nullptr, *d_out = nullptr;
  // observables
  double* d_obs = nullptr;
  long obs_slots = 0;
  int obs_stride = 0;
  double* d_res = nullptr;
  long res_slots = 0;
  // field arrays and the step counter
  double *d_A = nullptr, *d_Aave = nullptr;
  long A_cap = 0;
  int* d_step = nullptr;
  // step graphs: one per chunk length of kChunkSteps, captured together the
  // first time a run needs them (for the field / observables buffers in
  // graph_key)
  hipGraphExec_t gexec[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  bool graphs_ready = false;
  const void* graph_key[3] = {nullptr, nullptr, nullptr};
  // run opened by tdhf_rt_begin: field length and steps done so far
  bool rt_open = false;
  long rt_nt = 0, rt_done = 0;
  // launch shape of the apply kernel
  int nw = 8;
  size_t lds = 0;
  // kernel timing: per apply launch (timing) or per graph-replayed block of
  // steps (block_timing, block_launches[i] apply launches in event pair i)
  bool timing = false;
  bool block_timing = false;
  std::vector<long> block_launches;
  long launches = 0;
  double total_ms = 0.0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  // ground-state scratch
  tdhf::GsWork gs;
};

namespace {

int fail(tdhf_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  g_err = buf;
  return code;
}

#define HIPCHK(c, expr)                                                                      \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail((c), TDHF_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                  __FILE__, __LINE__);                                                       \
  } while (0)

#define CHKRET(expr)          \
  do {                        \
    int r_ = (expr);          \
    if (r_ != 0) return r_;   \
  } while (0)

size_t nstate(const tdhf_ctx* c) { return (size_t)c->ng * c->nb * c->nk; }

int check_device(tdhf_ctx* c) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return fail(c, TDHF_EDEVICE, "no HIP device available (%s)",
                e == hipSuccess ? "device count 0" : hipGetErrorString(e));
  HIPCHK(c, hipGetDevice(&c->device));
  hipDeviceProp_t prop;
  HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(c, TDHF_EDEVICE, "libtdhf is built for gfx950, device is %s", prop.gcnArchName);
  return 0;
}

// Device address of the allocation-free step's scratch array on the current
// device (which is c->device), taken once per device.
int noalloc_scratch_init(tdhf_ctx* c) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= c->device) g_scratch_ptr.resize(c->device + 1, nullptr);
  if (g_scratch_ptr[c->device]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(c, TDHF_EDEVICE, "hipGetSymbolAddress of the step scratch failed: %s",
                hipGetErrorString(e));
  g_scratch_ptr[c->device] = static_cast<double*>(ptr);
  return 0;
}

double* noalloc_scratch(const tdhf_ctx* c) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return c->device < (int)g_scratch_ptr.size() ? g_scratch_ptr[c->device] : nullptr;
}

template <typename T>
int dalloc(tdhf_ctx* c, T** p, size_t n) {
  if (*p) return 0;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T) + 16);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, TDHF_ENOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                hipGetErrorString(e));
  }
  return 0;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// Apply-kernel dispatch over the supported grid sizes.
// LDS of an apply_kernel launch
size_t lds_for(int ng, int nw) {
  switch (ng) {
    case 16: return tdhf::apply_lds_bytes<16>(nw);
    case 32: return tdhf::apply_lds_bytes<32>(nw);
    case 64: return tdhf::apply_lds_bytes<64>(nw);
    case 128: return tdhf::apply_lds_bytes<128>(nw);
  }
  return 0;
}

int choose_nw(int ng, int nk, int nbocc) {
  const int L = ng == 16 ? 4 : ng == 32 ? 4 : ng == 64 ? 8 : 16;
  const int gpw = 64 / L;
  const int items = nk * nbocc;
  int nw = (items + gpw - 1) / gpw;
  if (nw > 8) nw = 8;
  if (nw < 1) nw = 1;
  const int need = (ng + 63) / 64;  // epilogue uses one thread per grid point
  if (nw < need) nw = need;
  return nw;
}

// Compute units of the current device (queried once).
int device_cus() {
  static const int n = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    return v;
  }();
  return n;
}

template <int NG>
void launch_apply_t(const tdhf::ApplyArgs& a, int rows, int ntraj, int nw, size_t lds,
                    hipStream_t s) {
  tdhf::ApplyArgs b = a;
  // wave-priority balancing pays only when the grid is one workgroup per CU
  b.balance = (long)rows * ntraj <= device_cus() ? 1 : 0;
  if constexpr (NG >= 64) {
    if (tdhf::apply_by_transforms<NG>(nw)) {
      hipLaunchKernelGGL((tdhf::apply_kernel<NG, true>), dim3(rows, ntraj), dim3(64 * nw), lds, s,
                         b);
      return;
    }
  }
  hipLaunchKernelGGL((tdhf::apply_kernel<NG, false>), dim3(rows, ntraj), dim3(64 * nw), lds, s, b);
}

// Raises the dynamic LDS limit of every apply_kernel variant of grid size ng
// to the largest launch any context or batch makes (8 waves).  The limit is a
// property of the kernel, shared by every context of the process, so it is
// never lowered below what another one uses.
hipError_t set_apply_lds_limit(int ng) {
  const int cap = (int)lds_for(ng, 8);
  auto set = [cap](const void* fn) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  };
  switch (ng) {
    case 16: return set((const void*)tdhf::apply_kernel<16, false>);
    case 32: return set((const void*)tdhf::apply_kernel<32, false>);
    case 64: {
      const hipError_t e = set((const void*)tdhf::apply_kernel<64, false>);
      return e != hipSuccess ? e : set((const void*)tdhf::apply_kernel<64, true>);
    }
    case 128: {
      const hipError_t e = set((const void*)tdhf::apply_kernel<128, false>);
      return e != hipSuccess ? e : set((const void*)tdhf::apply_kernel<128, true>);
    }
  }
  return hipErrorInvalidValue;
}

// rows workgroups per trajectory; ntraj > 1 for a uniform batch (grid y)
// Two events of the context's pool for one timed interval.
int take_events(tdhf_ctx* c, hipEvent_t* e0, hipEvent_t* e1) {
  if (c->ev_used + 2 > c->ev_pool.size()) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      HIPCHK(c, hipEventCreate(&e));
      c->ev_pool.push_back(e);
    }
  }
  *e0 = c->ev_pool[c->ev_used++];
  *e1 = c->ev_pool[c->ev_used++];
  return 0;
}

int launch_apply(tdhf_ctx* c, const tdhf::ApplyArgs& a, int rows, int ntraj = 1) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    CHKRET(take_events(c, &e0, &e1));
    HIPCHK(c, hipEventRecord(e0, c->stream));
  }
  // a uniform batch with at least two workgroups per CU runs the fewest
  // waves per workgroup (as the batched autotuner chooses): the workgroups
  // sharing a CU then overlap their fixed phases (NG=128, 8 pulses: -7 % per
  // stage against the context's own launch shape)
  const bool many = ntraj > 1 && (long)rows * ntraj >= 2L * device_cus();
  const int nw = many ? (c->ng + 63) / 64 : c->nw;
  const size_t lds = many ? lds_for(c->ng, nw) : c->lds;
  switch (c->ng) {
    case 16: launch_apply_t<16>(a, rows, ntraj, nw, lds, c->stream); break;
    case 32: launch_apply_t<32>(a, rows, ntraj, nw, lds, c->stream); break;
    case 64: launch_apply_t<64>(a, rows, ntraj, nw, lds, c->stream); break;
    case 128: launch_apply_t<128>(a, rows, ntraj, nw, lds, c->stream); break;
    default: return fail(c, TDHF_EINVAL, "unsupported NG %d", c->ng);
  }
  HIPCHK(c, hipGetLastError());
  if (c->timing) HIPCHK(c, hipEventRecord(e1, c->stream));
  return 0;
}

// Collects the timed launches recorded so far (after a stream sync).
int harvest_timing(tdhf_ctx* c) {
  if (!c->timing && !c->block_timing) return 0;
  for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev_pool[i], c->ev_pool[i + 1]));
    c->total_ms += ms;
    c->launches += c->block_timing ? c->block_launches[i / 2] : 1;
  }
  c->ev_used = 0;
  c->block_launches.clear();
  return 0;
}

tdhf::ApplyArgs base_args(const tdhf_ctx* c) {
  tdhf::ApplyArgs a{};
  a.nk = c->nk;
  a.nb = c->nb;
  a.nbocc = c->nbocc;
  a.nbw = c->nb;
  a.bx = c->bx;
  a.dv = c->dv;
  a.occ = c->occ;
  a.dt = c->dt;
  a.kx = c->d_kx;
  a.vion = c->d_vion;
  a.vtl = c->d_vtl;
  a.kc = c->d_kc;
  a.vhg = c->d_vhg;
  a.tw = c->d_tw;
  a.obs_stride = c->obs_stride;
  return a;
}

int ensure_obs(tdhf_ctx* c, long slots) {
  if (slots > c->obs_slots) {
    dfree(c->d_obs);
    CHKRET(dalloc(c, &c->d_obs, (size_t)slots * c->obs_stride));
    c->obs_slots = slots;
  }
  if (slots > c->res_slots) {
    dfree(c->d_res);
    CHKRET(dalloc(c, &c->d_res, (size_t)slots * 7));
    c->res_slots = slots;
  }
  return 0;
}

int ensure_field(tdhf_ctx* c, long n) {
  if (n > c->A_cap) {
    dfree(c->d_A);
    dfree(c->d_Aave);
    CHKRET(dalloc(c, &c->d_A, (size_t)n));
    CHKRET(dalloc(c, &c->d_Aave, (size_t)n));
    c->A_cap = n;
  }
  return 0;
}

// Chunk lengths of the step graphs: a run of n steps is replayed as the
// binary decomposition of n over these (plus whole 32-step chunks), so any
// step count after the first run replays captured graphs only.
constexpr int kChunkSteps[6] = {32, 16, 8, 4, 2, 1};
constexpr int kNChunks = 6;

void destroy_graph(tdhf_ctx* c) {
  for (auto& g : c->gexec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  c->graphs_ready = false;
}

// Enqueues steps [*d_step + j0, *d_step + j0 + n) of the RK4 loop.
int enqueue_steps(tdhf_ctx* c, int j0, int n) {
  const int rows = c->nb * c->nk;
  tdhf::ApplyArgs a = base_args(c);
  a.Aarr = c->d_A;
  a.Aavearr = c->d_Aave;
  a.step_base = c->d_step;
  a.obs = c->d_obs;
  a.u0 = c->d_u0;
  a.ksum = c->d_ksum;
  for (int j = j0; j < j0 + n; ++j) {
    a.jstep = j;
    // stage 1: input u0 -> sa
    a.mode = tdhf::M_S1;
    a.u = a.w = c->d_u0;
    a.unext = c->d_sa;
    CHKRET(launch_apply(c, a, rows));
    a.mode = tdhf::M_S2;
    a.u = a.w = c->d_sa;
    a.unext = c->d_sb;
    CHKRET(launch_apply(c, a, rows));
    a.mode = tdhf::M_S3;
    a.u = a.w = c->d_sb;
    a.unext = c->d_sa;
    CHKRET(launch_apply(c, a, rows));
    a.mode = tdhf::M_S4;
    a.u = a.w = c->d_sa;
    a.unext = nullptr;
    CHKRET(launch_apply(c, a, rows));
  }
  return 0;
}

// Chunk length of the batched path's step graph.
constexpr int kGraphSteps = 25;

int run_steps_inner(tdhf_ctx* c, int nt);

// nt RK4 steps; with block timing the whole sequence (graph-replayed chunks,
// 4 apply launches per step) is one timed interval.
int run_steps(tdhf_ctx* c, int nt) {
  if (!c->block_timing || nt < 1) return run_steps_inner(c, nt);
  hipEvent_t b0 = nullptr, b1 = nullptr;
  CHKRET(take_events(c, &b0, &b1));
  HIPCHK(c, hipEventRecord(b0, c->stream));
  CHKRET(run_steps_inner(c, nt));
  HIPCHK(c, hipEventRecord(b1, c->stream));
  c->block_launches.push_back(4L * nt);
  return 0;
}

// Captures the graph of every chunk length: kChunkSteps[i] steps followed by
// the advance of the device step counter.
int capture_graphs(tdhf_ctx* c) {
  destroy_graph(c);
  for (int i = 0; i < kNChunks; ++i) {
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_steps(c, 0, kChunkSteps[i]);
    if (rc == 0)
      hipLaunchKernelGGL(tdhf::step_advance_kernel, dim3(1), dim3(1), 0, c->stream, c->d_step,
                         kChunkSteps[i]);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != 0 || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      destroy_graph(c);
      return rc != 0 ? rc : fail(c, TDHF_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&c->gexec[i], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->gexec[i] = nullptr;
      destroy_graph(c);
      return fail(c, TDHF_EDEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
    }
  }
  c->graphs_ready = true;
  c->graph_key[0] = c->d_A;
  c->graph_key[1] = c->d_obs;
  c->graph_key[2] = c->d_Aave;
  return 0;
}

int run_steps_inner(tdhf_ctx* c, int nt) {
  if (nt < 1) return 0;
  if (c->timing) {
    // an event pair around every launch: plain launches
    CHKRET(enqueue_steps(c, 0, nt));
    hipLaunchKernelGGL(tdhf::step_advance_kernel, dim3(1), dim3(1), 0, c->stream, c->d_step, nt);
    HIPCHK(c, hipGetLastError());
    return 0;
  }
  if (c->graphs_ready && (c->graph_key[0] != (const void*)c->d_A ||
                          c->graph_key[1] != (const void*)c->d_obs ||
                          c->graph_key[2] != (const void*)c->d_Aave))
    destroy_graph(c);
  if (!c->graphs_ready) CHKRET(capture_graphs(c));
  int done = 0;
  for (int i = 0; i < kNChunks; ++i)
    while (nt - done >= kChunkSteps[i]) {
      HIPCHK(c, hipGraphLaunch(c->gexec[i], c->stream));
      done += kChunkSteps[i];
    }
  return 0;
}

int upload(tdhf_ctx* c, zd* dst, const double* src) {
  HIPCHK(c, hipMemcpyAsync(dst, src, nstate(c) * sizeof(zd), hipMemcpyHostToDevice, c->stream));
  return 0;
}

int download(tdhf_ctx* c, double* dst, const zd* src) {
  HIPCHK(c, hipMemcpyAsync(dst, src, nstate(c) * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int sync(tdhf_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return harvest_timing(c);
}

// Observables of the state in buffer u (rows nb) at field A into slot 0 of
// the observables area; result (7 doubles) copied to res7.
int observables_of(tdhf_ctx* c, const zd* u, double A, double* res7) {
  CHKRET(ensure_obs(c, 1));
  tdhf::ApplyArgs a = base_args(c);
  a.mode = tdhf::M_OBS;
  a.u = a.w = u;
  a.A = A;
  a.Aobs = A;
  a.obs = c->d_obs;
  a.obs_slot = 0;
  CHKRET(launch_apply(c, a, c->nb * c->nk));
  hipLaunchKernelGGL(tdhf::obs_reduce_kernel, dim3(1), dim3(64), 0, c->stream, c->d_obs,
                     c->obs_stride, c->nb * c->nk, 1, c->dv, c->vcell, c->d_res);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(res7, c->d_res, 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int valid_params(const tdhf_params* p, tdhf_ctx* c) {
  if (!p) return fail(c, TDHF_EINVAL, "params is NULL");
  if (p->ng != 16 && p->ng != 32 && p->ng != 64 && p->ng != 128)
    return fail(c, TDHF_EINVAL, "NG=%d unsupported (16, 32, 64 or 128)", p->ng);
  if (p->nk < 1 || p->nb < 1 || p->nbocc < 1 || p->nbocc > p->nb)
    return fail(c, TDHF_EINVAL, "invalid NK=%d NB=%d NBocc=%d", p->nk, p->nb, p->nbocc);
  if (!(p->acell > 0.0) || !std::isfinite(p->acell))
    return fail(c, TDHF_EINVAL, "invalid acell=%g", p->acell);
  if (!std::isfinite(p->dt)) return fail(c, TDHF_EINVAL, "invalid dt");
  return 0;
}

long kvck_nsteps(double acell, int ng, int nk) {
  // the floating-point loop bound of k_vCk, src/Functions.f90:42-48
  const double dx = acell / (double)ng / 20.0;
  const double L = (double)nk * acell;
  long ix = 0;
  while ((double)ix * dx < L) ++ix;
  return ix;
}

int relayout(tdhf_ctx* c, const double* raw, int nrows, double scale, double* out) {
  const int n = nrows * c->ng;
  const dim3 b(256), g((n + 255) / 256);
  switch (c->ng) {
    case 16: hipLaunchKernelGGL(tdhf::relayout_kernel<16>, g, b, 0, c->stream, raw, nrows, scale, out); break;
    case 32: hipLaunchKernelGGL(tdhf::relayout_kernel<32>, g, b, 0, c->stream, raw, nrows, scale, out); break;
    case 64: hipLaunchKernelGGL(tdhf::relayout_kernel<64>, g, b, 0, c->stream, raw, nrows, scale, out); break;
    case 128: hipLaunchKernelGGL(tdhf::relayout_kernel<128>, g, b, 0, c->stream, raw, nrows, scale, out); break;
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

int build_tables(tdhf_ctx* c, const double* vcg, const double* vcgkk) {
  const int NG = c->ng, NK = c->nk, ND = 2 * NK - 1;
  std::vector<double> gx(NG);
  for (int i = 0; i < NG / 2; ++i) gx[i] = (double)i * c->bx;
  for (int i = NG / 2; i < NG; ++i) gx[i] = (double)(i - NG) * c->bx;
  CHKRET(dalloc(c, &c->d_vcg, NG));
  CHKRET(dalloc(c, &c->d_vq, (size_t)NG * ND));
  if (!vcg || !vcgkk) {
    // arguments G + kx(ikr) - kx(ikc) (src/Initialization.f90:130-137), one
    // representative pair per difference d = ikr - ikc, then G alone (:143-146)
    std::vector<double> q((size_t)NG * ND + NG);
    for (int d = -(NK - 1); d <= NK - 1; ++d) {
      const int ikr = d >= 0 ? d : 0, ikc = d >= 0 ? 0 : -d;
      for (int ig = 0; ig < NG; ++ig)
        q[(size_t)(d + NK - 1) * NG + ig] = gx[ig] + c->kx[ikr] - c->kx[ikc];
    }
    for (int ig = 0; ig < NG; ++ig) q[(size_t)NG * ND + ig] = gx[ig];
    double *d_q = nullptr, *d_v = nullptr;
    CHKRET(dalloc(c, &d_q, q.size()));
    int rc = dalloc(c, &d_v, q.size());
    if (rc) { dfree(d_q); return rc; }
    hipError_t e = hipMemcpyAsync(d_q, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream);
    const double dx = c->acell / (double)NG / 20.0;
    const long ns = kvck_nsteps(c->acell, NG, NK);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(tdhf::kvck_kernel, dim3((unsigned)q.size()), dim3(256), 0, c->stream,
                         d_q, dx, ns, d_v);
      e = hipGetLastError();
    }
    if (e == hipSuccess && !vcgkk)
      e = hipMemcpyAsync(c->d_vq, d_v, (size_t)NG * ND * sizeof(double),
                         hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess && !vcg)
      e = hipMemcpyAsync(c->d_vcg, d_v + (size_t)NG * ND, NG * sizeof(double),
                         hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(d_q);
    dfree(d_v);
    if (e != hipSuccess) return fail(c, TDHF_EDEVICE, "k_vCk table: %s", hipGetErrorString(e));
    if (!vcg) {
      const double zero = 0.0;  // vCG(1) = 0, src/Initialization.f90:146
      HIPCHK(c, hipMemcpy(c->d_vcg, &zero, sizeof zero, hipMemcpyHostToDevice));
    }
  }
  if (vcg) HIPCHK(c, hipMemcpy(c->d_vcg, vcg, NG * sizeof(double), hipMemcpyHostToDevice));
  if (vcgkk) {
    // keep one row per k-difference: row d from the pair (ikr, ikc) above
    std::vector<double> rows((size_t)NG * ND);
    for (int d = -(NK - 1); d <= NK - 1; ++d) {
      const int ikr = d >= 0 ? d : 0, ikc = d >= 0 ? 0 : -d;
      std::memcpy(&rows[(size_t)(d + NK - 1) * NG], vcgkk + (size_t)NG * (ikr * NK + ikc),
                  NG * sizeof(double));
    }
    HIPCHK(c, hipMemcpy(c->d_vq, rows.data(), rows.size() * sizeof(double),
                        hipMemcpyHostToDevice));
  }
  CHKRET(dalloc(c, &c->d_vtl, (size_t)NG * ND));
  CHKRET(dalloc(c, &c->d_tw, 2 * NG));
  switch (NG) {
    case 16: hipLaunchKernelGGL(tdhf::twiddle_kernel<16>, dim3(1), dim3(256), 0, c->stream, c->d_tw); break;
    case 32: hipLaunchKernelGGL(tdhf::twiddle_kernel<32>, dim3(1), dim3(256), 0, c->stream, c->d_tw); break;
    case 64: hipLaunchKernelGGL(tdhf::twiddle_kernel<64>, dim3(1), dim3(256), 0, c->stream, c->d_tw); break;
    case 128: hipLaunchKernelGGL(tdhf::twiddle_kernel<128>, dim3(1), dim3(256), 0, c->stream, c->d_tw); break;
  }
  HIPCHK(c, hipGetLastError());
  {
    // circulant kernels of the Hartree and kinetic round trips:
    // K_H(d) = 1/NG sum_g vCG(g) cos(2 pi g d/NG), K_2(d) = 1/NG sum_g G^2/2
    // cos(2 pi g d/NG), K_1(d) = 1/NG sum_g G exp(2 pi i g d/NG), G in FFT order
    // (src/Initialization.f90:97-102); summed in long double.
    std::vector<double> vh(NG), kc(4 * (size_t)NG);
    HIPCHK(c, hipMemcpy(vh.data(), c->d_vcg, NG * sizeof(double), hipMemcpyDeviceToHost));
    const long double tpi = 6.283185307179586476925286766559L;
    for (int d = 0; d < NG; ++d) {
      long double kh = 0.0L, k2 = 0.0L, k1r = 0.0L, k1i = 0.0L;
      for (int g = 0; g < NG; ++g) {
        const long double th = tpi * (long double)((g * d) % NG) / (long double)NG;
        const long double cs = cosl(th), sn = sinl(th);
        const long double G = (long double)gx[g];
        kh += (long double)vh[g] * cs;
        k2 += 0.5L * G * G * cs;
        k1r += G * cs;
        k1i += G * sn;
      }
      kc[d] = (double)(kh / NG);
      kc[NG + d] = (double)(k2 / NG);
      kc[2 * NG + 2 * d] = (double)(k1r / NG);
      kc[2 * NG + 2 * d + 1] = (double)(k1i / NG);
    }
    CHKRET(dalloc(c, &c->d_kc, kc.size()));
    HIPCHK(c, hipMemcpy(c->d_kc, kc.data(), kc.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  CHKRET(relayout(c, c->d_vq, ND, -1.0 / ((double)NG * (double)NK), c->d_vtl));
  CHKRET(dalloc(c, &c->d_vhg, NG));
  CHKRET(relayout(c, c->d_vcg, 1, 1.0 / (double)NG, c->d_vhg));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

void release(tdhf_ctx* c) {
  destroy_graph(c);
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  c->ev_pool.clear();
  dfree(c->d_kx); dfree(c->d_vion); dfree(c->d_vcg);
  dfree(c->d_vq); dfree(c->d_vtl); dfree(c->d_tw); dfree(c->d_kc); dfree(c->d_vhg);
  dfree(c->d_u0); dfree(c->d_ksum); dfree(c->d_sa); dfree(c->d_sb);
  dfree(c->d_w); dfree(c->d_out);
  dfree(c->d_obs); dfree(c->d_res); dfree(c->d_A); dfree(c->d_Aave); dfree(c->d_step);
  tdhf::gs_free(c->gs);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// ground-state helpers shared with tdhf_gs.h
// ---------------------------------------------------------------------------
namespace tdhf {

int ctx_apply(tdhf_ctx* c, const zd* u, const zd* w, int nbw, double A, zd* out) {
  ApplyArgs a = base_args(c);
  a.mode = M_HPSI;
  a.u = u;
  a.w = w;
  a.nbw = nbw;
  a.A = A;
  a.Aobs = A;
  a.out = out;
  return launch_apply(c, a, nbw * c->nk);
}

int ctx_fail(tdhf_ctx* c, int code, const char* msg) { return fail(c, code, "%s", msg); }

}  // namespace tdhf

#include "tdhf_gs_impl.h"
#include "tdhf_adjoint_impl.h"

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* tdhf_last_error(const tdhf_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int tdhf_create(const tdhf_params* p, const double* vion, const double* vcg, const double* vcgkk,
                tdhf_ctx** out) {
  if (!out) return fail(nullptr, TDHF_EINVAL, "out is NULL");
  *out = nullptr;
  CHKRET(valid_params(p, nullptr));
  if (!vion) return fail(nullptr, TDHF_EINVAL, "vion is NULL");
  tdhf_ctx* c = new tdhf_ctx();
  int rc = check_device(c);
  if (rc) {
    g_err = c->err;
    delete c;
    return rc;
  }
  c->p = *p;
  c->ng = p->ng;
  c->nk = p->nk;
  c->nb = p->nb;
  c->nbocc = p->nbocc;
  c->acell = p->acell;
  c->dt = p->dt;
  c->vcell = std::fabs(p->acell);
  c->dv = c->vcell / (double)c->ng;
  c->bx = 2.0 * (4.0 * std::atan(1.0)) / p->acell;
  c->occ = 2.0 / (double)c->nk;
  c->kx.resize(c->nk);
  for (int i = 0; i < c->nk; ++i)
    c->kx[i] = (-0.5 + (2.0 * (i + 1) - 1.0) / (double)(2 * c->nk)) * c->bx;
  c->obs_stride = 3 * c->nb * c->nk + 3;
  c->nw = choose_nw(c->ng, c->nk, c->nbocc);
  c->lds = lds_for(c->ng, c->nw);
  auto bail = [&](int code) {
    g_err = c->err;
    release(c);
    delete c;
    return code;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    fail(c, TDHF_EDEVICE, "hipStreamCreate failed");
    return bail(TDHF_EDEVICE);
  }
  const size_t n = nstate(c);
  if ((rc = dalloc(c, &c->d_kx, c->nk)) || (rc = dalloc(c, &c->d_vion, c->ng)) ||
      (rc = dalloc(c, &c->d_u0, n)) || (rc = dalloc(c, &c->d_ksum, n)) ||
      (rc = dalloc(c, &c->d_sa, n)) || (rc = dalloc(c, &c->d_sb, n)) ||
      (rc = dalloc(c, &c->d_w, n)) || (rc = dalloc(c, &c->d_out, n)) ||
      (rc = dalloc(c, &c->d_step, 1)))
    return bail(rc);
  if (hipMemcpy(c->d_kx, c->kx.data(), c->nk * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(c->d_vion, vion, c->ng * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    fail(c, TDHF_EDEVICE, "table upload failed");
    return bail(TDHF_EDEVICE);
  }
  if ((rc = build_tables(c, vcg, vcgkk))) return bail(rc);
  {
    const hipError_t e = set_apply_lds_limit(c->ng);
    if (e != hipSuccess && c->lds > 64 * 1024) {
      fail(c, TDHF_EDEVICE, "cannot raise LDS limit to %zu bytes: %s", lds_for(c->ng, 8),
           hipGetErrorString(e));
      return bail(TDHF_EDEVICE);
    }
  }
  if ((rc = noalloc_scratch_init(c))) return bail(rc);
  *out = c;
  return 0;
}

int tdhf_destroy(tdhf_ctx* c) {
  if (!c) return 0;
  (void)hipStreamSynchronize(c->stream);
  adj_release(c);
  release(c);
  delete c;
  return 0;
}

int tdhf_ion_potential(const tdhf_params* p, int nion, const double* z, const double* alpha,
                       const double* rion, double* vion) {
  CHKRET(valid_params(p, nullptr));
  if (nion < 0 || (nion > 0 && (!z || !alpha || !rion)) || !vion)
    return fail(nullptr, TDHF_EINVAL, "invalid ion arguments");
  tdhf_ctx tmp;
  CHKRET(check_device(&tmp));
  double *d = nullptr;
  const size_t n = (size_t)3 * (nion > 0 ? nion : 1) + p->ng;
  if (hipMalloc((void**)&d, n * sizeof(double)) != hipSuccess)
    return fail(nullptr, TDHF_ENOMEM, "hipMalloc failed");
  double *dz = d, *da = d + (nion > 0 ? nion : 1), *dr = da + (nion > 0 ? nion : 1),
         *dv = dr + (nion > 0 ? nion : 1);
  hipError_t e = hipSuccess;
  if (nion > 0) {
    e = hipMemcpy(dz, z, nion * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(da, alpha, nion * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dr, rion, nion * sizeof(double), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(tdhf::ion_kernel, dim3(1), dim3(128), 0, 0, p->ng, p->nk, p->acell, nion,
                       dz, da, dr, dv);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(vion, dv, p->ng * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(nullptr, TDHF_EDEVICE, "ion potential: %s", hipGetErrorString(e));
  return 0;
}

int tdhf_get_tables(tdhf_ctx* c, double* vcg, double* vq) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (vcg) HIPCHK(c, hipMemcpy(vcg, c->d_vcg, c->ng * sizeof(double), hipMemcpyDeviceToHost));
  if (vq)
    HIPCHK(c, hipMemcpy(vq, c->d_vq, (size_t)c->ng * (2 * c->nk - 1) * sizeof(double),
                        hipMemcpyDeviceToHost));
  return 0;
}

int tdhf_hpsi(tdhf_ctx* c, const double* u, const double* w, double A, double* hw) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u || !w || !hw) return fail(c, TDHF_EINVAL, "NULL array");
  CHKRET(upload(c, c->d_sa, u));
  CHKRET(upload(c, c->d_w, w));
  CHKRET(tdhf::ctx_apply(c, c->d_sa, c->d_w, c->nb, A, c->d_out));
  CHKRET(download(c, hw, c->d_out));
  return sync(c);
}

int tdhf_observables(tdhf_ctx* c, const double* u, double A, double* etot, double* ecomps,
                     double* nele, double* je) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u) return fail(c, TDHF_EINVAL, "NULL array");
  CHKRET(upload(c, c->d_sa, u));
  double r[7];
  CHKRET(observables_of(c, c->d_sa, A, r));
  if (etot) *etot = r[0];
  if (nele) *nele = r[1];
  if (je) *je = r[2];
  if (ecomps) {
    ecomps[0] = r[3];
    ecomps[1] = r[4];
    ecomps[2] = r[5];
    ecomps[3] = r[6];
  }
  return 0;
}

int tdhf_state_upload(tdhf_ctx* c, const double* u) {
  if (!c || !u) return fail(c, TDHF_EINVAL, "NULL argument");
  CHKRET(upload(c, c->d_u0, u));
  return sync(c);
}

int tdhf_state_download(tdhf_ctx* c, double* u) {
  if (!c || !u) return fail(c, TDHF_EINVAL, "NULL argument");
  return download(c, u, c->d_u0);
}

int tdhf_rt_begin(tdhf_ctx* c, int nt, const double* A, const double* Aave) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (nt < 0 || (nt > 0 && (!A || !Aave))) return fail(c, TDHF_EINVAL, "invalid field arrays");
  c->rt_open = false;
  CHKRET(ensure_obs(c, (long)nt + 1));
  if (nt > 0) {
    CHKRET(ensure_field(c, nt));
    HIPCHK(c, hipMemcpyAsync(c->d_A, A, nt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_Aave, Aave, nt * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
  }
  const int zero = 0;
  HIPCHK(c, hipMemcpyAsync(c->d_step, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->rt_open = true;
  c->rt_nt = nt;
  c->rt_done = 0;
  return 0;
}

int tdhf_rt_advance(tdhf_ctx* c, int n) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!c->rt_open) return fail(c, TDHF_EINVAL, "tdhf_rt_advance without tdhf_rt_begin");
  if (n < 0 || c->rt_done + n > c->rt_nt)
    return fail(c, TDHF_EINVAL, "advance by %d steps: %ld of %ld steps done", n, c->rt_done,
                c->rt_nt);
  CHKRET(run_steps(c, n));
  CHKRET(sync(c));
  c->rt_done += n;
  return 0;
}

int tdhf_rt_end(tdhf_ctx* c, double* etot, double* nele, double* jtot, double* init3) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!c->rt_open) return fail(c, TDHF_EINVAL, "tdhf_rt_end without tdhf_rt_begin");
  c->rt_open = false;
  const long done = c->rt_done;
  if (done == 0) {
    double r[7];
    CHKRET(observables_of(c, c->d_u0, 0.0, r));
    if (init3) { init3[0] = r[0]; init3[1] = r[1]; init3[2] = r[2]; }
    return 0;
  }
  const int rows = c->nb * c->nk;
  // observables of the final state: step = *d_step - 1 = done - 1, slot done
  // (those of the earlier states came from stage 1 of the next step)
  tdhf::ApplyArgs a = base_args(c);
  a.mode = tdhf::M_OBS;
  a.u = a.w = c->d_u0;
  a.Aarr = c->d_A;
  a.Aavearr = c->d_Aave;
  a.step_base = c->d_step;
  a.jstep = -1;
  a.obs = c->d_obs;
  CHKRET(launch_apply(c, a, rows));
  const long ns = done + 1;
  hipLaunchKernelGGL(tdhf::obs_reduce_kernel, dim3((unsigned)((ns + 63) / 64)), dim3(64), 0,
                     c->stream, c->d_obs, c->obs_stride, rows, (int)ns, c->dv, c->vcell, c->d_res);
  HIPCHK(c, hipGetLastError());
  std::vector<double> res((size_t)7 * ns);
  HIPCHK(c, hipMemcpyAsync(res.data(), c->d_res, res.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  CHKRET(sync(c));
  if (init3) { init3[0] = res[0]; init3[1] = res[1]; init3[2] = res[2]; }
  for (long it = 0; it < done; ++it) {
    const double* r = &res[(size_t)7 * (it + 1)];
    if (etot) etot[it] = r[0];
    if (nele) nele[it] = r[1];
    if (jtot) jtot[it] = r[2];
  }
  return 0;
}

int tdhf_rt_propagate(tdhf_ctx* c, int nt, const double* A, const double* Aave, double* etot,
                      double* nele, double* jtot, double* init3) {
  CHKRET(tdhf_rt_begin(c, nt, A, Aave));
  CHKRET(tdhf_rt_advance(c, nt));
  return tdhf_rt_end(c, etot, nele, jtot, init3);
}

int tdhf_rt_run(tdhf_ctx* c, double* u, int nt, const double* A, const double* Aave, double* etot,
                double* nele, double* jtot, double* init3) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u) return fail(c, TDHF_EINVAL, "NULL array");
  CHKRET(tdhf_state_upload(c, u));
  CHKRET(tdhf_rt_propagate(c, nt, A, Aave, etot, nele, jtot, init3));
  return tdhf_state_download(c, u);
}

int tdhf_rk4_step(tdhf_ctx* c, double* u, double Aave, double A) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u) return fail(c, TDHF_EINVAL, "NULL array");
  CHKRET(tdhf_state_upload(c, u));
  CHKRET(ensure_obs(c, 2));
  CHKRET(ensure_field(c, 1));
  HIPCHK(c, hipMemcpyAsync(c->d_A, &A, sizeof A, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_Aave, &Aave, sizeof Aave, hipMemcpyHostToDevice, c->stream));
  const int zero = 0;
  HIPCHK(c, hipMemcpyAsync(c->d_step, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  CHKRET(enqueue_steps(c, 0, 1));
  CHKRET(sync(c));
  return tdhf_state_download(c, u);
}

int tdhf_rk4_step_noalloc(tdhf_ctx* c, double* u, double Aave, double A, void* stream) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u) return fail(c, TDHF_EINVAL, "NULL array");
  const size_t n = nstate(c);
  if (n > kNoAllocMaxState)
    return fail(c, TDHF_EINVAL,
                "state of %zu complex elements (NG*NB*NK) exceeds the allocation-free step's "
                "limit of %zu",
                n, kNoAllocMaxState);
  double* sc = noalloc_scratch(c);
  if (!sc) return fail(c, TDHF_EDEVICE, "step scratch not set up on device %d", c->device);
  zd* ksum = reinterpret_cast<zd*>(sc);
  zd* sa = ksum + kNoAllocMaxState;
  zd* sb = sa + kNoAllocMaxState;
  zd* u0 = reinterpret_cast<zd*>(u);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  // the four stages of RT_RK4_ubk with the fields as kernel arguments
  // (no step counter, no field arrays, no observables)
  tdhf::ApplyArgs a = base_args(c);
  a.u0 = u0;
  a.ksum = ksum;
  const int rows = c->nb * c->nk;
  const int modes[4] = {tdhf::M_S1, tdhf::M_S2, tdhf::M_S3, tdhf::M_S4};
  const double fields[4] = {2.0 * Aave - A, Aave, Aave, A};
  zd* ins[4] = {u0, sa, sb, sa};
  zd* nexts[4] = {sa, sb, sa, nullptr};
  for (int s = 0; s < 4; ++s) {
    a.mode = modes[s];
    a.A = a.Aobs = fields[s];
    a.u = a.w = ins[s];
    a.unext = nexts[s];
    switch (c->ng) {
      case 16: launch_apply_t<16>(a, rows, 1, c->nw, c->lds, st); break;
      case 32: launch_apply_t<32>(a, rows, 1, c->nw, c->lds, st); break;
      case 64: launch_apply_t<64>(a, rows, 1, c->nw, c->lds, st); break;
      case 128: launch_apply_t<128>(a, rows, 1, c->nw, c->lds, st); break;
    }
    HIPCHK(c, hipGetLastError());
  }
  return 0;
}

int tdhf_gram_schmidt(tdhf_ctx* c, double* u) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u) return fail(c, TDHF_EINVAL, "NULL array");
  CHKRET(upload(c, c->d_sa, u));
  hipLaunchKernelGGL(tdhf::gram_schmidt_kernel, dim3(c->nk), dim3(c->ng), 0, c->stream, c->d_sa,
                     c->ng, c->nb, c->vcell);
  HIPCHK(c, hipGetLastError());
  return download(c, u, c->d_sa);
}

int tdhf_set_kernel_timing(tdhf_ctx* c, int enable) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (enable < 0 || enable > 2) return fail(c, TDHF_EINVAL, "timing mode %d (0, 1 or 2)", enable);
  CHKRET(sync(c));
  c->timing = enable == 1;
  c->block_timing = enable == 2;
  c->block_launches.clear();
  c->launches = 0;
  c->total_ms = 0.0;
  return 0;
}

int tdhf_kernel_stats(tdhf_ctx* c, long* launches, double* total_ms) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  CHKRET(sync(c));
  if (launches) *launches = c->launches;
  if (total_ms) *total_ms = c->total_ms;
  return 0;
}

}  // extern "C"

// batched trajectories (tdhf_batch_*)
#include "tdhf_batch_impl.h"
#include "tdhf_post_impl.h"
#include "tdhf_probe.h"
