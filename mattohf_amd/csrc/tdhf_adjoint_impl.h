// tdhf_adjoint_impl.h -- host side of the forward/backward pair (included by
// tdhf.hip after the single-trajectory helpers).
//
// tdhf_fwd propagates nt RK4 steps and keeps what the backward sweep needs:
//   - when 4*nt stage states fit the memory budget, every stage state y1..y4
//     of every step (the forward kernels write them straight into the tape);
//   - otherwise the state at every S-th step, S = ceil(sqrt(nt)) (or the
//     least-memory spacing ~sqrt(nt/4) if that does not fit), and tdhf_bwd
//     recomputes each segment once, writing its stage states into a segment
//     tape.
// tdhf_bwd starts from dE/du at the final state and runs the reverse RK4
// stages (adjoint_kernel, two exchange transforms per pair), collecting
// dL/dA per step; A enters step it as A(it-1) (stage 1), Aave(it) (stages 2,
// 3) and A(it) (stage 4), with Aave(it) = (A(it) + A(it-1))/2.
#pragma once

struct AdjState {
  bool have_fwd = false;
  int nt = 0;
  int seg = 0;            // steps per segment (== nt when the whole tape is kept)
  bool full = false;
  std::vector<double> A, Aave;
  double jN = 0.0;
  zd* tape = nullptr;     // 4 * seg states
  zd* ckpt = nullptr;     // ceil(nt/seg) states (segment starts), checkpoint mode
  zd* buf = nullptr;      // UA | UB | GA | GB
  double* abar = nullptr; // nt * 4 * rows
  size_t tape_states = 0, ckpt_states = 0;
  long abar_cap = 0;
};

namespace {

std::map<const tdhf_ctx*, AdjState>& adj_states() {
  static std::map<const tdhf_ctx*, AdjState> m;
  return m;
}

void adj_free(AdjState& s) {
  dfree(s.tape);
  dfree(s.ckpt);
  dfree(s.buf);
  dfree(s.abar);
  s = AdjState();
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

int launch_adj(tdhf_ctx* c, const tdhf::AdjArgs& a) {
  const int rows = c->nb * c->nk;
  int nw = choose_nw(c->ng, c->nk, c->nb);
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
  const void* fn = nullptr;
  switch (c->ng) {
    case 16: fn = (const void*)tdhf::adjoint_kernel<16>; break;
    case 32: fn = (const void*)tdhf::adjoint_kernel<32>; break;
    case 64: fn = (const void*)tdhf::adjoint_kernel<64>; break;
    case 128: fn = (const void*)tdhf::adjoint_kernel<128>; break;
  }
  if (lds > 64 * 1024)
    HIPCHK(c, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  switch (c->ng) {
    case 16: hipLaunchKernelGGL(tdhf::adjoint_kernel<16>, dim3(rows), dim3(64 * nw), lds, c->stream, a); break;
    case 32: hipLaunchKernelGGL(tdhf::adjoint_kernel<32>, dim3(rows), dim3(64 * nw), lds, c->stream, a); break;
    case 64: hipLaunchKernelGGL(tdhf::adjoint_kernel<64>, dim3(rows), dim3(64 * nw), lds, c->stream, a); break;
    case 128: hipLaunchKernelGGL(tdhf::adjoint_kernel<128>, dim3(rows), dim3(64 * nw), lds, c->stream, a); break;
  }
  HIPCHK(c, hipGetLastError());
  if (c->timing) HIPCHK(c, hipEventRecord(e1, c->stream));
  return 0;
}

tdhf::AdjArgs adj_base(const tdhf_ctx* c) {
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
  a.vcgl = c->d_vcgl;
  a.vtl = c->d_vtl;
  a.tw = c->d_tw;
  return a;
}

// Forward steps [s0, s0+n) on d_u0 with the stage states written to tape
// (tape slot 0 = step s0) or, when tape == nullptr, untaped.
int adj_forward_steps(tdhf_ctx* c, int s0, int n, zd* tape) {
  const int rows = c->nb * c->nk;
  tdhf::ApplyArgs a = base_args(c);
  a.Aarr = c->d_A;
  a.Aavearr = c->d_Aave;
  a.step_base = c->d_step;
  a.obs = c->d_obs;
  a.u0 = c->d_u0;
  a.ksum = c->d_ksum;
  a.tape = tape;
  a.tape_base = s0;
  a.nstate = nstate(c);
  hipLaunchKernelGGL(tdhf::step_set_kernel, dim3(1), dim3(1), 0, c->stream, c->d_step, s0);
  HIPCHK(c, hipGetLastError());
  for (int j = 0; j < n; ++j) {
    a.jstep = j;
    a.mode = tdhf::M_S1; a.u = a.w = c->d_u0; a.unext = c->d_sa;
    CHKRET(launch_apply(c, a, rows));
    a.mode = tdhf::M_S2; a.u = a.w = c->d_sa; a.unext = c->d_sb;
    CHKRET(launch_apply(c, a, rows));
    a.mode = tdhf::M_S3; a.u = a.w = c->d_sb; a.unext = c->d_sa;
    CHKRET(launch_apply(c, a, rows));
    a.mode = tdhf::M_S4; a.u = a.w = c->d_sa; a.unext = nullptr;
    CHKRET(launch_apply(c, a, rows));
  }
  return 0;
}

// Reverse sweep over steps [s0, s0+n) whose stage states are in tape (slot 0
// = step s0).  ub/ua/ga/gb rotate as described in tdhf_adjoint.h.
int adj_backward_steps(tdhf_ctx* c, AdjState& s, int s0, int n, zd** ub, zd** ua, zd** ga,
                       zd** gb) {
  const size_t ns = nstate(c);
  const int rows = c->nb * c->nk;
  tdhf::AdjArgs a = adj_base(c);
  for (int it = s0 + n - 1; it >= s0; --it) {
    zd* tp = s.tape + (size_t)(it - s0) * 4 * ns;
    const double fa[4] = {2.0 * s.Aave[it] - s.A[it], s.Aave[it], s.Aave[it], s.A[it]};
    a.ub = *ub;
    a.ua = *ua;
    for (int st = 4; st >= 1; --st) {
      a.mode = st;
      a.y = tp + (size_t)(st - 1) * ns;
      a.a = fa[st - 1];
      a.gb = *ga;
      a.gout = *gb;
      a.abar = s.abar + ((size_t)it * 4 + (st - 1)) * rows;
      CHKRET(launch_adj(c, a));
      std::swap(*ga, *gb);
    }
    std::swap(*ub, *ua);  // the accumulated cotangent becomes ubar of the previous step
  }
  return 0;
}

}  // namespace

extern "C" int tdhf_fwd(tdhf_ctx* c, double* u, int nt, const double* A, double eref,
                        double mem_budget, double* L, double* etot, double* nele, double* jtot) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u || nt < 1 || !A) return fail(c, TDHF_EINVAL, "invalid forward arguments");
  if (c->nb != c->nbocc)
    return fail(c, TDHF_EINVAL, "the adjoint path propagates occupied bands only (NB must equal NBocc)");
  AdjState& s = adj_states()[c];
  adj_free(s);
  s.nt = nt;
  s.A.assign(A, A + nt);
  s.Aave.resize(nt);
  for (int it = 0; it < nt; ++it) s.Aave[it] = 0.5 * (A[it] + (it > 0 ? A[it - 1] : 0.0));
  const size_t ns = nstate(c), sb = ns * sizeof(zd);
  double budget = mem_budget;
  if (!(budget > 0.0)) {
    size_t fr = 0, tot = 0;
    HIPCHK(c, hipMemGetInfo(&fr, &tot));
    budget = 0.25 * (double)fr;
  }
  if (4.0 * nt * sb <= budget) {
    s.full = true;
    s.seg = nt;
  } else {
    // sqrt(nt) spacing; if that does not fit, the spacing of least memory,
    // ceil(nt/S) + 4 S states with S ~ sqrt(nt/4)
    auto need = [&](int S) { return ((double)((nt + S - 1) / S) + 4.0 * S) * (double)sb; };
    int S = (int)std::ceil(std::sqrt((double)nt));
    if (need(S) > budget) S = std::max(1, (int)std::ceil(std::sqrt(nt / 4.0)));
    if (need(S) > budget)
      return fail(c, TDHF_ENOMEM, "memory budget %.3g B too small for %d steps", budget, nt);
    s.full = false;
    s.seg = S;
  }
  s.tape_states = (size_t)4 * s.seg;
  CHKRET(dalloc(c, &s.tape, s.tape_states * ns));
  if (!s.full) {
    s.ckpt_states = (size_t)((nt + s.seg - 1) / s.seg);
    CHKRET(dalloc(c, &s.ckpt, s.ckpt_states * ns));
  }
  CHKRET(ensure_obs(c, (long)nt + 1));
  CHKRET(ensure_field(c, nt));
  HIPCHK(c, hipMemcpyAsync(c->d_A, s.A.data(), nt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_Aave, s.Aave.data(), nt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CHKRET(upload(c, c->d_u0, u));
  if (s.full) {
    CHKRET(adj_forward_steps(c, 0, nt, s.tape));
  } else {
    for (int s0 = 0; s0 < nt; s0 += s.seg) {
      HIPCHK(c, hipMemcpyAsync(s.ckpt + (size_t)(s0 / s.seg) * ns, c->d_u0, sb,
                               hipMemcpyDeviceToDevice, c->stream));
      CHKRET(adj_forward_steps(c, s0, std::min(s.seg, nt - s0), nullptr));
    }
  }
  // observables after the last step (slot nt) and the per-step reduction
  const int rows = c->nb * c->nk;
  hipLaunchKernelGGL(tdhf::step_set_kernel, dim3(1), dim3(1), 0, c->stream, c->d_step, nt);
  HIPCHK(c, hipGetLastError());
  tdhf::ApplyArgs a = base_args(c);
  a.mode = tdhf::M_OBS;
  a.u = a.w = c->d_u0;
  a.Aarr = c->d_A;
  a.Aavearr = c->d_Aave;
  a.step_base = c->d_step;
  a.jstep = -1;
  a.obs = c->d_obs;
  CHKRET(launch_apply(c, a, rows));
  const int nsl = nt + 1;
  hipLaunchKernelGGL(tdhf::obs_reduce_kernel, dim3((nsl + 63) / 64), dim3(64), 0, c->stream,
                     c->d_obs, c->obs_stride, rows, nsl, c->dv, c->vcell, c->d_res);
  HIPCHK(c, hipGetLastError());
  std::vector<double> res((size_t)7 * nsl);
  HIPCHK(c, hipMemcpyAsync(res.data(), c->d_res, res.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  CHKRET(sync(c));
  for (int it = 0; it < nt; ++it) {
    const double* r = &res[(size_t)7 * (it + 1)];
    if (etot) etot[it] = r[0];
    if (nele) nele[it] = r[1];
    if (jtot) jtot[it] = r[2];
  }
  s.jN = res[(size_t)7 * nt + 2];
  if (L) *L = res[(size_t)7 * nt] - eref;
  s.have_fwd = true;
  return download(c, u, c->d_u0);
}

extern "C" int tdhf_bwd(tdhf_ctx* c, double* dLdA, double* dLdu0) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  auto itS = adj_states().find(c);
  if (itS == adj_states().end() || !itS->second.have_fwd)
    return fail(c, TDHF_EINVAL, "tdhf_bwd needs a preceding tdhf_fwd on this context");
  AdjState& s = itS->second;
  if (!dLdA) return fail(c, TDHF_EINVAL, "dLdA is NULL");
  const size_t ns = nstate(c);
  const int rows = c->nb * c->nk;
  const int nt = s.nt;
  CHKRET(dalloc(c, &s.buf, 4 * ns));
  CHKRET(dalloc(c, &s.abar, (size_t)nt * 4 * rows));
  zd *ua = s.buf, *ub = s.buf + ns, *ga = s.buf + 2 * ns, *gb = s.buf + 3 * ns;
  // dE/du at the final state (d_u0 holds it after the forward run)
  tdhf::AdjArgs e = adj_base(c);
  e.mode = tdhf::A_EGRAD;
  e.y = c->d_u0;
  e.a = s.A[nt - 1];
  e.cout = ub;
  e.gout = ga;
  CHKRET(launch_adj(c, e));
  if (s.full) {
    CHKRET(adj_backward_steps(c, s, 0, nt, &ub, &ua, &ga, &gb));
  } else {
    for (int s0 = ((nt - 1) / s.seg) * s.seg; s0 >= 0; s0 -= s.seg) {
      const int n = std::min(s.seg, nt - s0);
      HIPCHK(c, hipMemcpyAsync(c->d_u0, s.ckpt + (size_t)(s0 / s.seg) * ns, ns * sizeof(zd),
                               hipMemcpyDeviceToDevice, c->stream));
      CHKRET(adj_forward_steps(c, s0, n, s.tape));
      CHKRET(adj_backward_steps(c, s, s0, n, &ub, &ua, &ga, &gb));
    }
  }
  std::vector<double> ab((size_t)nt * 4 * rows);
  HIPCHK(c, hipMemcpyAsync(ab.data(), s.abar, ab.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  if (dLdu0) HIPCHK(c, hipMemcpyAsync(dLdu0, ub, ns * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
  CHKRET(sync(c));
  for (int it = 0; it < nt; ++it) dLdA[it] = 0.0;
  for (int it = 0; it < nt; ++it) {
    double st[4];
    for (int k = 0; k < 4; ++k) {
      double acc = 0.0;
      for (int r = 0; r < rows; ++r) acc += ab[((size_t)it * 4 + k) * rows + r];
      st[k] = acc;
    }
    if (it > 0) dLdA[it - 1] += st[0] + 0.5 * (st[1] + st[2]);
    dLdA[it] += st[3] + 0.5 * (st[1] + st[2]);
  }
  dLdA[nt - 1] += -s.jN * c->vcell;  // dEtot/dA at the final state
  return 0;
}

extern "C" int tdhf_adjoint_info(tdhf_ctx* c, int* full_tape, int* segment) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  auto it = adj_states().find(c);
  if (it == adj_states().end() || !it->second.have_fwd)
    return fail(c, TDHF_EINVAL, "no forward run recorded");
  if (full_tape) *full_tape = it->second.full ? 1 : 0;
  if (segment) *segment = it->second.seg;
  return 0;
}

namespace {
void adj_release(const tdhf_ctx* c) {
  auto it = adj_states().find(c);
  if (it != adj_states().end()) {
    adj_free(it->second);
    adj_states().erase(it);
  }
}
}  // namespace
