// tdhf_gs_impl.h -- host side of the ground-state iteration (included by
// tdhf.hip after the context helpers).  Mirrors GS_calculation
// (src/GS_calculation.f90:26-57): per iteration Gram_Schmidt, KS_ubk2_SVD,
// Gram_Schmidt, Subspace_diag, then the total energy at A=0.
#pragma once

namespace tdhf {

void gs_free(GsWork& g) {
  if (g.W) (void)hipFree(g.W);
  if (g.HW) (void)hipFree(g.HW);
  if (g.eps) (void)hipFree(g.eps);
  if (g.nonconv) (void)hipFree(g.nonconv);
  g.W = g.HW = nullptr;
  g.eps = nullptr;
  g.nonconv = nullptr;
  g.ncg = 0;
}

}  // namespace tdhf

namespace {

int gs_alloc(tdhf_ctx* c, int ncg) {
  if (c->gs.ncg >= ncg) return 0;
  tdhf::gs_free(c->gs);
  const size_t n = nstate(c) * (size_t)ncg;
  CHKRET(dalloc(c, &c->gs.W, n));
  CHKRET(dalloc(c, &c->gs.HW, n));
  CHKRET(dalloc(c, &c->gs.eps, (size_t)c->nb * c->nk * 2));
  CHKRET(dalloc(c, &c->gs.nonconv, 1));
  c->gs.ncg = ncg;
  return 0;
}

// Clears the eigensolver's non-convergence flag (before a GS run).
int gs_flag_clear(tdhf_ctx* c) {
  HIPCHK(c, hipMemsetAsync(c->gs.nonconv, 0, sizeof(int), c->stream));
  return 0;
}

// Fails if a Rayleigh-Ritz eigensolve since the last clear did not converge.
int gs_flag_check(tdhf_ctx* c) {
  int f = 0;
  HIPCHK(c, hipMemcpyAsync(&f, c->gs.nonconv, sizeof f, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (f)
    return fail(c, TDHF_ENUMERIC,
                "subspace eigensolver (cyclic Jacobi) did not converge in 100 sweeps at k-point %d",
                f - 1);
  return 0;
}

int launch_gs_orth(tdhf_ctx* c, zd* u) {
  hipLaunchKernelGGL(tdhf::gram_schmidt_kernel, dim3(c->nk), dim3(c->ng), 0, c->stream, u, c->ng,
                     c->nb, c->vcell);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_rr(tdhf_ctx* c, const zd* W, const zd* HW, size_t blk, int m, zd* uout) {
  const size_t lds = (size_t)2 * m * m * sizeof(zd) + (size_t)m * sizeof(int) + 16;
  if (lds > 48 * 1024)
    HIPCHK(c, hipFuncSetAttribute((const void*)tdhf::rr_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(tdhf::rr_kernel, dim3(c->nk), dim3(256), lds, c->stream, W, HW, blk, c->ng,
                     c->nb, m, c->dv, uout, (double*)nullptr, c->gs.nonconv);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// KS_ubk2_SVD (src/GS_calculation.f90:124-194) on the state in d_u0.
int gs_krylov(tdhf_ctx* c, int ncg) {
  const size_t blk = nstate(c);
  zd* W = c->gs.W;
  zd* HW = c->gs.HW;
  HIPCHK(c, hipMemcpyAsync(W, c->d_u0, blk * sizeof(zd), hipMemcpyDeviceToDevice, c->stream));
  for (int icg = 1; icg < ncg; ++icg)
    CHKRET(tdhf::ctx_apply(c, c->d_u0, W + (icg - 1) * blk, c->nb, 0.0, W + icg * blk));
  const int m = c->nb * ncg;
  hipLaunchKernelGGL(tdhf::orth_kernel, dim3(c->nk), dim3(c->ng), 0, c->stream, W, blk, c->ng,
                     c->nb, m, c->dv);
  HIPCHK(c, hipGetLastError());
  for (int icg = 0; icg < ncg; ++icg)
    CHKRET(tdhf::ctx_apply(c, c->d_u0, W + icg * blk, c->nb, 0.0, HW + icg * blk));
  return launch_rr(c, W, HW, blk, m, c->d_u0);
}

// Subspace_diag (src/GS_calculation.f90:197-237) on the state in d_u0.
int gs_subspace(tdhf_ctx* c) {
  const size_t blk = nstate(c);
  CHKRET(tdhf::ctx_apply(c, c->d_u0, c->d_u0, c->nb, 0.0, c->gs.HW));
  HIPCHK(c, hipMemcpyAsync(c->gs.W, c->d_u0, blk * sizeof(zd), hipMemcpyDeviceToDevice,
                           c->stream));
  return launch_rr(c, c->gs.W, c->gs.HW, blk, c->nb, c->d_u0);
}

}  // namespace

extern "C" int tdhf_gs_run(tdhf_ctx* c, double* u, int nscf, int ncg, double tol,
                           double* etot_hist, double* epsbk, double* epserr, int* niter) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u || nscf < 0 || ncg < 1) return fail(c, TDHF_EINVAL, "invalid GS arguments");
  const int m = c->nb * ncg;
  if (m > 64) return fail(c, TDHF_EINVAL, "NB*NCG=%d exceeds 64", m);
  if (m > c->ng) return fail(c, TDHF_EINVAL, "NB*NCG=%d exceeds NG=%d", m, c->ng);
  CHKRET(gs_alloc(c, ncg));
  CHKRET(gs_flag_clear(c));
  CHKRET(upload(c, c->d_u0, u));
  double r[7];
  CHKRET(observables_of(c, c->d_u0, 0.0, r));
  double eprev = r[0];
  if (etot_hist) etot_hist[0] = eprev;
  int it = 0;
  for (it = 1; it <= nscf; ++it) {
    CHKRET(launch_gs_orth(c, c->d_u0));
    CHKRET(gs_krylov(c, ncg));
    CHKRET(launch_gs_orth(c, c->d_u0));
    CHKRET(gs_subspace(c));
    CHKRET(observables_of(c, c->d_u0, 0.0, r));
    if (etot_hist) etot_hist[it] = r[0];
    if (!std::isfinite(r[0])) return fail(c, TDHF_ENUMERIC, "GS energy is not finite at iteration %d", it);
    const bool conv = tol > 0.0 && std::fabs(r[0] - eprev) < tol;
    eprev = r[0];
    if (conv) break;
  }
  CHKRET(gs_flag_check(c));
  if (niter) *niter = it > nscf ? nscf : it;
  if (epsbk || epserr) {
    CHKRET(tdhf::ctx_apply(c, c->d_u0, c->d_u0, c->nb, 0.0, c->gs.HW));
    const int rows = c->nb * c->nk;
    hipLaunchKernelGGL(tdhf::epsbk_kernel, dim3(rows), dim3(c->ng), 0, c->stream, c->d_u0,
                       c->gs.HW, c->ng, c->dv, c->gs.eps, c->gs.eps + rows);
    HIPCHK(c, hipGetLastError());
    std::vector<double> h((size_t)2 * rows);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->gs.eps, h.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    CHKRET(sync(c));
    if (epsbk) std::memcpy(epsbk, h.data(), rows * sizeof(double));
    if (epserr) std::memcpy(epserr, h.data() + rows, rows * sizeof(double));
  }
  return download(c, u, c->d_u0);
}
