// tdhf_post_impl.h -- C-ABI of the Taylor propagators, the ground state with
// diagnostics, Make_DoS and the spectra (included at the end of tdhf.hip).
#pragma once

namespace {

int te_launch_update(tdhf_ctx* c, const zd* h, const zd* h2, double f, zd* w, zd* y) {
  const size_t n = nstate(c);
  hipLaunchKernelGGL(tdhf::te_update_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     c->stream, n, h, h2, f, w, y);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// y += sum_{k=1..4} w_k, w_0 = x, w_k = -i dt H w_{k-1} / k with H = H[ua]
// (or (H[ua] + H[ub])/2), as the order loops of src/RT_calculation.f90:71-151.
int te_series(tdhf_ctx* c, const zd* ua, const zd* ub, const zd* x, double Aave, zd* y) {
  const size_t bytes = nstate(c) * sizeof(zd);
  HIPCHK(c, hipMemcpyAsync(c->d_w, x, bytes, hipMemcpyDeviceToDevice, c->stream));
  for (int k = 1; k <= 4; ++k) {
    CHKRET(tdhf::ctx_apply(c, ua, c->d_w, c->nb, Aave, c->d_out));
    if (ub) CHKRET(tdhf::ctx_apply(c, ub, c->d_w, c->nb, Aave, c->d_ksum));
    CHKRET(te_launch_update(c, c->d_out, ub ? c->d_ksum : nullptr, c->dt / (double)k, c->d_w, y));
  }
  return 0;
}

int post_device_check() {
  tdhf_ctx tmp;
  return check_device(&tmp);
}

}  // namespace

extern "C" int tdhf_te_step(tdhf_ctx* c, double* u, double Aave, int variant) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u) return fail(c, TDHF_EINVAL, "NULL array");
  if (variant < 0 || variant > 2) return fail(c, TDHF_EINVAL, "unknown Taylor variant %d", variant);
  const size_t bytes = nstate(c) * sizeof(zd);
  CHKRET(upload(c, c->d_u0, u));
  // ubk0 in sa; the state accumulates in u0
  HIPCHK(c, hipMemcpyAsync(c->d_sa, c->d_u0, bytes, hipMemcpyDeviceToDevice, c->stream));
  if (variant == 0) {
    CHKRET(te_series(c, c->d_sa, nullptr, c->d_sa, Aave, c->d_u0));
  } else {
    // predictor ubk1 in sb
    HIPCHK(c, hipMemcpyAsync(c->d_sb, c->d_sa, bytes, hipMemcpyDeviceToDevice, c->stream));
    CHKRET(te_series(c, c->d_sa, nullptr, c->d_sa, Aave, c->d_sb));
    if (variant == 1) {
      CHKRET(te_series(c, c->d_sa, c->d_sb, c->d_sa, Aave, c->d_u0));
    } else {
      const size_t n = nstate(c);
      hipLaunchKernelGGL(tdhf::te_mid_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                         c->stream, n, c->d_sb, c->d_sa);
      HIPCHK(c, hipGetLastError());
      CHKRET(te_series(c, c->d_sb, nullptr, c->d_sa, Aave, c->d_u0));
    }
  }
  return download(c, u, c->d_u0);
}

extern "C" int tdhf_gs_run_diag(tdhf_ctx* c, double* u, int nscf, int ncg, double* etot_hist,
                                double* epsbk, double* epserr, double* diag) {
  if (!c) return fail(nullptr, TDHF_EINVAL, "ctx is NULL");
  if (!u || nscf < 0 || ncg < 1 || (nscf > 0 && !diag))
    return fail(c, TDHF_EINVAL, "invalid GS arguments");
  const int m = c->nb * ncg;
  if (m > 64) return fail(c, TDHF_EINVAL, "NB*NCG=%d exceeds 64", m);
  if (m > c->ng) return fail(c, TDHF_EINVAL, "NB*NCG=%d exceeds NG=%d", m, c->ng);
  CHKRET(gs_alloc(c, ncg));
  const int rows = c->nb * c->nk;
  double *dmix = nullptr, *ddiag = nullptr;
  CHKRET(dalloc(c, &dmix, c->ng));
  int rc = dalloc(c, &ddiag, (size_t)5 * (nscf > 0 ? nscf : 1));
  if (rc) {
    dfree(dmix);
    return rc;
  }
  auto done = [&](int r) {
    dfree(dmix);
    dfree(ddiag);
    return r;
  };
  auto step_diag = [&](int mode, int it) -> int {
    CHKRET(tdhf::ctx_apply(c, c->d_u0, c->d_u0, c->nb, 0.0, c->gs.HW));
    hipLaunchKernelGGL(tdhf::epsbk_kernel, dim3(rows), dim3(c->ng), 0, c->stream, c->d_u0,
                       c->gs.HW, c->ng, c->dv, c->gs.eps, c->gs.eps + rows);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(tdhf::gs_diag_kernel, dim3(1), dim3(256), 0, c->stream, c->d_u0, c->ng,
                       c->nb, c->nk, c->nbocc, c->occ, c->gs.eps, c->gs.eps + rows, dmix,
                       ddiag + (size_t)5 * (it > 0 ? it - 1 : 0), mode);
    HIPCHK(c, hipGetLastError());
    return 0;
  };
  if ((rc = gs_flag_clear(c)) || (rc = upload(c, c->d_u0, u))) return done(rc);
  double r[7];
  if ((rc = observables_of(c, c->d_u0, 0.0, r))) return done(rc);
  if (etot_hist) etot_hist[0] = r[0];
  if ((rc = step_diag(0, 0))) return done(rc);
  for (int it = 1; it <= nscf; ++it) {
    if ((rc = launch_gs_orth(c, c->d_u0)) || (rc = gs_krylov(c, ncg)) ||
        (rc = launch_gs_orth(c, c->d_u0)) || (rc = gs_subspace(c)))
      return done(rc);
    if ((rc = observables_of(c, c->d_u0, 0.0, r))) return done(rc);
    if (etot_hist) etot_hist[it] = r[0];
    if (!std::isfinite(r[0]))
      return done(fail(c, TDHF_ENUMERIC, "GS energy is not finite at iteration %d", it));
    if ((rc = step_diag(1, it))) return done(rc);
  }
  std::vector<double> h((size_t)2 * rows);
  hipError_t e = hipMemcpyAsync(h.data(), c->gs.eps, h.size() * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && nscf > 0)
    e = hipMemcpyAsync(diag, ddiag, (size_t)5 * nscf * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream);
  if (e != hipSuccess) return done(fail(c, TDHF_EDEVICE, "copy: %s", hipGetErrorString(e)));
  if ((rc = sync(c)) || (rc = gs_flag_check(c))) return done(rc);
  if (epsbk) std::memcpy(epsbk, h.data(), rows * sizeof(double));
  if (epserr) std::memcpy(epserr, h.data() + rows, rows * sizeof(double));
  return done(download(c, u, c->d_u0));
}

extern "C" int tdhf_make_dos(int nb, int nk, int nbocc, const double* epsbk, int ndos,
                             double* epsdos, double* dos, double* dosocc) {
  if (nb < 1 || nk < 1 || nbocc < 0 || nbocc > nb || ndos < 2 || !epsbk || !epsdos || !dos ||
      !dosocc)
    return fail(nullptr, TDHF_EINVAL, "invalid DoS arguments");
  CHKRET(post_device_check());
  const size_t ne = (size_t)nb * nk;
  double* d = nullptr;
  if (hipMalloc((void**)&d, (ne + 3 * (size_t)ndos) * sizeof(double)) != hipSuccess)
    return fail(nullptr, TDHF_ENOMEM, "hipMalloc failed");
  hipError_t e = hipMemcpy(d, epsbk, ne * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(tdhf::dos_kernel, dim3((unsigned)((ndos + 255) / 256)), dim3(256), 0, 0,
                       nb, nk, nbocc, d, ndos, d + ne, d + ne + ndos, d + ne + 2 * ndos);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(epsdos, d + ne, ndos * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(dos, d + ne + ndos, ndos * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(dosocc, d + ne + 2 * ndos, ndos * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(nullptr, TDHF_EDEVICE, "DoS: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int tdhf_fourier(int nt, double dt, const double* time, const double* J, double j0,
                            const double* E, const double* A, int nomega, double omegam,
                            int window, double tpulse, int nenvelope, double* out) {
  if (nt < 1 || nomega < 1 || !time || !J || !E || !A || !out || (window != 0 && window != 1) ||
      (window == 1 && !(tpulse > 0.0)) || nenvelope < 0)
    return fail(nullptr, TDHF_EINVAL, "invalid Fourier arguments");
  CHKRET(post_device_check());
  double* d = nullptr;
  const size_t n = (size_t)nt;
  if (hipMalloc((void**)&d, (4 * n + 7 * (size_t)nomega) * sizeof(double)) != hipSuccess)
    return fail(nullptr, TDHF_ENOMEM, "hipMalloc failed");
  hipError_t e = hipMemcpy(d, time, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + n, J, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + 2 * n, E, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + 3 * n, A, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(tdhf::fourier_kernel, dim3(nomega), dim3(256), 0, 0, nt, dt, d, d + n, j0,
                       d + 2 * n, d + 3 * n, nomega, omegam, window, tpulse, nenvelope, d + 4 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpy(out, d + 4 * n, 7 * (size_t)nomega * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(nullptr, TDHF_EDEVICE, "Fourier: %s", hipGetErrorString(e));
  return 0;
}
