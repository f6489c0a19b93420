// tdhf_post.h -- device kernels around the propagation: the Taylor
// propagators' vector updates, the ground-state diagnostics of
// GS_calculation, Make_DoS and the spectra of FourierTr / WindowFourierTr.
#pragma once
#include <cfloat>
#include <cmath>

#include "tdhf_kernels.h"

namespace tdhf {

// One order of the Taylor series of RT_TE_ubk / RT_TEPC_ubk / RT_TEPC2_ubk
// (src/RT_calculation.f90:71-151): w = -i f (h or (h + h2)/2), y += w.
__global__ void te_update_kernel(size_t n, const zd* h, const zd* h2, double f, zd* w, zd* y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  zd hv = h[i];
  if (h2) {
    const zd b = h2[i];
    hv = mk(0.5 * (hv.x + b.x), 0.5 * (hv.y + b.y));
  }
  const zd wv = mk(f * hv.y, -f * hv.x);  // -i f h
  w[i] = wv;
  const zd yv = y[i];
  y[i] = mk(yv.x + wv.x, yv.y + wv.y);
}

// a = (a + b)/2 (RT_TEPC2_ubk's midpoint state)
__global__ void te_mid_kernel(size_t n, zd* a, const zd* b) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const zd x = a[i], y = b[i];
  a[i] = mk(0.5 * (x.x + y.x), 0.5 * (x.y + y.y));
}

// Fixed-order sum over a 256-thread block.
__device__ __forceinline__ double sum256(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double min256(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) red[t] = fmin(red[t], red[t + h]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Per-iteration diagnostics of GS_calculation (src/GS_calculation.f90:44-51),
// one 256-thread block: the density dns = sum occ |u|^2 (ubk_dns, src/
// Functions.f90:55-71) is mixed into dmix with weight 1/2 (mode 1; mode 0 only
// sets dmix = dns), and diag[0..5) = mean |dmix_new - dmix_old|, mean and max
// epserrorbk, fundamental gap, smallest same-k gap (DBL_MAX without empty
// bands, as Fortran's minval of an empty set).
__global__ void __launch_bounds__(256) gs_diag_kernel(const zd* u, int ng, int nb, int nk,
                                                      int nbocc, double occ, const double* eps,
                                                      const double* err, double* dmix,
                                                      double* diag, int mode) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double dd = 0.0;
  for (int x = t; x < ng; x += 256) {
    double d = 0.0;
    for (int ik = 0; ik < nk; ++ik)
      for (int ib = 0; ib < nbocc; ++ib) d += occ * zabs2(u[(size_t)ng * (ib + nb * ik) + x]);
    if (mode == 0) {
      dmix[x] = d;
    } else {
      const double nx = 0.5 * d + 0.5 * dmix[x];
      dd += fabs(nx - dmix[x]);
      dmix[x] = nx;
    }
  }
  if (mode == 0) return;
  const int rows = nb * nk;
  double se = 0.0, me = -DBL_MAX, vtop = -DBL_MAX, cbot = DBL_MAX, opt = DBL_MAX;
  for (int r = t; r < rows; r += 256) {
    const int ib = r % nb, ik = r / nb;
    se += err[r];
    me = fmax(me, err[r]);
    if (ib < nbocc) vtop = fmax(vtop, eps[r]);
    else cbot = fmin(cbot, eps[r]);
    if (nbocc > 0 && ib == nbocc) opt = fmin(opt, eps[r] - eps[nbocc - 1 + nb * ik]);
  }
  const double sdd = sum256(dd, red), sse = sum256(se, red);
  const double mme = -min256(-me, red), mvt = -min256(-vtop, red);
  const double mcb = min256(cbot, red), mop = min256(opt, red);
  if (t == 0) {
    diag[0] = sdd / (double)ng;
    diag[1] = sse / (double)rows;
    diag[2] = mme;
    diag[3] = nb > nbocc ? mcb - mvt : DBL_MAX;
    diag[4] = mop;
  }
}

// Make_DoS (src/GS_calculation.f90:239-265) on the energy grid of
// GS_calculation (:93-96); one thread per grid point, every block first
// finds min/max of eps (fixed order).
__global__ void __launch_bounds__(256) dos_kernel(int nb, int nk, int nbocc, const double* eps,
                                                  int ndos, double* epsdos, double* dos,
                                                  double* dosocc) {
  __shared__ double red[256];
  double lo = DBL_MAX, hi = -DBL_MAX;
  for (int r = threadIdx.x; r < nb * nk; r += 256) {
    lo = fmin(lo, eps[r]);
    hi = fmax(hi, eps[r]);
  }
  const double emin = min256(lo, red), emax = -min256(-hi, red);
  const double a = emin - 0.1 * (emax - emin), b = emax + 0.1 * (emax - emin);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ndos) return;
  const double e = a + (b - a) * (double)i / (double)ndos;
  const double e1 = a + (b - a) * 1.0 / (double)ndos;
  const double ew = 5.0 * (e1 - a);
  const double occ = 2.0 / (double)nk, occmax = nbocc > 0 ? occ : 0.0;
  double d = 0.0, q = 0.0;
  for (int ik = 0; ik < nk; ++ik)
    for (int ib = 0; ib < nb; ++ib) {
      const double x = eps[ib + nb * ik] - e;
      const double v = exp(-x * x / (ew * ew));
      d += occmax * v;
      q += (ib < nbocc ? occ : 0.0) * v;
    }
  const double norm = ew * sqrt(M_PI);
  epsdos[i] = e;
  dos[i] = d / norm;
  dosocc[i] = q / norm;
}

// FourierTr (window 0) / WindowFourierTr (window 1), src/RT_calculation.f90
// :246-351: block iw computes omega = (iw+1)/nomega * omegam and the filtered
// transforms JF, EF, AF = sum_t e^{i omega t} filter (J - j0 | E | A) dt.
// out[7 iw + ...] = omega, Re/Im JF, Re/Im EF, Re/Im AF.
__global__ void __launch_bounds__(256) fourier_kernel(int nt, double dt, const double* time,
                                                      const double* J, double j0,
                                                      const double* E, const double* A,
                                                      int nomega, double omegam, int window,
                                                      double tpulse, int nenvelope, double* out) {
  __shared__ double red[256];
  const int iw = blockIdx.x;
  const double w = (double)(iw + 1) / (double)nomega * omegam;
  const double tn = time[nt - 1];
  int ntp = 0;
  if (window == 1) ntp = (int)fmin(tpulse / dt, (double)nt);
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int it = threadIdx.x; it < nt; it += 256) {
    double f;
    if (window == 0) {
      const double r = time[it] / tn;
      f = 1.0 - 3.0 * r * r + 2.0 * r * r * r;
    } else if (it < ntp) {
      const double sn = sin(M_PI * time[it] / tpulse);
      f = 1.0;
      for (int k = 0; k < nenvelope; ++k) f *= sn;
    } else {
      f = 0.0;
    }
    double sn, cs;
    sincos(w * time[it], &sn, &cs);
    const double jt = f * (J[it] - j0), et = f * E[it], at = f * A[it];
    s[0] += cs * jt;
    s[1] += sn * jt;
    s[2] += cs * et;
    s[3] += sn * et;
    s[4] += cs * at;
    s[5] += sn * at;
  }
  double r[6];
  for (int k = 0; k < 6; ++k) r[k] = sum256(s[k], red);
  if (threadIdx.x == 0) {
    double* o = out + (size_t)7 * iw;
    o[0] = w;
    for (int k = 0; k < 6; ++k) o[1 + k] = r[k] * dt;
  }
}

}  // namespace tdhf
