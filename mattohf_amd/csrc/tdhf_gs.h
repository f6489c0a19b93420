// tdhf_gs.h -- device kernels of the ground-state (Krylov / Rayleigh-Ritz)
// iteration: GS_calculation, KS_ubk2_SVD and Subspace_diag
// (src/GS_calculation.f90:26-57,124-237).
//
// The H applications reuse apply_kernel (M_HPSI).  The small per-k dense
// steps run one workgroup per k-point:
//   orth_kernel : dvcell-weighted modified Gram-Schmidt (two passes) of the
//                 NB*NCG Krylov columns.  It replaces the left singular
//                 vectors of ZGESVD('O') rescaled by 1/sqrt(dvcell)
//                 (:161-166); both are a dvcell-orthonormal basis of the same
//                 span, and the Rayleigh-Ritz step depends on the span only.
//   rr_kernel   : subspace matrix <w_r|H|w_c> dvcell (:173-181 / :217-224),
//                 cyclic complex Jacobi eigensolver standing in for ZHEEV
//                 (ascending eigenvalues), and the rotation into the lowest
//                 NB Ritz vectors (:184-190 / :227-233).
#pragma once
#include "tdhf_fft.h"

struct tdhf_ctx;

namespace tdhf {

struct GsWork {
  zd* W = nullptr;   // ncg blocks of (NG,NB,NK)
  zd* HW = nullptr;  // ncg blocks
  double* eps = nullptr;
  int* nonconv = nullptr;  // set by rr_kernel when its eigensolver does not converge
  int ncg = 0;
};

// column c of the Krylov basis at k-point ik
__device__ __forceinline__ zd* kcol(zd* base, size_t blk, int ng, int nb, int ik, int c) {
  return base + (size_t)(c / nb) * blk + (size_t)ng * ((c % nb) + (size_t)nb * ik);
}

__device__ __forceinline__ double bsum(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int h = blockDim.x >> 1; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// blockDim.x == ng (power of two <= 128)
__global__ void orth_kernel(zd* W, size_t blk, int ng, int nb, int m, double dv) {
  __shared__ double red[2][128];
  const int ik = blockIdx.x, x = threadIdx.x;
  for (int pass = 0; pass < 2; ++pass)
    for (int c = 0; c < m; ++c) {
      zd* xc = kcol(W, blk, ng, nb, ik, c);
      zd xv = xc[x];
      for (int b = 0; b < c; ++b) {
        const zd yb = kcol(W, blk, ng, nb, ik, b)[x];
        double re = bsum(yb.x * xv.x + yb.y * xv.y, red[0]) * dv;
        double im = bsum(yb.x * xv.y - yb.y * xv.x, red[1]) * dv;
        xv = mk(xv.x - (re * yb.x - im * yb.y), xv.y - (re * yb.y + im * yb.x));
      }
      const double s = sqrt(bsum(xv.x * xv.x + xv.y * xv.y, red[0]) * dv);
      xc[x] = mk(xv.x / s, xv.y / s);
      __syncthreads();
    }
}

// Rayleigh-Ritz at one k-point per workgroup.  m basis columns in W (block
// stride blk), their H images in HW.  Writes the lowest nb Ritz vectors to
// uout (NG,NB,NK) and the Ritz values to eig (nb per k).  If the cyclic
// Jacobi iteration has not converged after 100 sweeps, *nonconv is set to
// 1 + ik (the caller reports it).  Dynamic LDS: 2*m*m complex + m ints.
__global__ void __launch_bounds__(256) rr_kernel(const zd* W, const zd* HW, size_t blk, int ng,
                                                 int nb, int m, double dv, zd* uout, double* eig,
                                                 int* nonconv) {
  extern __shared__ zd sh[];
  zd* A = sh;
  zd* V = sh + (size_t)m * m;
  int* idx = reinterpret_cast<int*>(V + (size_t)m * m);
  __shared__ double red[256];
  __shared__ double rot[6];
  const int ik = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  zd* Wm = const_cast<zd*>(W);
  zd* HWm = const_cast<zd*>(HW);
  // subspace matrix, lower triangle and diagonal
  for (int e = t; e < m * m; e += nt) {
    const int r = e % m, cc = e / m;
    if (cc > r) continue;
    const zd* xr = kcol(Wm, blk, ng, nb, ik, r);
    const zd* hc = kcol(HWm, blk, ng, nb, ik, cc);
    double re = 0.0, im = 0.0;
    for (int x = 0; x < ng; ++x) {
      const zd a = xr[x], b = hc[x];
      re += a.x * b.x + a.y * b.y;
      im += a.x * b.y - a.y * b.x;
    }
    if (cc == r)
      A[r + m * cc] = mk(re * dv, 0.0);
    else
      A[r + m * cc] = mk(re * dv, im * dv);
  }
  for (int e = t; e < m * m; e += nt) V[e] = mk((e % m) == (e / m) ? 1.0 : 0.0, 0.0);
  __syncthreads();
  for (int e = t; e < m * m; e += nt) {
    const int r = e % m, cc = e / m;
    if (cc > r) A[r + m * cc] = mk(A[cc + m * r].x, -A[cc + m * r].y);
  }
  __syncthreads();
  // cyclic Jacobi (same rotation as the oracle's orc_heev)
  bool converged = false;
  for (int sweep = 0; sweep <= 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int e = t; e < m * m; e += nt) {
      const double v = zabs2(A[e]);
      tot += v;
      if ((e % m) != (e / m)) off += v;
    }
    off = bsum(off, red);
    tot = bsum(tot, red);
    if (off <= 1e-32 * tot) {
      converged = true;
      break;
    }
    if (sweep == 100) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        if (t == 0) {
          const zd apq = A[p + m * q];
          const double r = sqrt(zabs2(apq));
          if (r == 0.0) {
            rot[0] = 0.0;
          } else {
            const double app = A[p + m * p].x, aqq = A[q + m * q].x;
            const double tau = (aqq - app) / (2.0 * r);
            const double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            const double cs = 1.0 / sqrt(1.0 + tt * tt);
            rot[0] = 1.0;
            rot[1] = cs;
            rot[2] = tt * cs;
            rot[3] = apq.x / r;
            rot[4] = apq.y / r;
          }
        }
        __syncthreads();
        if (rot[0] != 0.0) {
          const double cs = rot[1], sn = rot[2];
          const zd ph = mk(rot[3], rot[4]);
          const zd eph = mk(rot[3], -rot[4]);
          for (int i = t; i < m; i += nt) {
            const zd aip = A[i + m * p], aiq = A[i + m * q];
            const zd ea = zmul(eph, aiq);
            A[i + m * p] = mk(cs * aip.x - sn * ea.x, cs * aip.y - sn * ea.y);
            A[i + m * q] = mk(sn * aip.x + cs * ea.x, sn * aip.y + cs * ea.y);
            const zd vip = V[i + m * p], viq = V[i + m * q];
            const zd ev = zmul(eph, viq);
            V[i + m * p] = mk(cs * vip.x - sn * ev.x, cs * vip.y - sn * ev.y);
            V[i + m * q] = mk(sn * vip.x + cs * ev.x, sn * vip.y + cs * ev.y);
          }
          __syncthreads();
          for (int j = t; j < m; j += nt) {
            const zd apj = A[p + m * j], aqj = A[q + m * j];
            const zd pa = zmul(ph, aqj);
            A[p + m * j] = mk(cs * apj.x - sn * pa.x, cs * apj.y - sn * pa.y);
            A[q + m * j] = mk(sn * apj.x + cs * pa.x, sn * apj.y + cs * pa.y);
          }
          __syncthreads();
          if (t == 0) {
            A[p + m * q] = mk(0.0, 0.0);
            A[q + m * p] = mk(0.0, 0.0);
            A[p + m * p].y = 0.0;
            A[q + m * q].y = 0.0;
          }
        }
        __syncthreads();
      }
  }
  if (t == 0 && !converged && nonconv) *nonconv = 1 + ik;
  if (t == 0) {
    for (int i = 0; i < m; ++i) idx[i] = i;
    for (int i = 0; i < m; ++i) {
      int b = i;
      for (int j = i + 1; j < m; ++j)
        if (A[idx[j] + m * idx[j]].x < A[idx[b] + m * idx[b]].x) b = j;
      const int tmp = idx[i];
      idx[i] = idx[b];
      idx[b] = tmp;
    }
  }
  __syncthreads();
  if (eig && t < nb) eig[(size_t)nb * ik + t] = A[idx[t] + m * idx[t]].x;
  // rotation: u(:,ib,ik) = sum_c w_c * V(c, idx(ib))
  for (int e = t; e < ng * nb; e += nt) {
    const int x = e % ng, ib = e / ng;
    const int col = idx[ib];
    zd s = mk(0.0, 0.0);
    for (int cc = 0; cc < m; ++cc) {
      const zd w = kcol(Wm, blk, ng, nb, ik, cc)[x];
      s = zadd(s, zmul(w, V[cc + m * col]));
    }
    uout[(size_t)ng * (ib + (size_t)nb * ik) + x] = s;
  }
}

// ubk_epsbk / ubk_epserrorbk (src/Functions.f90:321-358) from u and H u:
// one workgroup of ng threads per orbital (ib,ik).
__global__ void epsbk_kernel(const zd* u, const zd* hu, int ng, double dv, double* eps,
                             double* err) {
  __shared__ double red[128];
  const int row = blockIdx.x, x = threadIdx.x;
  const zd a = u[(size_t)ng * row + x], h = hu[(size_t)ng * row + x];
  const double e = bsum(a.x * h.x + a.y * h.y, red) * dv;
  const zd r = mk(h.x - e * a.x, h.y - e * a.y);
  const double s = bsum(sqrt(r.x * r.x + r.y * r.y), red);
  if (x == 0) {
    if (eps) eps[row] = e;
    if (err) err[row] = sqrt(s * dv);
  }
}

int ctx_apply(tdhf_ctx* c, const zd* u, const zd* w, int nbw, double A, zd* out);
int ctx_fail(tdhf_ctx* c, int code, const char* msg);
void gs_free(GsWork& g);

}  // namespace tdhf
