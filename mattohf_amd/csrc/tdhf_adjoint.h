// tdhf_adjoint.h -- backward (adjoint) RK4/TDHF kernels (no counterpart in
// the reference; they differentiate RT_RK4_ubk, src/RT_calculation.f90
// :154-173, and ubkwbk_hwbk, src/Functions.f90:187-209).
//
// Cotangents are Euclidean: dL = Re sum conj(c) du.  For f(y) = -i H_a[y] y
// and a cotangent kbar of f, gbar = i kbar is the cotangent of g = H y and
//   J^T gbar = T_a gbar + (vion + vH) gbar + 2 occ (K s) y
//            + sum_{q,j} y_qj C_kq(gbar_ki conj y_qj + y_ki conj gbar_qj)
//            + sum_{q,j} C'_kq(y_ki conj y_qj) gbar_qj
//            + sum_{q,j} D_kq [FFT(y_ki conj gbar_qj)](NG/2) (-1)^x y_qj
// with s = Re sum conj(gbar) y, K the Hartree operator, C_kq the exchange
// convolution (table vt_kq) and C'_kq the same with the table at reversed G,
// vt'_kq(g) = vt_qk(-g), so that conj(C_kq f) = C'_qk(conj f).  Because the
// grid holds -NG/2 but not +NG/2, C' differs from C only at g = NG/2:
// C'_kq f = C_kq f + D_kq fhat(NG/2) (-1)^x, D_kq = vt_qk(NG/2) - vt_kq(NG/2).
// So each pair costs two transforms; the Nyquist terms are O(NG).  The field
// derivative of f is -i P_a y (P_a = p + k + a), giving
//   dL/da = Re sum_G (G + k + a) conj(ghat(G)) yhat(G) / NG.
#pragma once
#include "tdhf_kernels.h"

namespace tdhf {

enum AdjMode : int { A_EGRAD = 0, A_S1 = 1, A_S2 = 2, A_S3 = 3, A_S4 = 4 };

struct AdjArgs {
  int nk, nb, nbocc;
  double bx, dv, occ, dt;
  // field of the stage, from the per-step arrays of each trajectory: stage 1
  // 2 Aave[step] - A[step] (= A(step-1)), stages 2, 3 Aave[step], stage 4 and
  // A_EGRAD A[step]
  const double* Aarr;
  const double* Aavearr;
  int step;
  const double* kx;
  const double* vion;
  const double* vtl;
  const double* kc;         // circulant kernels K_H, K_2, K_1 (see apply_kernel)
  const zd* tw;
  const zd* y;              // stage state (all rows)
  const zd* gb;             // cotangent gbar of H y (all rows); A_EGRAD: unused
  const zd* ub;             // cotangent of the state after the step (read only)
  zd* ua;                   // cotangent accumulator of the state before the step
  zd* gout;                 // next gbar (A_EGRAD: gbar of stage 4 of the last step)
  zd* cout;                 // A_EGRAD: dE/du (Euclidean cotangent)
  double* abar;             // per-row field derivative of this stage, or nullptr
  int mode;
  // graph-replayable form: step = *step_base - jstep, y = stage block of that
  // step in the tape (slot step % tape_seg, blocks of blk states), abar the
  // matching slot of the per-row derivatives (4 * rows per step)
  const int* step_base;
  int jstep;
  const zd* tape;
  int tape_seg;
  size_t blk;
  // uniform batch (grid y = trajectory): offsets of trajectory t are
  // t*tstride (states, zd), t*fstride (fields), t*abar_tstride (abar)
  size_t tstride;
  int fstride;
  size_t abar_tstride;
  // exchange part of J^T gbar, density and s from the pair tiles
  // (adjtile_kernel; adjoint_kernel<NG, true>): nblk partial sums per row,
  // trajectory t at t * xstride, and per trajectory nblk density partials
  // followed by nblk partials of s, at t * dstride
  const zd* xpart;
  const double* dpart;
  int nblk;
  size_t xstride, dstride;
};

// Per-group LDS slot of adjoint_kernel (complex elements): the transpose
// tile during the pair loop; afterwards the slot area holds the wave partials
// of the reduction (4 NG doubles per wave), the circulant-product partials and
// the field-derivative staging.  At least 2 NG elements, 8 mod 16 as in
// apply_slot.
template <int NG>
__host__ __device__ constexpr int adjoint_slot() {
  constexpr int need = GroupFFT<NG>::TSZ >= 2 * NG ? GroupFFT<NG>::TSZ : 2 * NG;
  return need + ((8 - need % 16) + 16) % 16;
}

// Slot of the launches that take the exchange part from the pair tiles
// (adjoint_kernel<NG, true>: no transposes, no reduction partials): just the
// circulant-product partials and the staging at the smallest launch, so that
// more workgroups share a CU.
template <int NG>
__host__ __device__ constexpr int adjoint_epi_slot() {
  constexpr int nwmin = NG > 64 ? NG / 64 : 1, gpw = 64 / GroupFFT<NG>::L;
  constexpr int need = (8 * 64 * nwmin + NG + 2 * gpw * nwmin - 1) / (2 * gpw * nwmin);
  return need + ((8 - need % 16) + 16) % 16;
}

// LDS of adjoint_kernel<NG> with nw waves (bytes): the groups' tiles (pair
// transposes, then reduction partials, circulant-product partials and the
// field derivative staging), y, gbar, the exchange sum and K_1 (complex NG
// each), rho, s, K_H, K_2 (NG doubles each), a small scratch and the chunk-sum
// scratch (epilogue_only: the smaller slots of adjoint_kernel<NG, true>).
// NG = 32 reads the transforms' lane twiddles from an LDS copy of the table:
// its geometry (two register sets of twiddles) would otherwise push the
// kernel's registers into spills.
template <int NG>
__host__ __device__ constexpr bool adjoint_lds_twiddles() {
  return NG == 32;
}

template <int NG>
__host__ __device__ constexpr size_t adjoint_lds_bytes(int nw, bool epilogue_only = false) {
  return (size_t)nw * (64 / GroupFFT<NG>::L) *
             (epilogue_only ? adjoint_epi_slot<NG>() : adjoint_slot<NG>()) * sizeof(zd) +
         4 * NG * sizeof(zd) + 4 * NG * sizeof(double) + 16 * sizeof(double) +
         ((size_t)64 * nw > 4 * NG ? (size_t)64 * nw : 4 * NG) * sizeof(double) +
         (adjoint_lds_twiddles<NG>() ? GroupFFT<NG>::TWN * sizeof(zd) : 0);
}

// XP: the exchange part, the density and s come from the pair tiles' partial
// sums (adjtile_kernel): the pair loop is compiled out.
template <int NG, bool XP = false>
__global__ void __launch_bounds__(512) adjoint_kernel(AdjArgs a0) {
  // per-trajectory view of the arguments
  AdjArgs a = a0;
  const int tr = blockIdx.y;
  if (a0.step_base) {
    a.step = *a0.step_base - a0.jstep;
    const size_t slot = (size_t)(a.step % a0.tape_seg) * 4 + (a0.mode - 1);
    a.y = a0.tape + slot * a0.blk;
    if (a.abar) a.abar = a0.abar + slot * (size_t)(a0.nk * a0.nb);
  }
  {
    const size_t so = (size_t)tr * a0.tstride;
    a.y += so;
    if (a.gb) a.gb += so;
    if (a.ub) a.ub += so;
    if (a.ua) a.ua += so;
    if (a.gout) a.gout += so;
    if (a.cout) a.cout += so;
    if (a.abar) a.abar += (size_t)tr * a0.abar_tstride;
  }
  double fa;
  {
    const double At = a.Aarr[(size_t)tr * a.fstride + a.step];
    const double Av = a.Aavearr[(size_t)tr * a.fstride + a.step];
    fa = a.mode == A_S1 ? 2.0 * Av - At : (a.mode == A_S2 || a.mode == A_S3) ? Av : At;
  }
  using F = GroupFFT<NG>;
  constexpr int R = F::R, L = F::L, GPW = 64 / L;
  constexpr int JNYQ = F::NYQ_REG;  // register of lane 0 holding G index NG/2
  extern __shared__ zd smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = blockDim.x >> 6;
  const int l = lane % L, gw = lane / L;
  const int grp = wave * GPW + gw, ngrp = nw * GPW;
  const int row = blockIdx.x;
  const int ik = row / a.nb, ib = row - ik * a.nb;
  const bool egrad = a.mode == A_EGRAD;
  const zd* gbar = egrad ? a.y : a.gb;

  constexpr int SSZ = XP ? adjoint_epi_slot<NG>() : adjoint_slot<NG>();
  zd* T = smem + grp * SSZ;
  zd* sY = smem + ngrp * SSZ;      // y_ki
  zd* sG = sY + NG;                // gbar_ki
  zd* sK1 = sG + NG;
  zd* sACC = sK1 + NG;             // exchange part of J^T gbar, then rho and s
  double* sRho = reinterpret_cast<double*>(sACC + NG);
  double* sS = sRho + NG;
  double* sKH = sS + NG;
  double* sK2 = sKH + NG;
  double* sRed = sK2 + NG;
  double* sPart = sRed + 16;  // max(blockDim.x, 4 NG) doubles (chunk sums)
  constexpr bool LTW = adjoint_lds_twiddles<NG>();
  zd* sTw = reinterpret_cast<zd*>(sPart + (64 * nw > 4 * NG ? 64 * nw : 4 * NG));
  static_assert(8 * 64 * (NG > 64 ? NG / 64 : 1) + NG <= 2 * SSZ * GPW * (NG > 64 ? NG / 64 : 1),
                "slot area too small for the epilogue scratch");
  // XP: the runs of partial sums, 4 NG doubles per thread of a grid point
  // (256 per wave), also fit the slot area
  static_assert(!XP || 128 <= SSZ * GPW, "slot area too small for the partial-sum runs");

  const size_t o = (size_t)NG * row + (tid < NG ? tid : 0);
  zd yv = mk(0.0, 0.0), gv = mk(0.0, 0.0);
  double vionv = 0.0;
  if (tid < NG) {
    yv = gload(a.y + o);
    gv = gload(gbar + o);
    sY[tid] = yv;
    sG[tid] = gv;
    vionv = gload(a.vion + tid);
    sKH[tid] = gload(a.kc + tid);
    sK2[tid] = gload(a.kc + NG + tid);
    sK1[tid] = gload(reinterpret_cast<const zd*>(a.kc + 2 * NG) + tid);
  }
  const double kxv = gload(a.kx + ik);
  F fft;
  if constexpr (LTW) {
    for (int i = tid; i < F::TWN; i += blockDim.x) sTw[i] = gload(a.tw + i);
  } else if constexpr (!XP) {
    fft.init(a.tw, l);
  }
  // pair tiles: the exchange part, the density and s of this row from their
  // partials.  The H = blockDim / NG threads of a grid point each add a
  // contiguous run of the blocks in order, eight blocks in flight, into the
  // slot area (free: no pair loop); the runs are added in order after the
  // barrier below.
  if constexpr (XP) {
    const int H = blockDim.x / NG, x = tid % NG, h = tid / NG;
    if (h < H) {
      const int nblk = a.nblk, c1 = nblk * (h + 1) / H;
      const zd* xp = a.xpart + (size_t)tr * a.xstride + (size_t)row * nblk * NG + x;
      const double* dp = a.dpart + (size_t)tr * a.dstride + x;
      zd xacc = mk(0.0, 0.0);
      double xrho = 0.0, xs = 0.0;
      int c = nblk * h / H;
      for (; c + 8 <= c1; c += 8) {
        zd q[8];
        double d[8], e[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          q[i] = gload(xp + (size_t)(c + i) * NG);
          d[i] = gload(dp + (size_t)(c + i) * NG);
          e[i] = gload(dp + (size_t)(nblk + c + i) * NG);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          xacc = zadd(xacc, q[i]);
          xrho += d[i];
          xs += e[i];
        }
      }
      for (; c < c1; ++c) {
        xacc = zadd(xacc, gload(xp + (size_t)c * NG));
        xrho += gload(dp + (size_t)c * NG);
        xs += gload(dp + (size_t)(nblk + c) * NG);
      }
      double* hp = reinterpret_cast<double*>(smem) + (size_t)4 * NG * h;
      hp[x] = xacc.x;
      hp[NG + x] = xacc.y;
      hp[2 * NG + x] = xrho;
      hp[3 * NG + x] = xs;
    }
  }
  auto fwd = [&](zd (&v)[R]) {
    if constexpr (LTW) fft.fwd_lds(v, T, l, sTw); else fft.fwd(v, T, l);
  };
  auto inv = [&](zd (&v)[R]) {
    if constexpr (LTW) fft.inv_lds(v, T, l, sTw); else fft.inv(v, T, l);
  };
  const double sgn = (l & 1) ? -1.0 : 1.0;  // (-1)^x for x = L*m + l, L even
  const bool out_occ = ib < a.nbocc;

  zd acc[R];
  double den[R], sv[R];
  static_for<0, R>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    acc[m] = mk(0.0, 0.0);
    den[m] = 0.0;
    sv[m] = 0.0;
  });
  lds_barrier();

  // ---- pair loop over all partner rows (q, j) ----
  const int nitem = XP ? 0 : a.nk * a.nb;  // pair tiles: no pairs here
  for (int item = grp; item < nitem; item += ngrp) {
    const int iq = item / a.nb, jb = item - iq * a.nb;
    const size_t prow = (size_t)NG * (jb + a.nb * iq);
    const bool p_occ = jb < a.nbocc;
    zd uq[R], gq[R], v[R];
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      uq[m] = gload(a.y + prow + L * m + l);
      gq[m] = egrad ? mk(0.0, 0.0) : gload(a.gb + prow + L * m + l);
      den[m] += p_occ ? zabs2(uq[m]) : 0.0;
      sv[m] += gq[m].x * uq[m].x + gq[m].y * uq[m].y;
    });
    const double* vt = a.vtl + (size_t)(ik - iq + a.nk - 1) * NG + l * R;
    // D_kq = vt_qk(NG/2) - vt_kq(NG/2): G-layout position L/2 of rows (q-k), (k-q)
    const double dnyq = gload(a.vtl + (size_t)(iq - ik + a.nk - 1) * NG + JNYQ) -
                        gload(a.vtl + (size_t)(ik - iq + a.nk - 1) * NG + JNYQ);
    double vtr[R];
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      vtr[j] = gload(vt + j);
    });
    // T1: phi' = C'_kq(y_ki conj y_qj); adds phi' gbar_qj (or, for dE/du,
    // (phi + phi') y_qj)
    if (out_occ && (p_occ || !egrad)) {
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = zmulc(sY[L * m + l], uq[m]);
      });
      fwd(v);
      // rho_hat(NG/2) sits in lane 0 (of the group), register JNYQ
      const double nx = __shfl(v[JNYQ].x, gw * L, 64);
      const double ny = __shfl(v[JNYQ].y, gw * L, 64);
      static_for<0, R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        v[j] = zscal(v[j], vtr[j]);
      });
      inv(v);
      const zd fix = mk(sgn * dnyq * nx, sgn * dnyq * ny);
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const zd ph = v[m];                 // C_kq rho
        const zd php = zadd(ph, fix);       // C'_kq rho
        if (egrad) {
          const zd s2 = zadd(ph, php);
          acc[m] = zadd(acc[m], zmul(s2, uq[m]));
        } else {
          acc[m] = zadd(acc[m], zmul(php, gq[m]));
        }
      });
    }
    // T2: psi = C_kq(gbar_ki conj y_qj + y_ki conj gbar_qj) + D fhat_C(NG/2) (-1)^x
    if (!egrad) {
      double cx = 0.0, cy = 0.0;
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const zd fb = p_occ ? zmulc(sG[L * m + l], uq[m]) : mk(0.0, 0.0);
        const zd fc = out_occ ? zmulc(sY[L * m + l], gq[m]) : mk(0.0, 0.0);
        cx += fc.x;
        cy += fc.y;
        v[m] = zadd(fb, fc);
      });
      cx *= sgn;
      cy *= sgn;
#pragma unroll
      for (int off = 1; off < L; off <<= 1) {
        cx += __shfl_xor(cx, off);
        cy += __shfl_xor(cy, off);
      }
      fwd(v);
      static_for<0, R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        v[j] = zscal(v[j], vtr[j]);
      });
      inv(v);
      const zd fix = mk(sgn * dnyq * cx, sgn * dnyq * cy);
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        acc[m] = zadd(acc[m], zmul(zadd(v[m], fix), uq[m]));
      });
    }
  }

  // ---- one reduction over the groups (fixed order): exchange sum, density
  // and s land in sACC and, right after it, sRho and sS ----
  if constexpr (XP) {
    if (tid < NG) {
      const int H = blockDim.x / NG;
      const double* hp = reinterpret_cast<const double*>(smem);
      zd xa = mk(0.0, 0.0);
      double xr = 0.0, xv = 0.0;
      for (int h = 0; h < H; ++h, hp += 4 * NG) {
        xa = zadd(xa, mk(hp[tid], hp[NG + tid]));
        xr += hp[2 * NG + tid];
        xv += hp[3 * NG + tid];
      }
      sACC[tid] = xa;
      sRho[tid] = xr;
      sS[tid] = xv;
    }
  } else {
    static_assert((4 * R) % GPW == 0, "four sums per point split over the groups");
    // the groups of a wave are summed in registers, then the nw wave partials
    // through LDS, each parked in its own wave's slots as [re | im | rho | s][x]
    constexpr int NV = 4 * R, VF = NV / GPW;
    double v4[NV];
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      v4[4 * m] = acc[m].x;
      v4[4 * m + 1] = acc[m].y;
      v4[4 * m + 2] = den[m];
      v4[4 * m + 3] = sv[m];
    });
    wave_reduce_scatter<L, NV>(v4);
    double* wp = reinterpret_cast<double*>(smem + (size_t)wave * GPW * SSZ);
    static_for<0, VF>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      const int vi = k + VF * gw;  // value 4 m + c of the groups, m = point
      wp[(vi % 4) * NG + L * (vi / 4) + l] = v4[k];
    });
    lds_barrier();
    double* outd = reinterpret_cast<double*>(sACC);
    for (int t = tid; t < 4 * NG; t += blockDim.x) {
      const int c = t / NG, x = t - c * NG;
      double sum = 0.0;
      for (int w2 = 0; w2 < nw; ++w2)
        sum += reinterpret_cast<const double*>(smem + (size_t)w2 * GPW * SSZ)[t];
      outd[c < 2 ? 2 * x + c : NG * c + x] = sum;
    }
  }
  lds_barrier();

  // ---- circulant products over all threads: K_H * rho, K_H * s, K_2 * gbar,
  // K_1 * gbar, K_1 * y (Hartree potential and response, T_a gbar, P_a y) ----
  {
    const int P = blockDim.x / NG;
    const int x = tid % NG, p = tid / NG;
    if (p < P) {
      double vh = 0.0, ks = 0.0;
      zd c2 = mk(0.0, 0.0), c1 = mk(0.0, 0.0), py = mk(0.0, 0.0);
      for (int t = p; t < NG; t += P) {
        const int d = (x - t) & (NG - 1);
        const double kh = sKH[d], k2 = sK2[d];
        const zd k1 = sK1[d], gt = sG[t], yt = sY[t];
        vh = fma(kh, sRho[t], vh);
        ks = fma(kh, sS[t], ks);
        c2.x = fma(k2, gt.x, c2.x);
        c2.y = fma(k2, gt.y, c2.y);
        c1 = zadd(c1, zmul(k1, gt));
        py = zadd(py, zmul(k1, yt));
      }
      double* q = reinterpret_cast<double*>(smem) + 8 * (size_t)tid;
      q[0] = vh;
      q[1] = ks;
      q[2] = c2.x;
      q[3] = c2.y;
      q[4] = c1.x;
      q[5] = c1.y;
      q[6] = py.x;
      q[7] = py.y;
    }
  }
  lds_barrier();

  // ---- epilogue ----
  double* stg = reinterpret_cast<double*>(smem) + 8 * (size_t)blockDim.x;  // abar terms
  if (tid < NG) {
    double vh = 0.0, ks = 0.0;
    zd c2 = mk(0.0, 0.0), c1 = mk(0.0, 0.0), py = mk(0.0, 0.0);
    const int P = blockDim.x / NG;
    for (int p = 0; p < P; ++p) {
      const double* q = reinterpret_cast<const double*>(smem) + 8 * (size_t)(p * NG + tid);
      vh += q[0];
      ks += q[1];
      c2.x += q[2];
      c2.y += q[3];
      c1.x += q[4];
      c1.y += q[5];
      py.x += q[6];
      py.y += q[7];
    }
    vh *= a.occ;  // the density sum holds sum |y|^2
    const double kA = kxv + fa;
    // T_a gbar = K_2 * gbar + (k+a) K_1 * gbar + (k+a)^2/2 gbar
    const zd tg = mk(c2.x + kA * c1.x + 0.5 * kA * kA * gv.x, c2.y + kA * c1.y + 0.5 * kA * kA * gv.y);
    // dL/da term: Re conj(gbar) P_a y, P_a y = K_1 * y + (k+a) y
    stg[tid] = gv.x * (py.x + kA * yv.x) + gv.y * (py.y + kA * yv.y);
    const double occr = out_occ ? a.occ : 0.0;
    if (egrad) {
      // dE/du = occ dv [2 (T y + (vion + vH) y) + (V_F + V'_F) y]
      const zd h = zadd(tg, zscal(yv, vionv + vh));
      const zd c = zscal(zadd(zscal(h, 2.0), sACC[tid]), occr * a.dv);
      gstore(a.cout + o, c);
      const double w = a.dt / 6.0;  // gbar_4 = i (dt/6) c
      gstore(a.gout + o, mk(-w * c.y, w * c.x));
    } else {
      zd vj = zadd(tg, zscal(gv, vionv + vh));
      vj = zadd(vj, zscal(yv, 2.0 * a.occ * (ib < a.nbocc ? 1.0 : 0.0) * ks));
      vj = zadd(vj, sACC[tid]);
      const zd ubv = gload(a.ub + o);
      zd uav, kb;
      if (a.mode == A_S4) {
        uav = zadd(ubv, vj);
        kb = zadd(zscal(ubv, a.dt / 3.0), zscal(vj, a.dt));
      } else if (a.mode == A_S3) {
        uav = zadd(gload(a.ua + o), vj);
        kb = zadd(zscal(ubv, a.dt / 3.0), zscal(vj, 0.5 * a.dt));
      } else if (a.mode == A_S2) {
        uav = zadd(gload(a.ua + o), vj);
        kb = zadd(zscal(ubv, a.dt / 6.0), zscal(vj, 0.5 * a.dt));
      } else {  // A_S1: cotangent of the state before the step; next step's gbar_4
        uav = zadd(gload(a.ua + o), vj);
        kb = zscal(uav, a.dt / 6.0);
      }
      gstore(a.ua + o, uav);
      gstore(a.gout + o, mk(-kb.y, kb.x));  // i * kbar
    }
  }
  if (!egrad && a.abar) {
    // fixed-order sum of the NG terms: chunks of 8, then the chunks
    constexpr int NCH = (NG + 7) / 8;
    lds_barrier();
    if (tid < NCH) {
      double s8 = 0.0;
      for (int i = 0; i < 8 && 8 * tid + i < NG; ++i) s8 += stg[8 * tid + i];
      sPart[tid] = s8;
    }
    lds_barrier();
    if (tid == 0) {
      double t = 0.0;
      for (int c = 0; c < NCH; ++c) t += sPart[c];
      gstore(a.abar + row, t);
    }
  }
}

// Current-squared objective L = dt sum_n J_n^2 (ubk_Je, src/Functions.f90
// :289-318): for every row of every trajectory of a uniform batch,
//   c = 2 dt J[step] (-2 occ dv / vcell) P_A y,   P_A y = K_1 * y + (k + A) y
// with A = A[step], is the cotangent dL/dy of the state y after step 'step'.
// init != 0: ua = c and gb = i (dt/6) c (start of the reverse sweep);
// otherwise both are added (the state cotangent and the stage-4 cotangent of
// the step before, which the stage-1 adjoint launch has just formed).
struct JgradArgs {
  int nb, nbocc;
  double dt, dv, vcell, occ;
  const double* kx;
  const double* kc;         // circulant kernels; K_1 (complex) at kc + 2 NG
  const double* Aarr;       // per-step A of each trajectory
  const double* J;          // per-step J of each trajectory
  int step;
  const zd* y;              // state after step 'step' (all trajectories)
  zd* ua;
  zd* gb;
  int init;
  size_t tstride;
  int fstride;
  // graph-replayable form: the state after step it - 1 is stage 1 of step
  // it = *step_base - jstep in the tape (nothing to add for it = 0)
  const int* step_base;
  int jstep;
  const zd* tape;
  int tape_seg;
  size_t blk;
};

template <int NG>
__global__ void __launch_bounds__(NG) jgrad_kernel(JgradArgs a) {
  __shared__ zd sy[NG];
  __shared__ zd sk[NG];
  const int x = threadIdx.x, row = blockIdx.x, tr = blockIdx.y;
  const int ik = row / a.nb, ib = row - ik * a.nb;
  const size_t o = (size_t)tr * a.tstride + (size_t)NG * row + x;
  int step = a.step;
  const zd* y = a.y;
  if (a.step_base) {
    const int it = *a.step_base - a.jstep;
    if (it <= 0) return;  // uniform over the grid
    step = it - 1;
    y = a.tape + (size_t)(it % a.tape_seg) * 4 * a.blk;
  }
  sy[x] = gload(y + o);
  sk[x] = gload(reinterpret_cast<const zd*>(a.kc + 2 * NG) + x);
  __syncthreads();
  const double A = gload(a.Aarr + (size_t)tr * a.fstride + step);
  const double J = gload(a.J + (size_t)tr * a.fstride + step);
  zd p = zscal(sy[x], gload(a.kx + ik) + A);
  for (int t = 0; t < NG; ++t) p = zadd(p, zmul(sk[(x - t) & (NG - 1)], sy[t]));
  const double occr = ib < a.nbocc ? a.occ : 0.0;
  const zd c = zscal(p, 2.0 * a.dt * J * (-2.0 * occr * a.dv / a.vcell));
  const double w = a.dt / 6.0;
  const zd g = mk(-w * c.y, w * c.x);
  if (a.init) {
    gstore(a.ua + o, c);
    gstore(a.gb + o, g);
  } else {
    gstore(a.ua + o, zadd(gload(a.ua + o), c));
    gstore(a.gb + o, zadd(gload(a.gb + o), g));
  }
}

// out[i] = sum_r in[i*rows + r] (fixed order), for the per-row field
// derivatives of the adjoint stages.
__global__ void row_sum_kernel(const double* in, int rows, long n, double* out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = in + (size_t)i * rows;
  double s0 = 0.0, s1 = 0.0;
  int r = 0;
  for (; r + 1 < rows; r += 2) {
    s0 += p[r];
    s1 += p[r + 1];
  }
  if (r < rows) s0 += p[r];
  out[i] = s0 + s1;
}

}  // namespace tdhf
