// tdhf_kernels.h -- device kernels of the TDHF path (gfx950, FP64).
//
// The dominant kernel is apply_kernel<NG>: one workgroup per output orbital
// (k,i) ("row").  Its groups of L lanes stream the occupied orbitals u(q,j)
// of every k-point q and, for each, do the exchange pair round trip of
// ubkwbk_vFwbk (src/Functions.f90:247-257):
//     rho = w(k,i) * conj(u(q,j))  ->  FFT  ->  * -v(G+k-q)/(NG*NK)  ->  IFFT
//     acc += result * u(q,j)
// entirely in registers plus one LDS transpose per transform.  The same pass
// accumulates the density sum_{q,j} |u(q,j)|^2 (ubk_dns, src/Functions.f90
// :55-71), so no separate density kernel is needed.  After a workgroup-wide
// reduction, the Hartree round trip (dns_vH, :147-162) and the kinetic round
// trip of w (ubk_tubk, :212-234) are evaluated either as the circulant
// products they are, vH = K_H * rho and T w = K_2 * w + (k+A) K_1 * w +
// (k+A)^2/2 w, with
//     K_H(d) = 1/NG sum_G vCG(G) cos(2 pi g d/NG),
//     K_2(d) = 1/NG sum_G G^2/2 cos(2 pi g d/NG),
//     K_1(d) = 1/NG sum_G G exp(2 pi i g d/NG)
// (G in FFT order, Nyquist at -NG/2 b), by all threads of the workgroup, or,
// when that would leave each thread 32 terms or more (workgroups of one or
// two waves, NG = 128), as one transform round trip each,
//     vH = occ IFFT[vCG/NG . FFT[rho]],   T w = IFFT[(G+k+A)^2/(2 NG) . FFT[w]],
// on two groups of the first wave.  The epilogue forms
// H w = T w + (vH + vion) w + V_F w (ubkwbk_hwbk, :187-209) and applies the
// mode-specific update (plain H|w>, one RK4 stage of RT_RK4_ubk,
// src/RT_calculation.f90:154-173, or observables only).  Per-row partial sums
// of Ekin and J (src/Functions.f90:94-119,289-318) follow by Parseval, from
// the circulant products, sum_x conj(w) (K_f * w) = sum_G f(G)|w_G|^2/NG, or
// from FFT[w]; Ex (:165-183) from V_F w.
#pragma once
#include "tdhf_fft.h"

namespace tdhf {

enum ApplyMode : int {
  M_HPSI = 0,  // out = H[u] w
  M_S1 = 1,    // RK4 stage 1 (+ observables of the stage input)
  M_S2 = 2,
  M_S3 = 3,
  M_S4 = 4,
  M_OBS = 5    // observables only
};

// One trajectory of a batch: its k-mesh, tables, RK4 buffers, field and
// observables slots.  The batched launch covers the rows of all trajectories.
struct TrajDesc {
  int nk;
  int obs_stride;           // 3*NB*nk + 3
  double occ;               // 2/nk
  const double* kx;
  const double* vion;
  const double* vtl;
  const double* kc;
  zd* u0;
  zd* ksum;
  zd* sa;
  zd* sb;
  const double* A;
  const double* Aave;
  double* obs;
  // exchange from the pair tiles (tdhf_pairtile.h, read by
  // pair_epilogue_kernel): nblk partial sums per row (row r, block c at
  // (r * nblk + c) * NG) and nblk density partials (block b at b * NG)
  zd* xpart;
  double* dpart;
  int nblk;
  const double* vhg;  // vCG(G)/NG in the G layout of GroupFFT<NG> (pair-tile epilogue)
};

// Offsets inside one observables slot: rows*3 partials then (N, Eion, EH).
struct ApplyArgs {
  int nk, nb, nbocc, nbw;  // nbw = columns of w / out per k-point
  double bx, dv, occ, dt;
  double A, Aobs;           // used when Aarr == nullptr
  const double* Aarr;       // per-step A(it), or nullptr
  const double* Aavearr;    // per-step Aave(it)
  const int* step_base;     // device step counter (graph replay), or nullptr
  int jstep;                // step = *step_base + jstep
  int obs_slot;             // slot when step_base == nullptr
  int obs_stride;           // doubles per observables slot
  const double* kx;         // NK
  const double* vion;       // NG
  const double* vtl;        // (2NK-1)*NG, G layout, -v(G+k-q)/(NG*NK)
  const double* kc;         // 4*NG: circulant kernels K_H, K_2, K_1 (complex)
  const double* vhg;        // NG: vCG(G)/NG in the G layout of GroupFFT<NG>
  const zd* tw;             // twiddle table of GroupFFT<NG> (2*NG)
  const zd* u;              // (NG,NB,NK) orbitals defining rho and V_F
  const zd* w;              // (NG,nbw,NK) orbitals H acts on
  zd* out;                  // M_HPSI output (NG,nbw,NK)
  zd* u0;                   // RK4: state at the start of the step
  zd* ksum;                 // RK4: k1 + 2 k2 + 2 k3
  zd* unext;                // RK4: next stage input
  double* obs;              // observables slots, or nullptr
  int mode;
  const TrajDesc* td;       // batched launch: trajectory table, or nullptr
  const int2* rowmap;       // batched launch: (trajectory, row) per workgroup
  zd* tape;                 // RK4 stage states y1..y4 per step (adjoint), or nullptr
  int tape_seg;             // steps held by the tape: step it in slot it % tape_seg
  size_t nstate;            // tape stride of one stage block (all trajectories)
  // uniform batch (grid y = trajectory, same NK and tables): offsets of
  // trajectory t are t*tstride (states, zd), t*fstride (A, Aave), t*ostride
  // (observables)
  size_t tstride;
  int fstride;
  size_t ostride;
  int balance;  // 1: balance the waves of a SIMD by issue priority (set by the launcher)
};

__device__ __forceinline__ double gvalue(int g, int NG, double bx) {
  // Gx in FFT order, src/Initialization.f90:97-102
  return g < NG / 2 ? (double)g * bx : (double)(g - NG) * bx;
}

// Exchange-and-add steps of wave_reduce_scatter for a pair of values (x, y)
// over one lane bit: lanes with the bit clear return x summed over the lane
// pair, lanes with the bit set return y summed over the pair.
// Bit 5: v_permlane32_swap exchanges the upper half of x with the lower half
// of y, so x' + y' is (x_lo + x_hi | y_lo + y_hi).
__device__ __forceinline__ double pair_sum_bit5(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x),
                                                   (unsigned)__double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x),
                                                   (unsigned)__double2hiint(y), false, false);
  return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}
// Bit 4: v_permlane16_swap exchanges the odd 16-lane rows of x with the even
// rows of y.
__device__ __forceinline__ double pair_sum_bit4(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x),
                                                   (unsigned)__double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x),
                                                   (unsigned)__double2hiint(y), false, false);
  return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}
// Bit 3: DPP row_ror:8 reads lane (i ^ 8) within each 16-lane row.
__device__ __forceinline__ double pair_sum_bit3(double x, double y, bool bit) {
  const double send = bit ? x : y, keep = bit ? y : x;
  const int rl = __builtin_amdgcn_update_dpp(0, __double2loint(send), 0x128, 0xf, 0xf, false);
  const int rh = __builtin_amdgcn_update_dpp(0, __double2hiint(send), 0x128, 0xf, 0xf, false);
  return keep + __hiloint2double(rh, rl);
}

// Bit 2: ds_bpermute (__shfl_xor) from lane i ^ 4.
__device__ __forceinline__ double pair_sum_bit2(double x, double y, bool bit) {
  const double send = bit ? x : y, keep = bit ? y : x;
  return keep + __shfl_xor(send, 4);
}

// Sums the NV values v[] of every lane over the groups of L lanes of a wave
// (lane bits log2(L)..5) without LDS, as a reduce-scatter: each step hands
// half of the values to the partner lane of one lane bit and keeps the other
// half, summed.  Afterwards v[k], k < NV L / 64, holds the sum over the
// wave's groups of value k + (NV L / 64) gw, where gw = lane / L.
template <int L, int NV>
__device__ __forceinline__ void wave_reduce_scatter(double* v) {
  static_assert(L == 4 || L == 8 || L == 16, "lane groups of 4, 8 or 16");
  static_assert((NV * L) % 64 == 0, "values must split evenly over the groups");
  const int lane = threadIdx.x & 63;
  static_for<0, NV / 2>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    v[k] = pair_sum_bit5(v[k], v[k + NV / 2]);
  });
  static_for<0, NV / 4>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    v[k] = pair_sum_bit4(v[k], v[k + NV / 4]);
  });
  if constexpr (L <= 8) {
    const bool b3 = (lane & 8) != 0;
    static_for<0, NV / 8>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      v[k] = pair_sum_bit3(v[k], v[k + NV / 8], b3);
    });
  }
  if constexpr (L <= 4) {
    const bool b2 = (lane & 4) != 0;
    static_for<0, NV / 16>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      v[k] = pair_sum_bit2(v[k], v[k + NV / 16], b2);
    });
  }
}

// Sets the issue priority (0..3) of the executing wave; p must be the same in
// every lane of the wave.
__device__ __forceinline__ void set_wave_priority(int p) {
  switch (__builtin_amdgcn_readfirstlane(p)) {
    case 3: __builtin_amdgcn_s_setprio(3); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    default: __builtin_amdgcn_s_setprio(0); break;
  }
}

// Scratch doubles after the fixed LDS arrays of apply_kernel (the chunk sums
// of the observables), max(threads, 3 NG).
__host__ __device__ constexpr size_t apply_part_doubles(int nw, int NG) {
  return (size_t)64 * nw > (size_t)3 * NG ? (size_t)64 * nw : (size_t)3 * NG;
}

// Per-group LDS slot of apply_kernel (complex elements): the transpose tile
// during the pair loop; afterwards the slot area holds the wave partials of
// the reduction (3 NG doubles per wave), the circulant-product partials or the
// tiles of the Hartree and kinetic transforms, and the observables staging.  At least 3 NG / 2 elements, rounded up to 8 mod 16
// elements, so that consecutive groups start half a bank cycle (32 of the 64
// four-byte banks) apart and the transposes of two groups served in one LDS
// pass do not collide.
template <int NG>
__host__ __device__ constexpr int apply_slot() {
  constexpr int need = GroupFFT<NG>::TSZ >= (3 * NG + 1) / 2 ? GroupFFT<NG>::TSZ : (3 * NG + 1) / 2;
  return need + ((8 - need % 16) + 16) % 16;
}

// LDS footprint of apply_kernel<NG> with nw waves (bytes): the groups' slots
// (also the circulant-product partials and observables staging after the
// pair loop), then w, V_F w (followed by rho), K_1 or T w (complex NG each),
// rho, K_H or vH, K_2 or vCG/NG in the G layout (NG doubles each), a small
// scratch and the reduction scratch.
template <int NG>
__host__ __device__ constexpr size_t apply_lds_bytes(int nw) {
  return (size_t)nw * (64 / GroupFFT<NG>::L) * apply_slot<NG>() * sizeof(zd) +
         3 * NG * sizeof(zd) + 3 * NG * sizeof(double) + 16 * sizeof(double) +
         apply_part_doubles(nw, NG) * sizeof(double);
}

// TE: the Hartree and kinetic round trips as transforms (else circulant
// products); the launcher chooses (apply_by_transforms).
template <int NG>
__host__ __device__ constexpr bool apply_by_transforms(int nw) {
  return NG >= 64 && NG >= 32 * (64 * nw / NG);
}

template <int NG, bool TE>
__global__ void __launch_bounds__(512) apply_kernel(ApplyArgs a) {
  using F = GroupFFT<NG>;
  constexpr int R = F::R, L = F::L, GPW = 64 / L;
  constexpr int SSZ = apply_slot<NG>();
  // after the reduction the slot area holds 5 partials per thread and the
  // observables staging (7 NG doubles at 5 * blockDim); checked at the
  // smallest launch (one wave, or NG/64 waves)
  static_assert(5 * 64 * (NG > 64 ? NG / 64 : 1) + 7 * NG <=
                    2 * SSZ * GPW * (NG > 64 ? NG / 64 : 1),
                "slot area too small for the epilogue scratch");
  extern __shared__ zd smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = blockDim.x >> 6;
  const int l = lane % L, gw = lane / L;
  const int grp = wave * GPW + gw, ngrp = nw * GPW;
  // per-workgroup operands: from the arguments, or from the trajectory table
  int row = blockIdx.x;
  int nk = a.nk, obs_stride = a.obs_stride;
  double occ = a.occ;
  const double *kx = a.kx, *vion = a.vion, *vtl = a.vtl, *kc = a.kc, *vhg = a.vhg;
  const double *Aarr = a.Aarr, *Aavearr = a.Aavearr;
  const zd *u = a.u, *w = a.w;
  zd *u0 = a.u0, *ksum = a.ksum, *unext = a.unext, *out = a.out;
  double* obs = a.obs;
  const int tr = blockIdx.y;  // trajectory of a uniform batch (0 otherwise)
  if (tr > 0) {
    Aarr += (size_t)tr * a.fstride;
    Aavearr += (size_t)tr * a.fstride;
  }
  if (a.td) {
    const int2 rm = a.rowmap[blockIdx.x];
    if (rm.x < 0) return;  // padding entry of the schedule (whole workgroup)
    const TrajDesc& t = a.td[rm.x];
    row = rm.y;
    nk = t.nk;
    obs_stride = t.obs_stride;
    occ = t.occ;
    kx = t.kx;
    vion = t.vion;
    vtl = t.vtl;
    kc = t.kc;
    vhg = t.vhg;
    Aarr = t.A;
    Aavearr = t.Aave;
    u0 = t.u0;
    ksum = t.ksum;
    obs = t.obs;
    switch (a.mode) {
      case M_S1: u = t.u0; unext = t.sa; break;
      case M_S2: u = t.sa; unext = t.sb; break;
      case M_S3: u = t.sb; unext = t.sa; break;
      case M_S4: u = t.sa; unext = nullptr; break;
      default: u = t.u0; unext = nullptr; break;
    }
    w = u;
  }
  const int ik = row / a.nbw, ib = row - ik * a.nbw;

  zd* T = smem + grp * SSZ;
  zd* sW = smem + ngrp * SSZ;
  zd* sVF = sW + NG;
  double* sRho = reinterpret_cast<double*>(sVF + NG);  // follows V_F w (one reduction)
  const int P = blockDim.x / NG;  // >= 1: at least NG threads
  zd* sK1 = reinterpret_cast<zd*>(sRho + NG);          // K_1, or T w
  double* sKH = reinterpret_cast<double*>(sK1 + NG);   // K_H, or vH
  double* sK2 = sKH + NG;                              // K_2, or vCG/NG (G layout)
  double* sRed = sK2 + NG;
  double* sPart = sRed + 16;  // apply_part_doubles(nw, NG)

  // taped forward: the stage states live in the tape slots of this step
  zd* tape0 = nullptr;
  if (a.tape && a.step_base) {
    const int tstep = *a.step_base + a.jstep;
    zd* tp = a.tape + (size_t)(tstep % a.tape_seg) * 4 * a.nstate;
    switch (a.mode) {
      case M_S1: tape0 = tp; unext = tp + a.nstate; break;
      case M_S2: u = w = tp + a.nstate; unext = tp + 2 * a.nstate; break;
      case M_S3: u = w = tp + 2 * a.nstate; unext = tp + 3 * a.nstate; break;
      case M_S4: u = w = tp + 3 * a.nstate; break;
      default: break;
    }
  }
  if (tr > 0) {
    const size_t so = (size_t)tr * a.tstride;
    u += so;
    w += so;
    if (u0) u0 += so;
    if (ksum) ksum += so;
    if (unext) unext += so;
    if (out) out += so;
    if (tape0) tape0 += so;
    if (obs) obs += (size_t)tr * a.ostride;
  }

  // ---- prologue: every operand load of the first pass is issued before the
  // first wait, so the loop starts after one memory round trip: the lane
  // twiddles, the first u(q,j), then w of this row and the kinetic kernels,
  // which go to LDS ----
  const size_t o = (size_t)NG * row + (tid < NG ? tid : 0);
  F fft;
  fft.init(a.tw, l);

  zd acc[R], un[R];
  double den[R];

  // ---- exchange pair loop (ubkwbk_vFwbk, src/Functions.f90:247-257) ----
  // u(q,j) of the next pair is fetched while the current pair is transformed;
  // the multipliers of the current pair are loaded before that prefetch, so
  // waiting for them leaves the prefetch in flight.
  const int nitem = nk * a.nbocc;
  // item = jb + nbocc iq; a group's items are ngrp apart
  const int dq = ngrp / a.nbocc, dj = ngrp - dq * a.nbocc;
  int cq = grp / a.nbocc, cj = grp - cq * a.nbocc;  // item of the current pass
  auto fetch = [&](int iq, int jb) {
    const zd* urow = u + (size_t)NG * (jb + a.nb * iq);
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      un[m] = gload(urow + L * m + l);
    });
  };
  if (grp < nitem) fetch(cq, cj);
  zd w_o = mk(0.0, 0.0), k1_o = mk(0.0, 0.0);
  double k2_o = 0.0;
  if (tid < NG) {
    w_o = gload(w + o);
    if constexpr (TE) {
      k2_o = gload(vhg + tid);
    } else {
      k2_o = gload(kc + NG + tid);
      k1_o = gload(reinterpret_cast<const zd*>(kc + 2 * NG) + tid);
    }
  }
  const double kxv = gload(kx + ik);
  static_for<0, R>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    acc[m] = mk(0.0, 0.0);
    den[m] = 0.0;
  });
  if (tid < NG) {
    sW[tid] = w_o;
    sK2[tid] = k2_o;
    sK1[tid] = k1_o;
  }
  // When the whole grid is one workgroup per CU, the two waves sharing a SIMD
  // belong to this workgroup and would otherwise finish far apart (the older
  // one is issued first whenever both are ready, and the other then runs the
  // rest of its items alone), while the workgroup waits for the later one.
  // With a.balance each wave lowers its issue priority as it passes quarters
  // of its items, so a wave that falls behind is issued first until it
  // catches up.  With more workgroups than CUs the partner wave belongs to
  // another workgroup and the default (oldest first) order is faster.
  const int wave_items = (nitem - wave * GPW + ngrp - 1) / ngrp;  // of its first group
  lds_barrier();  // sW visible

  int it = 0;
  for (int item = grp; item < nitem; item += ngrp, ++it) {
    if (a.balance) set_wave_priority(3 - (4 * it) / (wave_items > 0 ? wave_items : 1));
    zd uq[R], v[R];
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      uq[m] = un[m];
    });
    const double* vt = vtl + (size_t)(ik - cq + nk - 1) * NG + l * R;
    double vtr[R];
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      vtr[j] = gload(vt + j);
    });
    // unconditional (the last pass re-reads its own item), so the wait for
    // vtr above counts exactly the loads issued after it
    int nq = cq + dq, nj = cj + dj;
    if (nj >= a.nbocc) {
      nj -= a.nbocc;
      ++nq;
    }
    if (item + ngrp >= nitem) {
      nq = cq;
      nj = cj;
    }
    fetch(nq, nj);
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      v[m] = zmulc(sW[L * m + l], uq[m]);
      den[m] += zabs2(uq[m]);
    });
    fft.fwd(v, T, l);
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      v[j] = zscal(v[j], vtr[j]);
    });
    fft.inv(v, T, l);
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      acc[m].x = fma(v[m].x, uq[m].x, acc[m].x);
      acc[m].x = fma(-v[m].y, uq[m].y, acc[m].x);
      acc[m].y = fma(v[m].x, uq[m].y, acc[m].y);
      acc[m].y = fma(v[m].y, uq[m].x, acc[m].y);
    });
    cq = nq;
    cj = nj;
  }
  if (a.balance) set_wave_priority(0);

  // field values of this launch: read only now, so that the device step
  // counter and the field arrays (a chain of two dependent loads) do not
  // hold back the operand loads of the pair loop
  int slot = a.obs_slot;
  double Akin = a.A, Aobs = a.Aobs;
  if (a.step_base) {
    const int step = *a.step_base + a.jstep;
    const double At = gload(Aarr + step), Aav = gload(Aavearr + step);
    if (a.mode == M_S1) {
      Akin = 2.0 * Aav - At;  // A(it-1) as in RT_RK4_ubk
      Aobs = step > 0 ? gload(Aarr + step - 1) : 0.0;
      slot = step;
    } else if (a.mode == M_S2 || a.mode == M_S3) {
      Akin = Aav;
    } else if (a.mode == M_S4) {
      Akin = At;
    } else {  // M_OBS after step 'step'
      Akin = At;
      Aobs = At;
      slot = step + 1;
    }
  }

  // epilogue operands: loaded now, used after the reduction (whose barriers
  // order LDS only, so these loads stay in flight meanwhile)
  zd u0v = mk(0.0, 0.0), ksv = mk(0.0, 0.0);
  double vionv = 0.0, khv = 0.0;
  if (tid < NG) {
    if constexpr (!TE) khv = gload(kc + tid);
    vionv = gload(vion + tid);
    if (a.mode >= M_S1 && a.mode <= M_S4) {
      u0v = gload(u0 + o);
      if (a.mode != M_S1) ksv = gload(ksum + o);
    }
  }

  // ---- one reduction over all groups (fixed order) of V_F w and the density
  // sum; the sums land in sVF and, right after it, sRho (sum |u|^2, without
  // occ) ----
  {
    // the groups of a wave are summed in registers (values padded with zeros
    // to a multiple of the group count), then the nw wave partials through
    // LDS, each parked in its own wave's slots as [re | im | rho][x]
    constexpr int NV = (3 * R + GPW - 1) / GPW * GPW, VF = NV / GPW;
    double v3[NV];
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      v3[3 * m] = acc[m].x;
      v3[3 * m + 1] = acc[m].y;
      v3[3 * m + 2] = den[m];
    });
    static_for<3 * R, NV>([&](auto kc) { v3[decltype(kc)::value] = 0.0; });
    wave_reduce_scatter<L, NV>(v3);
    double* wp = reinterpret_cast<double*>(smem + (size_t)wave * GPW * SSZ);
    static_for<0, VF>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      const int vi = k + VF * gw;  // value 3 m + c of the groups, m = point
      if (vi < 3 * R) wp[(vi % 3) * NG + L * (vi / 3) + l] = v3[k];
    });
    lds_barrier();
    for (int t = tid; t < 3 * NG; t += blockDim.x) {
      const int c = t / NG, x = t - c * NG;
      double sum = 0.0;
      for (int w2 = 0; w2 < nw; ++w2)
        sum += reinterpret_cast<const double*>(smem + (size_t)w2 * GPW * SSZ)[t];
      if (c < 2)
        reinterpret_cast<double*>(sVF)[2 * x + c] = sum;
      else
        sRho[x] = sum;
    }
  }
  if (!TE && tid < NG) sKH[tid] = khv;
  lds_barrier();  // V_F w and rho complete; the slot area is free

  const bool want_obs = (a.mode == M_S1 || a.mode == M_OBS) && obs != nullptr;
  if constexpr (TE) {
    // ---- one transform round trip each: group 0 of the first wave on rho,
    // group 1 on w (the wave's other groups repeat group 1's work and store
    // nothing); group 1 also sums Ekin and J of the row by Parseval ----
    if (wave == 0) {
      const double kA = kxv + Akin, kAo = kxv + Aobs;
      zd v[R];
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const int x = L * m + l;
        v[m] = gw == 0 ? mk(sRho[x], 0.0) : sW[x];
      });
      fft.fwd(v, T, l);
      double ek = 0.0, jc = 0.0;
      static_for<0, R>([&](auto jc_) {
        constexpr int j = decltype(jc_)::value;
        const double G = gvalue(F::gindex(l, j), NG, a.bx);
        if (want_obs) {
          const double p2 = zabs2(v[j]), q = G + kAo;
          ek = fma(0.5 * q * q, p2, ek);
          jc = fma(q, p2, jc);
        }
        const double q = G + kA;
        v[j] = zscal(v[j], gw == 0 ? sK2[l * R + j] : (0.5 / NG) * q * q);
      });
      fft.inv(v, T, l);
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const int x = L * m + l;
        if (gw == 0) sKH[x] = occ * v[m].x;  // the density sum holds sum |u|^2
        else if (gw == 1) sK1[x] = v[m];
      });
      if (want_obs) {
#pragma unroll
        for (int off = 1; off < L; off <<= 1) {
          ek += __shfl_xor(ek, off);
          jc += __shfl_xor(jc, off);
        }
        if (gw == 1 && l == 0) {
          sRed[0] = ek / NG;
          sRed[1] = jc / NG;
        }
      }
    }
  } else {
    // ---- circulant products over all threads: K_H * rho, K_2 * w, K_1 * w;
    // thread (p, x) sums the terms y = p, p + P, ... into partials parked in
    // the slot area (free after the reduction), then thread x adds the P
    // partials in order ----
    const int x = tid % NG, p = tid / NG;
    if (p < P) {
      double vh = 0.0;
      zd c2 = mk(0.0, 0.0), c1 = mk(0.0, 0.0);
      for (int y = p; y < NG; y += P) {
        const int d = (x - y) & (NG - 1);
        const zd wy = sW[y];
        const double k2 = sK2[d];
        const zd k1 = sK1[d];
        vh = fma(sKH[d], sRho[y], vh);
        c2.x = fma(k2, wy.x, c2.x);
        c2.y = fma(k2, wy.y, c2.y);
        c1.x = fma(k1.x, wy.x, c1.x);
        c1.x = fma(-k1.y, wy.y, c1.x);
        c1.y = fma(k1.x, wy.y, c1.y);
        c1.y = fma(k1.y, wy.x, c1.y);
      }
      double* q = reinterpret_cast<double*>(smem) + 5 * (size_t)tid;
      q[0] = vh;
      q[1] = c2.x;
      q[2] = c2.y;
      q[3] = c1.x;
      q[4] = c1.y;
    }
  }
  lds_barrier();

  // ---- epilogue: H w and the mode-specific update ----
  if (tid < NG) {
    const zd wv = sW[tid];
    const zd vf = sVF[tid];
    double vh, ekt = 0.0, jt = 0.0;  // per-point Ekin and J terms (circulant form)
    zd tw;
    if constexpr (TE) {
      vh = sKH[tid];
      tw = sK1[tid];
    } else {
      const double* cp = reinterpret_cast<const double*>(smem);
      zd c2 = mk(0.0, 0.0), c1 = mk(0.0, 0.0);
      vh = 0.0;
      for (int p = 0; p < P; ++p) {
        const double* q = cp + 5 * (size_t)(p * NG + tid);
        vh += q[0];
        c2.x += q[1];
        c2.y += q[2];
        c1.x += q[3];
        c1.y += q[4];
      }
      vh *= occ;  // the density sum holds sum |u|^2
      const double kA = kxv + Akin;
      // T w = K_2 * w + (k+A) K_1 * w + (k+A)^2/2 w
      tw = mk(c2.x + kA * c1.x + 0.5 * kA * kA * wv.x, c2.y + kA * c1.y + 0.5 * kA * kA * wv.y);
      if (want_obs) {
        const double kAo = kxv + Aobs;
        const double w2 = zabs2(wv);
        const double wc1 = wv.x * c1.x + wv.y * c1.y;  // Re conj(w) (K_1 * w)
        const double wc2 = wv.x * c2.x + wv.y * c2.y;
        ekt = wc2 + kAo * wc1 + 0.5 * kAo * kAo * w2;
        jt = wc1 + kAo * w2;
      }
    }
    zd h = zadd(tw, zscal(wv, vh + vionv));
    h = zadd(h, vf);
    if (a.mode == M_HPSI) {
      gstore(out + o, h);
    } else if (a.mode >= M_S1 && a.mode <= M_S4) {
      const zd kk = mk(h.y, -h.x);  // -i * H w
      const double hdt = 0.5 * a.dt;
      if (a.mode == M_S1) {
        if (tape0) gstore(tape0 + o, u0v);
        gstore(ksum + o, kk);
        gstore(unext + o, mk(u0v.x + hdt * kk.x, u0v.y + hdt * kk.y));
      } else if (a.mode == M_S2 || a.mode == M_S3) {
        gstore(ksum + o, mk(ksv.x + 2.0 * kk.x, ksv.y + 2.0 * kk.y));
        const double c = (a.mode == M_S2) ? hdt : a.dt;
        gstore(unext + o, mk(u0v.x + c * kk.x, u0v.y + c * kk.y));
      } else {
        gstore(u0 + o,
               mk(u0v.x + (ksv.x + kk.x) * a.dt / 6.0, u0v.y + (ksv.y + kk.y) * a.dt / 6.0));
      }
    }
    if (want_obs) {
      // per-point terms of Ekin, J (circulant form), Ex (this row) and N,
      // Eion, EH (row 0)
      const double rho = occ * sRho[tid];
      double* st = reinterpret_cast<double*>(smem) + 5 * (size_t)blockDim.x;
      st[0 * NG + tid] = ekt;
      st[1 * NG + tid] = jt;
      st[2 * NG + tid] = wv.x * vf.x + wv.y * vf.y;
      st[3 * NG + tid] = rho;
      st[4 * NG + tid] = vionv * rho;
      st[5 * NG + tid] = vh * rho;
    }
  }
  if (want_obs) {
    // six fixed-order sums over x: chunks of 8 points, then the chunks
    constexpr int NCH = (NG + 7) / 8;
    const int nv = row == 0 ? 6 : 3;
    lds_barrier();
    const double* st = reinterpret_cast<const double*>(smem) + 5 * (size_t)blockDim.x;
    double* ch = sPart;
    if (tid < nv * NCH) {
      const int v = tid / NCH, c = tid - v * NCH;
      const double* src = st + v * NG + 8 * c;
      double s = 0.0;
      for (int i = 0; i < 8 && 8 * c + i < NG; ++i) s += src[i];
      ch[tid] = s;
    }
    lds_barrier();
    if (tid == 0) {
      double t[6];
      for (int v = 0; v < nv; ++v) {
        double s = 0.0;
        for (int c = 0; c < NCH; ++c) s += ch[v * NCH + c];
        t[v] = s;
      }
      if constexpr (TE) {  // Ekin and J from the transform of w
        t[0] = sRed[0];
        t[1] = sRed[1];
      }
      const double occr = ib < a.nbocc ? occ : 0.0;
      double* p = obs + (size_t)slot * obs_stride;
      gstore(p + 3 * row + 0, occr * t[0] * a.dv);
      gstore(p + 3 * row + 1, occr * t[1]);
      gstore(p + 3 * row + 2, occr * t[2] * a.dv);
      if (row == 0) {
        const int rows = a.nbw * nk;
        gstore(p + 3 * rows + 0, t[3] * a.dv);
        gstore(p + 3 * rows + 1, t[4] * a.dv);
        gstore(p + 3 * rows + 2, 0.5 * t[5] * a.dv);
      }
    }
  }
}

// Twiddle tables of GroupFFT<NG> (2*NG complex).
template <int NG>
__global__ void twiddle_kernel(zd* tw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 * NG) tw[i] = GroupFFT<NG>::table_value(i);
}

// Compensated (Neumaier) sum: the rounding error of every addition is kept in
// c, so the result carries about one rounding of the total instead of one per
// term.  The energy objective of the adjoint is a difference of two such
// totals, and its finite-difference check divides that difference by h.
struct CompSum {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

// Sums the per-row partials of one observables slot p (rows rows) into
// (Etot, N, J, Ekin, Eion, EH, Ex), src/Functions.f90:74-91 and :289-318.
__device__ __forceinline__ void obs_reduce_slot(const double* p, int rows, double dv, double vcell,
                                                double* o) {
  CompSum ek, jj, ex;
  for (int r = 0; r < rows; ++r) {
    ek.add(p[3 * r + 0]);
    jj.add(p[3 * r + 1]);
    ex.add(p[3 * r + 2]);
  }
  const double n = p[3 * rows + 0], ei = p[3 * rows + 1], eh = p[3 * rows + 2];
  CompSum et;
  et.add(ek.s);
  et.add(ei);
  et.add(eh);
  et.add(0.5 * ex.s);
  et.add(ek.c);
  et.add(0.5 * ex.c);
  o[0] = et.value();
  o[1] = n;
  o[2] = -(jj.value() * dv / vcell);
  o[3] = ek.value();
  o[4] = ei;
  o[5] = eh;
  o[6] = 0.5 * ex.value();
}

__global__ void obs_reduce_kernel(const double* obs, int stride, int rows, int nslots, double dv,
                                  double vcell, double* out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  obs_reduce_slot(obs + (size_t)s * stride, rows, dv, vcell, out + (size_t)7 * s);
}

// Batched variant: one thread per (trajectory, slot); out[(t*nslots + s)*7].
__global__ void obs_reduce_batch_kernel(const TrajDesc* td, int ntraj, int nb, int nslots,
                                        double dv, double vcell, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntraj * nslots) return;
  const int t = i / nslots, s = i - t * nslots;
  obs_reduce_slot(td[t].obs + (size_t)s * td[t].obs_stride, nb * td[t].nk, dv, vcell,
                  out + (size_t)7 * i);
}

__global__ void step_advance_kernel(int* step_base, int by) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step_base += by;
}

__global__ void step_set_kernel(int* step_base, int value) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step_base = value;
}

// k_vCk (src/Functions.f90:33-52) for a list of arguments: one workgroup per
// argument, the nsteps trapezoid terms split over the threads and summed by
// a fixed-shape tree (deterministic).
__global__ void __launch_bounds__(256) kvck_kernel(const double* q, double dx, long nsteps,
                                                   double* out) {
  __shared__ double red[256];
  const double k = q[blockIdx.x];
  double s = 0.0;
  for (long ix = threadIdx.x; ix < nsteps; ix += blockDim.x) {
    const double x0 = (double)ix * dx, x1 = (double)(ix + 1) * dx;
    s += 0.5 * (cos(k * x0) / sqrt(x0 * x0 + 1.0) + cos(k * x1) / sqrt(x1 * x1 + 1.0));
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = 2.0 * red[0] * dx;
}

// Rearranges a raw table t[d*NG + g] into the G layout of GroupFFT<NG>,
// scaled: out[d*NG + l*R + j] = scale * t[d*NG + gindex(l,j)].
template <int NG>
__global__ void relayout_kernel(const double* t, int nrows, double scale, double* out) {
  using F = GroupFFT<NG>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows * NG) return;
  const int d = i / NG, r = i - d * NG;
  const int l = r / F::R, j = r - l * F::R;
  out[i] = scale * t[(size_t)d * NG + F::gindex(l, j)];
}

// Prep_periodic_ion_potential (src/Prep_periodic_ion_potential.f90:10-20);
// one workgroup of NG threads.
__global__ void ion_kernel(int ng, int nk, double ax, int nion, const double* z,
                           const double* alpha, const double* rion, double* vion) {
  __shared__ double red[128];
  const int x = threadIdx.x;
  double v = 0.0;
  if (x < ng) {
    const double rx = (double)x / (double)ng * ax;
    for (int ii = 0; ii < nion; ++ii)
      for (int icell = -4 * nk; icell <= 4 * nk; ++icell) {
        const double dist = fabs(rx - ax * rion[ii] - ax * (double)icell);
        v = v - z[ii] / sqrt(dist * dist + alpha[ii] * alpha[ii]);
      }
    red[x] = v;
  }
  __syncthreads();
  if (x == 0) {
    double top = red[0];
    for (int i = 1; i < ng; ++i) top = red[i] > top ? red[i] : top;
    red[127] = top;
  }
  __syncthreads();
  const double top = red[127];
  __syncthreads();
  if (x < ng) vion[x] = v - top;
}

// Block-wide deterministic sum over blockDim.x (power of two, <= 128).
__device__ __forceinline__ double block_sum(double v, double* red) {
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

// Gram_Schmidt (src/mATTOHF.f90:134-157): one workgroup (NG threads) per
// k-point, modified Gram-Schmidt with the vcell/NG weight.
__global__ void gram_schmidt_kernel(zd* u, int ng, int nb, double vcell) {
  __shared__ double red[2][128];
  const int ik = blockIdx.x, x = threadIdx.x;
  zd* uk = u + (size_t)ng * nb * ik;
  for (int ib = 0; ib < nb; ++ib) {
    zd ui = uk[(size_t)ng * ib + x];
    for (int ibt = 0; ibt < ib; ++ibt) {
      const zd ut = uk[(size_t)ng * ibt + x];
      double re = block_sum(ut.x * ui.x + ut.y * ui.y, red[0]);
      double im = block_sum(ut.x * ui.y - ut.y * ui.x, red[1]);
      re = re * vcell / (double)ng;
      im = im * vcell / (double)ng;
      ui = mk(ui.x - (ut.x * re - ut.y * im), ui.y - (ut.x * im + ut.y * re));
    }
    double s = block_sum(ui.x * ui.x + ui.y * ui.y, red[0]);
    s = s * vcell / (double)ng;
    const double r = sqrt(s);
    ui = mk(ui.x / r, ui.y / r);
    uk[(size_t)ng * ib + x] = ui;
    __syncthreads();
  }
}

}  // namespace tdhf
