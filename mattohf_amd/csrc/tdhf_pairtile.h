// tdhf_pairtile.h -- exchange of batched RK4 stages with the pair symmetry
// (gfx950, FP64; NG = 64, NB == NBocc).
//
// In an RK4 stage the exchange acts on the state itself (w = u in
// ubkwbk_vFwbk, src/Functions.f90:237-261), so the pair convolutions of
// rows r = (k,i) and s = (q,j),
//     c_rs = IFFT[ t_kq(G) FFT[u_r conj(u_s)] ]      (t: the scaled table
//                                                      -v(G+k-q)/(NG NK)),
// come in conjugate pairs: v is even, so t_qk(G) = t_kq(-G) for every grid G
// except the Nyquist one (the grid has -NG/2 b but not +NG/2 b), and
//     c_sr(x) = conj(c_rs(x)) + D_kq conj(P(NG/2)) (-1)^x,
//     D_kq = t_qk(NG/2) - t_kq(NG/2),  P = FFT[u_r conj(u_s)].
// One transform pair therefore serves row r (c_rs u_s) and row s (c_sr u_r).
//
// pairtile_kernel<W> computes, for one trajectory and one tile of TR x TR
// rows (TR = 8 W; row blocks bi <= bj; W = 8, 4 or 2 waves), every pair once
// and writes per-row partial sums; pair_epilogue_kernel (below) then adds a
// row's partials (one per row block, in block order) and applies the Hartree
// and kinetic terms and the stage update.
// Smaller tiles give more workgroups for small batches and more partials per
// row; the batch autotuner chooses.
//
// Tile schedule (every ordered pair of a tile exactly once): group g = 8 w + gw
// (wave w, group gw of the wave) owns row r = TR bi + g and keeps its sum in
// registers.  In pass (p1, p0) it pairs r with s = TR bj + 8 ((w + p1) % W)
// + (gw + p0) % 8, whose sum lives in LDS.  For a fixed p1 (a window of
// eight passes) wave w alone touches the LDS rows of block (w + p1) % W, in
// program order, so the LDS sums need no atomics and their order is fixed;
// a workgroup barrier separates the windows.  Off-diagonal tiles (bi < bj)
// run all 8 W offsets two-sided.  Diagonal tiles run the offsets o with
// o < -o (mod W, mod 8) two-sided and the four self-inverse ones ((0,0): the
// self pairs; (0,4), (W/2,0), (W/2,4)) one-sided: 4 W + 2 passes cover all
// TR x TR ordered pairs (34 at W = 8).
#pragma once
#include "tdhf_kernels.h"

namespace tdhf {

// One tile of the pair schedule: trajectory (x), row blocks bi (y) <= bj (z);
// x < 0 marks a padding entry.
struct PairTileArgs {
  const TrajDesc* td;
  const int4* tiles;
  const zd* tw;  // twiddle table of GroupFFT<64>
  const double* zrow;  // NG zeros: the table row used for partners past the last row
  int nb;              // NB (== NBocc)
  int mode;            // RK4 stage of the launch (selects the input state)
};

constexpr int kTilePitch = 72;  // complex elements per LDS row of the tile sums

// Tiles of W waves (W = 8, 4 or 2): 8 W rows per block (one per group), 8 W
// passes per off-diagonal tile.
__host__ __device__ constexpr int tile_rows(int W) { return 8 * W; }

__host__ __device__ constexpr size_t pairtile_lds_bytes(int W) {
  // one transpose tile per group and one row sum per row, both 72 complex
  return (size_t)2 * tile_rows(W) * kTilePitch * sizeof(zd);
}

// The passes of a diagonal tile of W waves of G groups each (G = 8 at NG = 64,
// 4 at NG = 128), as bit masks over p = G p1 + p0 (p1 < W, p0 < G): the
// offsets o = (p1, p0) with o <= -o (mod W, mod G) run, two-sided unless
// o == -o.
__host__ __device__ constexpr uint64_t diag_pass_mask(int W, bool two_sided, int G = 8) {
  uint64_t m = 0;
  for (int p = 0; p < G * W; ++p) {
    const int p1 = p / G, p0 = p % G, m1 = (W - p1) & (W - 1), m0 = (G - p0) & (G - 1);
    const bool self = p1 == m1 && p0 == m0;
    const bool run = self || p1 < m1 || (p1 == m1 && p0 < m0);
    if (run && (!two_sided || !self)) m |= uint64_t(1) << p;
  }
  return m;
}
// passes of a diagonal tile: 4 W + 2 (the four self-inverse offsets one-sided)
__host__ __device__ constexpr int diag_passes(int W) { return 4 * W + 2; }
static_assert(__builtin_popcountll(diag_pass_mask(8, false)) == diag_passes(8) &&
                  __builtin_popcountll(diag_pass_mask(4, false)) == diag_passes(4) &&
                  __builtin_popcountll(diag_pass_mask(2, false)) == diag_passes(2) &&
                  __builtin_popcountll(diag_pass_mask(8, true)) == diag_passes(8) - 4,
              "diagonal tiles: 4 W + 2 passes, four of them one-sided");
static_assert(__builtin_popcountll(diag_pass_mask(8, false, 4)) == 2 * 8 + 2 &&
                  __builtin_popcountll(diag_pass_mask(4, false, 4)) == 2 * 4 + 2 &&
                  __builtin_popcountll(diag_pass_mask(8, true, 4)) == 2 * 8 - 2,
              "diagonal tiles of 4 groups per wave: 2 W + 2 passes, four one-sided");

template <int W>
__global__ void __launch_bounds__(64 * W) pairtile_kernel(PairTileArgs a) {
  static_assert(W == 2 || W == 4 || W == 8, "tiles of 2, 4 or 8 waves");
  constexpr int NG = 64, TR = tile_rows(W);
  constexpr uint64_t kDiagRun = diag_pass_mask(W, false), kDiagBoth = diag_pass_mask(W, true);
  constexpr uint64_t kAll = W == 8 ? ~uint64_t(0) : (uint64_t(1) << (8 * W)) - 1;
  using F = GroupFFT<NG>;
  constexpr int R = F::R, L = F::L;  // 8 x 8
  constexpr int JNYQ = F::NYQ_REG;   // lane 0 of a group holds G index NG/2 here
  static_assert(F::TSZ <= kTilePitch, "transpose tile exceeds its LDS slot");
  extern __shared__ zd smem[];
  const int4 tile = a.tiles[blockIdx.x];
  if (tile.x < 0) return;  // padding entry (whole workgroup)
  const TrajDesc t = a.td[tile.x];  // by value: its fields stay in registers
  const int bi = tile.y, bj = tile.z;
  const bool diag = bi == bj;
  const int nb = a.nb, nk = t.nk, rows = nb * nk;
  const double* vtl = t.vtl;
  // k-point of a row, s / nb, by a float product: exact while s < 2^22
  // ((s + 1/2)/nb lies at least 1/(2 nb) from an integer)
  const float inb = 1.0f / (float)nb;
  auto kpt = [&](int s) { return __float2int_rd(((float)s + 0.5f) * inb); };
  const zd* u;
  switch (a.mode) {
    case M_S2: u = t.sa; break;
    case M_S3: u = t.sb; break;
    case M_S4: u = t.sa; break;
    default: u = t.u0; break;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l = lane % L, gw = lane / L;
  const int g = wave * 8 + gw;
  zd* T = smem + g * kTilePitch;           // this group's transpose tile
  zd* acj = smem + TR * kTilePitch;        // row sums of the J block
  const int r = TR * bi + g;
  const bool rv = r < rows;
  const int rc = rv ? r : rows - 1;        // clamped for table addresses
  const int kr = kpt(rc);
  const double sgn = (l & 1) ? -1.0 : 1.0;  // (-1)^x for x = 8 m + l

  F fft;
  fft.init(a.tw, l);
  zd ur[R], acc[R];
  static_for<0, R>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    ur[m] = rv ? gload(u + (size_t)NG * r + L * m + l) : mk(0.0, 0.0);
    acc[m] = mk(0.0, 0.0);
  });
  // zero the J sums
  for (int i = tid; i < TR * kTilePitch; i += blockDim.x) acj[i] = mk(0.0, 0.0);

  // passes: off-diagonal all 8 W offsets two-sided, diagonal those of kDiagRun
  auto partner = [&](int q) {
    return TR * bj + 8 * ((wave + (q >> 3)) & (W - 1)) + ((gw + q) & 7);
  };
  // a partner past the last row is read from the last row; its table row is
  // zeros (a.zrow), so it adds nothing to row r, and its own sum is never
  // written out
  auto fetch = [&](zd(&dst)[R], int s) {
    const zd* urow = u + (size_t)NG * (s < rows ? s : rows - 1);
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      dst[m] = gload(urow + L * m + l);
    });
  };
  uint64_t left = diag ? kDiagRun : kAll;
  const uint64_t two = diag ? kDiagBoth : kAll;
  int win = 0;  // window (p1) of the pass in flight

  // One pass: the next offset of `left` with the partner in us; the partner
  // of the pass after it is fetched into un (the last pass re-reads its own).
  auto pass = [&](zd(&us)[R], zd(&un)[R]) {
    const int p = __builtin_ctzll(left);
    left &= left - 1;
    // one barrier per window (empty windows included): in window p1 wave w
    // owns the J rows of block (w + p1) % W
    for (; win < (p >> 3); ++win) lds_barrier();
    const bool both = (two >> p) & 1;
    const int s = partner(p);
    const bool sv = s < rows;
    const int sc = sv ? s : rows - 1;
    const int ks = kpt(sc);
    // the Coulomb row of (k_r, k_s) and the Nyquist difference D
    const double* vt = (sv ? vtl + (size_t)(kr - ks + nk - 1) * NG : a.zrow) + l * R;
    double vtr[R];
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      vtr[j] = gload(vt + j);
    });
    const double dnyq = gload(vtl + (size_t)(ks - kr + nk - 1) * NG + JNYQ) -
                        gload(vtl + (size_t)(kr - ks + nk - 1) * NG + JNYQ);
    fetch(un, left ? partner(__builtin_ctzll(left)) : s);
    zd v[R];
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      v[m] = zmulc(ur[m], us[m]);
    });
    fft.fwd(v, T, l);
    const double nx = __shfl(v[JNYQ].x, gw * L, 64);
    const double ny = __shfl(v[JNYQ].y, gw * L, 64);
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      v[j] = zscal(v[j], vtr[j]);
    });
    fft.inv(v, T, l);
    // row r: c_rs u_s
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      acc[m].x = fma(v[m].x, us[m].x, acc[m].x);
      acc[m].x = fma(-v[m].y, us[m].y, acc[m].x);
      acc[m].y = fma(v[m].x, us[m].y, acc[m].y);
      acc[m].y = fma(v[m].y, us[m].x, acc[m].y);
    });
    if (both) {
      // row s: (conj(c_rs) + D conj(P(NG/2)) (-1)^x) u_r, into its LDS sum
      const double fx = sgn * dnyq * nx, fy = -sgn * dnyq * ny;
      zd* js = acj + (size_t)(s - TR * bj) * kTilePitch;
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const double cx = v[m].x + fx, cy = -v[m].y + fy;
        zd sm = js[L * m + l];
        sm.x = fma(cx, ur[m].x, sm.x);
        sm.x = fma(-cy, ur[m].y, sm.x);
        sm.y = fma(cx, ur[m].y, sm.y);
        sm.y = fma(cy, ur[m].x, sm.y);
        js[L * m + l] = sm;
      });
    }
  };

  // two partner buffers in turn (no copies between passes)
  zd ua[R], ub[R];
  fetch(ua, partner(0));  // pass 0 (offset (0,0)) runs in both kinds of tile
  lds_barrier();          // J sums zeroed
  while (left) {
    pass(ua, ub);
    if (!left) break;
    pass(ub, ua);
  }
  for (; win < W; ++win) lds_barrier();  // the last windows; J sums complete

  // ---- per-row partials: slot c of row r holds its pairs with block c ----
  const int nblk = t.nblk;
  if (rv) {
    zd* out = t.xpart + ((size_t)r * nblk + bj) * NG;
    const zd* own = acj + (size_t)g * kTilePitch;  // diagonal: this row's J sum
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      zd val = acc[m];
      if (diag) val = zadd(val, own[L * m + l]);
      gstore(out + L * m + l, val);
    });
  }
  if (!diag) {
    for (int i = tid; i < TR * NG; i += blockDim.x) {
      const int j = i / NG, x = i - j * NG;
      const int s = TR * bj + j;
      if (s < rows) gstore(t.xpart + ((size_t)s * nblk + bi) * NG + x, acj[(size_t)j * kTilePitch + x]);
    }
  } else {
    // density partial of the block: sum over its rows of |u_r|^2 (without
    // the occupation), first over the groups of a wave in registers, then
    // over the waves through LDS (the transpose tiles are free now)
    double den[R];
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      den[m] = zabs2(ur[m]);
    });
    wave_reduce_scatter<L, R>(den);  // lane (gw, l) holds point x = 8 gw + l
    lds_barrier();                   // everyone has read its J sum above
    double* wp = reinterpret_cast<double*>(smem);
    wp[wave * NG + L * gw + l] = den[0];
    lds_barrier();
    if (tid < NG) {
      double sden = 0.0;
      for (int w2 = 0; w2 < (int)(blockDim.x >> 6); ++w2) sden += wp[w2 * NG + tid];
      gstore(t.dpart + (size_t)bi * NG + tid, sden);
    }
  }
}

// Second launch of a pair-tile stage, for every row r = (k,i) of every
// trajectory: V_F w (the row's partial sums, added in block order) and the
// density (the trajectory's diagonal-block partials) come from the tiles;
// the Hartree and kinetic terms are round trips through the group's
// transform,
//     vH  = occ IFFT[vCG/NG . FFT[rho]]           (dns_vH, src/Functions.f90:147-162)
//     T w = IFFT[(G+k+A)^2/(2 NG) . FFT[w]]        (ubk_tubk, src/Functions.f90:212-234);
// then H w = T w + (vH + vion) w + V_F w (ubkwbk_hwbk, :187-209) and the RK4
// stage update of RT_RK4_ubk (src/RT_calculation.f90:154-173) or the
// observables only, as in apply_kernel.  The per-row observables follow by
// Parseval from FFT[w]:
//     sum_x conj(w) T_A w      = 1/NG sum_G (G+k+A)^2/2 |w_G|^2   (Ekin),
//     sum_x Re conj(w) P_A w   = 1/NG sum_G (G+k+A) |w_G|^2       (J).
// One group of 8 lanes per row, 8 rows per wave, one workgroup barrier (the
// twiddle table); rows[i] = (trajectory, row) for the nrows rows of the batch.
__global__ void __launch_bounds__(512) pair_epilogue_kernel(ApplyArgs a, const int2* rows,
                                                            int nrows) {
  constexpr int NG = 64;
  using F = GroupFFT<NG>;
  constexpr int R = F::R, L = F::L;  // 8 x 8, G layout: register j of lane l holds g = l + 8 j
  extern __shared__ zd smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l = lane % L, gw = lane / L;
  const int gi = (blockIdx.x * (int)(blockDim.x >> 6) + wave) * 8 + gw;
  // a group past the last row repeats the last row and stores nothing, so
  // that every lane of the wave takes part in the transforms
  const bool valid = gi < nrows;
  const int2 e = rows[valid ? gi : nrows - 1];
  const TrajDesc& t = a.td[e.x];
  const int row = e.y, nk = t.nk, nbw = a.nbw;
  const int ik = row / nbw, ib = row - ik * nbw;
  zd* T = smem + (wave * 8 + gw) * kTilePitch;
  const zd* w;
  zd* unext = nullptr;
  switch (a.mode) {
    case M_S1: w = t.u0; unext = t.sa; break;
    case M_S2: w = t.sa; unext = t.sb; break;
    case M_S3: w = t.sb; unext = t.sa; break;
    case M_S4: w = t.sa; break;
    default: w = t.u0; break;
  }
  const size_t o = (size_t)NG * row;
  const double occ = t.occ;
  // the lane twiddles of the transforms come from an LDS copy of the table
  // (no registers held across the partial-sum loads)
  zd* sTw = smem + (blockDim.x >> 3) * kTilePitch;
  for (int i = tid; i < F::TWN; i += blockDim.x) sTw[i] = gload(a.tw + i);

  // ---- operands: the row, its exchange partials, the density partials ----
  const F fft{};
  zd wv[R], vf[R];
  double rho[R];
  static_for<0, R>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    wv[m] = gload(w + o + L * m + l);
    vf[m] = mk(0.0, 0.0);
    rho[m] = 0.0;
  });
  // block c's partials are added in block order; block c + 1's are loaded
  // before that, so two blocks are in flight at a time
  const int nblk = t.nblk;
  {
    zd pa[R], pb[R];
    double da[R], db[R];
    auto ld = [&](zd(&p)[R], double(&d)[R], int c) {
      const zd* xp = t.xpart + ((size_t)row * nblk + c) * NG;
      const double* dp = t.dpart + (size_t)c * NG;
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        p[m] = gload(xp + L * m + l);
        d[m] = gload(dp + L * m + l);
      });
    };
    auto add = [&](const zd(&p)[R], const double(&d)[R]) {
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        vf[m] = zadd(vf[m], p[m]);
        rho[m] += d[m];
      });
    };
    if (nblk > 0) ld(pa, da, 0);
    for (int c = 0; c < nblk; c += 2) {
      if (c + 1 < nblk) ld(pb, db, c + 1);
      add(pa, da);
      if (c + 1 >= nblk) break;
      if (c + 2 < nblk) ld(pa, da, c + 2);
      add(pb, db);
    }
  }
  const double kxv = gload(t.kx + ik);
  lds_barrier();  // twiddle table in LDS

  // field values of this launch (as in apply_kernel)
  int slot = a.obs_slot;
  double Akin = a.A, Aobs = a.Aobs;
  if (a.step_base) {
    const int step = *a.step_base + a.jstep;
    const double At = gload(t.A + step), Aav = gload(t.Aave + step);
    if (a.mode == M_S1) {
      Akin = 2.0 * Aav - At;  // A(it-1) as in RT_RK4_ubk
      Aobs = step > 0 ? gload(t.A + step - 1) : 0.0;
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
  const bool want_obs = (a.mode == M_S1 || a.mode == M_OBS) && t.obs != nullptr;

  // ---- Hartree potential of the trajectory's density ----
  double vh[R];
  {
    zd hv[R];
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      hv[m] = mk(rho[m], 0.0);
    });
    fft.fwd_lds(hv, T, l, sTw);
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      hv[j] = zscal(hv[j], gload(t.vhg + l * R + j));
    });
    fft.inv_lds(hv, T, l, sTw);
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      vh[m] = occ * hv[m].x;
    });
  }

  // ---- kinetic term, and Ekin and J of the row by Parseval ----
  zd tw[R];
  double ek = 0.0, jc = 0.0;
  {
    const double kA = kxv + Akin, kAo = kxv + Aobs;
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      tw[m] = wv[m];
    });
    fft.fwd_lds(tw, T, l, sTw);
    static_for<0, R>([&](auto jc_) {
      constexpr int j = decltype(jc_)::value;
      const double G = gvalue(F::gindex(l, j), NG, a.bx);
      if (want_obs) {
        const double p2 = zabs2(tw[j]), q = G + kAo;
        ek = fma(0.5 * q * q, p2, ek);
        jc = fma(q, p2, jc);
      }
      const double q = G + kA;
      tw[j] = zscal(tw[j], (0.5 / NG) * q * q);
    });
    fft.inv_lds(tw, T, l, sTw);
  }

  // ---- H w and the stage update ----
  const int mode = a.mode;
  double ex = 0.0, sn = 0.0, si = 0.0, sh = 0.0;
  static_for<0, R>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const int x = L * m + l;
    const double vio = gload(t.vion + x);
    const zd h = zadd(zadd(tw[m], zscal(wv[m], vh[m] + vio)), vf[m]);
    if (want_obs) {
      const double r = occ * rho[m];
      ex += wv[m].x * vf[m].x + wv[m].y * vf[m].y;
      sn += r;
      si += vio * r;
      sh += vh[m] * r;
    }
    if (valid && mode >= M_S1 && mode <= M_S4) {
      const zd kk = mk(h.y, -h.x);  // -i * H w
      const double hdt = 0.5 * a.dt;
      const zd u0v = mode == M_S1 ? wv[m] : gload(t.u0 + o + x);
      if (mode == M_S1) {
        gstore(t.ksum + o + x, kk);
        gstore(unext + o + x, mk(u0v.x + hdt * kk.x, u0v.y + hdt * kk.y));
      } else {
        const zd ksv = gload(t.ksum + o + x);
        if (mode == M_S2 || mode == M_S3) {
          gstore(t.ksum + o + x, mk(ksv.x + 2.0 * kk.x, ksv.y + 2.0 * kk.y));
          const double c = mode == M_S2 ? hdt : a.dt;
          gstore(unext + o + x, mk(u0v.x + c * kk.x, u0v.y + c * kk.y));
        } else {
          gstore(t.u0 + o + x, mk(u0v.x + (ksv.x + kk.x) * a.dt / 6.0,
                                  u0v.y + (ksv.y + kk.y) * a.dt / 6.0));
        }
      }
    }
  });
  if (want_obs) {
    // sums over the group's 8 lanes (the same value in every lane)
#pragma unroll
    for (int off = 1; off < L; off <<= 1) {
      ek += __shfl_xor(ek, off);
      jc += __shfl_xor(jc, off);
      ex += __shfl_xor(ex, off);
      sn += __shfl_xor(sn, off);
      si += __shfl_xor(si, off);
      sh += __shfl_xor(sh, off);
    }
    if (valid && l == 0) {
      const double occr = ib < a.nbocc ? occ : 0.0;
      double* p = t.obs + (size_t)slot * t.obs_stride;
      gstore(p + 3 * row + 0, occr * (ek / NG) * a.dv);
      gstore(p + 3 * row + 1, occr * (jc / NG));
      gstore(p + 3 * row + 2, occr * ex * a.dv);
      if (row == 0) {
        const int rows_t = nbw * nk;
        gstore(p + 3 * rows_t + 0, sn * a.dv);
        gstore(p + 3 * rows_t + 1, si * a.dv);
        gstore(p + 3 * rows_t + 2, 0.5 * sh * a.dv);
      }
    }
  }
}

// LDS of pair_epilogue_kernel with nw waves: one transpose tile per group
// and the twiddle table of GroupFFT<64>.
__host__ __device__ constexpr size_t pair_epilogue_lds_bytes(int nw) {
  return ((size_t)nw * 8 * kTilePitch + GroupFFT<64>::TWN) * sizeof(zd);
}

}  // namespace tdhf
