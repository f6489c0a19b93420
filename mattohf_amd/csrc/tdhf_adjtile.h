// tdhf_adjtile.h -- reverse (adjoint) RK4 stages with the pair symmetry
// (gfx950, FP64; NG = 128, NB == NBocc).  No counterpart in the reference:
// like adjoint_kernel (tdhf_adjoint.h) they differentiate ubkwbk_hwbk,
// src/Functions.f90:187-209, and its exchange, :237-261.
//
// The exchange part of J^T gbar for output row r = (k,i) is a sum over all
// rows s = (q,j) (every band occupied),
//     O_r = sum_s  y_s At_rs + gbar_s Bt_rs,
//     At_rs = C_kq f_rs + D_kq P_rs (-1)^x,     f_rs = gbar_r conj y_s + y_r conj gbar_s,
//                                                P_rs = sum_x (-1)^x y_r conj gbar_s,
//     Bt_rs = C_kq h_rs + D_kq hhat_rs(NG/2) (-1)^x,     h_rs = y_r conj y_s,
// with C_kq the exchange convolution by the table t_kq and
// D_kq = t_qk(NG/2) - t_kq(NG/2) (the grid holds -NG/2 but not +NG/2).
// Because f_sr = conj f_rs, h_sr = conj h_rs, D_qk = -D_kq and
// C_qk(conj g) = conj(C_kq g) + D_kq conj(ghat(NG/2)) (-1)^x, the reversed
// pair needs no transform of its own:
//     At_sr = conj(At_rs),      Bt_sr = conj(C_kq h_rs)   (its Nyquist terms cancel),
// so one pass of two transform pairs adds y_s At_rs + gbar_s Bt_rs to row r
// and y_r conj(At_rs) + gbar_r conj(C_kq h_rs) to row s.
//
// adjtile_kernel<W> runs the pair tiles of tdhf_pairtile.h with the 16-lane
// transform of NG = 128 (4 groups per wave, tiles of 4 W rows): group
// g = 4 w + gw owns row r = TR bi + g (y_r, gbar_r and its sum in registers)
// and in pass (p1, p0) pairs with s = TR bj + 4 ((w + p1) % W) + (gw + p0) % 4,
// whose sum lives in LDS; within the window of fixed p1 only wave w touches
// the LDS rows of sub-block (w + p1) % W.  Diagonal tiles run the offsets
// o <= -o (mod W, mod 4), two-sided unless o == -o (2 W + 2 passes).  Through
// the transforms a pass holds y_r, gbar_r, the row sum and the transform
// buffer only: y_s and gbar_s are read where they are consumed (again, from
// the cache, after a transform), which keeps the pass within the register
// budget of two waves per SIMD.  Each tile writes per-row partials (xpart, as the forward
// pair tiles) and, for a diagonal tile, the block's partials of the density
// sum |y|^2 and of s = Re conj(gbar) y; adjoint_kernel<128, true> then adds
// them in block order and finishes the reverse stage.
#pragma once
#include "tdhf_adjoint.h"
#include "tdhf_pairtile.h"

namespace tdhf {

struct AdjTileArgs {
  int nk, nb;
  const double* vtl;
  const zd* tw;          // twiddle table of GroupFFT<128>
  const double* zrow;    // NG zeros: the table row of partners past the last row
  const int4* tiles;     // (trajectory, bi, bj, 0); x < 0 marks a padding entry
  // stage state y and cotangent gbar of trajectory t at + t * tstride; in the
  // graph-replayable form y is the stage block of step *step_base - jstep in
  // the tape (slot step % tape_seg, blocks of blk states), as in AdjArgs
  const zd* y;
  const zd* gb;
  const int* step_base;
  int jstep;
  const zd* tape;
  int tape_seg;
  size_t blk;
  int mode;
  size_t tstride;
  zd* xpart;      // trajectory t at t * xstride: rows * nblk * NG
  double* dpart;  // trajectory t at t * dstride: nblk density, then nblk s partials
  size_t xstride, dstride;
  int nblk;
};

constexpr int kAdjTilePitch = 136;  // complex elements per LDS row of the J sums

__host__ __device__ constexpr int adjtile_rows(int W) { return 4 * W; }

__host__ __device__ constexpr size_t adjtile_lds_bytes(int W) {
  // one transpose tile per group, one J-side sum per row and a spare row
  return ((size_t)adjtile_rows(W) * GroupFFT<128>::TSZ +
          (size_t)(adjtile_rows(W) + 1) * kAdjTilePitch) * sizeof(zd);
}

template <int W>
__global__ void __launch_bounds__(64 * W) adjtile_kernel(AdjTileArgs a) {
  static_assert(W == 4 || W == 8, "tiles of 4 or 8 waves");
  constexpr int NG = 128, G = 4, TR = adjtile_rows(W);
  using F = GroupFFT<NG>;
  constexpr int R = F::R, L = F::L;   // 8 points x 16 lanes
  constexpr int JNYQ = F::NYQ_REG;    // lane 0 of a group holds G index NG/2 here
  constexpr int TS = F::TSZ;
  constexpr uint64_t kDiagRun = diag_pass_mask(W, false, G), kDiagBoth = diag_pass_mask(W, true, G);
  constexpr uint64_t kAll = (uint64_t(1) << (G * W)) - 1;
  extern __shared__ zd smem[];
  const int4 tile = a.tiles[blockIdx.x];
  if (tile.x < 0) return;  // padding entry (whole workgroup)
  const int tr = tile.x, bi = tile.y, bj = tile.z;
  const bool diag = bi == bj;
  const int nb = a.nb, nk = a.nk, rows = nb * nk;
  const double* vtl = a.vtl;
  const zd* y = a.y;
  if (a.step_base) {
    const int step = *a.step_base - a.jstep;
    const size_t slot = (size_t)(step % a.tape_seg) * 4 + (a.mode - 1);
    y = a.tape + slot * a.blk;
  }
  y += (size_t)tr * a.tstride;
  const zd* gb = a.gb + (size_t)tr * a.tstride;
  zd* xpart = a.xpart + (size_t)tr * a.xstride;
  double* dpart = a.dpart + (size_t)tr * a.dstride;
  const int nblk = a.nblk;
  // k-point of a row, s / nb, by a float product: exact while s < 2^22
  const float inb = 1.0f / (float)nb;
  auto kpt = [&](int s) { return __float2int_rd(((float)s + 0.5f) * inb); };

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l = lane % L, gw = lane / L;
  const int g = G * wave + gw;
  zd* T = smem + g * TS;                     // this group's transpose tile
  zd* acj = smem + TR * TS;                  // row sums of the J block
  const int r = TR * bi + g;
  const bool rv = r < rows;
  const int kr = kpt(rv ? r : rows - 1);
  const double sgn = (l & 1) ? -1.0 : 1.0;  // (-1)^x for x = 16 m + l

  F fft;
  fft.init(a.tw, l);
  zd yr[R], gr[R], acc[R];
  static_for<0, R>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    yr[m] = rv ? gload(y + (size_t)NG * r + L * m + l) : mk(0.0, 0.0);
    gr[m] = rv ? gload(gb + (size_t)NG * r + L * m + l) : mk(0.0, 0.0);
    acc[m] = mk(0.0, 0.0);
  });
  for (int i = tid; i < TR * kAdjTilePitch; i += blockDim.x) acj[i] = mk(0.0, 0.0);

  auto partner = [&](int q) {
    return TR * bj + G * ((wave + q / G) & (W - 1)) + ((gw + q) & (G - 1));
  };
  uint64_t left = diag ? kDiagRun : kAll;
  const uint64_t two = diag ? kDiagBoth : kAll;
  int win = 0;  // window (p1) of the pass in flight
  lds_barrier();  // J sums zeroed

  while (left) {
    const int p = __builtin_ctzll(left);
    left &= left - 1;
    // one barrier per window (empty windows included): in window p1 wave w
    // owns the J rows of sub-block (w + p1) % W
    for (; win < p / G; ++win) lds_barrier();
    const bool both = (two >> p) & 1;
    const int s = partner(p);
    const bool sv = s < rows;
    const int sc = sv ? s : rows - 1;  // a partner past the last row: zero table, D = 0
    const int ks = kpt(sc);
    const zd* ysp = y + (size_t)NG * sc;
    const zd* gsp = gb + (size_t)NG * sc;
    const double* vt = (sv ? vtl + (size_t)(kr - ks + nk - 1) * NG : a.zrow) + l * R;
    const double dnyq = sv ? gload(vtl + (size_t)(ks - kr + nk - 1) * NG + JNYQ) -
                                 gload(vtl + (size_t)(kr - ks + nk - 1) * NG + JNYQ)
                           : 0.0;
    zd v[R];
    double px = 0.0, py = 0.0;
    // ---- At: f = gbar_r conj y_s + y_r conj gbar_s, P = sum (-1)^x y_r conj gbar_s
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const zd ys = gload(ysp + L * m + l);
      const zd gs = gload(gsp + L * m + l);
      const zd fc = zmulc(yr[m], gs);
      px += fc.x;
      py += fc.y;
      v[m] = zadd(zmulc(gr[m], ys), fc);
    });
    double vtr[R];
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      vtr[j] = gload(vt + j);
    });
    px *= sgn;  // (-1)^x is (-1)^l in every register of the lane
    py *= sgn;
#pragma unroll
    for (int off = 1; off < L; off <<= 1) {
      px += __shfl_xor(px, off);
      py += __shfl_xor(py, off);
    }
    fft.fwd(v, T, l);
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      v[j] = zscal(v[j], vtr[j]);
    });
    fft.inv(v, T, l);
    // y_s is read again below (a cache hit) rather than held through the
    // transforms, and h = y_r conj y_s replaces At in place.  Row s's sum: a
    // one-sided pass (diagonal tiles) writes a spare row that is never read
    // instead, so the update needs no branch.
    zd* js = acj + (size_t)(both ? s - TR * bj : TR) * kAdjTilePitch;
    {
      const double fx = sgn * dnyq * px, fy = sgn * dnyq * py;
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const zd ys = gload(ysp + L * m + l);
        const zd at = mk(v[m].x + fx, v[m].y + fy);
        acc[m] = zadd(acc[m], zmul(at, ys));  // row r: y_s At
        {  // row s: y_r conj(At)
          zd sm = js[L * m + l];
          sm = zadd(sm, zmulc(yr[m], at));
          js[L * m + l] = sm;
        }
        v[m] = zmulc(yr[m], ys);
      });
    }
    // ---- Bt: h = y_r conj y_s (the table row read again before the
    // transform, so that the read is covered by it)
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      vtr[j] = gload(vt + j);
    });
    fft.fwd(v, T, l);
    const double nx = __shfl(v[JNYQ].x, gw * L, 64);
    const double ny = __shfl(v[JNYQ].y, gw * L, 64);
    static_for<0, R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      v[j] = zscal(v[j], vtr[j]);
    });
    fft.inv(v, T, l);
    {
      const double fx = sgn * dnyq * nx, fy = sgn * dnyq * ny;
      static_for<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const zd gs = gload(gsp + L * m + l);
        acc[m] = zadd(acc[m], zmul(mk(v[m].x + fx, v[m].y + fy), gs));  // row r: gbar_s Bt
        {  // row s: gbar_r conj(C h)
          zd sm = js[L * m + l];
          sm = zadd(sm, zmulc(gr[m], v[m]));
          js[L * m + l] = sm;
        }
      });
    }
  }
  for (; win < W; ++win) lds_barrier();  // the last windows; J sums complete

  // ---- per-row partials: slot c of row r holds its pairs with block c ----
  if (rv) {
    zd* out = xpart + ((size_t)r * nblk + bj) * NG;
    const zd* own = acj + (size_t)g * kAdjTilePitch;  // diagonal: this row's J sum
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
      if (s < rows)
        gstore(xpart + ((size_t)s * nblk + bi) * NG + x, acj[(size_t)j * kAdjTilePitch + x]);
    }
  } else {
    // the block's partials of the density sum |y|^2 and of s = Re conj(gbar) y,
    // summed over its rows in row order through the (now free) transpose tiles
    double* dq = reinterpret_cast<double*>(T);
    static_for<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      dq[L * m + l] = zabs2(yr[m]);
      dq[NG + L * m + l] = gr[m].x * yr[m].x + gr[m].y * yr[m].y;
    });
    lds_barrier();
    for (int i = tid; i < 2 * NG; i += blockDim.x) {
      double sum = 0.0;
      for (int q = 0; q < TR && TR * bi + q < rows; ++q)
        sum += reinterpret_cast<const double*>(smem + (size_t)q * TS)[i];
      const int c = i / NG, x = i - c * NG;
      gstore(dpart + ((size_t)c * nblk + bi) * NG + x, sum);
    }
  }
}

}  // namespace tdhf
