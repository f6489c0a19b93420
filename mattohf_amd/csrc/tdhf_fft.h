// tdhf_fft.h -- register-resident FP64 FFT building blocks for gfx950.
//
// A length-NG transform is owned by a group of L consecutive lanes of a
// wavefront (NG = R*L, R points per lane, R >= L, R % L == 0).  Lane l holds
// the spatial points x = L*m + l (m = 0..R-1) in registers v[m].  The
// forward transform is an R-point DFT in registers, a twiddle by
// exp(-2 pi i l g0 / NG), one transpose through the group's LDS tile, and
// R/L L-point DFTs in registers; the result stays in "G layout": register
// j = s*L + g1 of lane l holds G index g = (l*R/L + s) + R*g1.  The inverse
// transform consumes the G layout and returns the spatial layout, so a
// forward/inverse pair (a convolution) needs no digit reversal at all.
//
// Conventions are FFTW's, as used by the reference (src/Initialization.f90
// :270-273): forward = sum_x f(x) exp(-2 pi i x g / NG), backward with +i,
// both unnormalised.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace tdhf {

using zd = double2;

__device__ __forceinline__ zd mk(double x, double y) {
  zd r;
  r.x = x;
  r.y = y;
  return r;
}
__device__ __forceinline__ zd zadd(zd a, zd b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ zd zsub(zd a, zd b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ zd zmul(zd a, zd b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ zd zmulc(zd a, zd b) {
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ zd zscal(zd a, double s) { return mk(a.x * s, a.y * s); }
__device__ __forceinline__ double zabs2(zd a) { return a.x * a.x + a.y * a.y; }

// Loads and stores through the global address space.  Buffer pointers that
// come from tables in device memory are generic, and plain dereferences of
// them compile to flat instructions; the LDS wait counter also counts flat
// accesses, so an LDS wait inside a loop would also wait for global loads
// issued ahead (the prefetch of the next operand).  Every pointer passed here
// points into device global memory.
typedef double gd2 __attribute__((ext_vector_type(2)));
#define TDHF_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ zd gload(const zd* p) {
  const gd2 v = *(const TDHF_GLOBAL gd2*)(p);
  return mk(v.x, v.y);
}
__device__ __forceinline__ double gload(const double* p) {
  return *(const TDHF_GLOBAL double*)(p);
}
__device__ __forceinline__ void gstore(zd* p, zd v) {
  gd2 t;
  t.x = v.x;
  t.y = v.y;
  *(TDHF_GLOBAL gd2*)(p) = t;
}
__device__ __forceinline__ void gstore(double* p, double v) { *(TDHF_GLOBAL double*)(p) = v; }

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// cos and sin of 2*pi*j/16
template <int J>
__device__ __forceinline__ constexpr double cos16() {
  constexpr double t[16] = {1.0,
                            0.92387953251128675613,
                            0.70710678118654752440,
                            0.38268343236508977173,
                            0.0,
                            -0.38268343236508977173,
                            -0.70710678118654752440,
                            -0.92387953251128675613,
                            -1.0,
                            -0.92387953251128675613,
                            -0.70710678118654752440,
                            -0.38268343236508977173,
                            0.0,
                            0.38268343236508977173,
                            0.70710678118654752440,
                            0.92387953251128675613};
  return t[J & 15];
}
template <int J>
__device__ __forceinline__ constexpr double sin16() {
  return cos16<(J + 12) & 15>();  // sin(t) = cos(t - pi/2)
}

// a * exp(S * 2 pi i K / N) for compile-time K, N (N divides 16); the
// multiplications by 1, -1 and +-i are done without arithmetic.
template <int N, int S, int K>
__device__ __forceinline__ zd twm(zd a) {
  constexpr int j = (((K % N) + N) % N) * (16 / N);
  if constexpr (j == 0) {
    return a;
  } else if constexpr (j == 8) {
    return mk(-a.x, -a.y);
  } else if constexpr (j == 4) {
    if constexpr (S > 0)
      return mk(-a.y, a.x);
    else
      return mk(a.y, -a.x);
  } else if constexpr (j == 12) {
    if constexpr (S > 0)
      return mk(a.y, -a.x);
    else
      return mk(-a.y, a.x);
  } else {
    constexpr double c = cos16<j>();
    constexpr double s = S * sin16<j>();
    return mk(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// out[k] = sum_{j<N} in[IO + IS*j] * exp(S*2*pi*i*j*k/N), radix-2 DIT
// recursion, fully unrolled; natural order in and out.
template <int N, int S, int IS, int IO, int M>
__device__ __forceinline__ void dft(const zd (&in)[M], zd (&out)[N]) {
  if constexpr (N == 1) {
    out[0] = in[IO];
  } else {
    zd e[N / 2], o[N / 2];
    dft<N / 2, S, IS * 2, IO, M>(in, e);
    dft<N / 2, S, IS * 2, IO + IS, M>(in, o);
    static_for<0, N / 2>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      const zd t = twm<N, S, k>(o[k]);
      out[k] = zadd(e[k], t);
      out[k + N / 2] = zsub(e[k], t);
    });
  }
}

// Orders the memory operations of one wavefront around an LDS exchange
// between its lanes (lanes of a wavefront execute LDS instructions in order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup barrier that orders LDS accesses only: global loads issued
// before it stay in flight (a __syncthreads would wait for them).  For
// workgroups that exchange data through LDS and never through global memory.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int NG>
struct Geo;
template <>
struct Geo<16> {
  static constexpr int R = 4, L = 4;
};
template <>
struct Geo<32> {
  static constexpr int R = 8, L = 4;
};
template <>
struct Geo<64> {
  static constexpr int R = 8, L = 8;
};
template <>
struct Geo<128> {
  static constexpr int R = 8, L = 16;  // GroupFFT<128> below
};

template <int NG>
struct GroupFFT {
  static constexpr int R = Geo<NG>::R;
  static constexpr int L = Geo<NG>::L;
  static constexpr int Q = R / L;
  static constexpr int PAD = L + 1;      // padded LDS row (bank spread)
  static constexpr int TSZ = R * PAD;    // complex elements of one group's tile
  static constexpr int NYQ_REG = L / 2;  // lane 0 holds G index NG/2 in this register
  static_assert(R * L == NG && R % L == 0, "bad FFT geometry");

  zd fw[R];                    // exp(-2 pi i l g0 / NG), g0 = 0..R-1
  zd iw[(Q == 1) ? 1 : R];     // exp(+2 pi i x0 (l*Q+s) / NG) at s*L + x0 (Q > 1)

  // Loads this lane's twiddles from the table built by twiddle_kernel:
  // tw[l*R + g0] = fw of lane l, tw[NG + l*R + i] = iw of lane l.
  __device__ __forceinline__ void init(const zd* tw, int l) {
    static_for<0, R>([&](auto gc) {
      constexpr int g0 = decltype(gc)::value;
      fw[g0] = gload(tw + l * R + g0);
    });
    if constexpr (Q > 1) {
      static_for<0, R>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        iw[i] = gload(tw + NG + l * R + i);
      });
    }
  }

  // Value of the twiddle tables (used once, by twiddle_kernel).
  __device__ static zd table_value(int idx) {
    double sn, cs;
    if (idx < NG) {
      const int l = idx / R, g0 = idx % R;
      const int e = (l * g0) % NG;
      sincospi(2.0 * (double)e / (double)NG, &sn, &cs);
      return mk(cs, -sn);
    }
    const int r = idx - NG, l = r / R, i = r % R;
    const int s = i / L, x0 = i % L;
    const int e = (x0 * (l * Q + s)) % NG;
    sincospi(2.0 * (double)e / (double)NG, &sn, &cs);
    return mk(cs, sn);
  }

  static constexpr int TWN = 2 * NG;  // entries of the twiddle table

  // spatial layout -> G layout (forward, exp(-i...)), T = this group's tile;
  // the lane twiddles come from registers (fwd, after init), from a copy of
  // the twiddle table (fwd_lds, e.g. in LDS: no registers held) or from an
  // accessor (fwd_t: fwf(g0) = fw[g0] of this lane)
  __device__ __forceinline__ void fwd(zd (&v)[R], zd* T, int l) const {
    fwd_t(v, T, l, [&](int g0) { return fw[g0]; });
  }
  __device__ __forceinline__ void fwd_lds(zd (&v)[R], zd* T, int l, const zd* tw) const {
    fwd_t(v, T, l, [&](int g0) { return tw[l * R + g0]; });
  }
  template <class FW>
  __device__ __forceinline__ void fwd_t(zd (&v)[R], zd* T, int l, FW&& fwf) const {
    zd y[R];
    dft<R, -1, 1, 0, R>(v, y);
    static_for<1, R>([&](auto gc) {
      constexpr int g0 = decltype(gc)::value;
      y[g0] = zmul(y[g0], fwf(g0));
    });
    static_for<0, R>([&](auto gc) {
      constexpr int g0 = decltype(gc)::value;
      T[g0 * PAD + l] = y[g0];
    });
    wave_sync();
    zd z[R];
    static_for<0, R>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int s = i / L, x0 = i % L;
      z[i] = T[(l * Q + s) * PAD + x0];
    });
    wave_sync();
    static_for<0, Q>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      zd t[L];
      dft<L, -1, 1, s * L, R>(z, t);
      static_for<0, L>([&](auto gc) {
        constexpr int g1 = decltype(gc)::value;
        v[s * L + g1] = t[g1];
      });
    });
  }

  // G layout -> spatial layout (backward, exp(+i...)); inv_t as fwd_t, with
  // iwf(i) = iw[i] for Q > 1
  __device__ __forceinline__ void inv(zd (&v)[R], zd* T, int l) const {
    inv_t(v, T, l, [&](int g0) { return fw[g0]; },
          [&](int i) { return iw[(Q == 1) ? 0 : i]; });
  }
  __device__ __forceinline__ void inv_lds(zd (&v)[R], zd* T, int l, const zd* tw) const {
    inv_t(v, T, l, [&](int g0) { return tw[l * R + g0]; },
          [&](int i) { return tw[NG + l * R + i]; });
  }
  template <class FW, class IW>
  __device__ __forceinline__ void inv_t(zd (&v)[R], zd* T, int l, FW&& fwf, IW&& iwf) const {
    zd z[R];
    static_for<0, Q>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      zd t[L];
      dft<L, 1, 1, s * L, R>(v, t);
      static_for<0, L>([&](auto xc) {
        constexpr int x0 = decltype(xc)::value;
        z[s * L + x0] = t[x0];
      });
    });
    if constexpr (Q == 1) {
      static_for<1, R>([&](auto xc) {
        constexpr int x0 = decltype(xc)::value;
        z[x0] = zmulc(z[x0], fwf(x0));
      });
    } else {
      static_for<0, R>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        z[i] = zmul(z[i], iwf(i));
      });
    }
    static_for<0, R>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int s = i / L, x0 = i % L;
      T[(l * Q + s) * PAD + x0] = z[i];
    });
    wave_sync();
    zd y[R];
    static_for<0, R>([&](auto gc) {
      constexpr int g0 = decltype(gc)::value;
      y[g0] = T[g0 * PAD + l];
    });
    wave_sync();
    dft<R, 1, 1, 0, R>(y, v);
  }

  // G index held by lane l in register j of the G layout
  __host__ __device__ static constexpr int gindex(int l, int j) {
    return (l * Q + j / L) + R * (j % L);
  }
};

// NG = 128 on 16 lanes x 8 points (the 8 x 16 split of the generic engine
// would need 16 points per lane and does not fit the register file).
// 128 = 8 * 4 * 4 with x = 16 m + 4 l2 + l1 and g = g0 + 8 h1 + 32 h2:
//   exp(-2 pi i x g/128) = w8^(m g0) * w128^((4 l2 + l1) g0) * w4^(l2 h1)
//                          * w16^(l1 h1) * w4^(l1 h2)
// Forward: 8-point DFT over m in registers, twiddle w128^(l g0); transpose
// so that lane (a, l1) = (lane/4, lane%4) holds g0 = 2a+s (s = 0, 1) and
// l2 = 0..3; 4-point DFTs over l2, twiddle w16^(l1 h1); transpose within each
// quad of lanes so that lane (a, c) holds h1 = c and l1 = 0..3; 4-point DFTs
// over l1.  G layout: lane (a, c), register s*4 + h2 holds
// g = (2a + s) + 8 c + 32 h2.  The inverse runs the same steps backwards with
// conjugate twiddles.  Tile offsets are chosen so that neither transpose has
// bank conflicts within a group (rows of 18 for the first, 44 a + 19 s +
// 5 h1 + l1 for the second).
template <>
struct GroupFFT<128> {
  static constexpr int NG = 128, R = 8, L = 16, Q = 1;
  static constexpr int TSZ = 170;
  static constexpr int NYQ_REG = 2;  // g = 64: lane 0, s = 0, h2 = 2

  zd fw[8];   // w128^(l g0), g0 = 0..7 (lane l)
  zd fq[4];   // w16^(l1 h1), h1 = 0..3 (l1 = lane % 4)

  // tw[l*8 + g0] = w128^(l g0); tw[128 + l1*4 + h1] = w16^(l1 h1)
  __device__ __forceinline__ void init(const zd* tw, int l) {
    static_for<0, 8>([&](auto gc) {
      constexpr int g0 = decltype(gc)::value;
      fw[g0] = gload(tw + l * 8 + g0);
    });
    static_for<0, 4>([&](auto hc) {
      constexpr int h = decltype(hc)::value;
      fq[h] = gload(tw + 128 + (l & 3) * 4 + h);
    });
  }

  __device__ static zd table_value(int idx) {
    double sn, cs;
    int e;
    if (idx < 128) {
      e = ((idx / 8) * (idx % 8)) % 128;
    } else {
      const int r = idx - 128;
      e = r < 16 ? (((r / 4) * (r % 4)) % 16) * 8 : 0;
    }
    sincospi(2.0 * (double)e / 128.0, &sn, &cs);
    return mk(cs, -sn);
  }

  static __device__ __forceinline__ int t1(int g0, int l) { return g0 * 18 + l; }
  static __device__ __forceinline__ int t2(int a, int s, int h1, int l1) {
    return 44 * a + 19 * s + 5 * h1 + l1;
  }

  static constexpr int TWN = 144;  // entries of the twiddle table

  // fwd / inv with the twiddles in registers (after init), fwd_lds / inv_lds
  // with them read from a copy of the table (e.g. in LDS), fwd_t / inv_t with
  // accessors fwf(g0) = fw[g0], fqf(h) = fq[h] of this lane
  __device__ __forceinline__ void fwd(zd (&v)[8], zd* T, int l) const {
    fwd_t(v, T, l, [&](int g0) { return fw[g0]; }, [&](int h) { return fq[h]; });
  }
  __device__ __forceinline__ void fwd_lds(zd (&v)[8], zd* T, int l, const zd* tw) const {
    fwd_t(v, T, l, [&](int g0) { return tw[l * 8 + g0]; },
          [&](int h) { return tw[128 + (l & 3) * 4 + h]; });
  }
  __device__ __forceinline__ void inv(zd (&v)[8], zd* T, int l) const {
    inv_t(v, T, l, [&](int g0) { return fw[g0]; }, [&](int h) { return fq[h]; });
  }
  __device__ __forceinline__ void inv_lds(zd (&v)[8], zd* T, int l, const zd* tw) const {
    inv_t(v, T, l, [&](int g0) { return tw[l * 8 + g0]; },
          [&](int h) { return tw[128 + (l & 3) * 4 + h]; });
  }

  template <class FW, class FQ>
  __device__ __forceinline__ void fwd_t(zd (&v)[8], zd* T, int l, FW&& fwf, FQ&& fqf) const {
    const int a = l >> 2, l1 = l & 3;
    zd y[8];
    dft<8, -1, 1, 0, 8>(v, y);
    static_for<1, 8>([&](auto gc) {
      constexpr int g0 = decltype(gc)::value;
      y[g0] = zmul(y[g0], fwf(g0));
    });
    static_for<0, 8>([&](auto gc) {
      constexpr int g0 = decltype(gc)::value;
      T[t1(g0, l)] = y[g0];
    });
    wave_sync();
    zd z[8];
    static_for<0, 8>([&](auto ic) {
      constexpr int i = decltype(ic)::value, s = i / 4, l2 = i % 4;
      z[i] = T[t1(2 * a + s, 4 * l2 + l1)];
    });
    wave_sync();
    static_for<0, 2>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      zd t[4];
      dft<4, -1, 1, s * 4, 8>(z, t);
      static_for<0, 4>([&](auto hc) {
        constexpr int h1 = decltype(hc)::value;
        T[t2(a, s, h1, l1)] = h1 == 0 ? t[0] : zmul(t[h1], fqf(h1));
      });
    });
    wave_sync();
    static_for<0, 8>([&](auto ic) {
      constexpr int i = decltype(ic)::value, s = i / 4, k1 = i % 4;
      z[i] = T[t2(a, s, l1, k1)];  // lane (a, c = l1) gathers l1' = k1
    });
    wave_sync();
    static_for<0, 2>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      zd t[4];
      dft<4, -1, 1, s * 4, 8>(z, t);
      static_for<0, 4>([&](auto hc) {
        constexpr int h2 = decltype(hc)::value;
        v[s * 4 + h2] = t[h2];
      });
    });
  }

  template <class FW, class FQ>
  __device__ __forceinline__ void inv_t(zd (&v)[8], zd* T, int l, FW&& fwf, FQ&& fqf) const {
    const int a = l >> 2, c = l & 3;
    static_for<0, 2>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      zd t[4];
      dft<4, 1, 1, s * 4, 8>(v, t);  // over h2 -> l1
      static_for<0, 4>([&](auto kc) {
        constexpr int k1 = decltype(kc)::value;
        T[t2(a, s, c, k1)] = k1 == 0 ? t[0] : zmulc(t[k1], fqf(k1));  // w16^(-c l1)
      });
    });
    wave_sync();
    zd z[8];
    static_for<0, 8>([&](auto ic) {
      constexpr int i = decltype(ic)::value, s = i / 4, h1 = i % 4;
      z[i] = T[t2(a, s, h1, c)];  // lane (a, l1 = c) gathers h1
    });
    wave_sync();
    static_for<0, 2>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      zd t[4];
      dft<4, 1, 1, s * 4, 8>(z, t);  // over h1 -> l2
      static_for<0, 4>([&](auto kc) {
        constexpr int l2 = decltype(kc)::value;
        T[t1(2 * a + s, 4 * l2 + c)] = t[l2];
      });
    });
    wave_sync();
    zd y[8];
    static_for<0, 8>([&](auto gc) {
      constexpr int g0 = decltype(gc)::value;
      y[g0] = g0 == 0 ? T[t1(0, l)] : zmulc(T[t1(g0, l)], fwf(g0));
    });
    wave_sync();
    dft<8, 1, 1, 0, 8>(y, v);
  }

  __host__ __device__ static constexpr int gindex(int l, int j) {
    return (2 * (l >> 2) + j / 4) + 8 * (l & 3) + 32 * (j % 4);
  }
};

}  // namespace tdhf
