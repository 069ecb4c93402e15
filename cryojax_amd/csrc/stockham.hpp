// Stockham autosort FFT building blocks for float2 (FP32) and double2 (FP64) complex values.
//
// An L-point forward FFT (exponent -), L a power of two in 16..512, is a sequence of passes: radix
// 16 while it divides, then the rest (L / 16^p). Butterfly j of a pass with radix R after
// sub-transforms of size Ns takes x[j + q L/R], q < R, multiplies by e^{-2 pi i q k / (Ns R)} with
// k = j mod Ns, runs an R-point DFT in registers and stores output q at (j / Ns) Ns R + k + q Ns;
// the last pass leaves natural order. L/16 threads share one transform, each holding 16 values;
// passes exchange through a padded LDS copy of the transform.
#pragma once

#include <hip/hip_runtime.h>

namespace cryomi {

// cos and sin of 2 pi m / 16
constexpr double ST_C16[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                               0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                               -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                               0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
constexpr double ST_S16[16] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                               1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                               0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                               -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

template <typename C>
struct CScalar;
template <>
struct CScalar<float2> {
  typedef float T;
  __device__ __forceinline__ static float2 make(float x, float y) { return make_float2(x, y); }
};
template <>
struct CScalar<double2> {
  typedef double T;
  __device__ __forceinline__ static double2 make(double x, double y) { return make_double2(x, y); }
};

template <typename C>
__device__ __forceinline__ C zadd(C a, C b) {
  return CScalar<C>::make(a.x + b.x, a.y + b.y);
}
template <typename C>
__device__ __forceinline__ C zsub(C a, C b) {
  return CScalar<C>::make(a.x - b.x, a.y - b.y);
}
// a * (c - i s) = a * e^{-i theta} for c = cos theta, s = sin theta
template <typename C>
__device__ __forceinline__ C zrot(C a, double c, double s) {
  typedef typename CScalar<C>::T T;
  const T cc = (T)c, ss = (T)s;
  return CScalar<C>::make(fma(a.x, cc, a.y * ss), fma(a.y, cc, -(a.x * ss)));
}
// a * w for a table entry w = (cos, -sin)
template <typename C>
__device__ __forceinline__ C zmul(C a, C w) {
  return CScalar<C>::make(fma(a.x, w.x, -(a.y * w.y)), fma(a.x, w.y, a.y * w.x));
}
template <typename C>
__device__ __forceinline__ C zmul_negi(C a) {
  return CScalar<C>::make(a.y, -a.x);
}

// Forward R-point DFT (exponent -) of x[0..R) in registers, natural order in and out.
template <int R>
struct FwdDft;
template <>
struct FwdDft<1> {
  template <typename C>
  __device__ __forceinline__ static void run(C*) {}
};
template <>
struct FwdDft<2> {
  template <typename C>
  __device__ __forceinline__ static void run(C* x) {
    const C a = x[0];
    x[0] = zadd(a, x[1]);
    x[1] = zsub(a, x[1]);
  }
};
template <>
struct FwdDft<4> {
  template <typename C>
  __device__ __forceinline__ static void run(C* x) {
    const C t0 = zadd(x[0], x[2]), t1 = zsub(x[0], x[2]);
    const C t2 = zadd(x[1], x[3]), t3 = zmul_negi(zsub(x[1], x[3]));
    x[0] = zadd(t0, t2);
    x[1] = zadd(t1, t3);
    x[2] = zsub(t0, t2);
    x[3] = zsub(t1, t3);
  }
};
template <int R>
struct FwdDft {
  static constexpr int B = R / 4;  // R = 8 or 16: 4-point DFTs over x[B a + b], twiddle, B-point DFTs
  template <typename C>
  __device__ __forceinline__ static void run(C* x) {
    C y[4][B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      C t[4] = {x[b], x[B + b], x[2 * B + b], x[3 * B + b]};
      FwdDft<4>::run(t);
#pragma unroll
      for (int ka = 0; ka < 4; ++ka) {
        const int m = (16 / R) * b * ka;  // root index in sixteenths
        y[ka][b] = m == 0 ? t[ka] : zrot(t[ka], ST_C16[m], ST_S16[m]);
      }
    }
#pragma unroll
    for (int ka = 0; ka < 4; ++ka) {
      FwdDft<B>::run(y[ka]);
#pragma unroll
      for (int kb = 0; kb < B; ++kb) x[ka + 4 * kb] = y[ka][kb];
    }
  }
};

// radix of pass p (1: no such pass): 16 | 16, L/16 | 16, 16, L/256
__host__ __device__ constexpr int pass_radix(int L, int p) {
  return p == 0 ? 16 : (p == 1 ? (L >= 256 ? 16 : L / 16) : (L >= 256 ? L / 256 : 1));
}
__device__ __forceinline__ int fpad(int p) { return p + (p >> 4); }  // LDS index with one pad per 16 entries

// One Stockham pass on the 16 values a thread holds (16 / R butterflies j = t + u T): twiddle and DFT.
// tw holds e^{-2 pi i m / (TWS L)}.
template <int L, int R, int NS, int TWS, typename C>
__device__ __forceinline__ void stockham_compute(C (&v)[16], int t, const C* __restrict__ tw) {
#pragma unroll
  for (int u = 0; u < 16 / R; ++u) {
    const int j = t + u * (L / 16);
    if (NS > 1) {
      const int k = j % NS;
#pragma unroll
      for (int q = 1; q < R; ++q) v[u * R + q] = zmul(v[u * R + q], tw[q * k * (L / (NS * R)) * TWS]);
    }
    FwdDft<R>::run(v + u * R);
  }
}
// logical input / output indices of value (u, q) in a pass
template <int L, int R>
__device__ __forceinline__ int stockham_in(int t, int u, int q) {
  return t + u * (L / 16) + q * (L / R);
}
template <int L, int R, int NS>
__device__ __forceinline__ int stockham_out(int t, int u, int q) {
  const int j = t + u * (L / 16);
  return (j / NS) * NS * R + (j % NS) + q * NS;
}

template <int L, int R, int NS, typename C>
__device__ __forceinline__ void store_pass(const C (&v)[16], int t, C* __restrict__ z) {
#pragma unroll
  for (int u = 0; u < 16 / R; ++u)
#pragma unroll
    for (int q = 0; q < R; ++q) z[fpad(stockham_out<L, R, NS>(t, u, q))] = v[u * R + q];
}
template <int L, int R, typename C>
__device__ __forceinline__ void load_pass(C (&v)[16], int t, const C* __restrict__ z) {
#pragma unroll
  for (int u = 0; u < 16 / R; ++u)
#pragma unroll
    for (int q = 0; q < R; ++q) v[u * R + q] = z[fpad(stockham_in<L, R>(t, u, q))];
}

// Runs the passes of an L-point FFT on one transform whose pass-0 inputs the caller has loaded
// into v; intermediate passes go through the padded LDS transform z. On return the natural-order
// result is in z (STORE_LAST = false, after a closing barrier the caller issues) or still in v,
// at the last pass's output indices (STORE_LAST = true). Every thread of the workgroup calls it.
template <int L, int TWS, bool STORE_LAST, typename C>
__device__ __forceinline__ void fft_passes(C (&v)[16], int t, C* __restrict__ z, const C* __restrict__ tw) {
  constexpr int R0 = pass_radix(L, 0), R1 = pass_radix(L, 1), R2 = pass_radix(L, 2);
  stockham_compute<L, R0, 1, TWS>(v, t, tw);
  if constexpr (R1 == 1) {
    if constexpr (!STORE_LAST) store_pass<L, R0, 1>(v, t, z);
  } else {
    store_pass<L, R0, 1>(v, t, z);
    __syncthreads();
    load_pass<L, R1>(v, t, z);
    stockham_compute<L, R1, R0, TWS>(v, t, tw);
    if constexpr (R2 == 1) {
      if constexpr (!STORE_LAST) {
        __syncthreads();
        store_pass<L, R1, R0>(v, t, z);
      }
    } else {
      __syncthreads();
      store_pass<L, R1, R0>(v, t, z);
      __syncthreads();
      load_pass<L, R2>(v, t, z);
      stockham_compute<L, R2, R0 * R1, TWS>(v, t, tw);
      if constexpr (!STORE_LAST) {
        __syncthreads();
        store_pass<L, R2, R0 * R1>(v, t, z);
      }
    }
  }
}

template <int L>
constexpr int last_radix() {
  return pass_radix(L, 2) > 1 ? pass_radix(L, 2) : (pass_radix(L, 1) > 1 ? pass_radix(L, 1) : 16);
}
template <int L>
constexpr int last_ns() {
  return pass_radix(L, 2) > 1 ? 256 : (pass_radix(L, 1) > 1 ? 16 : 1);
}

// tw[m] = e^{-2 pi i m / n}, n a multiple of 4, by NTH threads: one FP64 sincospi per quarter
// turn, the other quadrants by exact multiplication with -i, -1 and i.
template <int NTH, typename C>
__device__ __forceinline__ void fill_twiddles(C* tw, int n) {
  typedef typename CScalar<C>::T T;
  const int q = n / 4;
  for (int m = threadIdx.x; m < q; m += NTH) {
    double s, c;
    sincospi(2.0 * (double)m / (double)n, &s, &c);
    const T cc = (T)c, ss = (T)s;
    tw[m] = CScalar<C>::make(cc, -ss);
    tw[m + q] = CScalar<C>::make(-ss, -cc);
    tw[m + 2 * q] = CScalar<C>::make(-cc, ss);
    tw[m + 3 * q] = CScalar<C>::make(ss, cc);
  }
}

}  // namespace cryomi
