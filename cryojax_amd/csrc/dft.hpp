// Register-resident small inverse DFTs (exponent +) for gfx950.
//
// dft<R>(x) transforms R complex values held in registers, natural order in
// and out: X[k] = sum_n x[n] exp(+2 pi i n k / R). Composite sizes use the
// four-step split R = A * B: A-point DFTs over x[B a + b], twiddles
// exp(+2 pi i b ka / R), then B-point DFTs, giving X[ka + A kb]. Every index is
// a compile-time constant after unrolling, so the values stay in VGPRs.
#pragma once

#include <hip/hip_runtime.h>

namespace cryomi {

// exp(+2 pi i m / R), m = 0..R-1, as float (rounded from double).
template <int R>
struct RootTable;
template <>
struct RootTable<6> {
  static constexpr float c[6] = {1.000000000e+00f, 5.000000000e-01f, -5.000000000e-01f, -1.000000000e+00f, -5.000000000e-01f, 5.000000000e-01f};
  static constexpr float s[6] = {0.000000000e+00f, 8.660254038e-01f, 8.660254038e-01f, 1.224646799e-16f, -8.660254038e-01f, -8.660254038e-01f};
};
template <>
struct RootTable<8> {
  static constexpr float c[8] = {1.000000000e+00f, 7.071067812e-01f, 6.123233996e-17f, -7.071067812e-01f, -1.000000000e+00f, -7.071067812e-01f, -1.836970199e-16f, 7.071067812e-01f};
  static constexpr float s[8] = {0.000000000e+00f, 7.071067812e-01f, 1.000000000e+00f, 7.071067812e-01f, 1.224646799e-16f, -7.071067812e-01f, -1.000000000e+00f, -7.071067812e-01f};
};
template <>
struct RootTable<10> {
  static constexpr float c[10] = {1.000000000e+00f, 8.090169944e-01f, 3.090169944e-01f, -3.090169944e-01f, -8.090169944e-01f, -1.000000000e+00f, -8.090169944e-01f, -3.090169944e-01f, 3.090169944e-01f, 8.090169944e-01f};
  static constexpr float s[10] = {0.000000000e+00f, 5.877852523e-01f, 9.510565163e-01f, 9.510565163e-01f, 5.877852523e-01f, 1.224646799e-16f, -5.877852523e-01f, -9.510565163e-01f, -9.510565163e-01f, -5.877852523e-01f};
};
template <>
struct RootTable<12> {
  static constexpr float c[12] = {1.000000000e+00f, 8.660254038e-01f, 5.000000000e-01f, 6.123233996e-17f, -5.000000000e-01f, -8.660254038e-01f, -1.000000000e+00f, -8.660254038e-01f, -5.000000000e-01f, -1.836970199e-16f, 5.000000000e-01f, 8.660254038e-01f};
  static constexpr float s[12] = {0.000000000e+00f, 5.000000000e-01f, 8.660254038e-01f, 1.000000000e+00f, 8.660254038e-01f, 5.000000000e-01f, 1.224646799e-16f, -5.000000000e-01f, -8.660254038e-01f, -1.000000000e+00f, -8.660254038e-01f, -5.000000000e-01f};
};
template <>
struct RootTable<16> {
  static constexpr float c[16] = {1.000000000e+00f, 9.238795325e-01f, 7.071067812e-01f, 3.826834324e-01f, 6.123233996e-17f, -3.826834324e-01f, -7.071067812e-01f, -9.238795325e-01f, -1.000000000e+00f, -9.238795325e-01f, -7.071067812e-01f, -3.826834324e-01f, -1.836970199e-16f, 3.826834324e-01f, 7.071067812e-01f, 9.238795325e-01f};
  static constexpr float s[16] = {0.000000000e+00f, 3.826834324e-01f, 7.071067812e-01f, 9.238795325e-01f, 1.000000000e+00f, 9.238795325e-01f, 7.071067812e-01f, 3.826834324e-01f, 1.224646799e-16f, -3.826834324e-01f, -7.071067812e-01f, -9.238795325e-01f, -1.000000000e+00f, -9.238795325e-01f, -7.071067812e-01f, -3.826834324e-01f};
};

__device__ __forceinline__ float2 f2add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 f2sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 f2mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 f2muli(float2 a) { return make_float2(-a.y, a.x); }

template <int R>
struct Dft {
  static constexpr int A = (R % 4 == 0) ? 4 : ((R % 2 == 0) ? 2 : ((R % 3 == 0) ? 3 : 5));
  static constexpr int B = R / A;
  __device__ __forceinline__ static void run(float2 (&x)[R]) {
    float2 y[A][B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      float2 t[A];
#pragma unroll
      for (int a = 0; a < A; ++a) t[a] = x[B * a + b];
      Dft<A>::run(t);
#pragma unroll
      for (int ka = 0; ka < A; ++ka) {
        const int m = b * ka;  // < R
        if (m == 0) {
          y[ka][b] = t[ka];
        } else {
          y[ka][b] = f2mul(t[ka], make_float2(RootTable<R>::c[m], RootTable<R>::s[m]));
        }
      }
    }
#pragma unroll
    for (int ka = 0; ka < A; ++ka) {
      float2 u[B];
#pragma unroll
      for (int b = 0; b < B; ++b) u[b] = y[ka][b];
      Dft<B>::run(u);
#pragma unroll
      for (int kb = 0; kb < B; ++kb) x[ka + A * kb] = u[kb];
    }
  }
};

template <>
struct Dft<1> {
  __device__ __forceinline__ static void run(float2 (&)[1]) {}
};

template <>
struct Dft<2> {
  __device__ __forceinline__ static void run(float2 (&x)[2]) {
    const float2 a = x[0], b = x[1];
    x[0] = f2add(a, b);
    x[1] = f2sub(a, b);
  }
};

template <>
struct Dft<3> {
  __device__ __forceinline__ static void run(float2 (&x)[3]) {
    const float h = 0.86602540378443864676f;  // sqrt(3) / 2
    const float2 s = f2add(x[1], x[2]);
    const float2 d = f2sub(x[1], x[2]);
    const float2 m = make_float2(x[0].x - 0.5f * s.x, x[0].y - 0.5f * s.y);
    const float2 id = make_float2(-h * d.y, h * d.x);  // i * sqrt(3)/2 * d
    x[0] = f2add(x[0], s);
    x[1] = f2add(m, id);
    x[2] = f2sub(m, id);
  }
};

// X_q = sum_p x_p e^{+2 pi i p q / 5}: with a_1 = x1 + x4, b_1 = x1 - x4, a_2 = x2 + x3,
// b_2 = x2 - x3 and c_m + i s_m = e^{2 pi i m / 5}, X_1,4 = x0 + c1 a_1 + c2 a_2 +- i (s1 b_1 + s2 b_2),
// X_2,3 = x0 + c2 a_1 + c1 a_2 +- i (s2 b_1 - s1 b_2).
template <>
struct Dft<5> {
  __device__ __forceinline__ static void run(float2 (&x)[5]) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    const float2 a1 = f2add(x[1], x[4]), b1 = f2sub(x[1], x[4]), a2 = f2add(x[2], x[3]), b2 = f2sub(x[2], x[3]);
    const float2 t1 = make_float2(x[0].x + c1 * a1.x + c2 * a2.x, x[0].y + c1 * a1.y + c2 * a2.y);
    const float2 t2 = make_float2(x[0].x + c2 * a1.x + c1 * a2.x, x[0].y + c2 * a1.y + c1 * a2.y);
    const float2 u1 = f2muli(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    const float2 u2 = f2muli(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    x[0] = f2add(x[0], f2add(a1, a2));
    x[1] = f2add(t1, u1);
    x[2] = f2add(t2, u2);
    x[3] = f2sub(t2, u2);
    x[4] = f2sub(t1, u1);
  }
};

template <>
struct Dft<4> {
  __device__ __forceinline__ static void run(float2 (&x)[4]) {
    const float2 t0 = f2add(x[0], x[2]), t1 = f2sub(x[0], x[2]);
    const float2 t2 = f2add(x[1], x[3]), t3 = f2muli(f2sub(x[1], x[3]));
    x[0] = f2add(t0, t2);
    x[1] = f2add(t1, t3);
    x[2] = f2sub(t0, t2);
    x[3] = f2sub(t1, t3);
  }
};

}  // namespace cryomi
