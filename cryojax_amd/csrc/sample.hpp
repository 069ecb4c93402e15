// Slice sampling shared by the image kernels: per-image constants, grid
// layouts and the branch-free trilinear gather (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"

namespace cryomi {

// Per-image constants, derived once per workgroup in FP64.
struct ImageConsts {
  double r00, r01, r10, r11, r20, r21;  // columns 0, 1 of R
  double half;                          // N / 2 (logical offset of zero frequency)
  double ca, cb, cc, cd, cphase;        // CTF phase polynomial in cycles
  double tx, ty;                        // shift phase in cycles per frequency index
  float scale;                          // amplitude (x 1/N^2 for real output)
  bool wrap;                            // corner indices in [-N, -1] wrap (else every index outside [0, N) fills 0)
  bool finish;                          // spectrum input: apply fix-ups, CTF and shift (finish_sample)
};

// `flags` only selects the out-of-range corner rule (CRYOMI_GATHER_FILL).
__device__ __forceinline__ ImageConsts make_consts(const cryomi_image_params& p, int n, bool real_out,
                                                   unsigned flags = 0u) {
  ImageConsts k;
  k.wrap = (flags & CRYOMI_GATHER_FILL) == 0u;
  k.finish = false;
  k.r00 = p.rotation[0];
  k.r01 = p.rotation[1];
  k.r10 = p.rotation[3];
  k.r11 = p.rotation[4];
  k.r20 = p.rotation[6];
  k.r21 = p.rotation[7];
  k.half = (double)(n / 2);
  // chi / (2 pi) = -lambda/2 [df k^2 + astig/2 ((qy^2-qx^2) cos 2a + 2 qx qy sin 2a)]
  //               + Cs lambda^3 k^4 / 4, with q = (j, i') / (N * pixel_size);
  // cos/sin(2 theta) from theta = atan2(qx, qy) are (qy^2-qx^2)/k^2, 2 qx qy/k^2.
  const double inv = 1.0 / ((double)n * p.pixel_size);
  const double s = inv * inv;
  const double lam = p.wavelength;
  const double a2 = 2.0 * p.astigmatism_angle * (M_PI / 180.0);
  const double cs_a = p.spherical_aberration_mm * 1e7;
  k.ca = -0.5 * lam * p.defocus * s;
  k.cb = -0.25 * lam * p.astigmatism * cos(a2) * s;
  k.cc = -0.5 * lam * p.astigmatism * sin(a2) * s;
  k.cd = 0.25 * cs_a * lam * lam * lam * s * s;
  k.cphase = p.phase_shift / (2.0 * M_PI);
  k.tx = p.offset_x * inv;
  k.ty = p.offset_y * inv;
  double sc = p.amplitude;
  if (real_out) sc /= (double)n * (double)n;
  k.scale = (float)sc;
  return k;
}

// 16-byte vector of two complex64 values at 8-byte alignment.
typedef float float4a8 __attribute__((ext_vector_type(4), aligned(8)));

// Grid layouts.
// Standard: element (z, y, x) at (z N + y) N + x.
// Bricked (CRYOMI_GRID_BRICKED): the grid cut into 2 x 2 x 4 (z, y, x) bricks
// of 16 complex64 = one 128-byte cache line, bricks row-major over (z/2, y/2,
// b), elements (z&1, y&1, e) inside, where brick b of a row holds the four x
// entries x = 3b - 1 + e, e = 0..3: consecutive bricks overlap by one entry,
// so every x pair (x, x + 1), x in [-1, N - 1], lies inside brick (x + 1) / 3
// and is a single 16-byte load (the wavefront never diverges on it). The
// entry x = -1 holds the grid's own wrap-around x = N - 1 (JAX normalises
// indices in [-N, -1] to i + N) and entries x >= N hold the fill value 0.
// N/3 + 1 bricks per row: 4/3 of the grid's bytes, and a tilted plane touches
// about 23% fewer lines than with two copies of non-overlapping bricks.
__host__ __device__ constexpr int brick_row_bricks(int n) { return n / 3 + 1; }
__host__ __device__ inline size_t bricked_elems(int n) { return (size_t)(n / 2) * (n / 2) * brick_row_bricks(n) * 16; }
// offset of entry e of brick b in the (z, y) row
__host__ __device__ __forceinline__ size_t brick_index(int z, int y, int b, int e, int n) {
  return ((((size_t)(z >> 1) * (size_t)(n >> 1) + (size_t)(y >> 1)) * (size_t)brick_row_bricks(n) + (size_t)b) << 4) +
         (size_t)(((z & 1) << 3) | ((y & 1) << 2) | e);
}
// offset of the pair (z, y, x), (z, y, x + 1): x in [0, n - 2] for the
// standard layout, x in [-1, n - 1] for the bricked one
__device__ __forceinline__ size_t pair_offset(int z, int y, int x, int n, bool bricked) {
  if (!bricked) return ((size_t)z * n + (size_t)y) * n + (size_t)x;
  const int xs = x + 1;
  const int b = xs / 3;
  return brick_index(z, y, b, xs - 3 * b, n);
}
// offset of element (z, y, x), x in [0, n) (standard layout)
__device__ __forceinline__ size_t elem_offset(int z, int y, int x, int n) {
  return ((size_t)z * n + (size_t)y) * n + (size_t)x;
}

// Byte strides of the bricked layout for a box known at compile time (the
// register former's boxes: the grid stays below 2^31 bytes, so offsets are
// 32-bit). The offset of the pair at corner row (z, y) is
// zterm(z) + yterm(y) + xterm(x).
template <int N>
struct BrickGeom {
  static constexpr unsigned SY = (unsigned)brick_row_bricks(N) * 128u;  // y pair to y pair
  static constexpr unsigned SZ = (unsigned)(N / 2) * SY;                // z pair to z pair
  static_assert((unsigned long long)(N / 2) * SZ < (1ull << 31), "bricked box too large for 32-bit offsets");
  // (z >> 1) SZ + (z & 1) 64 = z SZ/2 + (z & 1) (64 - SZ/2), modulo 2^32
  static __device__ __forceinline__ unsigned zterm(unsigned z) {
    return z * (SZ / 2u) + (z & 1u) * (64u - SZ / 2u);
  }
  // (y >> 1) SY + (y & 1) 32
  static __device__ __forceinline__ unsigned yterm(unsigned y) {
    return y * (SY / 2u) + (y & 1u) * (32u - SY / 2u);
  }
  // brick b = (x + 1) / 3, entry e = (x + 1) - 3 b: b 128 + e 8, x in [-1, N - 1]
  static __device__ __forceinline__ unsigned xterm(int x) {
    const unsigned xs = (unsigned)(x + 1);
    const unsigned b = xs / 3u;
    return b * 128u + (xs - 3u * b) * 8u;
  }
};

// One corner of the trilinear stencil. With `wrap` (JAX's gather with
// mode="fill": negative indices are normalised first), indices in [-N, -1]
// wrap to i + N and anything else outside [0, N) reads 0; without it every
// index outside [0, N) reads 0.
__device__ __forceinline__ bool wrap_index(int idx, int n, int& out, bool wrap = true) {
  int w = (wrap && idx < 0) ? idx + n : idx;
  out = w;
  return (unsigned)w < (unsigned)n;
}

// Sample (i, j) of the rfft half-plane: i in [0, N) rows (y), j in [0, N/2]
// cols. Split in two so callers can put several samples' loads in flight:
// gather_issue computes the stencil and issues the loads, gather_finish
// accumulates and applies fix-ups, CTF and shift.
struct Gather {
  float4a8 g[4];  // x pairs of the four (z, y) rows
  float w0[4], w1[4];
  float2 seam_v[4];  // standard layout only: the wrapped N-1 corner
  float seam_w[4];
  int i, j;
};

// The trilinear stencil at logical coordinates (cz, cy, cx): corner weights
// and the four 16-byte x-pair loads, issued together.
template <bool BRICKED>
__device__ __forceinline__ void stencil_issue(const float2* __restrict__ grid, int n, double cz, double cy, double cx,
                                              const ImageConsts& k, Gather& st) {
  const double fz0 = floor(cz), fy0 = floor(cy), fx0 = floor(cx);
  const int lz = (int)fz0, ly = (int)fy0, lx = (int)fx0;
  const float wz1 = (float)(cz - fz0), wy1 = (float)(cy - fy0), wx1 = (float)(cx - fx0);
  const float wz0 = 1.0f - wz1, wy0 = 1.0f - wy1, wx0 = 1.0f - wx1;
  int z0, z1, y0, y1, x0, x1;
  const bool vz0 = wrap_index(lz, n, z0, k.wrap), vz1 = wrap_index(lz + 1, n, z1, k.wrap);
  const bool vy0 = wrap_index(ly, n, y0, k.wrap), vy1 = wrap_index(ly + 1, n, y1, k.wrap);
  const bool vx0 = wrap_index(lx, n, x0, k.wrap), vx1 = wrap_index(lx + 1, n, x1, k.wrap);
  // Branch-free gather: every (z, y) row issues one 16-byte load of an x
  // pair at an always-valid address, and out-of-range corners get a zero
  // weight, so the loads of a sample are in flight together.
  const bool vzy[4] = {vz0 && vy0, vz0 && vy1, vz1 && vy0, vz1 && vy1};
  const int zs[4] = {vz0 ? z0 : 0, vz0 ? z0 : 0, vz1 ? z1 : 0, vz1 ? z1 : 0};
  const int ys[4] = {vy0 ? y0 : 0, vy1 ? y1 : 0, vy0 ? y0 : 0, vy1 ? y1 : 0};
  const float wzy[4] = {vzy[0] ? wz0 * wy0 : 0.f, vzy[1] ? wz0 * wy1 : 0.f, vzy[2] ? wz1 * wy0 : 0.f,
                        vzy[3] ? wz1 * wy1 : 0.f};
  constexpr bool bricked = BRICKED;
  float wlo, whi;
  int xq;
  bool seam = false;
  if (BRICKED) {
    // The bricked rows store the wrapped x = -1 value and the x = N fill (0)
    // beside the grid edge, so for lx in [-1, N-1] the pair (lx, lx + 1) is
    // exactly the two corners; lx in [-N, -2] wraps both corners to lx + N;
    // lx >= N fills. Without wrapping, the x = -1 corner (lx = -1) and every
    // lx < -1 fill.
    const bool inside = lx >= -1 && lx <= n - 1;
    const bool wraps = k.wrap && lx < -1 && lx >= -n;
    xq = inside ? lx : (wraps ? lx + n : 0);
    wlo = ((inside && (k.wrap || lx >= 0)) || wraps) ? wx0 : 0.f;
    whi = (inside || wraps) ? wx1 : 0.f;
  } else {
    // The pair base covers both x corners except at the wrap seam (x0 -> N-1,
    // x1 -> 0), whose N-1 corner is fetched separately.
    const bool split = vx0 && vx1 && x1 != x0 + 1;  // x0 = N-1 (wrapped), x1 = 0
    xq = split ? 0 : (vx0 ? (x0 < n - 1 ? x0 : n - 2) : (vx1 ? (x1 > 0 ? x1 - 1 : 0) : 0));
    const bool c0lo = vx0 && x0 == xq, c0hi = vx0 && x0 == xq + 1;
    const bool c1lo = vx1 && x1 == xq, c1hi = vx1 && x1 == xq + 1;
    wlo = (c0lo ? wx0 : 0.f) + (c1lo ? wx1 : 0.f);
    whi = (c0hi ? wx0 : 0.f) + (c1hi ? wx1 : 0.f);
    seam = vx0 && !c0lo && !c0hi;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    st.g[q] = *reinterpret_cast<const float4a8*>(grid + pair_offset(zs[q], ys[q], xq, n, bricked));
    st.w0[q] = wzy[q] * wlo;
    st.w1[q] = wzy[q] * whi;
    if (!BRICKED) {
      st.seam_w[q] = 0.f;
      st.seam_v[q] = make_float2(0.f, 0.f);
    }
  }
  if (!BRICKED && seam) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      st.seam_v[q] = grid[elem_offset(zs[q], ys[q], x0, n)];
      st.seam_w[q] = wzy[q] * wx0;
    }
  }
}

template <bool BRICKED>
__device__ __forceinline__ void gather_issue(const float2* __restrict__ grid, int n, int i, int j,
                                             const ImageConsts& k, Gather& st) {
  const int m = n / 2;
  const int jx = (j < m) ? j : j - n;  // fftfreq(N)[j] * N: Nyquist column samples -N/2
  const int iy = (i < m) ? i : i - n;
  const double fx = (double)jx, fy = (double)iy;
  // logical coordinates c = R k N + N/2, gathered as grid[c_z][c_y][c_x]
  const double cz = fma(k.r20, fx, fma(k.r21, fy, k.half));
  const double cy = fma(k.r10, fx, fma(k.r11, fy, k.half));
  const double cx = fma(k.r00, fx, fma(k.r01, fy, k.half));
  stencil_issue<BRICKED>(grid, n, cz, cy, cx, k, st);
  st.i = i;
  st.j = j;
}

// A sample of the bricked grid in flight: the four x-pair loads, the three
// fractional weights and which corners lie in range (BrickGather::Z0 ...).
// The corner weights are formed when the sample is summed (brick_sum), with
// the same expressions stencil_issue uses, so the sums are the same.
struct BrickGather {
  float4a8 g[4];  // x pairs of the (z0, y0), (z0, y1), (z1, y0), (z1, y1) rows
  float wz1, wy1, wx1;
  unsigned ok;  // Z0 | Z1 | Y0 | Y1 | XLO | XHI
  static constexpr unsigned Z0 = 1u, Z1 = 2u, Y0 = 4u, Y1 = 8u, XLO = 16u, XHI = 32u;
};

// stencil_issue for the bricked layout of a box known at compile time, with
// 32-bit byte offsets from the grid base: row offsets from BrickGeom's z, y
// and x terms. `lo` and `range` carry the corner rule: an index t is in range
// when (unsigned)(t - lo) < range (lo = -N, range = 2N when indices in
// [-N, -1] wrap; lo = 0, range = N when they fill); a wrapped index is t mod N.
template <int N>
__device__ __forceinline__ void brick_issue(const char* __restrict__ gb, double cz, double cy, double cx, bool wrap,
                                            int lo, unsigned range, BrickGather& st) {
  using G = BrickGeom<N>;
  const double fz0 = floor(cz), fy0 = floor(cy), fx0 = floor(cx);
  const int lz = (int)fz0, ly = (int)fy0, lx = (int)fx0;
  st.wz1 = (float)(cz - fz0);
  st.wy1 = (float)(cy - fy0);
  st.wx1 = (float)(cx - fx0);
  const unsigned tz = (unsigned)(lz - lo), ty = (unsigned)(ly - lo);
  // t mod N for t in [-N, N); indices out of range get any row in [0, N) (zero weight)
  auto wrapped = [](int t) -> unsigned {
    if ((N & (N - 1)) == 0) return (unsigned)t & (unsigned)(N - 1);
    const int w = t < 0 ? t + N : t;
    return (unsigned)min(max(w, 0), N - 1);
  };
  const unsigned zt0 = G::zterm(wrapped(lz)), zt1 = G::zterm(wrapped(lz + 1));
  const unsigned yt0 = G::yterm(wrapped(ly)), yt1 = G::yterm(wrapped(ly + 1));
  // x, as in stencil_issue's bricked branch
  const bool inside = lx >= -1 && lx <= N - 1;
  const bool wraps = wrap && lx < -1 && lx >= -N;
  const int xq = inside ? lx : (wraps ? lx + N : 0);
  st.ok = (tz < range ? BrickGather::Z0 : 0u) | (tz + 1u < range ? BrickGather::Z1 : 0u) |
          (ty < range ? BrickGather::Y0 : 0u) | (ty + 1u < range ? BrickGather::Y1 : 0u) |
          (((inside && (wrap || lx >= 0)) || wraps) ? BrickGather::XLO : 0u) |
          ((inside || wraps) ? BrickGather::XHI : 0u);
  const unsigned xt = G::xterm(xq);
  st.g[0] = *reinterpret_cast<const float4a8*>(gb + (zt0 + yt0 + xt));
  st.g[1] = *reinterpret_cast<const float4a8*>(gb + (zt0 + yt1 + xt));
  st.g[2] = *reinterpret_cast<const float4a8*>(gb + (zt1 + yt0 + xt));
  st.g[3] = *reinterpret_cast<const float4a8*>(gb + (zt1 + yt1 + xt));
}

// The trilinear sum of a BrickGather (stencil_sum's corner order and weights).
__device__ __forceinline__ float2 brick_sum(const BrickGather& st) {
  const float wz0 = 1.0f - st.wz1, wy0 = 1.0f - st.wy1, wx0 = 1.0f - st.wx1;
  const unsigned ok = st.ok;
  const bool vz0 = ok & BrickGather::Z0, vz1 = ok & BrickGather::Z1;
  const bool vy0 = ok & BrickGather::Y0, vy1 = ok & BrickGather::Y1;
  const float wzy[4] = {(vz0 && vy0) ? wz0 * wy0 : 0.f, (vz0 && vy1) ? wz0 * st.wy1 : 0.f,
                        (vz1 && vy0) ? st.wz1 * wy0 : 0.f, (vz1 && vy1) ? st.wz1 * st.wy1 : 0.f};
  const float wlo = (ok & BrickGather::XLO) ? wx0 : 0.f;
  const float whi = (ok & BrickGather::XHI) ? st.wx1 : 0.f;
  float2 acc = make_float2(0.f, 0.f);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float w0 = wzy[q] * wlo, w1 = wzy[q] * whi;
    acc.x = fmaf(w1, st.g[q].z, fmaf(w0, st.g[q].x, acc.x));
    acc.y = fmaf(w1, st.g[q].w, fmaf(w0, st.g[q].y, acc.y));
  }
  return acc;
}

// The trilinear sum of a gathered stencil: corners in itertools.product order
// over (z, y) rows, x lower then upper.
template <bool BRICKED>
__device__ __forceinline__ float2 stencil_sum(const Gather& st) {
  float2 acc = make_float2(0.f, 0.f);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!BRICKED) {  // the wrapped lower x corner at the seam comes first, as in the pair order
      acc.x = fmaf(st.seam_w[q], st.seam_v[q].x, acc.x);
      acc.y = fmaf(st.seam_w[q], st.seam_v[q].y, acc.y);
    }
    acc.x = fmaf(st.w1[q], st.g[q].z, fmaf(st.w0[q], st.g[q].x, acc.x));
    acc.y = fmaf(st.w1[q], st.g[q].w, fmaf(st.w0[q], st.g[q].y, acc.y));
  }
  return acc;
}

// sin(pi q) for |q| <= 1/2: the Taylor series of sin through the q^11 term,
// whose remainder is below 5.7e-8 there (pi^13 / 13! / 2^13).
__device__ __forceinline__ float sinpi_half(float q) {
  const float q2 = q * q;
  float p = -7.370430945714348e-03f;       // -pi^11 / 11!
  p = fmaf(p, q2, 8.214588661112819e-02f);  // pi^9 / 9!
  p = fmaf(p, q2, -5.992645293207919e-01f); // -pi^7 / 7!
  p = fmaf(p, q2, 2.550164039877345e+00f);  // pi^5 / 5!
  p = fmaf(p, q2, -5.167712780049969e+00f); // -pi^3 / 3!
  p = fmaf(p, q2, 3.141592653589793e+00f);  // pi
  return p * q;
}

// Per-column constants of a thread that forms the samples of one rfft column
// j in [1, N/2) (the register former's phase A): the coordinate bases, the CTF
// phase as a polynomial in the row frequency, and the column's factor of the
// separable shift phase. The self-conjugate fix-ups never apply there.
struct ColumnConsts {
  double bz, by, bx;  // R[:, 0] j + N/2
  double p0, p1, p2;  // twice the CTF phase in cycles: p0 + p1 fy + p2 fy^2 + p4 fy^4
  float2 ph;          // (-1)^j exp(-2 pi i tx j) amplitude (/ N^2): fftshift sign, shift and scale
};

// Twice the CTF phase chi / (2 pi) is
//   2 [ca kk + cb (fy^2 - j^2) + cc fy j + cd kk^2 - cphase],  kk = j^2 + fy^2,
// = p0 + p1 fy + p2 fy^2 + p4 fy^4 with the coefficients below (FP64; j the
// rfftfreq index, fy the fftfreq row index).
__device__ __forceinline__ double ctf_p4(const ImageConsts& k) { return 2.0 * k.cd; }

__device__ __forceinline__ ColumnConsts column_consts(const ImageConsts& k, int j, unsigned flags) {
  ColumnConsts c;
  const double fx = (double)j;
  c.bz = fma(k.r20, fx, k.half);
  c.by = fma(k.r10, fx, k.half);
  c.bx = fma(k.r00, fx, k.half);
  const double j2 = fx * fx;
  c.p0 = 2.0 * fma(k.cd, j2 * j2, fma(k.ca - k.cb, j2, -k.cphase));
  c.p1 = 2.0 * (k.cc * fx);
  c.p2 = 2.0 * fma(2.0 * k.cd, j2, k.ca + k.cb);
  float sc = (j & 1) ? -k.scale : k.scale;
  c.ph = make_float2(sc, 0.f);
  if (flags & CRYOMI_APPLY_TRANSLATION) {
    const double v = k.tx * fx;
    const double r = v - rint(v);
    float sn, cs;
    sincospif(2.0f * (float)r, &sn, &cs);
    c.ph = make_float2(sc * cs, -sc * sn);  // exp(-2 pi i v) = cs - i sn
  }
  return c;
}

// The row factor (-1)^i exp(-2 pi i ty fy) of the separable shift phase.
__device__ __forceinline__ float2 row_phase_factor(const ImageConsts& k, int n, int i, unsigned flags) {
  const float s = (i & 1) ? -1.f : 1.f;
  if (!(flags & CRYOMI_APPLY_TRANSLATION)) return make_float2(s, 0.f);
  const int m = n / 2;
  const double v = k.ty * (double)((i < m) ? i : i - n);
  const double r = v - rint(v);
  float sn, cs;
  sincospif(2.0f * (float)r, &sn, &cs);
  return make_float2(s * cs, -s * sn);
}

// CTF and shift of a phase-A sample sum at row frequency fy: acc sin(chi)
// ph(j) ph(i), with ph(i) the row factor. Equals gather_finish with the
// fftshift sign up to rounding (the CTF sine by sinpi_half of the FP64-reduced
// phase, the shift phase as the product of its row and column factors).
__device__ __forceinline__ float2 column_finish(float2 acc, const ColumnConsts& c, double p4, double fy, float2 rowph,
                                                unsigned flags) {
  if (flags & CRYOMI_APPLY_CTF) {
    const double t = fy * fy;
    const double w = fma(t, fma(p4, t, c.p2), fma(c.p1, fy, c.p0));  // 2 chi / (2 pi)
    const double nw = rint(w);
    float s = sinpi_half((float)(w - nw));  // sin(pi w) = (-1)^nw sin(pi (w - nw))
    if ((int)nw & 1) s = -s;
    acc.x *= s;
    acc.y *= s;
  }
  const float2 p = cmul(c.ph, rowph);
  return cmul(acc, p);
}

// Fix-ups, CTF and shift of rfft sample (i, j) of an N-point half-plane: the
// self-conjugate bins (N/2, 0), (0, N/2), (N/2, N/2) keep their real part (zero
// when translating), then the amplitude / CTF factor and exp(-2 pi i q . t).
__device__ __forceinline__ float2 finish_sample(float2 acc, int n, int i, int j, const ImageConsts& k,
                                                unsigned flags) {
  const int m = n / 2;
  if ((i == m || i == 0) && (j == 0 || j == m) && !(i == 0 && j == 0)) {
    acc.y = 0.f;
    if (flags & CRYOMI_APPLY_TRANSLATION) acc.x = 0.f;
  }
  const double fy = (double)((i < m) ? i : i - n);
  float sc = k.scale;
  if (flags & CRYOMI_APPLY_CTF) {
    const double jj = (double)j;  // CTF uses rfftfreq: +1/2 at the Nyquist column
    const double j2 = jj * jj, i2 = fy * fy, ij = fy * jj;
    const double kk = j2 + i2;
    const double u = fma(k.ca, kk, fma(k.cb, i2 - j2, fma(k.cc, ij, fma(k.cd, kk * kk, -k.cphase))));
    const double r = u - rint(u);
    sc *= sinpif(2.0f * (float)r);
  }
  acc.x *= sc;
  acc.y *= sc;
  if (flags & CRYOMI_APPLY_TRANSLATION) {
    const double v = fma(k.tx, (double)j, k.ty * fy);
    const double r = v - rint(v);
    float sn, cs;
    sincospif(2.0f * (float)r, &sn, &cs);
    // multiply by exp(-2 pi i v) = cs - i sn
    const float2 a = acc;
    acc.x = a.x * cs + a.y * sn;
    acc.y = a.y * cs - a.x * sn;
  }
  return acc;
}

template <bool BRICKED>
__device__ __forceinline__ float2 gather_finish(int n, const ImageConsts& k, unsigned flags, const Gather& st) {
  // self-conjugate bins: real part at extraction (convert_fftn_to_rfftn
  // mode="real"), zeroed before the shift when translating
  return finish_sample(stencil_sum<BRICKED>(st), n, st.i, st.j, k, flags);
}

__device__ __forceinline__ float2 slice_sample(const float2* __restrict__ grid, int n, int i, int j,
                                               const ImageConsts& k, unsigned flags) {
  Gather st;
  if (flags & CRYOMI_GRID_BRICKED) {  // uniform per launch
    gather_issue<true>(grid, n, i, j, k, st);
    return gather_finish<true>(n, k, flags, st);
  }
  gather_issue<false>(grid, n, i, j, k, st);
  return gather_finish<false>(n, k, flags, st);
}


}  // namespace cryomi
