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
// Bricked (CRYOMI_GRID_BRICKED): two copies of the grid cut into 2 x 2 x 4
// (z, y, x) bricks of 16 complex64 = one 128-byte cache line, bricks
// row-major over (z/2, y/2, x/4), elements (z&1, y&1, x&3) inside. Copy B
// is the same with x shifted by +2 (N/4 + 1 bricks per row), so the x pair
// (x, x+1) lies inside one brick row of copy A unless x % 4 == 3, and then
// inside one brick row of copy B: every pair is a single 16-byte load and
// the wavefront never diverges on it. A tilted plane touches ~30% fewer
// lines than with x rows.
__host__ __device__ inline size_t brick_copy_b_offset(int n) { return (size_t)n * n * n; }
__host__ __device__ inline size_t bricked_elems(int n) {
  return brick_copy_b_offset(n) + (size_t)(n / 2) * (n / 2) * (n / 4 + 1) * 16;
}
__device__ __forceinline__ size_t brick_index(int z, int y, int xs, int n, int bricks_x) {
  return ((((size_t)(z >> 1) * (size_t)(n >> 1) + (size_t)(y >> 1)) * (size_t)bricks_x + (size_t)(xs >> 2)) << 4) +
         (size_t)(((z & 1) << 3) | ((y & 1) << 2) | (xs & 3));
}
// offset of element (z, y, x) (copy A in the bricked layout)
__device__ __forceinline__ size_t elem_offset(int z, int y, int x, int n, bool bricked) {
  if (!bricked) return ((size_t)z * n + (size_t)y) * n + (size_t)x;
  return brick_index(z, y, x, n, n >> 2);
}
// offset of the pair (z, y, x), (z, y, x + 1): x in [0, n - 2] for the
// standard layout, x in [-1, n - 1] for the bricked one (copy B's padding)
__device__ __forceinline__ size_t pair_offset(int z, int y, int x, int n, bool bricked) {
  if (!bricked) return ((size_t)z * n + (size_t)y) * n + (size_t)x;
  const bool use_b = (x & 3) == 3;
  const int xs = use_b ? x + 2 : x;
  const int bx = use_b ? (n >> 2) + 1 : (n >> 2);
  return (use_b ? brick_copy_b_offset(n) : (size_t)0) + brick_index(z, y, xs, n, bx);
}

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
    // Copy B stores the wrapped x = -1 value and the x = N fill (0) beside the
    // grid edge, so for lx in [-1, N-1] the pair (lx, lx + 1) is exactly the
    // two corners; lx in [-N, -2] wraps both corners to lx + N; lx >= N fills.
    // Without wrapping, the x = -1 corner (lx = -1) and every lx < -1 fill.
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
      st.seam_v[q] = grid[elem_offset(zs[q], ys[q], x0, n, bricked)];
      st.seam_w[q] = wzy[q] * wx0;
    }
  }
  st.i = i;
  st.j = j;
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
  const int i = st.i, j = st.j;
  float2 acc = make_float2(0.f, 0.f);
  // corners accumulate in itertools.product order over (z, y) rows, x lower
  // then upper
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!BRICKED) {  // the wrapped lower x corner at the seam comes first, as in the pair order
      acc.x = fmaf(st.seam_w[q], st.seam_v[q].x, acc.x);
      acc.y = fmaf(st.seam_w[q], st.seam_v[q].y, acc.y);
    }
    acc.x = fmaf(st.w1[q], st.g[q].z, fmaf(st.w0[q], st.g[q].x, acc.x));
    acc.y = fmaf(st.w1[q], st.g[q].w, fmaf(st.w0[q], st.g[q].y, acc.y));
  }
  // self-conjugate bins: real part at extraction (convert_fftn_to_rfftn
  // mode="real"), zeroed before the shift when translating
  return finish_sample(acc, n, i, j, k, flags);
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
