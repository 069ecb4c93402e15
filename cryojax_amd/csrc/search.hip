// Exhaustive NCC pose search on gfx950: FP32 MFMA contraction of CTF- and
// shift-folded observed spectra against slice templates, fused NCC + top-2
// epilogue, candidate selection and FP64 re-scoring of near-ties.
//
// Kernels
//   prepare_obs_kernel  : one row per (image, offset): A = w Y CTF conj(shift)
//                         as interleaved (re, im) float32, K = 2 N (N/2 + 1)
//                         padded to a multiple of 16.
//   prepare_ctf_kernel  : one row per image: D = w CTF^2 (duplicated per re/im
//                         slot) and yn2 = sum w |Y|^2 (FP64).
//   template_kernel     : slice templates P (FP32 trilinear, no CTF / shift)
//                         for a chunk of rotations, same K layout.
//   gemm_nt_kernel      : 128 x 128 tiles, K steps of 16 through LDS (double
//                         buffered), 4 waves x (4 x 4) v_mfma_f32_16x16x4f32.
//                         Modes: plain store; D . (P o P)^T = sum w CTF^2 |P|^2;
//                         A . P^T with the NCC epilogue keeping, per (row,
//                         128-column block), the two best (score, rotation).
//   select_kernel       : per image, the FP32 best and every grid point within
//                         the tie window (whole blocks whose second best is in
//                         the window).
//   rescore_kernel      : per image, FP64 NCC of every candidate from the
//                         complex128 grid; lowest raveled index wins ties.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <string>

#include "../../include/cryomi_search.h"
#include "common.hpp"
#include "sample.hpp"

namespace cryomi {
namespace {

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int KT = 16;
constexpr int GEMM_THREADS = 256;
constexpr int LDS_STRIDE = BM + 16;  // 144 floats: the 4 k rows of one MFMA operand hit disjoint banks
constexpr int ROT_CHUNK = 4096;
constexpr int CAND_CAP = 1024;
constexpr int NT = 256;

typedef float floatx4 __attribute__((ext_vector_type(4)));

struct Top2 {
  float s1;
  int c1;
  float s2;
  int c2;
};

__device__ __forceinline__ bool before(float sa, int ca, float sb, int cb) {
  return sa < sb || (sa == sb && ca < cb);
}

__device__ __forceinline__ void top2_push(Top2& t, float s, int c) {
  if (before(s, c, t.s1, t.c1)) {
    t.s2 = t.s1;
    t.c2 = t.c1;
    t.s1 = s;
    t.c1 = c;
  } else if (before(s, c, t.s2, t.c2)) {
    t.s2 = s;
    t.c2 = c;
  }
}

__device__ __forceinline__ Top2 top2_empty() { return Top2{INFINITY, INT_MAX, INFINITY, INT_MAX}; }

// rfft half-plane weight of bin (i, j): 1 on columns 0 and M, 2 elsewhere, 0
// at DC and at the self-conjugate bins (M, 0), (0, M), (M, M).
__device__ __forceinline__ double bin_weight(int i, int j, int m) {
  if ((i == 0 || i == m) && (j == 0 || j == m)) return 0.0;
  return (j == 0 || j == m) ? 1.0 : 2.0;
}

// CTF of an image at bin (i, j) in FP64 (same polynomial as the forward).
__device__ __forceinline__ double ctf_f64(const ImageConsts& k, int i, int j, int n) {
  const int m = n / 2;
  const double fy = (double)((i < m) ? i : i - n);
  const double jj = (double)j;
  const double j2 = jj * jj, i2 = fy * fy, ij = fy * jj;
  const double kk = j2 + i2;
  const double u = fma(k.ca, kk, fma(k.cb, i2 - j2, fma(k.cc, ij, fma(k.cd, kk * kk, -k.cphase))));
  return sinpi(2.0 * (u - rint(u)));
}

// exp(+2 pi i (q . t)) = conj of the shift operator, in FP64.
__device__ __forceinline__ double2 conj_shift_f64(double tx, double ty, int i, int j, int n) {
  const int m = n / 2;
  const double fy = (double)((i < m) ? i : i - n);
  const double v = fma(tx, (double)j, ty * fy);
  double s, c;
  sincospi(2.0 * (v - rint(v)), &s, &c);
  return make_double2(c, s);
}

__device__ __forceinline__ double block_sum_f64(double v, double* red) {
  const int tid = threadIdx.x;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(NT) void fill_zero_kernel(float* p, size_t n) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < n; e += (size_t)gridDim.x * NT) p[e] = 0.f;
}

// A[row = img * n_off + off][k] = w Y CTF conj(shift)   (rows >= n_rows are zero)
__global__ __launch_bounds__(NT) void prepare_obs_kernel(const double2* __restrict__ y, int n, int n_img, int n_off,
                                                         const cryomi_image_params* __restrict__ prm,
                                                         const double* __restrict__ offsets, int kp2,
                                                         float2* __restrict__ a) {
  const int row = blockIdx.x;
  const int m = n / 2;
  const int nh = n * (m + 1);
  float2* dst = a + (size_t)row * kp2;
  if (row >= n_img * n_off) {
    for (int k = threadIdx.x; k < kp2; k += NT) dst[k] = make_float2(0.f, 0.f);
    return;
  }
  const int img = row / n_off, off = row % n_off;
  const ImageConsts kc = make_consts(prm[img], n, false);
  const double inv = 1.0 / ((double)n * prm[img].pixel_size);
  const double tx = offsets[2 * off] * inv, ty = offsets[2 * off + 1] * inv;
  const double2* yi = y + (size_t)img * nh;
  for (int k = threadIdx.x; k < kp2; k += NT) {
    float2 v = make_float2(0.f, 0.f);
    if (k < nh) {
      const int i = k / (m + 1), j = k % (m + 1);
      const double w = bin_weight(i, j, m);
      if (w != 0.0) {
        const double c = w * ctf_f64(kc, i, j, n);
        const double2 s = conj_shift_f64(tx, ty, i, j, n);
        const double2 yy = yi[k];
        v = make_float2((float)(c * (yy.x * s.x - yy.y * s.y)), (float)(c * (yy.x * s.y + yy.y * s.x)));
      }
    }
    dst[k] = v;
  }
}

// D[img][2k] = D[img][2k+1] = w CTF^2; yn2[img] = sum w |Y|^2
__global__ __launch_bounds__(NT) void prepare_ctf_kernel(const double2* __restrict__ y, int n, int n_img,
                                                         const cryomi_image_params* __restrict__ prm, int kp2,
                                                         float2* __restrict__ d, double* __restrict__ yn2) {
  __shared__ double red[NT / 64];
  const int img = blockIdx.x;
  const int m = n / 2;
  const int nh = n * (m + 1);
  float2* dst = d + (size_t)img * kp2;
  if (img >= n_img) {
    for (int k = threadIdx.x; k < kp2; k += NT) dst[k] = make_float2(0.f, 0.f);
    return;
  }
  const ImageConsts kc = make_consts(prm[img], n, false);
  const double2* yi = y + (size_t)img * nh;
  double acc = 0.0;
  for (int k = threadIdx.x; k < kp2; k += NT) {
    float v = 0.f;
    if (k < nh) {
      const int i = k / (m + 1), j = k % (m + 1);
      const double w = bin_weight(i, j, m);
      if (w != 0.0) {
        const double c = ctf_f64(kc, i, j, n);
        v = (float)(w * c * c);
        const double2 yy = yi[k];
        acc += w * (yy.x * yy.x + yy.y * yy.y);
      }
    }
    dst[k] = make_float2(v, v);
  }
  acc = block_sum_f64(acc, red);
  if (threadIdx.x == 0) yn2[img] = acc;
}

__device__ __forceinline__ ImageConsts rotation_consts(const double* r, int n) {
  ImageConsts k{};
  k.r00 = r[0];
  k.r01 = r[1];
  k.r10 = r[3];
  k.r11 = r[4];
  k.r20 = r[6];
  k.r21 = r[7];
  k.half = (double)(n / 2);
  k.scale = 1.0f;
  return k;
}

// B[r][k] = P_rot(k) for rotations [rot0, rot0 + count); rows >= count are zero
__global__ __launch_bounds__(NT) void template_kernel(const float2* __restrict__ grid, int n,
                                                     const double* __restrict__ rotations, int rot0, int count,
                                                     int kp2, float2* __restrict__ b) {
  const int r = blockIdx.y;
  const int k = blockIdx.x * NT + threadIdx.x;
  if (k >= kp2) return;
  const int m = n / 2;
  const int nh = n * (m + 1);
  float2 v = make_float2(0.f, 0.f);
  if (r < count && k < nh) {
    const ImageConsts kc = rotation_consts(rotations + 9 * (size_t)(rot0 + r), n);
    v = slice_sample(grid, n, k / (m + 1), k % (m + 1), kc, 0u);
  }
  b[(size_t)r * kp2 + k] = v;
}

enum { GEMM_STORE = 0, GEMM_SQUARE_B = 1, GEMM_NCC = 2 };

struct NccEpilogue {
  int n_off;
  const double* yn2;   // [n_img]
  const float* tn2;    // [n_img][ld_tn2]: sum w CTF^2 |P|^2 of this chunk
  int ld_tn2;
  int col_base;        // global rotation index of column 0
  int cb_base;         // global column-block index of block column 0
  int n_cb;            // column blocks in the whole search
  Top2* partial;       // [rows][n_cb]
};

// C = A . B^T (A: M x K, B: N x K, row-major, K % 16 == 0).
template <int MODE>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_nt_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                               int M, int N, int K, float* __restrict__ C, int ldc,
                                                               NccEpilogue ep) {
  __shared__ float As[2][KT][LDS_STRIDE];
  __shared__ float Bs[2][KT][LDS_STRIDE];
  __shared__ Top2 red[2][BM];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int row0 = blockIdx.y * BM, col0 = blockIdx.x * BN;
  // global -> LDS mapping: tile row tid / 2, k segment (tid & 1) * 8
  const int lr = tid >> 1, lk = (tid & 1) * 8;
  const bool aval = row0 + lr < M, bval = col0 + lr < N;
  const float* ap = A + (size_t)(aval ? row0 + lr : 0) * K + lk;
  const float* bp = B + (size_t)(bval ? col0 + lr : 0) * K + lk;
  floatx4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  float ra[8], rb[8];
  auto gload = [&](int k0) {
    const float4 a0 = *reinterpret_cast<const float4*>(ap + k0);
    const float4 a1 = *reinterpret_cast<const float4*>(ap + k0 + 4);
    const float4 b0 = *reinterpret_cast<const float4*>(bp + k0);
    const float4 b1 = *reinterpret_cast<const float4*>(bp + k0 + 4);
    const float am = aval ? 1.f : 0.f, bm = bval ? 1.f : 0.f;
    ra[0] = a0.x * am; ra[1] = a0.y * am; ra[2] = a0.z * am; ra[3] = a0.w * am;
    ra[4] = a1.x * am; ra[5] = a1.y * am; ra[6] = a1.z * am; ra[7] = a1.w * am;
    rb[0] = b0.x * bm; rb[1] = b0.y * bm; rb[2] = b0.z * bm; rb[3] = b0.w * bm;
    rb[4] = b1.x * bm; rb[5] = b1.y * bm; rb[6] = b1.z * bm; rb[7] = b1.w * bm;
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      As[buf][lk + q][lr] = ra[q];
      Bs[buf][lk + q][lr] = (MODE == GEMM_SQUARE_B) ? rb[q] * rb[q] : rb[q];
    }
  };
  const int nk = K / KT;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * KT);
#pragma unroll
    for (int ks = 0; ks < KT / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[cur][kk][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[cur][kk][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }
  // D element (row 4 (lane / 16) + r, column lane % 16) of each 16 x 16 tile
  if (MODE != GEMM_NCC) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row0 + wm * 64 + i * 16 + 4 * (lane >> 4) + r;
          const int col = col0 + wn * 64 + j * 16 + (lane & 15);
          if (row < M && col < N) C[(size_t)row * ldc + col] = acc[i][j][r];
        }
    return;
  }
  // NCC epilogue: score = -num / sqrt(yn2 * tn2), top-2 per row over this block's columns
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rl = wm * 64 + i * 16 + 4 * (lane >> 4) + r;
      const int row = row0 + rl;
      Top2 t = top2_empty();
      if (row < M) {
        const int img = row / ep.n_off;
        const double y2 = ep.yn2[img];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int col = col0 + wn * 64 + j * 16 + (lane & 15);
          if (col < N) {
            const float t2 = ep.tn2[(size_t)img * ep.ld_tn2 + col];
            float s = (float)(-(double)acc[i][j][r] / sqrt(y2 * (double)t2));
            if (!(s == s)) s = INFINITY;  // empty template / image
            top2_push(t, s, ep.col_base + col);
          }
        }
      }
      // reduce over the 16 lanes holding the same row
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        Top2 u;
        u.s1 = __shfl_xor(t.s1, o, 64);
        u.c1 = __shfl_xor(t.c1, o, 64);
        u.s2 = __shfl_xor(t.s2, o, 64);
        u.c2 = __shfl_xor(t.c2, o, 64);
        top2_push(t, u.s1, u.c1);
        top2_push(t, u.s2, u.c2);
      }
      if ((lane & 15) == 0) red[wn][rl] = t;
    }
  __syncthreads();
  if (tid < BM) {
    const int row = row0 + tid;
    if (row < M) {
      Top2 t = red[0][tid];
      top2_push(t, red[1][tid].s1, red[1][tid].c1);
      top2_push(t, red[1][tid].s2, red[1][tid].c2);
      ep.partial[(size_t)row * ep.n_cb + ep.cb_base + blockIdx.x] = t;
    }
  }
}

// Per image: FP32 best, then every candidate within the tie window.
__global__ __launch_bounds__(NT) void select_kernel(const Top2* __restrict__ partial, int n_off, int n_cb, int n_rot,
                                                    float window, long long* __restrict__ cand,
                                                    int* __restrict__ cand_count, int* __restrict__ overflow) {
  __shared__ float bs[NT];
  __shared__ long long bi[NT];
  __shared__ int count;
  __shared__ int over;
  const int img = blockIdx.x;
  const int tid = threadIdx.x;
  const int total = n_off * n_cb;
  float best = INFINITY;
  long long bidx = LLONG_MAX;
  for (int e = tid; e < total; e += NT) {
    const int off = e / n_cb, cb = e % n_cb;
    const Top2 t = partial[((size_t)img * n_off + off) * n_cb + cb];
    if (t.c1 != INT_MAX) {
      const long long ridx = (long long)t.c1 * n_off + off;
      if (t.s1 < best || (t.s1 == best && ridx < bidx)) {
        best = t.s1;
        bidx = ridx;
      }
    }
  }
  bs[tid] = best;
  bi[tid] = bidx;
  if (tid == 0) {
    count = 0;
    over = 0;
  }
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) {
      if (bs[tid + s] < bs[tid] || (bs[tid + s] == bs[tid] && bi[tid + s] < bi[tid])) {
        bs[tid] = bs[tid + s];
        bi[tid] = bi[tid + s];
      }
    }
    __syncthreads();
  }
  const float thr = bs[0] + window;
  long long* my = cand + (size_t)img * CAND_CAP;
  for (int e = tid; e < total; e += NT) {
    const int off = e / n_cb, cb = e % n_cb;
    const Top2 t = partial[((size_t)img * n_off + off) * n_cb + cb];
    if (t.c1 != INT_MAX && t.s1 <= thr) {
      if (t.c2 != INT_MAX && t.s2 <= thr) {
        // more than one close score in this block: re-score every rotation in it
        const int r0 = cb * BN;
        const int r1 = (r0 + BN) < n_rot ? r0 + BN : n_rot;
        for (int r = r0; r < r1; ++r) {
          const int slot = atomicAdd(&count, 1);
          if (slot < CAND_CAP) my[slot] = (long long)r * n_off + off;
          else over = 1;
        }
      } else {
        const int slot = atomicAdd(&count, 1);
        if (slot < CAND_CAP) my[slot] = (long long)t.c1 * n_off + off;
        else over = 1;
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    if (over || count == 0) {  // keep at least the FP32 best
      my[0] = bi[0];
      cand_count[img] = 1;
    } else {
      cand_count[img] = count;
    }
    overflow[img] = over;
  }
}

// FP64 trilinear sample of a complex128 grid at the rotated bin (i, j), with
// the fill / wrap semantics of the FP32 gather.
__device__ __forceinline__ double2 sample_f64(const double2* __restrict__ g, int n, int i, int j,
                                              const double* __restrict__ r) {
  const int m = n / 2;
  const double fx = (double)((j < m) ? j : j - n), fy = (double)((i < m) ? i : i - n);
  const double h = (double)m;
  const double c[3] = {fma(r[6], fx, fma(r[7], fy, h)), fma(r[3], fx, fma(r[4], fy, h)),
                       fma(r[0], fx, fma(r[1], fy, h))};
  int lo[3];
  double fr[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double f = floor(c[a]);
    lo[a] = (int)f;
    fr[a] = c[a] - f;
  }
  double2 acc = make_double2(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
    int z, y, x;
    const bool v = wrap_index(lo[0] + dz, n, z) && wrap_index(lo[1] + dy, n, y) && wrap_index(lo[2] + dx, n, x);
    const double w = (dz ? fr[0] : 1.0 - fr[0]) * (dy ? fr[1] : 1.0 - fr[1]) * (dx ? fr[2] : 1.0 - fr[2]);
    if (v) {
      const double2 e = g[((size_t)z * n + y) * n + x];
      acc.x = fma(w, e.x, acc.x);
      acc.y = fma(w, e.y, acc.y);
    }
  }
  return acc;
}

__global__ __launch_bounds__(NT) void rescore_kernel(const double2* __restrict__ g, int n,
                                                     const double2* __restrict__ y,
                                                     const cryomi_image_params* __restrict__ prm,
                                                     const double* __restrict__ rotations,
                                                     const double* __restrict__ offsets, int n_off,
                                                     const double* __restrict__ yn2,
                                                     const long long* __restrict__ cand,
                                                     const int* __restrict__ cand_count,
                                                     long long* __restrict__ best_index,
                                                     double* __restrict__ best_score) {
  __shared__ double red[NT / 64];
  const int img = blockIdx.x;
  const int m = n / 2;
  const int nh = n * (m + 1);
  const ImageConsts kc = make_consts(prm[img], n, false);
  const double inv = 1.0 / ((double)n * prm[img].pixel_size);
  const double2* yi = y + (size_t)img * nh;
  double best = INFINITY;
  long long bidx = LLONG_MAX;
  const int cnt = cand_count[img];
  for (int c = 0; c < cnt; ++c) {
    const long long ridx = cand[(size_t)img * CAND_CAP + c];
    const int rot = (int)(ridx / n_off), off = (int)(ridx % n_off);
    const double* r = rotations + 9 * (size_t)rot;
    const double tx = offsets[2 * off] * inv, ty = offsets[2 * off + 1] * inv;
    double num = 0.0, tn = 0.0;
    for (int k = threadIdx.x; k < nh; k += NT) {
      const int i = k / (m + 1), j = k % (m + 1);
      const double w = bin_weight(i, j, m);
      if (w == 0.0) continue;
      const double ct = ctf_f64(kc, i, j, n);
      const double2 p = sample_f64(g, n, i, j, r);
      const double2 s = conj_shift_f64(tx, ty, i, j, n);
      const double2 yy = yi[k];
      // Re(w Y CTF conj(shift) conj(P))
      const double yr = yy.x * s.x - yy.y * s.y, yim = yy.x * s.y + yy.y * s.x;
      num += w * ct * (yr * p.x + yim * p.y);
      tn += w * ct * ct * (p.x * p.x + p.y * p.y);
    }
    num = block_sum_f64(num, red);
    tn = block_sum_f64(tn, red);
    double sc = -num / sqrt(yn2[img] * tn);
    if (!(sc == sc)) sc = INFINITY;
    // MinimumSearchMethod's rule: strictly smaller wins, the lower raveled
    // index wins an exact tie
    if (sc < best || (sc == best && ridx < bidx)) {
      best = sc;
      bidx = ridx;
    }
  }
  if (threadIdx.x == 0) {
    best_index[img] = bidx;
    best_score[img] = best;
  }
}

struct SearchLayout {
  size_t a, d, yn2, b, tn2, partial, cand, cand_count, total;
  int kp, kp2, rows, rows_p, img_p, chunk, chunk_p, n_cb;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static SearchLayout layout(int n, int n_img, int n_rot, int n_off) {
  SearchLayout L{};
  const int nh = n * (n / 2 + 1);
  L.kp = ((2 * nh + KT - 1) / KT) * KT;
  L.kp2 = L.kp / 2;
  L.rows = n_img * n_off;
  L.rows_p = ((L.rows + BM - 1) / BM) * BM;
  L.img_p = ((n_img + BM - 1) / BM) * BM;
  L.chunk = n_rot < ROT_CHUNK ? n_rot : ROT_CHUNK;
  L.chunk_p = ((L.chunk + BN - 1) / BN) * BN;
  L.n_cb = (n_rot + BN - 1) / BN;
  size_t o = 0;
  L.a = o;
  o = align256(o + (size_t)L.rows_p * L.kp * sizeof(float));
  L.d = o;
  o = align256(o + (size_t)L.img_p * L.kp * sizeof(float));
  L.yn2 = o;
  o = align256(o + (size_t)n_img * sizeof(double));
  L.b = o;
  o = align256(o + (size_t)L.chunk_p * L.kp * sizeof(float));
  L.tn2 = o;
  o = align256(o + (size_t)L.img_p * L.chunk_p * sizeof(float));
  L.partial = o;
  o = align256(o + (size_t)L.rows * L.n_cb * sizeof(Top2));
  L.cand = o;
  o = align256(o + (size_t)n_img * CAND_CAP * sizeof(long long));
  L.cand_count = o;
  o = align256(o + (size_t)n_img * sizeof(int));
  L.total = o;
  return L;
}

}  // namespace
}  // namespace cryomi

using namespace cryomi;

extern "C" {

size_t cryomi_ncc_search_workspace_bytes(int box, int n_images, int n_rotations, int n_offsets) {
  if (!box_supported(box) || n_images <= 0 || n_rotations <= 0 || n_offsets <= 0) return 0;
  return layout(box, n_images, n_rotations, n_offsets).total;
}

int cryomi_sgemm_nt(const float* A, const float* B, float* C, int M, int N, int K, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(M >= 0 && N >= 0 && K >= 0 && K % KT == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_sgemm_nt: need M, N >= 0 and K a non-negative multiple of 16");
  if (M == 0 || N == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(A && B && C, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_sgemm_nt: NULL pointer");
  if (K == 0) {
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1024), dim3(NT), 0, stream, C, (size_t)M * N);
    return check_launch("fill_zero_kernel");
  }
  NccEpilogue ep{};
  hipLaunchKernelGGL(gemm_nt_kernel<GEMM_STORE>, dim3((N + BN - 1) / BN, (M + BM - 1) / BM), dim3(GEMM_THREADS), 0,
                     stream, A, B, M, N, K, C, N, ep);
  return check_launch("gemm_nt_kernel");
}

int cryomi_ncc_search(const cryomi_ncc_search_args* a, void* workspace, size_t workspace_bytes, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(a != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_search: NULL args");
  CRYOMI_REQUIRE(box_supported(a->box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_ncc_search: box must be even, 8 <= box <= 512 and of the form 2^a 3^b");
  CRYOMI_REQUIRE(a->n_images >= 0 && a->n_rotations >= 0 && a->n_offsets >= 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_ncc_search: negative count");
  if (a->n_images == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(a->n_rotations > 0 && a->n_offsets > 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_ncc_search: empty rotation or offset grid");
  CRYOMI_REQUIRE((long long)a->n_rotations * a->n_offsets < (1ll << 62) &&
                     (long long)a->n_images * a->n_offsets < (1ll << 31),
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_search: grid too large");
  CRYOMI_REQUIRE(a->grid_c64 && a->grid_c128 && a->observed_rfft && a->image_params && a->rotations && a->offsets &&
                     a->best_index && a->best_score && a->overflow,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_search: NULL pointer");
  const int n = a->box;
  const SearchLayout L = layout(n, a->n_images, a->n_rotations, a->n_offsets);
  CRYOMI_REQUIRE(workspace && workspace_bytes >= L.total, CRYOMI_ERR_WORKSPACE,
                 "cryomi_ncc_search: workspace smaller than cryomi_ncc_search_workspace_bytes()");
  char* ws = static_cast<char*>(workspace);
  float* A = reinterpret_cast<float*>(ws + L.a);
  float* D = reinterpret_cast<float*>(ws + L.d);
  double* yn2 = reinterpret_cast<double*>(ws + L.yn2);
  float* Bt = reinterpret_cast<float*>(ws + L.b);
  float* tn2 = reinterpret_cast<float*>(ws + L.tn2);
  Top2* partial = reinterpret_cast<Top2*>(ws + L.partial);
  long long* cand = reinterpret_cast<long long*>(ws + L.cand);
  int* cand_count = reinterpret_cast<int*>(ws + L.cand_count);
  const double2* y = static_cast<const double2*>(a->observed_rfft);
  const float window = a->tie_window > 0 ? (float)a->tie_window : 2e-4f;

  hipLaunchKernelGGL(prepare_obs_kernel, dim3(L.rows_p), dim3(NT), 0, stream, y, n, a->n_images, a->n_offsets,
                     a->image_params, a->offsets, L.kp2, reinterpret_cast<float2*>(A));
  int st = check_launch("prepare_obs_kernel");
  if (st) return st;
  hipLaunchKernelGGL(prepare_ctf_kernel, dim3(L.img_p), dim3(NT), 0, stream, y, n, a->n_images, a->image_params,
                     L.kp2, reinterpret_cast<float2*>(D), yn2);
  if ((st = check_launch("prepare_ctf_kernel"))) return st;
  for (int r0 = 0; r0 < a->n_rotations; r0 += L.chunk) {
    const int cnt = (a->n_rotations - r0) < L.chunk ? (a->n_rotations - r0) : L.chunk;
    hipLaunchKernelGGL(template_kernel, dim3((L.kp2 + NT - 1) / NT, L.chunk_p), dim3(NT), 0, stream,
                       static_cast<const float2*>(a->grid_c64), n, a->rotations, r0, cnt, L.kp2,
                       reinterpret_cast<float2*>(Bt));
    if ((st = check_launch("template_kernel"))) return st;
    NccEpilogue none{};
    hipLaunchKernelGGL(gemm_nt_kernel<GEMM_SQUARE_B>, dim3((cnt + BN - 1) / BN, L.img_p / BM), dim3(GEMM_THREADS), 0,
                       stream, D, Bt, a->n_images, cnt, L.kp, tn2, L.chunk_p, none);
    if ((st = check_launch("gemm_nt_kernel<square>"))) return st;
    NccEpilogue ep{a->n_offsets, yn2, tn2, L.chunk_p, r0, r0 / BN, L.n_cb, partial};
    hipLaunchKernelGGL(gemm_nt_kernel<GEMM_NCC>, dim3((cnt + BN - 1) / BN, L.rows_p / BM), dim3(GEMM_THREADS), 0,
                       stream, A, Bt, L.rows, cnt, L.kp, nullptr, 0, ep);
    if ((st = check_launch("gemm_nt_kernel<ncc>"))) return st;
  }
  hipLaunchKernelGGL(select_kernel, dim3(a->n_images), dim3(NT), 0, stream, partial, a->n_offsets, L.n_cb,
                     a->n_rotations, window, cand, cand_count, a->overflow);
  if ((st = check_launch("select_kernel"))) return st;
  hipLaunchKernelGGL(rescore_kernel, dim3(a->n_images), dim3(NT), 0, stream,
                     static_cast<const double2*>(a->grid_c128), n, y, a->image_params, a->rotations, a->offsets,
                     a->n_offsets, yn2, cand, cand_count, a->best_index, a->best_score);
  return check_launch("rescore_kernel");
}

}  // extern "C"
