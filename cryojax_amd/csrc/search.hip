// Exhaustive NCC pose search on gfx950: FP32 MFMA contraction of CTF- and
// shift-folded observed spectra against slice templates, fused NCC + top-2
// epilogue, candidate selection and FP64 re-scoring of near-ties.
//
// Kernels
//   prepare_obs_kernel  : rows (image, offset): A = w Y CTF conj(shift) as
//                         interleaved (re, im) float32, K = 2 N (N/2 + 1)
//                         padded to a multiple of 16; one thread per (image,
//                         bin) writes that bin of all the image's rows.
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
//   rescore_flat_kernel : FP64 NCC of every selected candidate from the
//                         complex128 grid, one wave per (image, candidate)
//                         from a queue; best_of_kernel then keeps per image
//                         the lowest score, the lowest raveled index on ties.
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
constexpr int NCC_SPLIT = 8;  // K splits of the few-row NCC contraction

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

// A[row = img * n_off + off][k] = w Y CTF conj(shift); rows n_img * n_off .. n_rows - 1 are zero.
// Block (chunk, img): one thread per bin k loads Y and forms w CTF once (FP64), then writes
// the bin of every offset's row (coalesced along k). The shift phase q . t is formed and
// reduced mod 1 in FP64, and only the reduced angle goes through the FP32 sincospi: the rows
// are FP32, and the FP64 re-scoring recomputes the shift itself. Grid y = n_img + 1; the last
// row of blocks zero-fills the padding rows.
__global__ __launch_bounds__(NT) void prepare_obs_kernel(const double2* __restrict__ y, int n, int n_img, int n_off,
                                                         int n_rows, const cryomi_image_params* __restrict__ prm,
                                                         const double* __restrict__ offsets, int kp2, int ld2,
                                                         float2* __restrict__ a) {
  const int img = blockIdx.y;
  const int k = blockIdx.x * NT + threadIdx.x;
  if (k >= kp2) return;
  if (img >= n_img) {
    for (int row = n_img * n_off; row < n_rows; ++row) a[(size_t)row * ld2 + k] = make_float2(0.f, 0.f);
    return;
  }
  const int m = n / 2;
  const int nh = n * (m + 1);
  float2* dst = a + (size_t)img * n_off * ld2 + k;
  float2 cy = make_float2(0.f, 0.f);
  int i = 0, j = 0;
  if (k < nh) {
    i = k / (m + 1);
    j = k % (m + 1);
    const double w = bin_weight(i, j, m);
    if (w != 0.0) {
      const double c = w * ctf_f64(make_consts(prm[img], n, false), i, j, n);
      const double2 yy = y[(size_t)img * nh + k];
      cy = make_float2((float)(c * yy.x), (float)(c * yy.y));
    }
  }
  const double inv = 1.0 / ((double)n * prm[img].pixel_size);
  const double fy = (double)((i < m) ? i : i - n), fx = (double)j;
  for (int off = 0; off < n_off; ++off) {
    const double v = fma(offsets[2 * off] * inv, fx, offsets[2 * off + 1] * inv * fy);
    float sn, cs;
    sincospif((float)(2.0 * (v - rint(v))), &sn, &cs);
    dst[(size_t)off * ld2] = make_float2(cy.x * cs - cy.y * sn, cy.x * sn + cy.y * cs);
  }
}

// D[img][2k] = D[img][2k+1] = w CTF^2; yn2[img] = sum w |Y|^2
__global__ __launch_bounds__(NT) void prepare_ctf_kernel(const double2* __restrict__ y, int n, int n_img,
                                                         const cryomi_image_params* __restrict__ prm, int kp2,
                                                         int ld2, float2* __restrict__ d,
                                                         double* __restrict__ yn2) {
  __shared__ double red[NT / 64];
  const int img = blockIdx.x;
  const int m = n / 2;
  const int nh = n * (m + 1);
  float2* dst = d + (size_t)img * ld2;
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

__device__ __forceinline__ ImageConsts rotation_consts(const double* r, int n, bool wrap) {
  ImageConsts k{};
  k.wrap = wrap;
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
                                                     int wrap, int kp2, int ld2, float2* __restrict__ b) {
  const int r = blockIdx.y;
  const int k = blockIdx.x * NT + threadIdx.x;
  if (k >= kp2) return;
  const int m = n / 2;
  const int nh = n * (m + 1);
  float2 v = make_float2(0.f, 0.f);
  if (r < count && k < nh) {
    const ImageConsts kc = rotation_consts(rotations + 9 * (size_t)(rot0 + r), n, wrap != 0);
    v = slice_sample(grid, n, k / (m + 1), k % (m + 1), kc, 0u);
  }
  b[(size_t)r * ld2 + k] = v;
}

enum { GEMM_STORE = 0, GEMM_SQUARE_B = 1, GEMM_NCC = 2, GEMM_STORE_NORMS = 3 };
constexpr int NORM_MAX_IMG = 16;  // images whose template norms ride along a few-row contraction

// GEMM_STORE_NORMS: the first row of blocks also forms, from the B tiles in LDS, the template
// norms sum_k d[img][k] B[col][k]^2 of its K range for n_img <= NORM_MAX_IMG images, stored at
// out + z split + img ldo + col.
struct NormSide {
  const float* d;
  int ldd;
  int n_img;
  float* out;
  int ldo;
  long long split;
};

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

// C = A . B^T (A: M x K, B: N x K, row-major, K % 16 == 0). Grid z splits K: block z
// contracts k in [z ksplit, min((z + 1) ksplit, K)) and stores to C + z c_split (ksplit a
// multiple of 16; gridDim.z = 1 with ksplit = K for the whole contraction).
template <int MODE>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_nt_kernel(const float* __restrict__ A, int lda,
                                                               const float* __restrict__ B, int ldb, int M, int N,
                                                               int K, int ksplit, long long c_split,
                                                               float* __restrict__ C, int ldc, NccEpilogue ep,
                                                               NormSide ns) {
  __shared__ float As[2][KT][LDS_STRIDE];
  __shared__ float Bs[2][KT][LDS_STRIDE];
  __shared__ Top2 red[2][BM];
  __shared__ float Ds[2][MODE == GEMM_STORE_NORMS ? NORM_MAX_IMG : 1][KT];
  const bool norms = MODE == GEMM_STORE_NORMS && blockIdx.y == 0;  // block-uniform
  float nacc[MODE == GEMM_STORE_NORMS ? NORM_MAX_IMG : 1];
#pragma unroll
  for (int i = 0; i < (MODE == GEMM_STORE_NORMS ? NORM_MAX_IMG : 1); ++i) nacc[i] = 0.f;
  float rdn = 0.f;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int row0 = blockIdx.y * BM, col0 = blockIdx.x * BN;
  // global -> LDS mapping: tile row tid / 2, k segment (tid & 1) * 8
  const int lr = tid >> 1, lk = (tid & 1) * 8;
  const bool aval = row0 + lr < M, bval = col0 + lr < N;
  const int kbeg = blockIdx.z * ksplit;
  const float* ap = A + (size_t)(aval ? row0 + lr : 0) * lda + kbeg + lk;
  const float* bp = B + (size_t)(bval ? col0 + lr : 0) * ldb + kbeg + lk;
  if (MODE != GEMM_NCC) C += blockIdx.z * c_split;
  floatx4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  float ra[8], rb[8];
  auto gload = [&](int k0) {
    if (norms) {  // d[img][kbeg + k0 + kq] for slot tid = img KT + kq (clamped, masked)
      const int slot = min(tid, ns.n_img * KT - 1);
      rdn = ns.d[(size_t)(slot / KT) * ns.ldd + kbeg + k0 + slot % KT] * (tid < ns.n_img * KT ? 1.f : 0.f);
    }
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
    if (norms && tid < NORM_MAX_IMG * KT) Ds[buf][tid / KT][tid % KT] = rdn;
  };
  const int nk = min(ksplit, K - kbeg) / KT;
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
    if (norms) {  // column tid % BN, k half tid / BN of this tile
      const int col = tid % BN, kh = tid / BN;
      float b2[KT / 2];
#pragma unroll
      for (int q = 0; q < KT / 2; ++q) {
        const float b = Bs[cur][kh * (KT / 2) + q][col];
        b2[q] = b * b;
      }
#pragma unroll
      for (int i = 0; i < NORM_MAX_IMG; ++i) {
        if (i < ns.n_img) {
          float t = 0.f;
#pragma unroll
          for (int q = 0; q < KT / 2; ++q) t = fmaf(Ds[cur][i][kh * (KT / 2) + q], b2[q], t);
          nacc[i] += t;
        }
      }
    }
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }
  if (norms) {  // combine the two k halves (the A tiles are free now) and store this split's norms
    float* half = &As[0][0][0];
    const int col = tid % BN, kh = tid / BN;
    if (kh == 1)
#pragma unroll
      for (int i = 0; i < NORM_MAX_IMG; ++i)
        if (i < ns.n_img) half[i * BN + col] = nacc[i];
    __syncthreads();
    if (kh == 0 && col0 + col < N)
#pragma unroll
      for (int i = 0; i < NORM_MAX_IMG; ++i)
        if (i < ns.n_img) ns.out[blockIdx.z * ns.split + (size_t)i * ns.ldo + col0 + col] = nacc[i] + half[i * BN + col];
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

// Few images (at most NORM_MAX_IMG): tn2[img][col] = sum_k D[img][k] B[col][k]^2 with one
// wave per template row, each row streamed once; the MFMA GEMM would pad the few image rows
// to a 128-row tile and run 235 workgroups at a time for a 30000-row bank.
__global__ __launch_bounds__(NT) void template_norms_kernel(const float* __restrict__ D, int ldd,
                                                            const float* __restrict__ B, int ldb, int n_img, int N,
                                                            int K, float* __restrict__ tn2, int ld_tn2) {
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  if (col >= N) return;  // whole wave
  const float* b = B + (size_t)col * ldb;
  float acc[NORM_MAX_IMG];
#pragma unroll
  for (int i = 0; i < NORM_MAX_IMG; ++i) acc[i] = 0.f;
  // four bank loads in flight per lane (clamped addresses, zero weight past K)
  constexpr int U = 4;
  for (int k0 = 4 * lane; k0 < K; k0 += 256 * U) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = *reinterpret_cast<const float4*>(b + min(k0 + 256 * u, K - 4));
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = min(k0 + 256 * u, K - 4);
      const float live = k0 + 256 * u < K ? 1.f : 0.f;
      const float4 q = make_float4(v[u].x * v[u].x * live, v[u].y * v[u].y * live, v[u].z * v[u].z * live,
                                   v[u].w * v[u].w * live);
#pragma unroll
      for (int i = 0; i < NORM_MAX_IMG; ++i) {
        if (i < n_img) {
          const float4 d = *reinterpret_cast<const float4*>(D + (size_t)i * ldd + k);
          acc[i] = fmaf(d.x, q.x, fmaf(d.y, q.y, fmaf(d.z, q.z, fmaf(d.w, q.w, acc[i]))));
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NORM_MAX_IMG; ++i) {
    if (i < n_img) {
      float t = acc[i];
      for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
      if (lane == 0) tn2[(size_t)i * ld_tn2 + col] = t;
    }
  }
}

// Few rows: the K-split partial products of the NCC contraction (gemm_nt_kernel<GEMM_STORE>
// with grid z), summed and given the epilogue of gemm_nt_kernel<GEMM_NCC>: per row, the two
// best over each BN-column block.
__global__ __launch_bounds__(BN) void ncc_splitk_kernel(const float* __restrict__ part, int n_split,
                                                        long long split_stride, int ldc, int N, NccEpilogue ep,
                                                        const float* __restrict__ norm_part, long long norm_split) {
  __shared__ Top2 red[BN / 64];
  const int row = blockIdx.y, col = blockIdx.x * BN + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  Top2 t = top2_empty();
  if (col < N) {
    float acc = 0.f;
    for (int z = 0; z < n_split; ++z) acc += part[z * split_stride + (size_t)row * ldc + col];
    const int img = row / ep.n_off;
    float t2 = 0.f;
    if (norm_part) {  // the split contraction's own template norms
      for (int z = 0; z < n_split; ++z) t2 += norm_part[z * norm_split + (size_t)img * ldc + col];
    } else {
      t2 = ep.tn2[(size_t)img * ep.ld_tn2 + col];
    }
    float sc = (float)(-(double)acc / sqrt(ep.yn2[img] * (double)t2));
    if (!(sc == sc)) sc = INFINITY;  // empty template / image
    top2_push(t, sc, ep.col_base + col);
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    Top2 u;
    u.s1 = __shfl_xor(t.s1, o, 64);
    u.c1 = __shfl_xor(t.c1, o, 64);
    u.s2 = __shfl_xor(t.s2, o, 64);
    u.c2 = __shfl_xor(t.c2, o, 64);
    top2_push(t, u.s1, u.c1);
    top2_push(t, u.s2, u.c2);
  }
  if (lane == 0) red[wave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    Top2 r = red[0];
    for (int w = 1; w < BN / 64; ++w) {
      top2_push(r, red[w].s1, red[w].c1);
      top2_push(r, red[w].s2, red[w].c2);
    }
    ep.partial[(size_t)row * ep.n_cb + ep.cb_base + blockIdx.x] = r;
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

// Bricked complex128 grid for the FP64 re-scoring (written once per search call): each (z, y)
// row is padded to x = -1 .. N (x = -1 holds the wrapped voxel N - 1, or the fill zero under
// CRYOMI_GATHER_FILL; x = N the fill zero), so that every trilinear x pair with x0 in
// [-1, N - 1] is contiguous, and pairs with x0 in [-N, -2] (both corners wrapped) lie in the
// wrapped part; pairs start at row entry p = x0 + 1 (x0 >= -1) or x0 + N + 1. The FP64 sample
// has the FP32 gather's corner rule (sample.hpp). Two copies, the second shifted by one entry, are cut
// into 2 x 2 x 2 (z, y, entry) bricks of 8 complex128 (one 128-byte line); the copy in which
// the pair starts at an even entry holds it in one brick. A sample's eight corners then touch
// one to four lines (2.25 on average) instead of about 4.5 in the plain layout.
__host__ __device__ inline size_t c128_brick_copy(int n) { return (size_t)n * n * (n + 4); }  // entries per copy

__global__ __launch_bounds__(NT) void c128_to_bricks_kernel(const double2* __restrict__ g, int n, int wrap,
                                                           double2* __restrict__ gb) {
  const size_t per = c128_brick_copy(n);
  const int pb = (n + 4) / 2;  // bricks per padded row
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < 2 * per; e += (size_t)gridDim.x * NT) {
    const int copy = (int)(e / per);
    const size_t r = e - (size_t)copy * per;
    const size_t brick = r >> 3;
    const int in = (int)(r & 7);
    const int bp = (int)(brick % pb), byz = (int)(brick / pb);
    const int by = byz % (n / 2), bz = byz / (n / 2);
    const int z = 2 * bz + (in >> 2), y = 2 * by + ((in >> 1) & 1), p = 2 * bp + (in & 1);
    const int x = p - 1 - copy;  // row entry p of copy c holds x = p - 1 - c
    double2 v = make_double2(0.0, 0.0);
    const int xs = x < 0 ? x + n : x;  // [-N, -1] wraps (or fills); N and beyond fill zero
    if ((wrap ? x >= -n : x >= 0) && x < n) v = g[((size_t)z * n + y) * n + xs];
    gb[e] = v;
  }
}

// sample_f64 on the bricked copies (the same trilinear sum, grouped by (z, y) rows)
__device__ __forceinline__ double2 sample_f64_bricked(const double2* __restrict__ gb, int n, int i, int j,
                                                      const double* __restrict__ r, bool wrap) {
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
  const int x0 = lo[2];
  // else both x corners fill zero; x0 = -N - 1 (whose x1 = -N wraps) cannot occur: |r f| <= N / sqrt(2).
  // Without wrapping, x0 = -1 reads the zero entry x = -1 and x0 < -1 fills both corners.
  const bool xok = (wrap ? x0 >= -n : x0 >= -1) && x0 <= n - 1;
  const int p = x0 >= -1 ? x0 + 1 : x0 + n + 1;
  const int copy = p & 1;
  const int pp = p + copy;  // even entry of the pair in its copy
  const double2* base = gb + (size_t)copy * c128_brick_copy(n);
  const int pb = (n + 4) / 2;
  double2 acc = make_double2(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int dz = q >> 1, dy = q & 1;
    int z, y;
    const bool v = xok && wrap_index(lo[0] + dz, n, z, wrap) && wrap_index(lo[1] + dy, n, y, wrap);
    const size_t at = v ? ((((size_t)(z >> 1) * m + (y >> 1)) * pb + (pp >> 1)) * 8 + ((z & 1) * 2 + (y & 1)) * 2) : 0;
    const double2 e0 = base[at], e1 = base[at + 1];
    const double w = v ? (dz ? fr[0] : 1.0 - fr[0]) * (dy ? fr[1] : 1.0 - fr[1]) : 0.0;
    const double w0 = w * (1.0 - fr[2]), w1 = w * fr[2];
    acc.x = fma(w1, e1.x, fma(w0, e0.x, acc.x));
    acc.y = fma(w1, e1.y, fma(w0, e0.y, acc.y));
  }
  return acc;
}

// ---------------------------------------------------------------------------
// work lists and flat FP64 re-scoring
// ---------------------------------------------------------------------------
constexpr int SCAN_NT = 1024;

// off[i] = sum_{j < i} ceil(min(counts[j], cap) / per), off[n] = total. One
// workgroup.
__global__ __launch_bounds__(SCAN_NT) void work_prefix_kernel(const int* __restrict__ counts, int n, int cap, int per,
                                                              int* __restrict__ off) {
  __shared__ int sums[SCAN_NT / 64];
  __shared__ int carry_s;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < n; base += SCAN_NT) {
    const int i = base + tid;
    int v = 0;
    if (i < n) {
      const int c = min(max(counts[i], 0), cap);
      v = (c + per - 1) / per;
    }
    // inclusive scan within the wave, then across waves
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) sums[wave] = x;
    __syncthreads();
    int before = carry_s;
    for (int w = 0; w < wave; ++w) before += sums[w];
    if (i < n) off[i] = before + x - v;
    __syncthreads();
    if (tid == SCAN_NT - 1) carry_s = before + x;
    __syncthreads();
  }
  if (tid == 0) off[n] = carry_s;
}

// Largest i in [0, n) with off[i] <= item (off nondecreasing, item < off[n]).
__device__ __forceinline__ int owner_of(const int* __restrict__ off, int n, int item) {
  int lo = 0, hi = n;  // off[lo] <= item < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Persistent workgroups, items (image, candidate) dealt round-robin
// (blockIdx.x, blockIdx.x + gridDim.x, ...): every item costs the same, and
// the loop control stays uniform with no per-thread ticket. Each item is one
// FP64 -NCC from the complex128 grid.
__global__ __launch_bounds__(NT) void rescore_flat_kernel(const double2* __restrict__ g, int n,
                                                          const double2* __restrict__ y,
                                                          const cryomi_image_params* __restrict__ prm,
                                                          const double* __restrict__ rotations,
                                                          const double* __restrict__ offsets, int n_off,
                                                          const double* __restrict__ yn2,
                                                          const long long* __restrict__ cand, int n_img,
                                                          const int* __restrict__ flat_off, int wrap,
                                                          double* __restrict__ flat_score) {
  __shared__ double red[NT / 64];
  const int tid = threadIdx.x;
  const int m = n / 2;
  const int nh = n * (m + 1);
  const int total = flat_off[n_img];
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int img = owner_of(flat_off, n_img, item);
    const long long ridx = cand[(size_t)img * CAND_CAP + (item - flat_off[img])];
    const int rot = (int)(ridx / n_off), off = (int)(ridx % n_off);
    const ImageConsts kc = make_consts(prm[img], n, false);
    const double inv = 1.0 / ((double)n * prm[img].pixel_size);
    const double tx = offsets[2 * off] * inv, ty = offsets[2 * off + 1] * inv;
    const double* r = rotations + 9 * (size_t)rot;
    const double2* yi = y + (size_t)img * nh;
    double num = 0.0, tn = 0.0;
    for (int k = tid; k < nh; k += NT) {
      const int i = k / (m + 1), j = k % (m + 1);
      const double w = bin_weight(i, j, m);
      if (w == 0.0) continue;
      const double ct = ctf_f64(kc, i, j, n);
      const double2 p = sample_f64_bricked(g, n, i, j, r, wrap != 0);
      const double2 sh = conj_shift_f64(tx, ty, i, j, n);
      const double2 yy = yi[k];
      // Re(w Y CTF conj(shift) conj(P))
      const double yr = yy.x * sh.x - yy.y * sh.y, yim = yy.x * sh.y + yy.y * sh.x;
      num += w * ct * (yr * p.x + yim * p.y);
      tn += w * ct * ct * (p.x * p.x + p.y * p.y);
    }
    num = block_sum_f64(num, red);
    tn = block_sum_f64(tn, red);
    double sc = -num / sqrt(yn2[img] * tn);
    if (!(sc == sc)) sc = INFINITY;
    if (tid == 0) flat_score[item] = sc;
  }
}

// Per image (one wave): the best re-scored candidate under MinimumSearchMethod's
// rule (strictly smaller, then the lower raveled index), merged with the
// existing result when asked.
__global__ __launch_bounds__(64) void best_of_kernel(const long long* __restrict__ cand, const int* __restrict__ flat_off,
                                                     const double* __restrict__ flat_score, int merge,
                                                     long long* __restrict__ best_index,
                                                     double* __restrict__ best_score) {
  const int img = blockIdx.x, lane = threadIdx.x;
  const int f0 = flat_off[img], f1 = flat_off[img + 1];
  double best = INFINITY;
  long long bidx = LLONG_MAX;
  for (int f = f0 + lane; f < f1; f += 64) {
    const double sc = flat_score[f];
    const long long ridx = cand[(size_t)img * CAND_CAP + (f - f0)];
    if (sc < best || (sc == best && ridx < bidx)) {
      best = sc;
      bidx = ridx;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const long long oi = __shfl_xor(bidx, o, 64);
    if (ob < best || (ob == best && oi < bidx)) {
      best = ob;
      bidx = oi;
    }
  }
  if (lane == 0) {
    if (merge) {
      const double ob = best_score[img];
      const long long oi = best_index[img];
      if (!(best < ob || (best == ob && bidx < oi))) {
        best = ob;
        bidx = oi;
      }
    }
    best_index[img] = bidx;
    best_score[img] = best;
  }
}

// Resident workgroups for a persistent launch of `kernel`.
static int persistent_blocks(const void* kernel, int threads) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess || nb <= 0) nb = 1;
  return device_cu_count() * nb;
}

// FP64 re-scoring of every image's candidate list (cand, cand_count), then
// the per-image best (merged into the existing best when `merge`).
static int launch_rescore(const double2* g, int n, const double2* y, const cryomi_image_params* prm,
                          const double* rotations, const double* offsets, int n_off, const double* yn2,
                          const long long* cand, const int* cand_count, int n_img, int merge, int wrap,
                          int* flat_off, double* flat_score, long long* best_index, double* best_score,
                          hipStream_t stream) {
  hipLaunchKernelGGL(work_prefix_kernel, dim3(1), dim3(SCAN_NT), 0, stream, cand_count, n_img, CAND_CAP, 1, flat_off);
  int st = check_launch("work_prefix_kernel");
  if (st) return st;
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cached[dev] == 0) cached[dev] = persistent_blocks((const void*)rescore_flat_kernel, NT);
  const int blocks = cached[dev];
  hipLaunchKernelGGL(rescore_flat_kernel, dim3(blocks), dim3(NT), 0, stream, g, n, y, prm, rotations, offsets, n_off,
                     yn2, cand, n_img, flat_off, wrap, flat_score);
  if ((st = check_launch("rescore_flat_kernel"))) return st;
  hipLaunchKernelGGL(best_of_kernel, dim3(n_img), dim3(64), 0, stream, cand, flat_off, flat_score, merge, best_index,
                     best_score);
  return check_launch("best_of_kernel");
}

// ---------------------------------------------------------------------------
// candidate-restricted search (learned proposals)
// ---------------------------------------------------------------------------
constexpr int SC_BN = 64;   // candidate templates per workgroup
constexpr int SC_BM = 32;   // offset rows per workgroup (n_off <= 32)
constexpr int MAX_CAP = 8192;

// One image's n_off folded rows against 64 of its candidate templates: FP32
// MFMA contraction with the NCC epilogue keeping, per (row, 64-candidate
// block), the two best. Within every 16-k chunk, lane group g (lane / 16)
// holds k = 4 g .. 4 g + 3 of its row / column as one float4, and MFMA step
// s of the chunk contracts k = 4 g + s over the four groups (a permutation
// of k, identical for both operands). Each wave streams its own 16 candidate
// rows of the bank through registers (SC_U chunks per buffer, double
// buffered); the image's 32 observed rows and its w CTF^2 row, which all four
// waves need, are staged once per buffer in LDS by the whole workgroup with
// row-contiguous loads (zero past the rows and past K, so no masking in the
// contraction). The same bank loads accumulate sum w CTF^2 |P|^2 per template.
constexpr int SC_U = 4;              // 16-k chunks per buffer
constexpr int SC_AK = SC_U * 16;     // k per buffer
constexpr int SC_AS = SC_AK + 4;     // LDS row stride of the staged observed rows (floats)

struct SubsetLds {
  float a[2][SC_BM][SC_AS];
  float d[2][SC_AK];
  Top2 red[NT / 64][SC_BM];
};

__device__ __forceinline__ void subset_block(const float* __restrict__ A, const float* __restrict__ D,
                                             const double* __restrict__ yn2, const float* __restrict__ bank, int kp,
                                             int ld, const int* __restrict__ cand, const int* __restrict__ cand_count,
                                             int cap, int n_rot, int n_off, int n_cb, Top2* __restrict__ partial,
                                             int img, int cb, SubsetLds& sm, int tid, int lane, int wave) {
  const int cnt = min(cand_count[img], cap);
  const int c0 = cb * SC_BN;
  Top2* out = partial + (size_t)img * n_off * n_cb + cb;  // row r at out[r * n_cb]
  const int g = lane >> 4, li = lane & 15;
  const int col = wave * 16 + li;
  // candidate indices outside [0, n_rot) are treated as empty columns
  const int crot = c0 + col < cnt ? cand[(size_t)img * cap + c0 + col] : -1;
  const bool cval = crot >= 0 && crot < n_rot;
  const float* bp = bank + (size_t)(cval ? crot : 0) * ld + 4 * g;
  const float* arow = A + (size_t)img * n_off * ld;
  const float* drow = D + (size_t)img * ld;
  const int nchunk = kp / 16;
  // staging loads: slot e = tid + NT q of the 32 rows x (SC_AK / 4) float4, from clamped
  // addresses (always in bounds); out-of-range slots are zeroed at the LDS store
  float4 ra[SC_BM * SC_AK / 4 / NT];
  float4 rd;
  auto a_issue = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < SC_BM * SC_AK / 4 / NT; ++q) {
      const int e = tid + NT * q;
      const int r = e / (SC_AK / 4), k = k0 + 4 * (e % (SC_AK / 4));
      ra[q] = *reinterpret_cast<const float4*>(arow + (size_t)min(r, n_off - 1) * ld + min(k, kp - 4));
    }
    rd = *reinterpret_cast<const float4*>(drow + min(k0 + 4 * (tid % (SC_AK / 4)), kp - 4));
  };
  // component-wise selects: a select of whole float4 values goes through private memory
  auto keep = [](bool ok, float4 v) __attribute__((always_inline)) {
    return make_float4(ok ? v.x : 0.f, ok ? v.y : 0.f, ok ? v.z : 0.f, ok ? v.w : 0.f);
  };
  auto a_store = [&](int buf, int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < SC_BM * SC_AK / 4 / NT; ++q) {
      const int e = tid + NT * q;
      const int r = e / (SC_AK / 4), c4 = e % (SC_AK / 4);
      *reinterpret_cast<float4*>(&sm.a[buf][r][4 * c4]) = keep(r < n_off && k0 + 4 * c4 < kp, ra[q]);
    }
    if (tid < SC_AK / 4) *reinterpret_cast<float4*>(&sm.d[buf][4 * tid]) = keep(k0 + 4 * tid < kp, rd);
  };
  floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
  float tn = 0.f;
  float4 xb[2][SC_U];
  auto b_issue = [&](int buf, int c) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < SC_U; ++u)
      xb[buf][u] = *reinterpret_cast<const float4*>(bp + (size_t)min(c + u, nchunk - 1) * 16);  // clamped
  };
  auto compute = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < SC_U; ++u) {
      const float4 p = *reinterpret_cast<const float4*>(&sm.a[buf][li][16 * u + 4 * g]);
      const float4 q = *reinterpret_cast<const float4*>(&sm.a[buf][16 + li][16 * u + 4 * g]);
      const float4 d = *reinterpret_cast<const float4*>(&sm.d[buf][16 * u + 4 * g]);
      const float4 b = xb[buf][u];
      tn = fmaf(d.x * b.x, b.x, tn);
      tn = fmaf(d.y * b.y, b.y, tn);
      tn = fmaf(d.z * b.z, b.z, tn);
      tn = fmaf(d.w * b.w, b.w, tn);
      acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(p.x, b.x, acc0, 0, 0, 0);
      acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(p.y, b.y, acc0, 0, 0, 0);
      acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(p.z, b.z, acc0, 0, 0, 0);
      acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(p.w, b.w, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(q.x, b.x, acc1, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(q.y, b.y, acc1, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(q.z, b.z, acc1, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(q.w, b.w, acc1, 0, 0, 0);
    }
  };
  // Pipeline per pair of buffers: issue the next buffer's bank and staging loads, contract the
  // current buffer, then store the staged rows and synchronise (all waves are past the buffer
  // being overwritten: the previous barrier followed their last read of it). Buffers past the
  // end contract zero rows (clamped bank loads, zeroed staging). Scheduling barriers keep the
  // loads ahead of the contraction.
  // A wave none of whose 16 columns holds a candidate (partial last blocks, refinement
  // rounds) only helps stage the observed rows: it skips its bank loads and contraction
  // (wave-uniform, so the barriers still line up).
  const bool busy = __ballot(cval) != 0ull;
  a_issue(0);
  if (busy) b_issue(0, 0);
  a_store(0, 0);
  __syncthreads();
  for (int c = 0; c < nchunk; c += 2 * SC_U) {
    a_issue(16 * (c + SC_U));
    if (busy) b_issue(1, c + SC_U);
    __builtin_amdgcn_sched_barrier(0);
    if (busy) compute(0);
    __builtin_amdgcn_sched_barrier(0);
    a_store(1, 16 * (c + SC_U));
    __syncthreads();
    a_issue(16 * (c + 2 * SC_U));
    if (busy) b_issue(0, c + 2 * SC_U);
    __builtin_amdgcn_sched_barrier(0);
    if (busy) compute(1);
    __builtin_amdgcn_sched_barrier(0);
    a_store(0, 16 * (c + 2 * SC_U));
    __syncthreads();
  }
  // the four lane groups of one column hold partial sums of its norm
  tn += __shfl_xor(tn, 16, 64);
  tn += __shfl_xor(tn, 32, 64);
  const int rot = cval ? crot : INT_MAX;
  const double y2 = yn2[img];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = i * 16 + 4 * g + r;
      Top2 t = top2_empty();
      if (row < n_off && cval) {
        const float v = i == 0 ? acc0[r] : acc1[r];
        float sc = (float)(-(double)v / sqrt(y2 * (double)tn));
        if (!(sc == sc)) sc = INFINITY;  // empty template / image
        top2_push(t, sc, rot);
      }
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
      if (li == 0) sm.red[wave][row] = t;
    }
  __syncthreads();
  if (tid < n_off) {
    Top2 t = sm.red[0][tid];
    for (int w = 1; w < NT / 64; ++w) {
      top2_push(t, sm.red[w][tid].s1, sm.red[w][tid].c1);
      top2_push(t, sm.red[w][tid].s2, sm.red[w][tid].c2);
    }
    out[(size_t)tid * n_cb] = t;
  }
}

__global__ __launch_bounds__(NT) void subset_ncc_kernel(const float* __restrict__ A, const float* __restrict__ D,
                                                        const double* __restrict__ yn2, const float* __restrict__ bank,
                                                        int kp, int ld, const int* __restrict__ cand,
                                                        const int* __restrict__ cand_count, int cap, int n_rot,
                                                        int n_off, int n_cb, int n_img,
                                                        const int* __restrict__ work_off,
                                                        Top2* __restrict__ partial) {
  __shared__ SubsetLds sm;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int total = work_off[n_img];
  // persistent: (image, 64-candidate block) items dealt round-robin in
  // image-major order; only non-empty blocks exist, so no workgroup idles
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int img = owner_of(work_off, n_img, item);
    subset_block(A, D, yn2, bank, kp, ld, cand, cand_count, cap, n_rot, n_off, n_cb, partial, img,
                 item - work_off[img], sm, tid, lane, wave);
    __syncthreads();  // the staging buffers and red are reused by the next item
  }
}

// Per image: FP32 best over the candidate blocks, then every (rotation,
// offset) within the tie window; a block whose second best is also in the
// window contributes all of its candidates at that offset. When merging, the
// image's current best is left out: its FP64 score is the stored best score
// (re-scoring is deterministic), so re-scoring it cannot change the result.
__global__ __launch_bounds__(NT) void select_subset_kernel(const Top2* __restrict__ partial, int n_off, int n_cb,
                                                           const int* __restrict__ cand_rot,
                                                           const int* __restrict__ cand_count, int cap, int n_rot,
                                                           float window, int merge,
                                                           const long long* __restrict__ best_index,
                                                           long long* __restrict__ cand, int* __restrict__ count_out,
                                                           int* __restrict__ overflow) {
  __shared__ float bs[NT];
  __shared__ long long bi[NT];
  __shared__ int count;
  __shared__ int over;
  const int img = blockIdx.x;
  const int tid = threadIdx.x;
  const int cnt = min(max(cand_count[img], 0), cap);
  const int nbv = (cnt + SC_BN - 1) / SC_BN;  // blocks the scoring kernel wrote
  const int total = n_off * nbv;
  const Top2* pi = partial + (size_t)img * n_off * n_cb;
  float best = INFINITY;
  long long bidx = LLONG_MAX;
  for (int e = tid; e < total; e += NT) {
    const int off = e / nbv;
    const Top2 t = pi[(size_t)off * n_cb + e % nbv];
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
    if (tid < s && (bs[tid + s] < bs[tid] || (bs[tid + s] == bs[tid] && bi[tid + s] < bi[tid]))) {
      bs[tid] = bs[tid + s];
      bi[tid] = bi[tid + s];
    }
    __syncthreads();
  }
  const float thr = bs[0] + window;
  long long* my = cand + (size_t)img * CAND_CAP;
  const long long known = merge ? best_index[img] : -1;  // the current best (merging), or none
  for (int e = tid; e < total; e += NT) {
    const int off = e / nbv, cb = e % nbv;
    const Top2 t = pi[(size_t)off * n_cb + cb];
    if (t.c1 != INT_MAX && t.s1 <= thr) {
      if (t.c2 != INT_MAX && t.s2 <= thr) {
        const int p1 = min(cnt, (cb + 1) * SC_BN);
        for (int p = cb * SC_BN; p < p1; ++p) {
          const int r = cand_rot[(size_t)img * cap + p];
          if (r < 0 || r >= n_rot) continue;
          const long long ridx = (long long)r * n_off + off;
          if (ridx == known) continue;
          const int slot = atomicAdd(&count, 1);
          if (slot < CAND_CAP) my[slot] = ridx;
          else over = 1;
        }
      } else {
        const long long ridx = (long long)t.c1 * n_off + off;
        if (ridx != known) {
          const int slot = atomicAdd(&count, 1);
          if (slot < CAND_CAP) my[slot] = ridx;
          else over = 1;
        }
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    if (bi[0] == LLONG_MAX) {
      count_out[img] = 0;  // no candidates for this image
    } else if (over) {
      my[0] = bi[0];
      count_out[img] = bi[0] == known ? 0 : 1;
    } else if (count == 0) {
      // every in-window candidate was the current best, or (not merging) none was listed
      my[0] = bi[0];
      count_out[img] = bi[0] == known ? 0 : 1;
    } else {
      count_out[img] = count;
    }
    overflow[img] = merge ? (overflow[img] | over) : over;
  }
}

// Grid rotations within the angle threshold (trace(R_h^T G) >= thr) of any
// hypothesis, compacted in ascending index order (wave ballots + prefix).
__global__ __launch_bounds__(NT) void pose_candidates_kernel(const float* __restrict__ hyp, long long hyp_stride,
                                                             int n_hyp, const long long* __restrict__ center,
                                                             int n_off, const double* __restrict__ rot, int n_rot,
                                                             float thr, int cap, int* __restrict__ cand,
                                                             int* __restrict__ cand_count,
                                                             int* __restrict__ overflow) {
  __shared__ float R[8][9];
  __shared__ int wcount[NT / 64];
  const int img = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < n_hyp) {
    float m[9];
    if (center) {
      // a missing result (negative or out of range) matches nothing
      const long long ci = center[img];
      const bool ok = ci >= 0 && ci / n_off < n_rot;
      const double* g = rot + 9 * (size_t)(ok ? ci / n_off : 0);
#pragma unroll
      for (int q = 0; q < 9; ++q) m[q] = ok ? (float)g[q] : NAN;
    } else {
      const float* h = hyp + img * hyp_stride + 9 * tid;
      // Gram-Schmidt on columns 0 and 1, third column = their cross product
      float a1[3] = {h[0], h[3], h[6]}, a2[3] = {h[1], h[4], h[7]};
      const float n1 = rsqrtf(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
      float b1[3] = {a1[0] * n1, a1[1] * n1, a1[2] * n1};
      const float d = b1[0] * a2[0] + b1[1] * a2[1] + b1[2] * a2[2];
      float u2[3] = {a2[0] - d * b1[0], a2[1] - d * b1[1], a2[2] - d * b1[2]};
      const float n2 = rsqrtf(u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]);
      float b2[3] = {u2[0] * n2, u2[1] * n2, u2[2] * n2};
      float b3[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        m[3 * i] = b1[i];
        m[3 * i + 1] = b2[i];
        m[3 * i + 2] = b3[i];
      }
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) R[tid][q] = m[q];
  }
  __syncthreads();
  // the (block-uniform) hypotheses in registers: no LDS reads in the loop over the grid
  float Rr[8][9];
#pragma unroll
  for (int h = 0; h < 8; ++h)
#pragma unroll
    for (int q = 0; q < 9; ++q) Rr[h][q] = h < n_hyp ? R[h][q] : 0.f;
  int total = 0;
  for (int base = 0; base < n_rot; base += NT) {
    const int gi = base + tid;
    bool sel = false;
    if (gi < n_rot) {
      float gm[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) gm[q] = (float)rot[9 * (size_t)gi + q];
#pragma unroll
      for (int h = 0; h < 8; ++h) {
        float tr = 0.f;
#pragma unroll
        for (int q = 0; q < 9; ++q) tr = fmaf(Rr[h][q], gm[q], tr);
        sel = sel || (h < n_hyp && tr >= thr);
      }
    }
    const unsigned long long mask = __ballot(sel);
    const int pos = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wave] = __popcll(mask);
    __syncthreads();
    int before = total, sum = 0;
    for (int w = 0; w < NT / 64; ++w) {
      if (w < wave) before += wcount[w];
      sum += wcount[w];
    }
    if (sel && before + pos < cap) cand[(size_t)img * cap + before + pos] = gi;
    total += sum;
    __syncthreads();
  }
  if (tid == 0) {
    cand_count[img] = total < cap ? total : cap;
    overflow[img] = total > cap ? 1 : 0;
  }
}

struct SearchLayout {
  size_t a, d, yn2, b, tn2, partial, cand, cand_count, flat_off, flat_score, gb, total;
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
  L.flat_off = o;
  o = align256(o + (size_t)(n_img + 1) * sizeof(int));
  L.flat_score = o;
  o = align256(o + (size_t)n_img * CAND_CAP * sizeof(double));
  L.gb = o;
  o = align256(o + 2 * c128_brick_copy(n) * sizeof(double2));
  L.total = o;
  return L;
}

__global__ __launch_bounds__(NT) void copy_count_kernel(const int* __restrict__ src, int n, int* __restrict__ dst) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Candidate-path rows: K = 2 N (N/2+1) padded to 16 floats, stored with a
// row stride of K + 32 floats. K alone is 2^a 520-ish floats for the usual
// boxes, a stride that maps every row's same k onto two of the sixteen L2
// channels (128-byte interleave); the extra 128 bytes rotate rows across all
// of them.
static int cand_kp(int n) {
  const int nh = n * (n / 2 + 1);
  return ((2 * nh + KT - 1) / KT) * KT;
}
static int cand_ld(int n) { return cand_kp(n) + 32; }

struct CandLayout {
  size_t a, d, yn2, partial, cand, cand_count, work_off, flat_off, flat_score, gb, total;
  int kp, ld, n_cb;
};

static CandLayout cand_layout(int n, int n_img, int n_off, int cap) {
  CandLayout L{};
  L.kp = cand_kp(n);
  L.ld = cand_ld(n);
  L.n_cb = (cap + SC_BN - 1) / SC_BN;
  size_t o = 0;
  L.a = o;
  o = align256(o + (size_t)n_img * n_off * L.ld * sizeof(float));
  L.d = o;
  o = align256(o + (size_t)n_img * L.ld * sizeof(float));
  L.yn2 = o;
  o = align256(o + (size_t)n_img * sizeof(double));
  L.gb = o;  // with a, d and yn2: what a call with CRYOMI_NCC_REUSE_PREPARED finds in place (cap-independent)
  o = align256(o + 2 * c128_brick_copy(n) * sizeof(double2));
  L.partial = o;
  o = align256(o + (size_t)n_img * n_off * L.n_cb * sizeof(Top2));
  L.cand = o;
  o = align256(o + (size_t)n_img * CAND_CAP * sizeof(long long));
  L.cand_count = o;
  o = align256(o + (size_t)n_img * sizeof(int));
  L.work_off = o;
  o = align256(o + (size_t)(n_img + 1) * sizeof(int));
  L.flat_off = o;
  o = align256(o + (size_t)(n_img + 1) * sizeof(int));
  L.flat_score = o;
  o = align256(o + (size_t)n_img * CAND_CAP * sizeof(double));
  L.total = o;
  return L;
}

int launch_c128_bricks(const double2* g, int n, int wrap, double2* gb, hipStream_t stream) {
  const size_t total = 2 * c128_brick_copy(n);
  const unsigned blocks = (unsigned)std::min<size_t>((total + NT - 1) / NT, 65536);
  hipLaunchKernelGGL(c128_to_bricks_kernel, dim3(blocks), dim3(NT), 0, stream, g, n, wrap, gb);
  return check_launch("c128_to_bricks_kernel");
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
                     stream, A, K, B, K, M, N, K, K, 0ll, C, N, ep, NormSide{});
  return check_launch("gemm_nt_kernel");
}

int cryomi_ncc_search(const cryomi_ncc_search_args* a, void* workspace, size_t workspace_bytes, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(a != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_search: NULL args");
  CRYOMI_REQUIRE(box_supported(a->box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_ncc_search: box must be even, 8 <= box <= 512 and of the form 2^a 3^b 5^c");
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
  CRYOMI_REQUIRE((a->flags & ~CRYOMI_GATHER_FILL) == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_ncc_search: args->flags may only hold CRYOMI_GATHER_FILL");
  const int wrap = (a->flags & CRYOMI_GATHER_FILL) ? 0 : 1;
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

  hipLaunchKernelGGL(prepare_obs_kernel, dim3((unsigned)((L.kp2 + NT - 1) / NT), (unsigned)(a->n_images + 1)),
                     dim3(NT), 0, stream, y, n, a->n_images, a->n_offsets, L.rows_p, a->image_params, a->offsets,
                     L.kp2, L.kp2, reinterpret_cast<float2*>(A));
  int st = check_launch("prepare_obs_kernel");
  if (st) return st;
  hipLaunchKernelGGL(prepare_ctf_kernel, dim3(L.img_p), dim3(NT), 0, stream, y, n, a->n_images, a->image_params,
                     L.kp2, L.kp2, reinterpret_cast<float2*>(D), yn2);
  if ((st = check_launch("prepare_ctf_kernel"))) return st;
  // Templates: built per chunk of ROT_CHUNK rotations into the workspace, or
  // read from a caller's bank (cryomi_ncc_template_bank rows). With a bank the
  // template buffer is free, so it holds sum w CTF^2 |P|^2 for chunks as large
  // as it allows (one chunk for the usual sizes): fewer, fuller launches.
  const bool banked = a->bank != nullptr;
  const int ldb = banked ? cand_ld(n) : L.kp;
  int chunk = L.chunk, ld_tn2 = L.chunk_p;
  float* tn2_buf = tn2;
  if (banked) {
    const size_t room = (size_t)L.chunk_p * L.kp / (size_t)L.img_p;  // floats per image row of tn2
    long long c = (long long)(room / BN) * BN;
    if (c > a->n_rotations) c = ((a->n_rotations + BN - 1) / BN) * BN;
    if (c < L.chunk) c = L.chunk;  // never below the ordinary chunk (then tn2 keeps its own buffer)
    if (c > L.chunk) {
      chunk = (int)c;
      ld_tn2 = (int)c;
      tn2_buf = Bt;
    }
  }
  // Few rows (a fallback batch of up to 20 images, with a bank): the GEMMs would run at most four
  // 128-row tiles over the whole bank, too few workgroups to fill the GPU. The norms then come
  // with the contraction (or from template_norms_kernel) and the contraction is split along K
  // into the free template buffer.
  const bool skinny = banked && L.rows_p <= 4 * BM;
  const size_t split_used = tn2_buf == Bt ? align256((size_t)L.img_p * ld_tn2 * sizeof(float)) : 0;
  float* split_buf = reinterpret_cast<float*>(reinterpret_cast<char*>(Bt) + split_used);
  const size_t split_room = ((size_t)L.chunk_p * L.kp * sizeof(float) - split_used) / sizeof(float);
  for (int r0 = 0; r0 < a->n_rotations; r0 += chunk) {
    const int cnt = (a->n_rotations - r0) < chunk ? (a->n_rotations - r0) : chunk;
    const float* B = Bt;
    if (banked) {
      B = a->bank + (size_t)r0 * ldb;
    } else {
      hipLaunchKernelGGL(template_kernel, dim3((L.kp2 + NT - 1) / NT, L.chunk_p), dim3(NT), 0, stream,
                         static_cast<const float2*>(a->grid_c64), n, a->rotations, r0, cnt, wrap, L.kp2, L.kp2,
                         reinterpret_cast<float2*>(Bt));
      if ((st = check_launch("template_kernel"))) return st;
    }
    NccEpilogue none{};
    const int ldc = ((cnt + BN - 1) / BN) * BN;
    const long long split_stride = (long long)L.rows_p * ldc;
    const long long norm_stride = (long long)NORM_MAX_IMG * ldc;
    // few rows: the contraction is split along K; with few images its first row of blocks also
    // forms the template norms (no separate pass over the bank)
    const bool split = skinny && (size_t)NCC_SPLIT * (split_stride + norm_stride) <= split_room;
    const bool fused_norms = split && a->n_images <= NORM_MAX_IMG;
    float* norm_part = split_buf + (size_t)NCC_SPLIT * split_stride;
    if (fused_norms) {
      // norms come with the contraction below
    } else if (skinny && a->n_images <= NORM_MAX_IMG) {
      hipLaunchKernelGGL(template_norms_kernel, dim3((unsigned)((cnt + NT / 64 - 1) / (NT / 64))), dim3(NT), 0,
                         stream, D, L.kp, B, ldb, a->n_images, cnt, L.kp, tn2_buf, ld_tn2);
      if ((st = check_launch("template_norms_kernel"))) return st;
    } else {
      hipLaunchKernelGGL(gemm_nt_kernel<GEMM_SQUARE_B>, dim3((cnt + BN - 1) / BN, L.img_p / BM), dim3(GEMM_THREADS),
                         0, stream, D, L.kp, B, ldb, a->n_images, cnt, L.kp, L.kp, 0ll, tn2_buf, ld_tn2, none,
                         NormSide{});
      if ((st = check_launch("gemm_nt_kernel<square>"))) return st;
    }
    NccEpilogue ep{a->n_offsets, yn2, tn2_buf, ld_tn2, r0, r0 / BN, L.n_cb, partial};
    if (split) {
      // K split across grid z into partial products (and norms), then summed with the NCC epilogue
      const int ksplit = ((L.kp / KT + NCC_SPLIT - 1) / NCC_SPLIT) * KT;
      const int nz = (L.kp + ksplit - 1) / ksplit;
      const dim3 grid((cnt + BN - 1) / BN, L.rows_p / BM, nz);
      if (fused_norms) {
        const NormSide ns{D, L.kp, a->n_images, norm_part, ldc, norm_stride};
        hipLaunchKernelGGL(gemm_nt_kernel<GEMM_STORE_NORMS>, grid, dim3(GEMM_THREADS), 0, stream, A, L.kp, B, ldb,
                           L.rows, cnt, L.kp, ksplit, split_stride, split_buf, ldc, none, ns);
      } else {
        hipLaunchKernelGGL(gemm_nt_kernel<GEMM_STORE>, grid, dim3(GEMM_THREADS), 0, stream, A, L.kp, B, ldb, L.rows,
                           cnt, L.kp, ksplit, split_stride, split_buf, ldc, none, NormSide{});
      }
      if ((st = check_launch("gemm_nt_kernel<split>"))) return st;
      hipLaunchKernelGGL(ncc_splitk_kernel, dim3((cnt + BN - 1) / BN, L.rows), dim3(BN), 0, stream, split_buf, nz,
                         split_stride, ldc, cnt, ep, fused_norms ? norm_part : nullptr, norm_stride);
      if ((st = check_launch("ncc_splitk_kernel"))) return st;
    } else {
      hipLaunchKernelGGL(gemm_nt_kernel<GEMM_NCC>, dim3((cnt + BN - 1) / BN, L.rows_p / BM), dim3(GEMM_THREADS), 0,
                         stream, A, L.kp, B, ldb, L.rows, cnt, L.kp, L.kp, 0ll, nullptr, 0, ep, NormSide{});
      if ((st = check_launch("gemm_nt_kernel<ncc>"))) return st;
    }
  }
  hipLaunchKernelGGL(select_kernel, dim3(a->n_images), dim3(NT), 0, stream, partial, a->n_offsets, L.n_cb,
                     a->n_rotations, window, cand, cand_count, a->overflow);
  if ((st = check_launch("select_kernel"))) return st;
  double2* gb = reinterpret_cast<double2*>(ws + L.gb);
  if ((st = launch_c128_bricks(static_cast<const double2*>(a->grid_c128), n, wrap, gb, stream))) return st;
  return launch_rescore(gb, n, y, a->image_params, a->rotations, a->offsets,
                        a->n_offsets, yn2, cand, cand_count, a->n_images, 0, wrap, reinterpret_cast<int*>(ws + L.flat_off),
                        reinterpret_cast<double*>(ws + L.flat_score), a->best_index, a->best_score, stream);
}

size_t cryomi_ncc_bank_row_floats(int box) {
  if (!box_supported(box)) return 0;
  return (size_t)cand_ld(box);
}

int cryomi_ncc_template_bank(const void* grid_c64, int box, const double* rotations, int n_rotations,
                             unsigned flags, float* bank, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE((flags & ~CRYOMI_GATHER_FILL) == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_ncc_template_bank: flags may only hold CRYOMI_GATHER_FILL");
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_ncc_template_bank: box must be even, 8 <= box <= 512 and of the form 2^a 3^b 5^c");
  CRYOMI_REQUIRE(n_rotations >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_template_bank: n_rotations < 0");
  if (n_rotations == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(grid_c64 && rotations && bank, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_template_bank: NULL pointer");
  const int kp2 = cand_kp(box) / 2, ld2 = cand_ld(box) / 2;
  const int chunk = 32768;
  for (int r0 = 0; r0 < n_rotations; r0 += chunk) {
    const int cnt = (n_rotations - r0) < chunk ? (n_rotations - r0) : chunk;
    hipLaunchKernelGGL(template_kernel, dim3((kp2 + NT - 1) / NT, cnt), dim3(NT), 0, stream,
                       static_cast<const float2*>(grid_c64), box, rotations, r0, cnt,
                       (flags & CRYOMI_GATHER_FILL) ? 0 : 1, kp2, ld2,
                       reinterpret_cast<float2*>(bank) + (size_t)r0 * ld2);
    const int st = check_launch("template_kernel");
    if (st) return st;
  }
  return CRYOMI_OK;
}

int cryomi_pose_candidates(const float* hyp, long long hyp_stride, int n_hyp, const long long* center_index,
                           int n_offsets, int n_images, const double* rotations, int n_rotations,
                           double max_angle_deg, int cap, int* cand, int* cand_count, int* overflow,
                           ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n_images >= 0 && n_rotations >= 0 && cap >= 1 && cap <= MAX_CAP, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_pose_candidates: need n_images, n_rotations >= 0 and 1 <= cap <= 8192");
  CRYOMI_REQUIRE(center_index ? (n_offsets >= 1) : (n_hyp >= 1 && n_hyp <= 8 && hyp_stride >= 9 * (long long)n_hyp),
                 CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_pose_candidates: need 1 <= n_hyp <= 8 with hyp_stride >= 9 n_hyp, or n_offsets >= 1");
  CRYOMI_REQUIRE(max_angle_deg >= 0.0 && max_angle_deg <= 180.0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_pose_candidates: max_angle_deg must be in [0, 180]");
  if (n_images == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE((hyp || center_index) && rotations && cand && cand_count && overflow, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_pose_candidates: NULL pointer");
  // trace(R^T G) = 1 + 2 cos(angle)
  const float thr = (float)(1.0 + 2.0 * cos(max_angle_deg * M_PI / 180.0));
  hipLaunchKernelGGL(pose_candidates_kernel, dim3(n_images), dim3(NT), 0, stream, hyp, hyp_stride,
                     center_index ? 1 : n_hyp, center_index, n_offsets, rotations, n_rotations, thr, cap, cand,
                     cand_count, overflow);
  return check_launch("pose_candidates_kernel");
}

size_t cryomi_ncc_search_candidates_workspace_bytes(int box, int n_images, int n_offsets, int cap) {
  if (!box_supported(box) || n_images <= 0 || n_offsets <= 0 || n_offsets > SC_BM || cap < 1 || cap > MAX_CAP) return 0;
  return cand_layout(box, n_images, n_offsets, cap).total;
}

int cryomi_ncc_search_candidates(const cryomi_ncc_search_args* a, const cryomi_ncc_candidate_args* c,
                                 void* workspace, size_t workspace_bytes, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(a != nullptr && c != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_search_candidates: NULL args");
  CRYOMI_REQUIRE(box_supported(a->box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_ncc_search_candidates: box must be even, 8 <= box <= 512 and of the form 2^a 3^b 5^c");
  CRYOMI_REQUIRE(a->n_images >= 0 && a->n_rotations > 0 && a->n_offsets >= 1 && a->n_offsets <= SC_BM,
                 CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_ncc_search_candidates: need n_images >= 0, n_rotations > 0, 1 <= n_offsets <= 32");
  CRYOMI_REQUIRE(c->cap >= 1 && c->cap <= MAX_CAP && (c->flags & ~(CRYOMI_NCC_REUSE_PREPARED | CRYOMI_NCC_MERGE)) == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_search_candidates: need 1 <= cap <= 8192 and known flags");
  if (a->n_images == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(a->grid_c128 && a->observed_rfft && a->image_params && a->rotations && a->offsets && a->best_index &&
                     a->best_score && a->overflow && c->bank && c->cand && c->cand_count,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_search_candidates: NULL pointer");
  CRYOMI_REQUIRE((long long)a->n_rotations * a->n_offsets < (1ll << 62) &&
                     (long long)a->n_images * a->n_offsets < (1ll << 31),
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_ncc_search_candidates: grid too large");
  CRYOMI_REQUIRE((a->flags & ~CRYOMI_GATHER_FILL) == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_ncc_search_candidates: args->flags may only hold CRYOMI_GATHER_FILL");
  const int wrap = (a->flags & CRYOMI_GATHER_FILL) ? 0 : 1;
  const int n = a->box;
  const CandLayout L = cand_layout(n, a->n_images, a->n_offsets, c->cap);
  CRYOMI_REQUIRE(workspace && workspace_bytes >= L.total, CRYOMI_ERR_WORKSPACE,
                 "cryomi_ncc_search_candidates: workspace smaller than cryomi_ncc_search_candidates_workspace_bytes()");
  char* ws = static_cast<char*>(workspace);
  float* A = reinterpret_cast<float*>(ws + L.a);
  float* D = reinterpret_cast<float*>(ws + L.d);
  double* yn2 = reinterpret_cast<double*>(ws + L.yn2);
  Top2* partial = reinterpret_cast<Top2*>(ws + L.partial);
  long long* cand = reinterpret_cast<long long*>(ws + L.cand);
  int* cand_count = reinterpret_cast<int*>(ws + L.cand_count);
  const double2* y = static_cast<const double2*>(a->observed_rfft);
  const float window = a->tie_window > 0 ? (float)a->tie_window : 2e-4f;
  const int rows = a->n_images * a->n_offsets;
  int st;
  double2* gb = reinterpret_cast<double2*>(ws + L.gb);
  if (!(c->flags & CRYOMI_NCC_REUSE_PREPARED)) {
    if ((st = launch_c128_bricks(static_cast<const double2*>(a->grid_c128), n, wrap, gb, stream))) return st;
    hipLaunchKernelGGL(prepare_obs_kernel, dim3((unsigned)((L.kp / 2 + NT - 1) / NT), (unsigned)(a->n_images + 1)),
                       dim3(NT), 0, stream, y, n, a->n_images, a->n_offsets, rows, a->image_params, a->offsets,
                       L.kp / 2, L.ld / 2, reinterpret_cast<float2*>(A));
    if ((st = check_launch("prepare_obs_kernel"))) return st;
    hipLaunchKernelGGL(prepare_ctf_kernel, dim3(a->n_images), dim3(NT), 0, stream, y, n, a->n_images, a->image_params,
                       L.kp / 2, L.ld / 2, reinterpret_cast<float2*>(D), yn2);
    if ((st = check_launch("prepare_ctf_kernel"))) return st;
  }
  int* work_off = reinterpret_cast<int*>(ws + L.work_off);
  int* flat_off = reinterpret_cast<int*>(ws + L.flat_off);
  double* flat_score = reinterpret_cast<double*>(ws + L.flat_score);
  hipLaunchKernelGGL(work_prefix_kernel, dim3(1), dim3(SCAN_NT), 0, stream, c->cand_count, a->n_images, c->cap, SC_BN,
                     work_off);
  if ((st = check_launch("work_prefix_kernel"))) return st;
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cached[dev] == 0) cached[dev] = persistent_blocks((const void*)subset_ncc_kernel, NT);
  hipLaunchKernelGGL(subset_ncc_kernel, dim3(cached[dev]), dim3(NT), 0, stream, A, D, yn2, c->bank, L.kp, L.ld,
                     c->cand, c->cand_count, c->cap, a->n_rotations, a->n_offsets, L.n_cb, a->n_images, work_off,
                     partial);
  if ((st = check_launch("subset_ncc_kernel"))) return st;
  const int merge = (c->flags & CRYOMI_NCC_MERGE) ? 1 : 0;
  hipLaunchKernelGGL(select_subset_kernel, dim3(a->n_images), dim3(NT), 0, stream, partial, a->n_offsets, L.n_cb,
                     c->cand, c->cand_count, c->cap, a->n_rotations, window, merge, a->best_index, cand, cand_count,
                     a->overflow);
  if ((st = check_launch("select_subset_kernel"))) return st;
  if ((st = launch_rescore(gb, n, y, a->image_params, a->rotations,
                           a->offsets, a->n_offsets, yn2, cand, cand_count, a->n_images, merge, wrap, flat_off,
                           flat_score, a->best_index, a->best_score, stream)))
    return st;
  if (c->rescored) {
    hipLaunchKernelGGL(copy_count_kernel, dim3((a->n_images + NT - 1) / NT), dim3(NT), 0, stream, cand_count,
                       a->n_images, c->rescored);
    return check_launch("copy_count_kernel");
  }
  return CRYOMI_OK;
}

}  // extern "C"
