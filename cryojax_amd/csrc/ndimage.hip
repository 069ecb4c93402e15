// Fourier-crop (Hartley) downsampling of real images on gfx950 as a chain of
// batched FP32 MFMA contractions with separable cas operators.
//
// With centred indices k~ = k - K//2 on an axis of length K, cryoJAX's
// fftshift(fftn(ifftshift(x))) is X[q] = sum_m x[m] exp(-2 pi i q~ m~ / N)
// per axis, so the Hartley array Re X - Im X of an image is a sum of cas
// products, cas(a + b) = cos a cas b + sin a cas(-b). The downsampled image
// (h, w) -> (oh, ow) is, per image x,
//   Zt  = [Ce . E(x)^T ; So . O(x)^T]                           (2 ow, h)
//   H   = sum_b U_b . Zt_b^T       U = [Cr + Sr ; Cr - Sr]      (2 x (oh, h))
//   T   = Q . H^T                  Q = [Qc + Qs ; Qc - Qs]      (2 ow, ow)
//   y   = sum_a P_a . T_a^T        P = s [Pc ; Ps]              (2 x (oh, oh))
// where E(x)[r][t] = x[r][t~ = t] + x[r][t~ = -t] and O(x)[r][t] =
// x[r][t~ = t] - x[r][t~ = -t] (t < w/2 + 1) fold each image row about its
// centre (the cosine is even and the sine odd in m~, so the first
// contraction runs over half the columns; entries without a mirror are taken
// once), Ce / So the column cosines and sines on the folded index, H the
// centre-cropped
// Hartley array (x Cc^T = E(x) Ce^T and x Sc^T = O(x) So^T recombine in U),
// Qc / Pc the coarse-grid cosines and s = 1 / (oh ow) (1 / (h w) when the
// mean is preserved).
// rfftn(ifftshift(y)) adds
//   Gt  = R . y^T                  R = [cos ; sin] of k2 p2~    (2 kw, ow)
//   F   = sum_s E_s . Gt_s^T       E_0 = [Ca ; -Sa], E_1 = [-Sa ; -Ca]
// with the real parts in rows [0, oh) and imaginary parts in rows [oh, 2 oh)
// of F, stored interleaved as complex64.
//
// Every step is C_i = sum_s A_{i,s} B_{i,s}^T, one workgroup per (image,
// BM x BN tile), 4 waves as 2 x 2 of v_mfma_f32_16x16x4f32 sub-tiles, K
// steps of 16 staged through LDS with a register prefetch of the next step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/cryomi_ndimage.h"
#include "common.hpp"
#include "stockham.hpp"

namespace cryomi {
namespace {

constexpr int NT = 256;
constexpr int KT = 16;
constexpr int MAX_DIM = 4096;
constexpr size_t CHUNK_FLOATS = size_t(1) << 26;  // intermediate floats per chunk (256 MiB)

typedef float floatx4 __attribute__((ext_vector_type(4)));

struct GemmDesc {
  const float* a;
  long long a_batch, a_seg;
  int lda;
  const float* b;
  long long b_batch, b_seg;
  int ldb;
  float* c;
  long long c_batch;
  int ldc;
  int m, n, k, nseg, batch;
  int cplx;    // > 0: rows [0, cplx) real, [cplx, 2 cplx) imaginary parts, stored interleaved (row stride 2 n)
  int a_pad4;  // A rows are zero-padded to a multiple of 4 floats: whole float4 groups may be loaded at the K tail
};

enum { B_PLAIN = 0, B_VEC = 1 };

// P consecutive entries kk .. kk + P - 1 of one operand row, zero past k or on
// rows outside the operand, in two steps: issue_run loads unconditionally (an
// invalid run reads the row's first entries, which exist since k >= 1 and a
// float4 run needs k >= 4), mask_run zeroes the invalid entries once the
// data is needed. A branch around a load, or arithmetic on its result right
// after it, makes the compiler wait for the load there, which would
// serialise the prefetch of the next K step with this step's MFMAs.
template <int P, bool VEC>
__device__ __forceinline__ void issue_run(const float* __restrict__ p, bool row_ok, int kk, int k, float* r) {
  if constexpr (VEC) {
#pragma unroll
    for (int q = 0; q < P; q += 4) {
      const bool ok = row_ok && kk + q < k;
      const float4 v = *reinterpret_cast<const float4*>(p + (ok ? kk + q : 0));
      r[q] = v.x;
      r[q + 1] = v.y;
      r[q + 2] = v.z;
      r[q + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < P; ++q) r[q] = p[(row_ok && kk + q < k) ? kk + q : 0];
  }
}

template <int P, bool VEC>
__device__ __forceinline__ void mask_run(bool row_ok, int kk, int k, float* r) {
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const bool ok = row_ok && (VEC ? kk + (q & ~3) : kk + q) < k;
    r[q] = ok ? r[q] : 0.f;
  }
}

template <int BM, int BN, bool VA, int BMODE>
__global__ __launch_bounds__(NT) void bgemm_nt_kernel(GemmDesc g) {
  constexpr int PA = BM / 16, PB = BN / 16;  // floats per thread per K step
  constexpr int TA = 16 / PA, TB = 16 / PB;  // threads per tile row
  constexpr int TI = BM / 32, TJ = BN / 32;  // MFMA sub-tiles per wave
  __shared__ float As[2][KT][BM + 16];
  __shared__ float Bs[2][KT][BN + 16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tiles_m = (g.m + BM - 1) / BM, tiles_n = (g.n + BN - 1) / BN;
  const int per_img = tiles_m * tiles_n;
  const int img = blockIdx.x / per_img;
  const int rem = blockIdx.x - img * per_img;
  const int row0 = (rem / tiles_n) * BM, col0 = (rem % tiles_n) * BN;
  const int ar = tid / TA, ak = (tid % TA) * PA;
  const int br = tid / TB, bk = (tid % TB) * PB;
  const bool a_ok = row0 + ar < g.m, b_ok = col0 + br < g.n;
  const float* abase = g.a + img * g.a_batch + (long long)(a_ok ? row0 + ar : 0) * g.lda;
  const float* bbase = g.b + img * g.b_batch + (long long)(b_ok ? col0 + br : 0) * g.ldb;
  const int ktiles = (g.k + KT - 1) / KT;
  const int nt = ktiles * g.nseg;

  floatx4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  constexpr bool VB = BMODE == B_VEC;
  const int ka = g.a_pad4 ? (g.k + 3) & ~3 : g.k;
  float ra[PA], rb[PB];
  auto gload = [&](int t) {
    const int s = t / ktiles, k0 = (t - s * ktiles) * KT;
    issue_run<PA, VA>(abase + s * g.a_seg, a_ok, k0 + ak, ka, ra);
    issue_run<PB, VB>(bbase + s * g.b_seg, b_ok, k0 + bk, g.k, rb);
  };
  auto sstore = [&](int t, int buf) {
    const int k0 = (t % ktiles) * KT;
    mask_run<PA, VA>(a_ok, k0 + ak, ka, ra);
    mask_run<PB, VB>(b_ok, k0 + bk, g.k, rb);
#pragma unroll
    for (int q = 0; q < PA; ++q) As[buf][ak + q][ar] = ra[q];
#pragma unroll
    for (int q = 0; q < PB; ++q) Bs[buf][bk + q][br] = rb[q];
  };
  gload(0);
  sstore(0, 0);
  __syncthreads();
  for (int t = 0; t < nt; ++t) {
    const int cur = t & 1;
    gload(t + 1 < nt ? t + 1 : t);  // unconditional: a branch here would make the MFMAs wait for the loads
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ks = 0; ks < KT / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      float av[TI], bv[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) av[i] = As[cur][kk][wm * (BM / 2) + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TJ; ++j) bv[j] = Bs[cur][kk][wn * (BN / 2) + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (t + 1 < nt) sstore(t + 1, cur ^ 1);
    __syncthreads();
  }
  // D element (row 4 (lane / 16) + r, column lane % 16) of each 16 x 16 sub-tile
  float* cbase = g.c + img * g.c_batch;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wm * (BM / 2) + i * 16 + 4 * (lane >> 4) + r;
        const int col = col0 + wn * (BN / 2) + j * 16 + (lane & 15);
        if (row < g.m && col < g.n) {
          if (g.cplx > 0) {
            const int part = row >= g.cplx ? 1 : 0;
            cbase[((long long)(row - part * g.cplx) * g.n + col) * 2 + part] = acc[i][j][r];
          } else {
            cbase[(long long)row * g.ldc + col] = acc[i][j][r];
          }
        }
      }
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <int BM, int BN, bool VA>
int launch_tiles(const GemmDesc& g, int bmode, hipStream_t stream) {
  const long long blocks = (long long)g.batch * ((g.m + BM - 1) / BM) * ((g.n + BN - 1) / BN);
  CRYOMI_REQUIRE(blocks <= 0x7fffffffLL, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_bgemm_nt: too many tiles");
  const dim3 grid((unsigned)blocks), block(NT);
  switch (bmode) {
    case B_VEC: hipLaunchKernelGGL((bgemm_nt_kernel<BM, BN, VA, B_VEC>), grid, block, 0, stream, g); break;
    default: hipLaunchKernelGGL((bgemm_nt_kernel<BM, BN, VA, B_PLAIN>), grid, block, 0, stream, g); break;
  }
  return check_launch("bgemm_nt_kernel");
}

template <int BM, int BN>
int launch_va(const GemmDesc& g, bool va, int bmode, hipStream_t stream) {
  return va ? launch_tiles<BM, BN, true>(g, bmode, stream) : launch_tiles<BM, BN, false>(g, bmode, stream);
}

int launch_gemm(const GemmDesc& g, hipStream_t stream) {
  if (g.batch == 0 || g.m == 0 || g.n == 0) return CRYOMI_OK;
  const bool va = (g.k % 4 == 0 || g.a_pad4) && g.lda % 4 == 0 && g.a_batch % 4 == 0 && g.a_seg % 4 == 0 &&
                  aligned16(g.a);
  const bool vb = g.k % 4 == 0 && g.ldb % 4 == 0 && g.b_batch % 4 == 0 && g.b_seg % 4 == 0 && aligned16(g.b);
  const int bmode = vb ? B_VEC : B_PLAIN;
  const bool wide_m = g.m > 64, wide_n = g.n > 64;
  if (wide_m && wide_n) return launch_va<128, 128>(g, va, bmode, stream);
  if (wide_m) return launch_va<128, 64>(g, va, bmode, stream);
  if (wide_n) return launch_va<64, 128>(g, va, bmode, stream);
  return launch_va<64, 64>(g, va, bmode, stream);
}

// First contraction with the row fold: for image i and coarse column j < ow,
//   Zt[j][r]      = sum_t Ce[j][t] E[r][t],  E[r][t] = x[r][c + t] + x[r][c - t] (t > 0; x[r][c] at t = 0)
//   Zt[ow + j][r] = sum_t So[j][t] O[r][t],  O[r][t] = x[r][c + t] - x[r][c - t]
// with c = w / 2. For odd w every t <= c pairs two columns (K = c + 1). For
// even w the contraction runs over t < c (K = c) and the unpaired column
// x[r][0] (m~ = -w/2, where the cosine is cos(pi j~) = +-1 and the sine 0)
// is added in the epilogue. One workgroup per (image, 64 coarse columns, 128
// image rows): the image tile is read once, folded both ways into LDS, waves
// 0-1 contract Ce with E and waves 2-3 So with O. VEC: w % 8 == 0 and aligned
// rows, so each mirrored run is three aligned float4 loads. LDS is single
// buffered (the next step's global loads are in flight during the MFMAs) so
// that four workgroups fit a CU.
constexpr int FBM = 64, FBN = 128;

// Raw loads of one folded run t = kk .. kk + 7 of an image row (centre c = w / 2).
// VEC (w % 8 == 0, so c and kk are multiples of 4 and kk < k = c keeps every
// ascending run inside the row): ascending c + kk .. + 7 as two float4,
// descending c - kk - 8 .. c - kk + 3 as three; invalid runs read at c.
// Scalar: r[0..7] ascending, r[8..15] descending.
struct FoldRaw {
  float r[20];
};

template <bool VEC>
__device__ __forceinline__ void fold_issue(const float* __restrict__ row, bool row_ok, int kk, int k, int w,
                                           FoldRaw& f) {
  const int c = w / 2;
  if constexpr (VEC) {
    const bool ok = row_ok && kk < k;
    const float4 a0 = *reinterpret_cast<const float4*>(row + c + (ok ? kk : 0));
    const float4 a1 = *reinterpret_cast<const float4*>(row + c + (ok ? kk + 4 : 0));
    const float4 d0 = *reinterpret_cast<const float4*>(row + (ok ? c - kk : c));
    const float4 d1 = *reinterpret_cast<const float4*>(row + (ok && c - kk - 4 >= 0 ? c - kk - 4 : c));
    const float4 d2 = *reinterpret_cast<const float4*>(row + (ok && c - kk - 8 >= 0 ? c - kk - 8 : c));
    const float4 v[5] = {a0, a1, d0, d1, d2};
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      f.r[4 * i] = v[i].x;
      f.r[4 * i + 1] = v[i].y;
      f.r[4 * i + 2] = v[i].z;
      f.r[4 * i + 3] = v[i].w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int t = kk + q;
      const bool ok = row_ok && t < k;  // t < k <= c + 1 keeps c - t >= 0 and c + t < w
      f.r[q] = row[ok ? c + t : c];
      f.r[8 + q] = row[ok ? c - t : c];
    }
  }
}

// Folded entries e = asc + dsc (t > 0), o = asc - dsc, zero past k / outside the image
template <bool VEC>
__device__ __forceinline__ void fold_finish(const FoldRaw& f, bool row_ok, int kk, int k, int w, float* e, float* o) {
  const int c = w / 2;
  float asc[8], dsc[8];
  if constexpr (VEC) {
    const bool ok = row_ok && kk < k;
    const bool ok1 = ok && c - kk - 4 >= 0, ok2 = ok && c - kk - 8 >= 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) asc[q] = ok ? f.r[q] : 0.f;
    dsc[0] = ok ? f.r[8] : 0.f;  // d0.x: c - kk
#pragma unroll
    for (int q = 1; q <= 4; ++q) dsc[q] = ok1 ? f.r[12 + 4 - q] : 0.f;  // d1 holds c - kk - 4 .. - 1
#pragma unroll
    for (int q = 5; q < 8; ++q) dsc[q] = ok2 ? f.r[16 + 8 - q] : 0.f;   // d2 holds c - kk - 8 .. - 5
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bool ok = row_ok && kk + q < k;
      asc[q] = ok ? f.r[q] : 0.f;
      dsc[q] = ok ? f.r[8 + q] : 0.f;
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int t = kk + q;
    const bool ok = t < k;
    e[q] = ok ? asc[q] + (t > 0 ? dsc[q] : 0.f) : 0.f;
    o[q] = ok ? asc[q] - dsc[q] : 0.f;
  }
}

template <bool VEC>
__global__ __launch_bounds__(NT) void fold_gemm_kernel(const float* __restrict__ ce, const float* __restrict__ so,
                                                       int lde, const float* __restrict__ x, int h, int w, int ow,
                                                       float* __restrict__ zt, long long zt_batch) {
  __shared__ float As[KT][2 * FBM + 16];
  __shared__ float Es[KT][FBN + 16];
  __shared__ float Os[KT][FBN + 16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;  // wm 0: Ce x E, 1: So x O
  const int tiles_m = (ow + FBM - 1) / FBM, tiles_n = (h + FBN - 1) / FBN;
  const int per_img = tiles_m * tiles_n;
  const int img = blockIdx.x / per_img;
  const int rem = blockIdx.x - img * per_img;
  const int row0 = (rem / tiles_n) * FBM, col0 = (rem % tiles_n) * FBN;
  const bool even = (w & 1) == 0;
  const int k = even ? w / 2 : w / 2 + 1;
  // A: tile rows 0-63 from Ce, 64-127 from So; 8 K entries per thread
  const int ar = tid >> 1, ak = (tid & 1) * 8;
  const int arow = row0 + (ar & (FBM - 1));
  const bool a_ok = arow < ow;
  const float* abase = (ar < FBM ? ce : so) + (size_t)(a_ok ? arow : 0) * lde;
  // B: 128 image rows, 8 folded K entries per thread
  const int br = tid >> 1, bk = (tid & 1) * 8;
  const bool b_ok = col0 + br < h;
  const float* ximg = x + (size_t)img * h * w;
  const float* xrow = ximg + (size_t)(b_ok ? col0 + br : 0) * w;
  const int nt = (k + KT - 1) / KT;
  const int kpad = (k + 3) & ~3;  // operator rows are zero-padded to a multiple of 4 (lde >= kpad)

  floatx4 acc[4][4];  // per wave: 64 coarse columns x 64 image rows
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  float ra[8];
  FoldRaw fr;
  auto gload = [&](int t) {
    const int k0 = t * KT;
    issue_run<8, true>(abase, a_ok, k0 + ak, kpad, ra);
    fold_issue<VEC>(xrow, b_ok, k0 + bk, k, w, fr);
  };
  gload(0);
  for (int t = 0; t < nt; ++t) {
    {
      float re[8], ro[8];
      mask_run<8, true>(a_ok, t * KT + ak, kpad, ra);
      fold_finish<VEC>(fr, b_ok, t * KT + bk, k, w, re, ro);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        As[ak + q][ar] = ra[q];
        Es[bk + q][br] = re[q];
        Os[bk + q][br] = ro[q];
      }
    }
    __syncthreads();
    gload(t + 1 < nt ? t + 1 : t);  // unconditional: a branch here would make the MFMAs wait for the loads
    __builtin_amdgcn_sched_barrier(0);
    const float(*bs)[FBN + 16] = wm ? Os : Es;
#pragma unroll
    for (int ks = 0; ks < KT / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      float av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[kk][wm * FBM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = bs[kk][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
  }
  float* out = zt + img * zt_batch + (size_t)wm * ow * h;
  const bool nyq = even && wm == 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = col0 + wn * 64 + j * 16 + (lane & 15);
    const float xv = ximg[(size_t)(col < h ? col : 0) * w];
    const float x0 = (nyq && col < h) ? xv : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int jr = row0 + i * 16 + 4 * (lane >> 4) + r;
        // cos(pi j~) with j~ = jr - ow / 2
        const float sgn = ((jr - ow / 2) & 1) ? -1.f : 1.f;
        if (jr < ow && col < h) out[(size_t)jr * h + col] = acc[i][j][r] + sgn * x0;
      }
  }
}

int launch_fold(const float* ce, const float* so, int lde, const float* x, int nb, int h, int w, int ow, float* zt,
                long long zt_batch, hipStream_t stream) {
  const long long blocks = (long long)nb * ((ow + FBM - 1) / FBM) * ((h + FBN - 1) / FBN);
  CRYOMI_REQUIRE(blocks <= 0x7fffffffLL, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_fourier_crop_downsample: too many tiles");
  const bool vec = w % 8 == 0 && aligned16(x);
  if (vec)
    hipLaunchKernelGGL(fold_gemm_kernel<true>, dim3((unsigned)blocks), dim3(NT), 0, stream, ce, so, lde, x, h, w, ow,
                       zt, zt_batch);
  else
    hipLaunchKernelGGL(fold_gemm_kernel<false>, dim3((unsigned)blocks), dim3(NT), 0, stream, ce, so, lde, x, h, w, ow,
                       zt, zt_batch);
  return check_launch("fold_gemm_kernel");
}

// ---------------------------------------------------------------- operators

struct OpLayout {
  size_t ce, so, u, q, p, r, e, total;  // offsets in floats
  int ne, lde;                          // folded column count and its padded row stride
};

size_t up64(size_t x) { return (x + 63) & ~size_t(63); }

OpLayout op_layout(int h, int w, int oh, int ow, unsigned flags) {
  OpLayout L{};
  const size_t kw = ow / 2 + 1;
  L.ne = w / 2 + 1;
  L.lde = (L.ne + 3) & ~3;
  size_t o = 0;
  L.ce = o;
  o += up64((size_t)ow * L.lde);
  L.so = o;
  o += up64((size_t)ow * L.lde);
  L.u = o;
  o += up64(2 * (size_t)oh * h);
  L.q = o;
  o += up64(2 * (size_t)ow * ow);
  L.p = o;
  o += up64(2 * (size_t)oh * oh);
  L.r = L.e = o;
  if (flags & CRYOMI_DS_OUTPUT_FOURIER) {
    L.r = o;
    o += up64(2 * kw * ow);
    L.e = o;
    o += up64(4 * (size_t)oh * oh);
  }
  L.total = o;
  return L;
}

// ---------------------------------------------------------------------------
// FFT path: square N x N images, N in {128, 256}, to O x O, O in {32, 64}, O <= N/2, either
// output space. One image per workgroup iteration (persistent), all in LDS:
//   A (rows)    : each real row, read at its ifftshifted columns as packed complex values, takes
//                 an N/2-point FFT and is unpacked into the bins kx = 0..O/2 only (window W1);
//   B (columns) : the O/2 + 1 kept columns take N-point FFTs along y (logical row n read at
//                 (n + N/2) mod N), and the bins ky = -O/2..O/2 are written back over the
//                 column's own W1 entries: F(ky, kx) at W1[(ky + O/2) K + kx];
//   C           : the cropped Hartley values H[i][j] = Re F - Im F at (ky, kx) = (i, j) - O/2,
//                 F(ky, -kx) = F*(-ky, kx) for the negative kx, take the centred O x O FFT (rows,
//                 then columns, each with its logical index n at (n + O/2) mod O on load and
//                 store) and leave as (Re - Im) times 1 / O^2 (1 / N^2 with preserve_mean).
// The next image's first rows are loaded into registers during stages B and C.
// ---------------------------------------------------------------------------
constexpr int DS_NT = 512;

template <int N, int O>
struct DsPlan {
  static constexpr int M = N / 2, K = O / 2 + 1;
  static constexpr int TA = M / 16, RA = DS_NT / TA, RSA = M + M / 16, ROUNDS_A = N / RA;
  static constexpr int NCH = RA * N / 4 / DS_NT;  // float4 chunks per thread per row round
  // stage B transforms columns kx = KX0..O/2 by FFT. When the K columns need one round more than
  // K - 1 would, the real DC column goes by direct sums instead (KX0 = 1).
  static constexpr int TB = N / 16, CB = DS_NT / TB, RSB = N + N / 16 + 1;
  static constexpr bool DC_DIRECT = (K + CB - 1) / CB > (K - 1 + CB - 1) / CB;
  static constexpr int KX0 = DC_DIRECT ? 1 : 0, ROUNDS_B = (K - KX0 + CB - 1) / CB;
  static constexpr int DC_OUT = O / 2 + 1, DC_SPLIT = DS_NT / DC_OUT;  // outputs ky = 0..O/2, threads each
  static constexpr int MC = O / 2, TC = MC / 16, RSC = MC + MC / 16;
  static constexpr int TD = O / 16, RSD = O + O / 16 + 1, GS = O + 1;
  static constexpr int cmax(int a, int b) { return a > b ? a : b; }
  static constexpr int X = cmax(cmax(RA * RSA, CB * RSB), cmax((O + 1) * RSC, (O + 1) * RSD));
  static constexpr int W1 = N * K;
  static constexpr size_t LDS = (size_t)(X + W1 + N) * sizeof(float2);
  static_assert(NCH * DS_NT * 4 == RA * N && ROUNDS_A * RA == N, "row rounds");
  static_assert(O * GS <= W1 && O * TC <= DS_NT && O * TD <= DS_NT && MC >= 16, "stage C fits");
  static_assert(DC_OUT * DC_SPLIT <= X && DC_SPLIT >= 1, "DC partial sums fit the exchange buffer");
};

// FOURIER: the output is rfftn(ifftshift(y)) of the real-space result y, complex64 (O, O/2 + 1),
// by a stage D on y (kept in LDS): packed real rows at their ifftshifted columns, O/2-point FFTs
// unpacked into bins 0..O/2, then O-point column FFTs with the logical row n at (n + O/2) mod O,
// stored in natural order.
template <int N, int O, bool FOURIER>
__global__ __launch_bounds__(DS_NT) void ds_fft_kernel(const float* __restrict__ x, int n, float scale,
                                                       float* __restrict__ out) {
  using P = DsPlan<N, O>;
  constexpr int M = P::M, K = P::K;
  extern __shared__ float2 ds_lds[];
  float2* X = ds_lds;
  float2* W1 = ds_lds + P::X;
  float2* tw = W1 + P::W1;  // e^{-2 pi i m / N}
  fill_twiddles<DS_NT>(tw, N);
  const int tid = threadIdx.x;
  auto load_round = [&](long long img, int round, float4 (&buf)[P::NCH]) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < P::NCH; ++i) {
      const int e = tid + i * DS_NT, rr = e / (N / 4), ch = e % (N / 4);
      buf[i] = reinterpret_cast<const float4*>(x + (img * N + round * P::RA + rr) * (long long)N)[ch];
    }
  };
  long long img = blockIdx.x;
  float4 pf[P::NCH];
  if (img < n) load_round(img, 0, pf);
  __syncthreads();  // twiddle table
  for (; img < n; img += gridDim.x) {
    // ---- A: rows -> W1[r][0..O/2]
    const int ra = tid / P::TA, ta = tid % P::TA;
    float2* zr = X + ra * P::RSA;
    for (int round = 0; round < P::ROUNDS_A; ++round) {
#pragma unroll
      for (int i = 0; i < P::NCH; ++i) {
        const int e = tid + i * DS_NT, rr = e / (N / 4), ch = e % (N / 4);
        const int m = ((ch * 4 - N / 2 + N) % N) / 2;
        X[rr * P::RSA + fpad(m)] = make_float2(pf[i].x, pf[i].y);
        X[rr * P::RSA + fpad(m + 1)] = make_float2(pf[i].z, pf[i].w);
      }
      if (round + 1 < P::ROUNDS_A)
        load_round(img, round + 1, pf);
      else if (img + gridDim.x < n)
        load_round(img + gridDim.x, 0, pf);
      __syncthreads();
      float2 v[16];
      load_pass<M, 16>(v, ta, zr);
      __syncthreads();
      fft_passes<M, 2, false>(v, ta, zr, tw);
      __syncthreads();
      const int row = round * P::RA + ra;
      for (int k = ta; k < K; k += P::TA) {
        const float2 a = zr[fpad(k)];
        const float2 bc = zr[fpad((M - k) & (M - 1))];
        const float2 e = make_float2(0.5f * (a.x + bc.x), 0.5f * (a.y - bc.y));
        const float2 od = make_float2(0.5f * (a.y + bc.y), -0.5f * (a.x - bc.x));  // (a - b*) / 2i
        W1[row * K + k] = zadd(e, zmul(od, tw[k]));
      }
      __syncthreads();
    }
    // ---- B: columns kx = KX0..O/2, N-point centred along y; keep ky = -O/2..O/2
    {
      constexpr int RL = last_radix<N>(), NL = last_ns<N>();
      const int cb = tid % P::CB, tb = tid / P::CB;
      for (int rb = 0; rb < P::ROUNDS_B; ++rb) {
        const int kx = rb * P::CB + cb + P::KX0;
        const bool act = kx < K;
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)
          v[q] = act ? W1[((stockham_in<N, 16>(tb, 0, q) + N / 2) & (N - 1)) * K + kx] : make_float2(0.f, 0.f);
        fft_passes<N, 1, true>(v, tb, X + cb * P::RSB, tw);
        if (act) {
#pragma unroll
          for (int u = 0; u < 16 / RL; ++u)
#pragma unroll
            for (int q = 0; q < RL; ++q) {
              const int k = stockham_out<N, RL, NL>(tb, u, q);
              const int ky = k < N / 2 ? k : k - N;
              if (ky >= -O / 2 && ky <= O / 2) W1[(ky + O / 2) * K + kx] = v[u * RL + q];
            }
        }
        __syncthreads();
      }
    }
    // ---- B, DC column: c0[r] = W1[r][0] is real (the row sums), so
    //   F(ky, 0) = sum_n c0[(n + N/2) mod N] e^{-2 pi i ky n / N} = (-1)^ky sum_r c0[r] e^{-2 pi i ky r / N}
    // for ky = 0..O/2 by DC_SPLIT partial sums each, and F(-ky, 0) = F*(ky, 0).
    if constexpr (P::DC_DIRECT) {
      const int ky = tid / P::DC_SPLIT, part = tid % P::DC_SPLIT;
      if (ky < P::DC_OUT) {
        float2 acc = make_float2(0.f, 0.f);
        for (int r = part; r < N; r += P::DC_SPLIT) {
          const float c = W1[r * K].x;
          const float2 w = tw[(ky * r) & (N - 1)];
          acc.x = fmaf(c, w.x, acc.x);
          acc.y = fmaf(c, w.y, acc.y);
        }
        X[ky * P::DC_SPLIT + part] = acc;
      }
      __syncthreads();
      if (tid < P::DC_OUT) {
        float2 f = make_float2(0.f, 0.f);
        for (int q = 0; q < P::DC_SPLIT; ++q) f = zadd(f, X[tid * P::DC_SPLIT + q]);
        if (tid & 1) f = make_float2(-f.x, -f.y);
        W1[(tid + O / 2) * K] = f;
        if (tid > 0) W1[(O / 2 - tid) * K] = make_float2(f.x, -f.y);
      }
      __syncthreads();
    }
    // ---- C rows: packed real Hartley rows, O/2-point FFT, full centred rows into G (over W1)
    float2* G = W1;
    {
      auto hval = [&](int i, int j) __attribute__((always_inline)) -> float {
        const int ky = i - O / 2, kx = j - O / 2;
        float2 f;
        if (kx >= 0) {
          f = W1[(ky + O / 2) * K + kx];
        } else {
          const float2 g = W1[(O / 2 - ky) * K - kx];
          f = make_float2(g.x, -g.y);
        }
        return f.x - f.y;
      };
      constexpr int MC = P::MC, TC = P::TC;
      const bool crow = tid < O * TC;
      const int ic = crow ? tid / TC : O, tc = tid % TC;  // idle threads share a spare row
      float2* zc = X + ic * P::RSC;
      if (crow)
        for (int m = tc; m < MC; m += TC)
          zc[fpad(m)] = make_float2(hval(ic, (2 * m + O / 2) & (O - 1)), hval(ic, (2 * m + 1 + O / 2) & (O - 1)));
      __syncthreads();
      float2 v[16];
      load_pass<MC, 16>(v, tc, zc);
      __syncthreads();
      fft_passes<MC, 2 * N / O, false>(v, tc, zc, tw);
      __syncthreads();
      if (crow)
        for (int k = tc; k <= MC; k += TC) {
          const float2 a = zc[fpad(k & (MC - 1))];
          const float2 bc = zc[fpad((MC - k) & (MC - 1))];
          const float2 e = make_float2(0.5f * (a.x + bc.x), 0.5f * (a.y - bc.y));
          const float2 od = make_float2(0.5f * (a.y + bc.y), -0.5f * (a.x - bc.x));
          const float2 xk = zadd(e, zmul(od, tw[k * (N / O)]));
          G[ic * P::GS + ((k + O / 2) & (O - 1))] = xk;
          if (k > 0 && k < MC) G[ic * P::GS + (O / 2 - k)] = make_float2(xk.x, -xk.y);
        }
      __syncthreads();
    }
    // ---- C columns: O-point centred FFT down each column of G, (Re - Im) * scale out
    {
      constexpr int RL = last_radix<O>(), NL = last_ns<O>();
      const bool ccol = tid < O * P::TD;
      const int cd = tid % O, td = ccol ? tid / O : 0;
      float2 v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = G[((stockham_in<O, 16>(td, 0, q) + O / 2) & (O - 1)) * P::GS + cd];
      fft_passes<O, N / O, true>(v, td, X + (ccol ? cd : O) * P::RSD, tw);
      // every G read happened before the barrier inside fft_passes, so y may overwrite G
      float* yl = reinterpret_cast<float*>(W1);
      if (ccol) {
        float* o = out + img * O * (long long)O;
#pragma unroll
        for (int u = 0; u < 16 / RL; ++u)
#pragma unroll
          for (int q = 0; q < RL; ++q) {
            const int p = (stockham_out<O, RL, NL>(td, u, q) + O / 2) & (O - 1);
            const float y = (v[u * RL + q].x - v[u * RL + q].y) * scale;
            if constexpr (FOURIER)
              yl[p * (O + 1) + cd] = y;
            else
              o[p * O + cd] = y;
          }
      }
      __syncthreads();
    }
    if constexpr (FOURIER) {
      // ---- D rows: y rows packed at their ifftshifted columns, O/2-point FFT, bins 0..O/2
      constexpr int MC = P::MC, TC = P::TC, KO = O / 2 + 1;
      const float* yl = reinterpret_cast<const float*>(W1);
      float2* Gd = W1 + (O * (O + 1) + 1) / 2;  // (O, KO) complex after y
      {
        const bool crow = tid < O * TC;
        const int ic = crow ? tid / TC : O, tc = tid % TC;
        float2* zc = X + ic * P::RSC;
        if (crow)
          for (int m = tc; m < MC; m += TC)
            zc[fpad(m)] = make_float2(yl[ic * (O + 1) + ((2 * m + O / 2) & (O - 1))],
                                      yl[ic * (O + 1) + ((2 * m + 1 + O / 2) & (O - 1))]);
        __syncthreads();
        float2 v[16];
        load_pass<MC, 16>(v, tc, zc);
        __syncthreads();
        fft_passes<MC, 2 * N / O, false>(v, tc, zc, tw);
        __syncthreads();
        if (crow)
          for (int k = tc; k <= MC; k += TC) {
            const float2 a = zc[fpad(k & (MC - 1))];
            const float2 bc = zc[fpad((MC - k) & (MC - 1))];
            const float2 e = make_float2(0.5f * (a.x + bc.x), 0.5f * (a.y - bc.y));
            const float2 od = make_float2(0.5f * (a.y + bc.y), -0.5f * (a.x - bc.x));
            Gd[ic * KO + k] = zadd(e, zmul(od, tw[k * (N / O)]));
          }
        __syncthreads();
      }
      // ---- D columns: O-point FFT down each of the KO columns, logical row n at (n + O/2) mod O
      {
        constexpr int RL = last_radix<O>(), NL = last_ns<O>();
        const bool ccol = tid < KO * P::TD;
        const int cd = ccol ? tid % KO : 0, td = ccol ? tid / KO : 0;
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = Gd[((stockham_in<O, 16>(td, 0, q) + O / 2) & (O - 1)) * KO + cd];
        fft_passes<O, N / O, true>(v, td, X + (ccol ? cd : O) * P::RSD, tw);
        if (ccol) {
          float2* o = reinterpret_cast<float2*>(out) + img * O * (long long)KO;
#pragma unroll
          for (int u = 0; u < 16 / RL; ++u)
#pragma unroll
            for (int q = 0; q < RL; ++q) o[stockham_out<O, RL, NL>(td, u, q) * KO + cd] = v[u * RL + q];
        }
        __syncthreads();
      }
    }
  }
}

bool ds_fft_ok(int h, int w, int oh, int ow) {
  return h == w && oh == ow && (h == 128 || h == 256) && (oh == 32 || oh == 64) && 2 * oh <= h;
}

template <int N, int O, bool FOURIER>
int launch_ds_fft_t(const float* x, int n, float scale, float* out, hipStream_t stream) {
  using P = DsPlan<N, O>;
  static_assert(!FOURIER || (O * (O + 1) + 1) / 2 + O * (O / 2 + 1) <= P::W1, "stage D fits the window");
  const hipError_t e = hipFuncSetAttribute((const void*)ds_fft_kernel<N, O, FOURIER>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)P::LDS);
  if (e != hipSuccess) {
    set_error(std::string("ds_fft_kernel: LDS request too large: ") + hipGetErrorString(e));
    return CRYOMI_ERR_UNSUPPORTED;
  }
  const int cus = device_cu_count();
  const unsigned blocks = (unsigned)(n < cus ? n : cus);
  hipLaunchKernelGGL((ds_fft_kernel<N, O, FOURIER>), dim3(blocks), dim3(DS_NT), P::LDS, stream, x, n, scale, out);
  return check_launch("ds_fft_kernel");
}

template <bool FOURIER>
int launch_ds_fft_f(const float* x, int n, int h, int oh, float scale, float* out, hipStream_t stream) {
  if (h == 128) return oh == 32 ? launch_ds_fft_t<128, 32, FOURIER>(x, n, scale, out, stream)
                                : launch_ds_fft_t<128, 64, FOURIER>(x, n, scale, out, stream);
  return oh == 32 ? launch_ds_fft_t<256, 32, FOURIER>(x, n, scale, out, stream)
                  : launch_ds_fft_t<256, 64, FOURIER>(x, n, scale, out, stream);
}

int launch_ds_fft(const float* x, int n, int h, int oh, unsigned flags, float* out, hipStream_t stream) {
  const float scale = (flags & CRYOMI_DS_PRESERVE_MEAN) ? 1.0f / ((float)h * (float)h) : 1.0f / ((float)oh * (float)oh);
  return (flags & CRYOMI_DS_OUTPUT_FOURIER) ? launch_ds_fft_f<true>(x, n, h, oh, scale, out, stream)
                                            : launch_ds_fft_f<false>(x, n, h, oh, scale, out, stream);
}

bool sizes_ok(int h, int w, int oh, int ow) {
  return h >= 1 && w >= 1 && h <= MAX_DIM && w <= MAX_DIM && oh >= 1 && ow >= 1 && oh <= h && ow <= w;
}

// cos and sin of 2 pi (a b mod k) / k, the product reduced exactly in integers
__device__ __forceinline__ void cs(long long a, long long b, int k, double& c, double& s) {
  long long r = (a * b) % k;
  if (r < 0) r += k;
  sincospi(2.0 * (double)r / (double)k, &s, &c);
}

__global__ __launch_bounds__(NT) void ds_operators_kernel(int h, int w, int oh, int ow, int fourier, double scale,
                                                          OpLayout L, float* __restrict__ ops) {
  const int kw = ow / 2 + 1;
  const long long nv = (long long)ow * L.lde, nu = (long long)oh * h, nq = (long long)ow * ow, np = (long long)oh * oh;
  const long long nr = fourier ? (long long)kw * ow : 0, ne = fourier ? (long long)oh * oh : 0;
  const long long total = nv + nu + nq + np + nr + ne;
  for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < total; e += (long long)gridDim.x * NT) {
    double c, s;
    long long i = e;
    if (i < nv) {  // Ce / So: forward along columns on the folded index t = |m~| (zero past ne)
      const int row = (int)(i / L.lde), col = (int)(i % L.lde);
      cs(row - ow / 2, col, w, c, s);
      ops[L.ce + i] = col < L.ne ? (float)c : 0.f;
      ops[L.so + i] = col < L.ne ? (float)s : 0.f;
      continue;
    }
    i -= nv;
    if (i < nu) {  // U: forward along rows, cas and cas(-)
      const int row = (int)(i / h), col = (int)(i % h);
      cs(row - oh / 2, col - h / 2, h, c, s);
      ops[L.u + i] = (float)(c + s);
      ops[L.u + nu + i] = (float)(c - s);
      continue;
    }
    i -= nu;
    if (i < nq) {  // Q: inverse along coarse columns
      const int row = (int)(i / ow), col = (int)(i % ow);
      cs(row - ow / 2, col - ow / 2, ow, c, s);
      ops[L.q + i] = (float)(c + s);
      ops[L.q + nq + i] = (float)(c - s);
      continue;
    }
    i -= nq;
    if (i < np) {  // P: inverse along coarse rows, with the normalisation
      const int row = (int)(i / oh), col = (int)(i % oh);
      cs(row - oh / 2, col - oh / 2, oh, c, s);
      ops[L.p + i] = (float)(scale * c);
      ops[L.p + np + i] = (float)(scale * s);
      continue;
    }
    i -= np;
    if (i < nr) {  // R: rfft along columns of ifftshift(y)
      const int row = (int)(i / ow), col = (int)(i % ow);
      cs(row, col - ow / 2, ow, c, s);
      ops[L.r + i] = (float)c;
      ops[L.r + nr + i] = (float)s;
      continue;
    }
    i -= nr;
    {  // E: fft along rows, real rows then imaginary rows, two segments
      const int row = (int)(i / oh), col = (int)(i % oh);
      cs(row, col - oh / 2, oh, c, s);
      const size_t seg = 2 * (size_t)oh * oh;
      ops[L.e + (size_t)row * oh + col] = (float)c;
      ops[L.e + (size_t)(row + oh) * oh + col] = (float)(-s);
      ops[L.e + seg + (size_t)row * oh + col] = (float)(-s);
      ops[L.e + seg + (size_t)(row + oh) * oh + col] = (float)(-c);
    }
  }
}

struct WsLayout {
  size_t zt, hh, t, y, gt, per_image;  // floats
  int chunk;
};

WsLayout ws_layout(int n, int h, int w, int oh, int ow, unsigned flags) {
  (void)w;
  WsLayout W{};
  const size_t kw = ow / 2 + 1;
  W.zt = 0;
  W.hh = up64(2 * (size_t)ow * h);
  W.t = W.hh + up64((size_t)oh * ow);
  W.y = W.t + up64(2 * (size_t)ow * oh);
  W.gt = W.y;
  W.per_image = W.y;
  if (flags & CRYOMI_DS_OUTPUT_FOURIER) {
    W.gt = W.y + up64((size_t)oh * ow);
    W.per_image = W.gt + up64(2 * kw * oh);
  }
  size_t chunk = CHUNK_FLOATS / W.per_image;
  if (chunk < 1) chunk = 1;
  if (chunk > (size_t)n) chunk = (size_t)(n > 0 ? n : 1);
  W.chunk = (int)chunk;
  return W;
}

}  // namespace
}  // namespace cryomi

using namespace cryomi;

extern "C" {

int cryomi_bgemm_nt(const cryomi_bgemm_args* a, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(a != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_bgemm_nt: NULL args");
  CRYOMI_REQUIRE(a->m >= 0 && a->n >= 0 && a->k >= 1 && a->nseg >= 1 && a->batch >= 0 && a->lda >= a->k &&
                     a->ldb >= a->k && a->ldc >= a->n,
                 CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_bgemm_nt: need m, n, batch >= 0, k, nseg >= 1, lda, ldb >= k and ldc >= n");
  if (a->m == 0 || a->n == 0 || a->batch == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(a->a && a->b && a->c, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_bgemm_nt: NULL pointer");
  GemmDesc g{a->a, a->a_batch, a->a_seg, a->lda, a->b, a->b_batch, a->b_seg, a->ldb, a->c, a->c_batch, a->ldc,
             a->m, a->n, a->k, a->nseg, a->batch, 0, 0};
  return launch_gemm(g, stream);
}

size_t cryomi_fourier_crop_operator_floats(int h, int w, int oh, int ow, unsigned flags) {
  if (!sizes_ok(h, w, oh, ow) || (flags & ~(CRYOMI_DS_PRESERVE_MEAN | CRYOMI_DS_OUTPUT_FOURIER))) return 0;
  return op_layout(h, w, oh, ow, flags).total;
}

int cryomi_fourier_crop_operators(int h, int w, int oh, int ow, unsigned flags, float* ops, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(sizes_ok(h, w, oh, ow), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_fourier_crop_operators: need 1 <= oh <= h <= 4096 and 1 <= ow <= w <= 4096");
  CRYOMI_REQUIRE(!(flags & ~(CRYOMI_DS_PRESERVE_MEAN | CRYOMI_DS_OUTPUT_FOURIER)), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_fourier_crop_operators: unknown flag bits");
  CRYOMI_REQUIRE(ops != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_fourier_crop_operators: NULL pointer");
  const OpLayout L = op_layout(h, w, oh, ow, flags);
  const double scale =
      (flags & CRYOMI_DS_PRESERVE_MEAN) ? 1.0 / ((double)h * (double)w) : 1.0 / ((double)oh * (double)ow);
  const size_t entries = L.total / 2 + 1;
  unsigned blocks = (unsigned)((entries + NT - 1) / NT);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(ds_operators_kernel, dim3(blocks), dim3(NT), 0, stream, h, w, oh, ow,
                     (flags & CRYOMI_DS_OUTPUT_FOURIER) ? 1 : 0, scale, L, ops);
  return check_launch("ds_operators_kernel");
}

size_t cryomi_fourier_crop_downsample_workspace_bytes(int n, int h, int w, int oh, int ow, unsigned flags) {
  if (n < 0 || !sizes_ok(h, w, oh, ow)) return 0;
  const WsLayout W = ws_layout(n, h, w, oh, ow, flags);
  return W.per_image * (size_t)W.chunk * sizeof(float);
}

int cryomi_fourier_crop_downsample(const float* images, int n, int h, int w, int oh, int ow, unsigned flags,
                                   const float* ops, void* workspace, size_t workspace_bytes, void* out,
                                   ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && sizes_ok(h, w, oh, ow), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_fourier_crop_downsample: need n >= 0, 1 <= oh <= h <= 4096 and 1 <= ow <= w <= 4096");
  CRYOMI_REQUIRE(!(flags & ~(CRYOMI_DS_PRESERVE_MEAN | CRYOMI_DS_OUTPUT_FOURIER)), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_fourier_crop_downsample: unknown flag bits");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(images && ops && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_fourier_crop_downsample: NULL pointer");
  CRYOMI_REQUIRE(workspace && workspace_bytes >= cryomi_fourier_crop_downsample_workspace_bytes(n, h, w, oh, ow, flags),
                 CRYOMI_ERR_WORKSPACE,
                 "cryomi_fourier_crop_downsample: workspace smaller than cryomi_fourier_crop_downsample_workspace_bytes()");
  // the FFT kernel reads 16-byte vectors: other alignments take the contractions
  if ((reinterpret_cast<uintptr_t>(images) & 15) == 0 && ds_fft_ok(h, w, oh, ow))
    return launch_ds_fft(images, n, h, oh, flags, static_cast<float*>(out), stream);
  const OpLayout L = op_layout(h, w, oh, ow, flags);
  const WsLayout W = ws_layout(n, h, w, oh, ow, flags);
  const bool fourier = flags & CRYOMI_DS_OUTPUT_FOURIER;
  const int kw = ow / 2 + 1;
  float* ws = static_cast<float*>(workspace);
  for (int i0 = 0; i0 < n; i0 += W.chunk) {
    const int nb = (n - i0 < W.chunk) ? n - i0 : W.chunk;
    const long long per = (long long)W.per_image;
    float* zt = ws + W.zt;
    float* hh = ws + W.hh;
    float* tt = ws + W.t;
    float* yy = fourier ? ws + W.y : static_cast<float*>(out) + (size_t)i0 * oh * ow;
    // chunk buffers are image-major with stride `per` floats per image
    const float* x0 = images + (size_t)i0 * h * w;
    int st = launch_fold(ops + L.ce, ops + L.so, L.lde, x0, nb, h, w, ow, zt, per, stream);
    if (st) return st;
    const GemmDesc steps[3] = {
        {ops + L.u, 0, (long long)oh * h, h, zt, per, (long long)ow * h, h, hh, per, ow, oh, ow, h, 2, nb, 0, 0},
        {ops + L.q, 0, 0, ow, hh, per, 0, ow, tt, per, oh, 2 * ow, oh, ow, 1, nb, 0, 0},
        {ops + L.p, 0, (long long)oh * oh, oh, tt, per, (long long)ow * oh, oh, yy, fourier ? per : (long long)oh * ow,
         ow, oh, ow, oh, 2, nb, 0, 0},
    };
    for (const GemmDesc& g : steps) {
      st = launch_gemm(g, stream);
      if (st) return st;
    }
    if (fourier) {
      float* gt = ws + W.gt;
      float* fo = static_cast<float*>(out) + (size_t)i0 * oh * kw * 2;
      const GemmDesc f[2] = {
          {ops + L.r, 0, 0, ow, yy, per, 0, ow, gt, per, oh, 2 * kw, oh, ow, 1, nb, 0, 0},
          {ops + L.e, 0, 2LL * oh * oh, oh, gt, per, (long long)kw * oh, oh, fo, (long long)oh * kw * 2, kw, 2 * oh, kw,
           oh, 2, nb, oh, 0},
      };
      for (const GemmDesc& g : f) {
        int st = launch_gemm(g, stream);
        if (st) return st;
      }
    }
  }
  return CRYOMI_OK;
}

}  // extern "C"
