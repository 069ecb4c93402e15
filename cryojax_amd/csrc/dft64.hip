// FP64 centred DFTs of real volumes and images on gfx950 as batched FP64
// MFMA contractions, one axis at a time.
//
// For an axis with input length s and output length Q (of a transform of
// length P), the operator rows are
//   C[q][m] = cos(2 pi qq mm / P),  S[q][m] = sin(2 pi qq mm / P),
// mm = m - s/2 (ifftshift, and pad_to_shape's centre-aligned zero padding
// when s < P) and qq = q - P/2 (fftshift) or qq = q (natural order). The
// first axis takes real x to [Re; Im] = [C; -S] . x^T; each further axis
// takes the [Re; Im] planes to
//   Re' = C Re + S Im,   Im' = C Im - S Re
// as two K segments: A_0 = [C; -S] on the Re rows, A_1 = [S; C] on the Im
// rows. Every contraction C_i = sum_s A_s . B_{i,s}^T leaves its output axis
// first, so the next axis to contract is the contiguous one; the last one
// stores (Re, Im) interleaved as complex128 and optionally complex64.
//
// GEMM: 64 x 64 output tile per workgroup, 4 waves as 2 x 2 of
// v_mfma_f64_16x16x4f64 sub-tiles (32 x 32 per wave), K steps of 16 through
// double-buffered LDS with the next step's loads in flight during the MFMAs.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/cryomi_fft.h"
#include "common.hpp"

namespace cryomi {
namespace {

constexpr int NT = 256;
constexpr int KT = 16;
constexpr int TILE = 64;
constexpr int MAX_VOL = 1024;
constexpr int MAX_IMG = 4096;
constexpr size_t CHUNK_DOUBLES = size_t(1) << 26;  // 512 MiB of intermediates per image chunk

typedef double doublex4 __attribute__((ext_vector_type(4)));

struct Gemm64 {
  const double* a;
  long long a_seg;
  int lda;
  const void* b;  // double, or float when the kernel's B_F32
  long long b_batch, b_seg;
  int ldb;
  double* c;    // may be NULL when only c32 is written (cplx mode)
  float* c32;   // cplx mode: optional complex64 copy
  long long c_batch;  // in scalars of one part (doubles / floats per image = 2 * cplx * n)
  int ldc;
  int m, n, k, nseg, batch;
  int cplx;  // > 0: rows [0, cplx) real parts, [cplx, 2 cplx) imaginary parts, stored interleaved
};

template <bool BF32>
__device__ __forceinline__ double load_b(const void* p, long long i) {
  if constexpr (BF32)
    return (double)static_cast<const float*>(p)[i];
  else
    return static_cast<const double*>(p)[i];
}

// 4 consecutive K entries per thread; loads are unconditional (an invalid
// entry reads entry 0 of its row, which exists since k >= 1) and masked when
// stored, so no branch separates the prefetch from the MFMAs.
template <bool BF32>
__global__ __launch_bounds__(NT) void gemm64_kernel(Gemm64 g) {
  __shared__ double As[2][KT][TILE + 4];
  __shared__ double Bs[2][KT][TILE + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tiles_m = (g.m + TILE - 1) / TILE, tiles_n = (g.n + TILE - 1) / TILE;
  const int per_img = tiles_m * tiles_n;
  const int img = blockIdx.x / per_img;
  const int rem = blockIdx.x - img * per_img;
  const int row0 = (rem / tiles_n) * TILE, col0 = (rem % tiles_n) * TILE;
  const int lr = tid >> 2, lk = (tid & 3) * 4;
  const bool a_ok = row0 + lr < g.m, b_ok = col0 + lr < g.n;
  const long long arow = (long long)(a_ok ? row0 + lr : 0) * g.lda;
  const long long brow = img * g.b_batch + (long long)(b_ok ? col0 + lr : 0) * g.ldb;
  const int ktiles = (g.k + KT - 1) / KT;
  const int nt = ktiles * g.nseg;

  doublex4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};
  double ra[4], rb[4];
  auto gload = [&](int t) {
    const int s = t / ktiles, k0 = (t - s * ktiles) * KT + lk;
    const double* ap = g.a + s * g.a_seg + arow;
    const long long bp = s * g.b_seg + brow;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool ok = k0 + q < g.k;
      ra[q] = ap[(a_ok && ok) ? k0 + q : 0];
      rb[q] = load_b<BF32>(g.b, bp + ((b_ok && ok) ? k0 + q : 0));
    }
  };
  auto sstore = [&](int t, int buf) {
    const int k0 = (t % ktiles) * KT + lk;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool ok = k0 + q < g.k;
      As[buf][lk + q][lr] = (a_ok && ok) ? ra[q] : 0.0;
      Bs[buf][lk + q][lr] = (b_ok && ok) ? rb[q] : 0.0;
    }
  };
  gload(0);
  sstore(0, 0);
  __syncthreads();
  for (int t = 0; t < nt; ++t) {
    const int cur = t & 1;
    gload(t + 1 < nt ? t + 1 : t);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ks = 0; ks < KT / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      double av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = As[cur][kk][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j] = Bs[cur][kk][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (t + 1 < nt) sstore(t + 1, cur ^ 1);
    __syncthreads();
  }
  // v_mfma_f64_16x16x4f64 leaves D element (row lane / 16 + 4 r, column lane % 16) of each 16 x 16
  // sub-tile in register r (unlike the FP32 16x16x4 form, whose rows are 4 (lane / 16) + r)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
        const int col = col0 + wn * 32 + j * 16 + (lane & 15);
        if (row < g.m && col < g.n) {
          const double v = acc[i][j][r];
          if (g.cplx > 0) {
            const int part = row >= g.cplx ? 1 : 0;
            const long long o = img * g.c_batch + ((long long)(row - part * g.cplx) * g.n + col) * 2 + part;
            if (g.c) g.c[o] = v;
            if (g.c32) g.c32[o] = (float)v;
          } else {
            g.c[img * g.c_batch + (long long)row * g.ldc + col] = v;
          }
        }
      }
}

int launch64(const Gemm64& g, bool bf32, hipStream_t stream) {
  if (g.batch == 0 || g.m == 0 || g.n == 0) return CRYOMI_OK;
  const long long blocks = (long long)g.batch * ((g.m + TILE - 1) / TILE) * ((g.n + TILE - 1) / TILE);
  CRYOMI_REQUIRE(blocks <= 0x7fffffffLL, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi FP64 DFT: too many tiles");
  if (bf32)
    hipLaunchKernelGGL(gemm64_kernel<true>, dim3((unsigned)blocks), dim3(NT), 0, stream, g);
  else
    hipLaunchKernelGGL(gemm64_kernel<false>, dim3((unsigned)blocks), dim3(NT), 0, stream, g);
  return check_launch("gemm64_kernel");
}

// Operator pair for one axis: A_0 = [C; -S] then A_1 = [S; C], each (2 q_rows x s)
__global__ __launch_bounds__(NT) void dft_operator_kernel(int p, int s, int q_rows, int centred, double* __restrict__ op) {
  const long long total = (long long)q_rows * s;
  const long long seg = 2 * total;
  for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < total; e += (long long)gridDim.x * NT) {
    const int q = (int)(e / s), m = (int)(e % s);
    const long long qq = centred ? q - p / 2 : q, mm = m - s / 2;
    long long r = (qq * mm) % p;
    if (r < 0) r += p;
    double sn, cs;
    sincospi(2.0 * (double)r / (double)p, &sn, &cs);
    op[e] = cs;
    op[total + e] = -sn;
    op[seg + e] = sn;
    op[seg + total + e] = cs;
  }
}

int build_operator(int p, int s, int q_rows, bool centred, double* op, hipStream_t stream) {
  const long long total = (long long)q_rows * s;
  unsigned blocks = (unsigned)((total + NT - 1) / NT);
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(dft_operator_kernel, dim3(blocks), dim3(NT), 0, stream, p, s, q_rows, centred ? 1 : 0, op);
  return check_launch("dft_operator_kernel");
}

size_t up32(size_t x) { return (x + 31) & ~size_t(31); }

struct VolLayout {
  size_t ow, oh, od, c1, c2, total;  // offsets in doubles
};

VolLayout vol_layout(int d, int h, int w, int pd, int ph, int pw) {
  VolLayout L{};
  size_t o = 0;
  L.ow = o;
  o += up32(4 * (size_t)pw * w);
  L.oh = o;
  o += up32(4 * (size_t)ph * h);
  L.od = o;
  o += up32(4 * (size_t)pd * d);
  L.c1 = o;
  o += up32(2 * (size_t)pw * d * h);
  L.c2 = o;
  o += up32(2 * (size_t)ph * pw * d);
  L.total = o;
  return L;
}

bool vol_ok(int d, int h, int w, int pd, int ph, int pw) {
  return d >= 1 && h >= 1 && w >= 1 && pd >= d && ph >= h && pw >= w && pd <= MAX_VOL && ph <= MAX_VOL &&
         pw <= MAX_VOL;
}

struct ImgLayout {
  size_t ow, oh, c1, per_image, total;
  int chunk;
};

ImgLayout img_layout(int n, int h, int w) {
  ImgLayout L{};
  const size_t kw = w / 2 + 1;
  L.ow = 0;
  L.oh = up32(4 * kw * w);
  L.c1 = L.oh + up32(4 * (size_t)h * h);
  L.per_image = up32(2 * kw * h);
  size_t chunk = CHUNK_DOUBLES / L.per_image;
  if (chunk < 1) chunk = 1;
  if (chunk > (size_t)n) chunk = (size_t)(n > 0 ? n : 1);
  L.chunk = (int)chunk;
  L.total = L.c1 + L.per_image * chunk;
  return L;
}

}  // namespace
}  // namespace cryomi

using namespace cryomi;

extern "C" {

size_t cryomi_fourier_voxel_grid_workspace_bytes(int d, int h, int w, int pd, int ph, int pw) {
  if (!vol_ok(d, h, w, pd, ph, pw)) return 0;
  return vol_layout(d, h, w, pd, ph, pw).total * sizeof(double);
}

int cryomi_fourier_voxel_grid(const void* real_grid, int input_type, int d, int h, int w, int pd, int ph, int pw,
                              void* out_c128, void* out_c64, void* workspace, size_t workspace_bytes,
                              ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(vol_ok(d, h, w, pd, ph, pw), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_fourier_voxel_grid: need 1 <= size <= padded size <= 1024 on every axis");
  CRYOMI_REQUIRE(input_type == CRYOMI_REAL_F32 || input_type == CRYOMI_REAL_F64, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_fourier_voxel_grid: input_type must be CRYOMI_REAL_F32 or CRYOMI_REAL_F64");
  CRYOMI_REQUIRE(real_grid && (out_c128 || out_c64), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_fourier_voxel_grid: NULL input, or no output");
  CRYOMI_REQUIRE(workspace && workspace_bytes >= cryomi_fourier_voxel_grid_workspace_bytes(d, h, w, pd, ph, pw),
                 CRYOMI_ERR_WORKSPACE,
                 "cryomi_fourier_voxel_grid: workspace smaller than cryomi_fourier_voxel_grid_workspace_bytes()");
  const VolLayout L = vol_layout(d, h, w, pd, ph, pw);
  double* ws = static_cast<double*>(workspace);
  int st = build_operator(pw, w, pw, true, ws + L.ow, stream);
  if (!st) st = build_operator(ph, h, ph, true, ws + L.oh, stream);
  if (!st) st = build_operator(pd, d, pd, true, ws + L.od, stream);
  if (st) return st;
  double* c1 = ws + L.c1;
  double* c2 = ws + L.c2;
  // x (d h rows, w) -> c1 (Re/Im, pw, d, h)
  const Gemm64 a{ws + L.ow, 0, w, real_grid, 0, 0, w, c1, nullptr, 0, d * h, 2 * pw, d * h, w, 1, 1, 0};
  // c1 (Re/Im, pw d rows, h) -> c2 (Re/Im, ph, pw, d)
  const Gemm64 b{ws + L.oh, 4LL * ph * h / 2, h, c1, 0, (long long)pw * d * h, h, c2, nullptr, 0, pw * d,
                 2 * ph, pw * d, h, 2, 1, 0};
  // c2 (Re/Im, ph pw rows, d) -> out (pd, ph, pw) complex
  const Gemm64 c{ws + L.od, 4LL * pd * d / 2, d, c2, 0, (long long)ph * pw * d, d,
                 static_cast<double*>(out_c128), static_cast<float*>(out_c64), 0, 0, 2 * pd, ph * pw, d, 2, 1, pd};
  st = launch64(a, input_type == CRYOMI_REAL_F32, stream);
  if (!st) st = launch64(b, false, stream);
  if (!st) st = launch64(c, false, stream);
  return st;
}

size_t cryomi_rfft2_f64_workspace_bytes(int n, int h, int w) {
  if (n < 0 || h < 1 || w < 1 || h > MAX_IMG || w > MAX_IMG) return 0;
  return img_layout(n, h, w).total * sizeof(double);
}

int cryomi_rfft2_f64(const void* images, int input_type, int n, int h, int w, void* out_c128, void* workspace,
                     size_t workspace_bytes, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && h >= 1 && w >= 1 && h <= MAX_IMG && w <= MAX_IMG, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_rfft2_f64: need n >= 0 and 1 <= h, w <= 4096");
  CRYOMI_REQUIRE(input_type == CRYOMI_REAL_F32 || input_type == CRYOMI_REAL_F64, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_rfft2_f64: input_type must be CRYOMI_REAL_F32 or CRYOMI_REAL_F64");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(images && out_c128, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_rfft2_f64: NULL pointer");
  CRYOMI_REQUIRE(workspace && workspace_bytes >= cryomi_rfft2_f64_workspace_bytes(n, h, w), CRYOMI_ERR_WORKSPACE,
                 "cryomi_rfft2_f64: workspace smaller than cryomi_rfft2_f64_workspace_bytes()");
  const ImgLayout L = img_layout(n, h, w);
  const int kw = w / 2 + 1;
  double* ws = static_cast<double*>(workspace);
  int st = build_operator(w, w, kw, false, ws + L.ow, stream);
  if (!st) st = build_operator(h, h, h, false, ws + L.oh, stream);
  if (st) return st;
  const bool f32 = input_type == CRYOMI_REAL_F32;
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  for (int i0 = 0; i0 < n; i0 += L.chunk) {
    const int nb = (n - i0 < L.chunk) ? n - i0 : L.chunk;
    const void* x = static_cast<const char*>(images) + (size_t)i0 * h * w * esz;
    double* c1 = ws + L.c1;
    double* out = static_cast<double*>(out_c128) + (size_t)i0 * h * kw * 2;
    // image (h rows, w) -> c1 (Re/Im, kw, h)
    const Gemm64 a{ws + L.ow, 0, w, x, (long long)h * w, 0, w, c1, nullptr, (long long)L.per_image, h, 2 * kw, h,
                   w, 1, nb, 0};
    // c1 (Re/Im, kw rows, h) -> out (h, kw) complex
    const Gemm64 b{ws + L.oh, 2LL * h * h, h, c1, (long long)L.per_image, (long long)kw * h, h, out, nullptr,
                   2LL * h * kw, 0, 2 * h, kw, h, 2, nb, h};
    st = launch64(a, f32, stream);
    if (!st) st = launch64(b, false, stream);
    if (st) return st;
  }
  return CRYOMI_OK;
}

}  // extern "C"
