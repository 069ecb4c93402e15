// Pose-proposal CNN primitives on gfx950: BF16 MFMA GEMM with fused
// bias / ReLU / mask epilogues and split-k accumulation, stride-2 3x3
// im2col / col2im, bias gradients, the multi-hypothesis pose loss, Adam, and
// the input pipeline (spectrum crop + CTF phase flip, noise + normalisation).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/cryomi_cnn.h"
#include "common.hpp"
#include "sample.hpp"

namespace cryomi {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

constexpr int NT = 256;

__device__ __forceinline__ bf16 to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float((unsigned)h << 16); }

// ---------------------------------------------------------------------------
// GEMM: 64 x 64 tiles, k steps of 32 through double-buffered LDS, 4 waves of
// 32 x 32 (2 x 2 MFMA 16x16x16 bf16 tiles). LDS rows hold 32 k values (+8
// padding) so that an MFMA operand (4 consecutive k of one row) is one 8-byte
// read.
// ---------------------------------------------------------------------------
constexpr int GBM = 64, GBN = 64, GBK = 32, GSTR = GBK + 8;

struct GemmArgs {
  const bf16* a;
  long long lda_m, lda_k;
  const bf16* b;
  long long ldb_k, ldb_n;
  int m, n, k;
  float alpha;
  const float* bias;
  const bf16* mask;
  long long ld_mask;
  float* c32;
  long long ldc32;
  bf16* c16;
  long long ldc16;
  unsigned flags;
  int k_per_split;
};

// Loads one operand tile (64 rows x 32 k) into registers: 8 bf16 per thread.
// K_CONTIG: row r = tid / 4, k segment (tid % 4) * 8; else k = tid / 8, row
// segment (tid % 8) * 8 (8 consecutive rows at one k).
template <bool K_CONTIG>
__device__ __forceinline__ uint4 load_tile(const bf16* __restrict__ p, long long ld_row, long long ld_k, int row0,
                                           int rows, int k0, int kend, int tid) {
  uint4 v = make_uint4(0u, 0u, 0u, 0u);
  if (K_CONTIG) {
    const int r = tid >> 2, ks = (tid & 3) * 8;
    if (row0 + r < rows && k0 + ks < kend) v = *reinterpret_cast<const uint4*>(p + (row0 + r) * ld_row + (k0 + ks));
  } else {
    const int kk = tid >> 3, rs = (tid & 7) * 8;
    if (row0 + rs < rows && k0 + kk < kend) v = *reinterpret_cast<const uint4*>(p + (long long)(k0 + kk) * ld_k + (row0 + rs));
  }
  return v;
}

template <bool K_CONTIG>
__device__ __forceinline__ void store_tile(bf16 (*s)[GSTR], uint4 v, int tid) {
  if (K_CONTIG) {
    const int r = tid >> 2, ks = (tid & 3) * 8;
    *reinterpret_cast<uint4*>(&s[r][ks]) = v;
  } else {
    const int kk = tid >> 3, rs = (tid & 7) * 8;
    const bf16* e = reinterpret_cast<const bf16*>(&v);
#pragma unroll
    for (int q = 0; q < 8; ++q) s[rs + q][kk] = e[q];
  }
}

template <bool A_K, bool B_K>
__global__ __launch_bounds__(NT) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __align__(16) bf16 As[2][GBM][GSTR];
  __shared__ __align__(16) bf16 Bs[2][GBN][GSTR];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
  const int kb = blockIdx.z * g.k_per_split;
  const int ke = min(g.k, kb + g.k_per_split);
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int nk = (ke - kb + GBK - 1) / GBK;
  if (nk > 0) {
    uint4 ra = load_tile<A_K>(g.a, g.lda_m, g.lda_k, m0, g.m, kb, ke, tid);
    uint4 rb = load_tile<B_K>(g.b, g.ldb_n, g.ldb_k, n0, g.n, kb, ke, tid);
    store_tile<A_K>(As[0], ra, tid);
    store_tile<B_K>(Bs[0], rb, tid);
    __syncthreads();
    for (int t = 0; t < nk; ++t) {
      const int cur = t & 1;
      if (t + 1 < nk) {
        const int k0 = kb + (t + 1) * GBK;
        ra = load_tile<A_K>(g.a, g.lda_m, g.lda_k, m0, g.m, k0, ke, tid);
        rb = load_tile<B_K>(g.b, g.ldb_n, g.ldb_k, n0, g.n, k0, ke, tid);
      }
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int kk = s * 16 + 4 * (lane >> 4);
        bf16x4 av[2], bv[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) av[i] = *reinterpret_cast<const bf16x4*>(&As[cur][wm * 32 + i * 16 + (lane & 15)][kk]);
#pragma unroll
        for (int j = 0; j < 2; ++j) bv[j] = *reinterpret_cast<const bf16x4*>(&Bs[cur][wn * 32 + j * 16 + (lane & 15)][kk]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[i], bv[j], acc[i][j], 0, 0, 0);
      }
      if (t + 1 < nk) {
        store_tile<A_K>(As[cur ^ 1], ra, tid);
        store_tile<B_K>(Bs[cur ^ 1], rb, tid);
      }
      __syncthreads();
    }
  }
  // D element (4 (lane / 16) + r, lane % 16) of each 16 x 16 tile
  const bool accumulate = (g.flags & CRYOMI_GEMM_ACCUMULATE) != 0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row >= g.m || col >= g.n) continue;
        float v = g.alpha * acc[i][j][r];
        if (accumulate) {
          unsafeAtomicAdd(g.c32 + (long long)row * g.ldc32 + col, v);
          continue;
        }
        if (g.bias) v += g.bias[col];
        if (g.flags & CRYOMI_GEMM_RELU) v = fmaxf(v, 0.f);
        if (g.mask) {
          const bf16 mk = g.mask[(long long)row * g.ld_mask + col];
          if ((mk & 0x8000u) || mk == 0) v = 0.f;
        }
        if (g.c32) g.c32[(long long)row * g.ldc32 + col] = v;
        if (g.c16) g.c16[(long long)row * g.ldc16 + col] = to_bf16(v);
      }
}

// ---------------------------------------------------------------------------
// im2col / col2im for 3x3, stride 2, padding 1
// ---------------------------------------------------------------------------
// One thread per 8 consecutive entries of an im2col row (one 16-byte store), 32-bit index
// arithmetic (the host bounds the element count). With c % 8 == 0 the 8 entries are 8
// channels of one tap, read with one 16-byte load (channels are innermost in x).
__global__ __launch_bounds__(NT) void im2col_kernel(const bf16* __restrict__ x, int n, int h, int w, int c, int kp,
                                                    bf16* __restrict__ out) {
  const int ho = h / 2, wo = w / 2;
  const int kv = kp / 8;
  const int total = n * ho * wo * kv;
  for (int t = blockIdx.x * NT + threadIdx.x; t < total; t += gridDim.x * NT) {
    const int row = t / kv, kk = 8 * (t - row * kv);
    const int ox = row % wo, oy = (row / wo) % ho, b = row / (wo * ho);
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if ((c & 7) == 0) {
      if (kk < 9 * c) {
        const int tap = kk / c, ci = kk - tap * c;
        const int iy = 2 * oy - 1 + tap / 3, ix = 2 * ox - 1 + tap % 3;
        if (iy >= 0 && iy < h && ix >= 0 && ix < w)
          v = *reinterpret_cast<const uint4*>(x + ((size_t)(b * h + iy) * w + ix) * c + ci);
      }
    } else {
      unsigned short e8[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        e8[q] = 0;
        const int k = kk + q;
        if (k < 9 * c) {
          const int tap = k / c, ci = k - tap * c;
          const int iy = 2 * oy - 1 + tap / 3, ix = 2 * ox - 1 + tap % 3;
          if (iy >= 0 && iy < h && ix >= 0 && ix < w) e8[q] = x[((size_t)(b * h + iy) * w + ix) * c + ci];
        }
      }
      v = make_uint4(e8[0] | ((unsigned)e8[1] << 16), e8[2] | ((unsigned)e8[3] << 16), e8[4] | ((unsigned)e8[5] << 16),
                     e8[6] | ((unsigned)e8[7] << 16));
    }
    *reinterpret_cast<uint4*>(out + (size_t)row * kp + kk) = v;
  }
}

// dx = sum over the (at most four) taps that read each input pixel, times the ReLU mask of x
// (x > 0; no mask when x is NULL). One thread per 8 channels of one pixel when c % 8 == 0
// (16-byte mask load and store, two 16-byte loads per tap), else per element.
__global__ __launch_bounds__(NT) void col2im_kernel(const float* __restrict__ dcol, const bf16* __restrict__ x, int n,
                                                    int h, int w, int c, int kp, bf16* __restrict__ dx) {
  const int ho = h / 2, wo = w / 2;
  const int cg = (c & 7) == 0 ? c / 8 : c;  // channel groups per pixel
  const int width = (c & 7) == 0 ? 8 : 1;
  const int total = n * h * w * cg;
  for (int t = blockIdx.x * NT + threadIdx.x; t < total; t += gridDim.x * NT) {
    const int pix = t / cg, ci = (t - pix * cg) * width;
    const int ix = pix % w, iy = (pix / w) % h, b = pix / (w * h);
    const size_t e = (size_t)pix * c + ci;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ty = 0; ty < 3; ++ty) {
      const int yy = iy + 1 - ty;
      if (yy < 0 || (yy & 1) || (yy >> 1) >= ho) continue;
#pragma unroll
      for (int tx = 0; tx < 3; ++tx) {
        const int xx = ix + 1 - tx;
        if (xx < 0 || (xx & 1) || (xx >> 1) >= wo) continue;
        const float* src = dcol + (size_t)((b * ho + (yy >> 1)) * wo + (xx >> 1)) * kp + (ty * 3 + tx) * c + ci;
        if (width == 8) {
          const float4 lo = *reinterpret_cast<const float4*>(src), hi = *reinterpret_cast<const float4*>(src + 4);
          s[0] += lo.x; s[1] += lo.y; s[2] += lo.z; s[3] += lo.w;
          s[4] += hi.x; s[5] += hi.y; s[6] += hi.z; s[7] += hi.w;
        } else {
          s[0] += src[0];
        }
      }
    }
    if (width == 8) {
      unsigned short m8[8];
      if (x) {
        const uint4 xv = *reinterpret_cast<const uint4*>(x + e);
        const unsigned xw[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          m8[2 * q] = (unsigned short)(xw[q] & 0xffffu);
          m8[2 * q + 1] = (unsigned short)(xw[q] >> 16);
        }
      }
      unsigned short o8[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const bool keep = !x || (m8[q] != 0 && !(m8[q] & 0x8000u));
        o8[q] = keep ? to_bf16(s[q]) : to_bf16(0.f);
      }
      *reinterpret_cast<uint4*>(dx + e) =
          make_uint4(o8[0] | ((unsigned)o8[1] << 16), o8[2] | ((unsigned)o8[3] << 16), o8[4] | ((unsigned)o8[5] << 16),
                     o8[6] | ((unsigned)o8[7] << 16));
    } else {
      const bool keep = !x || (x[e] != 0 && !(x[e] & 0x8000u));
      dx[e] = keep ? to_bf16(s[0]) : to_bf16(0.f);
    }
  }
}

// grad[j] += sum over a chunk of `rows` rows of dy[i, j]: 4 row groups x 64 columns per
// block, the groups' partial sums combined in LDS, one atomic per column and block. The host
// sizes the chunk (at least CS_ROWS rows) so that at most CS_BLOCKS blocks add to one column:
// atomics to one address serialise (a 524288 x 32 gradient in 64-row chunks took 112 us).
constexpr int CS_ROWS = 64;
constexpr int CS_BLOCKS = 1024;
__global__ __launch_bounds__(NT) void column_sums_kernel(const bf16* __restrict__ dy, int m, int n, long long ld,
                                                         int rows, float* __restrict__ grad) {
  __shared__ float part[NT / 64][64];
  const int r0 = blockIdx.x * rows;
  const int r1 = min(m, r0 + rows);
  const int cj = threadIdx.x & 63, rg = threadIdx.x >> 6;
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j = j0 + cj;
    float sum = 0.f;
    if (j < n) {
#pragma unroll 8
      for (int r = r0 + rg; r < r1; r += NT / 64) sum += from_bf16(dy[(long long)r * ld + j]);
    }
    part[rg][cj] = sum;
    __syncthreads();
    if (rg == 0 && j < n) {
      float t = part[0][cj];
      for (int g = 1; g < NT / 64; ++g) t += part[g][cj];
      unsafeAtomicAdd(grad + j, t);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(NT) void to_bf16_kernel(const float* __restrict__ x, size_t n, bf16* __restrict__ out) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < n; e += (size_t)gridDim.x * NT) out[e] = to_bf16(x[e]);
}

__global__ __launch_bounds__(NT) void fill_zero_f32_kernel(float* __restrict__ x, size_t n) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < n; e += (size_t)gridDim.x * NT) x[e] = 0.f;
}

// ---------------------------------------------------------------------------
// pose loss: one thread per image
// ---------------------------------------------------------------------------
constexpr int NH = 8, NOUT = 96, ROT0 = 0, OFF0 = 72, LOG0 = 88;

__global__ __launch_bounds__(NT) void pose_loss_kernel(const float* __restrict__ out, const float* __restrict__ rot,
                                                       const float* __restrict__ off, int n, float eps,
                                                       float off_weight, bf16* __restrict__ dout,
                                                       float* __restrict__ dout32, float* __restrict__ terms) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const float* o = out + (size_t)i * NOUT;
  const float* R = rot + (size_t)i * 9;
  const float tx = off[2 * i], ty = off[2 * i + 1];
  float d[NH], e[NH];
  int w = 0;
  float dmin = INFINITY;
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const float t = o[ROT0 + 9 * h + q] - R[q];
      s = fmaf(t, t, s);
    }
    d[h] = s;
    const float ex = o[OFF0 + 2 * h] - tx, ey = o[OFF0 + 2 * h + 1] - ty;
    e[h] = ex * ex + ey * ey;
    if (s < dmin) {  // first minimum wins ties
      dmin = s;
      w = h;
    }
  }
  // softmax over the confidence logits, cross-entropy against the winner
  float mx = o[LOG0];
#pragma unroll
  for (int h = 1; h < NH; ++h) mx = fmaxf(mx, o[LOG0 + h]);
  float z = 0.f, p[NH];
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    p[h] = expf(o[LOG0 + h] - mx);
    z += p[h];
  }
  const float inv_n = 1.0f / (float)n;
  float loss = 0.f;
  float g[NOUT];
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const float ch = (h == w) ? 1.0f - eps : eps / (float)(NH - 1);
    loss += ch * (d[h] + off_weight * e[h]);
#pragma unroll
    for (int q = 0; q < 9; ++q) g[ROT0 + 9 * h + q] = 2.f * ch * (o[ROT0 + 9 * h + q] - R[q]) * inv_n;
    g[OFF0 + 2 * h] = 2.f * ch * off_weight * (o[OFF0 + 2 * h] - tx) * inv_n;
    g[OFF0 + 2 * h + 1] = 2.f * ch * off_weight * (o[OFF0 + 2 * h + 1] - ty) * inv_n;
    p[h] /= z;
    g[LOG0 + h] = (p[h] - (h == w ? 1.f : 0.f)) * inv_n;
  }
  float pw = 0.f;
#pragma unroll
  for (int h = 0; h < NH; ++h)
    if (h == w) pw = p[h];
  loss += -logf(fmaxf(pw, 1e-30f));
#pragma unroll
  for (int q = 0; q < NOUT; ++q) {
    dout[(size_t)i * NOUT + q] = to_bf16(g[q]);
    if (dout32) dout32[(size_t)i * NOUT + q] = g[q];
  }
  if (terms) {
    terms[3 * i] = loss;
    terms[3 * i + 1] = dmin;
    terms[3 * i + 2] = (float)w;
  }
}

// ---------------------------------------------------------------------------
// Adam
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ g, bf16* __restrict__ p16, size_t n,
                                                  float lr, float b1, float b2, float eps, float bc1, float bc2) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < n; e += (size_t)gridDim.x * NT) {
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    const float pi = p[e] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[e] = pi;
    if (p16) p16[e] = to_bf16(pi);
  }
}

// ---------------------------------------------------------------------------
// input pipeline
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < NT / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

template <typename C>
__device__ __forceinline__ float2 load_c(const C* p, size_t e);
template <>
__device__ __forceinline__ float2 load_c<double2>(const double2* p, size_t e) {
  const double2 v = p[e];
  return make_float2((float)v.x, (float)v.y);
}
template <>
__device__ __forceinline__ float2 load_c<float2>(const float2* p, size_t e) {
  return p[e];
}

// One workgroup per image.
template <typename C>
__global__ __launch_bounds__(NT) void crop_spectrum_kernel(const C* __restrict__ spec, int box, int out_box,
                                                           const cryomi_image_params* __restrict__ prm,
                                                           const float* __restrict__ snr, unsigned flags,
                                                           float2* __restrict__ out) {
  __shared__ double red[NT / 64];
  const int img = blockIdx.x;
  const int m = box / 2, mo = out_box / 2;
  const size_t nh = (size_t)box * (m + 1);
  const C* s = spec + (size_t)img * nh;
  float scale = 1.f;
  float dc = 0.f;
  if (flags & CRYOMI_CROP_NORMALIZE) {
    // Parseval over the half-plane: sum x^2 = (1/N^2) sum_k c_k |X_k|^2
    double acc = 0.0;
    for (size_t e = threadIdx.x; e < nh; e += NT) {
      const int j = (int)(e % (m + 1));
      const float2 v = load_c<C>(s, e);
      acc += (double)((j == 0 || j == m) ? 1.f : 2.f) * ((double)v.x * v.x + (double)v.y * v.y);
    }
    const double nn = (double)box * box;
    const double sumsq = block_sum(acc, red) / nn;
    const float2 x0 = load_c<C>(s, 0);
    const double mean = (double)x0.x / nn;
    const double var = sumsq / nn - mean * mean;
    scale = (float)(sqrt((double)snr[img]) / sqrt(var > 0.0 ? var : 1.0));
    dc = x0.x;  // the mean's spectrum: removed before scaling
  }
  ImageConsts k{};
  const bool flip = (flags & CRYOMI_CROP_PHASE_FLIP) != 0;
  if (flip) k = make_consts(prm[img], box, false);
  const int ho = out_box, wo = mo + 1;
  float2* o = out + (size_t)img * ho * wo;
  for (int e = threadIdx.x; e < ho * wo; e += NT) {
    const int io = e / wo, jo = e % wo;
    const int ky = io < mo ? io : io - out_box;
    const int i = ky >= 0 ? ky : ky + box;
    float2 v = load_c<C>(s, (size_t)i * (m + 1) + jo);
    if (i == 0 && jo == 0) v.x -= dc;
    v.x *= scale;
    v.y *= scale;
    if (flip) {
      // CTF on the box grid, the same FP64 phase polynomial as the forward
      const double fy = (double)ky, fx = (double)jo;
      const double kk = fx * fx + fy * fy;
      const double u = fma(k.ca, kk, fma(k.cb, fy * fy - fx * fx, fma(k.cc, fx * fy, fma(k.cd, kk * kk, -k.cphase))));
      const double ctf = sinpi(2.0 * (u - rint(u)));
      if (ctf < 0.0) v = make_float2(-v.x, -v.y);
    }
    o[e] = v;
  }
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter (c0..c3), key (k0, k1).
__device__ __forceinline__ uint4 philox(uint4 c, uint2 key) {
  const unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const unsigned hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    const unsigned hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    c = make_uint4(hi1 ^ c.y ^ key.x, lo1, hi0 ^ c.w ^ key.y, lo0);
    key.x += W0;
    key.y += W1;
  }
  return c;
}

// One workgroup per image: add noise, then normalise.
__global__ __launch_bounds__(NT) void noise_normalize_kernel(const float* __restrict__ x, int s, float noise_std,
                                                             unsigned long long seed, bf16* __restrict__ out) {
  __shared__ double red[NT / 64];
  const int img = blockIdx.x;
  const int np = s * s;
  const float* xi = x + (size_t)img * np;
  bf16* oi = out + (size_t)img * np;
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  // pass 1: moments of x + noise (the noise is regenerated identically in pass 2)
  double s1 = 0.0, s2 = 0.0;
  for (int p0 = threadIdx.x * 4; p0 < np; p0 += NT * 4) {
    const uint4 r = philox(make_uint4((unsigned)p0, (unsigned)img, 0x5eedu, 0u), key);
    const float u1 = ((r.x >> 8) + 1) * (1.0f / 16777216.0f), u2 = (r.y >> 8) * (1.0f / 16777216.0f);
    const float u3 = ((r.z >> 8) + 1) * (1.0f / 16777216.0f), u4 = (r.w >> 8) * (1.0f / 16777216.0f);
    const float ra = sqrtf(-2.f * logf(u1)), rb = sqrtf(-2.f * logf(u3));
    float sa, ca, sb, cb;
    sincospif(2.f * u2, &sa, &ca);
    sincospif(2.f * u4, &sb, &cb);
    const float nz[4] = {ra * ca, ra * sa, rb * cb, rb * sb};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (p0 + q < np) {
        const double v = (double)xi[p0 + q] + (double)noise_std * nz[q];
        s1 += v;
        s2 += v * v;
      }
    }
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  const double mean = s1 / np;
  const double var = s2 / np - mean * mean;
  const float inv = (float)(1.0 / sqrt(var > 1e-30 ? var : 1.0));
  const float mu = (float)mean;
  for (int p0 = threadIdx.x * 4; p0 < np; p0 += NT * 4) {
    const uint4 r = philox(make_uint4((unsigned)p0, (unsigned)img, 0x5eedu, 0u), key);
    const float u1 = ((r.x >> 8) + 1) * (1.0f / 16777216.0f), u2 = (r.y >> 8) * (1.0f / 16777216.0f);
    const float u3 = ((r.z >> 8) + 1) * (1.0f / 16777216.0f), u4 = (r.w >> 8) * (1.0f / 16777216.0f);
    const float ra = sqrtf(-2.f * logf(u1)), rb = sqrtf(-2.f * logf(u3));
    float sa, ca, sb, cb;
    sincospif(2.f * u2, &sa, &ca);
    sincospif(2.f * u4, &sb, &cb);
    const float nz[4] = {ra * ca, ra * sa, rb * cb, rb * sb};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (p0 + q < np) oi[p0 + q] = to_bf16((xi[p0 + q] + noise_std * nz[q] - mu) * inv);
  }
}

static unsigned grid_for(long long total) {
  long long b = (total + NT - 1) / NT;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace
}  // namespace cryomi

using namespace cryomi;

extern "C" {

int cryomi_gemm_bf16(const cryomi_gemm_bf16_args* a, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(a != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_gemm_bf16: NULL args");
  CRYOMI_REQUIRE(a->m >= 0 && a->n >= 0 && a->k >= 0 && a->k % 8 == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gemm_bf16: need m, n >= 0 and k a non-negative multiple of 8");
  const bool a_k = a->lda_k == 1, b_k = a->ldb_k == 1;
  CRYOMI_REQUIRE((a_k || a->lda_m == 1) && (b_k || a->ldb_n == 1), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gemm_bf16: each operand needs a unit stride");
  const long long sa = a_k ? a->lda_m : a->lda_k, sb = b_k ? a->ldb_n : a->ldb_k;
  CRYOMI_REQUIRE(sa % 8 == 0 && sb % 8 == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gemm_bf16: non-unit strides must be multiples of 8");
  CRYOMI_REQUIRE((a_k || a->m % 8 == 0) && (b_k || a->n % 8 == 0), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gemm_bf16: m (n) must be a multiple of 8 when A (B) is contiguous along it");
  const bool accumulate = (a->flags & CRYOMI_GEMM_ACCUMULATE) != 0;
  CRYOMI_REQUIRE((a->flags & ~(CRYOMI_GEMM_RELU | CRYOMI_GEMM_ACCUMULATE)) == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gemm_bf16: unknown flag bits");
  CRYOMI_REQUIRE(a->split_k >= 1 && (a->split_k == 1 || accumulate), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gemm_bf16: split_k > 1 needs CRYOMI_GEMM_ACCUMULATE");
  CRYOMI_REQUIRE(!accumulate || (a->c32 && !a->c16 && !a->bias && !a->mask && !(a->flags & CRYOMI_GEMM_RELU)),
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_gemm_bf16: accumulate mode takes c32 only, no epilogue");
  if (a->m == 0 || a->n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(a->a && a->b && (a->c32 || a->c16), CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_gemm_bf16: NULL pointer");
  CRYOMI_REQUIRE(reinterpret_cast<uintptr_t>(a->a) % 16 == 0 && reinterpret_cast<uintptr_t>(a->b) % 16 == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_gemm_bf16: a and b must be 16-byte aligned");
  GemmArgs g{static_cast<const bf16*>(a->a), a->lda_m, a->lda_k, static_cast<const bf16*>(a->b), a->ldb_k,
             a->ldb_n, a->m, a->n, a->k, a->alpha, a->bias, static_cast<const bf16*>(a->mask), a->ld_mask,
             a->c32, a->ldc32, static_cast<bf16*>(a->c16), a->ldc16, a->flags, 0};
  const int splits = a->split_k;
  int kps = (a->k + splits - 1) / splits;
  kps = ((kps + GBK - 1) / GBK) * GBK;
  g.k_per_split = kps > 0 ? kps : GBK;
  const int zs = a->k > 0 ? (a->k + g.k_per_split - 1) / g.k_per_split : 1;
  const dim3 grid((unsigned)((a->n + GBN - 1) / GBN), (unsigned)((a->m + GBM - 1) / GBM), (unsigned)zs);
  if (a_k && b_k) hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, dim3(NT), 0, stream, g);
  else if (a_k) hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, dim3(NT), 0, stream, g);
  else if (b_k) hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, dim3(NT), 0, stream, g);
  else hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, dim3(NT), 0, stream, g);
  return check_launch("gemm_bf16_kernel");
}

int cryomi_im2col_s2(const void* x, int n, int h, int w, int c, int kp, void* out, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && h > 0 && w > 0 && c > 0 && h % 2 == 0 && w % 2 == 0 && kp >= 9 * c && kp % 8 == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_im2col_s2: need even h, w and kp >= 9 c, kp % 8 == 0");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(x && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_im2col_s2: NULL pointer");
  const long long total = (long long)n * (h / 2) * (w / 2) * kp;
  CRYOMI_REQUIRE(total < (1ll << 31) && (long long)n * h * w * c < (1ll << 31), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_im2col_s2: more than 2^31 elements");
  hipLaunchKernelGGL(im2col_kernel, dim3(grid_for(total / 8)), dim3(NT), 0, stream, static_cast<const bf16*>(x), n,
                     h, w, c, kp, static_cast<bf16*>(out));
  return check_launch("im2col_kernel");
}

int cryomi_col2im_s2(const float* dcol, const void* x, int n, int h, int w, int c, int kp, void* dx,
                     ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && h > 0 && w > 0 && c > 0 && h % 2 == 0 && w % 2 == 0 && kp >= 9 * c,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_col2im_s2: need even h, w and kp >= 9 c");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(dcol && dx, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_col2im_s2: NULL pointer");
  const long long total = (long long)n * h * w * c;
  CRYOMI_REQUIRE(total < (1ll << 31) && (long long)n * (h / 2) * (w / 2) * kp < (1ll << 31),
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_col2im_s2: more than 2^31 elements");
  CRYOMI_REQUIRE(c % 8 != 0 || kp % 4 == 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_col2im_s2: need kp % 4 == 0");
  hipLaunchKernelGGL(col2im_kernel, dim3(grid_for(c % 8 == 0 ? total / 8 : total)), dim3(NT), 0, stream, dcol,
                     static_cast<const bf16*>(x), n, h, w, c, kp, static_cast<bf16*>(dx));
  return check_launch("col2im_kernel");
}

int cryomi_column_sums_bf16(const void* dy, int m, int n, long long ld, float* grad, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(m >= 0 && n >= 0 && ld >= n, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_column_sums_bf16: bad shape");
  if (m == 0 || n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(dy && grad, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_column_sums_bf16: NULL pointer");
  int rows = (m + CS_BLOCKS - 1) / CS_BLOCKS;
  rows = rows < CS_ROWS ? CS_ROWS : ((rows + 3) / 4) * 4;
  hipLaunchKernelGGL(column_sums_kernel, dim3((unsigned)((m + rows - 1) / rows)), dim3(NT), 0, stream,
                     static_cast<const bf16*>(dy), m, n, ld, rows, grad);
  return check_launch("column_sums_kernel");
}

int cryomi_to_bf16(const float* x, size_t n, void* out, ihipStream_t* stream) {
  clear_error();
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(x && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_to_bf16: NULL pointer");
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for((long long)n)), dim3(NT), 0, stream, x, n, static_cast<bf16*>(out));
  return check_launch("to_bf16_kernel");
}

int cryomi_fill_zero(float* x, size_t n, ihipStream_t* stream) {
  clear_error();
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(x != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_fill_zero: NULL pointer");
  hipLaunchKernelGGL(fill_zero_f32_kernel, dim3(grid_for((long long)n)), dim3(NT), 0, stream, x, n);
  return check_launch("fill_zero_f32_kernel");
}

int cryomi_pose_loss(const float* out, const float* rot, const float* off, int n, float eps, float off_weight,
                     void* dout, float* dout32, float* terms, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && eps >= 0.f && eps < 1.f && off_weight >= 0.f, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_pose_loss: need n >= 0, 0 <= eps < 1, off_weight >= 0");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(out && rot && off && dout, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_pose_loss: NULL pointer");
  hipLaunchKernelGGL(pose_loss_kernel, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, stream, out, rot, off, n,
                     eps, off_weight, static_cast<bf16*>(dout), dout32, terms);
  return check_launch("pose_loss_kernel");
}

int cryomi_adam(float* p, float* m, float* v, const float* g, void* p16, size_t n, float lr, float beta1, float beta2,
                float eps, float bias_c1, float bias_c2, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(bias_c1 > 0.f && bias_c2 > 0.f && eps > 0.f, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_adam: bias corrections and eps must be positive");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(p && m && v && g, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_adam: NULL pointer");
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for((long long)n)), dim3(NT), 0, stream, p, m, v, g,
                     static_cast<bf16*>(p16), n, lr, beta1, beta2, eps, bias_c1, bias_c2);
  return check_launch("adam_kernel");
}

int cryomi_crop_spectrum(const void* spec, int complex128, int n, int box, int out_box,
                         const cryomi_image_params* params, const float* snr, unsigned flags, void* out,
                         ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && box >= 2 && box % 2 == 0 && out_box >= 2 && out_box % 2 == 0 && out_box <= box,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_crop_spectrum: need even 2 <= out_box <= box");
  CRYOMI_REQUIRE((flags & ~(CRYOMI_CROP_PHASE_FLIP | CRYOMI_CROP_NORMALIZE)) == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_crop_spectrum: unknown flag bits");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(spec && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_crop_spectrum: NULL pointer");
  CRYOMI_REQUIRE(!(flags & CRYOMI_CROP_PHASE_FLIP) || params, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_crop_spectrum: phase flip needs params");
  CRYOMI_REQUIRE(!(flags & CRYOMI_CROP_NORMALIZE) || snr, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_crop_spectrum: normalisation needs snr");
  if (complex128)
    hipLaunchKernelGGL(crop_spectrum_kernel<double2>, dim3((unsigned)n), dim3(NT), 0, stream,
                       static_cast<const double2*>(spec), box, out_box, params, snr, flags, static_cast<float2*>(out));
  else
    hipLaunchKernelGGL(crop_spectrum_kernel<float2>, dim3((unsigned)n), dim3(NT), 0, stream,
                       static_cast<const float2*>(spec), box, out_box, params, snr, flags, static_cast<float2*>(out));
  return check_launch("crop_spectrum_kernel");
}

int cryomi_noise_normalize(const float* x, int n, int s, float noise_std, unsigned long long seed, void* out,
                           ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && s > 0 && noise_std >= 0.f, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_noise_normalize: need n >= 0, s > 0, noise_std >= 0");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(x && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_noise_normalize: NULL pointer");
  hipLaunchKernelGGL(noise_normalize_kernel, dim3((unsigned)n), dim3(NT), 0, stream, x, s, noise_std, seed,
                     static_cast<bf16*>(out));
  return check_launch("noise_normalize_kernel");
}

}  // extern "C"
