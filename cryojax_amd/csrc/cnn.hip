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
  const __bf16 h = (__bf16)f;  // v_cvt_pk_bf16_f32, round to nearest even
  return __builtin_bit_cast(bf16, h);
}
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float((unsigned)h << 16); }

// ---------------------------------------------------------------------------
// GEMM on v_mfma_f32_16x16x32_bf16: BM x BN tiles (128 or 64 each), k steps of
// 32 through double-buffered LDS fed from GPD register stages (the loads of
// the next GPD - 1 steps stay in flight while one is multiplied), 4 waves of
// (BM/2) x (BN/2) (16 x 16 MFMA tiles, one MFMA per tile and k step). LDS rows hold 32 k values (+8
// padding), so an MFMA operand (8 consecutive k of one row) is one 16-byte
// read. Operands come from strided matrices or, for the 3x3 / stride-2 /
// pad-1 convolutions, straight from the NHWC activations (no im2col buffer):
//   LD_K / LD_R   : dense, contiguous along k / along rows;
//   LD_IM2COL_A   : A(m = output pixel, k = patch column (tap, ci)) of x;
//   LD_IM2COL_B   : B(k = output pixel, n = patch column) of x (weight grad);
//   LD_DYG        : (data grad) A(m = input pixel of one parity class, k = j co + o)
//                   = dy at the output pixel that the class's j-th tap feeds. Input
//                   pixel (iy, ix) gets taps ty = 1 for even iy, ty in {0, 2} for
//                   odd iy (tx likewise), so the four classes (iy, ix mod 2) carry
//                   1, 2, 2 and 4 taps: blockIdx.z is the class, whose k range is
//                   its taps x co, and the rows map back to the class's pixels;
//   LD_WT         : B(k = j co + o, n = ci) = w[o kp + tap_j c + n] (data grad).
// Every load is 8 consecutive bf16 (16 bytes): along k for LD_K / IM2COL_A /
// DYG, along rows for LD_R / IM2COL_B / WT.
// ---------------------------------------------------------------------------
constexpr int GBK = 32, GSTR = GBK + 8, GPD = 4;
enum : int { LD_K = 0, LD_R = 1, LD_IM2COL_A = 2, LD_IM2COL_B = 3, LD_DYG = 4, LD_WT = 5 };
__host__ __device__ constexpr bool k_contig(int mode) { return mode == LD_K || mode == LD_IM2COL_A || mode == LD_DYG; }

typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));

// Geometry of a 3x3 / stride-2 / pad-1 layer: input (n, h, w, c) NHWC, output (n, h/2, w/2, co);
// h, w, c and co are powers of two (lh = log2 h, ...), so pixel and patch-column indices split with
// shifts and masks.
struct ConvGeom {
  int n, h, w, c, co, kp;
  int lh, lw, lc, lco;
  int py, px;  // data grad: the parity class of the input pixels (set per workgroup)
};

// Data grad: tap (ty, tx) of the class's j-th tap (j = jy * ntx + jx).
__device__ __forceinline__ void class_tap(const ConvGeom& cv, int j, int& ty, int& tx) {
  const int jy = cv.px ? j >> 1 : j, jx = cv.px ? j & 1 : 0;
  ty = cv.py ? 2 * jy : 1;
  tx = cv.px ? 2 * jx : 1;
}
// Data grad: class row r -> input pixel (b, iy, ix) with iy = 2 y' + py, ix = 2 x' + px.
__device__ __forceinline__ void class_pixel(const ConvGeom& cv, int r, int& b, int& iy, int& ix) {
  const int lwo = cv.lw - 1, lho = cv.lh - 1;
  ix = 2 * (r & ((1 << lwo) - 1)) + cv.px;
  iy = 2 * ((r >> lwo) & ((1 << lho) - 1)) + cv.py;
  b = r >> (lwo + lho);
}

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
  ConvGeom cv;
  float* part;            // if set: plain store of layer z's partial product at part + z part_slice + row n + col
  long long part_slice;
  unsigned a_bytes, b_bytes;  // operand extents (buffer-load bounds), < 2^31
  int vec_out;                // bf16-only output with n, ldc16, ld_mask multiples of 8 (16-byte rows)
};

// Operand loads are raw buffer loads: a segment outside the matrix or the image gets the byte
// offset SEG_NONE, past the end of every operand (the host keeps operands below 2^31 bytes), and
// the hardware returns zeros for it. No load sits behind a branch, so the loads of the later
// register stages stay in flight until their data is needed.
constexpr unsigned SEG_NONE = 0x80000000u;
constexpr int BUF_DWORD3 = 0x00020000;  // raw buffer, 32-bit data format
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t operand_rsrc(const bf16* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<bf16*>(p), (short)0, (int)bytes, BUF_DWORD3);
}

// Byte offset of the 8 elements of segment (r, k): (r, k .. k+7) for a k-contiguous mode, (r .. r+7, k)
// otherwise; SEG_NONE outside [0, rows) x [0, kend) and, for the convolution modes, outside the image.
template <int MODE>
__device__ __forceinline__ unsigned seg_offset(int ld_row, int ld_k, const ConvGeom& cv, int r, int k, int rows,
                                               int kend) {
  const bool in = r < rows && k < kend;
  int e = 0;
  bool ok = in;
  if (MODE == LD_K) {
    e = r * ld_row + k;
  } else if (MODE == LD_R) {
    e = k * ld_k + r;
  } else if (MODE == LD_IM2COL_A || MODE == LD_IM2COL_B) {
    const int pix = MODE == LD_IM2COL_A ? r : k, col = MODE == LD_IM2COL_A ? k : r;
    const int lwo = cv.lw - 1, lho = cv.lh - 1;
    const int ox = pix & ((1 << lwo) - 1), oy = (pix >> lwo) & ((1 << lho) - 1), b = pix >> (lwo + lho);
    const int tap = col >> cv.lc, ci = col & (cv.c - 1);
    const int ty = (tap * 11) >> 5, tx = tap - 3 * ty;  // tap / 3, tap % 3 for tap < 9
    const int iy = 2 * oy - 1 + ty, ix = 2 * ox - 1 + tx;
    ok = in && tap < 9 && (unsigned)iy < (unsigned)cv.h && (unsigned)ix < (unsigned)cv.w;
    e = ((((b << cv.lh) + iy) << cv.lw) | ix) * cv.c + ci;
  } else if (MODE == LD_DYG) {
    const int ho = cv.h >> 1, wo = cv.w >> 1;
    int b, iy, ix, ty, tx;
    class_pixel(cv, r, b, iy, ix);
    class_tap(cv, k >> cv.lco, ty, tx);
    const int o = k & (cv.co - 1);
    const int oy = (iy + 1 - ty) >> 1, ox = (ix + 1 - tx) >> 1;  // iy + 1 - ty is even and >= 0
    ok = in && oy < ho && ox < wo;
    e = (((b * ho + oy) * wo + ox) << cv.lco) + o;
  } else {  // LD_WT
    int ty, tx;
    class_tap(cv, k >> cv.lco, ty, tx);
    const int o = k & (cv.co - 1);
    e = o * cv.kp + ((ty * 3 + tx) << cv.lc) + r;
  }
  return ok ? 2u * (unsigned)e : SEG_NONE;
}

template <int MODE, int ROWS>
__device__ __forceinline__ void load_tile(__amdgpu_buffer_rsrc_t rsrc, int ld_row, int ld_k, const ConvGeom& cv,
                                          int row0, int rows, int k0, int kend, int tid, uint4 (&v)[ROWS / 64]) {
#pragma unroll
  for (int q = 0; q < ROWS / 64; ++q) {
    const int s = tid + 256 * q;
    const unsigned off = k_contig(MODE)
                             ? seg_offset<MODE>(ld_row, ld_k, cv, row0 + (s >> 2), k0 + (s & 3) * 8, rows, kend)
                             : seg_offset<MODE>(ld_row, ld_k, cv, row0 + (s % (ROWS / 8)) * 8, k0 + s / (ROWS / 8),
                                                rows, kend);
    const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)off, 0, 0);
    v[q] = make_uint4(w.x, w.y, w.z, w.w);
  }
}

// LDS tiles. A k-contiguous operand is stored row-major, rows of GBK k (+8 padding): an MFMA
// operand (8 consecutive k of one row) is one 16-byte read. The other modes load 8 consecutive
// rows of one k; they are stored k-major, GBK rows of ROWS (+16 padding) elements, each segment
// one 16-byte write, and the MFMA operand is assembled by two ds_read_b64_tr_b16 transposing
// reads: in each group of 16 lanes, lane j receives element j % 4 of the 8-byte chunks addressed
// by lanes j / 4 + 4 q, q = 0..3. So lane l of the group addresses (k0 + l / 4, m0 + 4 (l % 4)) and
// receives the 4 consecutive k at k0 of row m0 + l. The padding puts the 4 rows a group reads
// 8 banks apart (conflict-free).
__host__ __device__ constexpr int kmajor_stride(int rows) { return rows + 16; }
__host__ __device__ constexpr int lds_tile_elems(int rows) { return rows * GSTR; }  // >= GBK * kmajor_stride(rows)
static_assert(GBK * kmajor_stride(64) <= lds_tile_elems(64) && GBK * kmajor_stride(128) <= lds_tile_elems(128),
              "k-major tiles must fit the row-major allocation");

typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) bf16x4v lds_bf16x4;

template <int MODE, int ROWS>
__device__ __forceinline__ bf16x8v read_frag(const bf16* sm, int row0, int lane) {
  if (k_contig(MODE)) return *reinterpret_cast<const bf16x8v*>(sm + (row0 + (lane & 15)) * GSTR + 8 * (lane >> 4));
  constexpr int S = kmajor_stride(ROWS);
  const int j = lane & 15;
  const bf16* p = sm + (8 * (lane >> 4) + (j >> 2)) * S + row0 + 4 * (j & 3);
  const bf16x4v lo = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)p);
  const bf16x4v hi = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(p + 4 * S));
  return bf16x8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// LD_IM2COL_B: column 9 c of B is all ones (within the k range), so that column of the product is
// the bias gradient; its segment (9 c .. 9 c + 7, k) gets element 0 = 1 when it is written to LDS.
template <int MODE, int ROWS>
__device__ __forceinline__ void store_tile(bf16* sm, const uint4 (&v)[ROWS / 64], int tid, const ConvGeom& cv,
                                           int row0, int rows, int k0, int kend) {
#pragma unroll
  for (int q = 0; q < ROWS / 64; ++q) {
    const int s = tid + 256 * q;
    if (k_contig(MODE)) {
      *reinterpret_cast<uint4*>(sm + (s >> 2) * GSTR + (s & 3) * 8) = v[q];
    } else {
      const int rs = (s % (ROWS / 8)) * 8, kk = s / (ROWS / 8);
      uint4 u = v[q];
      if (MODE == LD_IM2COL_B && row0 + rs == 9 * cv.c && row0 + rs < rows && k0 + kk < kend) u.x |= 0x3F80u;
      *reinterpret_cast<uint4*>(sm + kk * kmajor_stride(ROWS) + rs) = u;
    }
  }
}

template <int AM, int BMODE, int BM, int BN>
__global__ __launch_bounds__(NT) void gemm_bf16_kernel(GemmArgs g) {
  // one LDS block: the double-buffered A and B tiles, reused by the staged epilogue
  __shared__ __align__(16) bf16 lds[2 * (lds_tile_elems(BM) + lds_tile_elems(BN))];
  static_assert(BM * (BN + 8) <= 2 * (lds_tile_elems(BM) + lds_tile_elems(BN)), "epilogue tile must fit");
  bf16(*As)[lds_tile_elems(BM)] = reinterpret_cast<bf16(*)[lds_tile_elems(BM)]>(lds);
  bf16(*Bs)[lds_tile_elems(BN)] = reinterpret_cast<bf16(*)[lds_tile_elems(BN)]>(lds + 2 * lds_tile_elems(BM));
  constexpr int TI = BM / 32, TJ = BN / 32;  // 16 x 16 tiles per wave along m, n
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  ConvGeom cv = g.cv;
  int kb, ke;
  if (AM == LD_DYG) {  // blockIdx.z is the parity class; its k range is its taps x co
    cv.py = blockIdx.z >> 1;
    cv.px = blockIdx.z & 1;
    kb = 0;
    ke = (cv.py ? 2 : 1) * (cv.px ? 2 : 1) * cv.co;
  } else {
    kb = blockIdx.z * g.k_per_split;
    ke = min(g.k, kb + g.k_per_split);
  }
  floatx4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int nk = (ke - kb + GBK - 1) / GBK;
  if (nk > 0) {
    // GPD register stages: k step t sits in stage t % GPD from its load until it is written to
    // LDS at the end of step t - 1, so the loads of GPD - 1 later steps are in flight meanwhile.
    const __amdgpu_buffer_rsrc_t sa = operand_rsrc(g.a, g.a_bytes), sb = operand_rsrc(g.b, g.b_bytes);
    const int lam = (int)g.lda_m, lak = (int)g.lda_k, lbn = (int)g.ldb_n, lbk = (int)g.ldb_k;
    uint4 ra[GPD][BM / 64], rb[GPD][BN / 64];
#pragma unroll
    for (int q = 0; q < GPD; ++q) {  // steps past the k range load zeros (no memory access)
      load_tile<AM, BM>(sa, lam, lak, cv, m0, g.m, kb + q * GBK, ke, tid, ra[q]);
      load_tile<BMODE, BN>(sb, lbn, lbk, cv, n0, g.n, kb + q * GBK, ke, tid, rb[q]);
    }
    store_tile<AM, BM>(As[0], ra[0], tid, cv, m0, g.m, kb, ke);
    store_tile<BMODE, BN>(Bs[0], rb[0], tid, cv, n0, g.n, kb, ke);
    __syncthreads();
    for (int t0 = 0; t0 < nk; t0 += GPD) {
#pragma unroll
      for (int q = 0; q < GPD; ++q) {
        const int t = t0 + q;
        if (t < nk) {
          const int cur = q & 1;  // GPD is even, so t & 1 == q & 1
          bf16x8v av[TI], bv[TJ];
#pragma unroll
          for (int i = 0; i < TI; ++i)
            av[i] = read_frag<AM, BM>(As[cur], wm * (BM / 2) + i * 16, lane);
#pragma unroll
          for (int j = 0; j < TJ; ++j)
            bv[j] = read_frag<BMODE, BN>(Bs[cur], wn * (BN / 2) + j * 16, lane);
          {  // stage q is free (step t is in LDS): step t + GPD, zeros past the k range
            const int k0 = kb + (t + GPD) * GBK;
            load_tile<AM, BM>(sa, lam, lak, cv, m0, g.m, k0, ke, tid, ra[q]);
            load_tile<BMODE, BN>(sb, lbn, lbk, cv, n0, g.n, k0, ke, tid, rb[q]);
          }
#pragma unroll
          for (int i = 0; i < TI; ++i)
#pragma unroll
            for (int j = 0; j < TJ; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
          if (t + 1 < nk) {  // step t + 1 into the other LDS buffer (last read in step t - 1)
            const int k1 = kb + (t + 1) * GBK;
            store_tile<AM, BM>(As[cur ^ 1], ra[(q + 1) % GPD], tid, cv, m0, g.m, k1, ke);
            store_tile<BMODE, BN>(Bs[cur ^ 1], rb[(q + 1) % GPD], tid, cv, n0, g.n, k1, ke);
          }
          __syncthreads();
        }
      }
    }
  }
  // D element (4 (lane / 16) + r, lane % 16) of each 16 x 16 tile. Offsets are 32-bit (the host
  // keeps every output matrix below 2^31 elements); one row offset per (i, r), columns by adds.
  const int rb = m0 + wm * (BM / 2) + 4 * (lane >> 4);
  const int cb = n0 + wn * (BN / 2) + (lane & 15);
  if (g.part) {
    float* base = g.part + (long long)blockIdx.z * g.part_slice;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb + i * 16 + r;
        if (row >= g.m) continue;
        float* dst = base + row * g.n;
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          if (cb + j * 16 < g.n) dst[cb + j * 16] = g.alpha * acc[i][j][r];
      }
    return;
  }
  if (g.flags & CRYOMI_GEMM_ACCUMULATE) {
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb + i * 16 + r;
        if (row >= g.m) continue;
        float* dst = g.c32 + row * (int)g.ldc32;
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          if (cb + j * 16 < g.n) unsafeAtomicAdd(dst + cb + j * 16, g.alpha * acc[i][j][r]);
      }
    return;
  }
  float bj[TJ];
#pragma unroll
  for (int j = 0; j < TJ; ++j) bj[j] = (g.bias && cb + j * 16 < g.n) ? g.bias[cb + j * 16] : 0.f;
  const bool relu = (g.flags & CRYOMI_GEMM_RELU) != 0;
  if (g.vec_out) {
    // bf16 output only, n and the output / mask strides multiples of 8: the tile goes through LDS
    // (rows of BN + 8) so that each thread then handles 8 consecutive columns of a row with one
    // 16-byte mask load and one 16-byte store. The main loop ended with a barrier, so the LDS
    // block is free.
    constexpr int CS = BN + 8;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float v = g.alpha * acc[i][j][r] + bj[j];
          if (relu) v = fmaxf(v, 0.f);
          lds[(wm * (BM / 2) + i * 16 + 4 * (lane >> 4) + r) * CS + wn * (BN / 2) + j * 16 + (lane & 15)] = to_bf16(v);
        }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BM * BN / 8 / NT; ++q) {
      const int seg = tid + NT * q;
      const int lr = seg / (BN / 8), lc = (seg % (BN / 8)) * 8;
      int row = m0 + lr;
      const int col = n0 + lc;
      if (row >= g.m || col >= g.n) continue;
      if (AM == LD_DYG) {  // class row -> input pixel
        int b, iy, ix;
        class_pixel(cv, row, b, iy, ix);
        row = ((b << cv.lh) + iy) << cv.lw | ix;
      }
      uint4 v = *reinterpret_cast<const uint4*>(lds + lr * CS + lc);
      if (g.mask) {  // zero where the mask is not positive (sign bit set, or +0)
        const uint4 mk = *reinterpret_cast<const uint4*>(g.mask + row * (int)g.ld_mask + col);
        auto keep = [](unsigned val, unsigned m) {
          const unsigned lo = (m & 0x8000u) || !(m & 0xFFFFu) ? 0u : 0xFFFFu;
          const unsigned hi = (m & 0x80000000u) || !(m >> 16) ? 0u : 0xFFFF0000u;
          return val & (lo | hi);
        };
        v = make_uint4(keep(v.x, mk.x), keep(v.y, mk.y), keep(v.z, mk.z), keep(v.w, mk.w));
      }
      *reinterpret_cast<uint4*>(g.c16 + row * (int)g.ldc16 + col) = v;
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int row = rb + i * 16 + r;
      if (row >= g.m) continue;
      if (AM == LD_DYG) {  // class row -> input pixel
        int b, iy, ix;
        class_pixel(cv, row, b, iy, ix);
        row = ((b << cv.lh) + iy) << cv.lw | ix;
      }
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        const int col = cb + j * 16;
        if (col >= g.n) continue;
        float v = g.alpha * acc[i][j][r] + bj[j];
        if (relu) v = fmaxf(v, 0.f);
        if (g.mask) {
          const bf16 mk = g.mask[row * (int)g.ld_mask + col];
          if ((mk & 0x8000u) || mk == 0) v = 0.f;
        }
        if (g.c32) g.c32[row * (int)g.ldc32 + col] = v;
        if (g.c16) g.c16[row * (int)g.ldc16 + col] = to_bf16(v);
      }
    }
}

// Tile shape for an m x n product with zs workgroup layers: the largest of 128 x 128, 128 x 64,
// 64 x 128 that still gives at least 256 workgroups (one per CU) and spans no axis of 64 or
// fewer, else 64 x 64.
template <int AM, int BMODE>
static void launch_gemm(const GemmArgs& g0, unsigned zs, hipStream_t stream) {
  GemmArgs g = g0;
  auto al16 = [](const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
  g.vec_out = !g.part && !(g.flags & CRYOMI_GEMM_ACCUMULATE) && g.c16 && !g.c32 && g.n % 8 == 0 &&
              g.ldc16 % 8 == 0 && al16(g.c16) && (!g.mask || (g.ld_mask % 8 == 0 && al16(g.mask)));
  auto count = [&](int bm, int bn) { return (long long)((g.m + bm - 1) / bm) * ((g.n + bn - 1) / bn) * zs; };
  const bool m_long = g.m > 64, n_long = g.n > 64;  // never a 128 tile along an axis of 64 or fewer
  int BMv = 64, BNv = 64;
  if (m_long && n_long && count(128, 128) >= 256) BMv = BNv = 128;
  else if (m_long && g.m >= g.n && count(128, 64) >= 256) BMv = 128;
  else if (n_long && count(64, 128) >= 256) BNv = 128;
  else if (m_long && count(128, 64) >= 256) BMv = 128;
  const dim3 grid((unsigned)((g.n + BNv - 1) / BNv), (unsigned)((g.m + BMv - 1) / BMv), zs);
  if (BMv == 128 && BNv == 128) hipLaunchKernelGGL((gemm_bf16_kernel<AM, BMODE, 128, 128>), grid, dim3(NT), 0, stream, g);
  else if (BMv == 128) hipLaunchKernelGGL((gemm_bf16_kernel<AM, BMODE, 128, 64>), grid, dim3(NT), 0, stream, g);
  else if (BNv == 128) hipLaunchKernelGGL((gemm_bf16_kernel<AM, BMODE, 64, 128>), grid, dim3(NT), 0, stream, g);
  else hipLaunchKernelGGL((gemm_bf16_kernel<AM, BMODE, 64, 64>), grid, dim3(NT), 0, stream, g);
}

// k range per split-k layer (a multiple of the k step) and the number of layers.
static unsigned split_layers(GemmArgs& g, int splits) {
  int kps = (g.k + splits - 1) / splits;
  kps = ((kps + GBK - 1) / GBK) * GBK;
  g.k_per_split = kps > 0 ? kps : GBK;
  return g.k > 0 ? (unsigned)((g.k + g.k_per_split - 1) / g.k_per_split) : 1u;
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

// c == 1 with kp == 16 (the first layer): one thread per output pixel reads its 9 taps directly
// and writes the row with two 16-byte stores (the generic kernel's divisions by a runtime c and
// its two threads per row were most of its time on this layer).
__global__ __launch_bounds__(NT) void im2col_c1_kernel(const bf16* __restrict__ x, int n, int h, int w,
                                                       bf16* __restrict__ out) {
  const int ho = h / 2, wo = w / 2;
  const int total = n * ho * wo;
  for (int row = blockIdx.x * NT + threadIdx.x; row < total; row += gridDim.x * NT) {
    const int ox = row % wo, t = row / wo, oy = t % ho, b = t / ho;
    unsigned e[9];
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const int iy = 2 * oy - 1 + dy;
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        const int ix = 2 * ox - 1 + dx;
        e[3 * dy + dx] = (iy >= 0 && iy < h && ix >= 0 && ix < w) ? x[(size_t)(b * h + iy) * w + ix] : 0u;
      }
    }
    uint4* o = reinterpret_cast<uint4*>(out + (size_t)row * 16);
    o[0] = make_uint4(e[0] | (e[1] << 16), e[2] | (e[3] << 16), e[4] | (e[5] << 16), e[6] | (e[7] << 16));
    o[1] = make_uint4(e[8], 0u, 0u, 0u);
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
constexpr int CS8_BLOCKS = 256;
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

// The same sums for n % 8 == 0 and 16-byte aligned rows: each thread reads 8 consecutive columns
// of a row with one 16-byte load, n / 8 threads per row and RG = NT / (n / 8) row groups, so no
// lane idles on narrow gradients (n = 32: the lane-per-column kernel left half the wave idle and
// loaded 2 bytes per lane). Columns are taken in windows of at most NT * 8.
__global__ __launch_bounds__(NT) void column_sums8_kernel(const bf16* __restrict__ dy, int m, int n, long long ld,
                                                          int rows, float* __restrict__ grad) {
  __shared__ float part[NT * 8];
  const int r0 = blockIdx.x * rows;
  const int r1 = min(m, r0 + rows);
  for (int j0 = 0; j0 < n; j0 += NT * 8) {
    const int w = min(n - j0, NT * 8), chunks = w / 8, rg_count = NT / chunks;
    const int ch = threadIdx.x % chunks, rg = threadIdx.x / chunks;
    if (rg < rg_count) {
      float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
      for (int r = r0 + rg; r < r1; r += rg_count) {
        const uint4 v = *reinterpret_cast<const uint4*>(dy + (long long)r * ld + j0 + 8 * ch);
        const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          acc[2 * q] += __uint_as_float(u[q] << 16);
          acc[2 * q + 1] += __uint_as_float(u[q] & 0xffff0000u);
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) part[rg * w + 8 * ch + q] = acc[q];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < w; j += NT) {
      float t = part[j];
      for (int g = 1; g < rg_count; ++g) t += part[g * w + j];
      unsafeAtomicAdd(grad + j0 + j, t);
    }
    __syncthreads();
  }
}

// Weight gradient from its split-k partial slices part (layers, m, n): c32 (m, nw) and, when
// has_bias, bias (m) from the product's last column (the all-ones column of B), nw = n - has_bias.
// Written, or added to with `accumulate`. A workgroup takes WR_ELEMS consecutive elements with
// WR_SUB threads each: sub-thread u sums layers u, u + WR_SUB, ... in order, and the WR_SUB
// partial sums are added in order of u, so the result does not depend on the launch (a single
// thread per element summing ~90 layers serially was latency-bound).
constexpr int WR_SUB = 8, WR_ELEMS = NT / WR_SUB;
__global__ __launch_bounds__(NT) void wgrad_reduce_kernel(const float* __restrict__ part, int layers, long long slice,
                                                          int m, int n, int has_bias, int accumulate,
                                                          float* __restrict__ c32, float* __restrict__ bias) {
  __shared__ float sums[WR_SUB][WR_ELEMS];
  const int total = m * n, nw = n - has_bias;
  const int k = threadIdx.x % WR_ELEMS, u = threadIdx.x / WR_ELEMS;
  for (int e0 = blockIdx.x * WR_ELEMS; e0 < total; e0 += gridDim.x * WR_ELEMS) {
    const int e = e0 + k;
    float sum = 0.f;
    if (e < total) {
#pragma unroll 4
      for (int z = u; z < layers; z += WR_SUB) sum += part[(long long)z * slice + e];
    }
    sums[u][k] = sum;
    __syncthreads();
    if (u == 0 && e < total) {
      float t = sums[0][k];
#pragma unroll
      for (int q = 1; q < WR_SUB; ++q) t += sums[q][k];
      const int row = e / n, col = e - row * n;
      float* dst = col < nw ? c32 + (long long)row * nw + col : bias + row;
      *dst = accumulate ? *dst + t : t;
    }
    __syncthreads();  // sums is rewritten by the next element block
  }
}

__global__ __launch_bounds__(NT) void to_bf16_kernel(const float* __restrict__ x, size_t n, bf16* __restrict__ out) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < n; e += (size_t)gridDim.x * NT) out[e] = to_bf16(x[e]);
}

__global__ __launch_bounds__(NT) void fill_zero_f32_kernel(float* __restrict__ x, size_t n) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < n; e += (size_t)gridDim.x * NT) x[e] = 0.f;
}

// ---------------------------------------------------------------------------
// pose loss: one thread per image, LOSS_IMAGES images per workgroup. A thread's 96-float row is
// 384 bytes from its neighbour's, so the rows and the gradients pass through LDS (row stride 97
// words: conflict-free for the per-thread accesses) and global memory is read and written
// coalesced by the whole workgroup.
// ---------------------------------------------------------------------------
constexpr int NH = 8, NOUT = 96, ROT0 = 0, OFF0 = 72, LOG0 = 88;
constexpr int LOSS_IMAGES = 64, LOSS_LD = NOUT + 1;

__global__ __launch_bounds__(NT) void pose_loss_kernel(const float* __restrict__ out, const float* __restrict__ rot,
                                                       const float* __restrict__ off, int n, float eps,
                                                       float off_weight, bf16* __restrict__ dout,
                                                       float* __restrict__ dout32, float* __restrict__ terms) {
  __shared__ float rows_lds[LOSS_IMAGES * LOSS_LD];
  const int base = blockIdx.x * LOSS_IMAGES;
  const int rows = n - base < LOSS_IMAGES ? n - base : LOSS_IMAGES;
  for (int e = threadIdx.x; e < rows * NOUT; e += NT) {
    const int r = e / NOUT;
    rows_lds[r * LOSS_LD + (e - r * NOUT)] = out[(size_t)base * NOUT + e];
  }
  __syncthreads();
  if (threadIdx.x < rows) {
    const int i = base + threadIdx.x;
    float* o = rows_lds + threadIdx.x * LOSS_LD;  // this image's outputs, then its gradient
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
    for (int q = 0; q < NOUT; ++q) o[q] = g[q];
    if (terms) {
      terms[3 * i] = loss;
      terms[3 * i + 1] = dmin;
      terms[3 * i + 2] = (float)w;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < rows * NOUT; e += NT) {
    const int r = e / NOUT;
    const float v = rows_lds[r * LOSS_LD + (e - r * NOUT)];
    dout[(size_t)base * NOUT + e] = to_bf16(v);
    if (dout32) dout32[(size_t)base * NOUT + e] = v;
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

// One workgroup per image. The input half-plane is either the box image's full rfft layout
// (box rows in FFT order, box/2 + 1 columns) or the band layout (out_box + 1 rows, row r = ky
// r - out_box/2, out_box/2 + 1 columns); spec_at(ky, kx) reads any frequency of the window,
// negative kx through conjugate symmetry.
template <typename C>
__global__ __launch_bounds__(NT) void crop_spectrum_kernel(const C* __restrict__ spec, int box, int out_box,
                                                           const cryomi_image_params* __restrict__ prm,
                                                           const float* __restrict__ snr, unsigned flags,
                                                           float2* __restrict__ out) {
  __shared__ double red[NT / 64];
  const int img = blockIdx.x;
  const int m = box / 2, mo = out_box / 2;
  const bool band = (flags & CRYOMI_CROP_BAND_INPUT) != 0;
  const int rows_in = band ? out_box + 1 : box, cols_in = band ? mo + 1 : m + 1;
  const size_t nh = (size_t)rows_in * cols_in;
  const C* s = spec + (size_t)img * nh;
  auto row_of = [&](int ky) { return band ? ky + mo : (ky >= 0 ? ky : ky + box); };
  auto spec_at = [&](int ky, int kx) {
    if (kx >= 0) return load_c<C>(s, (size_t)row_of(ky) * cols_in + kx);
    const float2 v = load_c<C>(s, (size_t)row_of(-ky) * cols_in - kx);
    return make_float2(v.x, -v.y);
  };
  float scale = 1.f;
  float dc = 0.f;
  if (flags & CRYOMI_CROP_NORMALIZE) {
    // Parseval over the half-plane: sum x^2 = (1/N^2) sum_k c_k |X_k|^2 (c_k = 1 on the
    // self-conjugate columns 0 and box/2, else 2; the band's last column is not self-conjugate)
    double acc = 0.0;
    for (size_t e = threadIdx.x; e < nh; e += NT) {
      const int j = (int)(e % cols_in);
      const float2 v = load_c<C>(s, e);
      const bool self = j == 0 || (!band && j == m);
      acc += (double)(self ? 1.f : 2.f) * ((double)v.x * v.x + (double)v.y * v.y);
    }
    const double nn = (double)box * box;
    const double sumsq = block_sum(acc, red) / nn;
    const float2 x0 = spec_at(0, 0);
    const double mean = (double)x0.x / nn;
    const double var = sumsq / nn - mean * mean;
    scale = (float)(sqrt((double)snr[img]) / sqrt(var > 0.0 ? var : 1.0));
    dc = x0.x;  // the mean's spectrum: removed before scaling
  }
  ImageConsts k{};
  const bool flip = (flags & CRYOMI_CROP_PHASE_FLIP) != 0;
  const bool hartley = (flags & CRYOMI_CROP_HARTLEY) != 0;
  if (flip) k = make_consts(prm[img], box, false);
  const int ho = out_box, wo = mo + 1;
  float2* o = out + (size_t)img * ho * wo;
  for (int e = threadIdx.x; e < ho * wo; e += NT) {
    const int io = e / wo, jo = e % wo;
    const int ky = io < mo ? io : io - out_box;
    float2 v;
    if (hartley && (ky == -mo || jo == mo)) {
      // unpaired edge of the centred window [-mo, mo)^2: q with +mo read as -mo, q' = -q folded back
      auto fold = [&](int f) { return f == mo ? -mo : f; };
      const int qy = ky, qx = fold(jo);
      const int py = fold(-qy), px = fold(-qx);
      float2 a = spec_at(qy, qx), b = spec_at(py, px);
      if (qy == 0 && qx == 0) a.x -= dc;
      if (py == 0 && px == 0) b.x -= dc;
      const float ha = a.x - a.y, hb = b.x - b.y;
      v = make_float2(0.5f * (ha + hb), 0.5f * (hb - ha));
    } else {
      v = spec_at(ky, jo);
      if (ky == 0 && jo == 0) v.x -= dc;
    }
    v.x *= scale;
    v.y *= scale;
    if (flip) {
      // CTF on the box grid, the same FP64 phase polynomial as the forward
      const double fy = (double)ky, fx = (double)jo;
      const double kk = fx * fx + fy * fy;
      const double u = fma(k.ca, kk, fma(k.cb, fy * fy - fx * fx, fma(k.cc, fx * fy, fma(k.cd, kk * kk, -k.cphase))));
      // only the sign of sin(2 pi t), t = u - rint(u) in [-1/2, 1/2], is needed: negative exactly
      // for -1/2 < t < 0
      const double t = u - rint(u);
      if (t < 0.0 && t > -0.5) v = make_float2(-v.x, -v.y);
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

// One workgroup per image: add noise, then normalise. With CACHE (images of at most NT * 16
// pixels, the network's 64 x 64 input) each thread keeps its noisy values in registers between
// the moments and the normalisation; otherwise the noise is regenerated identically.
__device__ __forceinline__ void noise4(int p0, int img, uint2 key, float (&nz)[4]) {
  const uint4 r = philox(make_uint4((unsigned)p0, (unsigned)img, 0x5eedu, 0u), key);
  const float u1 = ((r.x >> 8) + 1) * (1.0f / 16777216.0f), u2 = (r.y >> 8) * (1.0f / 16777216.0f);
  const float u3 = ((r.z >> 8) + 1) * (1.0f / 16777216.0f), u4 = (r.w >> 8) * (1.0f / 16777216.0f);
  const float ra = sqrtf(-2.f * logf(u1)), rb = sqrtf(-2.f * logf(u3));
  float sa, ca, sb, cb;
  sincospif(2.f * u2, &sa, &ca);
  sincospif(2.f * u4, &sb, &cb);
  nz[0] = ra * ca, nz[1] = ra * sa, nz[2] = rb * cb, nz[3] = rb * sb;
}

constexpr int NN_CACHE_ITERS = 4;  // pixels per thread / 4 held in registers

template <bool CACHE>
__global__ __launch_bounds__(NT) void noise_normalize_kernel(const float* __restrict__ x, int s, float noise_std,
                                                             unsigned long long seed, bf16* __restrict__ out) {
  __shared__ double red[NT / 64];
  const int img = blockIdx.x;
  const int np = s * s;
  const float* xi = x + (size_t)img * np;
  bf16* oi = out + (size_t)img * np;
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  // pass 1: moments of x + noise (FP64 sums)
  double s1 = 0.0, s2 = 0.0;
  float vc[CACHE ? NN_CACHE_ITERS : 1][4];
  if constexpr (CACHE) {
#pragma unroll
    for (int it = 0; it < NN_CACHE_ITERS; ++it) {
      const int p0 = threadIdx.x * 4 + it * NT * 4;
      if (p0 < np) {
        float nz[4];
        noise4(p0, img, key, nz);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (p0 + q < np) {
            const float xv = xi[p0 + q];
            const double v = (double)xv + (double)noise_std * nz[q];
            s1 += v;
            s2 += v * v;
            vc[it][q] = xv + noise_std * nz[q];
          }
        }
      }
    }
  } else {
    for (int p0 = threadIdx.x * 4; p0 < np; p0 += NT * 4) {
      float nz[4];
      noise4(p0, img, key, nz);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (p0 + q < np) {
          const double v = (double)xi[p0 + q] + (double)noise_std * nz[q];
          s1 += v;
          s2 += v * v;
        }
      }
    }
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  const double mean = s1 / np;
  const double var = s2 / np - mean * mean;
  const float inv = (float)(1.0 / sqrt(var > 1e-30 ? var : 1.0));
  const float mu = (float)mean;
  if constexpr (CACHE) {
#pragma unroll
    for (int it = 0; it < NN_CACHE_ITERS; ++it) {
      const int p0 = threadIdx.x * 4 + it * NT * 4;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (p0 + q < np) oi[p0 + q] = to_bf16((vc[it][q] - mu) * inv);
    }
  } else {
    for (int p0 = threadIdx.x * 4; p0 < np; p0 += NT * 4) {
      float nz[4];
      noise4(p0, img, key, nz);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (p0 + q < np) oi[p0 + q] = to_bf16((xi[p0 + q] + noise_std * nz[q] - mu) * inv);
    }
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
  auto fits = [&](const void* p, long long ld) { return !p || (ld >= a->n && (long long)a->m * ld < (1ll << 31)); };
  CRYOMI_REQUIRE(fits(a->c32, a->ldc32) && fits(a->c16, a->ldc16) && fits(a->mask, a->ld_mask),
                 CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gemm_bf16: output and mask strides must be >= n with m * stride < 2^31");
  CRYOMI_REQUIRE(reinterpret_cast<uintptr_t>(a->a) % 16 == 0 && reinterpret_cast<uintptr_t>(a->b) % 16 == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_gemm_bf16: a and b must be 16-byte aligned");
  GemmArgs g{static_cast<const bf16*>(a->a), a->lda_m, a->lda_k, static_cast<const bf16*>(a->b), a->ldb_k,
             a->ldb_n, a->m, a->n, a->k, a->alpha, a->bias, static_cast<const bf16*>(a->mask), a->ld_mask,
             a->c32, a->ldc32, static_cast<bf16*>(a->c16), a->ldc16, a->flags, 0, ConvGeom{}};
  // operand extents in elements (the last element any segment reads, + 1)
  const long long ea = a->k == 0 ? 0 : a_k ? (long long)(a->m - 1) * a->lda_m + a->k : (long long)(a->k - 1) * a->lda_k + a->m;
  const long long eb = a->k == 0 ? 0 : b_k ? (long long)(a->n - 1) * a->ldb_n + a->k : (long long)(a->k - 1) * a->ldb_k + a->n;
  CRYOMI_REQUIRE(ea < (1ll << 30) && eb < (1ll << 30), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gemm_bf16: each operand must span fewer than 2^30 elements");
  g.a_bytes = (unsigned)(2 * ea), g.b_bytes = (unsigned)(2 * eb);
  const unsigned zs = split_layers(g, a->split_k);
  if (a_k && b_k) launch_gemm<LD_K, LD_K>(g, zs, stream);
  else if (a_k) launch_gemm<LD_K, LD_R>(g, zs, stream);
  else if (b_k) launch_gemm<LD_R, LD_K>(g, zs, stream);
  else launch_gemm<LD_R, LD_R>(g, zs, stream);
  return check_launch("gemm_bf16_kernel");
}

int cryomi_conv3x3_s2_bf16(const cryomi_conv_bf16_args* a, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(a != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_conv3x3_s2_bf16: NULL args");
  auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
  auto log2i = [](int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
  };
  CRYOMI_REQUIRE(a->n >= 0 && pow2(a->h) && pow2(a->w) && a->h >= 2 && a->w >= 2 && pow2(a->c) && a->c >= 8 &&
                     pow2(a->co) && a->co >= 8,
                 CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_conv3x3_s2_bf16: h, w (>= 2) and the channel counts c, co (>= 8) must be powers of two");
  CRYOMI_REQUIRE((long long)a->n * a->h * a->w * (a->c > a->co ? a->c : a->co) < (1ll << 30),
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_conv3x3_s2_bf16: 2^30 or more elements in a tensor");
  CRYOMI_REQUIRE(a->split_k >= 1 && (a->split_k == 1 || a->op == CRYOMI_CONV_WEIGHT_GRAD),
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_conv3x3_s2_bf16: split_k > 1 only for the weight gradient");
  CRYOMI_REQUIRE((a->flags & ~(CRYOMI_GEMM_RELU | CRYOMI_GEMM_ACCUMULATE)) == 0 &&
                     (!(a->flags & CRYOMI_GEMM_RELU) || a->op == CRYOMI_CONV_FORWARD) &&
                     (!(a->flags & CRYOMI_GEMM_ACCUMULATE) || a->op == CRYOMI_CONV_WEIGHT_GRAD),
                 CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_conv3x3_s2_bf16: CRYOMI_GEMM_RELU only in the forward op, CRYOMI_GEMM_ACCUMULATE only in the "
                 "weight gradient");
  CRYOMI_REQUIRE(!a->bias_grad || a->op == CRYOMI_CONV_WEIGHT_GRAD, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_conv3x3_s2_bf16: bias_grad only in the weight gradient");
  if (a->n == 0) return CRYOMI_OK;
  const int kp = 9 * a->c, npix_out = a->n * (a->h / 2) * (a->w / 2), npix_in = a->n * a->h * a->w;
  GemmArgs g{};
  g.cv = ConvGeom{a->n, a->h, a->w, a->c, a->co, kp, log2i(a->h), log2i(a->w), log2i(a->c), log2i(a->co)};
  g.alpha = 1.f;
  g.flags = a->flags;
  auto aligned = [](const void* p) { return p && reinterpret_cast<uintptr_t>(p) % 16 == 0; };
  const unsigned x_bytes = (unsigned)(2ll * npix_in * a->c), dy_bytes = (unsigned)(2ll * npix_out * a->co),
                 w_bytes = (unsigned)(2ll * a->co * kp);
  if (a->op == CRYOMI_CONV_FORWARD) {
    // y (npix_out, co) = patches(x) (npix_out, kp) . w^T
    CRYOMI_REQUIRE(aligned(a->x) && aligned(a->weights) && (a->c16 || a->c32), CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_conv3x3_s2_bf16: forward needs 16-byte aligned x, w and an output");
    g.a = static_cast<const bf16*>(a->x);
    g.b = static_cast<const bf16*>(a->weights);
    g.ldb_k = 1;
    g.ldb_n = kp;
    g.m = npix_out, g.n = a->co, g.k = kp;
    g.a_bytes = x_bytes, g.b_bytes = w_bytes;
    g.bias = a->bias;
    g.c16 = static_cast<bf16*>(a->c16), g.ldc16 = a->co;
    g.c32 = a->c32, g.ldc32 = a->co;
    launch_gemm<LD_IM2COL_A, LD_K>(g, split_layers(g, 1), stream);
  } else if (a->op == CRYOMI_CONV_WEIGHT_GRAD) {
    // [dW | db] (co, kp + 1) = dy^T (co, npix_out) . [patches(x) | 1] (npix_out, kp + 1), each of the
    // split_k pixel ranges into its own workspace slice, then summed in a fixed order.
    CRYOMI_REQUIRE(aligned(a->x) && aligned(a->dy) && a->c32 && !a->c16 && !a->bias && !a->mask,
                   CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_conv3x3_s2_bf16: the weight gradient needs aligned x, dy and c32 only");
    const int nb = a->bias_grad ? 1 : 0;
    g.a = static_cast<const bf16*>(a->dy);
    g.lda_m = 1, g.lda_k = a->co;
    g.b = static_cast<const bf16*>(a->x);
    g.m = a->co, g.n = kp + nb, g.k = npix_out;
    g.a_bytes = dy_bytes, g.b_bytes = x_bytes;
    const unsigned zs = split_layers(g, a->split_k);
    g.part = static_cast<float*>(a->workspace);
    g.part_slice = (long long)g.m * g.n;
    const size_t need = (size_t)zs * (size_t)g.part_slice * sizeof(float);
    CRYOMI_REQUIRE(a->workspace && reinterpret_cast<uintptr_t>(a->workspace) % 16 == 0 && a->workspace_bytes >= need,
                   CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_conv3x3_s2_bf16: the weight gradient needs a 16-byte aligned workspace of "
                   "split_k * co * (9 c + (bias_grad != NULL)) floats");
    g.flags = 0;
    launch_gemm<LD_R, LD_IM2COL_B>(g, zs, stream);
    const int st = check_launch("gemm_bf16_kernel");
    if (st) return st;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(grid_for((long long)g.part_slice * WR_SUB)), dim3(NT), 0, stream,
                       g.part, (int)zs, g.part_slice, g.m, g.n, nb, (a->flags & CRYOMI_GEMM_ACCUMULATE) ? 1 : 0,
                       a->c32, a->bias_grad);
    return check_launch("wgrad_reduce_kernel");
  } else if (a->op == CRYOMI_CONV_DATA_GRAD) {
    // dx (npix_in, c) = [mask > 0] gather(dy) (npix_in, 9 co) . w^T arranged (9 co, c)
    CRYOMI_REQUIRE(aligned(a->dy) && aligned(a->weights) && a->c16 && !a->c32 && !a->bias, CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_conv3x3_s2_bf16: the data gradient needs aligned dy, w and c16 only");
    // one layer of workgroups per parity class of the input pixels (npix_in / 4 rows each)
    g.a = static_cast<const bf16*>(a->dy);
    g.b = static_cast<const bf16*>(a->weights);
    g.m = npix_in / 4, g.n = a->c, g.k = 4 * a->co;
    g.a_bytes = dy_bytes, g.b_bytes = w_bytes;
    g.mask = static_cast<const bf16*>(a->mask), g.ld_mask = a->c;
    g.c16 = static_cast<bf16*>(a->c16), g.ldc16 = a->c;
    launch_gemm<LD_DYG, LD_WT>(g, 4u, stream);
  } else {
    set_error("cryomi_conv3x3_s2_bf16: unknown op");
    return CRYOMI_ERR_INVALID_ARGUMENT;
  }
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
  if (c == 1 && kp == 16) {
    hipLaunchKernelGGL(im2col_c1_kernel, dim3(grid_for(total / 16)), dim3(NT), 0, stream, static_cast<const bf16*>(x),
                       n, h, w, static_cast<bf16*>(out));
    return check_launch("im2col_c1_kernel");
  }
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
  if (n % 8 == 0 && ld % 8 == 0 && reinterpret_cast<uintptr_t>(dy) % 16 == 0) {
    // 16-byte lanes read a block's rows fast, so fewer blocks (fewer same-address atomics) pay
    int rows8 = (m + CS8_BLOCKS - 1) / CS8_BLOCKS;
    rows8 = rows8 < CS_ROWS ? CS_ROWS : rows8;
    hipLaunchKernelGGL(column_sums8_kernel, dim3((unsigned)((m + rows8 - 1) / rows8)), dim3(NT), 0, stream,
                       static_cast<const bf16*>(dy), m, n, ld, rows8, grad);
    return check_launch("column_sums8_kernel");
  }
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
  hipLaunchKernelGGL(pose_loss_kernel, dim3((unsigned)((n + LOSS_IMAGES - 1) / LOSS_IMAGES)), dim3(NT), 0, stream,
                     out, rot, off, n,
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
  CRYOMI_REQUIRE((flags & ~(CRYOMI_CROP_PHASE_FLIP | CRYOMI_CROP_NORMALIZE | CRYOMI_CROP_HARTLEY |
                             CRYOMI_CROP_BAND_INPUT)) == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_crop_spectrum: unknown flag bits");
  CRYOMI_REQUIRE(!(flags & CRYOMI_CROP_BAND_INPUT) || out_box < box, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_crop_spectrum: the band layout needs out_box < box");
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
  if ((long long)s * s <= (long long)NT * 4 * NN_CACHE_ITERS)
    hipLaunchKernelGGL(noise_normalize_kernel<true>, dim3((unsigned)n), dim3(NT), 0, stream, x, s, noise_std, seed,
                       static_cast<bf16*>(out));
  else
    hipLaunchKernelGGL(noise_normalize_kernel<false>, dim3((unsigned)n), dim3(NT), 0, stream, x, s, noise_std, seed,
                       static_cast<bf16*>(out));
  return check_launch("noise_normalize_kernel");
}

}  // extern "C"
