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
#include <cstdint>
#include <string>
#include <type_traits>

#include "../../include/cryomi_fft.h"
#include "common.hpp"
#include "stockham.hpp"

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

// ---------------------------------------------------------------------------
// FP64 rfft2 by Stockham FFTs (power-of-two sizes, 16 <= h <= 512, 32 <= w <= 512)
//
// rfftn(ifftshift(y)) in two in-place passes over the output, both HBM-bound:
//   rows    : each real row of w samples, read at the ifftshifted position (row
//             (r + h/2) mod h, column (c + w/2) mod w), is packed as M = w/2
//             complex values z[m] = x[2m] + i x[2m+1], transformed by an M-point
//             FFT, and unpacked into the w/2 + 1 rfft bins:
//               X[k] = E + e^{-2 pi i k / w} O,  E = (Z[k] + Z*[M-k]) / 2,
//               O = (Z[k] - Z*[M-k]) / 2i   (Z indices mod M);
//   columns : an h-point complex FFT down each of the w/2 + 1 columns.
// Each L-point FFT is a sequence of Stockham autosort passes (radix 16, the
// last one radix L / 16^p): butterfly j of a pass with radix R after sub-
// transforms of size Ns takes x[j + q L/R], q < R, multiplies by
// e^{-2 pi i q k / (Ns R)} with k = j mod Ns, runs an R-point DFT in registers
// and stores output q at (j / Ns) Ns R + k + q Ns; the last pass leaves natural
// order. L/16 threads share one transform, each holding 16 values.
// ---------------------------------------------------------------------------
constexpr int FNT = 128;  // threads per workgroup of the FFT kernels

constexpr int FFT_WG_PER_CU = 4;  // LDS-limited residency of the FFT kernels (about 40 KB each)

// Row pass: real rows of W samples (F32 or F64) -> W/2 + 1 bins. Persistent: each workgroup takes
// RPW consecutive rows per item, and the next item's rows are loaded into registers while the
// current item is transformed.
// FULL = false: rfft rows (w/2 + 1 bins), rows read at (r + h/2) mod h of their image.
// FULL = true: the x-pass of a centred 3D transform: rows read in place (the other axes' shifts
// belong to their passes) and written as all W bins at their fftshifted index (k + W/2) mod W,
// the negative frequencies as conjugates of the positive ones (the row is real).
template <int W, bool F32, bool FULL>
__global__ __launch_bounds__(FNT, 2) void rfft_rows_f64_kernel(const void* __restrict__ in, long long rows, int h,
                                                            double2* __restrict__ out) {
  constexpr int M = W / 2, T = M / 16, RPW = FNT / T, RS = M + M / 16;
  constexpr int CH = F32 ? W / 4 : W / 2;  // 16-byte chunks per row
  constexpr int NCH = RPW * CH / FNT;      // chunks per thread per item
  static_assert(NCH * FNT == RPW * CH, "whole chunks per thread");
  typedef typename std::conditional<F32, float4, double2>::type Chunk;
  __shared__ double2 z[RPW * RS];
  __shared__ double2 tw[W];
  fill_twiddles<FNT>(tw, W);
  const long long items = (rows + RPW - 1) / RPW;
  // chunk i of this thread: row rr = e / CH, chunk ch = e % CH, e = threadIdx.x + i FNT
  // the prefetch buffer is two arrays of at most 8 chunks: a 16-entry one stays in scratch memory
  constexpr int NH = NCH / 2;
  auto load_item = [&](long long it, Chunk (&b0)[NH], Chunk (&b1)[NH]) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < NCH; ++i) {
      const int e = threadIdx.x + i * FNT, rr = e / CH, ch = e % CH;
      long long g = it * RPW + rr;
      if (g >= rows) g = rows - 1;  // a clamped duplicate; never stored
      long long src_row = g;
      if constexpr (!FULL) {
        const long long img = g / h;
        const int r = (int)(g - img * h);
        src_row = img * h + (r + h / 2) % h;
      }
      const Chunk x = reinterpret_cast<const Chunk*>(static_cast<const char*>(in) + src_row * W * (F32 ? 4 : 8))[ch];
      if (i < NH)
        b0[i] = x;
      else
        b1[i - NH] = x;
    }
  };
  long long it = blockIdx.x;
  Chunk nb0[NH], nb1[NH];
  if (it < items) load_item(it, nb0, nb1);
  const int rr = threadIdx.x / T, t = threadIdx.x % T;
  double2* zr = z + rr * RS;
  for (; it < items; it += gridDim.x) {
    // stage this item's rows, rotated by the ifftshift, as packed complex values
#pragma unroll
    for (int i = 0; i < NCH; ++i) {
      const int e = threadIdx.x + i * FNT, r2 = e / CH, ch = e % CH;
      const Chunk x = i < NH ? nb0[i] : nb1[i - NH];
      if constexpr (F32) {
        const int m = ((ch * 4 - W / 2 + W) % W) / 2;
        z[r2 * RS + fpad(m)] = make_double2(x.x, x.y);
        z[r2 * RS + fpad(m + 1)] = make_double2(x.z, x.w);
      } else {
        z[r2 * RS + fpad(((ch * 2 - W / 2 + W) % W) / 2)] = x;
      }
    }
    if (it + gridDim.x < items) load_item(it + gridDim.x, nb0, nb1);
    __syncthreads();
    double2 v[16];
    load_pass<M, 16>(v, t, zr);
    __syncthreads();
    fft_passes<M, 2, false>(v, t, zr, tw);
    __syncthreads();
    const long long g = it * RPW + rr;
    if (g < rows) {
      double2* o = out + g * (FULL ? W : M + 1);
      for (int k = t; k <= M; k += T) {
        const double2 a = zr[fpad(k & (M - 1))];
        const double2 bc = zr[fpad((M - k) & (M - 1))];
        const double2 b = make_double2(bc.x, -bc.y);
        const double2 e = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
        const double2 od = make_double2(0.5 * (a.y - b.y), -0.5 * (a.x - b.x));  // (a - b) / 2i
        const double2 x = zadd(e, zmul(od, tw[k]));
        if constexpr (FULL) {
          o[(k + M) & (W - 1)] = x;
          if (k > 0 && k < M) o[M - k] = make_double2(x.x, -x.y);
        } else {
          o[k] = x;
        }
      }
    }
    __syncthreads();  // the next item's staging overwrites z
  }
}

// LDS stride (double2) of one column transform: lanes 0-15 of a wave (ds_*_b128 serves 16 lanes
// per cycle over 64 banks) hold min(CPW, 16) columns and 16 / that many threads of each, whose
// pass offsets advance by one double2 per thread (mod 16), so the column stride must spread the
// columns over distinct 16-byte bank groups: odd for 16 columns, 16 / CPW (mod 16) below that.
constexpr int col_stride(int h, int cpw) {
  int rs = h + h / 16;
  if (cpw >= 16) return rs % 2 ? rs : rs + 1;
  while (rs % 16 != 16 / cpw) ++rs;
  return rs;
}

// Column pass, in place: an H-point FFT down each of the wc columns of every image; CPW columns
// per item, column index fastest across threads, so each row segment read or written is CPW
// complex128 wide (at least 16 = 256 bytes: narrower segments measured slower, as the rows of
// w/2 + 1 bins are not line-aligned). Items run over the flattened (image, column) index, so an
// item may end one image's columns and start the next image's: with w/2 + 1 columns per image,
// per-image items would leave all but one column of every fifth item idle. Persistent, with the
// next item's column values loaded into registers during the current transform. Dynamic LDS: CPW
// transforms, then the twiddles.
template <int H>
struct ColPlan {
  static constexpr int T = H / 16;
  static constexpr int NT = H == 256 ? 512 : (H > FNT ? H : FNT);  // threads per workgroup
  static constexpr int CPW = NT / T;
  static constexpr int RS = col_stride(H, CPW);
  static constexpr size_t LDS = ((size_t)CPW * RS + H) * sizeof(double2);
  static constexpr int WG_PER_CU = H >= 256 ? 1 : FFT_WG_PER_CU;
};

// CENTRED: the axis is transformed as fftshift(fft(ifftshift(.))): logical index n sits at row
// (n + H/2) mod H on load and on store. C64 != nullptr: a complex64 copy is stored beside.
template <int H, bool CENTRED>
__global__ __launch_bounds__(ColPlan<H>::NT, 2) void fft_cols_f64_kernel(double2* __restrict__ data, int wc,
                                                                     long long columns, long long items,
                                                                     float2* __restrict__ c64) {
  using P = ColPlan<H>;
  constexpr int CPW = P::CPW, RS = P::RS;
  constexpr int RL = last_radix<H>(), NL = last_ns<H>();
  extern __shared__ double2 col_lds[];
  double2* z = col_lds;
  double2* tw = col_lds + CPW * RS;
  fill_twiddles<P::NT>(tw, H);
  const int c = threadIdx.x % CPW, t = threadIdx.x / CPW;
  auto column = [&](long long it, bool& active) __attribute__((always_inline)) -> long long {
    long long g = it * CPW + c;
    active = g < columns;
    if (!active) g = columns - 1;  // a clamped duplicate load; never stored
    const long long img = g / wc;
    return img * H * (long long)wc + (g - img * wc);
  };
  auto row = [&](int n) __attribute__((always_inline)) -> long long {
    return (long long)(CENTRED ? (n + H / 2) & (H - 1) : n) * wc;
  };
  auto load = [&](const double2* base, double2 (&v)[16]) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = base[row(stockham_in<H, 16>(t, 0, q))];
  };
  long long it = blockIdx.x;
  double2 nv[16];
  bool nact = false;
  long long noff = 0;
  if (it < items) {
    noff = column(it, nact);
    load(data + noff, nv);
  }
  __syncthreads();  // twiddle table
  for (; it < items; it += gridDim.x) {
    double2 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = nv[q];
    const long long off = noff;
    const bool active = nact;
    if (it + gridDim.x < items) {
      noff = column(it + gridDim.x, nact);
      load(data + noff, nv);
    }
    fft_passes<H, 1, true>(v, t, z + c * RS, tw);
    if (active) {
#pragma unroll
      for (int u = 0; u < 16 / RL; ++u)
#pragma unroll
        for (int q = 0; q < RL; ++q) {
          const long long at = off + row(stockham_out<H, RL, NL>(t, u, q));
          data[at] = v[u * RL + q];
          if (c64) c64[at] = make_float2((float)v[u * RL + q].x, (float)v[u * RL + q].y);
        }
    }
    __syncthreads();  // the next item's first LDS store follows this item's last LDS loads
  }
}

bool rfft2_fft_ok(int h, int w) {
  const bool ph = h >= 16 && h <= 512 && (h & (h - 1)) == 0;
  const bool pw = w >= 32 && w <= 512 && (w & (w - 1)) == 0;
  return ph && pw;
}

template <int W, bool FULL>
int launch_rows(const void* in, bool f32, long long rows, int h, double2* out, hipStream_t stream) {
  constexpr int RPW = FNT / (W / 32);
  const long long items = (rows + RPW - 1) / RPW;
  const long long cap = (long long)device_cu_count() * FFT_WG_PER_CU;
  const unsigned blocks = (unsigned)(items < cap ? items : cap);
  if (f32)
    hipLaunchKernelGGL((rfft_rows_f64_kernel<W, true, FULL>), dim3(blocks), dim3(FNT), 0, stream, in, rows, h, out);
  else
    hipLaunchKernelGGL((rfft_rows_f64_kernel<W, false, FULL>), dim3(blocks), dim3(FNT), 0, stream, in, rows, h, out);
  return check_launch("rfft_rows_f64_kernel");
}

template <int W, bool FULL>
int launch_rows_w(int w, const void* in, bool f32, long long rows, int h, double2* out, hipStream_t stream) {
  if constexpr (W < 512) {
    if (w != W) return launch_rows_w<2 * W, FULL>(w, in, f32, rows, h, out, stream);
  }
  return launch_rows<W, FULL>(in, f32, rows, h, out, stream);
}

template <int H, bool CENTRED>
int launch_cols(double2* data, int n, int wc, float2* c64, hipStream_t stream) {
  using P = ColPlan<H>;
  const long long columns = (long long)n * wc;
  const long long items = (columns + P::CPW - 1) / P::CPW;
  const long long cap = (long long)device_cu_count() * P::WG_PER_CU;
  const unsigned blocks = (unsigned)(items < cap ? items : cap);
  if (P::LDS > (size_t)65536) {
    const hipError_t e = hipFuncSetAttribute((const void*)fft_cols_f64_kernel<H, CENTRED>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)P::LDS);
    if (e != hipSuccess) {
      set_error(std::string("fft_cols_f64_kernel: LDS request too large: ") + hipGetErrorString(e));
      return CRYOMI_ERR_UNSUPPORTED;
    }
  }
  hipLaunchKernelGGL((fft_cols_f64_kernel<H, CENTRED>), dim3(blocks), dim3(P::NT), P::LDS, stream, data, wc,
                     columns, items, c64);
  return check_launch("fft_cols_f64_kernel");
}

template <int H, bool CENTRED>
int launch_cols_h(int h, double2* data, int n, int wc, float2* c64, hipStream_t stream) {
  if constexpr (H < 512) {
    if (h != H) return launch_cols_h<2 * H, CENTRED>(h, data, n, wc, c64, stream);
  }
  return launch_cols<H, CENTRED>(data, n, wc, c64, stream);
}

int rfft2_fft(const void* in, bool f32, int n, int h, int w, double2* out, hipStream_t stream) {
  const int st = launch_rows_w<32, false>(w, in, f32, (long long)n * h, h, out, stream);
  if (st) return st;
  return launch_cols_h<16, false>(h, out, n, w / 2 + 1, nullptr, stream);
}

// Unpadded power-of-two volumes: x-pass (real rows -> centred complex rows), then centred column
// FFTs along y (per z plane) and z (one plane stride), the last also storing complex64.
bool volume_fft_ok(int d, int h, int w, int pd, int ph, int pw) {
  return d == pd && h == ph && w == pw && rfft2_fft_ok(h, w) && rfft2_fft_ok(d, w);
}

int volume_fft(const void* in, bool f32, int d, int h, int w, double2* c128, float2* c64, hipStream_t stream) {
  int st = launch_rows_w<32, true>(w, in, f32, (long long)d * h, d * h, c128, stream);
  if (!st) st = launch_cols_h<16, true>(h, c128, d, w, nullptr, stream);
  if (!st) st = launch_cols_h<16, true>(d, c128, 1, h * w, c64, stream);
  return st;
}

}  // namespace
}  // namespace cryomi

using namespace cryomi;

extern "C" {

size_t cryomi_fourier_voxel_grid_workspace_bytes(int d, int h, int w, int pd, int ph, int pw) {
  if (!vol_ok(d, h, w, pd, ph, pw)) return 0;
  if (volume_fft_ok(d, h, w, pd, ph, pw)) return (size_t)d * h * w * sizeof(double2);  // complex128 if not output
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
  if (volume_fft_ok(d, h, w, pd, ph, pw)) {
    CRYOMI_REQUIRE(((reinterpret_cast<uintptr_t>(real_grid) | reinterpret_cast<uintptr_t>(out_c128) |
                     reinterpret_cast<uintptr_t>(out_c64) | reinterpret_cast<uintptr_t>(workspace)) & 15) == 0,
                   CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_fourier_voxel_grid: buffers must be 16-byte aligned");
    double2* c128 = out_c128 ? static_cast<double2*>(out_c128) : static_cast<double2*>(workspace);
    return volume_fft(real_grid, input_type == CRYOMI_REAL_F32, d, h, w, c128, static_cast<float2*>(out_c64),
                      stream);
  }
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
  if (rfft2_fft_ok(h, w)) return 256;  // the FFT path works in place in the output
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
  const bool f32 = input_type == CRYOMI_REAL_F32;
  if (rfft2_fft_ok(h, w)) {
    // the FFT kernels move 16-byte vectors (and the workspace query sized no contraction buffers)
    CRYOMI_REQUIRE(((reinterpret_cast<uintptr_t>(images) | reinterpret_cast<uintptr_t>(out_c128)) & 15) == 0,
                   CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_rfft2_f64: images and out_c128 must be 16-byte aligned");
    return rfft2_fft(images, f32, n, h, w, static_cast<double2*>(out_c128), stream);
  }
  const ImgLayout L = img_layout(n, h, w);
  const int kw = w / 2 + 1;
  double* ws = static_cast<double*>(workspace);
  int st = build_operator(w, w, kw, false, ws + L.ow, stream);
  if (!st) st = build_operator(h, h, h, false, ws + L.oh, stream);
  if (st) return st;
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
