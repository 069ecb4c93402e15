// Reverse mode of the fused forward (vector-Jacobian product) on gfx950.
//
// For images I = fftshift(irfft2(F)), F = shift * CTF * fixups(P), P the
// trilinear slice of the grid at R k N + N/2, and a cotangent dL/dI:
//   adj_rows_kernel   : per image row, the forward real FFT of dL/dI along x
//                       (real-pair packing through an N/2-point transform),
//                       times the C2R weights c = 1 (kx = 0, N/2), 2 else.
//   adj_cols_kernel   : per image column strip, the forward FFT along y,
//                       times (-1)^(ky+kx) / N^2 (fftshift, 1/N^2): this is
//                       g_F = dL/dRe F + i dL/dIm F.
//   adj_gather_kernel : per rfft sample, g_P = CTF conj(shift) g_F (zero at the
//                       self-conjugate bins), stored per sample, and per-image
//                       sums of dL/dR (through the interpolation weights), dL/dt
//                       and dL/d defocus (one FP64 atomic per block and value).
//   grid gradient     : the transpose of the trilinear gather, sum over images
//                       and samples of w(sample, voxel) g_P, accumulated brick by
//                       brick in LDS and added to the grid once per brick (no
//                       global float atomics). Boxes up to 256 bin the samples
//                       by brick first (bin_count / bin_scan, then the gather
//                       writes the brick lists, then bin_pull, below); larger
//                       boxes use adj_pull_kernel.
//
// Why gather per brick: the push form (each sample adding w g_P into its 8
// corners) needs 16 FP32 atomics per sample, and device-wide FP32 atomics run at
// about 23 G/s on this part whatever the address pattern (1024 x 128^2 images:
// 6.0 ms of atomics against 0.4 ms for everything else). Gathering each brick's
// contributions on chip writes every voxel once.
//
// Geometry. Sample k = (kx, ky) sits at grid position q = R (kx, ky, 0) + N/2
// (x, y, z components) and gives corner c the weight prod_a (1 - |q_a - c_a|)
// when |q_a - c_a| < 1 on every axis, with a corner index in [-N, -1] wrapping
// to c + N (JAX fill/wrap rule, as the forward), or, with CRYOMI_GATHER_FILL,
// every corner outside [0, N) dropped. Every kernel forms the forward's own
// stencil (FP64 position, floor, FP32 fractions).
//
// adj_pull_kernel (boxes above 256): a block is a brick of 8 x 8 x 4 voxels; it
// first culls, per image and wrap combination, the slices whose plane misses
// the brick; for each remaining slice, one thread per sample of the 16 x 16
// window of the plane around the brick adds w g_P for the corners inside the
// brick into an LDS accumulator (LDS float atomics). Summation orders, and so
// the last bits of the grid gradient, vary between runs on both paths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "common.hpp"
#include "dft.hpp"
#include "sample.hpp"

namespace cryomi {
namespace {

constexpr int NT = 256;
// rows per block in adj_rows_kernel: 16 (8 above N/2 = 128), LDS (N + 2 rows N/2) complex
__host__ __device__ inline int rows_per_block(int m) { return m <= 128 ? 16 : 8; }
constexpr int N_GRAD = 9;  // dR00 dR01 dR10 dR11 dR20 dR21 dtx dty ddefocus

// columns per block in adj_cols_kernel: the widest of 32 .. 1 dividing N/2 whose LDS
// (N + 2 N (w + 1) complex) stays within 48 KB; wide strips keep the strided row segments
// of the column pass long (16 columns = 128 B at box 128)
__host__ __device__ inline int strip_w(int n) {
  const int m = n / 2;
  for (int w : {32, 16, 8, 4, 2, 1})
    if (m % w == 0 && (size_t)(n + 2 * n * (w + 1)) * 8 <= 48 * 1024) return w;
  return 1;
}

__global__ __launch_bounds__(NT) void adj_rows_kernel(const float* __restrict__ ct, int n, int n_img,
                                                      float2* __restrict__ s) {
  extern __shared__ float2 smem[];
  const int m = n / 2;
  const int tid = threadIdx.x;
  const int RPB = rows_per_block(m);
  const int groups = (n + RPB - 1) / RPB;  // the last group is partial when RPB does not divide N
  const int img = blockIdx.x / groups;
  const int row0 = (blockIdx.x % groups) * RPB;
  const int rows = min(RPB, n - row0);
  float2* tw = smem;
  float2* bufa = tw + n;
  float2* bufb = bufa + RPB * m;
  fill_twiddles(tw, n, tid, NT);
  const float* src = ct + ((size_t)img * n + row0) * n;
  // z[t] = x[2t] + i x[2t+1] (rows past the image transform zeros and are not stored)
  for (int e = tid; e < RPB * m; e += NT) {
    const int r = e / m, t = e % m;
    bufa[r * m + t] = r < rows ? reinterpret_cast<const float2*>(src + (size_t)r * n)[t] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* z = stockham_inverse(bufa, bufb, m, RPB, 1, m, tw, 2, tid, NT);
  // X_inv[k] = E[k] + W^k O[k], E = (Z[k] + Z*[M-k]) / 2, O = (Z[k] - Z*[M-k]) / (2i);
  // the forward transform of the real row is conj(X_inv)
  for (int e = tid; e < rows * (m + 1); e += NT) {
    const int r = e / (m + 1), k = e % (m + 1);
    const float2 zk = z[r * m + (k % m)];
    const float2 zc = cconj(z[r * m + ((m - k) % m)]);
    const float2 ev = cscale(cadd(zk, zc), 0.5f);
    const float2 d = cscale(csub(zk, zc), 0.5f);  // = i O
    const float2 od = make_float2(d.y, -d.x);    // O = d / i
    const float2 wk = (k == m) ? make_float2(-1.f, 0.f) : tw[k];
    const float2 xi = cadd(ev, cmul(wk, od));
    const float c = (k == 0 || k == m) ? 1.f : 2.f;
    s[((size_t)img * n + row0 + r) * (m + 1) + k] = make_float2(c * xi.x, -c * xi.y);
  }
}

__global__ __launch_bounds__(NT) void adj_cols_kernel(const float2* __restrict__ s, int n, int n_img,
                                                      float2* __restrict__ gf) {
  extern __shared__ float2 smem[];
  const int m = n / 2;
  const int tid = threadIdx.x;
  const int sw = strip_w(n);
  const int nstrips = m / sw;
  const int img = blockIdx.x / nstrips;
  const int strip = blockIdx.x % nstrips;
  const int ncols = sw + (strip == 0 ? 1 : 0);
  float2* tw = smem;
  float2* bufa = tw + n;
  float2* bufb = bufa + n * (sw + 1);
  fill_twiddles(tw, n, tid, NT);
  const float2* si = s + (size_t)img * n * (m + 1);
  for (int e = tid; e < n * ncols; e += NT) {
    const int c = e % ncols, y = e / ncols;
    const int j = (c < sw) ? strip * sw + c : m;
    bufa[y * ncols + c] = cconj(si[(size_t)y * (m + 1) + j]);  // forward FFT = conj(IDFT(conj(.)))
  }
  __syncthreads();
  const float2* res = stockham_inverse(bufa, bufb, n, ncols, ncols, 1, tw, 1, tid, NT);
  const float inv = 1.0f / ((float)n * (float)n);
  float2* gi = gf + (size_t)img * n * (m + 1);
  for (int e = tid; e < n * ncols; e += NT) {
    const int c = e % ncols, ky = e / ncols;
    const int j = (c < sw) ? strip * sw + c : m;
    const float sg = ((ky + j) & 1) ? -inv : inv;
    const float2 v = res[ky * ncols + c];
    gi[(size_t)ky * (m + 1) + j] = make_float2(sg * v.x, -sg * v.y);
  }
}

// Both transforms of one image in LDS (boxes 64 and 128, whose half-plane of (N/2 + 1) complex
// columns fits): the rows as in adj_rows_kernel (M-point inverse FFT of the real-pair packing,
// four-step 16 x M/16 in registers, then the unpacking with the C2R weights), the columns as
// in adj_cols_kernel (N-point inverse FFT, four-step 16 x N/16, of conj(S), written as
// sign conj(.) / N^2). One workgroup per image.
constexpr int FFT_NT = 512;
template <int N>
__global__ __launch_bounds__(FFT_NT) void adj_fft_kernel(const float* __restrict__ ct, float2* __restrict__ gf) {
  constexpr int M = N / 2, M2 = M / 16, N2 = N / 16, RS = M + 4, HALF = M / 2 + 1;
  static_assert(N % 32 == 0 && (N * 16) % FFT_NT == 0, "unsupported box");
  __shared__ float2 tw[N];
  __shared__ float2 a[N * RS];  // row r at a + r RS: z (M values), then S without the conj (M + 1)
  const int tid = threadIdx.x;
  const size_t img = blockIdx.x;
  fill_twiddles(tw, N, tid, FFT_NT);
  // z[t] = x[2t] + i x[2t+1]
  const float2* src = reinterpret_cast<const float2*>(ct + img * N * N);
  for (int e = tid; e < N * M; e += FFT_NT) a[(e / M) * RS + e % M] = src[e];
  __syncthreads();
  // rows, step 1 (in place): thread (r, n2), 16-point DFT over z[M2 n1 + n2], twiddle
  // exp(2 pi i n2 k1 / M), back to k1 M2 + n2
  for (int t = tid; t < N * M2; t += FFT_NT) {
    const int n2 = t % M2, r = t / M2;
    float2* row = a + r * RS;
    float2 x[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) x[n1] = row[M2 * n1 + n2];
    Dft<16>::run(x);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) row[k1 * M2 + n2] = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], tw[2 * n2 * k1]);
  }
  __syncthreads();
  // rows, step 2: thread (r, k1), M2-point DFT -> Z[k1 + 16 k2] (all reads before the writes)
  constexpr int R2 = N * 16 / FFT_NT;
  float2 u[R2][M2];
#pragma unroll
  for (int q = 0; q < R2; ++q) {
    const int t = tid + q * FFT_NT, k1 = t % 16, r = t / 16;
#pragma unroll
    for (int n2 = 0; n2 < M2; ++n2) u[q][n2] = a[r * RS + k1 * M2 + n2];
    Dft<M2>::run(u[q]);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R2; ++q) {
    const int t = tid + q * FFT_NT, k1 = t % 16, r = t / 16;
#pragma unroll
    for (int k2 = 0; k2 < M2; ++k2) a[r * RS + k1 + 16 * k2] = u[q][k2];
  }
  __syncthreads();
  // unpacking, in place per (k, M - k) pair: X[k] = E[k] + W^k O[k] with E = (Z[k] + Z*[M-k]) / 2,
  // O = (Z[k] - Z*[M-k]) / (2i), W = exp(2 pi i / N); stored c X[k] (c = 1 at k = 0, M, else 2)
  auto unpack = [&](float2 zk, float2 zm, float2 w) {
    const float2 zc = cconj(zm);
    const float2 ev = cscale(cadd(zk, zc), 0.5f);
    const float2 d = cscale(csub(zk, zc), 0.5f);  // i O
    return cadd(ev, cmul(w, make_float2(d.y, -d.x)));
  };
  for (int e = tid; e < N * HALF; e += FFT_NT) {
    const int r = e / HALF, kk = e % HALF;
    float2* row = a + r * RS;
    const float2 zk = row[kk], zm = row[(M - kk) % M];
    const float ck = kk == 0 ? 1.f : 2.f;
    row[kk] = cscale(unpack(zk, zm, tw[kk]), ck);
    if (kk == 0) row[M] = unpack(zk, zm, make_float2(-1.f, 0.f));
    else if (2 * kk != M) row[M - kk] = cscale(unpack(zm, zk, tw[M - kk]), 2.f);
  }
  __syncthreads();
  // columns, step 1 (in place): thread (j, n2), 16-point DFT over rows N2 n1 + n2, twiddle
  // exp(2 pi i n2 k1 / N), back to rows k1 N2 + n2
  for (int t = tid; t < (M + 1) * N2; t += FFT_NT) {
    const int j = t % (M + 1), n2 = t / (M + 1);
    float2 x[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) x[n1] = a[(N2 * n1 + n2) * RS + j];
    Dft<16>::run(x);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1)
      a[(k1 * N2 + n2) * RS + j] = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], tw[n2 * k1]);
  }
  __syncthreads();
  // columns, step 2: thread (j, k1), N2-point DFT -> rows ky = k1 + 16 k2 of g_F
  constexpr float inv = 1.0f / ((float)N * (float)N);
  float2* go = gf + img * N * (M + 1);
  for (int t = tid; t < (M + 1) * 16; t += FFT_NT) {
    const int j = t % (M + 1), k1 = t / (M + 1);
    float2 v[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) v[n2] = a[(k1 * N2 + n2) * RS + j];
    Dft<N2>::run(v);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      const int ky = k1 + 16 * k2;
      const float sg = ((ky + j) & 1) ? -inv : inv;
      go[(size_t)ky * (M + 1) + j] = make_float2(sg * v[k2].x, -sg * v[k2].y);
    }
  }
}

// ---------------------------------------------------------------------------
// binned grid gradient (boxes up to 256)
// ---------------------------------------------------------------------------
//
// Each sample's eight corners touch at most eight bricks of voxels (usually
// one or two). The samples are grouped by brick with a counting sort, so that
// each brick visits exactly the samples that reach it. Which bricks a sample
// reaches depends on the rotations only, so the counts come before the gather:
//   bin_count_kernel  : per run of samples of one image (the gather block's run),
//                       the number of (sample, brick) pairs per brick in an LDS
//                       histogram; the block claims its place in each touched
//                       brick's list with one global atomic (rel, per block
//                       and brick);
//   bin_scan_kernel   : exclusive scan of the counts -> list offsets;
//   adj_gather_kernel : (with BIN) writes each of its pairs to the list, at
//                       offset + rel + its place in the block (LDS atomics);
//   bin_pull_kernel   : per brick, the forward's stencil of every listed
//                       sample, the corners inside the brick added into an LDS
//                       accumulator, and the brick added to the grid gradient
//                       once.
// A list entry is (image, row, column) packed in 32 bits; images are binned in
// chunks of at most BIN_CHUNK (fewer when the rel table would exceed
// REL_CAP_WORDS). Bricks are at most 16 x 16 x 8 voxels and at most 8192, so the
// histogram (32 KB) and the per-wave accumulators and claim words (at most
// 54 KB) fit in LDS.
constexpr int BIN_MAX_BOX = 256;
constexpr int BIN_MAX_BRICKS = 8192;
constexpr int BIN_MAX_VOX = 2048;
constexpr int BIN_CHUNK = 2048;            // images per binning pass at most (16-bit image field of an entry)
constexpr size_t REL_CAP_WORDS = 1u << 26;  // list places per (block, brick) of one pass, at most
constexpr unsigned BIN_HEAVY = 4096;  // a brick with more entries is pulled in segments by several workgroups
constexpr unsigned BIN_SEG = 1024;    // entries per segment of a heavy brick
// 32-bit words after the list places: counts, offsets (+1), the light brick list, the two list
// lengths; then the heavy segments ((segment << 13) | brick), heavy_words() of them at most
constexpr size_t BIN_META_WORDS = 3 * BIN_MAX_BRICKS + 64;
// Segments of the heavy bricks of a pass with at most `pairs` (sample, brick) pairs: each heavy
// brick has more than BIN_HEAVY of them, so sum ceil(c / BIN_SEG) <= pairs (1 / BIN_SEG + 1 / BIN_HEAVY).
__host__ __device__ inline size_t heavy_words(size_t pairs) { return pairs / BIN_SEG + pairs / BIN_HEAVY + 1; }
static_assert(BIN_MAX_BRICKS <= 8192, "a heavy segment's brick fits 13 bits");

struct BinGeom {
  int n, m, nh;
  int e[3];   // brick edge along x, y, z
  int nb[3];  // bricks along x, y, z
  float inv_e[3];
  int nbricks;
  int wrap;  // corner indices in [-N, -1] wrap (0: CRYOMI_GATHER_FILL, they drop)
};

BinGeom bin_geom(int n, bool wrap) {
  BinGeom g{};
  g.wrap = wrap ? 1 : 0;
  g.n = n;
  g.m = n / 2;
  g.nh = n * (n / 2 + 1);
  g.e[0] = g.e[1] = std::max(8, (n + 15) / 16);
  g.e[2] = std::max(4, (n + 31) / 32);
  for (int a = 0; a < 3; ++a) {
    g.nb[a] = (n + g.e[a] - 1) / g.e[a];
    g.inv_e[a] = 1.f / (float)g.e[a];
  }
  g.nbricks = g.nb[0] * g.nb[1] * g.nb[2];
  return g;
}

// Brick coordinate of voxel index w < 256 along an axis of edge e: floor((w + 1/2) / e)
// equals floor(w / e), and (w + 1/2) / e lies at least 1 / (2e) from an integer, far
// beyond the rounding of the FP32 product.
__device__ __forceinline__ int brick_of(int w, float inv_e) { return (int)(((float)w + 0.5f) * inv_e); }

// The forward's stencil of sample (row r, column j): lower corner and FP32 fractions,
// axes (z, y, x), with q = R (kx, ky, 0) + N/2 as the forward forms it.
__device__ __forceinline__ void bin_stencil(double r00, double r01, double r10, double r11, double r20, double r21,
                                            int n, int r, int j, int lo[3], float fr[3]) {
  const int m = n / 2;
  const double fx = (double)(j < m ? j : j - n), fy = (double)(r < m ? r : r - n);
  const double half = (double)m;
  const double cc[3] = {fma(r20, fx, fma(r21, fy, half)), fma(r10, fx, fma(r11, fy, half)),
                        fma(r00, fx, fma(r01, fy, half))};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double f = floor(cc[a]);
    lo[a] = (int)f;
    fr[a] = (float)(cc[a] - f);
  }
}

// Distinct bricks of the valid corner indices lo, lo + 1 along one axis (0, 1 or 2 of them).
__device__ __forceinline__ int axis_bricks(int lo, int n, float inv_e, bool wrap, int& b0, int& b1) {
  int w0, w1;
  const bool v0 = wrap_index(lo, n, w0, wrap), v1 = wrap_index(lo + 1, n, w1, wrap);
  const int c0 = brick_of(v0 ? w0 : 0, inv_e), c1 = brick_of(v1 ? w1 : 0, inv_e);
  b0 = v0 ? c0 : c1;
  b1 = c1;
  return (int)v0 + (int)(v1 && !(v0 && c1 == c0));
}

// Calls f(brick) for each brick reached by a sample with lower corner lo (axes z, y, x).
template <typename F>
__device__ __forceinline__ void for_each_brick_lo(const BinGeom& g, const int lo[3], F&& f) {
  int z0, z1, y0, y1, x0, x1;
  const bool wrap = g.wrap != 0;
  const int nz = axis_bricks(lo[0], g.n, g.inv_e[2], wrap, z0, z1);
  const int ny = axis_bricks(lo[1], g.n, g.inv_e[1], wrap, y0, y1);
  const int nx = axis_bricks(lo[2], g.n, g.inv_e[0], wrap, x0, x1);
  for (int a = 0; a < nz; ++a)
    for (int b = 0; b < ny; ++b)
      for (int c = 0; c < nx; ++c) f(((a ? z1 : z0) * g.nb[1] + (b ? y1 : y0)) * g.nb[0] + (c ? x1 : x0));
}

// Calls f(brick) for each brick that sample (r, j) of an image with rotation rot reaches.
template <typename F>
__device__ __forceinline__ void for_each_brick(const BinGeom& g, const double* rot, int r, int j, F&& f) {
  int lo[3];
  float fr[3];
  bin_stencil(rot[0], rot[1], rot[3], rot[4], rot[6], rot[7], g.n, r, j, lo, fr);
  for_each_brick_lo(g, lo, f);
}

// An image's nh samples in `parts` even runs of about `target` samples each (run k is
// [k spb, min((k + 1) spb, nh))): no run is left with a small remainder.
__host__ __device__ inline int sample_parts(int nh, int target) { return nh > target ? (nh + target / 2) / target : 1; }
__host__ __device__ inline int sample_run(int nh, int parts) { return (nh + parts - 1) / parts; }

// Words of a block's touched-brick bitmap (bit b & 31 of word b >> 5 for brick b).
__host__ __device__ inline int bitmap_words(int nbricks) { return (nbricks + 31) / 32; }

// The histogram of brick counts, cleared four words at a time up to a multiple of 32 bricks
// (whole bitmap words).
__device__ __forceinline__ void hist_clear(unsigned* hist, int nbricks) {
  uint4* h4 = reinterpret_cast<uint4*>(hist);
  for (int b = threadIdx.x; b < 8 * bitmap_words(nbricks); b += blockDim.x) h4[b] = make_uint4(0u, 0u, 0u, 0u);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < NT / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

// One block per run of about GATHER_SPB samples of one image (a loop of samples per thread), so the nine
// per-image gradient sums are reduced and added once per block. With BIN, the block also
// writes its (sample, brick) pairs into the brick lists: pair k of the block for brick b goes
// to entries[offsets[b] + rel[block][b] + k] (bin_count_kernel counted the same pairs and marked
// the block's bricks in its bitmap); the list places of the marked bricks are read into LDS first.
constexpr int GATHER_SPB = 2800;  // samples per gather (and count) block, target; images split evenly
template <bool BIN, bool BRICKED>
__global__ __launch_bounds__(NT) void adj_gather_kernel(const float2* __restrict__ grid, int n, int n_img,
                                                        const cryomi_image_params* __restrict__ prm,
                                                        unsigned flags, const float2* __restrict__ gf,
                                                        float2* __restrict__ gpbuf, double* __restrict__ gparams,
                                                        BinGeom bg, const unsigned* __restrict__ offsets,
                                                        const unsigned* __restrict__ rel,
                                                        const unsigned* __restrict__ bitmap,
                                                        unsigned* __restrict__ entries) {
  __shared__ double red[NT / 64];
  __shared__ __align__(16) unsigned hist[BIN ? BIN_MAX_BRICKS : 4];  // next list place per brick
  __shared__ unsigned bmw[BIN ? BIN_MAX_BRICKS / 32 : 1];
  const int m = n / 2;
  const int nh = n * (m + 1);
  const int parts = sample_parts(nh, GATHER_SPB), spb = sample_run(nh, parts);
  const int img = blockIdx.x / parts;
  const int s0 = (blockIdx.x % parts) * spb, s1 = min(s0 + spb, nh);
  const cryomi_image_params& p = prm[img];
  const ImageConsts k = make_consts(p, n, false, flags);
  if (BIN) {
    const int nw = bitmap_words(bg.nbricks);
    const unsigned* rb = rel + (size_t)blockIdx.x * bg.nbricks;
    for (int w = threadIdx.x; w < nw; w += NT) bmw[w] = bitmap[(size_t)blockIdx.x * nw + w];
    __syncthreads();
#pragma unroll 8
    for (int b = threadIdx.x; b < bg.nbricks; b += NT)
      if ((bmw[b >> 5] >> (b & 31)) & 1u) hist[b] = offsets[b] + rb[b];
    __syncthreads();
  }
  float acc[N_GRAD] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  // sample s = i (m + 1) + j, stepped by NT without a division per sample
  const int di = NT / (m + 1), dj = NT % (m + 1);
  int i = (s0 + (int)threadIdx.x) / (m + 1), j = (s0 + (int)threadIdx.x) % (m + 1);
  for (int s = s0 + threadIdx.x; s < s1; s += NT, i += di, j += dj) {
    if (j > m) {
      j -= m + 1;
      ++i;
    }
    const bool selfc = (i == m || i == 0) && (j == 0 || j == m) && !(i == 0 && j == 0);
    const bool translate = (flags & CRYOMI_APPLY_TRANSLATION) != 0;
    const int jx = (j < m) ? j : j - n;
    const int iy = (i < m) ? i : i - n;
    const double fy = (double)iy;
    // grid axes (z, y, x): the stencil bin_count_kernel bins by
    int lo[3];
    float fr[3];
    bin_stencil(k.r00, k.r01, k.r10, k.r11, k.r20, k.r21, n, i, j, lo, fr);
    // forward sample and its coordinate derivatives
    float2 pv = make_float2(0.f, 0.f);
    float2 dp[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    if (BRICKED) {
      // the forward's bricked gather (sample.hpp gather_issue<true>): one x pair per (z, y) row
      // from the brick (x + 1) / 3 that holds both corners; out-of-range corners masked
      int z0, z1, y0, y1;
      const bool vz0 = wrap_index(lo[0], n, z0, k.wrap), vz1 = wrap_index(lo[0] + 1, n, z1, k.wrap);
      const bool vy0 = wrap_index(lo[1], n, y0, k.wrap), vy1 = wrap_index(lo[1] + 1, n, y1, k.wrap);
      const int lx = lo[2];
      const bool inside = lx >= -1 && lx <= n - 1;
      const bool wraps = k.wrap && lx < -1 && lx >= -n;
      const int xq = inside ? lx : (wraps ? lx + n : 0);
      const float mlo = ((inside && (k.wrap || lx >= 0)) || wraps) ? 1.f : 0.f;
      const float mhi = (inside || wraps) ? 1.f : 0.f;
      float4a8 gq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const bool v = (q >> 1 ? vz1 : vz0) && (q & 1 ? vy1 : vy0);
        const int zq = v ? (q >> 1 ? z1 : z0) : 0, yq = v ? (q & 1 ? y1 : y0) : 0;
        gq[q] = *reinterpret_cast<const float4a8*>(grid + pair_offset(zq, yq, xq, n, true));
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int bz = q >> 1, by = q & 1;
        const bool v = (bz ? vz1 : vz0) && (by ? vy1 : vy0);
        const float ml = v ? mlo : 0.f, mh = v ? mhi : 0.f;
        const float2 lo_v = make_float2(ml * gq[q].x, ml * gq[q].y), hi_v = make_float2(mh * gq[q].z, mh * gq[q].w);
        const float2 xs = make_float2(fmaf(fr[2], hi_v.x - lo_v.x, lo_v.x), fmaf(fr[2], hi_v.y - lo_v.y, lo_v.y));
        const float2 xd = make_float2(hi_v.x - lo_v.x, hi_v.y - lo_v.y);
        const float wz = bz ? fr[0] : 1.f - fr[0], wy = by ? fr[1] : 1.f - fr[1];
        const float dz = bz ? 1.f : -1.f, dy = by ? 1.f : -1.f;
        const float wzy = wz * wy;
        pv = make_float2(fmaf(wzy, xs.x, pv.x), fmaf(wzy, xs.y, pv.y));
        dp[0] = make_float2(fmaf(dz * wy, xs.x, dp[0].x), fmaf(dz * wy, xs.y, dp[0].y));
        dp[1] = make_float2(fmaf(wz * dy, xs.x, dp[1].x), fmaf(wz * dy, xs.y, dp[1].y));
        dp[2] = make_float2(fmaf(wzy, xd.x, dp[2].x), fmaf(wzy, xd.y, dp[2].y));
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int b0 = q >> 2, b1 = (q >> 1) & 1, b2 = q & 1;
        int z, y, x;
        const bool vz = wrap_index(lo[0] + b0, n, z, k.wrap);
        const bool vy = wrap_index(lo[1] + b1, n, y, k.wrap);
        const bool vx = wrap_index(lo[2] + b2, n, x, k.wrap);
        const bool v = vz && vy && vx;
        const float w0 = b0 ? fr[0] : 1.f - fr[0], w1 = b1 ? fr[1] : 1.f - fr[1], w2 = b2 ? fr[2] : 1.f - fr[2];
        const float d0 = b0 ? 1.f : -1.f, d1 = b1 ? 1.f : -1.f, d2 = b2 ? 1.f : -1.f;
        const float wq = w0 * w1 * w2;
        const float2 g = v ? grid[((size_t)z * n + y) * n + x] : make_float2(0.f, 0.f);
        pv.x = fmaf(wq, g.x, pv.x);
        pv.y = fmaf(wq, g.y, pv.y);
        dp[0] = make_float2(fmaf(d0 * w1 * w2, g.x, dp[0].x), fmaf(d0 * w1 * w2, g.y, dp[0].y));
        dp[1] = make_float2(fmaf(w0 * d1 * w2, g.x, dp[1].x), fmaf(w0 * d1 * w2, g.y, dp[1].y));
        dp[2] = make_float2(fmaf(w0 * w1 * d2, g.x, dp[2].x), fmaf(w0 * w1 * d2, g.y, dp[2].y));
      }
    }
    if (selfc) {
      pv.y = 0.f;
      if (translate) pv.x = 0.f;
    }
    // CTF and shift: phases in FP64 (as the forward forms them), reduced to [-1/2, 1/2] in FP64,
    // then FP32 sin / cos of the reduced phase; the gradient terms in FP32 per thread (a few
    // samples each), summed over the block in FP64
    float ctf = 1.f, dctf = 0.f;
    const double inv = 1.0 / ((double)n * p.pixel_size);
    if (flags & CRYOMI_APPLY_CTF) {
      const double jj = (double)j;
      const double j2 = jj * jj, i2 = fy * fy, ij = fy * jj;
      const double kk = j2 + i2;
      const double u = fma(k.ca, kk, fma(k.cb, i2 - j2, fma(k.cc, ij, fma(k.cd, kk * kk, -k.cphase))));
      float sn, cs;
      sincospif(2.f * (float)(u - rint(u)), &sn, &cs);
      ctf = sn;
      // d chi / d defocus = -pi lambda k^2 with k in 1/Angstrom
      dctf = cs * (float)(-M_PI * p.wavelength * kk * inv * inv);
    }
    float sr = 1.f, si = 0.f;  // shift = exp(-2 pi i v)
    if (flags & CRYOMI_APPLY_TRANSLATION) {
      const double v = fma(k.tx, (double)j, k.ty * fy);
      float sn, cs;
      sincospif(2.f * (float)(v - rint(v)), &sn, &cs);
      sr = cs;
      si = -sn;
    }
    const float2 g = gf[(size_t)img * nh + s];
    const float amp = (float)p.amplitude;
    // g_P = amp CTF conj(shift) g_F
    const float ac = amp * ctf;
    float gpr = ac * (sr * g.x + si * g.y);
    float gpi = ac * (sr * g.y - si * g.x);
    if (selfc) {
      gpi = 0.f;
      if (translate) gpr = 0.f;
    }
    const float2 gpf = make_float2(gpr, gpi);
    gpbuf[(size_t)img * nh + s] = gpf;
    if (BIN) {
      const unsigned key = ((unsigned)img << 16) | ((unsigned)i << 8) | (unsigned)j;
      for_each_brick_lo(bg, lo, [&](int b) { entries[atomicAdd(&hist[b], 1u)] = key; });
    }
    {
      // dL/dc_a = Re(conj(g_P) dP/dc_a); c_(x,y,z) = R (jx, iy, 0) + N/2
      const float dlz = gpr * dp[0].x + gpi * dp[0].y;
      const float dly = gpr * dp[1].x + gpi * dp[1].y;
      const float dlx = gpr * dp[2].x + gpi * dp[2].y;
      const float ffx = (float)jx, ffy = (float)iy;
      acc[0] += dlx * ffx;
      acc[1] += dlx * ffy;
      acc[2] += dly * ffx;
      acc[3] += dly * ffy;
      acc[4] += dlz * ffx;
      acc[5] += dlz * ffy;
      // F = amp shift CTF P
      const float fr_ = ac * (sr * pv.x - si * pv.y);
      const float fi_ = ac * (sr * pv.y + si * pv.x);
      // dL/dt = Re(conj(g_F) (-2 pi i q) F)
      const float cross = g.x * fi_ - g.y * fr_;  // Re(conj(g) * (-i) * F)
      if (flags & CRYOMI_APPLY_TRANSLATION) {
        const float tw = (float)(2.0 * M_PI * inv);
        acc[6] += tw * (float)j * cross;
        acc[7] += tw * ffy * cross;
      }
      // dL/d defocus = Re(conj(g_F) amp shift P) dCTF
      acc[8] += amp * dctf * (g.x * (sr * pv.x - si * pv.y) + g.y * (sr * pv.y + si * pv.x));
    }
  }
  for (int a = 0; a < N_GRAD; ++a) {
    const double t = block_sum((double)acc[a], red);
    if (threadIdx.x == 0 && t != 0.0) atomicAdd(gparams + (size_t)img * N_GRAD + a, t);
  }
}


constexpr int PBX = 8, PBY = 8, PBZ = 4;  // pull brick (x, y, z); PBX * PBY * PBZ == NT
constexpr int PCH = 256;                   // images per culling chunk
constexpr int PTW = 16;                    // sample window edge: covers 2 (sqrt 3 + brick half-diagonal) + 1
constexpr int PB = 16;                     // items whose sample loads are in flight together
constexpr float SQRT3 = 1.7320508f;
static_assert(PBX * PBY * PBZ == NT && PTW * PTW == NT, "one thread per brick voxel and per window sample");

// one (image, wrap combination) whose slice can reach the brick; wrap w: x bit 0, y bit 1, z bit 2
struct PullItem {
  short img;
  short w;
};

__global__ __launch_bounds__(NT) void adj_pull_kernel(int n, int n_img, const cryomi_image_params* __restrict__ prm,
                                                      int wrap, const float2* __restrict__ gpbuf,
                                                      float2* __restrict__ ggrid) {
  __shared__ double rq[PCH][6];  // r00 r10 r20 r01 r11 r21: the slice's two in-plane axes, FP64 as the forward
  __shared__ float rf[PCH][9];   // FP32 copy of the whole rotation for culling
  __shared__ PullItem items[PCH * 8];
  __shared__ int wave_count[NT / 64];
  __shared__ float accs[NT][2];  // this brick's gradient, voxel (x + PBX (y + PBY z))
  const int m = n / 2;
  const int nh = n * (m + 1);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nbx = (n + PBX - 1) / PBX, nby = (n + PBY - 1) / PBY;
  const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
  const int x0 = bx * PBX, y0 = by * PBY, z0 = bz * PBZ;
  const float half = (float)m;
  // brick centre and half-diagonal (voxel positions b0 .. b0 + P - 1)
  const float cx = x0 + 0.5f * (PBX - 1), cy = y0 + 0.5f * (PBY - 1), cz = z0 + 0.5f * (PBZ - 1);
  const float hd = 0.5f * sqrtf((float)((PBX - 1) * (PBX - 1) + (PBY - 1) * (PBY - 1) + (PBZ - 1) * (PBZ - 1)));
  const float reach = SQRT3 + hd + 0.01f;
  // wrap combinations this brick can receive: virtual coordinate v - N must be within reach of a sample,
  // q_a >= N/2 - (N/2) sqrt(2) (|k| <= m sqrt 2), so the brick's top voxel needs v - N > N/2 - m sqrt 2 - 1
  const float lowest_q = half - (float)m * 1.41421356f - 1.01f;
  const int bmax[3] = {min(x0 + PBX - 1, n - 1), min(y0 + PBY - 1, n - 1), min(z0 + PBZ - 1, n - 1)};
  int allowed = 0;
  for (int w = 0; w < (wrap ? 8 : 1); ++w) {  // without wrapping only the brick itself (w = 0) receives
    bool ok = true;
    for (int a = 0; a < 3; ++a)
      if ((w >> a) & 1) ok = ok && ((float)(bmax[a] - n) > lowest_q);
    if (ok) allowed |= 1 << w;
  }
  accs[tid][0] = 0.f;
  accs[tid][1] = 0.f;
  for (int i0 = 0; i0 < n_img; i0 += PCH) {
    const int cnt = min(PCH, n_img - i0);
    __syncthreads();  // the previous chunk's tables are consumed (and accs is initialised)
    // rotations of this chunk, and which (image, wrap) pairs can reach the brick
    int mask = 0;
    if (tid < cnt) {
      const double* r = prm[i0 + tid].rotation;
      float rr[9];
#pragma unroll
      for (int e = 0; e < 9; ++e) rr[e] = (float)r[e];
#pragma unroll
      for (int e = 0; e < 9; ++e) rf[tid][e] = rr[e];
      rq[tid][0] = r[0];
      rq[tid][1] = r[3];
      rq[tid][2] = r[6];
      rq[tid][3] = r[1];
      rq[tid][4] = r[4];
      rq[tid][5] = r[7];
      for (int w = 0; w < 8; ++w) {
        if (!((allowed >> w) & 1)) continue;
        const float dx = cx - (float)(n * (w & 1)) - half, dy = cy - (float)(n * ((w >> 1) & 1)) - half,
                    dz = cz - (float)(n * ((w >> 2) & 1)) - half;
        // p = R^T d: in-plane coordinates (px, py) and distance to the slice plane pz
        const float px = rr[0] * dx + rr[3] * dy + rr[6] * dz;
        const float py = rr[1] * dx + rr[4] * dy + rr[7] * dz;
        const float pz = rr[2] * dx + rr[5] * dy + rr[8] * dz;
        // samples: kx in {-m} u [0, m), ky in [-m, m). A voxel receives only from samples within
        // |q - c|_inf < 1, so the slice must pass within sum_a |n_a| (brick half-extent_a + 1) of
        // the brick centre along its normal n = (r02, r12, r22)
        const float slab = fabsf(rr[2]) * (0.5f * (PBX - 1) + 1.f) + fabsf(rr[5]) * (0.5f * (PBY - 1) + 1.f) +
                           fabsf(rr[8]) * (0.5f * (PBZ - 1) + 1.f) + 0.01f;
        const bool hit = fabsf(pz) < slab && px > -(float)m - reach && px < (float)m + reach &&
                         py > -(float)m - reach && py < (float)m + reach;
        if (hit) mask |= 1 << w;
      }
    }
    // ordered compaction into (image, wrap) items: wave-level inclusive scan of the counts
    const int c = __popc(mask);
    int incl = c;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) wave_count[wave] = incl;
    __syncthreads();
    int base = 0, total = 0;
    for (int w = 0; w < NT / 64; ++w) {
      if (w < wave) base += wave_count[w];
      total += wave_count[w];
    }
    {
      int pos = base + incl - c;
      for (int w = 0; w < 8; ++w)
        if ((mask >> w) & 1) items[pos++] = PullItem{(short)tid, (short)w};
    }
    __syncthreads();
    // items in batches of PB: the batch's sample loads are issued together, so one memory latency
    // is exposed per batch (bricks near the grid centre are crossed by every slice and walk the
    // longest item lists)
    auto window = [&](int e, int& kx, int& ky, int& img, int& w) {
      const PullItem it = items[e];
      img = it.img;
      w = it.w;
      const float* rr = rf[img];
      // window of samples that can reach the brick: the brick centre's in-plane projection
      // +- (sqrt 3 + half-diagonal); one thread per window sample
      const float dcx = cx - (float)(n * (w & 1)) - half, dcy = cy - (float)(n * ((w >> 1) & 1)) - half,
                  dcz = cz - (float)(n * ((w >> 2) & 1)) - half;
      const float pcx = rr[0] * dcx + rr[3] * dcy + rr[6] * dcz;
      const float pcy = rr[1] * dcx + rr[4] * dcy + rr[7] * dcz;
      kx = (int)floorf(pcx - reach) + (tid % PTW);
      ky = (int)floorf(pcy - reach) + tid / PTW;
    };
    auto is_sample = [&](int kx, int ky) { return ky >= -m && ky < m && ((kx >= 0 && kx < m) || kx == -m); };
    for (int e0 = 0; e0 < total; e0 += PB) {
      float2 gvb[PB];
#pragma unroll
      for (int b = 0; b < PB; ++b) {
        int kx = 0, ky = 0, img = 0, w = 0;
        if (e0 + b < total) window(e0 + b, kx, ky, img, w);
        const bool smp = e0 + b < total && is_sample(kx, ky);
        const float2* g = gpbuf + (size_t)(i0 + img) * nh;
        const float2 v = g[smp ? (size_t)(ky < 0 ? ky + n : ky) * (m + 1) + (kx < 0 ? m : kx) : 0];
        gvb[b] = smp ? v : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int b = 0; b < PB; ++b) {
        const float2 gv = gvb[b];
        if (gv.x == 0.f && gv.y == 0.f) continue;  // also every slot past the list
        int kx, ky, img, w;
        window(e0 + b, kx, ky, img, w);
        const double* rd = rq[img];
        const int wx = n * (w & 1), wy = n * ((w >> 1) & 1), wz = n * ((w >> 2) & 1);
        // the forward's stencil: grid position q (x, y, z) in FP64, lower corner and FP32 fractions
        const double fx = (double)kx, fy = (double)ky;
        const double qx = fma(rd[0], fx, fma(rd[3], fy, (double)m));
        const double qy = fma(rd[1], fx, fma(rd[4], fy, (double)m));
        const double qz = fma(rd[2], fx, fma(rd[5], fy, (double)m));
        const double flx = floor(qx), fly = floor(qy), flz = floor(qz);
        const float frx = (float)(qx - flx), fry = (float)(qy - fly), frz = (float)(qz - flz);
        // corner index -> brick-local index (the virtual brick is shifted by -N on wrapped axes)
        const int lx = (int)flx - (x0 - wx), ly = (int)fly - (y0 - wy), lz = (int)flz - (z0 - wz);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int b0 = q >> 2, b1 = (q >> 1) & 1, b2 = q & 1;  // z, y, x as the forward's corner order
          const int iz = lz + b0, iy = ly + b1, ix = lx + b2;
          if ((unsigned)ix < (unsigned)PBX && (unsigned)iy < (unsigned)PBY && (unsigned)iz < (unsigned)PBZ &&
              x0 + ix < n && y0 + iy < n && z0 + iz < n) {
            const float wq = (b0 ? frz : 1.f - frz) * (b1 ? fry : 1.f - fry) * (b2 ? frx : 1.f - frx);
            float* acc = accs[(iz * PBY + iy) * PBX + ix];
            atomicAdd(acc, wq * gv.x);
            atomicAdd(acc + 1, wq * gv.y);
          }
        }
      }
    }
  }
  __syncthreads();
  const int vx = x0 + tid % PBX, vy = y0 + (tid / PBX) % PBY, vz = z0 + tid / (PBX * PBY);
  if (vx < n && vy < n && vz < n) {
    float2* o = ggrid + ((size_t)vz * n + vy) * n + vx;
    const float2 prev = *o;
    *o = make_float2(prev.x + accs[tid][0], prev.y + accs[tid][1]);
  }
}

__global__ __launch_bounds__(NT) void bin_zero_kernel(unsigned* __restrict__ p, int count) {
  for (int i = blockIdx.x * NT + threadIdx.x; i < count; i += gridDim.x * NT) p[i] = 0u;
}

// One block of 1024 threads: offsets = exclusive scan of counts (offsets[nbricks] = total),
// and the counts cleared for the next chunk. The non-empty bricks are listed
// in `lists`, their number in list_n[0]; those above BIN_HEAVY entries are cut into segments of
// BIN_SEG entries, listed in `heavy` as (segment << 13) | brick, their number in list_n[1]. The
// order within a list does not matter.
constexpr int SCAN_NT = 1024;
__global__ __launch_bounds__(SCAN_NT) void bin_scan_kernel(int nbricks, unsigned* __restrict__ counts,
                                                           unsigned* __restrict__ offsets,
                                                           unsigned* __restrict__ lists,
                                                           unsigned* __restrict__ heavy,
                                                           unsigned* __restrict__ list_n) {
  __shared__ unsigned wave_tot[SCAN_NT / 64];
  __shared__ unsigned n_light, n_heavy;
  if (threadIdx.x == 0) n_light = n_heavy = 0u;
  constexpr int PER = BIN_MAX_BRICKS / SCAN_NT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  unsigned v[PER], sum = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int b = tid * PER + k;
    v[k] = b < nbricks ? counts[b] : 0u;
    sum += v[k];
  }
  unsigned incl = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wave_tot[wave] = incl;
  __syncthreads();
  unsigned base = 0, total = 0;
  for (int w = 0; w < SCAN_NT / 64; ++w) {
    if (w < wave) base += wave_tot[w];
    total += wave_tot[w];
  }
  unsigned run = base + incl - sum;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int b = tid * PER + k;
    if (b < nbricks) {
      offsets[b] = run;
      counts[b] = 0u;
      if (v[k] > BIN_HEAVY) {
        const unsigned ns = (v[k] + BIN_SEG - 1) / BIN_SEG, at = atomicAdd(&n_heavy, ns);
        for (unsigned q = 0; q < ns; ++q) heavy[at + q] = (q << 13) | (unsigned)b;
      } else if (v[k] > 0u) {
        lists[atomicAdd(&n_light, 1u)] = (unsigned)b;
      }
    }
    run += v[k];
  }
  if (tid == 0) offsets[nbricks] = total;
  __syncthreads();
  if (tid == 0) {
    list_n[0] = n_light;
    list_n[1] = n_heavy;
  }
}

constexpr int COUNT_NT = 512;
__global__ __launch_bounds__(COUNT_NT) void bin_count_kernel(BinGeom g, const cryomi_image_params* __restrict__ prm,
                                                             int parts, unsigned* __restrict__ counts,
                                                             unsigned* __restrict__ rel,
                                                             unsigned* __restrict__ bitmap) {
  __shared__ __align__(16) unsigned hist[BIN_MAX_BRICKS];
  const int tid = threadIdx.x;
  const int spb = sample_run(g.nh, parts);
  const int img = blockIdx.x / parts, s0 = (blockIdx.x % parts) * spb, s1 = min(s0 + spb, g.nh);
  hist_clear(hist, g.nbricks);
  __syncthreads();
  const double* rot = prm[img].rotation;
  // sample s = r (m + 1) + j, stepped by COUNT_NT without a division per sample
  const int dr = COUNT_NT / (g.m + 1), dj = COUNT_NT % (g.m + 1);
  int r = (s0 + tid) / (g.m + 1), j = (s0 + tid) % (g.m + 1);
  for (int s = s0 + tid; s < s1; s += COUNT_NT, r += dr, j += dj) {
    if (j > g.m) {
      j -= g.m + 1;
      ++r;
    }
    for_each_brick(g, rot, r, j, [&](int b) { atomicAdd(&hist[b], 1u); });
  }
  __syncthreads();
  // this block's place in each touched brick's list
  unsigned* rb = rel + (size_t)blockIdx.x * g.nbricks;
  const uint4* h4 = reinterpret_cast<const uint4*>(hist);
  for (int b = tid; b < (g.nbricks + 3) / 4; b += COUNT_NT) {
    const uint4 h = h4[b];  // words past nbricks stay zero
    if (h.x) rb[4 * b] = atomicAdd(counts + 4 * b, h.x);
    if (h.y) rb[4 * b + 1] = atomicAdd(counts + 4 * b + 1, h.y);
    if (h.z) rb[4 * b + 2] = atomicAdd(counts + 4 * b + 2, h.z);
    if (h.w) rb[4 * b + 3] = atomicAdd(counts + 4 * b + 3, h.w);
  }
  // the bitmap of touched bricks (the histogram is zero past nbricks up to whole words)
  const int nw = bitmap_words(g.nbricks);
  for (int w = tid; w < nw; w += COUNT_NT) {
    unsigned bits = 0u;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint4 h = h4[8 * w + q];
      bits |= ((h.x != 0u) | (h.y != 0u) << 1 | (h.z != 0u) << 2 | (h.w != 0u) << 3) << (4 * q);
    }
    bitmap[(size_t)blockIdx.x * nw + w] = bits;
  }
}

// Each wave adds into its own copy of the brick, so no other wave touches it, with plain
// LDS read-modify-writes (LDS float atomics run per lane, far below the plain LDS rate on
// this part). Lanes of one wave collide only when two samples share their lower corner:
// for a fixed corner q the targets L + q of lanes with different lower corners L differ.
// So each pending lane writes its lane id to the claim word of its lower corner, reads it
// back, and the lanes whose id stayed (one per lower corner) add all their corners; the
// others retry in the next round. Every round settles at least one lane per contended
// corner, and adjacent samples rarely share one, so a wave takes one or two rounds.
// Workgroups take the bricks of `list` (list_n of them) in turn; W is the workgroup's waves,
// BIN_BATCH the entries per lane whose loads are in flight together (more hold more registers
// and leave fewer waves per SIMD). With SEG, the list holds segments of heavy bricks
// ((segment << 13) | brick, BIN_SEG entries each), which several workgroups add into one brick
// with FP32 atomics; otherwise a workgroup owns its brick and adds it with plain stores.
template <int W, int BIN_BATCH, bool SEG>
__global__ __launch_bounds__(64 * W) void bin_pull_kernel(BinGeom g, const cryomi_image_params* __restrict__ prm,
                                                          const float2* __restrict__ gpbuf,
                                                          const unsigned* __restrict__ offsets,
                                                          const unsigned* __restrict__ entries,
                                                          const unsigned* __restrict__ list,
                                                          const unsigned* __restrict__ list_n,
                                                          float2* __restrict__ ggrid) {
  extern __shared__ float2 bin_lds[];  // W accumulators of vox float2, then W claim tables of kvox int
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned nl = *list_n;
#pragma unroll 1
  for (unsigned li = blockIdx.x; li < nl; li += gridDim.x) {
  const unsigned item = list[li];
  const int brick = SEG ? (int)(item & 8191u) : (int)item;
  const unsigned e0 = offsets[brick] + (SEG ? (item >> 13) * BIN_SEG : 0u);
  const unsigned e1 = SEG ? min(offsets[brick + 1], e0 + BIN_SEG) : offsets[brick + 1];
  const int ex = g.e[0], ey = g.e[1], ez = g.e[2];
  const int x0 = (brick % g.nb[0]) * ex, y0 = ((brick / g.nb[0]) % g.nb[1]) * ey, z0 = brick / (g.nb[0] * g.nb[1]) * ez;
  const int vox = ex * ey * ez;
  float2* wacc = bin_lds + wave * vox;
  const int kvox = (ex + 1) * (ey + 1) * (ez + 1);
  int* wclaim = reinterpret_cast<int*>(bin_lds + W * vox) + wave * kvox;
  for (int v = lane; v < vox; v += 64) wacc[v] = make_float2(0.f, 0.f);
  const int n = g.n;
  constexpr unsigned STEP = 64u * W;
  for (unsigned eb = e0 + wave * 64 + lane; eb - lane < e1; eb += STEP * BIN_BATCH) {
    // batches this wave holds entries of (the list's tail leaves the later ones empty)
    const unsigned nbat = min((unsigned)BIN_BATCH, (e1 - (eb - lane) + STEP - 1) / STEP);
    unsigned key[BIN_BATCH];
#pragma unroll
    for (int b = 0; b < BIN_BATCH; ++b) {
      if ((unsigned)b >= nbat) break;
      const unsigned e = eb + b * STEP;
      key[b] = entries[e < e1 ? e : e0];
    }
    float2 gv[BIN_BATCH];
    double rot[BIN_BATCH][6];
#pragma unroll
    for (int b = 0; b < BIN_BATCH; ++b) {
      if ((unsigned)b >= nbat) break;
      const int img = key[b] >> 16, r = (key[b] >> 8) & 0xff, j = key[b] & 0xff;
      gv[b] = gpbuf[(size_t)img * g.nh + r * (g.m + 1) + j];
      const double* rr = prm[img].rotation;
      rot[b][0] = rr[0];
      rot[b][1] = rr[1];
      rot[b][2] = rr[3];
      rot[b][3] = rr[4];
      rot[b][4] = rr[6];
      rot[b][5] = rr[7];
    }
#pragma unroll
    for (int b = 0; b < BIN_BATCH; ++b) {
      if ((unsigned)b >= nbat) break;
      const bool live = eb + b * STEP < e1 && (gv[b].x != 0.f || gv[b].y != 0.f);
      const int r = (key[b] >> 8) & 0xff, j = key[b] & 0xff;
      int lo[3];
      float fr[3];
      bin_stencil(rot[b][0], rot[b][1], rot[b][2], rot[b][3], rot[b][4], rot[b][5], n, r, j, lo, fr);
      // brick-local index of each corner along each axis (-1: invalid or outside the brick)
      int lz[2], ly[2], lx[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        int w;
        const bool wr = g.wrap != 0;
        lz[c] = wrap_index(lo[0] + c, n, w, wr) && (unsigned)(w - z0) < (unsigned)ez ? w - z0 : -1;
        ly[c] = wrap_index(lo[1] + c, n, w, wr) && (unsigned)(w - y0) < (unsigned)ey ? w - y0 : -1;
        lx[c] = wrap_index(lo[2] + c, n, w, wr) && (unsigned)(w - x0) < (unsigned)ex ? w - x0 : -1;
      }
      int t[8];
      unsigned pend = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int b0 = q >> 2, b1 = (q >> 1) & 1, b2 = q & 1;
        const bool in = live && (lz[b0] | ly[b1] | lx[b2]) >= 0;
        t[q] = in ? (lz[b0] * ey + ly[b1]) * ex + lx[b2] : 0;
        pend |= (unsigned)in << q;
      }
      // claim word of the lower corner, per axis in [0, e]: l0 + 1 when the lower corner is in
      // the brick, else l1. Lanes sharing a target for one corner bit along an axis share the
      // word (index -1 wraps to N - 1 and 0, which lie in one brick only when a single brick
      // spans the axis; every lane with that lower corner then has the word N)
      const int kz = lz[0] >= 0 ? lz[0] + 1 : lz[1], ky = ly[0] >= 0 ? ly[0] + 1 : ly[1],
                kx = lx[0] >= 0 ? lx[0] + 1 : lx[1];
      const int key2 = (kz * (ey + 1) + ky) * (ex + 1) + kx;
      bool todo = pend != 0u;
      while (__ballot(todo)) {
        if (todo) __hip_atomic_store(wclaim + key2, lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        const bool won = todo && __hip_atomic_load(wclaim + key2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) == lane;
        if (won) {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            if ((pend >> q) & 1u) {
              const int b0 = q >> 2, b1 = (q >> 1) & 1, b2 = q & 1;
              const float wq = (b0 ? fr[0] : 1.f - fr[0]) * (b1 ? fr[1] : 1.f - fr[1]) * (b2 ? fr[2] : 1.f - fr[2]);
              float2 a = wacc[t[q]];
              a.x = fmaf(wq, gv[b].x, a.x);
              a.y = fmaf(wq, gv[b].y, a.y);
              wacc[t[q]] = a;
            }
          }
        }
        todo = todo && !won;
      }
    }
  }
  __syncthreads();
  for (int v = tid; v < vox; v += 64 * W) {
    const int ix = v % ex, iy = (v / ex) % ey, iz = v / (ex * ey);
    if (x0 + ix < n && y0 + iy < n && z0 + iz < n) {
      float2 a = bin_lds[v];
#pragma unroll
      for (int w = 1; w < W; ++w) {
        const float2 o = bin_lds[w * vox + v];
        a.x += o.x;
        a.y += o.y;
      }
      float2* o = ggrid + ((size_t)(z0 + iz) * n + (y0 + iy)) * n + (x0 + ix);
      if (SEG) {
        unsafeAtomicAdd(&o->x, a.x);
        unsafeAtomicAdd(&o->y, a.y);
      } else {
        const float2 prev = *o;
        *o = make_float2(prev.x + a.x, prev.y + a.y);
      }
    }
  }
  __syncthreads();  // the LDS is reused for the next brick
  }
}

// LDS bytes of a bin_pull_kernel<W> workgroup for bricks of vox voxels (kvox claim words per wave).
static size_t pull_lds(int w, int vox, int kvox) { return (size_t)w * ((size_t)vox * 8 + (size_t)kvox * 4); }

}  // namespace
}  // namespace cryomi

using namespace cryomi;

extern "C" {

// Images per binning pass: at most BIN_CHUNK, few enough that the list places of the pass's
// blocks (one word per block and brick) stay within REL_CAP_WORDS, and that a heavy segment's
// index (below 8 nh chunk / BIN_SEG) fits its 19 bits.
static int bin_chunk_images(int box, int n_images) {
  const int nh = box * (box / 2 + 1);
  const size_t per_image = (size_t)sample_parts(nh, GATHER_SPB) * bin_geom(box, true).nbricks;
  const int cap = (int)std::max<size_t>(1, REL_CAP_WORDS / per_image);
  const int seg_cap = (int)(((size_t)1 << 19) * BIN_SEG / (8 * (size_t)nh));
  return std::min({n_images, BIN_CHUNK, cap, seg_cap});
}

// Brick-list entries of one pass: at most 8 per sample (a sample reaches at most 8 bricks).
static size_t bin_list_bytes(int box, int chunk) {
  return 8 * (size_t)chunk * box * (box / 2 + 1) * sizeof(unsigned);
}

// List places per (block, brick) of one pass, then the blocks' touched-brick bitmaps.
static size_t bin_rel_bytes(int box, int chunk) {
  const int nh = box * (box / 2 + 1);
  const int nb = bin_geom(box, true).nbricks;
  return (size_t)chunk * sample_parts(nh, GATHER_SPB) * (nb + bitmap_words(nb)) * sizeof(unsigned);
}

// A second stream per device (default priority) on which the first pass's brick counts,
// which need the rotations only, run beside the two FFT passes; fork and join events order
// it after the caller's earlier work and before the gather. The per-device lock is held while
// a call enqueues, so concurrent host threads do not interleave their fork and join events.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
static std::mutex side_mu[64];
static SideStream side_streams[64];

static SideStream* side_stream(std::unique_lock<std::mutex>& lk) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  lk = std::unique_lock<std::mutex>(side_mu[dev]);
  SideStream& ss = side_streams[dev];
  if (!ss.s) {
    // during a graph capture these calls may be refused: the call then runs on one stream
    if (hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ss.join, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      if (ss.s) (void)hipStreamDestroy(ss.s);
      if (ss.fork) (void)hipEventDestroy(ss.fork);
      ss = SideStream{};
      return nullptr;
    }
  }
  return ss.s ? &ss : nullptr;
}

static int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return CRYOMI_OK;
  set_error(std::string("cryomi_simulate_vjp: ") + what + ": " + hipGetErrorString(e));
  return CRYOMI_ERR_LAUNCH;
}

size_t cryomi_simulate_vjp_workspace_bytes(int box, int n_images) {
  if (!box_supported(box) || n_images <= 0) return 0;
  const size_t per = (size_t)n_images * box * (box / 2 + 1) * sizeof(float2);  // one complex half-plane per image
  if (box > BIN_MAX_BOX) return 2 * per;
  // row transforms (later g_P), column transforms (g_F), then per pass of images the brick
  // lists, the blocks' list places and the counts, offsets and brick lists
  const int chunk = bin_chunk_images(box, n_images);
  return 2 * per + bin_list_bytes(box, chunk) + bin_rel_bytes(box, chunk) +
         (BIN_META_WORDS + heavy_words(bin_list_bytes(box, chunk) / sizeof(unsigned))) * sizeof(unsigned);
}

int cryomi_simulate_vjp(const void* fourier_voxel_grid, int box, int n_images, const cryomi_image_params* params,
                        unsigned flags, const float* cotangent, void* workspace, size_t workspace_bytes,
                        void* grid_grad, double* param_grads, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate_vjp: box must be even, 8 <= box <= 512 and of the form 2^a 3^b 5^c");
  CRYOMI_REQUIRE(n_images >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_vjp: n_images < 0");
  CRYOMI_REQUIRE((flags & ~(CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION | CRYOMI_GATHER_FILL |
                             CRYOMI_GRID_BRICKED)) == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_vjp: flags may only hold CRYOMI_APPLY_CTF, CRYOMI_APPLY_TRANSLATION, "
                 "CRYOMI_GATHER_FILL and CRYOMI_GRID_BRICKED");
  CRYOMI_REQUIRE(!(flags & CRYOMI_GRID_BRICKED) || box % 4 == 0, CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate_vjp: the bricked grid layout needs box % 4 == 0");
  const bool bricked = (flags & CRYOMI_GRID_BRICKED) != 0;
  if (n_images == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(fourier_voxel_grid && params && cotangent && grid_grad && param_grads,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_vjp: NULL pointer");
  CRYOMI_REQUIRE(workspace && workspace_bytes >= cryomi_simulate_vjp_workspace_bytes(box, n_images),
                 CRYOMI_ERR_WORKSPACE, "cryomi_simulate_vjp: workspace smaller than cryomi_simulate_vjp_workspace_bytes()");
  const int n = box, m = n / 2;
  float2* s = static_cast<float2*>(workspace);
  float2* gf = s + (size_t)n_images * n * (m + 1);
  const int nh = n * (m + 1);
  const int rpb = rows_per_block(m);
  const int sw = strip_w(n);
  auto ffts = [&]() -> int {
    if (n == 64 || n == 128) {
      if (n == 64)
        hipLaunchKernelGGL(adj_fft_kernel<64>, dim3((unsigned)n_images), dim3(FFT_NT), 0, stream, cotangent, gf);
      else
        hipLaunchKernelGGL(adj_fft_kernel<128>, dim3((unsigned)n_images), dim3(FFT_NT), 0, stream, cotangent, gf);
      return check_launch("adj_fft_kernel");
    }
    hipLaunchKernelGGL(adj_rows_kernel, dim3((unsigned)(n_images * ((n + rpb - 1) / rpb))), dim3(NT),
                       (size_t)(n + 2 * rpb * m) * sizeof(float2), stream, cotangent, n, n_images, s);
    int st = check_launch("adj_rows_kernel");
    if (st) return st;
    hipLaunchKernelGGL(adj_cols_kernel, dim3((unsigned)(n_images * (m / sw))), dim3(NT),
                       (size_t)(n + 2 * n * (sw + 1)) * sizeof(float2), stream, s, n, n_images, gf);
    return check_launch("adj_cols_kernel");
  };
  int st;
  // g_P per sample goes to the row-transform buffer `s`, which adj_cols_kernel has consumed
  if (n > BIN_MAX_BOX) {
    if ((st = ffts())) return st;
    const int gather_blocks = n_images * sample_parts(nh, GATHER_SPB);
    if (bricked)
      hipLaunchKernelGGL((adj_gather_kernel<false, true>), dim3((unsigned)gather_blocks), dim3(NT), 0, stream,
                         static_cast<const float2*>(fourier_voxel_grid), n, n_images, params, flags, gf, s,
                         param_grads, BinGeom{}, nullptr, nullptr, nullptr, nullptr);
    else
      hipLaunchKernelGGL((adj_gather_kernel<false, false>), dim3((unsigned)gather_blocks), dim3(NT), 0, stream,
                         static_cast<const float2*>(fourier_voxel_grid), n, n_images, params, flags, gf, s,
                         param_grads, BinGeom{}, nullptr, nullptr, nullptr, nullptr);
    if ((st = check_launch("adj_gather_kernel"))) return st;
    const unsigned bricks = (unsigned)(((n + PBX - 1) / PBX) * ((n + PBY - 1) / PBY) * ((n + PBZ - 1) / PBZ));
    hipLaunchKernelGGL(adj_pull_kernel, dim3(bricks), dim3(NT), 0, stream, n, n_images, params,
                       (flags & CRYOMI_GATHER_FILL) ? 0 : 1, s, static_cast<float2*>(grid_grad));
    return check_launch("adj_pull_kernel");
  }
  // per pass of at most `chunk` images: brick lists, list places and counts after the two
  // half-plane regions
  const BinGeom g = bin_geom(n, (flags & CRYOMI_GATHER_FILL) == 0);
  const int chunk = bin_chunk_images(n, n_images);
  const int parts = sample_parts(nh, GATHER_SPB);
  char* tail = static_cast<char*>(workspace) + 2 * (size_t)n_images * nh * sizeof(float2);
  unsigned* entries = reinterpret_cast<unsigned*>(tail);
  unsigned* rel = reinterpret_cast<unsigned*>(tail + bin_list_bytes(n, chunk));
  unsigned* bitmap = rel + (size_t)chunk * parts * g.nbricks;
  unsigned* counts = reinterpret_cast<unsigned*>(tail + bin_list_bytes(n, chunk) + bin_rel_bytes(n, chunk));
  unsigned* offsets = counts + BIN_MAX_BRICKS;
  unsigned* lists = offsets + BIN_MAX_BRICKS + 1;  // light bricks
  unsigned* list_n = lists + BIN_MAX_BRICKS;
  unsigned* heavy = counts + BIN_META_WORDS;       // heavy segments
  const size_t n_heavy_max = heavy_words(bin_list_bytes(n, chunk) / sizeof(unsigned));
  // counts of a pass: bin_zero_kernel clears them once, bin_scan_kernel again for the next pass
  auto count_pass = [&](int c0, hipStream_t on) -> int {
    const int cnt = std::min(chunk, n_images - c0);
    if (c0 == 0) {
      hipLaunchKernelGGL(bin_zero_kernel, dim3((unsigned)((BIN_MAX_BRICKS + NT - 1) / NT)), dim3(NT), 0, on, counts,
                         BIN_MAX_BRICKS);
      int st = check_launch("bin_zero_kernel");
      if (st) return st;
    }
    hipLaunchKernelGGL(bin_count_kernel, dim3((unsigned)(cnt * parts)), dim3(COUNT_NT), 0, on, g, params + c0, parts,
                       counts, rel, bitmap);
    int st = check_launch("bin_count_kernel");
    if (st) return st;
    hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(SCAN_NT), 0, on, g.nbricks, counts, offsets, lists, heavy,
                       list_n);
    return check_launch("bin_scan_kernel");
  };
  {
    std::unique_lock<std::mutex> lk;
    SideStream* side = side_stream(lk);
    if (side) {
      if ((st = hip_check(hipEventRecord(side->fork, stream), "hipEventRecord"))) return st;
      if ((st = hip_check(hipStreamWaitEvent(side->s, side->fork, 0), "hipStreamWaitEvent"))) return st;
      if ((st = count_pass(0, side->s))) return st;
      if ((st = hip_check(hipEventRecord(side->join, side->s), "hipEventRecord"))) return st;
      if ((st = ffts())) return st;
      if ((st = hip_check(hipStreamWaitEvent(stream, side->join, 0), "hipStreamWaitEvent"))) return st;
    } else {
      if ((st = ffts())) return st;
      if ((st = count_pass(0, stream))) return st;
    }
  }
  const int vox = g.e[0] * g.e[1] * g.e[2];
  const int kvox = (g.e[0] + 1) * (g.e[1] + 1) * (g.e[2] + 1);
  for (int c0 = 0; c0 < n_images; c0 += chunk) {
    const int cnt = std::min(chunk, n_images - c0);
    const cryomi_image_params* pc = params + c0;
    float2* gc = s + (size_t)c0 * nh;
    if (c0 > 0 && (st = count_pass(c0, stream))) return st;
    if (bricked)
      hipLaunchKernelGGL((adj_gather_kernel<true, true>), dim3((unsigned)(cnt * parts)), dim3(NT), 0, stream,
                         static_cast<const float2*>(fourier_voxel_grid), n, cnt, pc, flags, gf + (size_t)c0 * nh, gc,
                         param_grads + (size_t)c0 * N_GRAD, g, offsets, rel, bitmap, entries);
    else
      hipLaunchKernelGGL((adj_gather_kernel<true, false>), dim3((unsigned)(cnt * parts)), dim3(NT), 0, stream,
                         static_cast<const float2*>(fourier_voxel_grid), n, cnt, pc, flags, gf + (size_t)c0 * nh, gc,
                         param_grads + (size_t)c0 * N_GRAD, g, offsets, rel, bitmap, entries);
    if ((st = check_launch("adj_gather_kernel"))) return st;
    // light bricks: 4 waves (2 for large bricks) per workgroup, one workgroup per brick; heavy
    // segments: the same workgroups, persistent over the segment list
    float2* gg = static_cast<float2*>(grid_grad);
    const unsigned nb = (unsigned)g.nbricks;
    const unsigned heavy_grid = (unsigned)std::min<size_t>(n_heavy_max, 4 * (size_t)device_cu_count());
    if (vox <= BIN_MAX_VOX / 2) {
      hipLaunchKernelGGL((bin_pull_kernel<4, 1, false>), dim3(nb), dim3(256), pull_lds(4, vox, kvox), stream, g, pc,
                         gc, offsets, entries, lists, list_n, gg);
      if ((st = check_launch("bin_pull_kernel"))) return st;
      hipLaunchKernelGGL((bin_pull_kernel<4, 1, true>), dim3(heavy_grid), dim3(256), pull_lds(4, vox, kvox), stream,
                         g, pc, gc, offsets, entries, heavy, list_n + 1, gg);
    } else {
      hipLaunchKernelGGL((bin_pull_kernel<2, 1, false>), dim3(nb), dim3(128), pull_lds(2, vox, kvox), stream, g, pc,
                         gc, offsets, entries, lists, list_n, gg);
      if ((st = check_launch("bin_pull_kernel"))) return st;
      hipLaunchKernelGGL((bin_pull_kernel<2, 1, true>), dim3(heavy_grid), dim3(128), pull_lds(2, vox, kvox), stream,
                         g, pc, gc, offsets, entries, heavy, list_n + 1, gg);
    }
    if ((st = check_launch("bin_pull_kernel"))) return st;
  }
  return CRYOMI_OK;
}

}  // extern "C"
