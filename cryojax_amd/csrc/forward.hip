// Fused Fourier-slice + CTF + shift + irfft2 forward for gfx950.
//
// Kernels
//   fused_image_kernel : persistent; one workgroup forms whole images.
//                        Phase A, per strip of `sw` rfft columns (strip 0 also
//                        owns the Nyquist column): trilinear gather (FP64
//                        coordinates), self-conjugate fix-ups, CTF (FP64 phase,
//                        reduced before sin), shift phase, fftshift sign
//                        (-1)^(i+j) and 1/N^2, then the inverse FFT along y in
//                        LDS, stored to the workgroup's scratch slot (stays in
//                        L2 / Infinity Cache). Phase B, per group of rows: C2R
//                        along x via an N/2-point complex inverse FFT
//                        (real-pair packing), float32 rows to the output.
//   slice_fourier_kernel : outputs_real_space=False path; writes the rfft
//                        half-plane (zero frequency in the corner).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <atomic>
#include <mutex>
#include <string>

#include "common.hpp"
#include "dft.hpp"
#include "order.hpp"
#include "regfft.hpp"
#include "sample.hpp"

namespace cryomi {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return CRYOMI_ERR_LAUNCH;
  }
  return CRYOMI_OK;
}

constexpr int kThreads = 256;
constexpr int kRowsPerGroup = 8;

__host__ __device__ inline int strip_width(int m) {
  for (int w : {8, 4, 2, 1})
    if (m % w == 0) return w;
  return 1;
}

// LDS floats2 needed by one workgroup of the fused kernel for box n.
__host__ __device__ inline int fused_lds_elems(int n, int sw) {
  const int m = n / 2;
  const int a = 2 * n * (sw + 1);       // strip ping-pong buffers
  const int b = 2 * kRowsPerGroup * m;  // row ping-pong buffers
  return n + (a > b ? a : b);           // + twiddles exp(2 pi i m / n)
}

// Phase A for one strip of one image: fill columns [strip*sw, strip*sw+sw)
// (+ the Nyquist column for strip 0) with fftshift-signed, 1/N^2-scaled
// samples, inverse-FFT them along y in LDS, store the result into the image's
// scratch rows (row stride m + 1).
template <bool FROM_SPECTRUM>
__device__ __forceinline__ void strip_phase(const float2* __restrict__ grid, const float2* __restrict__ spec,
                                            const ImageConsts& k, unsigned flags, int n, int strip, int sw,
                                            const float2* tw, float2* bufa, float2* bufb,
                                            float2* __restrict__ scratch, int tid) {
  const int m = n / 2;
  const int ncols = sw + (strip == 0 ? 1 : 0);
  const int total = n * ncols;
  const float inv = 1.0f / ((float)n * (float)n);
  for (int s = tid; s < total; s += kThreads) {
    const int c = s % ncols;
    const int i = s / ncols;
    const int j = (c < sw) ? strip * sw + c : m;
    float2 v;
    if (FROM_SPECTRUM) {
      const float2 x = spec[(size_t)i * (m + 1) + j];
      v = k.finish ? finish_sample(x, n, i, j, k, flags) : cscale(x, inv);
      if ((i + j) & 1) v = make_float2(-v.x, -v.y);  // fftshift of the output image
    } else {
      v = slice_sample(grid, n, i, j, k, flags);
      if ((i + j) & 1) v = make_float2(-v.x, -v.y);  // fftshift of the output image
    }
    bufa[i * ncols + c] = v;
  }
  __syncthreads();
  const float2* res = stockham_inverse(bufa, bufb, n, ncols, ncols, 1, tw, 1, tid, kThreads);
  for (int s = tid; s < total; s += kThreads) {
    const int c = s % ncols;
    const int y = s / ncols;
    const int j = (c < sw) ? strip * sw + c : m;
    scratch[(size_t)y * (m + 1) + j] = res[y * ncols + c];
  }
  __syncthreads();  // LDS buffers are reused by the next strip
}

// Phase B for rows [row0, row0 + kRowsPerGroup) of one image: C2R along x of
// each scratch row via an m-point complex inverse FFT of
// Z[k] = (X[k] + X*[m-k]) + i e^{2 pi i k/N} (X[k] - X*[m-k]), where the
// imaginary parts of X[0] and X[m] are dropped (C2R semantics); then
// out[2t] = Re z[t], out[2t+1] = Im z[t].
__device__ __forceinline__ void row_phase(const float2* __restrict__ scratch, int n, int row0,
                                          const float2* tw, float2* bufa, float2* bufb,
                                          float* __restrict__ out, int tid) {
  const int m = n / 2;
  const int nrows = (n - row0) < kRowsPerGroup ? (n - row0) : kRowsPerGroup;
  for (int t = tid; t < nrows * m; t += kThreads) {
    const int r = t / m;
    const int kk = t - r * m;
    const float2* x = scratch + (size_t)(row0 + r) * (m + 1);
    float2 a = x[kk];
    float2 b = x[m - kk];
    if (kk == 0) {
      a.y = 0.f;
      b.y = 0.f;
    }
    const float2 bc = cconj(b);
    const float2 s = cadd(a, bc);
    const float2 wd = cmul(tw[kk], csub(a, bc));
    bufa[r * m + kk] = make_float2(s.x - wd.y, s.y + wd.x);
  }
  __syncthreads();
  const float2* res = stockham_inverse(bufa, bufb, m, kRowsPerGroup, 1, m, tw, 2, tid, kThreads);
  for (int t = tid; t < nrows * m; t += kThreads) {
    const int r = t / m;
    const int mm = t - r * m;
    reinterpret_cast<float2*>(out + (size_t)(row0 + r) * n)[mm] = res[r * m + mm];
  }
  __syncthreads();
}

// Persistent fused forward: workgroup b handles images b, b + G, b + 2G, ...
// (G = gridDim.x), each entirely (all strips, then all rows), using its own
// scratch slot of n * (n/2 + 1) complex values. Every workgroup runs the same
// number of barriers per image and exits after its last image.
template <bool FROM_SPECTRUM>
__global__ __launch_bounds__(kThreads) void fused_image_kernel(const float2* __restrict__ grid,
                                                               const float2* __restrict__ spec, int n,
                                                               int n_images,
                                                               const cryomi_image_params* __restrict__ params,
                                                               unsigned flags, float2* __restrict__ scratch,
                                                               float* __restrict__ out) {
  extern __shared__ float2 smem[];
  const int m = n / 2;
  const int tid = threadIdx.x;
  const int sw = strip_width(m);
  const int nstrips = m / sw;
  float2* tw = smem;
  float2* bufa = tw + n;
  float2* bufb_strip = bufa + n * (sw + 1);
  float2* bufb_row = bufa + kRowsPerGroup * m;
  fill_twiddles(tw, n, tid, kThreads);
  __syncthreads();
  float2* slot = scratch + (size_t)blockIdx.x * n * (m + 1);
  for (int img = blockIdx.x; img < n_images; img += gridDim.x) {
    ImageConsts k{};
    const float2* sp = nullptr;
    if (FROM_SPECTRUM) {
      sp = spec + (size_t)img * n * (m + 1);
      if (params) {  // a spectrum given per image, with its fix-ups, CTF and shift still to apply
        k = make_consts(params[img], n, true, flags);
        k.finish = true;
      }
    } else {
      k = make_consts(params[img], n, true, flags);
    }
    for (int strip = 0; strip < nstrips; ++strip)
      strip_phase<FROM_SPECTRUM>(grid, sp, k, flags, n, strip, sw, tw, bufa, bufb_strip, slot, tid);
    __threadfence_block();
    __syncthreads();
    float* dst = out + (size_t)img * n * n;
    for (int row0 = 0; row0 < n; row0 += kRowsPerGroup) row_phase(slot, n, row0, tw, bufa, bufb_row, dst, tid);
    __threadfence_block();
    __syncthreads();  // the slot is rewritten by the next image
  }
}

// Writes the bricked layout: brick b of row (z, y) holds x = 3b - 1 + e,
// e = 0..3, with the grid's own wrap-around at x = -1 (JAX normalises indices
// in [-N, -1] to i + N) and the fill value 0 for x >= N.
__global__ __launch_bounds__(kThreads) void to_bricks_kernel(const float2* __restrict__ in, int n,
                                                             float2* __restrict__ out) {
  const size_t total = bricked_elems(n);
  const int bx = brick_row_bricks(n);
  for (size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x; e < total; e += (size_t)gridDim.x * kThreads) {
    const int in_brick = (int)(e & 15);
    const size_t brick = e >> 4;
    const int b = (int)(brick % (size_t)bx);
    const size_t zy = brick / (size_t)bx;
    const int y = 2 * (int)(zy % (size_t)(n / 2)) + ((in_brick >> 2) & 1);
    const int z = 2 * (int)(zy / (size_t)(n / 2)) + (in_brick >> 3);
    const int x = 3 * b - 1 + (in_brick & 3);
    const int xw = x < 0 ? x + n : x;
    out[e] = (xw < n) ? in[((size_t)z * n + y) * n + xw] : make_float2(0.f, 0.f);
  }
}

__global__ __launch_bounds__(kThreads) void slice_fourier_kernel(const float2* __restrict__ grid, int n,
                                                                 int blocks_per_image,
                                                                 const cryomi_image_params* __restrict__ params,
                                                                 unsigned flags, float2* __restrict__ out) {
  const int m = n / 2;
  const int img = blockIdx.x / blocks_per_image;
  const int s = (blockIdx.x - img * blocks_per_image) * kThreads + threadIdx.x;
  const int nh = n * (m + 1);
  __shared__ ImageConsts kc;  // the image's constants, formed once per workgroup
  if (threadIdx.x == 0) kc = make_consts(params[img], n, false, flags);
  __syncthreads();
  if (s >= nh) return;
  const int i = s / (m + 1);
  const int j = s - i * (m + 1);
  const ImageConsts k = kc;
  out[(size_t)img * nh + s] = slice_sample(grid, n, i, j, k, flags);
}

// The low-frequency band of the Fourier output: rows ky = -band/2 .. band/2 (band + 1 rows, row
// r holds ky = r - band/2) and columns kx = 0 .. band/2 of the box-size rfft half-plane, each sample
// computed exactly as slice_fourier_kernel computes it. These are the frequencies a Fourier crop of
// the box image to band x band reads (the +band/2 row for the mirrors of its -band/2 edge).
__global__ __launch_bounds__(kThreads) void slice_band_kernel(const float2* __restrict__ grid, int n, int band,
                                                              int blocks_per_image,
                                                              const cryomi_image_params* __restrict__ params,
                                                              unsigned flags, float2* __restrict__ out) {
  const int mo = band / 2, w = mo + 1;
  const int nb = (band + 1) * w;
  const int img = blockIdx.x / blocks_per_image;
  const int s = (blockIdx.x - img * blocks_per_image) * kThreads + threadIdx.x;
  __shared__ ImageConsts kc;  // the image's constants, formed once per workgroup
  if (threadIdx.x == 0) kc = make_consts(params[img], n, false, flags);
  __syncthreads();
  if (s >= nb) return;
  const int r = s / w;
  const int j = s - r * w;
  const int ky = r - mo;
  const ImageConsts k = kc;
  out[(size_t)img * nb + s] = slice_sample(grid, n, ky >= 0 ? ky : ky + n, j, k, flags);
}

// Spectrum given per image: fix-ups, CTF and shift of every rfft sample, rfft layout out.
__global__ __launch_bounds__(kThreads) void spectrum_finish_kernel(const float2* __restrict__ spec, int n,
                                                                   int blocks_per_image,
                                                                   const cryomi_image_params* __restrict__ params,
                                                                   unsigned flags, float2* __restrict__ out) {
  const int m = n / 2;
  const int img = blockIdx.x / blocks_per_image;
  const int s = (blockIdx.x - img * blocks_per_image) * kThreads + threadIdx.x;
  const int nh = n * (m + 1);
  __shared__ ImageConsts kc;  // the image's constants, formed once per workgroup
  if (threadIdx.x == 0) kc = make_consts(params[img], n, false, flags);
  __syncthreads();
  if (s >= nh) return;
  const int i = s / (m + 1);
  const int j = s - i * (m + 1);
  const ImageConsts k = kc;
  out[(size_t)img * nh + s] = finish_sample(spec[(size_t)img * nh + s], n, i, j, k, flags);
}

// ---------------------------------------------------------------------------
// register-FFT image kernel for N in {64, 128, 160, 192, 256} (plan in regfft.hpp)
// ---------------------------------------------------------------------------

// Persistent: workgroup b forms images b, b + G, b + 2G, ... (G = gridDim.x)
// with one workgroup of N threads per image.
// NTH threads per image: N, or 256 at box 64 for small batches and spectra (reg_wide).
template <int N, bool FROM_SPECTRUM, bool BRICKED, int NTH = N>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(3, 8))) void reg_image_kernel(const float2* __restrict__ grid,
                                                                 const float2* __restrict__ spec, int n_images,
                                                                 const cryomi_image_params* __restrict__ params,
                                                                 unsigned flags, float2* __restrict__ scratch,
                                                                 const int* __restrict__ order,
                                                                 float* __restrict__ out) {
  constexpr int M = N / 2;
  extern __shared__ float2 smem[];
  float2* tw = smem;  // exp(+2 pi i m / N)
  float2* lds = tw + N;
  const int tid = threadIdx.x;
  fill_twiddles(tw, N, tid, NTH);
  __syncthreads();
  float2* slot = scratch + (size_t)blockIdx.x * N * (M + 1);
#pragma unroll 1
  for (int pos = blockIdx.x; pos < n_images; pos += gridDim.x) {
    const int img = order ? order[pos] : pos;  // image formed at this schedule position
    ImageConsts k{};
    const float2* sp = nullptr;
    if (FROM_SPECTRUM) {
      sp = spec + (size_t)img * N * (M + 1);
      if (params) {  // a spectrum given per image, with its fix-ups, CTF and shift still to apply
        k = make_consts(params[img], N, true, flags);
        k.finish = true;
      }
    } else {
      k = make_consts(params[img], N, true, flags);
    }
    reg_image<N, NTH, FROM_SPECTRUM, BRICKED>(grid, sp, k, flags, tw, lds, slot,
                                              reinterpret_cast<float2*>(out + (size_t)img * N * N), tid);
  }
}

// ---------------------------------------------------------------------------
// image schedule: slices with nearby planes on the same XCD (order.hpp)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kOrderThreads) void order_kernel(const cryomi_image_params* __restrict__ params,
                                                              int n_images, int* __restrict__ order) {
  order_chunk(params, n_images, blockIdx.x * kOrderChunk, order);
}

// ---------------------------------------------------------------------------
// launch configuration
// ---------------------------------------------------------------------------

// Tuning values, set through cryomi_set_tuning (the autotuner sets them):
// persistent workgroups per CU (0 = as many as the kernel's occupancy allows,
// queried once per kernel) for every box and per box (-1 = unset), and the
// grouped kernel's workgroup size and workgroups per CU.
static std::atomic<int> g_wg_per_cu{0};
static std::atomic<int> g_box_wg_per_cu[513];
static std::atomic<int> g_groups_threads{256};
static std::atomic<int> g_groups_wg_per_cu{0};
static std::once_flag g_tuning_once;

static void init_tuning() {
  std::call_once(g_tuning_once, [] {
    for (auto& v : g_box_wg_per_cu) v.store(-1);
  });
}

int tuned_wg_per_cu(int box) {
  init_tuning();
  if (box >= 0 && box <= 512) {
    const int v = g_box_wg_per_cu[box].load();
    if (v >= 0) return v;
  }
  return g_wg_per_cu.load();
}
int tuned_groups_threads() { return g_groups_threads.load(); }
int tuned_groups_wg_per_cu() { return g_groups_wg_per_cu.load(); }

int device_cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}

static size_t slot_bytes(int n) { return (size_t)n * (n / 2 + 1) * sizeof(float2); }

// Resident workgroups per CU for `kernel` at this block size and LDS request.
static int occupancy(const void* kernel, int threads, size_t lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, lds) != hipSuccess || nb <= 0) nb = 1;
  return nb;
}

template <int N, bool FROM_SPECTRUM, bool BRICKED, int NTH = N>
static int reg_wg_per_cu() {
  static int cached = 0;
  if (cached == 0)
    cached = occupancy((const void*)reg_image_kernel<N, FROM_SPECTRUM, BRICKED, NTH>, NTH,
                       (size_t)(N + RegPlan<N, NTH>::LDS) * sizeof(float2));
  return cached;
}

// Box 64 with one 64-thread wave per image leaves each image's barrier-separated passes without
// latency hiding when the batch puts few waves on a CU (512 images: two): four waves per image
// take the half-plane in one column group and one row group (512 images: 43 -> 20 us; spectra of
// 512 / 4096 images: 25 -> 12 / 90 -> 68 us). Large forward batches keep one wave per image
// (4096 images: 120 us, against 130 with four).
static bool reg_wide(int n, int n_images, bool from_spectrum) {
  return n == 64 && (from_spectrum || n_images < 8 * device_cu_count());
}

// Persistent grid size for a box / batch: the same value sizes the workspace
// and the launch.
template <int N>
static int reg_wg_per_cu_any(bool from_spectrum, bool bricked, bool wide) {
  if constexpr (N == 64) {
    if (wide) {
      if (from_spectrum) return reg_wg_per_cu<N, true, false, 256>();
      return bricked ? reg_wg_per_cu<N, false, true, 256>() : reg_wg_per_cu<N, false, false, 256>();
    }
  }
  if (from_spectrum) return reg_wg_per_cu<N, true, false>();
  return bricked ? reg_wg_per_cu<N, false, true>() : reg_wg_per_cu<N, false, false>();
}

static int grid_blocks(int n, int n_images, bool from_spectrum, bool bricked) {
  int per_cu = tuned_wg_per_cu(n);
  if (per_cu <= 0) {
    const bool wide = reg_wide(n, n_images, from_spectrum);
    switch (n) {
      case 64: per_cu = reg_wg_per_cu_any<64>(from_spectrum, bricked, wide); break;
      case 128: per_cu = reg_wg_per_cu_any<128>(from_spectrum, bricked, wide); break;
      case 192: per_cu = reg_wg_per_cu_any<192>(from_spectrum, bricked, wide); break;
      case 256: per_cu = reg_wg_per_cu_any<256>(from_spectrum, bricked, wide); break;
      case 160: per_cu = reg_wg_per_cu_any<160>(from_spectrum, bricked, wide); break;
      default: per_cu = 4; break;
    }
  }
  const long long g = (long long)device_cu_count() * per_cu;
  return n_images < g ? n_images : (int)g;
}

// max_blocks > 0 caps the persistent grid (the scratch then holds that many slots).
template <int N, bool FROM_SPECTRUM, int NTH = N>
static int launch_reg(const float2* grid, const float2* spec, int n_images, const cryomi_image_params* params,
                      unsigned flags, float2* scratch, int* order, float* out, hipStream_t stream,
                      int max_blocks) {
  if constexpr (N == 64 && NTH == N) {
    if (reg_wide(N, n_images, FROM_SPECTRUM))
      return launch_reg<N, FROM_SPECTRUM, 256>(grid, spec, n_images, params, flags, scratch, order, out, stream,
                                               max_blocks);
  }
  const size_t lds = (size_t)(N + RegPlan<N, NTH>::LDS) * sizeof(float2);
  const bool bricked = !FROM_SPECTRUM && (flags & CRYOMI_GRID_BRICKED);
  int blocks = grid_blocks(N, n_images, FROM_SPECTRUM, bricked);
  if (max_blocks > 0 && blocks > max_blocks) blocks = max_blocks;
  if (order) {
    hipLaunchKernelGGL(order_kernel, dim3((unsigned)((n_images + kOrderChunk - 1) / kOrderChunk)),
                       dim3(kOrderThreads), 0, stream, params, n_images, order);
    const int st = check_launch("order_kernel");
    if (st) return st;
  }
  if (FROM_SPECTRUM || !bricked) {
    hipLaunchKernelGGL((reg_image_kernel<N, FROM_SPECTRUM, false, NTH>), dim3((unsigned)blocks), dim3(NTH), lds,
                       stream, grid, spec, n_images, params, flags, scratch, order, out);
  } else {
    hipLaunchKernelGGL((reg_image_kernel<N, false, true, NTH>), dim3((unsigned)blocks), dim3(NTH), lds, stream, grid,
                       spec, n_images, params, flags, scratch, order, out);
  }
  return check_launch("reg_image_kernel");
}

template <bool FROM_SPECTRUM>
static int launch_fused(const float2* grid, const float2* spec, int n, int n_images,
                        const cryomi_image_params* params, unsigned flags, float2* scratch, int* order, float* out,
                        hipStream_t stream, int max_blocks = 0) {
  switch (n) {
    case 64:
      return launch_reg<64, FROM_SPECTRUM>(grid, spec, n_images, params, flags, scratch, order, out, stream,
                                           max_blocks);
    case 128:
      return launch_reg<128, FROM_SPECTRUM>(grid, spec, n_images, params, flags, scratch, order, out, stream,
                                            max_blocks);
    case 192:
      return launch_reg<192, FROM_SPECTRUM>(grid, spec, n_images, params, flags, scratch, order, out, stream,
                                            max_blocks);
    case 256:
      return launch_reg<256, FROM_SPECTRUM>(grid, spec, n_images, params, flags, scratch, order, out, stream,
                                            max_blocks);
    case 160:
      return launch_reg<160, FROM_SPECTRUM>(grid, spec, n_images, params, flags, scratch, order, out, stream,
                                            max_blocks);
    default:
      break;
  }
  const int sw = strip_width(n / 2);
  const size_t lds = (size_t)fused_lds_elems(n, sw) * sizeof(float2);
  int blocks = grid_blocks(n, n_images, FROM_SPECTRUM, (flags & CRYOMI_GRID_BRICKED) != 0);
  if (max_blocks > 0 && blocks > max_blocks) blocks = max_blocks;
  if (lds > (size_t)65536) {
    hipError_t e = hipFuncSetAttribute((const void*)fused_image_kernel<FROM_SPECTRUM>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      set_error(std::string("fused_image_kernel: LDS request too large: ") + hipGetErrorString(e));
      return CRYOMI_ERR_UNSUPPORTED;
    }
  }
  hipLaunchKernelGGL(fused_image_kernel<FROM_SPECTRUM>, dim3((unsigned)blocks), dim3(kThreads), lds, stream, grid,
                     spec, n, n_images, params, flags, scratch, out);
  return check_launch("fused_image_kernel");
}

// Real-space images of cryomi_simulate formed with a caller-given scratch region instead of the
// workspace layout: the persistent grid is capped to the slots the region holds and the images
// are formed in their natural order (no schedule table). The images are those of cryomi_simulate:
// every image is formed by one workgroup with the same arithmetic whatever the grid size.
// Arguments are validated by the caller.
int simulate_in_scratch(const float2* grid, int box, int n_images, const cryomi_image_params* params, unsigned flags,
                        void* scratch, size_t scratch_bytes, float* out, hipStream_t stream) {
  const size_t slot = slot_bytes(box);
  const size_t slots = scratch_bytes / slot;
  if (slots < 1) {
    set_error("cryomi_simulate_noalloc: one scratch slot of " + std::to_string(slot) + " bytes (box " +
              std::to_string(box) + ") exceeds the " + std::to_string(scratch_bytes) + "-byte static scratch");
    return CRYOMI_ERR_WORKSPACE;
  }
  const int cap = slots > (size_t)(1 << 30) ? (1 << 30) : (int)slots;
  return launch_fused<false>(grid, nullptr, box, n_images, params, flags, static_cast<float2*>(scratch), nullptr,
                             out, stream, cap);
}

}  // namespace cryomi

using namespace cryomi;

extern "C" {

const char* cryomi_last_error(void) { return g_last_error.c_str(); }

int cryomi_abi_version(void) { return 1; }

int cryomi_set_tuning(const char* key, int value) {
  clear_error();
  CRYOMI_REQUIRE(key != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_set_tuning: NULL key");
  init_tuning();
  const std::string k(key);
  const std::string per_box = "persistent_wg_per_cu.";
  if (k == "persistent_wg_per_cu" || k == "groups.wg_per_cu") {
    CRYOMI_REQUIRE(value >= 0 && value <= 64, CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_set_tuning: workgroups per CU must be in [0, 64]");
    (k == "groups.wg_per_cu" ? g_groups_wg_per_cu : g_wg_per_cu).store(value);
    return CRYOMI_OK;
  }
  if (k.compare(0, per_box.size(), per_box) == 0) {
    char* end = nullptr;
    const long box = std::strtol(k.c_str() + per_box.size(), &end, 10);
    CRYOMI_REQUIRE(end && *end == '\0' && box_supported((int)box), CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_set_tuning: '" + k + "' does not name a supported box");
    CRYOMI_REQUIRE(value >= -1 && value <= 64, CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_set_tuning: workgroups per CU must be in [-1, 64]");
    g_box_wg_per_cu[box].store(value);
    return CRYOMI_OK;
  }
  if (k == "groups.threads") {
    CRYOMI_REQUIRE(value == 256 || value == 512, CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_set_tuning: groups.threads must be 256 or 512");
    g_groups_threads.store(value);
    return CRYOMI_OK;
  }
  set_error("cryomi_set_tuning: unknown key '" + k + "'");
  return CRYOMI_ERR_INVALID_ARGUMENT;
}

int cryomi_get_tuning(const char* key) {
  if (!key) return -1;
  init_tuning();
  const std::string k(key);
  const std::string per_box = "persistent_wg_per_cu.";
  if (k == "persistent_wg_per_cu") return g_wg_per_cu.load();
  if (k == "groups.wg_per_cu") return g_groups_wg_per_cu.load();
  if (k == "groups.threads") return g_groups_threads.load();
  if (k.compare(0, per_box.size(), per_box) == 0) {
    char* end = nullptr;
    const long box = std::strtol(k.c_str() + per_box.size(), &end, 10);
    if (end && *end == '\0' && box_supported((int)box)) return g_box_wg_per_cu[box].load();
  }
  return -1;
}

// The plane-ordered schedule pays (a one-block sort per 4096 images) once every workgroup forms
// several images; it exists for the register-FFT boxes.
static bool uses_order(int box, int n_images, unsigned flags) {
  const bool reg = box == 64 || box == 128 || box == 160 || box == 192 || box == 256;
  return reg && n_images >= 2 * grid_blocks(box, n_images, false, (flags & CRYOMI_GRID_BRICKED) != 0);
}

static size_t slots_bytes(int box, int n_images, unsigned flags) {
  return (size_t)grid_blocks(box, n_images, false, (flags & CRYOMI_GRID_BRICKED) != 0) * slot_bytes(box);
}

size_t cryomi_simulate_workspace_bytes(int box, int n_images, unsigned flags) {
  if (flags & CRYOMI_OUTPUT_FOURIER) return 0;
  if (!box_supported(box) || n_images <= 0) return 0;
  return slots_bytes(box, n_images, flags) + (uses_order(box, n_images, flags) ? (size_t)n_images * sizeof(int) : 0);
}

int cryomi_simulate_band(const void* fourier_voxel_grid, int box, int n_images, const cryomi_image_params* params,
                         unsigned flags, int band, void* out, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate_band: box must be even, 8 <= box <= 512 and of the form 2^a 3^b 5^c");
  CRYOMI_REQUIRE(n_images >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_band: n_images < 0");
  CRYOMI_REQUIRE(band >= 2 && band % 2 == 0 && band < box, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_band: band must be even with 2 <= band < box");
  CRYOMI_REQUIRE((flags & ~(CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION | CRYOMI_OUTPUT_FOURIER |
                             CRYOMI_GRID_BRICKED | CRYOMI_GATHER_FILL)) == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_band: unknown flag bits");
  CRYOMI_REQUIRE(!(flags & CRYOMI_GRID_BRICKED) || box % 4 == 0, CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate_band: the bricked grid layout needs box % 4 == 0");
  if (n_images == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(fourier_voxel_grid && params && out, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_band: NULL grid, params or output pointer");
  const int nb = (band + 1) * (band / 2 + 1);
  const int bpi = (nb + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(slice_band_kernel, dim3((unsigned)((long long)bpi * n_images)), dim3(kThreads), 0, stream,
                     static_cast<const float2*>(fourier_voxel_grid), box, band, bpi, params, flags | CRYOMI_OUTPUT_FOURIER,
                     static_cast<float2*>(out));
  return check_launch("slice_band_kernel");
}

int cryomi_simulate(const void* fourier_voxel_grid, int box, int n_images, const cryomi_image_params* params,
                    unsigned flags, void* workspace, size_t workspace_bytes, void* out, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate: box must be even, 8 <= box <= 512 and of the form 2^a 3^b 5^c");
  CRYOMI_REQUIRE(n_images >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate: n_images < 0");
  CRYOMI_REQUIRE((flags & ~(CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION | CRYOMI_OUTPUT_FOURIER |
                             CRYOMI_GRID_BRICKED | CRYOMI_GATHER_FILL)) == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate: unknown flag bits");
  CRYOMI_REQUIRE(!(flags & CRYOMI_GRID_BRICKED) || box % 4 == 0, CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate: the bricked grid layout needs box % 4 == 0");
  if (n_images == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(fourier_voxel_grid && params && out, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate: NULL grid, params or output pointer");
  const float2* grid = static_cast<const float2*>(fourier_voxel_grid);
  const int m = box / 2;
  if (flags & CRYOMI_OUTPUT_FOURIER) {
    const int nh = box * (m + 1);
    const int bpi = (nh + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(slice_fourier_kernel, dim3((unsigned)((long long)bpi * n_images)), dim3(kThreads), 0,
                       stream, grid, box, bpi, params, flags, static_cast<float2*>(out));
    return check_launch("slice_fourier_kernel");
  }
  const size_t need = cryomi_simulate_workspace_bytes(box, n_images, flags);
  CRYOMI_REQUIRE(workspace && workspace_bytes >= need, CRYOMI_ERR_WORKSPACE,
                 "cryomi_simulate: workspace smaller than cryomi_simulate_workspace_bytes()");
  int* order = uses_order(box, n_images, flags)
                   ? reinterpret_cast<int*>(static_cast<char*>(workspace) + slots_bytes(box, n_images, flags))
                   : nullptr;
  return launch_fused<false>(grid, nullptr, box, n_images, params, flags, static_cast<float2*>(workspace), order,
                             static_cast<float*>(out), stream);
}

size_t cryomi_bricked_grid_bytes(int box) {
  if (!box_supported(box) || box % 4 != 0) return 0;
  return bricked_elems(box) * sizeof(float2);
}

int cryomi_grid_to_bricks(const void* fourier_voxel_grid, int box, void* out, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box) && box % 4 == 0, CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_grid_to_bricks: box must be supported and a multiple of 4");
  CRYOMI_REQUIRE(fourier_voxel_grid && out && fourier_voxel_grid != out, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_grid_to_bricks: NULL or aliased pointers");
  const size_t total = bricked_elems(box);
  size_t blocks = (total + kThreads - 1) / kThreads;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(to_bricks_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, stream,
                     static_cast<const float2*>(fourier_voxel_grid), box, static_cast<float2*>(out));
  return check_launch("to_bricks_kernel");
}

size_t cryomi_irfft2_workspace_bytes(int box, int n) {
  if (!box_supported(box) || n <= 0) return 0;
  return (size_t)grid_blocks(box, n, true, false) * slot_bytes(box);
}

int cryomi_irfft2(const void* in, int box, int n, void* workspace, size_t workspace_bytes, void* out,
                  ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_irfft2: box must be even, 8 <= box <= 512 and of the form 2^a 3^b 5^c");
  CRYOMI_REQUIRE(n >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_irfft2: n < 0");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(in && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_irfft2: NULL input or output pointer");
  const size_t need = cryomi_irfft2_workspace_bytes(box, n);
  CRYOMI_REQUIRE(workspace && workspace_bytes >= need, CRYOMI_ERR_WORKSPACE,
                 "cryomi_irfft2: workspace smaller than cryomi_irfft2_workspace_bytes()");
  return launch_fused<true>(nullptr, static_cast<const float2*>(in), box, n, nullptr, 0u,
                            static_cast<float2*>(workspace), nullptr, static_cast<float*>(out), stream);
}

int cryomi_simulate_from_spectrum(const void* spectrum, int box, int n_images, const cryomi_image_params* params,
                                  unsigned flags, void* workspace, size_t workspace_bytes, void* out,
                                  ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate_from_spectrum: box must be even, 8 <= box <= 512 and of the form 2^a 3^b 5^c");
  CRYOMI_REQUIRE(n_images >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_from_spectrum: n_images < 0");
  CRYOMI_REQUIRE((flags & ~(CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION | CRYOMI_OUTPUT_FOURIER)) == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_from_spectrum: flags may only hold CRYOMI_APPLY_CTF, CRYOMI_APPLY_TRANSLATION "
                 "and CRYOMI_OUTPUT_FOURIER");
  if (n_images == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(spectrum && params && out, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_from_spectrum: NULL spectrum, params or output pointer");
  const float2* spec = static_cast<const float2*>(spectrum);
  if (flags & CRYOMI_OUTPUT_FOURIER) {
    const int nh = box * (box / 2 + 1);
    const int bpi = (nh + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(spectrum_finish_kernel, dim3((unsigned)((long long)bpi * n_images)), dim3(kThreads), 0,
                       stream, spec, box, bpi, params, flags, static_cast<float2*>(out));
    return check_launch("spectrum_finish_kernel");
  }
  const size_t need = cryomi_irfft2_workspace_bytes(box, n_images);
  CRYOMI_REQUIRE(workspace && workspace_bytes >= need, CRYOMI_ERR_WORKSPACE,
                 "cryomi_simulate_from_spectrum: workspace smaller than cryomi_irfft2_workspace_bytes()");
  return launch_fused<true>(nullptr, spec, box, n_images, params, flags, static_cast<float2*>(workspace), nullptr,
                            static_cast<float*>(out), stream);
}

}  // extern "C"
