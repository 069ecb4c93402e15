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
#include <mutex>
#include <string>

#include "common.hpp"

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

// Per-image constants, derived once per workgroup in FP64.
struct ImageConsts {
  double r00, r01, r10, r11, r20, r21;  // columns 0, 1 of R
  double half;                          // N / 2 (logical offset of zero frequency)
  double ca, cb, cc, cd, cphase;        // CTF phase polynomial in cycles
  double tx, ty;                        // shift phase in cycles per frequency index
  float scale;                          // amplitude (x 1/N^2 for real output)
};

__device__ __forceinline__ ImageConsts make_consts(const cryomi_image_params& p, int n, bool real_out) {
  ImageConsts k;
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

// One corner of the trilinear stencil with JAX fill semantics: indices in
// [-N, -1] wrap to i + N, anything else outside [0, N) reads 0.
__device__ __forceinline__ bool wrap_index(int idx, int n, int& out) {
  int w = idx < 0 ? idx + n : idx;
  out = w;
  return (unsigned)w < (unsigned)n;
}

// Sample (i, j) of the rfft half-plane: i in [0, N) rows (y), j in [0, N/2] cols.
__device__ __forceinline__ float2 slice_sample(const float2* __restrict__ grid, int n, int i, int j,
                                               const ImageConsts& k, unsigned flags) {
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
  const bool vz0 = wrap_index(lz, n, z0), vz1 = wrap_index(lz + 1, n, z1);
  const bool vy0 = wrap_index(ly, n, y0), vy1 = wrap_index(ly + 1, n, y1);
  const bool vx0 = wrap_index(lx, n, x0), vx1 = wrap_index(lx + 1, n, x1);
  const size_t nn = (size_t)n;
  float2 acc = make_float2(0.f, 0.f);
  // corner order follows itertools.product over (z, y, x) lower/upper
#define CRYOMI_CORNER(VZ, ZI, WZ, VY, YI, WY, VX, XI, WX)                  \
  if ((VZ) && (VY) && (VX)) {                                              \
    const float w = (WZ) * (WY) * (WX);                                    \
    const float2 g = grid[((size_t)(ZI) * nn + (size_t)(YI)) * nn + (XI)]; \
    acc.x = fmaf(w, g.x, acc.x);                                           \
    acc.y = fmaf(w, g.y, acc.y);                                           \
  }
  CRYOMI_CORNER(vz0, z0, wz0, vy0, y0, wy0, vx0, x0, wx0)
  CRYOMI_CORNER(vz0, z0, wz0, vy0, y0, wy0, vx1, x1, wx1)
  CRYOMI_CORNER(vz0, z0, wz0, vy1, y1, wy1, vx0, x0, wx0)
  CRYOMI_CORNER(vz0, z0, wz0, vy1, y1, wy1, vx1, x1, wx1)
  CRYOMI_CORNER(vz1, z1, wz1, vy0, y0, wy0, vx0, x0, wx0)
  CRYOMI_CORNER(vz1, z1, wz1, vy0, y0, wy0, vx1, x1, wx1)
  CRYOMI_CORNER(vz1, z1, wz1, vy1, y1, wy1, vx0, x0, wx0)
  CRYOMI_CORNER(vz1, z1, wz1, vy1, y1, wy1, vx1, x1, wx1)
#undef CRYOMI_CORNER
  // self-conjugate bins (N/2, 0), (0, N/2), (N/2, N/2): real part at
  // extraction; zeroed before the shift when translating.
  if ((i == m || i == 0) && (j == 0 || j == m) && !(i == 0 && j == 0)) {
    acc.y = 0.f;
    if (flags & CRYOMI_APPLY_TRANSLATION) acc.x = 0.f;
  }
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
      v = cscale(spec[(size_t)i * (m + 1) + j], ((i + j) & 1) ? -inv : inv);
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
    ImageConsts k;
    const float2* sp = nullptr;
    if (FROM_SPECTRUM) {
      sp = spec + (size_t)img * n * (m + 1);
    } else {
      k = make_consts(params[img], n, true);
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

__global__ __launch_bounds__(kThreads) void slice_fourier_kernel(const float2* __restrict__ grid, int n,
                                                                 int blocks_per_image,
                                                                 const cryomi_image_params* __restrict__ params,
                                                                 unsigned flags, float2* __restrict__ out) {
  const int m = n / 2;
  const int img = blockIdx.x / blocks_per_image;
  const int s = (blockIdx.x - img * blocks_per_image) * kThreads + threadIdx.x;
  const int nh = n * (m + 1);
  if (s >= nh) return;
  const int i = s / (m + 1);
  const int j = s - i * (m + 1);
  const ImageConsts k = make_consts(params[img], n, false);
  out[(size_t)img * nh + s] = slice_sample(grid, n, i, j, k, flags);
}

// ---------------------------------------------------------------------------
// launch configuration
// ---------------------------------------------------------------------------

static int g_wg_per_cu = 4;  // persistent workgroups per CU (tunable)

static int device_cu_count() {
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

static int persistent_grid(int n_images) {
  const int g = device_cu_count() * g_wg_per_cu;
  return n_images < g ? n_images : g;
}

static size_t slot_bytes(int n) { return (size_t)n * (n / 2 + 1) * sizeof(float2); }

template <bool FROM_SPECTRUM>
static int launch_fused(const float2* grid, const float2* spec, int n, int n_images,
                        const cryomi_image_params* params, unsigned flags, float2* scratch, float* out,
                        hipStream_t stream) {
  const int sw = strip_width(n / 2);
  const size_t lds = (size_t)fused_lds_elems(n, sw) * sizeof(float2);
  const int blocks = persistent_grid(n_images);
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

}  // namespace cryomi

using namespace cryomi;

extern "C" {

const char* cryomi_last_error(void) { return g_last_error.c_str(); }

int cryomi_abi_version(void) { return 1; }

size_t cryomi_simulate_workspace_bytes(int box, int n_images, unsigned flags) {
  if (flags & CRYOMI_OUTPUT_FOURIER) return 0;
  if (!box_supported(box) || n_images <= 0) return 0;
  return (size_t)persistent_grid(n_images) * slot_bytes(box);
}

int cryomi_simulate(const void* fourier_voxel_grid, int box, int n_images, const cryomi_image_params* params,
                    unsigned flags, void* workspace, size_t workspace_bytes, void* out, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate: box must be even, 8 <= box <= 512 and of the form 2^a 3^b");
  CRYOMI_REQUIRE(n_images >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate: n_images < 0");
  CRYOMI_REQUIRE((flags & ~(CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION | CRYOMI_OUTPUT_FOURIER)) == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate: unknown flag bits");
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
  return launch_fused<false>(grid, nullptr, box, n_images, params, flags, static_cast<float2*>(workspace),
                             static_cast<float*>(out), stream);
}

size_t cryomi_irfft2_workspace_bytes(int box, int n) {
  if (!box_supported(box) || n <= 0) return 0;
  return (size_t)persistent_grid(n) * slot_bytes(box);
}

int cryomi_irfft2(const void* in, int box, int n, void* workspace, size_t workspace_bytes, void* out,
                  ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_irfft2: box must be even, 8 <= box <= 512 and of the form 2^a 3^b");
  CRYOMI_REQUIRE(n >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_irfft2: n < 0");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(in && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_irfft2: NULL input or output pointer");
  const size_t need = cryomi_irfft2_workspace_bytes(box, n);
  CRYOMI_REQUIRE(workspace && workspace_bytes >= need, CRYOMI_ERR_WORKSPACE,
                 "cryomi_irfft2: workspace smaller than cryomi_irfft2_workspace_bytes()");
  return launch_fused<true>(nullptr, static_cast<const float2*>(in), box, n, nullptr, 0u,
                            static_cast<float2*>(workspace), static_cast<float*>(out), stream);
}

}  // extern "C"
