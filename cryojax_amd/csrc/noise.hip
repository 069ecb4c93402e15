// Image normalisation, white-noise spectra, noisy-image assembly and the
// Gaussian log-likelihood (gfx950). One workgroup per image for the
// reductions (FP64 accumulation); elementwise work is grid-strided.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/cryomi_noise.h"
#include "common.hpp"

namespace cryomi {
namespace {

constexpr int NT = 256;

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

__global__ __launch_bounds__(NT) void normalize_kernel(float* __restrict__ img, int npix,
                                                       const unsigned char* __restrict__ region) {
  __shared__ double red[NT / 64];
  float* x = img + (size_t)blockIdx.x * npix;
  double s1 = 0.0, s2 = 0.0, cnt = 0.0;
  for (int p = threadIdx.x; p < npix; p += NT) {
    if (region && !region[p]) continue;
    const double v = x[p];
    s1 += v;
    s2 += v * v;
    cnt += 1.0;
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  cnt = block_sum(cnt, red);
  const double mean = s1 / cnt;
  const double var = fmax(s2 / cnt - mean * mean, 0.0);
  const double inv = 1.0 / sqrt(var);  // jnp.std (population); a flat image gives inf / nan as there
  for (int p = threadIdx.x; p < npix; p += NT) x[p] = (float)(((double)x[p] - mean) * inv);
}

// One workgroup per image: out(y, x) = in(y + h/2 - oh/2, x + w/2 - ow/2), 0 outside the input
// (each axis is a centred crop when it shrinks, a zero pad split as pad_to_shape's when it grows),
// then optionally (x - mean) / std over the output (or its region), FP64 statistics.
__global__ __launch_bounds__(NT) void resize_kernel(const float* __restrict__ in, int h, int w, int oh, int ow,
                                                    int normalize, const unsigned char* __restrict__ region,
                                                    float* __restrict__ out) {
  __shared__ double red[NT / 64];
  const float* x = in + (size_t)blockIdx.x * h * w;
  float* y = out + (size_t)blockIdx.x * oh * ow;
  const int dy = h / 2 - oh / 2, dx = w / 2 - ow / 2;
  const int npix = oh * ow;
  auto value = [&](int p) {
    const int sy = p / ow + dy, sx = p % ow + dx;
    return ((unsigned)sy < (unsigned)h && (unsigned)sx < (unsigned)w) ? x[(size_t)sy * w + sx] : 0.f;
  };
  double mean = 0.0, inv = 1.0;
  if (normalize) {
    double s1 = 0.0, s2 = 0.0, cnt = 0.0;
    for (int p = threadIdx.x; p < npix; p += NT) {
      if (region && !region[p]) continue;
      const double v = value(p);
      s1 += v;
      s2 += v * v;
      cnt += 1.0;
    }
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    cnt = block_sum(cnt, red);
    mean = s1 / cnt;
    inv = 1.0 / sqrt(fmax(s2 / cnt - mean * mean, 0.0));
  }
  for (int p = threadIdx.x; p < npix; p += NT)
    y[p] = normalize ? (float)(((double)value(p) - mean) * inv) : value(p);
}

__global__ __launch_bounds__(NT) void c128_to_c64_kernel(const double2* __restrict__ in, size_t count,
                                                         float2* __restrict__ out) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < count; e += (size_t)gridDim.x * NT)
    out[e] = make_float2((float)in[e].x, (float)in[e].y);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
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

// Four standard normals per Philox block (Box-Muller, double precision).
__device__ __forceinline__ void normals4(unsigned long long counter, unsigned long long seed, double z[4]) {
  const uint4 r = philox(make_uint4((unsigned)counter, (unsigned)(counter >> 32), 0x6e6f6973u, 0u),
                         make_uint2((unsigned)seed, (unsigned)(seed >> 32)));
  const double u1 = ((double)r.x + 1.0) * (1.0 / 4294967296.0), u2 = (double)r.y * (1.0 / 4294967296.0);
  const double u3 = ((double)r.z + 1.0) * (1.0 / 4294967296.0), u4 = (double)r.w * (1.0 / 4294967296.0);
  const double ra = sqrt(-2.0 * log(u1)), rb = sqrt(-2.0 * log(u3));
  double sa, ca, sb, cb;
  sincospi(2.0 * u2, &sa, &ca);
  sincospi(2.0 * u4, &sb, &cb);
  z[0] = ra * ca;
  z[1] = ra * sa;
  z[2] = rb * cb;
  z[3] = rb * sb;
}

__global__ __launch_bounds__(NT) void white_noise_kernel(size_t total, int nh, double std, unsigned long long seed,
                                                         float2* __restrict__ out) {
  const size_t nquad = (total + 3) / 4;
  for (size_t q = (size_t)blockIdx.x * NT + threadIdx.x; q < nquad; q += (size_t)gridDim.x * NT) {
    double z[4];
    normals4(q, seed, z);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t e = 4 * q + u;
      if (e >= total) break;
      // zero frequency of every image: .at[0, 0].set(0.0)
      const float v = (e % (size_t)nh == 0) ? 0.f : (float)(std * z[u]);
      out[e] = make_float2(v, 0.f);
    }
  }
}

__global__ __launch_bounds__(NT) void noisy_kernel(const float* __restrict__ s, const float* __restrict__ nz,
                                                   size_t count, float scale, float offset, float* __restrict__ out) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < count; e += (size_t)gridDim.x * NT)
    out[e] = fmaf(scale, s[e], offset) + (nz ? nz[e] : 0.f);
}

__global__ __launch_bounds__(NT) void loglik_kernel(const float* __restrict__ sim, const float* __restrict__ obs,
                                                    int npix, double variance, double* __restrict__ out) {
  __shared__ double red[NT / 64];
  const size_t base = (size_t)blockIdx.x * npix;
  double acc = 0.0;
  for (int p = threadIdx.x; p < npix; p += NT) {
    const double r = (double)sim[base + p] - (double)obs[base + p];
    acc += r * r;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) out[blockIdx.x] = -(acc / (2.0 * variance) + 0.5 * npix * log(2.0 * M_PI * variance));
}

// data[b][p] *= factor[p], b < n, p < npix. T: the data element (float, float2, double, double2);
// F: the factor element (real or complex of the same precision)
template <typename T> struct Scalar { using type = T; };
template <> struct Scalar<float2> { using type = float; };
template <> struct Scalar<double2> { using type = double; };
template <typename T> __device__ __forceinline__ T mul_elem(T a, typename Scalar<T>::type f) { return a * f; }
__device__ __forceinline__ float2 mul_elem(float2 a, float f) { return make_float2(a.x * f, a.y * f); }
__device__ __forceinline__ double2 mul_elem(double2 a, double f) { return make_double2(a.x * f, a.y * f); }
__device__ __forceinline__ float2 mul_elem(float2 a, float2 f) {
  return make_float2(a.x * f.x - a.y * f.y, a.x * f.y + a.y * f.x);
}
__device__ __forceinline__ double2 mul_elem(double2 a, double2 f) {
  return make_double2(a.x * f.x - a.y * f.y, a.x * f.y + a.y * f.x);
}

template <typename T, typename F>
__global__ __launch_bounds__(NT) void multiply_broadcast_kernel(T* __restrict__ data, size_t total, long long npix,
                                                                const F* __restrict__ factor) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < total; e += (size_t)gridDim.x * NT)
    data[e] = mul_elem(data[e], factor[(long long)(e % (size_t)npix)]);
}

unsigned grid_for(size_t work) {
  size_t b = (work + NT - 1) / NT;
  if (b > 65536) b = 65536;
  return (unsigned)(b ? b : 1);
}

}  // namespace
}  // namespace cryomi

using namespace cryomi;

extern "C" {

int cryomi_normalize_images(float* images, int n, int npix, const unsigned char* region, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && npix > 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_normalize_images: need n >= 0, npix > 0");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(images != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_normalize_images: NULL images");
  hipLaunchKernelGGL(normalize_kernel, dim3((unsigned)n), dim3(NT), 0, stream, images, npix, region);
  return check_launch("normalize_kernel");
}

int cryomi_resize_images(const float* in, int n, int h, int w, int oh, int ow, unsigned flags,
                         const unsigned char* region, float* out, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && h > 0 && w > 0 && oh > 0 && ow > 0 && (long long)oh * ow < (1ll << 31),
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_resize_images: need n >= 0 and positive sizes");
  CRYOMI_REQUIRE((flags & ~CRYOMI_RESIZE_NORMALIZE) == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_resize_images: unknown flag bits");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(in && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_resize_images: NULL input or output");
  hipLaunchKernelGGL(resize_kernel, dim3((unsigned)n), dim3(NT), 0, stream, in, h, w, oh, ow,
                     (flags & CRYOMI_RESIZE_NORMALIZE) ? 1 : 0, region, out);
  return check_launch("resize_kernel");
}

int cryomi_complex128_to_complex64(const void* in, size_t count, void* out, ihipStream_t* stream) {
  clear_error();
  if (count == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(in && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_complex128_to_complex64: NULL pointer");
  hipLaunchKernelGGL(c128_to_c64_kernel, dim3(grid_for(count)), dim3(NT), 0, stream,
                     static_cast<const double2*>(in), count, static_cast<float2*>(out));
  return check_launch("c128_to_c64_kernel");
}

int cryomi_white_noise_spectrum(int n, int box, double variance, unsigned long long seed, void* out,
                                ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && box >= 2 && box % 2 == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_white_noise_spectrum: need n >= 0 and an even box");
  CRYOMI_REQUIRE(variance > 0.0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_white_noise_spectrum: variance must be > 0");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(out != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_white_noise_spectrum: NULL output");
  const int nh = box * (box / 2 + 1);
  const size_t total = (size_t)n * nh;
  const double std = sqrt((double)box * box * variance);
  hipLaunchKernelGGL(white_noise_kernel, dim3(grid_for((total + 3) / 4)), dim3(NT), 0, stream, total, nh, std, seed,
                     static_cast<float2*>(out));
  return check_launch("white_noise_kernel");
}

int cryomi_noisy_images(const float* signal, const float* noise, size_t count, float scale, float offset, float* out,
                        ihipStream_t* stream) {
  clear_error();
  if (count == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(signal && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_noisy_images: NULL pointer");
  hipLaunchKernelGGL(noisy_kernel, dim3(grid_for(count)), dim3(NT), 0, stream, signal, noise, count, scale, offset,
                     out);
  return check_launch("noisy_kernel");
}

int cryomi_multiply_broadcast(void* data, long long n, long long npix, int data_type, const void* factor,
                              int factor_type, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && npix > 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_multiply_broadcast: need n >= 0, npix > 0");
  const bool single = data_type == CRYOMI_F32 || data_type == CRYOMI_C64;
  const bool single_f = factor_type == CRYOMI_F32 || factor_type == CRYOMI_C64;
  const bool dbl = data_type == CRYOMI_F64 || data_type == CRYOMI_C128;
  const bool dbl_f = factor_type == CRYOMI_F64 || factor_type == CRYOMI_C128;
  const bool real_d = data_type == CRYOMI_F32 || data_type == CRYOMI_F64;
  const bool cplx_f = factor_type == CRYOMI_C64 || factor_type == CRYOMI_C128;
  CRYOMI_REQUIRE((single && single_f) || (dbl && dbl_f), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_multiply_broadcast: data and factor must be known types of the same precision");
  CRYOMI_REQUIRE(!(real_d && cplx_f), CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_multiply_broadcast: a complex factor needs complex data");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(data && factor, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_multiply_broadcast: NULL pointer");
  const size_t total = (size_t)n * (size_t)npix;
  const dim3 grid(grid_for(total)), block(NT);
#define CRYOMI_MB(T, F)                                                                                     \
  hipLaunchKernelGGL((multiply_broadcast_kernel<T, F>), grid, block, 0, stream, static_cast<T*>(data), total, npix, \
                     static_cast<const F*>(factor))
  if (data_type == CRYOMI_F32)
    CRYOMI_MB(float, float);
  else if (data_type == CRYOMI_C64 && factor_type == CRYOMI_F32)
    CRYOMI_MB(float2, float);
  else if (data_type == CRYOMI_C64)
    CRYOMI_MB(float2, float2);
  else if (data_type == CRYOMI_F64)
    CRYOMI_MB(double, double);
  else if (factor_type == CRYOMI_F64)
    CRYOMI_MB(double2, double);
  else
    CRYOMI_MB(double2, double2);
#undef CRYOMI_MB
  return check_launch("multiply_broadcast_kernel");
}

int cryomi_gaussian_log_likelihood(const float* simulated, const float* observed, int n, int npix, double variance,
                                   double* out, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(n >= 0 && npix > 0 && variance > 0.0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gaussian_log_likelihood: need n >= 0, npix > 0, variance > 0");
  if (n == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(simulated && observed && out, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_gaussian_log_likelihood: NULL pointer");
  hipLaunchKernelGGL(loglik_kernel, dim3((unsigned)n), dim3(NT), 0, stream, simulated, observed, npix, variance, out);
  return check_launch("loglik_kernel");
}

}  // extern "C"
