// Spectrum crops and complex recombination for Fourier-crop downsampling of
// volumes and complex arrays (gfx950). The transforms themselves are the FP64
// centred DFTs of dft64.hip (real input); these kernels crop their complex128
// output (optionally to a Hartley transform), undo the centring, and assemble
// the transform of a complex array from the transforms of its two real parts:
//   T(x) = fftshift(fftn(ifftshift(x)))  is linear, so for x = a + i b,
//   T(x) = T(a) + i T(b), and the centred inverse
//   fftshift(ifftn(ifftshift(c))) = conj(T(conj c)) / n
//                                 = (conj T(Re c) + i conj T(Im c)) / n.
// All index maps are per axis; everything is grid-strided and HBM-bound.
#include <hip/hip_runtime.h>

#include "../../include/cryomi_ndimage.h"
#include "common.hpp"

namespace cryomi {
namespace {

constexpr int NT = 256;

unsigned grid_of(size_t work) {
  const size_t b = (work + NT - 1) / NT;
  return (unsigned)(b < 65536 ? (b ? b : 1) : 65536);
}

// Source index along an axis of input size n for output index o of output size on: a centred crop
// (o + n/2 - on/2, crop_to_shape's window) or, uncentring, the ifftshift (o + n/2) mod n.
__device__ __forceinline__ int src_index(int o, int n, int on, bool uncentre) {
  return uncentre ? (o + n / 2) % n : o + n / 2 - on / 2;
}

__global__ __launch_bounds__(NT) void spectrum_crop_kernel(const double2* __restrict__ in, int d, int h, int w, int od,
                                                           int oh, int ow, unsigned flags, double scale,
                                                           void* __restrict__ out) {
  const size_t total = (size_t)od * oh * ow;
  const bool unc = (flags & CRYOMI_SPEC_UNCENTRE) != 0;
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < total; e += (size_t)gridDim.x * NT) {
    const int x = (int)(e % ow), y = (int)((e / ow) % oh), z = (int)(e / ((size_t)ow * oh));
    const int sz = src_index(z, d, od, unc), sy = src_index(y, h, oh, unc), sx = src_index(x, w, ow, unc);
    const double2 v = in[((size_t)sz * h + sy) * w + sx];
    if (flags & CRYOMI_SPEC_HARTLEY) {
      const double r = scale * (v.x - v.y);
      if (flags & CRYOMI_SPEC_OUT_F32) static_cast<float*>(out)[e] = (float)r;
      else static_cast<double*>(out)[e] = r;
    } else {
      if (flags & CRYOMI_SPEC_OUT_F32) static_cast<float2*>(out)[e] = make_float2((float)(scale * v.x), (float)(scale * v.y));
      else static_cast<double2*>(out)[e] = make_double2(scale * v.x, scale * v.y);
    }
  }
}

template <typename C>
__global__ __launch_bounds__(NT) void complex_split_kernel(const C* __restrict__ in, size_t count,
                                                           double* __restrict__ re, double* __restrict__ im) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < count; e += (size_t)gridDim.x * NT) {
    const C v = in[e];
    re[e] = (double)v.x;
    im[e] = (double)v.y;
  }
}

__global__ __launch_bounds__(NT) void complex_combine_kernel(const double2* __restrict__ a, const double2* __restrict__ b,
                                                             size_t count, int inverse, double scale, int out_f32,
                                                             void* __restrict__ out) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < count; e += (size_t)gridDim.x * NT) {
    const double2 p = a[e], q = b[e];
    // forward: p + i q; inverse: conj(p) + i conj(q)
    const double re = inverse ? p.x + q.y : p.x - q.y;
    const double im = inverse ? q.x - p.y : p.y + q.x;
    if (out_f32) static_cast<float2*>(out)[e] = make_float2((float)(scale * re), (float)(scale * im));
    else static_cast<double2*>(out)[e] = make_double2(scale * re, scale * im);
  }
}

}  // namespace
}  // namespace cryomi

using namespace cryomi;

extern "C" {

int cryomi_spectrum_crop(const void* in, int d, int h, int w, int od, int oh, int ow, unsigned flags, double scale,
                         void* out, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(d > 0 && h > 0 && w > 0 && od > 0 && oh > 0 && ow > 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_spectrum_crop: sizes must be positive");
  CRYOMI_REQUIRE((flags & ~(CRYOMI_SPEC_HARTLEY | CRYOMI_SPEC_UNCENTRE | CRYOMI_SPEC_OUT_F32)) == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_spectrum_crop: unknown flag bits");
  CRYOMI_REQUIRE(od <= d && oh <= h && ow <= w, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_spectrum_crop: the output may not exceed the input on any axis");
  CRYOMI_REQUIRE(in && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_spectrum_crop: NULL pointer");
  const size_t total = (size_t)od * oh * ow;
  hipLaunchKernelGGL(spectrum_crop_kernel, dim3(grid_of(total)), dim3(NT), 0, stream, static_cast<const double2*>(in),
                     d, h, w, od, oh, ow, flags, scale, out);
  return check_launch("spectrum_crop_kernel");
}

int cryomi_complex_split(const void* in, int complex128, size_t count, double* re, double* im, ihipStream_t* stream) {
  clear_error();
  if (count == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(in && re && im, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_complex_split: NULL pointer");
  if (complex128)
    hipLaunchKernelGGL(complex_split_kernel<double2>, dim3(grid_of(count)), dim3(NT), 0, stream,
                       static_cast<const double2*>(in), count, re, im);
  else
    hipLaunchKernelGGL(complex_split_kernel<float2>, dim3(grid_of(count)), dim3(NT), 0, stream,
                       static_cast<const float2*>(in), count, re, im);
  return check_launch("complex_split_kernel");
}

int cryomi_complex_combine(const void* a, const void* b, size_t count, int inverse, double scale, int out_complex64,
                           void* out, ihipStream_t* stream) {
  clear_error();
  if (count == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(a && b && out, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_complex_combine: NULL pointer");
  hipLaunchKernelGGL(complex_combine_kernel, dim3(grid_of(count)), dim3(NT), 0, stream, static_cast<const double2*>(a),
                     static_cast<const double2*>(b), count, inverse ? 1 : 0, scale, out_complex64 ? 1 : 0, out);
  return check_launch("complex_combine_kernel");
}

}  // extern "C"
