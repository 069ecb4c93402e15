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
//                       self-conjugate bins), scatter-add of w g_P into the
//                       grid gradient (FP32 atomics), and per-image sums of
//                       dL/dR (through the interpolation weights), dL/dt and
//                       dL/d defocus (FP64 atomics).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "common.hpp"
#include "sample.hpp"

namespace cryomi {
namespace {

constexpr int NT = 256;
constexpr int RPB = 8;  // rows per block in adj_rows_kernel
constexpr int N_GRAD = 9;  // dR00 dR01 dR10 dR11 dR20 dR21 dtx dty ddefocus

__host__ __device__ inline int strip_w(int m) {
  for (int w : {8, 4, 2, 1})
    if (m % w == 0) return w;
  return 1;
}

__global__ __launch_bounds__(NT) void adj_rows_kernel(const float* __restrict__ ct, int n, int n_img,
                                                      float2* __restrict__ s) {
  extern __shared__ float2 smem[];
  const int m = n / 2;
  const int tid = threadIdx.x;
  const int groups = n / RPB;
  const int img = blockIdx.x / groups;
  const int row0 = (blockIdx.x % groups) * RPB;
  float2* tw = smem;
  float2* bufa = tw + n;
  float2* bufb = bufa + RPB * m;
  fill_twiddles(tw, n, tid, NT);
  const float* src = ct + ((size_t)img * n + row0) * n;
  // z[t] = x[2t] + i x[2t+1]
  for (int e = tid; e < RPB * m; e += NT) {
    const int r = e / m, t = e % m;
    bufa[r * m + t] = reinterpret_cast<const float2*>(src + (size_t)r * n)[t];
  }
  __syncthreads();
  const float2* z = stockham_inverse(bufa, bufb, m, RPB, 1, m, tw, 2, tid, NT);
  // X_inv[k] = E[k] + W^k O[k], E = (Z[k] + Z*[M-k]) / 2, O = (Z[k] - Z*[M-k]) / (2i);
  // the forward transform of the real row is conj(X_inv)
  for (int e = tid; e < RPB * (m + 1); e += NT) {
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
  const int sw = strip_w(m);
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

__global__ __launch_bounds__(NT) void adj_gather_kernel(const float2* __restrict__ grid, int n, int n_img,
                                                        const cryomi_image_params* __restrict__ prm,
                                                        unsigned flags, const float2* __restrict__ gf,
                                                        float* __restrict__ ggrid, double* __restrict__ gparams) {
  __shared__ double red[NT / 64];
  const int m = n / 2;
  const int nh = n * (m + 1);
  const int bpi = (nh + NT - 1) / NT;
  const int img = blockIdx.x / bpi;
  const int s = (blockIdx.x % bpi) * NT + threadIdx.x;
  const cryomi_image_params& p = prm[img];
  const ImageConsts k = make_consts(p, n, false);
  double acc[N_GRAD] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (s < nh) {
    const int i = s / (m + 1), j = s % (m + 1);
    const bool selfc = (i == m || i == 0) && (j == 0 || j == m) && !(i == 0 && j == 0);
    const bool translate = (flags & CRYOMI_APPLY_TRANSLATION) != 0;
    const int jx = (j < m) ? j : j - n;
    const int iy = (i < m) ? i : i - n;
    const double fx = (double)jx, fy = (double)iy;
    const double cc[3] = {fma(k.r20, fx, fma(k.r21, fy, k.half)), fma(k.r10, fx, fma(k.r11, fy, k.half)),
                          fma(k.r00, fx, fma(k.r01, fy, k.half))};  // grid axes (z, y, x)
    int lo[3];
    float fr[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double f = floor(cc[a]);
      lo[a] = (int)f;
      fr[a] = (float)(cc[a] - f);
    }
    // forward sample and its coordinate derivatives
    float2 pv = make_float2(0.f, 0.f);
    float2 dp[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    size_t off[8];
    bool val[8];
    float wgt[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int b0 = q >> 2, b1 = (q >> 1) & 1, b2 = q & 1;
      int z, y, x;
      const bool vz = wrap_index(lo[0] + b0, n, z);
      const bool vy = wrap_index(lo[1] + b1, n, y);
      const bool vx = wrap_index(lo[2] + b2, n, x);
      const bool v = vz && vy && vx;
      const float w0 = b0 ? fr[0] : 1.f - fr[0], w1 = b1 ? fr[1] : 1.f - fr[1], w2 = b2 ? fr[2] : 1.f - fr[2];
      const float d0 = b0 ? 1.f : -1.f, d1 = b1 ? 1.f : -1.f, d2 = b2 ? 1.f : -1.f;
      off[q] = v ? ((size_t)z * n + y) * n + x : 0;
      val[q] = v;
      wgt[q] = w0 * w1 * w2;
      const float2 g = v ? grid[off[q]] : make_float2(0.f, 0.f);
      pv.x = fmaf(wgt[q], g.x, pv.x);
      pv.y = fmaf(wgt[q], g.y, pv.y);
      dp[0] = make_float2(fmaf(d0 * w1 * w2, g.x, dp[0].x), fmaf(d0 * w1 * w2, g.y, dp[0].y));
      dp[1] = make_float2(fmaf(w0 * d1 * w2, g.x, dp[1].x), fmaf(w0 * d1 * w2, g.y, dp[1].y));
      dp[2] = make_float2(fmaf(w0 * w1 * d2, g.x, dp[2].x), fmaf(w0 * w1 * d2, g.y, dp[2].y));
    }
    if (selfc) {
      pv.y = 0.f;
      if (translate) pv.x = 0.f;
    }
    // CTF (FP64 phase), its defocus derivative, and the shift
    double ctf = 1.0, dctf = 0.0;
    if (flags & CRYOMI_APPLY_CTF) {
      const double jj = (double)j;
      const double j2 = jj * jj, i2 = fy * fy, ij = fy * jj;
      const double kk = j2 + i2;
      const double u = fma(k.ca, kk, fma(k.cb, i2 - j2, fma(k.cc, ij, fma(k.cd, kk * kk, -k.cphase))));
      double sn, cs;
      sincospi(2.0 * (u - rint(u)), &sn, &cs);
      ctf = sn;
      // d chi / d defocus = -pi lambda k^2 with k in 1/Angstrom
      const double inv = 1.0 / ((double)n * p.pixel_size);
      dctf = cs * (-M_PI * p.wavelength * kk * inv * inv);
    }
    double sr = 1.0, si = 0.0;  // shift = exp(-2 pi i v)
    const double inv = 1.0 / ((double)n * p.pixel_size);
    if (flags & CRYOMI_APPLY_TRANSLATION) {
      const double v = fma(k.tx, (double)j, k.ty * fy);
      double sn, cs;
      sincospi(2.0 * (v - rint(v)), &sn, &cs);
      sr = cs;
      si = -sn;
    }
    const float2 g = gf[(size_t)img * nh + s];
    const double amp = p.amplitude;
    // g_P = amp CTF conj(shift) g_F
    double gpr = amp * ctf * (sr * g.x + si * g.y);
    double gpi = amp * ctf * (sr * g.y - si * g.x);
    if (selfc) {
      gpi = 0.0;
      if (translate) gpr = 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (val[q] && (gpr != 0.0 || gpi != 0.0)) {
        unsafeAtomicAdd(ggrid + 2 * off[q], (float)(wgt[q] * gpr));
        unsafeAtomicAdd(ggrid + 2 * off[q] + 1, (float)(wgt[q] * gpi));
      }
    }
    {
      // dL/dc_a = Re(conj(g_P) dP/dc_a); c_(x,y,z) = R (jx, iy, 0) + N/2
      const double dlz = gpr * dp[0].x + gpi * dp[0].y;
      const double dly = gpr * dp[1].x + gpi * dp[1].y;
      const double dlx = gpr * dp[2].x + gpi * dp[2].y;
      acc[0] = dlx * fx;
      acc[1] = dlx * fy;
      acc[2] = dly * fx;
      acc[3] = dly * fy;
      acc[4] = dlz * fx;
      acc[5] = dlz * fy;
      // F = amp shift CTF P
      const double fr_ = amp * ctf * (sr * pv.x - si * pv.y);
      const double fi_ = amp * ctf * (sr * pv.y + si * pv.x);
      // dL/dt = Re(conj(g_F) (-2 pi i q) F)
      const double cross = g.x * fi_ - g.y * fr_;  // Re(conj(g) * (-i) * F)
      if (flags & CRYOMI_APPLY_TRANSLATION) {
        acc[6] = 2.0 * M_PI * (double)j * inv * cross;
        acc[7] = 2.0 * M_PI * fy * inv * cross;
      }
      // dL/d defocus = Re(conj(g_F) amp shift P) dCTF
      acc[8] = amp * dctf * (g.x * (sr * pv.x - si * pv.y) + g.y * (sr * pv.y + si * pv.x));
    }
  }
  for (int a = 0; a < N_GRAD; ++a) {
    const double t = block_sum(acc[a], red);
    if (threadIdx.x == 0 && t != 0.0) atomicAdd(gparams + (size_t)img * N_GRAD + a, t);
  }
}

}  // namespace
}  // namespace cryomi

using namespace cryomi;

extern "C" {

size_t cryomi_simulate_vjp_workspace_bytes(int box, int n_images) {
  if (!box_supported(box) || n_images <= 0) return 0;
  return 2 * (size_t)n_images * box * (box / 2 + 1) * sizeof(float2);
}

int cryomi_simulate_vjp(const void* fourier_voxel_grid, int box, int n_images, const cryomi_image_params* params,
                        unsigned flags, const float* cotangent, void* workspace, size_t workspace_bytes,
                        void* grid_grad, double* param_grads, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate_vjp: box must be even, 8 <= box <= 512 and of the form 2^a 3^b");
  CRYOMI_REQUIRE(n_images >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_vjp: n_images < 0");
  CRYOMI_REQUIRE((flags & ~(CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION)) == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_vjp: flags may only hold CRYOMI_APPLY_CTF and CRYOMI_APPLY_TRANSLATION");
  if (n_images == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(fourier_voxel_grid && params && cotangent && grid_grad && param_grads,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_vjp: NULL pointer");
  CRYOMI_REQUIRE(workspace && workspace_bytes >= cryomi_simulate_vjp_workspace_bytes(box, n_images),
                 CRYOMI_ERR_WORKSPACE, "cryomi_simulate_vjp: workspace smaller than cryomi_simulate_vjp_workspace_bytes()");
  const int n = box, m = n / 2;
  float2* s = static_cast<float2*>(workspace);
  float2* gf = s + (size_t)n_images * n * (m + 1);
  hipLaunchKernelGGL(adj_rows_kernel, dim3((unsigned)(n_images * (n / RPB))), dim3(NT),
                     (size_t)(n + 2 * RPB * m) * sizeof(float2), stream, cotangent, n, n_images, s);
  int st = check_launch("adj_rows_kernel");
  if (st) return st;
  const int sw = strip_w(m);
  hipLaunchKernelGGL(adj_cols_kernel, dim3((unsigned)(n_images * (m / sw))), dim3(NT),
                     (size_t)(n + 2 * n * (sw + 1)) * sizeof(float2), stream, s, n, n_images, gf);
  if ((st = check_launch("adj_cols_kernel"))) return st;
  const int nh = n * (m + 1);
  const int bpi = (nh + NT - 1) / NT;
  hipLaunchKernelGGL(adj_gather_kernel, dim3((unsigned)(n_images * bpi)), dim3(NT), 0, stream,
                     static_cast<const float2*>(fourier_voxel_grid), n, n_images, params, flags, gf,
                     static_cast<float*>(grid_grad), param_grads);
  return check_launch("adj_gather_kernel");
}

}  // extern "C"
