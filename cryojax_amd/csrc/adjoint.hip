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
//   adj_pull_kernel   : the grid gradient, brick by brick: the transpose of the
//                       trilinear gather, sum over images and samples of
//                       w(sample, voxel) g_P, accumulated for one brick of
//                       voxels in LDS from the samples that can reach it (no
//                       global atomics).
//
// Why pull: the push form (each sample adding w g_P into its 8 corners) needs
// 16 FP32 atomics per sample, and device-wide FP32 atomics run at about 23 G/s
// on this part whatever the address pattern (1024 x 128^2 images: 6.0 ms of
// atomics against 0.4 ms for everything else). Gathering each brick's
// contributions on chip writes every voxel once.
//
// Pull geometry. Sample k = (kx, ky) sits at grid position q = R (kx, ky, 0) + N/2
// (x, y, z components) and gives corner c the weight prod_a (1 - |q_a - c_a|)
// when |q_a - c_a| < 1 on every axis, with a corner index in [-N, -1] wrapping
// to c + N (JAX fill/wrap rule, as the forward). So voxel v receives from the
// samples near each virtual position u = v - N w, w in {0,1}^3: those within
// sqrt 3 of u, whose in-plane coordinates lie within sqrt 3 of the projection
// of u. A block is a brick of 8 x 8 x 4 voxels; it first culls, per image and
// wrap combination, the slices whose
// plane misses the brick; for each remaining slice, one thread per sample of
// the 16 x 16 window of the plane around the brick forms the forward's own
// stencil (FP64 position, floor, FP32 fractions) and adds w g_P for the
// corners inside the brick into an LDS accumulator (LDS float atomics, so the
// summation order, and the last bits, vary between runs). The brick is then
// added to the grid gradient once, without global atomics.
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
                                                        float2* __restrict__ gpbuf, double* __restrict__ gparams) {
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
      const float wq = w0 * w1 * w2;
      const float2 g = v ? grid[((size_t)z * n + y) * n + x] : make_float2(0.f, 0.f);
      pv.x = fmaf(wq, g.x, pv.x);
      pv.y = fmaf(wq, g.y, pv.y);
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
    gpbuf[(size_t)img * nh + s] = make_float2((float)gpr, (float)gpi);
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
                                                      const float2* __restrict__ gpbuf, float2* __restrict__ ggrid) {
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
  for (int w = 0; w < 8; ++w) {
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
  // g_P per sample goes to the row-transform buffer `s`, which adj_cols_kernel has consumed
  hipLaunchKernelGGL(adj_gather_kernel, dim3((unsigned)(n_images * bpi)), dim3(NT), 0, stream,
                     static_cast<const float2*>(fourier_voxel_grid), n, n_images, params, flags, gf, s,
                     param_grads);
  if ((st = check_launch("adj_gather_kernel"))) return st;
  const unsigned bricks = (unsigned)(((n + PBX - 1) / PBX) * ((n + PBY - 1) / PBY) * ((n + PBZ - 1) / PBZ));
  hipLaunchKernelGGL(adj_pull_kernel, dim3(bricks), dim3(NT), 0, stream, n, n_images, params, s,
                     static_cast<float2*>(grid_grad));
  return check_launch("adj_pull_kernel");
}

}  // extern "C"
