// Register-FFT image former for N in {64, 128, 160, 192, 256}: one workgroup of
// NTH threads forms one whole image (slice samples or a given half-plane
// spectrum -> fftshift(irfft2)) through a per-workgroup scratch slot.
//
//   edge pass   : samples of columns 0 and N/2 for every row -> LDS.
//   phase A     : for each group of CG rfft columns, a four-step inverse FFT
//                 along y: pass 0 stores the samples in LDS; pass 1 has thread
//                 (c, n2) run a 16-point DFT over x[N2 n1 + n2] in registers,
//                 twiddle by exp(2 pi i n2 k1 / N) and write back in place;
//                 pass 2 runs the N2-point DFTs and stores rows k1 + 16 k2 to
//                 the scratch slot. Columns 0 and N/2 travel together as one
//                 complex column h0 + i hM, with h the Hermitian part of each
//                 column, whose transform is Re G[., 0] + i Re G[., N/2] -
//                 exactly the parts the C2R keeps.
//   phase B     : for each group of RG rows, the C2R along x as an M-point
//                 complex inverse FFT (M = N/2) of the real-pair packing
//                 Z[k] = (X[k] + X*[M-k]) + i e^{2 pi i k/N} (X[k] - X*[M-k]),
//                 again four-step (16 x M2) with one LDS pass; output rows
//                 are written as float2 pairs (x[2m], x[2m+1]).
//
// NTH (the workgroup size) sets the tile shape: CG = the largest divisor of
// M with CG * N/16 <= NTH, RG = the largest divisor of N with RG * M/16 <= NTH.
// NTH = N gives 16 columns per group.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "dft.hpp"
#include "sample.hpp"

namespace cryomi {

constexpr int largest_divisor_le(int n, int cap) {
  for (int d = cap; d > 1; --d)
    if (n % d == 0) return d;
  return 1;
}

template <int N, int NTH>
struct RegPlan {
  static constexpr int M = N / 2;
  static constexpr int N2 = N / 16;
  static constexpr int CG = largest_divisor_le(M, NTH / N2);  // columns per group
  static constexpr int PA = CG * N2;                          // threads busy in pass 1 of phase A
  static constexpr int NGROUPS = M / CG;
  static constexpr int M2 = M / 16;
  static constexpr int RG = largest_divisor_le(N, NTH / M2);  // rows per group
  static constexpr int NROWGROUPS = N / RG;
  static constexpr int LDS_A = N * CG;
  static constexpr int LDS_EDGE = 2 * N;
  static constexpr int LDS_ROW = N;  // row factors of the shift phase (phase A)
  static constexpr int RS = M + 8;   // padded LDS row stride of phase B
  static constexpr int LDS_B = RG * RS;
  // LDS float2 elements behind the twiddle table
  static constexpr int LDS_PHASE_A = LDS_A + LDS_EDGE + LDS_ROW;
  static constexpr int LDS = LDS_PHASE_A > LDS_B ? LDS_PHASE_A : LDS_B;
  static_assert(N % 32 == 0 && NGROUPS * CG == M && NROWGROUPS * RG == N, "unsupported register plan");
  static_assert(PA <= NTH && RG * M2 <= NTH, "register plan exceeds the workgroup");
  static_assert((N * CG) % (2 * NTH) == 0, "pass 0 pairs samples");
  static_assert(NTH % CG == 0, "pass 0 keeps one column per thread");
};

template <int N, bool FROM_SPECTRUM, bool BRICKED>
__device__ __forceinline__ float2 signed_sample(const float2* __restrict__ grid, const float2* __restrict__ sp,
                                                const ImageConsts& k, unsigned flags, int i, int j) {
  constexpr int M = N / 2;
  float2 v;
  if (FROM_SPECTRUM) {
    constexpr float inv = 1.0f / ((float)N * (float)N);
    const float2 x = sp[(size_t)i * (M + 1) + j];
    v = k.finish ? finish_sample(x, N, i, j, k, flags) : cscale(x, inv);
  } else {
    Gather st;
    gather_issue<BRICKED>(grid, N, i, j, k, st);
    v = gather_finish<BRICKED>(N, k, flags, st);
  }
  if ((i + j) & 1) v = make_float2(-v.x, -v.y);  // fftshift of the output image
  return v;
}

// Column 0 of phase A: 0.5 (a + conj a(-i)) + i * 0.5 (b + conj b(-i)) with a,
// b the (signed) samples of columns 0 and N/2 kept in `edge`.
template <int N>
__device__ __forceinline__ float2 combined_edge(const float2* edge, int i) {
  const int ni = (N - i) % N;
  const float2 a = edge[i], an = edge[ni], b = edge[N + i], bn = edge[N + ni];
  return make_float2(0.5f * (a.x + an.x) - 0.5f * (b.y - bn.y), 0.5f * (a.y - an.y) + 0.5f * (b.x + bn.x));
}

// One image. `tw` holds exp(+2 pi i m / N), m < N; `lds` has RegPlan::LDS
// elements; `slot` N * M complex values of scratch; `outimg` the N x N float32
// image viewed as float2 pairs. Every thread of the workgroup calls this with
// the same arguments (all barriers are workgroup-wide); the LDS and the slot
// are free again when it returns.
// An empty asm statement the compiler must assume changes `v`: values derived
// from it are formed after this point, not hoisted out of enclosing loops.
__device__ __forceinline__ int fresh(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <int N, int NTH, bool FROM_SPECTRUM, bool BRICKED>
__device__ __forceinline__ void reg_image(const float2* __restrict__ grid, const float2* __restrict__ sp,
                                          const ImageConsts& k, unsigned flags, const float2* tw, float2* lds,
                                          float2* __restrict__ slot, float2* __restrict__ outimg, int tid_in) {
  using P = RegPlan<N, NTH>;
  // Per-thread indices and addresses are formed per image: hoisted across the
  // persistent loop they would hold registers through every phase.
  const int tid = fresh(tid_in);
  constexpr int M = P::M, N2 = P::N2, CG = P::CG, M2 = P::M2, RG = P::RG;
  float2* edge = lds + P::LDS_A;                // [0, N): column 0, [N, 2N): column M
  float2* rowph = lds + P::LDS_A + P::LDS_EDGE;  // (-1)^i exp(-2 pi i ty fy(i))
  // ---- edge columns 0 and M, and the row factors of the shift phase
  if (FROM_SPECTRUM) {
#pragma unroll 1
    for (int t = tid; t < 2 * N; t += NTH) {
      const int i = t < N ? t : t - N;
      edge[t] = signed_sample<N, true, false>(grid, sp, k, flags, i, t < N ? 0 : M);
    }
  } else {
    // two samples' loads in flight per thread (rows of columns 0 and M when NTH = N)
#pragma unroll 1
    for (int t = tid; t < 2 * N; t += 2 * NTH) {
      const int tb = t + NTH;
      const int ia = t < N ? t : t - N, ja = t < N ? 0 : M;
      const int ib = tb < N ? tb : tb - N, jb = tb < N ? 0 : M;
      Gather sta, stb;
      gather_issue<BRICKED>(grid, N, ia, ja, k, sta);
      if (tb < 2 * N) gather_issue<BRICKED>(grid, N, ib, jb, k, stb);
      float2 va = gather_finish<BRICKED>(N, k, flags, sta);
      if ((ia + ja) & 1) va = make_float2(-va.x, -va.y);  // fftshift of the output image
      edge[t] = va;
      if (tb < 2 * N) {
        float2 vb = gather_finish<BRICKED>(N, k, flags, stb);
        if ((ib + jb) & 1) vb = make_float2(-vb.x, -vb.y);
        edge[tb] = vb;
      }
    }
  }
  if (!FROM_SPECTRUM) {
#pragma unroll 1
    for (int i = tid; i < N; i += NTH) rowph[i] = row_phase_factor(k, N, i, flags);
  }
  __syncthreads();
  // ---- phase A: y transforms, CG columns at a time
#pragma unroll 1
  for (int g = 0; g < P::NGROUPS; ++g) {
    // pass 0: samples of rows i, columns g*CG + c -> lds[i * CG + c]; two
    // samples per thread in flight
    if (FROM_SPECTRUM) {
#pragma unroll 1
      for (int s = tid; s < N * CG; s += 2 * NTH) {
        const int ca = s % CG, ia = s / CG;
        const int cb = (s + NTH) % CG, ib = (s + NTH) / CG;
        const int ja = g * CG + ca, jb = g * CG + cb;
        lds[s] = signed_sample<N, true, false>(grid, sp, k, flags, ia, ja);
        lds[s + NTH] = signed_sample<N, true, false>(grid, sp, k, flags, ib, jb);
        // column 0 carries the combined DC / Nyquist column
        if (ja == 0) lds[s] = combined_edge<N>(edge, ia);
        if (jb == 0) lds[s + NTH] = combined_edge<N>(edge, ib);
      }
    } else {
      // Every sample of this thread lies in column j (NTH is a multiple of
      // CG), so the coordinate bases, the CTF polynomial and the column's
      // shift factor are formed once per group. Column 0 (g = 0, c = 0) is
      // gathered and then replaced by the combined DC / Nyquist column.
      const int c = tid % CG;
      const int j = g * CG + c;
      const ColumnConsts cc = column_consts(k, j, flags);
      const double p4 = ctf_p4(k);
      const int lo = k.wrap ? -N : 0;
      const unsigned range = k.wrap ? 2u * N : (unsigned)N;
      auto store = [&](int s, int i, float2 acc) {
        const double fy = (double)(i < M ? i : i - N);
        const float2 v = column_finish(acc, cc, p4, fy, rowph[i], flags);
        lds[s] = j == 0 ? combined_edge<N>(edge, i) : v;
      };
      if (BRICKED) {
        // Two samples' loads in flight per thread. Issuing the next pair's
        // loads before summing this pair measured the same in the single-box
        // kernel and spilled in the grouped one.
        const char* gb = reinterpret_cast<const char*>(grid);
#pragma unroll 1
        for (int s = tid; s < N * CG; s += 2 * NTH) {
          const int ia = s / CG, ib = (s + NTH) / CG;
          const double fya = (double)(ia < M ? ia : ia - N), fyb = (double)(ib < M ? ib : ib - N);
          BrickGather sta, stb;
          brick_issue<N>(gb, fma(k.r21, fya, cc.bz), fma(k.r11, fya, cc.by), fma(k.r01, fya, cc.bx), k.wrap, lo,
                         range, sta);
          brick_issue<N>(gb, fma(k.r21, fyb, cc.bz), fma(k.r11, fyb, cc.by), fma(k.r01, fyb, cc.bx), k.wrap, lo,
                         range, stb);
          store(s, ia, brick_sum(sta));
          store(s + NTH, ib, brick_sum(stb));
        }
      } else {
#pragma unroll 1
        for (int s = tid; s < N * CG; s += 2 * NTH) {
          const int ia = s / CG, ib = (s + NTH) / CG;
          const double fya = (double)(ia < M ? ia : ia - N), fyb = (double)(ib < M ? ib : ib - N);
          Gather sta, stb;
          stencil_issue<false>(grid, N, fma(k.r21, fya, cc.bz), fma(k.r11, fya, cc.by), fma(k.r01, fya, cc.bx), k,
                               sta);
          stencil_issue<false>(grid, N, fma(k.r21, fyb, cc.bz), fma(k.r11, fyb, cc.by), fma(k.r01, fyb, cc.bx), k,
                               stb);
          store(s, ia, stencil_sum<false>(sta));
          store(s + NTH, ib, stencil_sum<false>(stb));
        }
      }
    }
    __syncthreads();
    // pass 1 (in place): thread (c, n2) reads rows N2 n1 + n2, 16-point DFT
    // over n1, twiddle, writes slot k1 N2 + n2 - the same 16 locations
    if (tid < P::PA) {
      const int tp = fresh(tid);
      const int c = tp % CG;
      const int n2 = tp / CG;
      float2 x[16];
#pragma unroll
      for (int n1 = 0; n1 < 16; ++n1) x[n1] = lds[(N2 * n1 + n2) * CG + c];
      Dft<16>::run(x);
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) {
        const float2 v = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], tw[n2 * k1]);
        lds[(k1 * N2 + n2) * CG + c] = v;
      }
    }
    __syncthreads();
    // pass 2: thread (c, k1) runs the N2-point DFT over n2 -> rows k1 + 16 k2
#pragma unroll 1
    for (int t = fresh(tid); t < CG * 16; t += NTH) {
      const int c = t % CG;
      const int k1 = t / CG;
      float2 u[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) u[n2] = lds[(k1 * N2 + n2) * CG + c];
      Dft<N2>::run(u);
      // streaming stores: each slot value is read back once, by phase B
      typedef float f2v __attribute__((ext_vector_type(2)));
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2)
        __builtin_nontemporal_store(f2v{u[k2].x, u[k2].y},
                                    reinterpret_cast<f2v*>(slot + (size_t)(k1 + 16 * k2) * M + g * CG + c));
    }
    __syncthreads();
  }
  __threadfence_block();
  __syncthreads();
  // ---- phase B: x C2R, RG rows at a time
#pragma unroll 1
  for (int rg = 0; rg < P::NROWGROUPS; ++rg) {
    // pass 0: pack scratch rows into Z -> lds[r * RS + kk]; one item loads
    // X[kk] and X[M-kk] once and forms both Z[kk] and Z[M-kk]
    constexpr int HALF = M / 2 + 1;
#pragma unroll 2
    for (int s = fresh(tid); s < RG * HALF; s += NTH) {
      const int r = s / HALF;
      const int kk = s % HALF;
      const float2* X = slot + (size_t)(rg * RG + r) * M;
      // streaming loads: the slot is read once
      typedef float f2l __attribute__((ext_vector_type(2)));
      const f2l la = __builtin_nontemporal_load(reinterpret_cast<const f2l*>(X + kk));
      const f2l lb = __builtin_nontemporal_load(reinterpret_cast<const f2l*>(X + (kk == 0 ? 0 : M - kk)));
      float2 a = make_float2(la.x, la.y);
      float2 b = make_float2(lb.x, lb.y);
      if (kk == 0) {  // X[0] = Re G[., 0] + i Re G[., M]
        b = make_float2(a.y, 0.f);
        a = make_float2(a.x, 0.f);
      }
      float2* row = lds + r * P::RS;
      {
        const float2 bc = cconj(b);
        const float2 sum = cadd(a, bc);
        const float2 wd = cmul(tw[kk], csub(a, bc));
        row[kk] = make_float2(sum.x - wd.y, sum.y + wd.x);
      }
      if (kk != 0 && 2 * kk != M) {
        const float2 ac = cconj(a);
        const float2 sum = cadd(b, ac);
        const float2 wd = cmul(tw[M - kk], csub(b, ac));
        row[M - kk] = make_float2(sum.x - wd.y, sum.y + wd.x);
      }
    }
    __syncthreads();
    // pass 1 (in place): thread (n2, r): 16-point DFT over n1 of Z[M2 n1 + n2],
    // twiddle exp(2 pi i n2 k1 / M), back to slot k1 M2 + n2
#pragma unroll 1
    for (int t = fresh(tid); t < RG * M2; t += NTH) {
      const int n2 = t % M2;
      const int r = t / M2;
      float2* row = lds + r * P::RS;
      float2 x[16];
#pragma unroll
      for (int n1 = 0; n1 < 16; ++n1) x[n1] = row[M2 * n1 + n2];
      Dft<16>::run(x);
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1)
        row[k1 * M2 + n2] = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], tw[2 * n2 * k1]);
    }
    __syncthreads();
    // pass 2: thread (k1, r): M2-point DFT over n2 -> z[k1 + 16 k2] = (x[2m], x[2m+1])
#pragma unroll 1
    for (int t = fresh(tid); t < RG * 16; t += NTH) {
      const int k1 = t % 16;
      const int r = t / 16;
      const float2* src = lds + r * P::RS + k1 * M2;
      float2 u[M2];
#pragma unroll
      for (int n2 = 0; n2 < M2; ++n2) u[n2] = src[n2];
      Dft<M2>::run(u);
      float2* orow = outimg + (size_t)(rg * RG + r) * M;
#pragma unroll
      for (int k2 = 0; k2 < M2; ++k2) {
        // streaming store: the images are not re-read by this kernel
        typedef float f2v __attribute__((ext_vector_type(2)));
        const f2v v = {u[k2].x, u[k2].y};
        __builtin_nontemporal_store(v, reinterpret_cast<f2v*>(orow + k1 + 16 * k2));
      }
    }
    __syncthreads();
  }
  __threadfence_block();
  __syncthreads();  // the slot and the LDS are rewritten by the next image
}

}  // namespace cryomi
