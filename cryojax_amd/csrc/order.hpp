// Image schedule of the persistent image kernels: slices with nearby planes on
// the same XCD (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"

namespace cryomi {

// Images touch the grid lines around their slice plane, which depends only on
// the plane's normal (the rotation's third column). The persistent kernel
// forms schedule position p on workgroup p mod G, and workgroups are dealt to
// the 8 XCDs round robin, so with G a multiple of 8 the XCD of position p is
// p mod 8. order_kernel groups each chunk of up to 4096 images into 256 cells
// of the normal (folded to one hemisphere, Lambert azimuthal projection, a
// 16 x 16 Morton-ordered grid) with a counting sort in LDS, and deals the
// grouped run to the schedule so that the images with p mod 8 == x are the
// x-th eighth of the run: each XCD's L2 then serves concurrent slices whose
// planes are a few degrees apart. The images formed are unchanged; only the
// order in which they are formed differs (within a cell it varies between
// runs).
constexpr int kOrderChunk = 4096;
constexpr int kOrderThreads = 1024;
constexpr int kOrderCells = 256;

__device__ __forceinline__ unsigned spread4(unsigned v) {
  v &= 0xfu;
  v = (v | (v << 2)) & 0x33u;
  v = (v | (v << 1)) & 0x55u;
  return v;
}

// One workgroup of kOrderThreads threads orders the chunk of up to kOrderChunk
// images starting at `base` (a multiple of kOrderChunk): order[base + p] for
// schedule position p of the chunk. Every thread of the workgroup calls it.
__device__ __forceinline__ void order_chunk(const cryomi_image_params* __restrict__ params, int n_images, int base,
                                            int* __restrict__ order) {
  __shared__ int count[kOrderCells];
  __shared__ int cursor[kOrderCells];
  __shared__ int sorted[kOrderChunk];
  constexpr int PER = kOrderChunk / kOrderThreads;
  const int cnt = min(kOrderChunk, n_images - base);
  const int tid = threadIdx.x;
  if (tid < kOrderCells) count[tid] = 0;
  __syncthreads();
  int cell[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = tid + r * kOrderThreads;
    cell[r] = -1;
    if (i < cnt) {
      const double* rot = params[base + i].rotation;
      double nx = rot[2], ny = rot[5], nz = rot[8];
      if (nz < 0.0) {
        nx = -nx;
        ny = -ny;
        nz = -nz;
      }
      const double sc = sqrt(2.0 / (1.0 + nz));
      const double fx = (nx * sc + 1.5) * (16.0 / 3.0), fy = (ny * sc + 1.5) * (16.0 / 3.0);  // Lambert, [0, 16)
      const unsigned qx = (fx > 0.0) ? (unsigned)fmin(fx, 15.0) : 0u;  // NaN -> 0
      const unsigned qy = (fy > 0.0) ? (unsigned)fmin(fy, 15.0) : 0u;
      cell[r] = (int)(spread4(qx) | (spread4(qy) << 1));
      atomicAdd(&count[cell[r]], 1);
    }
  }
  __syncthreads();
  if (tid < 64) {  // exclusive scan of the 256 counts by one wave, 4 per lane
    int v[4], sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[q] = count[tid * 4 + q];
      sum += v[q];
    }
    int incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (tid >= o) incl += t;
    }
    int run = incl - sum;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cursor[tid * 4 + q] = run;
      run += v[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = tid + r * kOrderThreads;
    if (cell[r] >= 0) sorted[atomicAdd(&cursor[cell[r]], 1)] = i;
  }
  __syncthreads();
  // schedule position p (residue x = p mod 8, t = p / 8) takes grouped rank start_x + t
  for (int p = tid; p < cnt; p += kOrderThreads) {
    const int x = p & 7, t = p >> 3;
    int start = 0;
    for (int y = 0; y < x; ++y) start += (cnt - y + 7) >> 3;
    order[base + p] = base + sorted[start + t];
  }
}

}  // namespace cryomi
