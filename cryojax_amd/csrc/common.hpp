// Shared device helpers for the cryomi kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/cryomi.h"

namespace cryomi {

// ---------------------------------------------------------------------------
// error reporting (thread-local last message)
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

#define CRYOMI_REQUIRE(cond, code, msg)     \
  do {                                       \
    if (!(cond)) {                           \
      ::cryomi::set_error(msg);              \
      return (code);                         \
    }                                        \
  } while (0)

int check_launch(const char* what);

// Tuning values (cryomi_set_tuning, defined in forward.hip).
int tuned_wg_per_cu(int box);  // 0 = occupancy limit
int tuned_groups_threads();    // 256 or 512
int tuned_groups_wg_per_cu();  // 0 = occupancy limit
int device_cu_count();

// ---------------------------------------------------------------------------
// complex helpers (float2 = complex64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cmul_i(float2 a) { return make_float2(-a.y, a.x); }

// exp(+2 pi i m / n) for m in [0, n), rounded once from double.
__device__ __forceinline__ void fill_twiddles(float2* tw, int n, int tid, int nthreads) {
  for (int m = tid; m < n; m += nthreads) {
    double s, c;
    sincospi(2.0 * (double)m / (double)n, &s, &c);
    tw[m] = make_float2((float)c, (float)s);
  }
}

// Radix choice for a length-n inverse Stockham pass: 4s, then 2s, then 3s, then 5s.
__device__ __forceinline__ int next_radix(int rem) {
  return (rem % 4 == 0) ? 4 : ((rem % 2 == 0) ? 2 : ((rem % 3 == 0) ? 3 : 5));
}

// Unnormalised inverse DFT (exponent +) of `nbatch` sequences of length n held
// in LDS; element r of sequence b lives at buf[r * es + b * bs]. Stockham
// autosort, ping-ponging between a and b; returns the buffer holding the
// result. `tw[m * tws]` must equal exp(+2 pi i m / n). n = 2^a 3^b 5^c.
// Every thread of the block must call this (it synchronises between stages).
__device__ __forceinline__ float2* stockham_inverse(float2* a, float2* b, int n, int nbatch, int es,
                                                    int bs, const float2* tw, int tws, int tid,
                                                    int nthreads) {
  int ns = 1;
  int rem = n;
  while (rem > 1) {
    const int R = next_radix(rem);
    const int nb = n / R;
    const int twstride = (n / (ns * R)) * tws;
    const int total = nb * nbatch;
    for (int t = tid; t < total; t += nthreads) {
      const int bb = t % nbatch;
      const int j = t / nbatch;
      const int k = j % ns;
      const float2* src = a + bb * bs;
      float2* dst = b + bb * bs;
      const int idxd = (j / ns) * ns * R + k;
      if (R == 4) {
        float2 v0 = src[(j)*es];
        float2 v1 = cmul(src[(j + nb) * es], tw[k * twstride]);
        float2 v2 = cmul(src[(j + 2 * nb) * es], tw[2 * k * twstride]);
        float2 v3 = cmul(src[(j + 3 * nb) * es], tw[3 * k * twstride]);
        float2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
        float2 t2 = cadd(v1, v3), t3 = cmul_i(csub(v1, v3));
        dst[(idxd)*es] = cadd(t0, t2);
        dst[(idxd + ns) * es] = cadd(t1, t3);
        dst[(idxd + 2 * ns) * es] = csub(t0, t2);
        dst[(idxd + 3 * ns) * es] = csub(t1, t3);
      } else if (R == 2) {
        float2 v0 = src[(j)*es];
        float2 v1 = cmul(src[(j + nb) * es], tw[k * twstride]);
        dst[(idxd)*es] = cadd(v0, v1);
        dst[(idxd + ns) * es] = csub(v0, v1);
      } else if (R == 5) {
        // X_q = sum_p v_p e^{+2 pi i p q / 5}: with a_1 = v1 + v4, b_1 = v1 - v4, a_2 = v2 + v3,
        // b_2 = v2 - v3, X_1,4 = v0 + c1 a_1 + c2 a_2 +- i (s1 b_1 + s2 b_2) and
        // X_2,3 = v0 + c2 a_1 + c1 a_2 +- i (s2 b_1 - s1 b_2), c_m + i s_m = e^{2 pi i m / 5}
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        const float2 v0 = src[(j)*es];
        const float2 v1 = cmul(src[(j + nb) * es], tw[k * twstride]);
        const float2 v2 = cmul(src[(j + 2 * nb) * es], tw[2 * k * twstride]);
        const float2 v3 = cmul(src[(j + 3 * nb) * es], tw[3 * k * twstride]);
        const float2 v4 = cmul(src[(j + 4 * nb) * es], tw[4 * k * twstride]);
        const float2 a1 = cadd(v1, v4), b1 = csub(v1, v4), a2 = cadd(v2, v3), b2 = csub(v2, v3);
        const float2 t1 = make_float2(v0.x + c1 * a1.x + c2 * a2.x, v0.y + c1 * a1.y + c2 * a2.y);
        const float2 t2 = make_float2(v0.x + c2 * a1.x + c1 * a2.x, v0.y + c2 * a1.y + c1 * a2.y);
        const float2 u1 = cmul_i(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
        const float2 u2 = cmul_i(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
        dst[(idxd)*es] = cadd(v0, cadd(a1, a2));
        dst[(idxd + ns) * es] = cadd(t1, u1);
        dst[(idxd + 2 * ns) * es] = cadd(t2, u2);
        dst[(idxd + 3 * ns) * es] = csub(t2, u2);
        dst[(idxd + 4 * ns) * es] = csub(t1, u1);
      } else {
        float2 v0 = src[(j)*es];
        float2 v1 = cmul(src[(j + nb) * es], tw[k * twstride]);
        float2 v2 = cmul(src[(j + 2 * nb) * es], tw[2 * k * twstride]);
        const float h = 0.86602540378443864676f;  // sqrt(3)/2
        float2 s = cadd(v1, v2), d = csub(v1, v2);
        float2 m = make_float2(v0.x - 0.5f * s.x, v0.y - 0.5f * s.y);
        float2 id = cscale(cmul_i(d), h);
        dst[(idxd)*es] = cadd(v0, s);
        dst[(idxd + ns) * es] = cadd(m, id);
        dst[(idxd + 2 * ns) * es] = csub(m, id);
      }
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    ns *= R;
    rem /= R;
  }
  return a;
}

// Host-side checks shared by the entry points.
inline bool box_supported(int n) {
  if (n < 8 || n > 512 || (n & 1)) return false;
  int m = n;
  while (m % 2 == 0) m /= 2;
  while (m % 3 == 0) m /= 3;
  while (m % 5 == 0) m /= 5;
  return m == 1;
}

}  // namespace cryomi
