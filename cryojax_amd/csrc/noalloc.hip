// Allocation-free variant of cryomi_simulate: the per-workgroup scratch slots of
// the persistent image kernels live in a file-scope device array instead of a
// caller workspace, so the call allocates nothing.
//
// The array's host-side name is only a stand-in for the device symbol: the
// device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and only that address is passed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"

namespace cryomi {
int simulate_in_scratch(const float2* grid, int box, int n_images, const cryomi_image_params* params, unsigned flags,
                        void* scratch, size_t scratch_bytes, float* out, hipStream_t stream);
}  // namespace cryomi

using namespace cryomi;

// Each GPU the library runs on holds this array until the process ends. The
// image kernels write every slot entry before reading it, so nothing in it
// has to start at zero.
static __device__ __align__(256) unsigned char scratch[CRYOMI_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first call per device).
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error(std::string("cryomi_simulate_noalloc: no usable current device: ") + hipGetErrorString(e));
    return CRYOMI_ERR_LAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("cryomi_simulate_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                hipGetErrorString(e));
      return CRYOMI_ERR_LAUNCH;
    }
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return CRYOMI_OK;
}

}  // namespace

extern "C" {

int cryomi_simulate_noalloc(const void* fourier_voxel_grid, int box, int n_images, const cryomi_image_params* params,
                            unsigned flags, void* out, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(box_supported(box), CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate_noalloc: box must be even, 8 <= box <= 512 and of the form 2^a 3^b 5^c");
  CRYOMI_REQUIRE(n_images >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_noalloc: n_images < 0");
  CRYOMI_REQUIRE((flags & ~(CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION | CRYOMI_OUTPUT_FOURIER |
                             CRYOMI_GRID_BRICKED | CRYOMI_GATHER_FILL)) == 0,
                 CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_noalloc: unknown flag bits");
  CRYOMI_REQUIRE(!(flags & CRYOMI_GRID_BRICKED) || box % 4 == 0, CRYOMI_ERR_UNSUPPORTED,
                 "cryomi_simulate_noalloc: the bricked grid layout needs box % 4 == 0");
  if (n_images == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(fourier_voxel_grid && params && out, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_noalloc: NULL grid, params or output pointer");
  if (flags & CRYOMI_OUTPUT_FOURIER)  // the Fourier-space path uses no scratch
    return cryomi_simulate(fourier_voxel_grid, box, n_images, params, flags, nullptr, 0, out, stream);
  void* ptr = nullptr;
  const int st = scratch_address(&ptr);
  if (st) return st;
  return simulate_in_scratch(static_cast<const float2*>(fourier_voxel_grid), box, n_images, params, flags, ptr,
                             (size_t)CRYOMI_NOALLOC_SCRATCH_BYTES, static_cast<float*>(out), stream);
}

}  // extern "C"
