// postprocess_noalloc.hip -- rtmh_postprocess_noalloc: the image
// post-processing of rtmh_postprocess (Laplacian filter and multi-shot
// normalisation) without any device allocation, for callers that cannot
// allocate (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Its staging lives in one file-scope device array per
// GPU, `scratch`, whose device address is looked up once per device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/rtm_hip.h"
#include "rtm_engine.h"
#include "rtm_launch.h"

// three images of at most RTMH_NOALLOC_MAX_CELLS cells: image1, illum, result
static __device__ __align__(256) float scratch[3 * (size_t)RTMH_NOALLOC_MAX_CELLS];

namespace {
std::mutex g_mu;
float* g_ptr[64] = {nullptr};   // per device: the address of `scratch` there

// the device address of `scratch` on the current device (looked up on the
// first call with that device current; the array's name is not a device
// address in host code)
int scratch_of(int device, float** out)
{
  if (device < 0 || device >= 64) { rtm::set_error("device index out of range"); return -1; }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      rtm::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return -10;
    }
    g_ptr[device] = (float*)p;
  }
  *out = g_ptr[device];
  return 0;
}
}  // namespace

extern "C" int rtmh_postprocess_noalloc(int device, const float* image1, const float* illum, int nx, int nz, float dx,
                                        float dz, float* image_lap, float* image_norm)
{
  if (!image1 || !illum || nx < 1 || nz < 1) { rtm::set_error("bad argument"); return -1; }
  const size_t n = (size_t)nx * nz;
  if (n > (size_t)RTMH_NOALLOC_MAX_CELLS) {
    rtm::set_error("image of " + std::to_string(n) + " cells exceeds the no-allocation limit of " +
                   std::to_string((size_t)RTMH_NOALLOC_MAX_CELLS) + " cells");
    return -3;
  }
  if (hipSetDevice(device) != hipSuccess) { rtm::set_error("hipSetDevice failed"); return -1; }
  float* base = nullptr;
  const int rc0 = scratch_of(device, &base);
  if (rc0) return rc0;
  float* a = base;                                  // image1
  float* b = base + (size_t)RTMH_NOALLOC_MAX_CELLS; // illum
  float* c = b + (size_t)RTMH_NOALLOC_MAX_CELLS;    // result
  int rc = 0;
  if (hipMemcpy(a, image1, n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b, illum, n * 4, hipMemcpyHostToDevice) != hipSuccess) rc = -10;
  if (!rc && image_lap) {
    if (rtm_launch_laplace(a, c, nx, nz, (int)dx, (int)dz, 0) != hipSuccess ||
        hipMemcpy(image_lap, c, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = -10;
  }
  if (!rc && image_norm) {
    if (rtm_launch_norm_multi(a, b, c, (long long)n, 0) != hipSuccess ||
        hipMemcpy(image_norm, c, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = -10;
  }
  if (rc) rtm::set_error("postprocess kernels failed");
  return rc;
}
