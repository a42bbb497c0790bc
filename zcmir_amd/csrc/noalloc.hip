// zc_moments_noalloc: the 45 overlap sums of a pose (the work of zc_moments) for callers that
// cannot allocate device memory — code under HIP graph capture, tight serving loops. The caller
// passes device pointers to both sides' feature planes and to the output; the only other memory
// the kernels touch is `scratch`, a fixed device array of this file that holds the per-workgroup
// partial sums. No hipMalloc, hipFree or allocating library routine is called.
//
// Every call on a device writes that device's one copy of `scratch` (each partial row is written
// in full before it is read, so nothing needs clearing between calls); calls that overlap on one
// device race on it and must be serialized by the caller.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "geometry.hpp"
#include "kernels.hpp"

constexpr int kMaxParts = 1024;   // moments_blocks() launches at most this many workgroups
constexpr int kSums = 45;
constexpr int kMaxDevices = 64;

// one row of 45 partial sums per workgroup of the moments kernel: 1024 x 45 doubles = 360 KiB,
// held by each device the library runs on until the process ends. In host code `scratch` names
// a host-side stand-in, never the device memory: its device address comes from
// hipGetSymbolAddress, per device.
static __device__ __align__(256) double scratch[kMaxParts * kSums];

namespace {

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (the current device), looked up once per device.
double* device_scratch(int dev) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch[dev] = static_cast<double*>(p);
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" {

// Returns 0 on success, -1 for an invalid argument (a null pointer, an empty image, a size whose
// moments launch would need more than kMaxParts partial rows), -2 when the device, the scratch
// lookup or a launch fails. The work is queued on `stream` (null: the default stream); `out` holds
// the 45 sums once it has run.
int32_t zc_moments_noalloc(const float* feat_a, uint32_t wa, uint32_t ha, const float* feat_b, uint32_t wb,
                           uint32_t hb, const double* H, double* out, void* stream) {
  if (!feat_a || !feat_b || !H || !out || wa == 0 || ha == 0 || wb == 0 || hb == 0) return -1;
  if ((uint64_t)wa * ha > (1ull << 30) || (uint64_t)wb * hb > (1ull << 30)) return -1;   // 32-bit byte offsets
  const int nb = zc::moments_blocks((int)wb, (int)hb);
  if (nb < 1 || nb > kMaxParts) return -1;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -2;
  double* part = device_scratch(dev);
  if (!part) return -2;
  const zc::Mat3 Hm = {H[0], H[1], H[2], H[3], H[4], H[5], H[6], H[7], H[8]};
  const zc::Mat3 Hi = zc::inv3(Hm);
  zc::PoseArgs pa;
  for (int i = 0; i < 9; ++i) pa.Hi[i] = Hi[i];
  hipStream_t st = static_cast<hipStream_t>(stream);
  zc::launch_pose_moments(feat_a, (int)wa, (int)ha, feat_b, (int)wb, (int)hb, pa, part, st);
  zc::launch_sum_parts(part, nb, kSums, out, st);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
