// fos_static.hip -- the JFA label schedule without a context: seeding and
// every pass (what fos_jfa_init followed by fos_jfa_passes computes) with
// the scratch those kernels use in one file-scope device array instead of a
// context's allocations.  For callers that cannot allocate device memory.
//
// The array `scratch` holds, at fixed offsets:
//   seed table   float4 [FOS_STATIC_MAX_N]                       4 MiB
//   counters     int [FOS_PASS_SLOTS][FOS_STRIPES][16]           256 KiB
// Every call on one device shares that device's copy (the last call's
// contents stay there); calls that overlap on one device race.
#include <string.h>

#include <mutex>

#include "fos_common.h"

namespace {

constexpr int64_t kSeedBytes = (int64_t)FOS_STATIC_MAX_N * 16;
constexpr int64_t kStripeBytes = (int64_t)FOS_PASS_SLOTS * FOS_STRIPE_PASS_INTS * 4;
constexpr int64_t kScratchBytes = kSeedBytes + kStripeBytes;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// The device address of this device's copy of `scratch`, looked up once per
// device with that device current.
int scratch_ptr(void **out) {
  int dev = 0;
  FOS_HIP(hipGetDevice(&dev));
  FOS_REQUIRE(dev >= 0 && dev < kMaxDevices, "fos_jfa_labels_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fos::hip_fail(e, "hipGetSymbolAddress(scratch)");
    FOS_REQUIRE(p, "hipGetSymbolAddress(scratch) returned NULL");
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

int fos_static_limits(int *max_n, int *max_res, int64_t *bytes) {
  if (max_n) *max_n = FOS_STATIC_MAX_N;
  if (max_res) *max_res = FOS_STATIC_MAX_RES;
  if (bytes) *bytes = kScratchBytes;
  return 0;
}

int fos_jfa_labels_noalloc(const float *pos, const float *rad, int n, int res, float L,
                           float voxel, float beta, int num_passes, int early_exit,
                           double th_pass_delta, int min_passes, int32_t *grid, int32_t *temp,
                           int32_t *info, void *stream) {
  FOS_REQUIRE(n >= 0 && n <= FOS_STATIC_MAX_N,
              "fos_jfa_labels_noalloc: n=%d exceeds the static scratch (at most %d particles)",
              n, FOS_STATIC_MAX_N);
  FOS_REQUIRE(res >= 1 && res <= FOS_STATIC_MAX_RES,
              "fos_jfa_labels_noalloc: res=%d exceeds the static scratch (at most %d)", res,
              FOS_STATIC_MAX_RES);
  void *base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc) return rc;
  unsigned char *b = (unsigned char *)base;
  // A context on the stack whose only buffers are views of the static array:
  // the seeding and pass kernels read nothing else from it.
  fos_ctx c;
  memset(&c, 0, sizeof(c));
  FOS_HIP(hipGetDevice(&c.device));
  c.max_n = n > 0 ? n : 1;  // picks the 16-bit candidate lists when ids fit
  c.max_res = FOS_STATIC_MAX_RES;
  c.max_cells = 27;
  c.seeds = (float4 *)b;
  c.stripes = (int *)(b + kSeedBytes);
  rc = fos_jfa_init(&c, pos, rad, n, res, L, voxel, beta, grid, nullptr, stream);
  if (rc) return rc;
  return fos_jfa_passes(&c, res, L, voxel, num_passes, early_exit, th_pass_delta, min_passes,
                        grid, temp, info, stream);
}

}  // extern "C"
