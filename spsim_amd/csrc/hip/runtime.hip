// Library runtime: error reporting, device checks and the per-device workspace.
#include <cstdio>
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "../host/cromer_mann.h"

namespace spsim {

namespace {
// Fixed-size storage, so that recording an error or the first call on a device
// allocates no host memory.
thread_local char g_last_error[1024];
constexpr int kMaxDevices = 64;

struct DeviceState {
  int checked = 0;  // 0 = not yet, 1 = ok, -1 = unusable
  int cus = 0;
  void *ws = nullptr;
  size_t ws_bytes = 0;
  int helpers = 0;  // 1 once hs and ev exist
  hipStream_t hs[kHelperStreams] = {};
  hipEvent_t ev[kHelperStreams + 1] = {};
};
std::mutex g_mu;
DeviceState g_dev[kMaxDevices];

DeviceState *state_for_current(int *dev_out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  if (dev_out) *dev_out = dev;
  return &g_dev[dev];
}
}  // namespace

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(SPSIM_HIP_ERR_HIP, "%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
}

int ensure_device(int *cus) {
  int dev = 0;
  DeviceState *st = state_for_current(&dev);
  if (!st) return fail(SPSIM_HIP_ERR_NODEVICE, "no HIP device is available");
  if (st->checked == 0) {
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      st->checked = -1;
      return fail(SPSIM_HIP_ERR_NODEVICE, "device %d is %s; this library is built for gfx950 (MI355X) only", dev,
                  prop.gcnArchName);
    }
    st->cus = prop.multiProcessorCount;
    st->checked = 1;
  }
  if (st->checked < 0) return fail(SPSIM_HIP_ERR_NODEVICE, "current device is not gfx950");
  if (cus) *cus = st->cus;
  return SPSIM_HIP_OK;
}

int workspace(size_t bytes, void **ptr) {
  int dev = 0;
  DeviceState *st = state_for_current(&dev);
  if (!st) return fail(SPSIM_HIP_ERR_NODEVICE, "no HIP device is available");
  if (st->ws_bytes < bytes) {
    if (st->ws) {
      // The previous workspace may still be in use by queued work on any stream.
      SPSIM_HIP_TRY(hipDeviceSynchronize());
      SPSIM_HIP_TRY(hipFree(st->ws));
      st->ws = nullptr;
      st->ws_bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&st->ws, want);
    if (e != hipSuccess) return fail(SPSIM_HIP_ERR_NOMEM, "workspace hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    st->ws_bytes = want;
  }
  *ptr = st->ws;
  return SPSIM_HIP_OK;
}

int helper_streams(hipStream_t streams[kHelperStreams], hipEvent_t events[kHelperStreams + 1]) {
  DeviceState *st = state_for_current(nullptr);
  if (!st) return fail(SPSIM_HIP_ERR_NODEVICE, "no HIP device is available");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!st->helpers) {
    for (int k = 0; k < kHelperStreams; ++k)
      if (!st->hs[k]) SPSIM_HIP_TRY(hipStreamCreateWithFlags(&st->hs[k], hipStreamNonBlocking));
    for (int k = 0; k <= kHelperStreams; ++k)
      if (!st->ev[k]) SPSIM_HIP_TRY(hipEventCreateWithFlags(&st->ev[k], hipEventDisableTiming));
    st->helpers = 1;
  }
  for (int k = 0; k < kHelperStreams; ++k) streams[k] = st->hs[k];
  for (int k = 0; k <= kHelperStreams; ++k) events[k] = st->ev[k];
  return SPSIM_HIP_OK;
}

void workspace_release() {
  int dev = 0;
  DeviceState *st = state_for_current(&dev);
  if (!st || !st->ws) return;
  (void)hipDeviceSynchronize();
  (void)hipFree(st->ws);
  st->ws = nullptr;
  st->ws_bytes = 0;
}

}  // namespace spsim

extern "C" {

const char *spsim_hip_last_error(void) { return spsim::g_last_error; }

int spsim_hip_device_count(int *count) {
  if (!count) return spsim::fail(SPSIM_HIP_ERR_INVALID, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return spsim::hip_fail(e, "hipGetDeviceCount");
  }
  *count = n;
  return SPSIM_HIP_OK;
}

int spsim_hip_cu_count(int *cus) {
  if (!cus) return spsim::fail(SPSIM_HIP_ERR_INVALID, "cus is NULL");
  return spsim::ensure_device(cus);
}

int spsim_hip_synchronize(void *stream) {
  SPSIM_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return SPSIM_HIP_OK;
}

int spsim_hip_default_form_factors(float *out) {
  if (!out) return spsim::fail(SPSIM_HIP_ERR_INVALID, "out is NULL");
  std::memcpy(out, spsim_cromer_mann, sizeof(float) * SPSIM_CM_ROWS * SPSIM_CM_COLS);
  float *unit = out + SPSIM_HIP_UNIT_SPECIES * SPSIM_HIP_FF_STRIDE;
  for (int i = 0; i < SPSIM_HIP_FF_STRIDE; ++i) unit[i] = 0.0f;
  unit[8] = 1.0f;  // f = c = 1 at every |q|
  return SPSIM_HIP_OK;
}

int spsim_hip_release(void) {
  spsim::workspace_release();
  return SPSIM_HIP_OK;
}

}  // extern "C"
