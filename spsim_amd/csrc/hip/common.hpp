// Internal helpers shared by the HIP translation units of libspsim_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>

#include "spsim_hip.h"

namespace spsim {

// Records a formatted message as the calling thread's last error and returns code.
int fail(int code, const char *fmt, ...);
// Records a HIP failure (with the runtime's message) and returns SPSIM_HIP_ERR_HIP.
int hip_fail(hipError_t e, const char *what);

// Checks that the current device is gfx950 and caches its CU count.
int ensure_device(int *cus);

// Per-device grow-only scratch allocation owned by the library.
int workspace(size_t bytes, void **ptr);
void workspace_release();

// Streams the library forks independent work onto, kHelperStreams per device,
// created at the first call at the default priority and kept until the process
// ends, with kHelperStreams + 1 events for the fork and the join (handles
// copied out). Calls that
// use them on one device at the same time race, as calls sharing the
// workspace do. The many-molecule NUFFT runs one molecule's chain of small
// launches per stream: with 4 helpers 512 molecules took 62.7 ms, with 7
// 61.0, 15 55.2 and 31 52.7 (the box exposes 4 hardware queues; the deeper
// software queues keep more independent chains ready).
constexpr int kHelperStreams = 31;
int helper_streams(hipStream_t streams[kHelperStreams], hipEvent_t events[kHelperStreams + 1]);

constexpr double kLog2e = 1.4426950408889634073599246810019;

}  // namespace spsim

#define SPSIM_HIP_TRY(call)                                 \
  do {                                                      \
    hipError_t spsim_e_ = (call);                           \
    if (spsim_e_ != hipSuccess) return spsim::hip_fail(spsim_e_, #call); \
  } while (0)

#define SPSIM_TRY(call)          \
  do {                           \
    int spsim_rc_ = (call);      \
    if (spsim_rc_ != SPSIM_HIP_OK) return spsim_rc_; \
  } while (0)
