// Status codes of the library, numbered as sankhya::gpu::GpuStatus
// (include/sankhya/gpu_runtime.hpp:19-30) with one addition for rejected arguments.
#pragma once

namespace sk {

enum : int {
  kStatusSuccess = 0,
  kStatusUnavailable = 1,      // built without HIP, or the HIP runtime cannot be used
  kStatusNoDevice = 2,         // no HIP device present
  kStatusInitFailed = 3,       // device, stream or event creation failed
  kStatusOutOfMemory = 4,      // hipMalloc failed
  kStatusLaunchFailed = 5,     // kernel or graph launch failed
  kStatusCopyFailed = 6,       // host <-> device copy failed
  kStatusSyncFailed = 7,       // synchronisation reported an error
  kStatusInvalidState = 8,     // invalid pointer, buffer or call order
  kStatusNonFinite = 9,        // NaN or Inf produced
  kStatusInvalidArgument = 10  // malformed model or argument (no counterpart in the reference)
};

}  // namespace sk
