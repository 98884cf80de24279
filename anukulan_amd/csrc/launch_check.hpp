// Host-side check of a kernel launch before it is enqueued: the grid and
// workgroup shape, and the timing events a dispatch records
// (hipExtLaunchKernel). Plain C++ with no HIP types, so the rules can be
// compiled and exercised on a machine without a GPU.
//
// The rules:
//  * every grid and workgroup dimension is at least 1;
//  * a workgroup holds at most 1024 work-items (gfx950's limit);
//  * grid.x is below 2^31, grid.y and grid.z below 2^16;
//  * timing events come as a pair recorded by this one dispatch, start and
//    stop, or not at all; the two are distinct. A dispatch that records only
//    one of them, or an interval split across two dispatches, is refused:
//    such an interval is bracketed with stream-recorded events instead.
#pragma once

#include <cstdint>

namespace anuk {

struct LaunchShape {
  std::uint64_t gx = 1, gy = 1, gz = 1;  // workgroups per dimension
  std::uint64_t bx = 1, by = 1, bz = 1;  // work-items per workgroup dimension
};

inline constexpr std::uint64_t kMaxWorkgroupItems = 1024;

// nullptr when the launch may be enqueued, else why not.
inline const char* launch_shape_error(const LaunchShape& s, const void* start_event,
                                      const void* stop_event) {
  if (s.gx < 1 || s.gy < 1 || s.gz < 1) return "empty grid";
  if (s.bx < 1 || s.by < 1 || s.bz < 1) return "empty workgroup";
  if (s.bx * s.by * s.bz > kMaxWorkgroupItems) return "workgroup larger than 1024 work-items";
  if (s.gx >= (std::uint64_t{1} << 31)) return "grid.x too large";
  if (s.gy >= (std::uint64_t{1} << 16) || s.gz >= (std::uint64_t{1} << 16))
    return "grid.y or grid.z too large";
  if ((start_event == nullptr) != (stop_event == nullptr))
    return "a dispatch records both timing events or neither";
  if (start_event != nullptr && start_event == stop_event)
    return "start and stop timing events are the same event";
  return nullptr;
}

}  // namespace anuk
