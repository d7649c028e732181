// What the run writer (wc_run.hip) offers the solver's entry points (wc_api.hip) for frames
// and statistics taken on the GPU.
#pragma once

#include <stddef.h>

#include "../../include/wavesim_run.h"
#include "wc_json.h"

namespace wc::run {

// One frame on its way from the device to the frames file: `device` holds the run's fields in
// order (nx * ny floats each), `pinned` is the host slot it is copied to, `scalars` the
// kernel's reductions (device) and `scalars_host` their pinned copy.
struct DeviceFrame {
  int slot = -1;
  size_t cells = 0;
  int eta = -1, breaking = -1;  // indices of the fields
  float* device = nullptr;
  float* pinned = nullptr;
  const float* elevation = nullptr;  // device copy of the run's bed elevation
  unsigned long long* scalars = nullptr;
  unsigned long long* scalars_host = nullptr;
};

// Waits for a free pinned slot (allocating the device buffers on `device` at first use) for a
// grid of nx * ny cells; fails if the run's grid or fields do not fit a device frame.
int begin_device_frame(wc_run_writer* w, int device, long long nx, long long ny,
                       DeviceFrame* f);
// Hands the slot to the writer's thread as the frame at `time`.
int commit_frame(wc_run_writer* w, const DeviceFrame& f, double time);
// Gives the slot back unwritten.
void abandon_frame(wc_run_writer* w, const DeviceFrame& f);
// waves[key] = value
void note(wc_run_writer* w, const char* key, json::Value value);

}  // namespace wc::run
