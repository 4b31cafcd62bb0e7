// hp2d_handle.hpp - access to a drop-in handle from the per-module sources.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "hp2d.h"

// per-module state a handle keeps until hp2d_destroy (which calls `del`)
struct HandleSlot {
    void *p = nullptr;
    void (*del)(void *) = nullptr;
};
constexpr int kHandleSlots = 2;
constexpr int kSlotSubgrid = 0;
HandleSlot *hp2d_handle_slot(hp2d_handle *h, int which);

// grid shape and precision of a handle
hp2d_status hp2d_handle_shape(hp2d_handle *h, int32_t *ny, int32_t *nx, int32_t *precision);
// the handle's HIP stream
hipStream_t hp2d_handle_stream(hp2d_handle *h);
// record the start (which = 0) or end (1) event of a timed drop-in's compute
// kernels on the handle's stream (read back by hp2d_last_kernel_time)
hipError_t hp2d_handle_mark(hp2d_handle *h, int which);
// record an error on the handle and return HP2D_E_HIP
hp2d_status hp2d_handle_fail(hp2d_handle *h, const std::string &msg);
// device workspace of at least `bytes` owned by the handle (grown on demand,
// kept until hp2d_destroy); NULL on allocation failure (error recorded)
void *hp2d_handle_workspace(hp2d_handle *h, size_t bytes);
// a second, independent workspace (a module that calls another module keeps
// its own arrays here while the callee uses the first)
void *hp2d_handle_workspace2(hp2d_handle *h, size_t bytes);
