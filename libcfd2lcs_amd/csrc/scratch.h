// scratch.h -- the fixed device workspace (scratch.hip) of the
// allocation-free entry point cfd2lcs_adjoint_forward_noalloc_c.
#pragma once
#include <cstddef>

namespace lcs {

// bytes of the per-device workspace: enough for the forward FTLE of a window
// on grids of about 466k ghosted nodes in FP32 with cubic B-splines (about
// 74^3, or 680^2 in 2-D), half that in FP64
constexpr size_t kScratchBytes = size_t(64) << 20;

// looks up the workspace's device address for the current device (once per
// device); false if the runtime cannot resolve it
bool scratch_setup();
// the workspace's device address on the current device (null before setup)
void* scratch_ptr();

}  // namespace lcs
