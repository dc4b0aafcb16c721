// Framebuffer views for capture inspection (rdp_view.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rdp_mi {

enum : uint32_t { VIEW_COLOR = 0, VIEW_DEPTH = 1, VIEW_COVERAGE = 2 };

struct ViewArgs {
	const uint8_t *rdram;
	const uint8_t *hidden;
	uint32_t rdram_size;
	uint32_t mode;
	uint32_t fb_addr, fb_size;  // colour image address and SetColorImage size field (1, 2 or 3)
	uint32_t depth_addr;
	uint32_t width, height;
};

// RGBA8 (R in the low byte) view of width x height pixels into d_out.
hipError_t launch_view(const ViewArgs &a, uint32_t *d_out, hipStream_t s);

}  // namespace rdp_mi
