// Framebuffer views for capture inspection: the colour, depth and coverage
// images of a framebuffer decoded from device RDRAM and hidden RDRAM into
// RGBA8, as the reference's capture replayer shows them
// (rdp_replayer.cpp:246-393).  One thread per output pixel; RDRAM addresses
// wrap at the RDRAM size.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rdp_view.hpp"

namespace rdp_mi {

__device__ __forceinline__ uint32_t rgba8(uint32_t r, uint32_t g, uint32_t b, uint32_t a)
{
	return r | (g << 8) | (b << 16) | (a << 24);
}

// 5-bit channel to 8 bits by bit replication.
__device__ __forceinline__ uint32_t expand5(uint32_t c) { return (c << 3) | (c >> 2); }

__global__ __launch_bounds__(256) void k_view(ViewArgs a, uint32_t *__restrict__ out)
{
	const uint32_t x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
	if (x >= a.width || y >= a.height) return;
	const uint32_t i = y * a.width + x;
	const uint16_t *r16 = reinterpret_cast<const uint16_t *>(a.rdram);
	const uint32_t *r32 = reinterpret_cast<const uint32_t *>(a.rdram);
	const uint32_t mask16 = (a.rdram_size >> 1) - 1, mask32 = (a.rdram_size >> 2) - 1, mask8 = a.rdram_size - 1;
	uint32_t v = 0;
	switch (a.mode) {
	case VIEW_COLOR:
		if (a.fb_size == 2) {
			// RGBA5551: 5-bit channels, alpha from bit 0.
			const uint32_t c = r16[(((a.fb_addr >> 1) + i) & mask16) ^ 1];
			v = rgba8(expand5((c >> 11) & 31), expand5((c >> 6) & 31), expand5((c >> 1) & 31), (c & 1) * 0xff);
		} else if (a.fb_size == 3) {
			const uint32_t c = r32[((a.fb_addr >> 2) + i) & mask32];
			v = rgba8(c >> 24, (c >> 16) & 0xff, (c >> 8) & 0xff, c & 0xff);
		} else if (a.fb_size == 1) {
			// 8-bit images as grey (the reference leaves these out).
			const uint32_t c = a.rdram[((a.fb_addr + i) & mask8) ^ 3];
			v = rgba8(c, c, c, 0xff);
		}
		break;
	case VIEW_DEPTH: {
		// The 14-bit depth word shown as RGBA5551, its 4-bit dz in alpha.
		const uint32_t idx = ((a.depth_addr >> 1) + i) & mask16;
		uint32_t c = r16[idx ^ 1];
		const uint32_t dz = (a.hidden[idx] | ((c & 3) << 2)) & 0xff;
		c &= ~3u;
		v = rgba8(expand5((c >> 11) & 31), expand5((c >> 6) & 31), expand5((c >> 1) & 31), dz);
		break;
	}
	default: {
		// Coverage of 16-bit images: bit 0 of the pixel over two hidden bits.
		if (a.fb_size != 2) break;
		const uint32_t idx = ((a.fb_addr >> 1) + i) & mask16;
		const uint32_t c = r16[idx ^ 1];
		uint32_t cov = ((c & 1) << 2) | (a.hidden[idx] & 3);
		cov = ((cov << 5) | (cov << 2) | (cov >> 1)) & 0xff;
		v = rgba8(cov, cov, cov, 0xff);
		break;
	}
	}
	out[i] = v;
}

hipError_t launch_view(const ViewArgs &a, uint32_t *d_out, hipStream_t s)
{
	if (a.width == 0 || a.height == 0) return hipSuccess;
	hipLaunchKernelGGL(k_view, dim3((a.width + 63) / 64, (a.height + 3) / 4), dim3(256), 0, s, a, d_out);
	return hipGetLastError();
}

}  // namespace rdp_mi
