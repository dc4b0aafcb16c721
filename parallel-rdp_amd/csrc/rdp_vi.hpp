// Video interface (VI) scanout on the device: register decode on the host,
// then two kernels (fetch filter, then divot + scale + gamma + blend).
// Follows parallel-rdp/video_interface.cpp:176-403, 796-998, 1126-1426 and
// shaders/extract_vram.comp, vi_fetch.frag, vi_divot.frag, vi_scale.frag.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rdp_mi.h"

namespace rdp_mi {

constexpr int kViMaxLines = 288;   // VI_MAX_OUTPUT_SCANLINES
constexpr int kViVEndMax = 621;    // VI_V_END_MAX

struct ViLine {
	int32_t h_start, h_start_clamp, h_end_clamp, x_start, x_add, y_start, y_add, y_base;
};

// Everything the two VI kernels need for one scanout.
struct ViParams {
	uint32_t status;
	int32_t fb_offset, fb_width, x_off, y_off;
	int32_t aw, ah, divot, fetch_bug;
	int32_t ow, oh, crop_l, crop_t;
	int32_t draw_x, draw_y, draw_w, draw_h;
	int32_t v_start, serrate_shift, serrate_mask, serrate_select, info_y_shift;
	int32_t init_y_add;
	uint32_t frame_count;
	int32_t blend, nrects, prev_w, prev_h;
	int32_t rects[3][4];
	uint32_t rdram_mask;
	int32_t scale_log2;  // upscaled scanout: the image is read from 4^scale_log2 sample planes
	uint32_t pad[2];
	ViLine lines[kViMaxLines];
};

// Where the scanout reads the frame.  scale_log2 = 0: rdram / hidden are
// the native memories.  scale_log2 > 0: the sample planes of the upscaled
// domain (rdram_size and rdram_size / 2 bytes each); before the fetch the
// domain's pre update runs over the range the VI reads, from the native
// memories and the reference copy below.
struct ViSource {
	const uint8_t *rdram, *hidden;
	uint32_t rdram_size, scale_log2;
	const uint8_t *native_rdram, *native_hidden;
	uint8_t *ref;
	uint8_t *planes, *hplanes;
};

// Per-processor VI state that persists between scanouts
// (video_interface.hpp:140-167).
struct ViState {
	uint32_t frame_count = 0, last_valid_frame_count = 0;
	bool previous_frame_blank = false;
	bool prev_valid = false;
	unsigned prev_w = 0, prev_h = 0;
	uint32_t *d_images[2] = { nullptr, nullptr };  // output, ping-pong with the previous one
	size_t image_words = 0;                         // capacity of each
	int cur = 0;
	uint32_t *d_filter = nullptr;                   // fetch-filter image, both layers
	size_t filter_words = 0;
	ViParams *h_params = nullptr;                   // pinned
	ViParams *d_params = nullptr;
	hipEvent_t params_done = nullptr;               // the upload of h_params has been read
	bool params_busy = false;

	uint32_t per_line_flags = 0;
	uint32_t h_start_latched = 0, x_scale_latched = 0;
	uint32_t h_start_lines[kViVEndMax] = {}, x_scale_lines[kViVEndMax] = {};
	unsigned per_line_line = 0;
	bool per_line_ended = false;

	void begin_per_scanline(const uint32_t *regs, uint32_t flags);
	void set_for_scanline(uint32_t reg_bit, uint32_t value);
	void latch_for_scanline(unsigned vi_line);
	void end_per_scanline();
	void release();
};

// Scans out the processor's current RDRAM through the VI into out
// (width x height RGBA8 words, byte order R, G, B, A), queued on `stream`
// behind the rendering.  With sync the call waits for the copy to `out`;
// otherwise the copy is still in flight on `stream` when it returns.
// Returns 0 or an RDP_MI_ERROR_* code.
// RDRAM bytes [offset, offset + length) (modulo the RDRAM size) a scanout
// with these registers fetches.
void vi_memory_range(const ViState &vi, const uint32_t *regs, uint32_t &offset, uint32_t &length);
// With upscaling the image is up to (640 x 576) << (2 * scale_log2) words.
int vi_scanout(ViState &vi, const uint32_t *regs, const ViSource &src, const uint8_t *d_gamma, hipStream_t stream,
               const rdp_mi_scanout_options &opts, uint32_t *out, unsigned *width, unsigned *height, bool sync);

// Gamma table (luts.hpp:1034-1555) from its closed form, 256 + 16384 bytes.
void vi_gamma_table(uint8_t *out);

}  // namespace rdp_mi
