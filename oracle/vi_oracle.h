/*
 * vi_oracle.h - CPU restatement of the paraLLEl-RDP video interface (VI)
 * scanout.  TEST INFRASTRUCTURE ONLY (see rdp_oracle.h).
 *
 * Internal interface between rdp_oracle.c (which owns RDRAM and the VI
 * registers) and vi_oracle.c.
 */
#ifndef VI_ORACLE_H
#define VI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "rdp_oracle.h"

/* VideoInterface per-frame state that persists between scanouts
 * (video_interface.hpp:149-163): frame counter, blank tracking, the previous
 * output image (only read with blend_previous_frame) and the per-scanline
 * register latches (video_interface.cpp:1126-1213). */
#define VI_ORACLE_V_END_MAX 621
typedef struct {
	uint32_t frame_count, last_valid_frame_count;
	int previous_frame_blank;
	int prev_valid;
	unsigned prev_width, prev_height;
	uint32_t *prev; /* previous output image, grown on demand */
	size_t prev_cap;

	uint32_t per_line_flags;
	uint32_t h_start_latched, x_scale_latched;
	uint32_t h_start_lines[VI_ORACLE_V_END_MAX], x_scale_lines[VI_ORACLE_V_END_MAX];
	unsigned per_line_line;
	int per_line_ended;
} vi_state_t;

void vi_state_init(vi_state_t *s);

void vi_begin_per_scanline(vi_state_t *s, const uint32_t regs[14], uint32_t flags);
void vi_set_for_scanline(vi_state_t *s, uint32_t reg_bit, uint32_t value);
void vi_latch_for_scanline(vi_state_t *s, unsigned vi_line);
void vi_end_per_scanline(vi_state_t *s);

void vi_state_free(vi_state_t *s);

/* Where the scanout reads the framebuffer.  scale_log2 = 0: rdram / hidden
 * are the native memories.  scale_log2 = L > 0: they are the 4^L sample
 * planes of the upscaled domain (planes of rdram_size and rdram_size / 2
 * bytes), the image comes out 2^L times larger in each direction, and
 * domain_pre(ctx, offset, length, size_log2) is called before the fetch with
 * the memory range the scanout reads (the reference's
 * submit_update_upscaled_domain_external). */
typedef struct {
	const uint8_t *rdram, *hidden;
	uint32_t rdram_size;
	unsigned scale_log2;
	void (*domain_pre)(void *ctx, uint32_t offset, uint32_t length, unsigned size_log2);
	void *ctx;
} vi_source_t;

/* out must hold (640 x 576) << (2 * scale_log2) words. */
int vi_scanout(vi_state_t *s, const uint32_t regs[14], const vi_source_t *src, const rdpo_scanout_options *opts,
               uint32_t *out, unsigned *width, unsigned *height);

#endif
