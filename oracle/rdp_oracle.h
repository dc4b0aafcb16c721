/*
 * rdp_oracle.h - CPU restatement of the paraLLEl-RDP rendering semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the correctness oracle for the MI355X
 * implementation under parallel-rdp_amd/.  Only the test-suite, the smoke
 * check and the CPU-baseline leg of the benchmark may load it; the product
 * path never calls into it.
 *
 * Draw commands are rasterised and shaded as they are decoded; their
 * framebuffer merge is kept in a per-pass pixel cache and written to RDRAM
 * when the render pass ends, at the same points where the reference's
 * renderer ends its passes (see the render-pass section of rdp_oracle.c).
 * It follows the
 * integer semantics of the reference's shaders and host setup code (citations
 * at each function in rdp_oracle.c), evaluated in plain 32-bit two's
 * complement arithmetic (build with -fwrapv).
 *
 * Memory layout matches the reference: RDRAM is an array of host-endian
 * 32-bit words holding big-endian N64 data, so byte address a lives at
 * host byte a^3 and halfword h at host halfword h^1.  Hidden RDRAM holds one
 * byte (2 significant bits) per RDRAM halfword.  TMEM is 2048 halfwords with
 * the same ^1 halfword swizzle.
 */
#ifndef RDP_ORACLE_H
#define RDP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rdpo rdpo;

/* rdram_size must be a power of two (4 or 8 MiB in practice).  Hidden RDRAM
 * is rdram_size/2 bytes.  RDRAM starts zeroed, hidden RDRAM starts at 0x03
 * per byte and TMEM at zero (reference: rdp_device.cpp:235-243). */
rdpo *rdpo_create(uint32_t rdram_size);
void rdpo_destroy(rdpo *o);

/* Upscaled rendering (the reference's COMMAND_PROCESSOR_FLAG_UPSCALING_*,
 * rdp_device.cpp:188-203): factor 2, 4 or 8, before the first command.
 * Returns 0, or -1 for another factor or a second call.  The upscaled
 * domain: factor^2 sample planes of RDRAM (rdram_size bytes each) and of
 * hidden RDRAM (rdram_size / 2 each); sample (sx, sy) of a pixel lives in
 * plane sy * factor + sx at the pixel's native address. */
int rdpo_enable_upscaling(rdpo *o, unsigned factor);
/* COMMAND_PROCESSOR_FLAG_SUPER_SAMPLED_READ_BACK_BIT (and _DITHER_BIT):
 * after rdpo_enable_upscaling, before the first command.  Every pass then
 * renders at scale only and is resolved into native RDRAM
 * (update_upscaled_domain_resolve.comp).  Returns 0, or -1 without
 * upscaling or on a second call. */
int rdpo_enable_super_sampled_readback(rdpo *o, int dither);
uint8_t *rdpo_upscaled_rdram(rdpo *o);
uint8_t *rdpo_upscaled_hidden_rdram(rdpo *o);
unsigned rdpo_upscaling_factor(const rdpo *o);

uint8_t *rdpo_rdram(rdpo *o);
uint8_t *rdpo_hidden_rdram(rdpo *o);
uint16_t *rdpo_tmem(rdpo *o);
uint32_t rdpo_rdram_size(const rdpo *o);

/* Number of 32-bit words the command starting with this word occupies
 * (2..44), or 0 for an opcode the RDP does not define. */
unsigned rdpo_command_length(uint32_t first_word);

/* Process a stream of RDP commands.  Returns the number of words consumed
 * (stops early on a truncated trailing command). */
size_t rdpo_process(rdpo *o, const uint32_t *words, size_t num_words);

/* End the open render pass: write back the colour and depth state of every
 * pixel it touched.  Equivalent to the reference's idle()/flush; RDRAM is
 * only complete after this (or after a SyncFull that closed the pass). */
void rdpo_flush(rdpo *o);

/* Render-pass log: the primitive counter (number of draw commands seen) at
 * the end of every pass that contained a draw.  Returns the total number of
 * entries; copies at most cap of them. */
size_t rdpo_pass_ends(const rdpo *o, uint64_t *out, size_t cap);

/* VI register file (reference: rdp_common.hpp VIRegister). */
void rdpo_set_vi_register(rdpo *o, unsigned reg, uint32_t value);

/* Scanout options (ScanoutOptions, video_interface.hpp:31-80; upscaling,
 * downscaling and external memory are not part of the path). */
typedef struct {
	unsigned crop_overscan_pixels;
	unsigned crop_left, crop_right, crop_top, crop_bottom;
	int crop_enable;
	int persist_frame_on_invalid_input;
	int blend_previous_frame;
	int upscale_deinterlacing;
	int vi_aa, vi_scale, vi_serrate, vi_dither_filter, vi_divot_filter, vi_gamma_dither;
} rdpo_scanout_options;

/* The reference's defaults: no crop, no persistence or blending, upscale
 * deinterlacing (fields come out progressive), every VI filter on. */
void rdpo_scanout_options_default(rdpo_scanout_options *opts);

/* Scan out the current frame through the VI filters (ends the open render
 * pass first).  out must hold 640 x 576 RGBA8 words (byte order R, G, B, A),
 * times factor^2 with upscaling, where the image is read from the sample
 * planes and comes out factor times larger in each direction;
 * width/height receive the output size, 0x0 for a blank or invalid frame.
 * rdpo_scanout uses the default options.  Returns 0. */
int rdpo_scanout(rdpo *o, uint32_t *out, unsigned *width, unsigned *height);
int rdpo_scanout_ex(rdpo *o, const rdpo_scanout_options *opts, uint32_t *out, unsigned *width, unsigned *height);

/* Per-scanline HStart / XScale registers (video_interface.cpp:1126-1213);
 * flags and reg_bit: 1 = HStart, 2 = XScale. */
void rdpo_begin_vi_register_per_scanline(rdpo *o, uint32_t flags);
void rdpo_set_vi_register_for_scanline(rdpo *o, uint32_t reg_bit, uint32_t value);
void rdpo_latch_vi_register_for_scanline(rdpo *o, unsigned vi_line);
void rdpo_end_vi_register_per_scanline(rdpo *o);

/* Instrumentation: number of (pixel, primitive) fragments that reached the
 * depth/blend stage, and number of draw commands seen. */
uint64_t rdpo_fragment_count(const rdpo *o);
/* The same count split by the draw's mode: 0 fill, 1 copy, 2 one-cycle
 * without texels, 3 one-cycle with texels, 4 two-cycle. */
#define RDPO_FRAGMENT_MODES 5
void rdpo_fragment_counts(const rdpo *o, uint64_t out[RDPO_FRAGMENT_MODES]);
uint64_t rdpo_primitive_count(const rdpo *o);

/* Digest of RDRAM, hidden RDRAM and TMEM with the definition the library's
 * rdp_mi_state_digest documents (rdp_mi.h), for per-frame checks. */
uint64_t rdpo_state_digest(const rdpo *o);

/* Generated lookup tables, exposed so tests can pin them. */
void rdpo_blender_divider_table(uint8_t out[0x8000]);
void rdpo_gamma_table(uint8_t out[256 + 256 * 64]);
void rdpo_perspective_table(int16_t out[64][2]);

/* Test hooks onto individual functions of the restatement. */
uint32_t rdpo_test_z_compress(int32_t z);
int32_t rdpo_test_z_decompress(uint32_t z);
int32_t rdpo_test_clamp_9bit(int32_t c);
int32_t rdpo_test_clamp_z(int32_t z);
int32_t rdpo_test_dz_compress_host(int32_t dz);
int32_t rdpo_test_normalize_dzpix(int32_t dz);
int32_t rdpo_test_combiner_equation(int32_t a, int32_t b, int32_t c, int32_t d);
uint32_t rdpo_test_noise(uint32_t x, uint32_t y, uint32_t prim);
void rdpo_test_perspective_divide(int32_t s, int32_t t, int32_t w, int32_t out[3]);
int rdpo_test_coverage(const uint16_t xleft[4], const uint16_t xright[4], int x);
void rdpo_test_rgb_dither(int32_t rgb[3], int dith);
#ifdef __cplusplus
}
#endif

#endif
