/*
 * rdp_oracle.c - CPU restatement of the paraLLEl-RDP rendering semantics.
 *
 * TEST INFRASTRUCTURE ONLY (see rdp_oracle.h).  Nothing in the product path
 * links or calls this file.
 *
 * Immediate-mode structure: each draw command is decoded, its per-scanline
 * span records are computed, and every covered pixel is shaded and merged into
 * RDRAM in primitive order.  Citations are to the reference tree
 * (Themaister/parallel-rdp), path:line.
 *
 * Build with -fwrapv: the reference's GLSL integer arithmetic wraps modulo
 * 2^32 and shifts of negative values are arithmetic.
 */
#include "rdp_oracle.h"
#include "vi_oracle.h"

#include <stdlib.h>
#include <string.h>

/* A build with -DRDPO_PLANT=n carries exactly one deliberate mistake, the
 * one marked PLANTED(n) below (or in vi_oracle.c).  tests/test_oracle_pins.py
 * and tests/test_oracle_upscale.py build such variants to show that the pin
 * of each function fails on a mistake in that function.  The normal build
 * (RDPO_PLANT 0) has none of them. */
#ifndef RDPO_PLANT
#define RDPO_PLANT 0
#endif
#define PLANTED(n) (RDPO_PLANT == (n))

/* ------------------------------------------------------------------------ */
/* Flag bits.  Values follow parallel-rdp/rdp_data_structures.hpp:33-90 so   */
/* that the state words read the same as the reference's.                   */
/* ------------------------------------------------------------------------ */
enum {
	TRI_FLIP = 1 << 0,
	TRI_DO_OFFSET = 1 << 1,
	TRI_SKIP_XFRAC = 1 << 2,
	TRI_INTERLACE_FIELD = 1 << 3,
	TRI_INTERLACE_KEEP_ODD = 1 << 4,
	TRI_DISABLE_UPSCALING = 1 << 5,
	TRI_NATIVE_LOD = 1 << 6,
	TRI_FILL_COPY_RASTER = 1 << 7,
};

enum {
	RAS_INTERLACE_FIELD = 1 << 0,
	RAS_INTERLACE_KEEP_ODD = 1 << 1,
	RAS_AA = 1 << 2,
	RAS_PERSPECTIVE = 1 << 3,
	RAS_TLUT = 1 << 4,
	RAS_TLUT_TYPE = 1 << 5,
	RAS_CVG_TIMES_ALPHA = 1 << 6,
	RAS_ALPHA_CVG_SELECT = 1 << 7,
	RAS_MULTI_CYCLE = 1 << 8,
	RAS_TEX_LOD = 1 << 9,
	RAS_SHARPEN_LOD = 1 << 10,
	RAS_DETAIL_LOD = 1 << 11,
	RAS_FILL = 1 << 12,
	RAS_COPY = 1 << 13,
	RAS_SAMPLE_QUAD = 1 << 14,
	RAS_ALPHA_TEST = 1 << 15,
	RAS_ALPHA_TEST_DITHER = 1 << 16,
	RAS_MID_TEXEL = 1 << 17,
	RAS_BILERP0 = 1 << 23,
	RAS_BILERP1 = 1 << 24,
	RAS_CONVERT_ONE = 1 << 22,
	/* Derived per draw by the host (deduce_static_texture_state and
	 * deduce_noise_state); part of the state-cache key only. */
	RAS_USES_TEXEL0 = 1 << 18,
	RAS_USES_TEXEL1 = 1 << 19,
	RAS_USES_LOD = 1 << 20,
	RAS_USES_PIPELINED_TEXEL1 = 1 << 21,
	RAS_NEED_NOISE_DUAL = 1 << 25,
	RAS_NEED_NOISE = 1 << 28,
	RAS_USE_STATIC_TEXTURE_SIZE_FORMAT = 1 << 29,
};

enum {
	DB_DEPTH_TEST = 1 << 0,
	DB_DEPTH_UPDATE = 1 << 1,
	DB_FORCE_BLEND = 1 << 3,
	DB_IMAGE_READ = 1 << 4,
	DB_COLOR_ON_CVG = 1 << 5,
	DB_MULTI_CYCLE = 1 << 6,
	DB_AA = 1 << 7,
	DB_DITHER = 1 << 8,
};

enum { FB_I4 = 0, FB_I8 = 1, FB_RGBA5551 = 2, FB_IA88 = 3, FB_RGBA8888 = 4 };
enum { TF_RGBA = 0, TF_YUV = 1, TF_CI = 2, TF_IA = 3, TF_I = 4 };
enum { UP_TILE = 0, UP_TLUT = 1, UP_BLOCK = 2 };
enum { TILE_CLAMP_S = 1, TILE_MIRROR_S = 2, TILE_CLAMP_T = 4, TILE_MIRROR_T = 8 };

#define MAX_FB_LINES 1024
#define MAX_FB_WIDTH 1024

#define COVERAGE_FILL 0x40
#define COVERAGE_COPY 0x20

typedef struct {
	int32_t xh, xm, xl;
	int32_t yh, ym, yl;
	int32_t dxhdy, dxmdy, dxldy;
	uint32_t flags;
	uint32_t tile;
} tri_t;

typedef struct {
	int32_t rgba[4], drgba_dx[4], drgba_de[4], drgba_dy[4];
	int32_t stzw[4], dstzw_dx[4], dstzw_de[4], dstzw_dy[4];
} attr_t;

typedef struct {
	uint32_t slo, shi, tlo, thi;
	uint32_t offset, stride;
	uint32_t fmt, size, palette;
	uint32_t mask_s, shift_s, mask_t, shift_t;
	uint32_t flags;
} tile_t;

typedef struct {
	int32_t rgba[4], stzw[4];
	int32_t xleft[4], xright[4]; /* stored as u16 in the reference */
	int32_t base_x, start_x, end_x;
	int32_t lodlength;            /* stored as i16 */
	int32_t valid;
} span_t;

/* Per-draw resolved state (what the reference packs into DerivedSetup,
 * StaticRasterizationState and DepthBlendState). */
typedef struct {
	uint32_t ras_flags;
	uint32_t dither;
	uint8_t comb[2][8]; /* rgb muladd, mulsub, mul, add; alpha muladd, mulsub, mul, add */
	int32_t const_muladd[2][4], const_mulsub[2][4], const_mul[2][4], const_add[2][4];
	uint8_t fog[4], blendc[4];
	uint32_t fill_color;
	int32_t dz, dz_compressed, min_lod;
	int32_t factors[4];
	uint32_t db_flags;
	uint8_t blend_modes[2][4]; /* 1a, 1b, 2a, 2b */
	uint32_t coverage_mode, z_mode;
	uint32_t prim_index; /* global primitive counter, seeds noise */
	int uses_texel0, uses_texel1, uses_pipelined_texel1, uses_lod;
	int need_noise, need_noise_dual;
} draw_t;

/* Keys of the reference renderer's per-pass state caches
 * (StaticRasterizationState, DepthBlendState, TileInfo;
 * rdp_data_structures.hpp:154-172, rdp_common.hpp:264-290).  Two draws share
 * a cache slot exactly when their keys compare equal byte for byte. */
typedef struct {
	uint8_t comb[2][8];
	uint32_t flags, dither, texture_size, texture_fmt;
} static_key_t;

typedef struct {
	uint8_t blend[2][4];
	uint32_t flags, coverage_mode, z_mode;
} depth_blend_key_t;

/* Bookkeeping of the open render pass, mirroring what Renderer::need_flush
 * and tmem_upload_needs_flush read (rdp_renderer.cpp:1464-1590, 3051-3098,
 * 3484-3485). */
typedef struct {
	static_key_t static_keys[64];
	depth_blend_key_t depth_blend_keys[64];
	tile_t tile_keys[256];
	unsigned num_static, num_depth_blend, num_tiles;
	unsigned num_prims, num_span_jobs, num_uploads;
	uint32_t max_shaded_tiles;
	uint32_t deduced_height;
	int color_write_pending, depth_write_pending;
} pass_track_t;

struct pixel_state;

struct rdpo {
	uint8_t *rdram;
	uint8_t *hidden;
	uint16_t tmem[2048];
	uint32_t rdram_size;

	struct { uint32_t xlo, ylo, xhi, yhi; } scissor;
	uint32_t ras_flags, dither;
	uint8_t comb[2][8];
	uint32_t db_flags;
	uint8_t blend[2][4];
	uint32_t coverage_mode, z_mode;

	struct { uint32_t addr, width, fmt, size; } teximg;
	tile_t tiles[8];

	uint32_t blend_color, fog_color, env_color, prim_color, fill_color;
	uint32_t min_level, prim_lod_frac;
	int32_t prim_depth;
	uint32_t prim_dz;
	int use_prim_depth;
	int32_t convert[6];
	uint32_t key_width[3], key_center[3], key_scale[3];

	uint32_t fb_addr, fb_width, fb_fmt, depth_addr;
	uint32_t quirks;

	uint64_t prim_counter;
	uint64_t fragments;
	uint64_t fragments_mode[RDPO_FRAGMENT_MODES]; /* by rdpo_fragment_counts' modes */

	uint32_t vi_regs[14];

	uint8_t blender_lut[0x8000];

	/* Open render pass: colour/depth state of every pixel it has touched,
	 * indexed y * 1024 + x, loaded on first touch and written back when the
	 * pass ends (the reference's init_tile/finish_tile, memory_interfacing.h:
	 * 286-362). */
	pass_track_t pass;
	struct pixel_state *px;
	uint8_t *px_loaded;
	uint32_t *px_touched;
	size_t num_touched;

	/* Upscaled domain (rdpo_enable_upscaling; the reference's
	 * Renderer::init_internal_upscaling_factor, rdp_renderer.cpp:545-601):
	 * up_log2 > 0 renders every pass whose framebuffer is RGBA5551 or
	 * RGBA8888 and at least 256 pixels wide a second time at 2^up_log2 times
	 * the resolution.  Sample (sx, sy) of a native pixel lives in plane
	 * sy * S + sx of up_rdram (planes of rdram_size bytes) and up_hidden
	 * (planes of rdram_size / 2); up_ref is the native RDRAM the domain last
	 * agreed with.  upx[native key] holds the open pass's S x S sample states
	 * of a native pixel, allocated on first touch. */
	uint32_t up_log2;
	uint8_t *up_rdram, *up_hidden, *up_ref;
	struct pixel_state **upx;
	uint32_t *upx_touched;
	size_t upx_num;
	/* Draws of the open pass, kept with the state they were drawn with
	 * (tile descriptors and TMEM included) when the pass qualifies for the
	 * upscaled render; replayed at scale when the pass ends. */
	struct up_draw *up_draws;
	unsigned up_num;
	/* Super-sampled readback (rdpo_enable_super_sampled_readback): every
	 * pass renders at scale only and is resolved into native RDRAM;
	 * up_wmask[native key] holds bit 0 colour written, bit 1 depth written
	 * by any sample of the pass. */
	int up_ssaa, up_ssaa_dither;
	uint8_t *up_wmask;

	/* Video interface state (frame counter, previous image, per-scanline
	 * register latches). */
	vi_state_t *vi;

	/* Primitive counter at the end of every pass that held a draw. */
	uint64_t *pass_log;
	size_t pass_log_len, pass_log_cap;
};

/* ------------------------------------------------------------------------ */
/* Small integer helpers (GLSL semantics).                                  */
/* ------------------------------------------------------------------------ */
static inline int32_t sext(int32_t v, int bits)
{
	return (int32_t)((uint32_t)v << (32 - bits)) >> (32 - bits);
}

static inline int find_msb(int32_t v)
{
	if (v < 0)
		v = ~v;
	if (v == 0)
		return -1;
	return 31 - __builtin_clz((uint32_t)v);
}

static inline int find_lsb(uint32_t v)
{
	return v ? __builtin_ctz(v) : -1;
}

static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t iclamp(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------------------ */
/* Lookup tables, generated from closed forms.                              */
/* ------------------------------------------------------------------------ */

/* Blender divider (reference data: luts.hpp:7-1032, used at blender.h:133-139).
 * The table is the output of a bit-serial non-restoring divider: 11-bit
 * dividend, 4-bit divisor, 3-bit partial remainder, 8 quotient bits.  The
 * first cycle always adds the divisor back regardless of the sign of the
 * initial trial subtraction, which is what produces the wrap-around values
 * for quotients >= 256 and the skew for divisors >= 9. */
void rdpo_blender_divider_table(uint8_t out[0x8000])
{
	for (unsigned i = 0; i < 0x8000; i++) {
		unsigned divisor = (i >> 11) & 0xf;
		unsigned dividend = i & 0x7ff;
		unsigned neg_divisor = (~divisor) & 0xf;
		unsigned partial = (neg_divisor + (dividend >> 8) + 1) & 7;
		unsigned quotient = 0;
		unsigned prev_bit = 0;
		for (unsigned k = 0; k < 8; k++) {
			unsigned next_bit = (dividend >> (7 - k)) & 1;
			unsigned t;
			if (prev_bit)
				t = neg_divisor + (partial << 1) + next_bit + 1;
			else
				t = divisor + (partial << 1) + next_bit;
			partial = t & 7;
			prev_bit = (t >> 4) & 1;
			quotient |= prev_bit << (7 - k);
		}
		out[i] = (uint8_t)quotient;
	}
}

static unsigned isqrt_u(unsigned v)
{
	unsigned r = 0;
	while ((r + 1) * (r + 1) <= v)
		r++;
	return r;
}

/* VI gamma table (reference data: luts.hpp:1034-1555).  Entries 0..255 are
 * the 8-bit gamma curve, sqrt(c/255)*255 truncated to even; the next 16384
 * entries serve gamma-dither with index (c << 6) + 6-bit noise. */
void rdpo_gamma_table(uint8_t out[256 + 256 * 64])
{
	for (unsigned i = 0; i < 256; i++)
		out[i] = (uint8_t)(isqrt_u(i * 256) & ~1u);
	for (unsigned j = 0; j < 256 * 64; j++)
		out[256 + j] = (uint8_t)(isqrt_u(j * 4) & ~1u);
}

static inline int32_t persp_base(int i);

/* Perspective reciprocal table (reference data: perspective.h:27-44).
 * Entry i: r_i = round(2^20 / (64 + i)), except r_6 which is one lower, and
 * slope -4 * (r_i - r_{i+1}). */
void rdpo_perspective_table(int16_t out[64][2])
{
	int32_t r[65];
	for (int i = 0; i <= 64; i++)
		r[i] = persp_base(i);
	for (int i = 0; i < 64; i++) {
		out[i][0] = (int16_t)r[i];
		out[i][1] = (int16_t)(-4 * (r[i] - r[i + 1]));
	}
}


/* Ordered dither matrices, dither.h:26-28 (magic square, Bayer). */
static const uint8_t dither_matrices[2][16] = {
	{ 0, 6, 1, 7, 4, 2, 5, 3, 3, 5, 2, 4, 7, 1, 6, 0 },
	{ 0, 4, 1, 5, 4, 0, 5, 1, 3, 7, 2, 6, 7, 3, 6, 2 },
};

/* ------------------------------------------------------------------------ */
/* RDRAM / TMEM access with the 32-bit host-endian swizzle.                 */
/* ------------------------------------------------------------------------ */
static inline uint8_t rd8(const rdpo *o, uint32_t addr)
{
	return o->rdram[(addr & (o->rdram_size - 1)) ^ 3];
}

static inline uint16_t rd16_index(const rdpo *o, uint32_t index)
{
	uint32_t mask = (o->rdram_size >> 1) - 1;
	const uint16_t *p = (const uint16_t *)o->rdram;
	return p[(index & mask) ^ 1];
}

static inline void wr16_index(rdpo *o, uint32_t index, uint16_t v)
{
	uint32_t mask = (o->rdram_size >> 1) - 1;
	uint16_t *p = (uint16_t *)o->rdram;
	p[(index & mask) ^ 1] = v;
}

static inline uint32_t rd32_index(const rdpo *o, uint32_t index)
{
	uint32_t mask = (o->rdram_size >> 2) - 1;
	const uint32_t *p = (const uint32_t *)o->rdram;
	return p[index & mask];
}

static inline void wr32_index(rdpo *o, uint32_t index, uint32_t v)
{
	uint32_t mask = (o->rdram_size >> 2) - 1;
	uint32_t *p = (uint32_t *)o->rdram;
	p[index & mask] = v;
}

/* 16-bit read at an arbitrary byte address (tmem_update.comp:441-445). */
static inline uint32_t rd16_addr(const rdpo *o, uint32_t addr)
{
	if ((addr & 1) == 0)
		return rd16_index(o, addr >> 1);
	return ((uint32_t)rd8(o, addr) << 8) | rd8(o, addr + 1);
}

static inline uint32_t tmem16(const rdpo *o, uint32_t index)
{
	return o->tmem[index & 0x7ff];
}

static inline uint32_t tmem8(const rdpo *o, uint32_t index)
{
	const uint8_t *p = (const uint8_t *)o->tmem;
	return p[index & 0xfff];
}

/* ------------------------------------------------------------------------ */
/* Creation.                                                                */
/* ------------------------------------------------------------------------ */
rdpo *rdpo_create(uint32_t rdram_size)
{
	if (rdram_size < 0x1000 || (rdram_size & (rdram_size - 1)))
		return NULL;
	rdpo *o = (rdpo *)calloc(1, sizeof(*o));
	if (!o)
		return NULL;
	o->rdram_size = rdram_size;
	o->rdram = (uint8_t *)calloc(rdram_size, 1);
	o->hidden = (uint8_t *)malloc(rdram_size / 2);
	if (!o->rdram || !o->hidden) {
		rdpo_destroy(o);
		return NULL;
	}
	memset(o->hidden, 0x03, rdram_size / 2);
	o->vi = (vi_state_t *)malloc(sizeof(vi_state_t));
	if (!o->vi) {
		rdpo_destroy(o);
		return NULL;
	}
	vi_state_init(o->vi);
	o->quirks = 1; /* native_resolution_tex_rect defaults on (rdp_device.hpp:80) */
	rdpo_blender_divider_table(o->blender_lut);
	return o;
}

static void upx_release(rdpo *o);

int rdpo_enable_upscaling(rdpo *o, unsigned factor)
{
	unsigned l = factor == 2 ? 1 : factor == 4 ? 2 : factor == 8 ? 3 : 0;
	if (!l || o->up_log2 || o->prim_counter)
		return -1;
	size_t planes = (size_t)1 << (2 * l);
	o->up_rdram = (uint8_t *)calloc(planes, o->rdram_size);
	o->up_hidden = (uint8_t *)malloc(planes * (o->rdram_size / 2));
	o->up_ref = (uint8_t *)calloc(o->rdram_size, 1);
	o->upx = (struct pixel_state **)calloc((size_t)MAX_FB_LINES * MAX_FB_WIDTH, sizeof(*o->upx));
	o->upx_touched = (uint32_t *)malloc(sizeof(uint32_t) * MAX_FB_LINES * MAX_FB_WIDTH);
	if (!o->up_rdram || !o->up_hidden || !o->up_ref || !o->upx || !o->upx_touched)
		abort();
	/* The upscaled hidden RDRAM starts at 0x03 like the native one
	 * (rdp_renderer.cpp:588-593); the domain and its reference start equal
	 * to the (zeroed) native RDRAM. */
	memset(o->up_hidden, 0x03, planes * (o->rdram_size / 2));
	o->up_log2 = l;
	return 0;
}

int rdpo_enable_super_sampled_readback(rdpo *o, int dither)
{
	if (!o->up_log2 || o->up_ssaa || o->prim_counter)
		return -1;
	o->up_wmask = (uint8_t *)calloc((size_t)MAX_FB_LINES * MAX_FB_WIDTH, 1);
	if (!o->up_wmask)
		abort();
	o->up_ssaa = 1;
	o->up_ssaa_dither = dither != 0;
	return 0;
}

uint8_t *rdpo_upscaled_rdram(rdpo *o) { return o->up_rdram; }
uint8_t *rdpo_upscaled_hidden_rdram(rdpo *o) { return o->up_hidden; }
unsigned rdpo_upscaling_factor(const rdpo *o) { return 1u << o->up_log2; }

void rdpo_destroy(rdpo *o)
{
	if (!o)
		return;
	if (o->upx)
		upx_release(o);
	free(o->up_draws);
	free(o->up_wmask);
	free(o->upx);
	free(o->upx_touched);
	free(o->up_rdram);
	free(o->up_hidden);
	free(o->up_ref);
	free(o->rdram);
	free(o->hidden);
	free(o->px);
	free(o->px_loaded);
	free(o->px_touched);
	free(o->pass_log);
	if (o->vi)
		vi_state_free(o->vi);
	free(o->vi);
	free(o);
}

uint8_t *rdpo_rdram(rdpo *o) { return o->rdram; }
uint8_t *rdpo_hidden_rdram(rdpo *o) { return o->hidden; }
uint16_t *rdpo_tmem(rdpo *o) { return o->tmem; }
uint32_t rdpo_rdram_size(const rdpo *o) { return o->rdram_size; }
uint64_t rdpo_fragment_count(const rdpo *o) { return o->fragments; }
void rdpo_fragment_counts(const rdpo *o, uint64_t out[RDPO_FRAGMENT_MODES])
{
	memcpy(out, o->fragments_mode, sizeof(o->fragments_mode));
}
uint64_t rdpo_primitive_count(const rdpo *o) { return o->prim_counter; }

/* ------------------------------------------------------------------------ */
/* Span setup: span_setup.comp:87-227.  L is SCALING_LOG2: 0 for the native */
/* pass, log2 of the upscaling factor for the upscaled pass, where y counts */
/* upscaled lines.                                                          */
/* ------------------------------------------------------------------------ */
static int32_t quantize_x(int32_t x)
{
	int32_t sticky = (x & 0xfff) != 0;
	if (PLANTED(1)) sticky = 0;
	return (x >> 12) | sticky;
}

/* interpolate_snapped (span_setup.comp:80-85): whole native lines step by
 * dvalue, the remaining sub-lines by dvalue >> L. */
static inline int32_t interpolate_snapped(int32_t dvalue, int dy, int L)
{
	return (dy >> L) * dvalue + (dy & ((1 << L) - 1)) * (dvalue >> L);
}

static void span_setup(const tri_t *setup, const attr_t *attr, const uint32_t scissor[4], int y, span_t *s, int L)
{
	const int S = 1 << L;
	/* scissor: xlo, ylo, xhi, yhi */
	int flip = (setup->flags & TRI_FLIP) != 0;
	int interlace_en = (setup->flags & TRI_INTERLACE_FIELD) != 0;
	int keep_odd = (setup->flags & TRI_INTERLACE_KEEP_ODD) != 0;

	/* Attribute interpolation to the scanline origin (span_setup.comp:108-149). */
	{
		int do_offset = (setup->flags & TRI_DO_OFFSET) != 0;
		int skip_xfrac = (setup->flags & TRI_SKIP_XFRAC) != 0;
		int y_base = (setup->yh >> 2) * S;
		int dy = y - y_base;
		int32_t xh = setup->xh * S + dy * (setup->dxhdy << 2);
		int32_t drgba_diff[4] = { 0, 0, 0, 0 }, dstzw_diff[4] = { 0, 0, 0, 0 };
		if (do_offset) {
			xh += (S * 3) * setup->dxhdy;
			for (int i = 0; i < 4; i++) {
				int32_t deh = attr->drgba_de[i] & ~0x1ff;
				int32_t dyh = attr->drgba_dy[i] & ~0x1ff;
				drgba_diff[i] = deh - (deh >> 2) - dyh + (dyh >> 2);
				deh = attr->dstzw_de[i] & ~0x1ff;
				dyh = attr->dstzw_dy[i] & ~0x1ff;
				dstzw_diff[i] = deh - (deh >> 2) - dyh + (dyh >> 2);
			}
		}
		int32_t base_x = xh >> 15;
		int32_t xfrac = skip_xfrac ? 0 : ((xh >> (PLANTED(3) ? 8 : 7)) & 0xff);
		for (int i = 0; i < 4; i++) {
			int32_t v = attr->rgba[i] + interpolate_snapped(attr->drgba_de[i], dy, L);
			s->rgba[i] = ((v & ~0x1ff) + drgba_diff[i] - interpolate_snapped((attr->drgba_dx[i] >> 8) & ~1, xfrac, L)) &
			             ~0x3ff;
			v = attr->stzw[i] + interpolate_snapped(attr->dstzw_de[i], dy, L);
			s->stzw[i] = ((v & ~0x1ff) + dstzw_diff[i] - interpolate_snapped((attr->dstzw_dx[i] >> 8) & ~1, xfrac, L)) &
			             ~0x3ff;
		}
		s->base_x = base_x;
	}

	/* X extents at the four sub-scanlines (span_setup.comp:151-225). */
	int yh_base = (PLANTED(2) ? setup->yh : (setup->yh & ~3)) * S;
	int ym_base = setup->ym * S;
	int ylo = imax(setup->yh, (int)scissor[1]) * S;
	int yhi = imin(setup->yl, (int)scissor[3]) * S;
	int32_t xleft[4], xright[4], invalid[4];
	int32_t lo_sc = S * (int32_t)(scissor[0] << 1), hi_sc = S * (int32_t)(scissor[2] << 1);
	int all_over = 1, all_under = 1, all_invalid = 1;
	for (int i = 0; i < 4; i++) {
		int ys = y * 4 + i;
		int clip_y = ys < ylo || ys >= yhi;
		int32_t xh = setup->xh * (PLANTED(31) ? 1 : S) + (ys - yh_base) * setup->dxhdy;
		int32_t xm = setup->xm * S + (ys - yh_base) * setup->dxmdy;
		int32_t xl = setup->xl * S + (ys - ym_base) * setup->dxldy;
		if (ys < S * setup->ym)
			xl = xm;
		xl = sext(xl, 27 + L);
		xh = sext(xh, 27 + L);
		int32_t xh_s = quantize_x(xh), xl_s = quantize_x(xl);
		int32_t l = flip ? xh_s : xl_s;
		int32_t r = flip ? xl_s : xh_s;
		invalid[i] = (l >> 1) > (r >> 1);
		if (!(imin(l, r) >= hi_sc))
			all_over = 0;
		if (!(imax(l, r) < lo_sc))
			all_under = 0;
		/* max() against the low edge, then min() against the high edge. */
		l = imin(imax(l, lo_sc), hi_sc);
		r = imin(imax(r, lo_sc), hi_sc);
		invalid[i] |= clip_y;
		if (invalid[i]) {
			l = 0xffff;
			r = 0;
		} else
			all_invalid = 0;
		xleft[i] = l;
		xright[i] = r;
	}
	int32_t start_x = imin(imin(xleft[0], xleft[1]), imin(xleft[2], xleft[3])) >> 3;
	int32_t end_x = imax(imax(xright[0], xright[1]), imax(xright[2], xright[3])) >> 3;
	for (int i = 0; i < 4; i++) {
		s->xleft[i] = xleft[i] & 0xffff;
		s->xright[i] = xright[i] & 0xffff;
	}
	s->start_x = start_x;
	s->end_x = end_x;
	s->valid = !all_invalid && !all_over && !all_under;
	if (interlace_en && (((y >> L) & 1) != keep_odd))
		s->valid = 0;
	s->lodlength = (int32_t)(int16_t)(flip ? (end_x - s->base_x) : (s->base_x - start_x));
}

/* ------------------------------------------------------------------------ */
/* Coverage, interpolation, clamping: coverage.h, interpolation.h,          */
/* clamping.h.                                                              */
/* ------------------------------------------------------------------------ */
static int compute_coverage(const span_t *s, int x)
{
	static const int offs[4] = { 0, 4, 2, 6 };
	int clip = 0;
	for (int k = 0; k < 8; k++) {
		int line = k >> 1;
		int xs = offs[(k & 1) | ((line & 1) << 1)] + (x << 3);
		if (xs < s->xleft[line] || xs >= s->xright[line])
			clip |= 1 << k;
	}
	return ~clip & 0xff;
}

static inline int32_t clamp_9bit(int32_t c)
{
	return iclamp(sext(c - 0x80, 9) + 0x80, 0, 0xff);
}

static inline int32_t clamp_z(int32_t z)
{
	z -= 1 << 17;
	z = sext(z, 19);
	z += 1 << 17;
	return iclamp(z, 0, 0x3ffff);
}

/* interpolation.h:30-53; L is SCALING_LOG2.  The reference keeps the
 * snapped value in 16 bits; only the low 9 bits of the result reach
 * clamp_9bit, and they are the same in 32-bit arithmetic. */
static void interpolate_rgba(const span_t *s, const attr_t *a, int dx, int coverage, int32_t out[4], int L)
{
	int first = find_lsb((uint32_t)coverage);
	int yoff = first >> 1;
	int xoff = ((first & 1) << 1) + (PLANTED(4) ? 0 : (yoff & 1));
	for (int i = 0; i < 4; i++) {
		int32_t v = s->rgba[i] + ((a->drgba_dx[i] & ~0x1f) >> (PLANTED(30) ? 0 : L)) * dx;
		int32_t snapped = v >> 14;
		snapped <<= 2 + L;
		snapped += xoff * (a->drgba_dx[i] >> 14) + yoff * (a->drgba_dy[i] >> 14);
		snapped >>= 4 + L;
		out[i] = clamp_9bit(snapped);
	}
}

/* perspective.h:46-112 */
static inline int32_t persp_base(int i)
{
	/* round(2^20 / (64 + i)); the hardware table's entry 6 is one lower. */
	return ((1 << 20) + (64 + i) / 2) / (64 + i) - (i == 6);
}

static void perspective_get_lut(int32_t w, int32_t *rcp, int32_t *shift)
{
	int sh = imin(14 - find_msb(w), 14);
	int32_t normout = (w << sh) & 0x3fff;
	int32_t wnorm = normout & 0xff;
	int i = normout >> 8;
	int32_t base = persp_base(i);
	int32_t slope = -4 * (base - persp_base(i + 1));
	*rcp = ((slope * wnorm) >> 10) + base;
	*shift = sh;
}

static void perspective_divide(const int32_t stw[3], int32_t st[2], int *overflow)
{
	int32_t w = stw[2];
	int w_carry = w <= 0;
	w &= 0x7fff;
	int32_t rcp, shift;
	perspective_get_lut(w, &rcp, &shift);
	int32_t prod[2] = { stw[0] * rcp, stw[1] * rcp };
	int32_t temp_mask = ((1 << 30) - 1) & -((1 << 29) >> shift);
	int32_t oob[2] = { prod[0] & temp_mask, prod[1] & temp_mask };
	int32_t temp[2];
	if (shift != 14) {
		prod[0] >>= (13 - shift);
		prod[1] >>= (13 - shift);
		temp[0] = prod[0];
		temp[1] = prod[1];
	} else {
		temp[0] = prod[0] << 1;
		temp[1] = prod[1] << 1;
	}
	for (int i = 0; i < 2; i++) {
		if (oob[i] != temp_mask && oob[i] != 0) {
			temp[i] = (prod[i] & (1 << 29)) == 0 ? 0x7fff : -0x8000;
			*overflow = 1;
		}
	}
	if (w_carry) {
		temp[0] = temp[1] = 0x7fff;
		*overflow = 1;
	}
	st[0] = iclamp(temp[0], -0x10000, 0xffff);
	st[1] = iclamp(temp[1], -0x10000, 0xffff);
}

static void stw_to_st(const int32_t stw[3], int perspective, int32_t st[2], int *overflow)
{
	int32_t v[3] = { stw[0] >> 16, stw[1] >> 16, stw[2] >> 16 };
	if (perspective)
		perspective_divide(v, st, overflow);
	else {
		st[0] = v[0];
		st[1] = v[1];
	}
}

/* interpolation.h:93-142; L is SCALING_LOG2. */
static void interpolate_stz(const span_t *s, const attr_t *a, int dx, int coverage, int perspective, int uses_lod,
                            int flip_dir, int32_t st[2], int32_t st_dx[2], int32_t st_dy[2], int32_t *z, int *ovf, int L)
{
	int32_t stw[3], stw_dx[3], stw_dy[3];
	static const int comp[3] = { 0, 1, 3 };
	for (int i = 0; i < 3; i++) {
		int c = comp[i];
		stw[i] = s->stzw[c] + ((a->dstzw_dx[c] & ~0x1f) >> (PLANTED(37) ? 0 : L)) * dx;
		if (uses_lod) {
			stw_dx[i] = stw[i] + flip_dir * ((a->dstzw_dx[c] & ~0x1f) >> L);
			if (L)
				stw_dy[i] = stw[i] + (flip_dir < 0 ? -flip_dir : flip_dir) * ((a->dstzw_dy[c] & ~0x7fff) >> L);
			else
				stw_dy[i] = stw[i] + (a->dstzw_dy[c] & ~0x7fff);
		}
	}
	stw_to_st(stw, perspective, st, ovf);
	if (uses_lod) {
		stw_to_st(stw_dx, perspective, st_dx, ovf);
		stw_to_st(stw_dy, perspective, st_dy, ovf);
	}

	/* Whole native pixels step by the full gradient, sub-pixels by the
	 * gradient >> L. */
	int32_t zz = s->stzw[2] + a->dstzw_dx[2] * (dx >> L) + (a->dstzw_dx[2] >> L) * (dx & ((1 << L) - 1));
	int32_t snapped = zz >> 10;
	int first = find_lsb((uint32_t)coverage);
	int yoff = first >> 1;
	int xoff = ((first & 1) << 1) + (yoff & 1);
	snapped <<= 2 + L;
	snapped += xoff * (a->dstzw_dx[2] >> 10) + yoff * (a->dstzw_dy[2] >> 10);
	snapped >>= (PLANTED(5) ? 4 : 5) + L;
	*z = clamp_z(snapped);
}

static void interpolate_st_single(const span_t *s, const attr_t *a, int dx, int perspective, int32_t st[2], int L)
{
	int32_t stw[3];
	static const int comp[3] = { 0, 1, 3 };
	for (int i = 0; i < 3; i++)
		stw[i] = s->stzw[comp[i]] + ((a->dstzw_dx[comp[i]] & ~0x1f) >> L) * dx;
	int ovf = 0;
	stw_to_st(stw, perspective, st, &ovf);
}

/* ------------------------------------------------------------------------ */
/* Noise: noise.h:29-69.                                                    */
/* ------------------------------------------------------------------------ */
static uint32_t reseed_noise(uint32_t x, uint32_t y, uint32_t prim)
{
	const uint32_t P = 1103515245u;
	uint32_t s0 = x, s1 = y, s2 = prim;
	for (int r = 0; r < 3; r++) {
		uint32_t n0 = ((s0 >> 8) ^ s1) * P;
		uint32_t n1 = ((s1 >> 8) ^ s2) * P;
		uint32_t n2 = ((s2 >> 8) ^ s0) * P;
		s0 = n0;
		s1 = n1;
		s2 = n2;
	}
	return (s0 >> 16) & 0xffff;
}

/* ------------------------------------------------------------------------ */
/* Texture sampling: texture.h.                                             */
/* ------------------------------------------------------------------------ */
static int texel_mask_s(const tile_t *t, int s)
{
	if (t->mask_s) {
		int mask = 1 << t->mask_s;
		if (t->flags & TILE_MIRROR_S)
			s ^= PLANTED(9) ? (s & mask) : imax((s & mask) - 1, 0);
		s &= mask - 1;
	}
	return s;
}

static int texel_mask_t(const tile_t *t, int v)
{
	if (t->mask_t) {
		int mask = 1 << t->mask_t;
		if (t->flags & TILE_MIRROR_T)
			v ^= imax((v & mask) - 1, 0);
		v &= mask - 1;
	}
	return v;
}

static void convert_rgba16(uint32_t w, int32_t out[4])
{
	for (int i = 0; i < 3; i++) {
		uint32_t c = (w >> (11 - 5 * i)) & 31;
		out[i] = (int32_t)((c << 3) | (PLANTED(6) ? 0 : (c >> 2)));
	}
	out[3] = (w & 1) ? 0xff : 0;
}

static void convert_ia16(uint32_t w, int32_t out[4])
{
	out[0] = out[1] = out[2] = (int32_t)(w >> 8);
	out[3] = (int32_t)(w & 0xff);
}

static inline void splat4(int32_t out[4], int32_t v) { out[0] = out[1] = out[2] = out[3] = v; }

/* Per-format texel fetches, texture.h:90-315. s/t are already masked. */
static uint32_t texel_index8(const tile_t *t, uint32_t s, uint32_t v, uint32_t wrap, uint32_t add)
{
	uint32_t byte_offset = t->offset + t->stride * v + add;
	byte_offset &= wrap;
	uint32_t index = byte_offset ^ ((v & 1) << 2);
	(void)s;
	return index ^ 3;
}

static uint32_t texel_index16(const tile_t *t, uint32_t s, uint32_t v, uint32_t wrap)
{
	uint32_t byte_offset = t->offset + t->stride * v + s * 2;
	byte_offset &= wrap;
	uint32_t index = byte_offset >> 1;
	if (!PLANTED(8)) index ^= (v & 1) << 1;
	return index ^ 1;
}

static void fetch_texel(const rdpo *o, const tile_t *t, uint32_t s, uint32_t v, int tlut, int tlut_type,
                        uint32_t lut_offset, uint32_t addr_xor, int32_t out[4])
{
	uint32_t fmt = t->fmt, size = t->size;
	if (tlut) {
		uint32_t word, lut_entry;
		if (size == 0) {
			uint32_t idx = texel_index8(t, s, v, 0x7ff, s >> 1);
			uint32_t shift = (~s & 1) * 4;
			word = (tmem8(o, idx) >> shift) & 0xf;
			word |= t->palette << 4;
			lut_entry = (word << 2) + lut_offset;
		} else if (size == 1) {
			uint32_t idx = texel_index8(t, s, v, 0x7ff, s);
			word = tmem8(o, idx);
			lut_entry = (word << 2) + lut_offset;
		} else {
			uint32_t idx = texel_index16(t, s, v, 0x7ff);
			word = tmem16(o, idx);
			lut_entry = ((word >> 6) & ~3u) + lut_offset;
		}
		if (!PLANTED(20)) lut_entry ^= addr_xor;
		word = tmem16(o, 0x400 | lut_entry);
		if (tlut_type)
			convert_ia16(word, out);
		else
			convert_rgba16(word, out);
		return;
	}

	switch (fmt) {
	case TF_RGBA:
		if (size == 0)
			goto i4;
		if (size == 1)
			goto i8;
		if (size == 2) {
			convert_rgba16(tmem16(o, texel_index16(t, s, v, 0xfff)), out);
			return;
		} else {
			uint32_t idx = texel_index16(t, s, v, 0x7ff);
			uint32_t lo = tmem16(o, idx), hi = tmem16(o, idx | 0x400);
			out[0] = (int32_t)(lo >> 8);
			out[1] = (int32_t)(lo & 0xff);
			out[2] = (int32_t)(hi >> 8);
			out[3] = (int32_t)(hi & 0xff);
			return;
		}
	case TF_CI:
		if (size == 0) {
			uint32_t idx = texel_index8(t, s, v, 0xfff, s >> 1);
			uint32_t w = (tmem8(o, idx) >> ((~s & 1) * 4)) & 0xf;
			w |= t->palette << 4;
			splat4(out, (int32_t)w);
			return;
		}
		if (size == 1)
			goto i8;
		goto ci32;
	case TF_IA:
		if (size == 0) {
			uint32_t idx = texel_index8(t, s, v, 0xfff, s >> 1);
			uint32_t w = (tmem8(o, idx) >> ((~s & 1) * 4)) & 0xf;
			uint32_t in = w & 0xe;
			in = (in << 4) | (in << 1) | (PLANTED(7) ? 0 : (in >> 2));
			out[0] = out[1] = out[2] = (int32_t)in;
			out[3] = (w & 1) ? 0xff : 0;
			return;
		}
		if (size == 1) {
			uint32_t w = tmem8(o, texel_index8(t, s, v, 0xfff, s));
			uint32_t in = w >> 4, al = w & 0xf;
			al |= al << 4;
			in |= in << 4;
			out[0] = out[1] = out[2] = (int32_t)in;
			out[3] = (int32_t)al;
			return;
		}
		if (size == 2) {
			convert_ia16(tmem16(o, texel_index16(t, s, v, 0xfff)), out);
			return;
		}
		goto ci32;
	case TF_I:
		if (size == 0)
			goto i4;
		if (size == 1)
			goto i8;
		goto ci32;
	default:
		/* YUV is handled by the caller; other values sample like RGBA 16. */
		splat4(out, 0);
		return;
	}

i4: {
		uint32_t idx = texel_index8(t, s, v, 0xfff, s >> 1);
		uint32_t w = (tmem8(o, idx) >> ((~s & 1) * 4)) & 0xf;
		w |= w << 4;
		splat4(out, (int32_t)w);
		return;
	}
i8: {
		uint32_t w = tmem8(o, texel_index8(t, s, v, 0xfff, s));
		splat4(out, (int32_t)w);
		return;
	}
ci32: {
		uint32_t w = tmem16(o, texel_index16(t, s, v, 0xfff));
		out[0] = out[2] = (int32_t)(w >> 8);
		out[1] = out[3] = (int32_t)(w & 0xff);
		return;
	}
}

static void fetch_yuv16(const rdpo *o, const tile_t *t, uint32_t s, uint32_t v, uint32_t chroma_x, int32_t out[4])
{
	uint32_t byte_offset = t->offset + t->stride * v;
	uint32_t bl = (byte_offset + s) & 0x7ff;
	uint32_t bc = (byte_offset + chroma_x * 2) & 0x7ff;
	uint32_t il = (bl ^ ((v & 1) << 2)) ^ 3;
	uint32_t ic = ((bc >> 1) ^ ((v & 1) << 1)) ^ 1;
	uint32_t luma = tmem8(o, il | 0x800);
	uint32_t chroma = tmem16(o, ic);
	out[0] = (int32_t)((chroma >> 8) & 0xff) - 0x80;
	out[1] = (int32_t)(chroma & 0xff) - 0x80;
	out[2] = (int32_t)luma;
	out[3] = (int32_t)luma;
}

/* texture.h:317-358 */
static int clamp_and_shift_coord(int clamp_bit, int coord, int lo, int hi, int shift)
{
	coord = iclamp(coord, -0x8000, 0x7fff);
	if (shift < 11)
		coord >>= shift;
	else {
		coord <<= (32 - shift);
		coord >>= 16;
	}
	if (clamp_bit) {
		int clamp_hi = PLANTED(10) ? (coord >> 3) > hi : (coord >> 3) >= hi;
		if (clamp_hi)
			coord = (((hi >> 2) - (lo >> 2)) & 0x3ff) << 5;
		else
			coord = imax(coord - (lo << 3), 0);
	} else
		coord -= lo << 3;
	return coord;
}

static int shift_coord(int coord, int lo, int shift)
{
	coord = iclamp(coord, -0x8000, 0x7fff);
	if (shift < 11)
		coord >>= shift;
	else {
		coord <<= (32 - shift);
		coord >>= 16;
	}
	return coord - (lo << 3);
}

static void texture_convert_factors(const int32_t in[4], const int32_t f[4], int32_t out[4])
{
	int32_t t[4];
	for (int i = 0; i < 4; i++)
		t[i] = sext(in[i], 9);
	int32_t r = t[2] + ((f[0] * t[1] + 0x80) >> 8);
	int32_t g = t[2] + ((f[1] * t[0] + f[2] * t[1] + 0x80) >> 8);
	int32_t b = t[2] + ((f[3] * t[0] + 0x80) >> 8);
	out[0] = r;
	out[1] = g;
	out[2] = b;
	out[3] = t[2];
}

static void bilinear_3tap2(const int32_t t00[2], const int32_t t10[2], const int32_t t01[2], const int32_t t11[2],
                           int fx, int fy, int32_t out[2])
{
	int sum = fx + fy;
	const int32_t *base = sum >= 32 ? t11 : t00;
	int ffx = sum >= 32 ? 32 - fy : fx;
	int ffy = sum >= 32 ? 32 - fx : fy;
	for (int i = 0; i < 2; i++) {
		int32_t acc = (t10[i] - base[i]) * ffx;
		acc += (t01[i] - base[i]) * ffy;
		acc += 0x10;
		acc >>= 5;
		out[i] = acc + base[i];
	}
}

/* texture.h:498-923 */
static void sample_texture(const rdpo *o, const tile_t *tile, const int32_t st_in[2], int tlut, int tlut_type,
                           int sample_quad, int mid_texel_state, int convert_one, int bilerp,
                           const int32_t factors[4], const int32_t prev[4], int32_t out[4])
{
	int32_t st[2];
	st[0] = clamp_and_shift_coord((tile->flags & TILE_CLAMP_S) != 0, st_in[0], (int)tile->slo, (int)tile->shi,
	                              (int)tile->shift_s);
	st[1] = clamp_and_shift_coord((tile->flags & TILE_CLAMP_T) != 0, st_in[1], (int)tile->tlo, (int)tile->thi,
	                              (int)tile->shift_t);
	int fx = 0, fy = 0;
	if (sample_quad || tlut) {
		fx = st[0] & 31;
		fy = st[1] & 31;
	}
	int sum_frac = fx + fy;
	st[0] >>= 5;
	st[1] >>= 5;

	int s0 = texel_mask_s(tile, st[0]);
	int t0 = texel_mask_t(tile, st[1]);
	int s1 = texel_mask_s(tile, st[0] + 1);
	int t1 = texel_mask_t(tile, st[1] + 1);

	int tdiff = imax(t1 - t0, -255);
	t1 = (t0 & 0xff) + tdiff;
	t0 &= 0xff;

	int32_t tb[4] = { 0, 0, 0, 0 }, t10[4] = { 0, 0, 0, 0 }, t01[4] = { 0, 0, 0, 0 }, t11[4] = { 0, 0, 0, 0 };
	int mid_texel = mid_texel_state && bilerp && fx == 0x10 && fy == 0x10;
	int upper_lut = sum_frac >= 0x20;
	if (mid_texel)
		sum_frac = 0;

	int yuv = tile->fmt == TF_YUV;
	int bs = sum_frac >= 0x20 ? s1 : s0;
	int bt = sum_frac >= 0x20 ? t1 : t0;
	int chroma_frac = ((s0 & 1) << 4) | (fx >> 1);

	if (tlut) {
		if (!sample_quad) {
			bs = s0;
			bt = t0;
			s1 = s0;
			t1 = t0;
		}
		if (tile->fmt != TF_YUV) {
			int upper = sum_frac >= 0x20;
			uint32_t addr_xor = upper_lut ? 2 : 1;
			fetch_texel(o, tile, (uint32_t)bs, (uint32_t)bt, 1, tlut_type, upper ? 3 : 0, addr_xor, tb);
			if (bilerp) {
				fetch_texel(o, tile, (uint32_t)s1, (uint32_t)t0, 1, tlut_type, 1, addr_xor, t10);
				fetch_texel(o, tile, (uint32_t)s0, (uint32_t)t1, 1, tlut_type, 2, addr_xor, t01);
			}
			if (mid_texel)
				fetch_texel(o, tile, (uint32_t)s1, (uint32_t)t1, 1, tlut_type, 3, addr_xor, t11);
		}
	} else if (yuv) {
		uint32_t cx0 = (uint32_t)(s0 >> 1);
		uint32_t cx1 = (uint32_t)((s1 + (s1 - s0)) >> 1);
		fetch_yuv16(o, tile, (uint32_t)s0, (uint32_t)t0, cx0, tb);
		if (sample_quad) {
			fetch_yuv16(o, tile, (uint32_t)s1, (uint32_t)t0, cx1, t10);
			fetch_yuv16(o, tile, (uint32_t)s0, (uint32_t)t1, cx0, t01);
			fetch_yuv16(o, tile, (uint32_t)s1, (uint32_t)t1, cx1, t11);
		}
	} else {
		fetch_texel(o, tile, (uint32_t)bs, (uint32_t)bt, 0, 0, 0, 0, tb);
		if (sample_quad) {
			fetch_texel(o, tile, (uint32_t)s1, (uint32_t)t0, 0, 0, 0, 0, t10);
			fetch_texel(o, tile, (uint32_t)s0, (uint32_t)t1, 0, 0, 0, 0, t01);
		}
		if (mid_texel)
			fetch_texel(o, tile, (uint32_t)s1, (uint32_t)t1, 0, 0, 0, 0, t11);
	}

	int32_t acc[4];
	if (convert_one) {
		int32_t ps[4];
		for (int i = 0; i < 4; i++)
			ps[i] = sext(prev[i], 9);
		if (sample_quad) {
			int mid_rg = yuv ? (mid_texel_state && chroma_frac == 0x10 && fy == 0x10) : mid_texel;
			int mid_ba = mid_texel;
			int upper_ba = sum_frac >= 32;
			int upper_rg = yuv ? ((chroma_frac + fy) >= 32 && !mid_rg) : upper_ba;
			int32_t frg0 = upper_rg ? ps[1] : ps[0], frg1 = upper_rg ? ps[0] : ps[1];
			int32_t fba0 = upper_ba ? ps[1] : ps[0], fba1 = upper_ba ? ps[0] : ps[1];
			int32_t conv[4];
			for (int i = 0; i < 2; i++) {
				if (mid_rg)
					conv[i] = frg0 * (t01[i] - t11[i]) + frg1 * (t10[i] - t11[i]) + ((tb[i] - t11[i]) << 6) + 0x80;
				else {
					int32_t base = (upper_rg && yuv) ? t11[i] : tb[i];
					conv[i] = frg0 * (t10[i] - base) + frg1 * (t01[i] - base) + 0x80;
				}
			}
			for (int i = 2; i < 4; i++) {
				if (mid_ba)
					conv[i] = fba0 * (t01[i] - t11[i]) + fba1 * (t10[i] - t11[i]) + ((tb[i] - t11[i]) << 6) + 0x80;
				else {
					int32_t base = (upper_ba && yuv) ? t11[i] : tb[i];
					conv[i] = fba0 * (t10[i] - base) + fba1 * (t01[i] - base) + 0x80;
				}
			}
			for (int i = 0; i < 4; i++)
				acc[i] = (conv[i] >> 8) + ps[2];
		} else
			splat4(acc, ps[2]);
	} else if (yuv) {
		if (sample_quad) {
			int32_t ac[2], al[2];
			if (bilerp) {
				int mid_chroma = mid_texel_state && chroma_frac == 0x10 && fy == 0x10;
				if (mid_chroma) {
					for (int i = 0; i < 2; i++)
						ac[i] = (tb[i] + t10[i] + t11[i] + t01[i] + 2) >> 2;
				} else
					bilinear_3tap2(tb, t10, t01, t11, chroma_frac, fy, ac);
				if (mid_texel) {
					for (int i = 0; i < 2; i++)
						al[i] = (tb[2 + i] + t10[2 + i] + t11[2 + i] + t01[2 + i] + 2) >> 2;
				} else
					bilinear_3tap2(tb + 2, t10 + 2, t01 + 2, t11 + 2, fx, fy, al);
			} else {
				const int32_t *l = (fx + fy >= 32) ? t11 : tb;
				const int32_t *c = (chroma_frac + fy >= 32) ? t11 : tb;
				al[0] = l[2];
				al[1] = l[3];
				ac[0] = c[0];
				ac[1] = c[1];
			}
			acc[0] = ac[0];
			acc[1] = ac[1];
			acc[2] = al[0];
			acc[3] = al[1];
		} else
			memcpy(acc, tb, sizeof(acc));
	} else if (mid_texel) {
		for (int i = 0; i < 4; i++)
			acc[i] = (tb[i] + t01[i] + t10[i] + t11[i] + 2) >> 2;
	} else if (bilerp && (sample_quad || tlut)) {
		int ffx = sum_frac >= 32 ? 32 - fy : fx;
		int ffy = sum_frac >= 32 ? 32 - fx : fy;
		for (int i = 0; i < 4; i++) {
			int32_t a = (t10[i] - tb[i]) * ffx;
			a += (t01[i] - tb[i]) * ffy;
			a += PLANTED(11) ? 0xf : 0x10;
			a >>= 5;
			acc[i] = a + tb[i];
		}
	} else
		memcpy(acc, tb, sizeof(acc));

	if (!bilerp && !convert_one)
		texture_convert_factors(acc, factors, out);
	else
		memcpy(out, acc, sizeof(acc));
}

/* texture.h:925-1003 */
static void compute_lod_2cycle(uint32_t *tile0, uint32_t *tile1, int32_t *lod_frac, uint32_t max_level, int min_lod,
                               const int32_t st[2], const int32_t st_dx[2], const int32_t st_dy[2], int ovf,
                               int tex_lod_en, int sharpen, int detail)
{
	int magnify = 0, distant = 0;
	uint32_t tile_offset = 0;
	if (ovf) {
		distant = 1;
		*lod_frac = 0xff;
	} else {
		int32_t dx0 = st_dx[0] - st[0], dx1 = st_dx[1] - st[1];
		int32_t dy0 = st_dy[0] - st[0], dy1 = st_dy[1] - st[1];
		dx0 ^= dx0 >> 31;
		dx1 ^= dx1 >> 31;
		dy0 ^= dy0 >> 31;
		dy1 ^= dy1 >> 31;
		int32_t max_d = imax(imax(dx0, dy0), imax(dx1, dy1));
		if (max_d >= 0x4000) {
			distant = 1;
			*lod_frac = 0xff;
			tile_offset = max_level;
		} else if (max_d < 32) {
			distant = max_level == 0;
			magnify = 1;
			if (!sharpen && !detail)
				*lod_frac = distant ? 0xff : 0;
			else
				*lod_frac = (imax(min_lod, max_d) << 3) + (sharpen ? -0x100 : 0);
		} else {
			int mip_base = imax(find_msb(max_d >> (PLANTED(12) ? 4 : 5)), 0);
			distant = (uint32_t)mip_base >= max_level;
			if (distant && !sharpen && !detail)
				*lod_frac = 0xff;
			else {
				*lod_frac = ((max_d << 3) >> mip_base) & 0xff;
				tile_offset = (uint32_t)mip_base;
			}
		}
	}

	if (tex_lod_en) {
		if (distant)
			tile_offset = max_level;
		if (!detail) {
			*tile0 = (*tile0 + tile_offset) & 7;
			if (distant || (!sharpen && magnify))
				*tile1 = *tile0;
			else
				*tile1 = (*tile0 + 1) & 7;
		} else {
			*tile1 = (*tile0 + tile_offset + ((distant || magnify) ? 1 : 2)) & 7;
			*tile0 = (*tile0 + tile_offset + (magnify ? 0 : 1)) & 7;
		}
	}
}

/* ------------------------------------------------------------------------ */
/* Copy pipe texture fetch: texture.h:360-472.                              */
/* ------------------------------------------------------------------------ */
static int sample_texture_copy_word(const rdpo *o, const tile_t *tile, int32_t st[2], int s_offset, int tlut)
{
	int high_word = s_offset < 2;
	int replicate_8bpp = high_word && tile->size != 2 && !tlut;
	int samp;
	int s_shamt = imin((int)tile->size, 2);
	int large = tile->size == 3;
	uint32_t idx_mask = (large || tlut) ? 0x3ff : 0x7ff;

	if (replicate_8bpp) {
		int sx = st[0] + 2 * s_offset;
		int ms[2] = { sx, sx + 1 };
		for (int i = 0; i < 2; i++) {
			if (tile->mask_s) {
				int mask = 1 << tile->mask_s;
				if (tile->flags & TILE_MIRROR_S)
					ms[i] ^= imax((ms[i] & mask) - 1, 0);
				ms[i] &= mask - 1;
			}
		}
		int t = texel_mask_t(tile, st[1]);
		uint32_t tbase = tile->offset + tile->stride * (uint32_t)t;
		uint32_t nib[2], sam[2];
		for (int i = 0; i < 2; i++) {
			nib[i] = (tbase * 2 + ((uint32_t)ms[i] << s_shamt)) & 0x1fff;
			nib[i] ^= ((uint32_t)t & 1) * 8;
			uint32_t index = (nib[i] >> 2) & idx_mask;
			sam[i] = tmem16(o, index ^ 1);
			if (tile->size == 1) {
				sam[i] >>= 8 - 4 * (nib[i] & 2);
				sam[i] &= 0xff;
			} else if (tile->size == 0) {
				sam[i] >>= 12 - 4 * (nib[i] & 3);
				sam[i] = (sam[i] & 0xf) * 0x11;
			} else
				sam[i] >>= 8;
		}
		samp = (int)((sam[0] << 8) | sam[1]);
	} else {
		int sx = st[0] + s_offset;
		int s = texel_mask_s(tile, sx);
		int t = texel_mask_t(tile, st[1]);
		uint32_t tbase = tile->offset + tile->stride * (uint32_t)t;
		uint32_t nib = (tbase * 2 + ((uint32_t)s << s_shamt)) & 0x1fff;
		nib ^= ((uint32_t)t & 1) * 8;
		uint32_t index = (nib >> 2) & idx_mask;
		samp = (int)tmem16(o, index ^ 1);
		if (tlut) {
			if (tile->size == 0) {
				samp >>= 12 - 4 * (nib & 3);
				samp &= 0xf;
				samp |= (int)tile->palette << 4;
				samp <<= 2;
				samp += s_offset;
			} else {
				samp >>= 8 - 4 * (nib & 2);
				samp &= 0xff;
				samp <<= 2;
				samp += s_offset;
			}
			samp = (int)tmem16(o, ((uint32_t)samp | 0x400) ^ 1);
		}
	}
	return samp;
}

static int sample_texture_copy(const rdpo *o, const tile_t *tile, int32_t st_in[2], int s_offset, int tlut, int fb_size)
{
	int32_t st[2];
	st[0] = shift_coord(st_in[0], (int)tile->slo, (int)tile->shift_s) >> 5;
	st[1] = shift_coord(st_in[1], (int)tile->tlo, (int)tile->shift_t) >> 5;
	int samp;
	if (fb_size == 0)
		samp = 0;
	else if (fb_size == 1) {
		samp = sample_texture_copy_word(o, tile, st, s_offset >> 1, tlut);
		samp >>= 8 - 8 * (s_offset & 1);
		samp &= 0xff;
	} else
		samp = sample_texture_copy_word(o, tile, st, s_offset, tlut);
	return samp;
}

/* ------------------------------------------------------------------------ */
/* Combiner: combiner.h.                                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
	const int32_t *const_muladd, *const_mulsub, *const_mul, *const_add;
	int32_t shade[4], combined[4], texel0[4], texel1[4];
	int32_t lod_frac, noise;
} comb_in_t;

static inline int32_t special_expand(int32_t v) { return sext(v - 0x80, 9) + 0x80; }

static void select4(const comb_in_t *in, int which, int sel_rgb, int sel_a, int32_t out[4])
{
	/* which: 0 muladd, 1 mulsub, 2 mul, 3 add.  Selector values follow
	 * rdp_common.hpp:83-174; unrecognised selectors read the constant slot,
	 * which the setup code fills for Prim/Env/Key/K4/K5 and leaves 0 otherwise. */
	const int32_t *k = which == 0 ? in->const_muladd : which == 1 ? in->const_mulsub : which == 2 ? in->const_mul : in->const_add;
	int r = 0;
	switch (which) {
	case 0:
		switch (sel_rgb) {
		case 0: memcpy(out, in->combined, 12); r = 1; break;
		case 1: memcpy(out, in->texel0, 12); r = 1; break;
		case 2: memcpy(out, in->texel1, 12); r = 1; break;
		case 4: memcpy(out, in->shade, 12); r = 1; break;
		case 7: out[0] = out[1] = out[2] = in->noise; r = 1; break;
		case 6: out[0] = out[1] = out[2] = 0x100; r = 1; break;
		default: break;
		}
		break;
	case 1:
		switch (sel_rgb) {
		case 0: memcpy(out, in->combined, 12); r = 1; break;
		case 1: memcpy(out, in->texel0, 12); r = 1; break;
		case 2: memcpy(out, in->texel1, 12); r = 1; break;
		case 4: memcpy(out, in->shade, 12); r = 1; break;
		case 7: out[0] = out[1] = out[2] = (k[1] << 8) | k[2]; r = 1; break;
		default: break;
		}
		break;
	case 2:
		switch (sel_rgb) {
		case 0: memcpy(out, in->combined, 12); r = 1; break;
		case 7: out[0] = out[1] = out[2] = in->combined[3]; r = 1; break;
		case 1: memcpy(out, in->texel0, 12); r = 1; break;
		case 2: memcpy(out, in->texel1, 12); r = 1; break;
		case 4: memcpy(out, in->shade, 12); r = 1; break;
		case 8: out[0] = out[1] = out[2] = in->texel0[3]; r = 1; break;
		case 9: out[0] = out[1] = out[2] = in->texel1[3]; r = 1; break;
		case 11: out[0] = out[1] = out[2] = PLANTED(13) ? in->texel0[3] : in->shade[3]; r = 1; break;
		case 13: out[0] = out[1] = out[2] = in->lod_frac; r = 1; break;
		case 15: out[0] = out[1] = out[2] = (k[1] << 8) | k[2]; r = 1; break;
		default: break;
		}
		break;
	default:
		switch (sel_rgb) {
		case 0: memcpy(out, in->combined, 12); r = 1; break;
		case 1: memcpy(out, in->texel0, 12); r = 1; break;
		case 2: memcpy(out, in->texel1, 12); r = 1; break;
		case 4: memcpy(out, in->shade, 12); r = 1; break;
		case 6: out[0] = out[1] = out[2] = 0x100; r = 1; break;
		default: break;
		}
		break;
	}
	if (!r) {
		out[0] = k[0];
		out[1] = k[1];
		out[2] = k[2];
	}

	if (which == 2) {
		switch (sel_a) {
		case 0: out[3] = in->lod_frac; break;
		case 1: out[3] = in->texel0[3]; break;
		case 2: out[3] = in->texel1[3]; break;
		case 4: out[3] = in->shade[3]; break;
		default: out[3] = k[3]; break;
		}
	} else {
		switch (sel_a) {
		case 0: out[3] = in->combined[3]; break;
		case 1: out[3] = in->texel0[3]; break;
		case 2: out[3] = in->texel1[3]; break;
		case 4: out[3] = in->shade[3]; break;
		case 6: out[3] = 0x100; break;
		default: out[3] = k[3]; break;
		}
	}
}

static void combiner_equation(const int32_t a[4], const int32_t b[4], const int32_t c[4], const int32_t d[4], int32_t out[4])
{
	for (int i = 0; i < 4; i++) {
		int32_t cc = sext(c[i], 9);
		int32_t aa = special_expand(a[i]);
		int32_t bb = special_expand(b[i]);
		int32_t dd = special_expand(d[i]);
		int32_t col = (aa - bb) * cc;
		col += 0x80;
		out[i] = (col >> 8) + dd;
	}
}

static void combine_select(const comb_in_t *in, const uint8_t sel[8], int32_t out[4])
{
	int32_t a[4], b[4], c[4], d[4];
	select4(in, 0, sel[0], sel[4], a);
	select4(in, 1, sel[1], sel[5], b);
	select4(in, 2, sel[2], sel[6], c);
	select4(in, 3, sel[3], sel[7], d);
	combiner_equation(a, b, c, d, out);
}

static void combiner_cycle0(const comb_in_t *in, const uint8_t sel[8], int alpha_dith, int coverage,
                            int cvg_times_alpha, int alpha_cvg_select, int alpha_test, int32_t out[4], int32_t *alpha_ref)
{
	combine_select(in, sel, out);
	if (alpha_test) {
		int32_t clamped = clamp_9bit(out[3]);
		int32_t expanded = clamped + ((clamped + 1) >> 8);
		if (alpha_cvg_select) {
			int32_t modulated;
			if (cvg_times_alpha)
				modulated = (expanded * coverage + 4) >> 3;
			else
				modulated = coverage << 5;
			expanded = modulated;
		} else
			expanded += alpha_dith;
		*alpha_ref = iclamp(expanded, 0, 0xff);
	} else
		*alpha_ref = 0;
}

static void combiner_cycle1(const comb_in_t *in, const uint8_t sel[8], int alpha_dith, int *coverage,
                            int cvg_times_alpha, int alpha_cvg_select, int32_t out[4])
{
	combine_select(in, sel, out);
	for (int i = 0; i < 4; i++)
		out[i] = clamp_9bit(out[i]);
	int32_t expanded = out[3] + ((out[3] + 1) >> 8);
	int32_t modulated;
	if (cvg_times_alpha) {
		modulated = (expanded * *coverage + 4) >> 3;
		*coverage = modulated >> 5;
	} else
		modulated = *coverage << 5;
	if (alpha_cvg_select)
		expanded = modulated;
	else
		expanded += alpha_dith;
	out[3] = iclamp(expanded, 0, 0xff);
}

/* dither.h:40-68 */
static void dither_coefficients(int x, int y, int mode_rgb, int mode_alpha, uint32_t noise, int *rgb_dith, int *alpha_dith)
{
	const int SPLAT = (1 << 0) | (1 << 3) | (1 << 6);
	if (mode_rgb < 2)
		*rgb_dith = dither_matrices[mode_rgb][(y & 3) * 4 + (x & 3)] * SPLAT;
	else if (mode_rgb == 2)
		*rgb_dith = (int)(noise & 0x1ff);
	else
		*rgb_dith = 0;

	if (mode_alpha == 3)
		*alpha_dith = 0;
	else if (mode_alpha == 2)
		*alpha_dith = (int)(noise & 7);
	else {
		int a = mode_rgb >= 2 ? dither_matrices[mode_rgb & 1][(y & 3) * 4 + (x & 3)] : (*rgb_dith & 7);
		if (mode_alpha == 1)
			a = ~a & 7;
		*alpha_dith = a;
	}
}

static void rgb_dither(int32_t rgb[3], int dith)
{
	for (int i = 0; i < 3; i++) {
		int d = (dith >> (3 * i)) & 7;
		int32_t orig = rgb[i];
		int32_t v = orig > 247 ? 255 : (orig & 0xf8) + 8;
		int32_t replace_sign = (d - (orig & 7)) >> 31;
		int32_t diff = v - orig;
		rgb[i] = (orig + (diff & replace_sign)) & 0xff;
	}
}

/* ------------------------------------------------------------------------ */
/* Z encoding, depth test, blender, coverage: z_encode.h, depth_test.h,     */
/* blender.h, coverage.h.                                                   */
/* ------------------------------------------------------------------------ */
static int32_t z_decompress(uint32_t z)
{
	int32_t exponent = (int32_t)(z >> 11);
	int32_t mantissa = (int32_t)(z & 0x7ff);
	int32_t shift = imax(6 - exponent, 0);
	int32_t base = 0x40000 - (0x40000 >> exponent);
	return (mantissa << shift) + base;
}

static uint32_t z_compress(int32_t z)
{
	int32_t inv = imax(0x3ffff - z, 1);
	int32_t exponent = 17 - find_msb(inv);
	exponent = iclamp(exponent, 0, 7);
	int32_t shift = imax(6 - exponent, 0);
	int32_t mantissa = (z >> shift) & 0x7ff;
	return (uint32_t)((exponent << 11) + mantissa) & 0xffff;
}

static int depth_test(int32_t z, int32_t dz, int32_t dz_compressed, uint32_t cur_depth, uint32_t cur_dz,
                      int *coverage_count, int cur_cvg, int z_compare, int z_mode, int force_blend, int aa,
                      int *blend_en, int *cvg_wrap, int blend_shift[2])
{
	int pass = 0;
	if (z_compare) {
		int32_t memory_z = z_decompress(cur_depth);
		int32_t memory_dz = 1 << cur_dz;
		int32_t precision = ((int32_t)cur_depth >> 11) & 0xf;
		int coplanar = 0;
		blend_shift[0] = iclamp(dz_compressed - (int32_t)cur_dz, 0, 4);
		blend_shift[1] = iclamp((int32_t)cur_dz - dz_compressed, 0, 4);
		if (precision < 3) {
			if (memory_dz != 0x8000)
				memory_dz = imax(memory_dz << 1, 16 >> precision);
			else {
				coplanar = 1;
				memory_dz = 0xffff;
			}
		}
		int32_t combined = dz | memory_dz;
		if (combined != 0)
			combined = 1 << find_msb(combined);
		int32_t combined_ip = combined;
		combined <<= 3;
		int farther = coplanar || ((z + combined) >= memory_z);
		int overflow = (*coverage_count + cur_cvg) >= 8;
		*blend_en = force_blend || (!overflow && aa && farther);
		*cvg_wrap = overflow;
		int max_z = memory_z == 0x3ffff;
		int front = z < memory_z;
		int nearer = coplanar || (PLANTED(14) ? (z - combined) < memory_z : (z - combined) <= memory_z);
		switch (z_mode) {
		case 0:
			pass = max_z || (overflow ? front : nearer);
			break;
		case 1:
			if (!front || !farther || !overflow)
				pass = max_z || (overflow ? front : nearer);
			else {
				int32_t c = imax(find_msb(combined_ip & 0xffff), 0);
				int32_t coeff = ((memory_z >> c) - (z >> c)) & 0xf;
				*coverage_count = imin((coeff * *coverage_count) >> 3, 8);
				pass = 1;
			}
			break;
		case 2:
			pass = front || max_z;
			break;
		default:
			pass = farther && nearer && !max_z;
			break;
		}
	} else {
		blend_shift[0] = 0;
		blend_shift[1] = imin(0xf - dz_compressed, 4);
		int overflow = (*coverage_count + cur_cvg) >= 8;
		*blend_en = force_blend || (!overflow && aa);
		*cvg_wrap = overflow;
		pass = 1;
	}
	return pass;
}

typedef struct {
	int32_t pixel[4], memory[4], fog[4], blendc[4];
	int32_t shade_alpha;
} blend_in_t;

static void blender(const rdpo *o, const blend_in_t *in, const uint8_t modes[4], int force_blend, int blend_en,
                    int color_on_cvg, int cvg_wrap, const int blend_shift[2], int final_cycle, int32_t out[3])
{
	const int32_t *src1;
	switch (modes[2]) {
	case 0: src1 = in->pixel; break;
	case 1: src1 = in->memory; break;
	case 2: src1 = in->blendc; break;
	default: src1 = in->fog; break;
	}
	if (final_cycle && color_on_cvg && !cvg_wrap) {
		memcpy(out, src1, 12);
		return;
	}
	const int32_t *src0;
	switch (modes[0]) {
	case 0: src0 = in->pixel; break;
	case 1: src0 = in->memory; break;
	case 2: src0 = in->blendc; break;
	default: src0 = in->fog; break;
	}
	if (final_cycle) {
		if (!blend_en || (modes[1] == 0 && modes[3] == 0 && in->pixel[3] == 0xff)) {
			memcpy(out, src0, 12);
			return;
		}
	}
	int32_t a0, a1;
	switch (modes[1]) {
	case 0: a0 = in->pixel[3]; break;
	case 1: a0 = in->fog[3]; break;
	case 2: a0 = in->shade_alpha; break;
	default: a0 = 0; break;
	}
	switch (modes[3]) {
	case 0: a1 = ~a0 & 0xff; break;
	case 1: a1 = in->memory[3]; break;
	case 2: a1 = 0xff; break;
	default: a1 = 0; break;
	}
	a0 >>= 3;
	a1 >>= 3;
	if (modes[3] == 1) {
		a0 = (a0 >> blend_shift[0]) & 0x3c;
		a1 = (a1 >> blend_shift[1]) | 3;
	}
	int32_t blended[3];
	for (int i = 0; i < 3; i++)
		blended[i] = src0[i] * a0 + src1[i] * (a1 + (PLANTED(15) ? 0 : 1));
	if (!final_cycle || force_blend) {
		for (int i = 0; i < 3; i++)
			out[i] = (blended[i] >> 5) & 0xff;
	} else {
		int32_t blend_sum = (a0 >> 2) + (a1 >> 2) + 1;
		for (int i = 0; i < 3; i++) {
			int32_t b = (blended[i] >> 2) & 0x7ff;
			uint32_t idx = ((uint32_t)blend_sum << 11) | (uint32_t)b;
			out[i] = idx < 0x8000 ? o->blender_lut[idx] : 0;
		}
	}
}

static int blend_coverage(int coverage, int memory_coverage, int blend_en, int mode)
{
	switch (mode) {
	case 0: return blend_en ? imin(7, memory_coverage + coverage) : ((coverage - 1) & 7);
	case 1: return (coverage + memory_coverage) & (PLANTED(19) ? 15 : 7);
	case 2: return 7;
	default: return memory_coverage;
	}
}

/* ------------------------------------------------------------------------ */
/* Framebuffer memory: memory_interfacing.h:56-282, 364-486.                */
/* ------------------------------------------------------------------------ */
typedef struct pixel_state {
	int32_t color[4];
	uint32_t depth, dz;
	int color_dirty, depth_dirty;
	uint32_t color_index, depth_index;
	int loaded; /* upscaled samples: loaded from their plane */
} pixel_state_t;

static void load_color(const rdpo *o, uint32_t fmt, uint32_t index, int32_t c[4])
{
	uint32_t hmask = (o->rdram_size >> 1) - 1;
	switch (fmt) {
	case FB_I4:
	case FB_I8: {
		index &= o->rdram_size - 1;
		uint32_t w = o->rdram[index ^ 3];
		c[0] = c[1] = c[2] = (int32_t)w;
		c[3] = o->hidden[(index >> 1) & hmask];
		break;
	}
	case FB_RGBA5551: {
		index &= hmask;
		uint32_t w = rd16_index(o, index);
		c[0] = (int32_t)((w >> 8) & 0xf8);
		c[1] = (int32_t)((w >> 3) & 0xf8);
		c[2] = (int32_t)((w << 2) & 0xf8);
		c[3] = (int32_t)(((o->hidden[index] << 5) | ((w & 1) << 7)) & 0xff);
		break;
	}
	case FB_IA88: {
		index &= hmask;
		uint32_t w = rd16_index(o, index);
		c[0] = c[1] = c[2] = (int32_t)((w >> 8) & 0xff);
		c[3] = (int32_t)(w & 0xff);
		break;
	}
	default: {
		uint32_t w = rd32_index(o, index);
		c[0] = (int32_t)(w >> 24);
		c[1] = (int32_t)((w >> 16) & 0xff);
		c[2] = (int32_t)((w >> 8) & 0xff);
		c[3] = (int32_t)(w & 0xff);
		break;
	}
	}
}

static void store_color(rdpo *o, uint32_t fmt, uint32_t index, const int32_t c[4])
{
	uint32_t hmask = (o->rdram_size >> 1) - 1;
	switch (fmt) {
	case FB_I4:
		index &= o->rdram_size - 1;
		o->rdram[index ^ 3] = 0;
		if (index & 1)
			o->hidden[(index >> 1) & hmask] = (uint8_t)c[3];
		break;
	case FB_I8: {
		index &= o->rdram_size - 1;
		uint32_t col = (uint32_t)((index & 1) ? c[1] : c[0]) & 0xff;
		o->rdram[index ^ 3] = (uint8_t)col;
		if (index & 1)
			o->hidden[(index >> 1) & hmask] = (uint8_t)((col & 1) * 3);
		break;
	}
	case FB_RGBA5551: {
		index &= hmask;
		uint32_t r = (uint32_t)c[0] & 0xf8, g = (uint32_t)c[1] & 0xf8, b = (uint32_t)c[2] & 0xf8;
		uint32_t cov = ((uint32_t)c[3] & 0xff) >> 5;
		uint32_t w = (r << 8) | (g << 3) | (b >> 2) | (PLANTED(16) ? (cov & 1) : (cov >> 2));
		wr16_index(o, index, (uint16_t)w);
		o->hidden[index] = (uint8_t)(cov & 3);
		break;
	}
	case FB_IA88: {
		index &= hmask;
		uint32_t w = (((uint32_t)c[0] & 0xff) << 8) | ((uint32_t)c[3] & 0xff);
		wr16_index(o, index, (uint16_t)w);
		o->hidden[index] = (uint8_t)(((uint32_t)c[3] & 1) * 3);
		break;
	}
	default: {
		uint32_t w = (((uint32_t)c[0] & 0xff) << 24) | (((uint32_t)c[1] & 0xff) << 16) | (((uint32_t)c[2] & 0xff) << 8) |
		             ((uint32_t)c[3] & 0xff);
		wr32_index(o, index, w);
		index &= (o->rdram_size >> 2) - 1;
		o->hidden[(2 * index) & hmask] = (uint8_t)(((uint32_t)c[1] & 1) * 3);
		o->hidden[(2 * index + 1) & hmask] = (uint8_t)(((uint32_t)c[3] & 1) * 3);
		break;
	}
	}
}

static void load_depth(const rdpo *o, uint32_t index, uint32_t *depth, uint32_t *dz)
{
	uint32_t hmask = (o->rdram_size >> 1) - 1;
	index &= hmask;
	uint32_t w = rd16_index(o, index);
	*depth = w >> 2;
	*dz = (o->hidden[index] | ((w & 3) << 2)) & 0xff;
}

static void store_depth(rdpo *o, uint32_t index, uint32_t depth, uint32_t dz)
{
	uint32_t hmask = (o->rdram_size >> 1) - 1;
	index &= hmask;
	wr16_index(o, index, (uint16_t)((depth << 2) | (dz >> 2)));
	o->hidden[index] = (uint8_t)(dz & 3);
}

static void alias_color_to_depth(pixel_state_t *p, uint32_t fmt)
{
	if (fmt == FB_RGBA5551) {
		p->dz = (((uint32_t)p->color[3] >> 3) | ((uint32_t)p->color[2] & 8)) & 0xff;
		uint32_t w = ((uint32_t)p->color[0] & 0xf8) << 6;
		w |= ((uint32_t)p->color[1] & 0xf8) << 1;
		w |= ((uint32_t)p->color[2] & 0xf8) >> 4;
		p->depth = w & 0xffff;
	} else if (fmt == FB_IA88) {
		uint32_t w = (((uint32_t)p->color[0] & 0xff) << 8) | ((uint32_t)p->color[3] & 0xff);
		p->depth = (w >> 2) & 0xffff;
		p->dz = (((w & 3) << 2) | ((w & 1) * 3)) & 0xff;
	}
}

static void alias_depth_to_color(pixel_state_t *p, uint32_t fmt)
{
	uint32_t w = (p->depth << 4) | p->dz;
	if (fmt == FB_RGBA5551) {
		p->color[0] = (int32_t)((w >> 10) & 0xf8);
		p->color[1] = (int32_t)((w >> 5) & 0xf8);
		p->color[2] = (int32_t)(w & 0xf8);
		p->color[3] = (int32_t)((w & 7) << 5);
	} else if (fmt == FB_IA88) {
		p->color[0] = (int32_t)((w >> 10) & 0xff);
		p->color[3] = (int32_t)((w >> 2) & 0xff);
	}
	p->color_dirty = 1;
}

static void write_color(pixel_state_t *p, uint32_t fmt, const int32_t c[4])
{
	if (fmt == FB_I4) {
		p->color[0] = c[0];
		p->color[1] = c[1];
		p->color[2] = c[2];
	} else
		memcpy(p->color, c, 16);
	p->color_dirty = 1;
}

/* ------------------------------------------------------------------------ */
/* Per-pixel shading, shading.h:38-359 (SCALING_FACTOR = 1).                */
/* ------------------------------------------------------------------------ */
typedef struct {
	int32_t combined[4];
	int32_t z_dith;
	int32_t coverage_count;
	int32_t shade_alpha;
} shaded_t;

typedef struct {
	const tri_t *tri;
	const attr_t *attr;
	const draw_t *d;
	const tile_t *tiles; /* snapshot at draw time */
	const span_t *spans; /* lines ylo .. yhi+1 (upscaled: S * ylo .. S * (yhi + 2) - 1) */
	int ylo, yhi;        /* native lines */
	int L;               /* SCALING_LOG2: 0 native, > 0 the upscaled pass */
	uint32_t fb_size; /* 0,1,2,4 */
	uint32_t dx_mask, dx_shift;
} prim_ctx_t;

static int shade_pixel(const rdpo *o, const prim_ctx_t *pc, int x, int y, shaded_t *out)
{
	const draw_t *d = pc->d;
	const attr_t *attr = pc->attr;
	const int L = pc->L, S = 1 << L;
	/* shading.h:38-56 */
	if (y < S * pc->ylo || y > pc->yhi * S + (S - 1))
		return 0;
	if (L && (pc->tri->flags & TRI_DISABLE_UPSCALING)) {
		x &= ~(S - 1);
		y &= ~(S - 1);
	}
	const span_t *span = &pc->spans[y - S * pc->ylo];
	if (!span->valid)
		return 0;

	uint32_t f = d->ras_flags;
	int tlut = (f & RAS_TLUT) != 0;
	int tlut_type = (f & RAS_TLUT_TYPE) != 0;
	int sample_quad = (f & RAS_SAMPLE_QUAD) != 0;
	int cvg_times_alpha = (f & RAS_CVG_TIMES_ALPHA) != 0;
	int alpha_cvg_select = (f & RAS_ALPHA_CVG_SELECT) != 0;
	int perspective = (f & RAS_PERSPECTIVE) != 0;
	int tex_lod_en = (f & RAS_TEX_LOD) != 0;
	int sharpen = (f & RAS_SHARPEN_LOD) != 0;
	int detail = (f & RAS_DETAIL_LOD) != 0;
	int aa = (f & RAS_AA) != 0;
	int multi_cycle = (f & RAS_MULTI_CYCLE) != 0;
	int interlace_en = (f & RAS_INTERLACE_FIELD) != 0;
	int fill_en = (f & RAS_FILL) != 0;
	int copy_en = (f & RAS_COPY) != 0;
	int alpha_test = (f & RAS_ALPHA_TEST) != 0;
	int alpha_test_dither = (f & RAS_ALPHA_TEST_DITHER) != 0;
	int mid_texel = (f & RAS_MID_TEXEL) != 0;
	int convert_one = (f & RAS_CONVERT_ONE) != 0;
	int bilerp0 = (f & RAS_BILERP0) != 0;
	int bilerp1 = (f & RAS_BILERP1) != 0;
	int uses_texel0 = d->uses_texel0, uses_texel1 = d->uses_texel1;
	int uses_pipelined_texel1 = d->uses_pipelined_texel1, uses_lod = d->uses_lod;

	uint32_t noise = 0;
	if (d->need_noise)
		noise = reseed_noise((uint32_t)x, (uint32_t)y, d->prim_index);

	int flip = (pc->tri->flags & TRI_FLIP) != 0;
	uint32_t setup_tile = pc->tri->tile;

	if (copy_en) {
		if (!(x >= span->start_x && x <= span->end_x))
			return 0;
		int dxp = flip ? (x - span->start_x) : (span->end_x - x);
		/* interpolate_st_copy: sub-pixels repeat their native pixel. */
		dxp >>= PLANTED(35) ? 0 : L;
		int snapped_dx = dxp & (int)pc->dx_mask;
		int s_offset = dxp - snapped_dx;
		int lerp_dx = (dxp >> pc->dx_shift) * (flip ? 1 : -1);
		int32_t stw[3] = {
			span->stzw[0] + (attr->dstzw_dx[0] & ~0x1f) * lerp_dx,
			span->stzw[1] + (attr->dstzw_dx[1] & ~0x1f) * lerp_dx,
			span->stzw[3] + (attr->dstzw_dx[3] & ~0x1f) * lerp_dx,
		};
		int32_t st[2];
		int ovf = 0;
		stw_to_st(stw, perspective, st, &ovf);
		const tile_t *t0 = &pc->tiles[setup_tile & 7];
		int texel0 = sample_texture_copy(o, t0, st, s_offset, tlut, (int)pc->fb_size);
		out->z_dith = texel0;
		out->coverage_count = COVERAGE_COPY;
		if (alpha_test && pc->fb_size == 2 && (texel0 & 1) == 0)
			return 0;
		return 1;
	} else if (fill_en) {
		out->coverage_count = COVERAGE_FILL;
		return x >= span->start_x && x <= span->end_x;
	}

	int coverage = compute_coverage(span, x);
	if (coverage == 0)
		return 0;
	int coverage_count = __builtin_popcount((unsigned)coverage);
	if (!aa && (coverage & 1) == 0)
		return 0;

	int dx = x - span->base_x;
	int interp_dir = flip ? 1 : -1;

	int32_t shade[4];
	interpolate_rgba(span, attr, dx, coverage, shade, L);

	/* shading.h:192-198: with the native-LOD quirk the LOD neighbours are a
	 * native pixel away. */
	int tex_dir = interp_dir;
	if (L && uses_lod && (pc->tri->flags & TRI_NATIVE_LOD))
		tex_dir *= S;
	int32_t st[2], st_dx[2] = { 0, 0 }, st_dy[2] = { 0, 0 }, z;
	int povf = 0;
	interpolate_stz(span, attr, dx, coverage, perspective, uses_lod, tex_dir, st, st_dx, st_dy, &z, &povf, L);

	uint32_t tile0 = setup_tile & 7;
	uint32_t tile1 = (tile0 + 1) & 7;
	uint32_t max_level = setup_tile >> 3;
	int min_lod = d->min_lod;
	int32_t lod_frac = 0;
	if (uses_lod)
		compute_lod_2cycle(&tile0, &tile1, &lod_frac, max_level, min_lod, st, st_dx, st_dy, povf, tex_lod_en, sharpen,
		                   detail);

	int32_t texel0[4] = { 0, 0, 0, 0 }, texel1[4] = { 0, 0, 0, 0 };
	if (uses_texel0)
		sample_texture(o, &pc->tiles[tile0], st, tlut, tlut_type, sample_quad, mid_texel, 0, bilerp0, d->factors, texel1,
		               texel0);

	if (uses_pipelined_texel1) {
		const span_t *next = &pc->spans[y - S * pc->ylo + 1];
		int long_span = span->lodlength >= 8;
		int end_span = x == (flip ? span->end_x : span->start_x);
		if (end_span && long_span && next->valid) {
			int32_t stw[3] = { next->stzw[0] >> 16, next->stzw[1] >> 16, next->stzw[3] >> 16 };
			if (perspective) {
				int ovf2 = 0;
				perspective_divide(stw, st, &ovf2);
			} else {
				st[0] = stw[0];
				st[1] = stw[1];
			}
		} else
			interpolate_st_single(span, attr, dx + interp_dir * S, perspective, st, L);
		tile1 = tile0;
		uses_texel1 = 1;
	}

	if (uses_texel1) {
		if (convert_one && !bilerp1)
			texture_convert_factors(texel0, d->factors, texel1);
		else
			sample_texture(o, &pc->tiles[tile1], st, tlut, tlut_type, sample_quad, mid_texel, convert_one, bilerp1,
			               d->factors, texel0, texel1);
	}

	int rgb_dith, alpha_dith;
	dither_coefficients(x, y >> interlace_en, (int)(d->dither >> 2), (int)(d->dither & 3), noise, &rgb_dith, &alpha_dith);

	int32_t combined[4];
	int32_t alpha_ref;
	comb_in_t ci;
	memset(&ci, 0, sizeof(ci));
	memcpy(ci.shade, shade, 16);
	memcpy(ci.texel0, texel0, 16);
	memcpy(ci.texel1, texel1, 16);
	ci.lod_frac = lod_frac;
	ci.noise = (int32_t)(((noise & 7) << 6) | 0x20);

	if (multi_cycle) {
		ci.const_muladd = d->const_muladd[0];
		ci.const_mulsub = d->const_mulsub[0];
		ci.const_mul = d->const_mul[0];
		ci.const_add = d->const_add[0];
		int32_t c0[4];
		combiner_cycle0(&ci, d->comb[0], alpha_dith, coverage_count, cvg_times_alpha, alpha_cvg_select, alpha_test, c0,
		                &alpha_ref);
		memcpy(ci.combined, c0, 16);
		ci.const_muladd = d->const_muladd[1];
		ci.const_mulsub = d->const_mulsub[1];
		ci.const_mul = d->const_mul[1];
		ci.const_add = d->const_add[1];
		int32_t tmp[4];
		memcpy(tmp, ci.texel0, 16);
		memcpy(ci.texel0, ci.texel1, 16);
		memcpy(ci.texel1, tmp, 16);
		if (d->need_noise_dual) {
			uint32_t n2 = reseed_noise((uint32_t)(x + 1023), (uint32_t)(y + 7), d->prim_index + 11);
			ci.noise = (int32_t)(((n2 & 7) << 6) | 0x20);
		}
		combiner_cycle1(&ci, d->comb[1], alpha_dith, &coverage_count, cvg_times_alpha, alpha_cvg_select, combined);
	} else {
		ci.const_muladd = d->const_muladd[1];
		ci.const_mulsub = d->const_mulsub[1];
		ci.const_mul = d->const_mul[1];
		ci.const_add = d->const_add[1];
		combiner_cycle1(&ci, d->comb[1], alpha_dith, &coverage_count, cvg_times_alpha, alpha_cvg_select, combined);
		alpha_ref = combined[3];
	}

	if (aa && coverage_count == 0)
		return 0;

	if (alpha_test) {
		int32_t threshold = alpha_test_dither ? (int32_t)(noise & 0xff) : d->blendc[3];
		if (alpha_ref < threshold)
			return 0;
	}

	memcpy(out->combined, combined, 16);
	out->z_dith = (z << 9) | rgb_dith;
	out->coverage_count = coverage_count;
	out->shade_alpha = imin(shade[3] + alpha_dith, 0xff);
	return 1;
}

/* memory_interfacing.h:488-571 */
static void depth_blend(const rdpo *o, const draw_t *d, uint32_t fmt, int alias, pixel_state_t *p, const shaded_t *sh)
{
	int32_t z = sh->z_dith >> 9;
	int dith = sh->z_dith & 0x1ff;
	int coverage_count = sh->coverage_count;

	int force_blend = (d->db_flags & DB_FORCE_BLEND) != 0;
	int z_compare = (d->db_flags & DB_DEPTH_TEST) != 0;
	int z_update = (d->db_flags & DB_DEPTH_UPDATE) != 0;
	int image_read = (d->db_flags & DB_IMAGE_READ) != 0;
	int color_on_cvg = (d->db_flags & DB_COLOR_ON_CVG) != 0;
	int multi = (d->db_flags & DB_MULTI_CYCLE) != 0;
	int aa = (d->db_flags & DB_AA) != 0;
	int dither_en = (d->db_flags & DB_DITHER) != 0;

	/* decode_memory_color */
	int32_t memory[4];
	int32_t mem_cov = image_read ? (p->color[3] & 0xe0) : 0xe0;
	switch (fmt) {
	case FB_I4:
		memory[0] = memory[1] = memory[2] = 0;
		mem_cov = 0xe0;
		break;
	case FB_I8:
		memory[0] = memory[1] = memory[2] = p->color[0] & 0xff;
		mem_cov = 0xe0;
		break;
	case FB_RGBA5551:
		memory[0] = p->color[0] & 0xf8;
		memory[1] = p->color[1] & 0xf8;
		memory[2] = p->color[2] & 0xf8;
		break;
	case FB_IA88:
		memory[0] = memory[1] = memory[2] = p->color[0] & 0xff;
		break;
	default:
		memory[0] = p->color[0] & 0xff;
		memory[1] = p->color[1] & 0xff;
		memory[2] = p->color[2] & 0xff;
		break;
	}
	memory[3] = mem_cov;
	int memory_coverage = mem_cov >> 5;

	int blend_en = 0, cvg_wrap = 0, blend_shift[2] = { 0, 0 };
	int z_pass = depth_test(z, d->dz, d->dz_compressed, p->depth, p->dz, &coverage_count, memory_coverage, z_compare,
	                        (int)d->z_mode, force_blend, aa, &blend_en, &cvg_wrap, blend_shift);

	if (z_pass && (!aa || coverage_count != 0)) {
		blend_in_t bi;
		memcpy(bi.pixel, sh->combined, 16);
		memcpy(bi.memory, memory, 16);
		for (int i = 0; i < 4; i++) {
			bi.fog[i] = d->fog[i];
			bi.blendc[i] = d->blendc[i];
		}
		bi.shade_alpha = sh->shade_alpha;
		const uint8_t *modes = d->blend_modes[0];
		int32_t rgb[3];
		if (multi) {
			blender(o, &bi, modes, force_blend, blend_en, color_on_cvg, cvg_wrap, blend_shift, 0, rgb);
			bi.pixel[0] = rgb[0];
			bi.pixel[1] = rgb[1];
			bi.pixel[2] = rgb[2];
			modes = d->blend_modes[1];
		}
		blender(o, &bi, modes, force_blend, blend_en, color_on_cvg, cvg_wrap, blend_shift, 1, rgb);
		if (dither_en)
			rgb_dither(rgb, dith);
		int new_cov = blend_coverage(coverage_count, memory_coverage, blend_en, (int)d->coverage_mode);
		int32_t c[4] = { rgb[0], rgb[1], rgb[2], new_cov << 5 };
		write_color(p, fmt, c);
		if (z_update) {
			p->depth = z_compress(z);
			p->dz = (uint32_t)d->dz_compressed;
			p->depth_dirty = 1;
			if (alias)
				alias_depth_to_color(p, fmt);
		} else if (alias)
			alias_color_to_depth(p, fmt);
	}
}

static void fill_color(pixel_state_t *p, uint32_t fmt, int alias, uint32_t col)
{
	int32_t c[4];
	switch (fmt) {
	case FB_RGBA8888:
		c[0] = (int32_t)((col >> 24) & 0xff);
		c[1] = (int32_t)((col >> 16) & 0xff);
		c[2] = (int32_t)((col >> 8) & 0xff);
		c[3] = (int32_t)(col & 0xff);
		write_color(p, fmt, c);
		break;
	case FB_RGBA5551:
		col >>= ((p->color_index & 1) ^ (PLANTED(17) ? 0 : 1)) * 16;
		c[0] = (int32_t)((col >> 8) & 0xf8);
		c[1] = (int32_t)((col >> 3) & 0xf8);
		c[2] = (int32_t)((col << 2) & 0xf8);
		c[3] = (int32_t)((col & 1) * 0xe0);
		write_color(p, fmt, c);
		break;
	case FB_IA88:
		col >>= ((p->color_index & 1) ^ 1) * 16;
		col &= 0xffff;
		c[0] = c[1] = c[2] = (int32_t)((col >> 8) & 0xff);
		c[3] = (int32_t)(col & 0xff);
		write_color(p, fmt, c);
		break;
	case FB_I8:
		col >>= ((p->color_index & 3) ^ 3) * 8;
		col &= 0xff;
		c[0] = c[1] = c[2] = c[3] = (int32_t)col;
		write_color(p, fmt, c);
		break;
	default:
		break;
	}
	if (alias)
		alias_color_to_depth(p, fmt);
}

static void copy_pipeline(pixel_state_t *p, uint32_t fmt, int alias, uint32_t word)
{
	int32_t c[4];
	switch (fmt) {
	case FB_I4:
		p->color[0] = p->color[1] = p->color[2] = p->color[3] = 0;
		p->color_dirty = 1;
		break;
	case FB_I8:
		word &= 0xff;
		c[0] = c[1] = c[2] = c[3] = (int32_t)word;
		write_color(p, fmt, c);
		break;
	case FB_RGBA5551:
		c[0] = (int32_t)((word >> 8) & 0xf8);
		c[1] = (int32_t)((word >> 3) & 0xf8);
		c[2] = (int32_t)((word << 2) & 0xf8);
		c[3] = (int32_t)(((PLANTED(18) ? (word >> 1) : word) & 1) * 0xe0);
		write_color(p, fmt, c);
		break;
	default:
		break;
	}
	if (alias)
		alias_color_to_depth(p, fmt);
}

/* ------------------------------------------------------------------------ */
/* Draw path: Renderer::draw_shaded_primitive and helpers,                  */
/* rdp_renderer.cpp:717-1554.                                              */
/* ------------------------------------------------------------------------ */
static int normalize_dzpix(int dz)
{
	if (dz >= 0x8000)
		return 0x8000;
	if (dz == 0)
		return 1;
	return 1 << (find_msb(dz) + 1);
}

/* Host-side dz compression, rdp_renderer.cpp:728-740. */
static int dz_compress_host(int dz)
{
	int val = 0;
	if (dz & 0xff00)
		val |= 8;
	if (dz & 0xf0f0)
		val |= 4;
	if (dz & 0xcccc)
		val |= 2;
	if (dz & 0xaaaa)
		val |= 1;
	return val;
}

static int accesses_texel0(const uint8_t c[8])
{
	return c[0] == 1 || c[1] == 1 || c[2] == 1 || c[3] == 1 || c[2] == 8 || c[4] == 1 || c[5] == 1 || c[6] == 1 ||
	       c[7] == 1;
}

static int accesses_texel1(const uint8_t c[8])
{
	return c[0] == 2 || c[1] == 2 || c[2] == 2 || c[3] == 2 || c[2] == 9 || c[4] == 2 || c[5] == 2 || c[6] == 2 ||
	       c[7] == 2;
}

static int accesses_lod_frac(const uint8_t c[8])
{
	return c[2] == 13 || c[6] == 0;
}

/* build_combiner_constants, rdp_renderer.cpp:754-908.  Constants that a
 * selector does not read stay zero. */
static void build_constants(const rdpo *o, draw_t *d, int cycle)
{
	const uint8_t *c = o->comb[cycle];
	int32_t *ma = d->const_muladd[cycle], *ms = d->const_mulsub[cycle], *mu = d->const_mul[cycle],
	        *ad = d->const_add[cycle];
	uint32_t prim = o->prim_color, env = o->env_color;
#define RGB3(dst, col) do { (dst)[0] = (int32_t)(((col) >> 24) & 0xff); (dst)[1] = (int32_t)(((col) >> 16) & 0xff); \
	(dst)[2] = (int32_t)(((col) >> 8) & 0xff); } while (0)
	switch (c[0]) {
	case 5: RGB3(ma, env); break;
	case 3: RGB3(ma, prim); break;
	default: break;
	}
	switch (c[1]) {
	case 5: RGB3(ms, env); break;
	case 3: RGB3(ms, prim); break;
	case 7: RGB3(ms, (uint32_t)o->convert[4] << 8); break;
	case 6:
		ms[0] = (int32_t)o->key_center[0];
		ms[1] = (int32_t)o->key_center[1];
		ms[2] = (int32_t)o->key_center[2];
		break;
	default: break;
	}
	switch (c[2]) {
	case 3: RGB3(mu, prim); break;
	case 5: RGB3(mu, env); break;
	case 10: RGB3(mu, 0x01010101u * (prim & 0xff)); break;
	case 12: RGB3(mu, 0x01010101u * (env & 0xff)); break;
	case 14: RGB3(mu, 0x01010101u * o->prim_lod_frac); break;
	case 15: RGB3(mu, (uint32_t)o->convert[5] << 8); break;
	case 6:
		mu[0] = (int32_t)o->key_scale[0];
		mu[1] = (int32_t)o->key_scale[1];
		mu[2] = (int32_t)o->key_scale[2];
		break;
	default: break;
	}
	switch (c[3]) {
	case 3: RGB3(ad, prim); break;
	case 5: RGB3(ad, env); break;
	default: break;
	}
#undef RGB3
	if (c[4] == 3) ma[3] = (int32_t)(prim & 0xff);
	else if (c[4] == 5) ma[3] = (int32_t)(env & 0xff);
	if (c[5] == 3) ms[3] = (int32_t)(prim & 0xff);
	else if (c[5] == 5) ms[3] = (int32_t)(env & 0xff);
	if (c[6] == 3) mu[3] = (int32_t)(prim & 0xff);
	else if (c[6] == 5) mu[3] = (int32_t)(env & 0xff);
	else if (c[6] == 6) mu[3] = (int32_t)(o->prim_lod_frac & 0xff);
	if (c[7] == 3) ad[3] = (int32_t)(prim & 0xff);
	else if (c[7] == 5) ad[3] = (int32_t)(env & 0xff);
}

static void resolve_draw_state(rdpo *o, const attr_t *attr, draw_t *d)
{
	memset(d, 0, sizeof(*d));
	uint32_t f = o->ras_flags;

	/* build_derived_attributes, rdp_renderer.cpp:910-948 */
	if (o->use_prim_depth) {
		d->dz = (int32_t)o->prim_dz;
		d->dz_compressed = dz_compress_host((int)o->prim_dz);
	} else {
		int dzdx = attr->dstzw_dx[2] >> 16;
		int dzdy = attr->dstzw_dy[2] >> 16;
		int dzpix = (dzdx < 0 ? (~dzdx & 0x7fff) : dzdx) + (dzdy < 0 ? (~dzdy & 0x7fff) : dzdy);
		dzpix = normalize_dzpix(dzpix);
		d->dz = dzpix;
		d->dz_compressed = dz_compress_host(dzpix);
	}
	build_constants(o, d, 0);
	build_constants(o, d, 1);
	for (int i = 0; i < 4; i++) {
		d->fog[i] = (uint8_t)(o->fog_color >> (24 - 8 * i));
		d->blendc[i] = (uint8_t)(o->blend_color >> (24 - 8 * i));
	}
	d->fill_color = o->fill_color;
	d->min_lod = (int32_t)o->min_level;
	for (int i = 0; i < 4; i++)
		d->factors[i] = (int32_t)(int16_t)o->convert[i];

	memcpy(d->comb, o->comb, sizeof(d->comb));
	d->dither = o->dither;
	d->db_flags = o->db_flags;
	memcpy(d->blend_modes, o->blend, sizeof(d->blend_modes));
	d->coverage_mode = o->coverage_mode;
	d->z_mode = o->z_mode;

	/* deduce_static_texture_state, rdp_renderer.cpp:1333-1400 (usage part) */
	int multi = (f & RAS_MULTI_CYCLE) != 0;
	if (!(f & (RAS_FILL | RAS_COPY))) {
		int t0, t1, pt1, lodf;
		if (multi) {
			t0 = accesses_texel0(o->comb[0]) || accesses_texel1(o->comb[1]);
			t1 = accesses_texel1(o->comb[0]) || accesses_texel0(o->comb[1]);
			pt1 = 0;
			lodf = accesses_lod_frac(o->comb[0]) || accesses_lod_frac(o->comb[1]);
		} else {
			t0 = accesses_texel0(o->comb[1]);
			t1 = 0;
			pt1 = accesses_texel1(o->comb[1]);
			lodf = 0;
		}
		if (t1 && (f & RAS_CONVERT_ONE))
			t0 = 1;
		d->uses_texel0 = t0;
		d->uses_texel1 = t1;
		d->uses_pipelined_texel1 = pt1;
		d->uses_lod = lodf || (f & RAS_TEX_LOD) != 0;
	}

	/* deduce_noise_state, rdp_renderer.cpp:1136-1169 */
	int need_noise = 0, need_dual = 0;
	if ((o->dither & 3) == 2 || ((o->dither >> 2) & 3) == 2)
		need_noise = 1;
	else if (!(f & (RAS_COPY | RAS_FILL))) {
		if (multi && o->comb[0][0] == 7)
			need_noise = 1;
		if (o->comb[1][0] == 7)
			need_noise = 1;
		if (multi && o->comb[0][0] == 7 && o->comb[1][0] == 7)
			need_dual = 1;
		if ((f & (RAS_ALPHA_TEST | RAS_ALPHA_TEST_DITHER)) == (RAS_ALPHA_TEST | RAS_ALPHA_TEST_DITHER))
			need_noise = 1;
	}

	/* normalize_static_state, rdp_renderer.cpp:1294-1331 */
	if (f & RAS_FILL) {
		f = RAS_FILL;
		need_noise = need_dual = 0;
		d->dither = 0;
	} else if (f & RAS_COPY) {
		f &= RAS_COPY | RAS_TLUT | RAS_TLUT_TYPE | RAS_TEX_LOD | RAS_DETAIL_LOD | RAS_PERSPECTIVE | RAS_ALPHA_TEST;
		need_noise = need_dual = 0;
		d->dither = 0;
	} else if (!(f & RAS_MULTI_CYCLE) && !d->uses_pipelined_texel1)
		f &= ~(uint32_t)(RAS_BILERP1 | RAS_CONVERT_ONE);
	d->ras_flags = f;
	d->need_noise = need_noise;
	d->need_noise_dual = need_dual;
}

/* ------------------------------------------------------------------------ */
/* Render passes.  The reference merges draws into render passes and keeps    */
/* each touched pixel's colour and depth in registers for the whole pass,     */
/* loading it once and writing it back once (depth_blend.comp:86-149 with     */
/* init_tile/finish_tile).  For I8 colour images and for colour/depth         */
/* aliasing that differs from writing every fragment through to RDRAM, so the */
/* oracle keeps the same pass-resident pixel state and ends passes where the  */
/* reference's Renderer does (rdp_renderer.cpp:679-693, 1464-1590,            */
/* 2909-2987, 3051-3111, 3484-3485; rdp_device.cpp:263-266).  Flushes the     */
/* reference issues when its command thread runs idle                          */
/* (maintain_queues_idle, rdp_renderer.cpp:2403-2411) depend on timing and are */
/* not modelled: the oracle behaves as enqueue_command_direct.                */
/* ------------------------------------------------------------------------ */

static void pass_write_back(rdpo *o)
{
	uint32_t fmt = o->fb_fmt;
	int alias = o->fb_addr == o->depth_addr;
	for (size_t i = 0; i < o->num_touched; i++) {
		uint32_t key = o->px_touched[i];
		pixel_state_t *p = &o->px[key];
		if (p->color_dirty)
			store_color(o, fmt, p->color_index, p->color);
		if (!alias && p->depth_dirty)
			store_depth(o, p->depth_index, p->depth, p->dz);
		o->px_loaded[key] = 0;
	}
	o->num_touched = 0;
}

static void up_pass_end(rdpo *o, void (*native_write_back)(rdpo *));

/* Renderer::flush_queues + reset_context. */
static void pass_flush(rdpo *o)
{
	if (o->up_log2)
		up_pass_end(o, pass_write_back);
	else
		pass_write_back(o);
	if (o->pass.num_prims) {
		if (o->pass_log_len == o->pass_log_cap) {
			o->pass_log_cap = o->pass_log_cap ? 2 * o->pass_log_cap : 256;
			o->pass_log = (uint64_t *)realloc(o->pass_log, o->pass_log_cap * sizeof(uint64_t));
			if (!o->pass_log)
				abort();
		}
		o->pass_log[o->pass_log_len++] = o->prim_counter;
	}
	memset(&o->pass, 0, sizeof(o->pass));
}

/* normalize_combiner, rdp_renderer.cpp:1171-1292: selectors that only read
 * per-draw constants fold into the "zero" value of their slot. */
static uint8_t keep_or(uint8_t v, uint32_t keep_mask, uint8_t zero)
{
	return (v < 32 && ((keep_mask >> v) & 1)) ? v : zero;
}

static void static_state_key(const rdpo *o, uint32_t setup_tile, static_key_t *k)
{
	memset(k, 0, sizeof(*k));
	uint32_t f = o->ras_flags;
	unsigned tile = setup_tile & 7;

	/* normalize_static_state, rdp_renderer.cpp:1294-1331 (fill and copy). */
	if (f & RAS_FILL) {
		k->flags = RAS_FILL;
		return;
	}
	if (f & RAS_COPY) {
		/* deduce_static_texture_state always finds a static format in copy
		 * mode; the USES_* bits are never set there (the stream state is
		 * re-copied from the command processor on every mode change). */
		k->flags = (f & (RAS_COPY | RAS_TLUT | RAS_TLUT_TYPE | RAS_TEX_LOD | RAS_DETAIL_LOD | RAS_PERSPECTIVE |
		                 RAS_ALPHA_TEST)) |
		           RAS_USE_STATIC_TEXTURE_SIZE_FORMAT;
		k->texture_fmt = o->tiles[tile].fmt;
		k->texture_size = o->tiles[tile].size;
		return;
	}

	/* deduce_static_texture_state, rdp_renderer.cpp:1333-1400. */
	int multi = (f & RAS_MULTI_CYCLE) != 0;
	int t0, t1, pt1, lodf;
	if (multi) {
		t0 = accesses_texel0(o->comb[0]) || accesses_texel1(o->comb[1]);
		t1 = accesses_texel1(o->comb[0]) || accesses_texel0(o->comb[1]);
		pt1 = 0;
		lodf = accesses_lod_frac(o->comb[0]) || accesses_lod_frac(o->comb[1]);
	} else {
		t0 = accesses_texel0(o->comb[1]);
		t1 = 0;
		pt1 = accesses_texel1(o->comb[1]);
		lodf = 0;
	}
	if (t1 && (f & RAS_CONVERT_ONE))
		t0 = 1;
	if (t0) f |= RAS_USES_TEXEL0;
	if (t1) f |= RAS_USES_TEXEL1;
	if (pt1) f |= RAS_USES_PIPELINED_TEXEL1;
	if (lodf || (f & RAS_TEX_LOD)) f |= RAS_USES_LOD;
	if (t0 || t1 || pt1) {
		int physical_t1 = t1 && (!(f & RAS_CONVERT_ONE) || (f & RAS_BILERP1));
		unsigned max_lod = setup_tile >> 3;
		if (!(f & RAS_TEX_LOD))
			max_lod = physical_t1 ? 1 : 0;
		if (f & RAS_DETAIL_LOD)
			max_lod++;
		if (max_lod > 7)
			max_lod = 7;
		int same = 1;
		for (unsigned i = 1; i <= max_lod && same; i++) {
			const tile_t *t = &o->tiles[(tile + i) & 7];
			same = t->fmt == o->tiles[tile].fmt && t->size == o->tiles[tile].size;
		}
		if (same) {
			f |= RAS_USE_STATIC_TEXTURE_SIZE_FORMAT;
			k->texture_fmt = o->tiles[tile].fmt;
			k->texture_size = o->tiles[tile].size;
		}
	}

	/* deduce_noise_state, rdp_renderer.cpp:1136-1169. */
	if ((o->dither & 3) == 2 || ((o->dither >> 2) & 3) == 2)
		f |= RAS_NEED_NOISE;
	else {
		if (multi && o->comb[0][0] == 7)
			f |= RAS_NEED_NOISE;
		if (o->comb[1][0] == 7)
			f |= RAS_NEED_NOISE;
		if (multi && o->comb[0][0] == 7 && o->comb[1][0] == 7)
			f |= RAS_NEED_NOISE_DUAL;
		if ((f & (RAS_ALPHA_TEST | RAS_ALPHA_TEST_DITHER)) == (RAS_ALPHA_TEST | RAS_ALPHA_TEST_DITHER))
			f |= RAS_NEED_NOISE;
	}

	/* normalize_static_state, rdp_renderer.cpp:1321-1330. */
	if (!(f & (RAS_MULTI_CYCLE | RAS_USES_PIPELINED_TEXEL1)))
		f &= ~(uint32_t)(RAS_BILERP1 | RAS_CONVERT_ONE);
	for (int c = 0; c < 2; c++) {
		const uint8_t *in = o->comb[c];
		uint8_t *out = k->comb[c];
		out[0] = keep_or(in[0], 0xd7u, 8);      /* combined texel0 texel1 shade one noise */
		out[1] = keep_or(in[1], 0x97u, 8);      /* combined texel0 texel1 shade k4 */
		out[2] = keep_or(in[2], 0xab97u, 16);   /* + combined/texel alphas, shade alpha, lod frac, k5 */
		out[3] = keep_or(in[3], 0x57u, 7);      /* combined texel0 texel1 shade one */
		out[4] = keep_or(in[4], 0x57u, 7);
		out[5] = keep_or(in[5], 0x57u, 7);
		out[6] = keep_or(in[6], 0x17u, 7);      /* lod frac, texel0/1 alpha, shade alpha */
		out[7] = keep_or(in[7], 0x57u, 7);
	}
	k->flags = f;
	k->dither = o->dither;
}

static unsigned cache_add(void *keys, unsigned *count, const void *key, size_t size)
{
	uint8_t *base = (uint8_t *)keys;
	for (unsigned i = *count; i-- > 0;)
		if (memcmp(base + i * size, key, size) == 0)
			return i;
	memcpy(base + (*count) * size, key, size);
	return (*count)++;
}

/* interpolate_x, rdp_renderer.cpp:962-992: edges at sub-scanline y of the
 * S-times domain (S = 1 natively). */
static void interpolate_x_64(const tri_t *s, int y, int flip, int64_t S, int64_t *left, int64_t *right)
{
	int64_t yh_base = (int64_t)(s->yh & ~3) * S;
	int64_t ym_base = (int64_t)s->ym * S;
	int64_t xh = S * s->xh + ((int64_t)y - yh_base) * s->dxhdy;
	int64_t xm = S * s->xm + ((int64_t)y - yh_base) * s->dxmdy;
	int64_t xl = S * s->xl + ((int64_t)y - ym_base) * s->dxldy;
	if (y < S * s->ym)
		xl = xm;
	int64_t hs = xh >> 15, ls = xl >> 15;
	int64_t l = flip ? hs : ls, r = flip ? ls : hs;
	if (l < INT32_MIN) l = INT32_MIN;
	if (l > INT32_MAX) l = INT32_MAX;
	if (r < INT32_MIN) r = INT32_MIN;
	if (r > INT32_MAX) r = INT32_MAX;
	*left = l;
	*right = r;
}

/* compute_conservative_max_num_tiles, rdp_renderer.cpp:994-1060: 8x8 tiles
 * of the S-times domain, S the upscaling factor (1 natively). */
static uint32_t conservative_tile_count(const rdpo *o, const tri_t *s)
{
	if (s->yl <= s->yh)
		return 0;
	const int S = 1 << o->up_log2;
	int start_y = imax(s->yh & ~3, (int)o->scissor.ylo) * S;
	int end_y = imin((s->yl - 1) | 3, (int)o->scissor.yhi - 1) * S;
	if (end_y < start_y)
		return 0;
	int flip = (s->flags & TRI_FLIP) != 0;
	int64_t l[4], r[4];
	interpolate_x_64(s, start_y, flip, S, &l[0], &r[0]);
	interpolate_x_64(s, end_y, flip, S, &l[1], &r[1]);
	l[2] = l[3] = l[0];
	r[2] = r[3] = r[0];
	int ym = S * s->ym;
	if (ym > start_y && ym < end_y) {
		interpolate_x_64(s, ym, flip, S, &l[2], &r[2]);
		interpolate_x_64(s, ym - 1, flip, S, &l[3], &r[3]);
	}
	int64_t sx = l[0], ex = r[0];
	for (int i = 1; i < 4; i++) {
		if (l[i] < sx) sx = l[i];
		if (r[i] > ex) ex = r[i];
	}
	int64_t range = (sx < 0 ? -sx : sx) > (ex < 0 ? -ex : ex) ? (sx < 0 ? -sx : sx) : (ex < 0 ? -ex : ex);
	if (range > 2047 * S) {
		sx = 0;
		ex = 0x7fffffff;
	}
	if (sx < (int64_t)S * (o->scissor.xlo >> 2))
		sx = (int64_t)S * (o->scissor.xlo >> 2);
	if (ex > (int64_t)S * ((o->scissor.xhi + 3) >> 2) - 1)
		ex = (int64_t)S * ((o->scissor.xhi + 3) >> 2) - 1;
	if (ex < sx)
		return 0;
	return (uint32_t)((ex / 8 - sx / 8 + 1) * (end_y / 32 - start_y / 32 + 1));
}

/* Per-draw part of draw_shaded_primitive + need_flush,
 * rdp_renderer.cpp:1464-1521, 1523-1590. */
static void pass_account_draw(rdpo *o, const tri_t *setup)
{
	pass_track_t *t = &o->pass;
	t->max_shaded_tiles += conservative_tile_count(o, setup);

	int max_sub = imin(setup->yl - 1, (int)o->scissor.yhi - 1);
	int height = imax((max_sub >> 2) + 1, 0);
	if ((uint32_t)height > t->deduced_height)
		t->deduced_height = (uint32_t)height;

	int min_line = imax(setup->yh, (int)o->scissor.ylo) >> 2;
	int max_line = max_sub >> 2;
	if (max_line >= min_line) {
		int h = imin(imax(max_line - min_line + 2, 0), 1024);
		t->num_span_jobs += (unsigned)((h + 63) / 64);
	}

	static_key_t sk;
	static_state_key(o, setup->tile, &sk);
	cache_add(t->static_keys, &t->num_static, &sk, sizeof(sk));
	depth_blend_key_t dk;
	memset(&dk, 0, sizeof(dk));
	memcpy(dk.blend, o->blend, sizeof(dk.blend));
	dk.flags = o->db_flags;
	dk.coverage_mode = o->coverage_mode;
	dk.z_mode = o->z_mode;
	cache_add(t->depth_blend_keys, &t->num_depth_blend, &dk, sizeof(dk));
	for (int i = 0; i < 8; i++)
		cache_add(t->tile_keys, &t->num_tiles, &o->tiles[i], sizeof(tile_t));

	t->color_write_pending = 1;
	if (o->db_flags & DB_DEPTH_UPDATE)
		t->depth_write_pending = 1;
	t->num_prims++;

	int cache_full = t->num_static == 64 || t->num_depth_blend == 64 || t->num_tiles + 8 > 256;
	int prims_full = t->num_prims == 256;
	int spans_full = t->num_span_jobs * 64 + 1024 > 32 * 1024;
	/* Tile-instance capacity scales with the factor squared
	 * (rdp_renderer.cpp:63-67). */
	const uint32_t S = 1u << o->up_log2;
	int tiles_full = t->max_shaded_tiles + (128 * S) * (128 * S) > S * S * 0x8000;
	if (cache_full || prims_full || spans_full || tiles_full)
		pass_flush(o);
}

/* Renderer::tmem_upload_needs_flush, rdp_renderer.cpp:3051-3097 (compares
 * only the start address of the texture image with the pending images). */
static int tmem_upload_needs_flush(const rdpo *o, uint32_t addr)
{
	const pass_track_t *t = &o->pass;
	uint32_t mask = o->rdram_size - 1;
	if (t->color_write_pending) {
		uint32_t offset = (addr - o->fb_addr) & mask;
		uint32_t pending = t->deduced_height * o->fb_width;
		if (o->fb_fmt == FB_RGBA5551 || o->fb_fmt == FB_I8)
			offset >>= 1;
		else if (o->fb_fmt == FB_RGBA8888)
			offset >>= 2;
		if (offset < pending)
			return 1;
	}
	if (t->depth_write_pending) {
		uint32_t offset = ((addr - o->depth_addr) & mask) >> 1;
		if (offset < t->deduced_height * o->fb_width)
			return 1;
	}
	return 0;
}

static pixel_state_t *pass_pixel(rdpo *o, uint32_t fmt, int x, int y, uint32_t ci, uint32_t di)
{
	if (!o->px) {
		o->px = (pixel_state_t *)malloc(sizeof(pixel_state_t) * MAX_FB_LINES * MAX_FB_WIDTH);
		o->px_loaded = (uint8_t *)calloc(MAX_FB_LINES * MAX_FB_WIDTH, 1);
		o->px_touched = (uint32_t *)malloc(sizeof(uint32_t) * MAX_FB_LINES * MAX_FB_WIDTH);
		if (!o->px || !o->px_loaded || !o->px_touched)
			abort();
	}
	uint32_t key = (uint32_t)y * MAX_FB_WIDTH + (uint32_t)x;
	pixel_state_t *p = &o->px[key];
	if (!o->px_loaded[key]) {
		memset(p, 0, sizeof(*p));
		p->color_index = ci;
		p->depth_index = di;
		load_color(o, fmt, ci, p->color);
		load_depth(o, di, &p->depth, &p->dz);
		o->px_loaded[key] = 1;
		o->px_touched[o->num_touched++] = key;
	}
	return p;
}

/* ------------------------------------------------------------------------ */
/* Upscaled domain.  Renderer::should_render_upscaled and                    */
/* render_pass_is_upscaled (rdp_renderer.cpp:2988-3010), the domain updates  */
/* (submit_update_upscaled_domain, rdp_renderer.cpp:1943-2033;               */
/* update_upscaled_domain_pre.comp / _post.comp) and the upscaled pass.      */
/* ------------------------------------------------------------------------ */
typedef struct up_draw {
	tri_t tri;
	attr_t attr;
	draw_t d;
	tile_t tiles[8];
	uint32_t scissor[4];
	int ylo, yhi;
	uint16_t tmem[2048];
} up_draw_t;

#define UP_MAX_DRAWS 256 /* a pass ends at 256 primitives */

/* Only 16- and 32-bit colour framebuffers of at least 256 pixels are
 * rendered upscaled. */
static int up_pass_eligible(const rdpo *o)
{
	return (o->fb_fmt == FB_RGBA5551 || o->fb_fmt == FB_RGBA8888) && o->fb_width >= 256;
}

static void up_record(rdpo *o, const tri_t *setup, const attr_t *attr, const draw_t *d, const uint32_t scissor[4],
                      int ylo, int yhi)
{
	if (!o->up_draws) {
		o->up_draws = (up_draw_t *)malloc(sizeof(up_draw_t) * UP_MAX_DRAWS);
		if (!o->up_draws)
			abort();
	}
	if (o->up_num >= UP_MAX_DRAWS)
		abort();
	up_draw_t *r = &o->up_draws[o->up_num++];
	r->tri = *setup;
	r->attr = *attr;
	r->d = *d;
	memcpy(r->tiles, o->tiles, sizeof(r->tiles));
	memcpy(r->scissor, scissor, sizeof(r->scissor));
	r->ylo = ylo;
	r->yhi = yhi;
	memcpy(r->tmem, o->tmem, sizeof(r->tmem));
}

/* Point the RDRAM / hidden RDRAM views at sample plane `slice` of the
 * upscaled domain (plane < 0: back to the native memory).  Plane addressing
 * is the native addressing plus slice * size (memory_interfacing.h:55-157),
 * so the native load/store functions serve the planes unchanged. */
static void up_select_plane(rdpo *o, int slice, uint8_t *native_rdram, uint8_t *native_hidden)
{
	if (slice < 0) {
		o->rdram = native_rdram;
		o->hidden = native_hidden;
	} else {
		o->rdram = o->up_rdram + (size_t)slice * o->rdram_size;
		o->hidden = o->up_hidden + (size_t)slice * (o->rdram_size / 2);
	}
}

static pixel_state_t *up_pixel(rdpo *o, uint32_t fmt, int x, int y, uint32_t ci, uint32_t di)
{
	const int L = (int)o->up_log2, S = 1 << L;
	uint32_t key = ((uint32_t)y >> L) * MAX_FB_WIDTH + ((uint32_t)x >> L);
	pixel_state_t *blk = o->upx[key];
	if (!blk) {
		blk = (pixel_state_t *)calloc((size_t)S * S, sizeof(pixel_state_t));
		if (!blk)
			abort();
		o->upx[key] = blk;
		o->upx_touched[o->upx_num++] = key;
	}
	int slice = PLANTED(32) ? (x & (S - 1)) * S + (y & (S - 1)) : (y & (S - 1)) * S + (x & (S - 1));
	pixel_state_t *p = &blk[slice];
	if (!p->loaded) {
		uint8_t *nr = o->rdram, *nh = o->hidden;
		p->color_index = ci;
		p->depth_index = di;
		up_select_plane(o, slice, nr, nh);
		load_color(o, fmt, ci, p->color);
		load_depth(o, di, &p->depth, &p->dz);
		up_select_plane(o, -1, nr, nh);
		p->loaded = 1;
	}
	return p;
}

static void upx_release(rdpo *o)
{
	for (size_t i = 0; i < o->upx_num; i++) {
		free(o->upx[o->upx_touched[i]]);
		o->upx[o->upx_touched[i]] = NULL;
	}
	o->upx_num = 0;
}

static void up_write_back(rdpo *o)
{
	const int S = 1 << o->up_log2;
	uint32_t fmt = o->fb_fmt;
	int alias = o->fb_addr == o->depth_addr;
	uint8_t *nr = o->rdram, *nh = o->hidden;
	for (size_t i = 0; i < o->upx_num; i++) {
		pixel_state_t *blk = o->upx[o->upx_touched[i]];
		for (int slice = 0; slice < S * S; slice++) {
			pixel_state_t *p = &blk[slice];
			if (!p->loaded)
				continue;
			/* emit_scaled_write_masks, memory_interfacing.h:305-336 */
			if (o->up_ssaa)
				o->up_wmask[o->upx_touched[i]] |= (uint8_t)((p->color_dirty ? 1 : 0) | (p->depth_dirty ? 2 : 0));
			up_select_plane(o, slice, nr, nh);
			if (p->color_dirty)
				store_color(o, fmt, p->color_index, p->color);
			if (!alias && p->depth_dirty)
				store_depth(o, p->depth_index, p->depth, p->dz);
		}
	}
	up_select_plane(o, -1, nr, nh);
	upx_release(o);
}

/* One element of update_upscaled_domain_pre.comp (post = 0: where native
 * RDRAM differs from the reference copy, broadcast the native element and
 * its hidden bits to every plane, then update the reference) or
 * update_upscaled_domain_post.comp (post = 1: reference := native).  index
 * counts host-order elements of 1 << size_log2 bytes. */
static void up_sync_element(rdpo *o, int post, unsigned size_log2, uint32_t index)
{
	const unsigned bytes = 1u << size_log2;
	const uint32_t planes = 1u << (2 * o->up_log2);
	const size_t hplane = o->rdram_size >> 1;
	index &= (o->rdram_size >> size_log2) - 1;
	size_t off = (size_t)index << size_log2;
	if (post) {
		memcpy(o->up_ref + off, o->rdram + off, bytes);
		return;
	}
	if (!memcmp(o->rdram + off, o->up_ref + off, bytes))
		return;
	for (uint32_t i = 0; i < planes; i++) {
		memcpy(o->up_rdram + (size_t)i * o->rdram_size + off, o->rdram + off, bytes);
		uint8_t *uh = o->up_hidden + (size_t)i * hplane;
		if (PLANTED(34)) {
		} else if (size_log2 == 0) {
			uint32_t m = index ^ 3u;
			if (m & 1u)
				uh[m >> 1] = o->hidden[m >> 1];
		} else if (size_log2 == 1) {
			uh[index ^ 1u] = o->hidden[index ^ 1u];
		} else {
			uh[2 * index] = o->hidden[2 * index];
			uh[2 * index + 1] = o->hidden[2 * index + 1];
		}
	}
	memcpy(o->up_ref + off, o->rdram + off, bytes);
}

/* submit_update_upscaled_domain (rdp_renderer.cpp:1943-2033) for the pre and
 * post stages: whole words only, colour elements of 1 << size_log2 bytes
 * from addr, depth halfwords from depth_addr unless the two coincide. */
static void up_domain_update(rdpo *o, int post, uint32_t addr, uint32_t depth_addr, uint32_t num_pixels,
                             unsigned size_log2)
{
	addr &= ~3u;
	depth_addr &= ~3u;
	uint32_t align = 4u >> size_log2;
	num_pixels = (num_pixels + align - 1u) & ~(align - 1u);
	int alias = addr == depth_addr;
	uint32_t color_index = addr >> size_log2, depth_index = depth_addr >> 1;
	for (uint32_t i = 0; i < num_pixels; i++) {
		up_sync_element(o, post, size_log2, color_index + i);
		if (!alias)
			up_sync_element(o, post, 1, depth_index + i);
	}
}

static unsigned fb_pixel_size_log2(uint32_t fmt)
{
	return fmt == FB_RGBA8888 ? 2 : (fmt == FB_RGBA5551 || fmt == FB_IA88) ? 1 : 0;
}

static uint16_t host16(const uint8_t *p, size_t index)
{
	uint16_t v;
	memcpy(&v, p + 2 * index, 2);
	return v;
}

static void set_host16(uint8_t *p, size_t index, uint16_t v) { memcpy(p + 2 * index, &v, 2); }

/* The 4x4 ordered-dither offsets of the 16-bit resolve
 * (update_upscaled_domain_resolve.comp:126-130). */
static const uint8_t ssaa_dither_lut[16] = { 0, 4, 1, 5, 4, 0, 5, 1, 3, 7, 2, 6, 7, 3, 6, 2 };

/* update_upscaled_domain_resolve.comp:165-276 over width x height native
 * pixels: a pixel any sample wrote gets the average of its samples in
 * native RDRAM and the domain's reference copy (8-bit: plane 0's byte;
 * depth: plane 0's value); hidden RDRAM is left alone. */
static void up_ssaa_resolve(rdpo *o, uint32_t width, uint32_t height)
{
	const uint32_t N = 1u << (2 * o->up_log2);
	const unsigned size_log2 = fb_pixel_size_log2(o->fb_fmt);
	const uint32_t fb_index = o->fb_addr >> size_log2, depth_index = o->depth_addr >> 1;
	const int alias = o->fb_addr == o->depth_addr;
	const size_t R = o->rdram_size;
	for (uint32_t y = 0; y < height; y++) {
		for (uint32_t x = 0; x < width; x++) {
			const uint32_t key = y * MAX_FB_WIDTH + x;
			const uint8_t m = o->up_wmask[key];
			o->up_wmask[key] = 0;
			const uint32_t index = y * width + x;
			if (m & 1) {
				if (size_log2 == 0) {
					const uint32_t ci = ((fb_index + index) & (uint32_t)(R - 1)) ^ 3u;
					uint32_t r = 0;
					for (uint32_t i = 0; i < N; i++)
						r += o->up_rdram[ci];
					r = (r + (N >> 1)) / N;
					o->rdram[ci] = o->up_ref[ci] = (uint8_t)r;
				} else if (size_log2 == 1) {
					const uint32_t ci = ((fb_index + index) & (uint32_t)(R / 2 - 1)) ^ 1u;
					uint32_t c[4] = { 0, 0, 0, 0 };
					for (uint32_t i = 0; i < N; i++) {
						const uint32_t w = host16(o->up_rdram + (size_t)i * R, ci);
						c[0] += (w >> 11) & 31;
						c[1] += (w >> 6) & 31;
						c[2] += (w >> 1) & 31;
						c[3] += w & 1;
					}
					for (int k = 0; k < 4; k++) {
						if (o->up_ssaa_dither)
							c[k] = (8 * c[k] + ssaa_dither_lut[PLANTED(39) ? (x & 3) * 4 + (y & 3) : (y & 3) * 4 + (x & 3)] * N) /
							       (8 * N);
						else
							c[k] = (c[k] + (PLANTED(38) ? 0 : (N >> 1))) / N;
					}
					const uint16_t e = (uint16_t)((c[0] << 11) | (c[1] << 6) | (c[2] << 1) | c[3]);
					set_host16(o->rdram, ci, e);
					set_host16(o->up_ref, ci, e);
				} else {
					const uint32_t ci = (fb_index + index) & (uint32_t)(R / 4 - 1);
					uint32_t c[4] = { 0, 0, 0, 0 };
					for (uint32_t i = 0; i < N; i++) {
						uint32_t w;
						memcpy(&w, o->up_rdram + (size_t)i * R + 4 * (size_t)ci, 4);
						for (int k = 0; k < 4; k++)
							c[k] += (w >> (24 - 8 * k)) & 0xff;
					}
					uint32_t e = 0;
					for (int k = 0; k < 4; k++)
						e |= ((c[k] + (N >> 1)) / N) << (24 - 8 * k);
					memcpy(o->rdram + 4 * (size_t)ci, &e, 4);
					memcpy(o->up_ref + 4 * (size_t)ci, &e, 4);
				}
			}
			if (!alias && (m & 2)) {
				const uint32_t di = ((depth_index + index) & (uint32_t)(R / 2 - 1)) ^ 1u;
				const uint16_t v = host16(o->up_rdram + (PLANTED(40) ? R : 0), di);
				set_host16(o->rdram, di, v);
				set_host16(o->up_ref, di, v);
			}
		}
	}
}

static void pass_discard(rdpo *o)
{
	for (size_t i = 0; i < o->num_touched; i++)
		o->px_loaded[o->px_touched[i]] = 0;
	o->num_touched = 0;
}

static void render_draw(rdpo *o, const tri_t *setup, const attr_t *attr, const draw_t *d, const tile_t *tiles,
                        const uint32_t scissor[4], int ylo, int yhi, int L);

/* The render-pass end of the upscaled domain (submit_render_pass and
 * submit_render_pass_upscaled, rdp_renderer.cpp:2223-2342, in the order
 * Renderer::flush_queues issues them, :2955-2985): before the native
 * raster, the pre update over the framebuffer region; after it, for a pass
 * that qualifies, the post update and the same draws rendered at scale into
 * the sample planes.  native_write_back runs the native raster's merge. */
static void up_render_recorded(rdpo *o)
{
	uint16_t saved[2048];
	memcpy(saved, o->tmem, sizeof(saved));
	for (unsigned i = 0; i < o->up_num; i++) {
		up_draw_t *r = &o->up_draws[i];
		memcpy(o->tmem, r->tmem, sizeof(saved));
		render_draw(o, &r->tri, &r->attr, &r->d, r->tiles, r->scissor, r->ylo, r->yhi, (int)o->up_log2);
	}
	memcpy(o->tmem, saved, sizeof(saved));
}

static void up_pass_end(rdpo *o, void (*native_write_back)(rdpo *))
{
	int need_render_pass = o->fb_width && o->pass.deduced_height && o->pass.num_span_jobs;
	uint32_t num_pixels = o->fb_width * o->pass.deduced_height;
	unsigned size_log2 = fb_pixel_size_log2(o->fb_fmt);
	if (o->up_ssaa) {
		/* Super-sampled readback (flush_queues without the native pass,
		 * rdp_renderer.cpp:2908-2985, then submit_render_pass_upscaled): a
		 * pass with span jobs or TMEM uploads runs the pre update, renders
		 * at scale with write masks and resolves into native RDRAM. */
		pass_discard(o);
		if (o->pass.num_span_jobs || o->pass.num_uploads) {
			up_domain_update(o, 0, o->fb_addr, o->depth_addr, num_pixels, size_log2);
			up_render_recorded(o);
			up_write_back(o);
			up_ssaa_resolve(o, o->fb_width, o->pass.deduced_height);
		}
		o->up_num = 0;
		return;
	}
	if (need_render_pass)
		up_domain_update(o, 0, o->fb_addr, o->depth_addr, num_pixels, size_log2);
	native_write_back(o);
	if (need_render_pass && up_pass_eligible(o)) {
		if (!PLANTED(33))
			up_domain_update(o, 1, o->fb_addr, o->depth_addr, num_pixels, size_log2);
		up_render_recorded(o);
		up_write_back(o);
	}
	o->up_num = 0;
}

/* The mode a fragment is counted under (rdpo_fragment_counts). */
static int fragment_mode(const draw_t *d)
{
	if (d->ras_flags & RAS_FILL) return 0;
	if (d->ras_flags & RAS_COPY) return 1;
	if (d->ras_flags & RAS_MULTI_CYCLE) return 4;
	return (d->uses_texel0 || d->uses_texel1) ? 3 : 2;
}

/* Rasterise, shade and merge one draw into the open pass's pixel states.
 * L = 0 is the native pass: the span walker over native lines ylo .. yhi
 * (plus the line after, read by the pipelined texel-1 path) and the pixel
 * cache of the native framebuffer.  L > 0 is the upscaled pass
 * (submit_render_pass_upscaled, rdp_renderer.cpp:2287-2342): span lines
 * S * ylo .. S * (yhi + 2) - 1 (span_setup.comp:89-95), pixels of the
 * S-times framebuffer, each merged into its own sample of the native pixel
 * (init_tile/finish_tile, memory_interfacing.h:286-362). */
static void render_draw(rdpo *o, const tri_t *setup, const attr_t *attr, const draw_t *d, const tile_t *tiles,
                        const uint32_t scissor[4], int ylo, int yhi, int L)
{
	const int S = 1 << L;
	int nlines = S * (yhi - ylo + 2);
	span_t *spans = (span_t *)malloc(sizeof(span_t) * (size_t)nlines);
	if (!spans)
		abort();
	for (int i = 0; i < nlines; i++)
		span_setup(setup, attr, scissor, S * ylo + i, &spans[i], L);

	prim_ctx_t pc;
	pc.tri = setup;
	pc.attr = attr;
	pc.d = d;
	pc.tiles = tiles;
	pc.spans = spans;
	pc.ylo = ylo;
	pc.yhi = yhi;
	pc.L = L;
	switch (o->fb_fmt) {
	case FB_I4: pc.fb_size = 0; pc.dx_mask = 0; pc.dx_shift = 0; break;
	case FB_I8: pc.fb_size = 1; pc.dx_mask = ~7u; pc.dx_shift = 3; break;
	case FB_RGBA5551:
	case FB_IA88: pc.fb_size = 2; pc.dx_mask = ~3u; pc.dx_shift = 2; break;
	default: pc.fb_size = 4; pc.dx_mask = ~1u; pc.dx_shift = 1; break;
	}

	uint32_t fmt = o->fb_fmt;
	uint32_t addr_index;
	switch (fmt) {
	case FB_I4:
	case FB_I8: addr_index = o->fb_addr; break;
	case FB_RGBA5551:
	case FB_IA88: addr_index = o->fb_addr >> 1; break;
	default: addr_index = o->fb_addr >> 2; break;
	}
	uint32_t depth_index = o->depth_addr >> 1;
	int alias = o->fb_addr == o->depth_addr;
	/* Native-resolution draws shade every sample as the pixel's first one,
	 * so a covered pixel's samples reach up to S - 1 past the span end. */
	int snap = L && (setup->flags & TRI_DISABLE_UPSCALING);

	for (int y = S * ylo; y <= S * yhi + (S - 1); y++) {
		const span_t *sp = &spans[(snap ? (y & ~(S - 1)) : y) - S * ylo];
		if (!sp->valid)
			continue;
		int x0 = imax(sp->start_x, 0);
		int x1 = imin(sp->end_x + (snap ? S - 1 : 0), S * (int)o->fb_width - 1);
		for (int x = x0; x <= x1; x++) {
			shaded_t sh;
			memset(&sh, 0, sizeof(sh));
			if (!shade_pixel(o, &pc, x, y, &sh))
				continue;
			if (!L) {
				o->fragments++;
				o->fragments_mode[fragment_mode(d)]++;
			}
			uint32_t nx = (uint32_t)x >> L, ny = (uint32_t)y >> L;
			uint32_t ci = addr_index + o->fb_width * ny + nx;
			uint32_t di = depth_index + o->fb_width * ny + nx;
			pixel_state_t *p = L ? up_pixel(o, fmt, x, y, ci, di) : pass_pixel(o, fmt, x, y, ci, di);
			if (sh.coverage_count & COVERAGE_FILL)
				fill_color(p, fmt, alias, d->fill_color);
			else if (sh.coverage_count & COVERAGE_COPY)
				copy_pipeline(p, fmt, alias, (uint32_t)sh.z_dith);
			else
				depth_blend(o, d, fmt, alias, p, &sh);
		}
	}
	free(spans);
}

static void draw_primitive(rdpo *o, tri_t *setup, const attr_t *attr_in)
{
	uint32_t prim_index = (uint32_t)o->prim_counter++;

	/* fixup_triangle_setup, rdp_renderer.cpp:1402-1423 */
	int start_y = setup->yh & ~3;
	if (setup->ym < start_y)
		setup->ym = 0x7fff;
	if (o->ras_flags & RAS_INTERLACE_FIELD) {
		setup->flags |= TRI_INTERLACE_FIELD;
		if (o->ras_flags & RAS_INTERLACE_KEEP_ODD)
			setup->flags |= TRI_INTERLACE_KEEP_ODD;
	}
	if (o->ras_flags & (RAS_COPY | RAS_FILL))
		setup->flags |= TRI_FILL_COPY_RASTER;

	attr_t attr = *attr_in;
	if (o->use_prim_depth) {
		attr.stzw[2] = o->prim_depth;
		attr.dstzw_dx[2] = attr.dstzw_de[2] = attr.dstzw_dy[2] = 0;
	}

	draw_t d;
	resolve_draw_state(o, attr_in, &d);
	d.prim_index = prim_index;

	uint32_t scissor[4] = { o->scissor.xlo, o->scissor.ylo, o->scissor.xhi, o->scissor.yhi };

	/* allocate_span_jobs, rdp_renderer.cpp:1523-1554 */
	int ylo = imax(setup->yh, (int)o->scissor.ylo) >> 2;
	int yhi = imin(setup->yl - 1, (int)o->scissor.yhi - 1) >> 2;
	if (o->fb_width == 0 || yhi < ylo) {
		pass_account_draw(o, setup);
		return;
	}

	render_draw(o, setup, &attr, &d, o->tiles, scissor, ylo, yhi, 0);
	if (o->up_log2 && (o->up_ssaa || up_pass_eligible(o)))
		up_record(o, setup, &attr, &d, scissor, ylo, yhi);
	pass_account_draw(o, setup);
}

/* ------------------------------------------------------------------------ */
/* TMEM uploads.  Host planning: Renderer::load_tile / load_tile_iteration, */
/* rdp_renderer.cpp:3099-3486.  Per-word update: tmem_update.comp:108-577,  */
/* with the float reciprocals replaced by exact integer division.           */
/* ------------------------------------------------------------------------ */
typedef struct {
	int32_t width, height;
	int32_t min_line_stride, max_line_stride; /* divisors behind min/max_t_mod */
	int32_t vram_addr, vram_width, vram_size, vram_effective_width;
	int32_t tmem_offset, tmem_stride_words, tmem_size, tmem_fmt;
	int32_t mode, dxt;
} upload_t;

static inline int div_floor_pos(int32_t num, int32_t den)
{
	/* compute_upload_t for the non-negative offsets it is called with */
	if (den <= 0)
		return 0;
	if (num >= 0)
		return num / den;
	return -((-num + den - 1) / den);
}

static int update_tmem_32(const rdpo *o, const upload_t *info, int idx, int upper, int yuv, uint32_t *value)
{
	int off16 = (info->tmem_offset & 0x7ff) >> 1;
	int stride = info->tmem_stride_words;
	int pixel_offset = (idx - off16) & 0x3ff;
	int ux, uy, ux_xor = 0;

	if (info->mode == UP_BLOCK) {
		int word_offset = pixel_offset >> 1;
		if (info->tmem_stride_words == 0) {
			int c0 = word_offset & ~1, c1 = c0 + 1;
			int t0 = (c0 * info->dxt) >> 16;
			int t1 = (c1 * info->dxt) >> 16;
			if (t0 != t1) {
				int w0 = c0 ^ (t0 & 1), w1 = c1 ^ (t1 & 1);
				if (w1 == word_offset)
					ux_xor = (t1 & 1) << 1;
				else if (w0 == word_offset)
					ux_xor = (t0 & 1) << 1;
				else
					return 0;
			} else
				ux_xor ^= (t0 & 1) << 1;
		} else {
			int min_t = div_floor_pos(word_offset & ~1, info->max_line_stride);
			int max_t = div_floor_pos(word_offset | 1, info->min_line_stride);
			int max_cand = (word_offset | 1) - stride * min_t;
			int min_cand = (word_offset & ~1) - stride * max_t;
			min_t = imax(min_t, (min_cand * info->dxt) >> 16);
			max_t = imin(max_t, (max_cand * info->dxt) >> 16);
			int found = 0;
			for (int t = max_t; t >= min_t; t--) {
				int cs0 = (word_offset & ~1) - stride * t;
				int cs1 = (word_offset | 1) - stride * t;
				int ct0 = (cs0 * info->dxt) >> 16;
				int ct1 = (cs1 * info->dxt) >> 16;
				if (((cs1 + ct1 * stride) ^ (ct1 & 1)) == word_offset) {
					found = 1;
					pixel_offset = (cs1 << 1) + (pixel_offset & 1);
					break;
				} else if (((cs0 + ct0 * stride) ^ (ct0 & 1)) == word_offset) {
					found = 1;
					pixel_offset = (cs0 << 1) + (pixel_offset & 1);
					break;
				}
			}
			if (!found)
				return 0;
		}
		ux = pixel_offset;
		uy = 0;
	} else if (stride == 0) {
		ux = pixel_offset;
		uy = info->height - 1;
	} else {
		uy = div_floor_pos(pixel_offset, stride);
		ux = pixel_offset - uy * stride;
		if (uy >= info->height) {
			ux += stride * (uy - info->height + 1);
			uy = info->height - 1;
		}
	}

	int last_line_x = ux ^ ((uy & 1) << 1);
	if (last_line_x >= info->width && uy > 0) {
		uy--;
		ux += stride;
	}

	int iter_off = 0;
	ux ^= ((uy & 1) << 1) | ux_xor;

	if (info->vram_size == 3 || yuv)
		iter_off = 4 * (ux & ~1);
	else if (info->vram_size == 2) {
		if (ux & 2) {
			if (uy > 0) {
				uy--;
				ux += stride;
				ux ^= 2;
			} else
				return 0;
		}
		iter_off = 2 * (ux & ~1);
	} else if (info->vram_size == 1) {
		for (int i = 0; i < 4 && uy > 0 && (ux & 6) != 0; i++) {
			uy--;
			ux += stride;
			ux ^= 2;
		}
		if (ux & 6)
			return 0;
		iter_off = ux & ~1;
	}

	if (ux >= info->width)
		return 0;

	int line_addr = info->vram_addr + ((uy * info->vram_width) << (info->vram_size - 1));
	uint32_t addr = (uint32_t)(line_addr + iter_off + 4 * (ux & 1));
	uint32_t word;
	if ((addr & 3) == 0)
		word = rd32_index(o, addr >> 2);
	else
		word = ((uint32_t)rd8(o, addr) << 24) | ((uint32_t)rd8(o, addr + 1) << 16) | ((uint32_t)rd8(o, addr + 2) << 8) |
		       rd8(o, addr + 3);

	if (yuv) {
		if (upper) {
			uint32_t y0 = (word >> 16) & 0xff, y1 = word & 0xff;
			word = (y0 << 8) | y1;
		} else {
			uint32_t u = (word >> 24) & 0xff, v = (word >> 8) & 0xff;
			word = (u << 8) | v;
		}
	} else {
		word >>= 16 - 16 * (uint32_t)upper;
		word &= 0xffff;
	}
	*value = word;
	return 1;
}

static int update_tmem_16(const rdpo *o, const upload_t *info, int idx, uint32_t *value)
{
	int off16 = (info->tmem_offset & 0xfff) >> 1;
	int stride = info->tmem_stride_words;
	int pixel_offset = (idx - off16) & 0x7ff;
	int ux, uy, ux_xor = 0;

	if (info->mode == UP_BLOCK) {
		int word_offset = pixel_offset >> 2;
		if (info->tmem_stride_words == 0)
			ux_xor = (((word_offset * info->dxt) >> 16) & 1) << 1;
		else {
			int min_t = div_floor_pos(word_offset, info->max_line_stride);
			int max_t = div_floor_pos(word_offset, info->min_line_stride);
			int max_cand = word_offset - stride * min_t;
			int min_cand = word_offset - stride * max_t;
			min_t = imax(min_t, (min_cand * info->dxt) >> 16);
			max_t = imin(max_t, (max_cand * info->dxt) >> 16);
			int found = 0;
			for (int t = max_t; t >= min_t; t--) {
				int cand = word_offset - stride * t;
				int ct = (cand * info->dxt) >> 16;
				if (cand + ct * stride == word_offset) {
					found = 1;
					ux_xor = (ct & 1) << 1;
					pixel_offset = (cand << 2) + (pixel_offset & 3);
				}
			}
			if (!found)
				return 0;
		}
		ux = pixel_offset;
		uy = 0;
	} else if (stride == 0) {
		ux = pixel_offset;
		uy = info->height - 1;
	} else {
		uy = div_floor_pos(pixel_offset, stride);
		ux = pixel_offset - uy * stride;
		if (uy >= info->height) {
			ux += stride * (uy - info->height + 1);
			uy = info->height - 1;
		}
	}

	int iter_off = 0;
	if (info->tmem_size != info->vram_size) {
		if (info->vram_size - info->tmem_size == 1) {
			iter_off = (ux & ~3) * 4;
			if ((ux & ~3) + 2 < (info->vram_effective_width >> (3 - info->vram_size)))
				iter_off += 8;
		} else if (info->tmem_size == 2 && info->vram_size == 1) {
			if (ux & 4) {
				if ((stride & 4) != 0 && uy > 0) {
					uy--;
					ux += stride;
				} else
					return 0;
			}
			iter_off = ux & ~3;
		}
	} else
		iter_off = (ux & ~3) * 2;

	if (ux >= info->width)
		return 0;

	int line_addr = info->vram_addr + ((uy * info->vram_width) << (info->vram_size - 1));
	ux ^= ((uy & 1) << 1) | ux_xor;
	uint32_t addr = (uint32_t)(line_addr + iter_off + 2 * (ux & 3));
	*value = rd16_addr(o, addr);
	return 1;
}

static int update_tmem_lut(const rdpo *o, const upload_t *info, int idx, uint32_t *value)
{
	int off16 = (info->tmem_offset & 0xfff) >> 1;
	int p = (idx - off16) & 0x7ff;
	int splat;
	int diff = info->vram_size - info->tmem_size;
	if (diff == 2) {
		splat = p >> 2;
		splat <<= info->vram_size - 2;
		if (splat >= info->vram_effective_width)
			return 0;
	} else if (diff == 1) {
		if ((p & 4) == 0) {
			int shamt = info->tmem_size + (info->vram_size == 2 ? 2 : 0);
			splat = (p & ~7) >> shamt;
			if (splat >= info->vram_effective_width)
				return 0;
		} else
			return 0;
	} else if (diff == 0) {
		if ((p & 0xc) == 0) {
			int shamt = info->tmem_size + (info->vram_size == 2 ? 2 : 0);
			splat = (p & ~3) >> shamt;
			if (splat >= info->vram_effective_width)
				return 0;
		} else
			return 0;
	} else if (diff == -1) {
		if ((p & 0x1c) == 0) {
			splat = (p >> info->tmem_size) & ~7;
			if (splat >= info->vram_effective_width)
				return 0;
		} else
			return 0;
	} else {
		int span_pixel = (p >> 2) * 4;
		if (span_pixel + 2 < info->vram_effective_width)
			span_pixel += 2;
		if (span_pixel >= info->vram_effective_width)
			return 0;
		splat = span_pixel;
	}
	uint32_t addr = (uint32_t)(info->vram_addr + (splat << (info->vram_size - 1)));
	addr += 2 * (addr & 1) * (uint32_t)(p & 3);
	*value = rd16_addr(o, addr);
	return 1;
}

static void apply_upload(rdpo *o, const upload_t *info)
{
	for (int g = 0; g < 2048; g++) {
		int idx = g ^ 1;
		int upper = idx >= 0x400;
		uint32_t v;
		int written;
		if (info->mode == UP_TLUT)
			written = update_tmem_lut(o, info, idx, &v);
		else {
			int yuv = info->tmem_fmt == TF_YUV;
			if (info->tmem_size == 3 || yuv)
				written = update_tmem_32(o, info, idx & 0x3ff, upper, yuv, &v);
			else
				written = update_tmem_16(o, info, idx, &v);
		}
		if (written)
			o->tmem[g] = (uint16_t)v;
	}
}

typedef struct {
	uint32_t tex_addr, tex_width, fmt, size;
	uint32_t slo, shi, tlo, thi;
	uint32_t mode;
} load_info_t;

static void load_tile_iteration(rdpo *o, uint32_t tile, const load_info_t *info, uint32_t tmem_offset)
{
	tile_t *t = &o->tiles[tile];
	t->slo = info->slo;
	t->shi = info->shi;
	t->tlo = info->tlo;
	t->thi = info->thi;

	if (t->fmt == TF_YUV && (t->size != 2 || info->size != 2))
		return;
	if (info->size == 0)
		return;
	if (t->size == 3 && t->fmt != TF_RGBA)
		return;
	if (info->mode == UP_TLUT && t->size == 3)
		return;
	if (info->mode != UP_TLUT) {
		if ((info->size == 3 && t->size == 1) || (info->size == 2 && t->size == 0) || (info->size == 3 && t->size == 0))
			return;
	}

	upload_t up;
	memset(&up, 0, sizeof(up));
	up.tmem_stride_words = (int32_t)(t->stride >> 1);
	uint32_t ux = 0, uy = 0;
	uint32_t mode = info->mode;

	if (mode == UP_BLOCK) {
		ux = info->slo;
		uy = info->tlo;
		unsigned pixel_count = (info->shi - info->slo + 1) & 0xfff;
		unsigned dt = info->thi;
		unsigned max_iter = (pixel_count - 1) >> (4u - info->size);
		unsigned max_t = (max_iter * dt) >> 11;
		if (max_t != 0) {
			unsigned max_elems = ((1u << 11) + dt - 1u) / dt;
			unsigned min_elems = (1u << 11) / dt;
			int uneven = max_elems != min_elems;
			if (uneven) {
				unsigned overflow_amt = dt * max_elems - (1u << 11);
				overflow_amt *= max_t;
				if (overflow_amt < dt) {
					min_elems = max_elems;
					uneven = 0;
				}
			}
			up.dxt = (int32_t)(dt << 5);
			if (t->size == 3 || t->fmt == TF_YUV) {
				up.tmem_stride_words <<= 1;
				if (uneven && t->size != info->size)
					return;
			}
			if (t->size > info->size) {
				unsigned sh = t->size - info->size;
				max_elems <<= sh;
				min_elems <<= sh;
				up.dxt >>= sh;
			} else if (info->size > t->size)
				return;
			up.max_line_stride = (int32_t)(max_elems + ((unsigned)up.tmem_stride_words >> 2));
			up.min_line_stride = (int32_t)(min_elems + ((unsigned)up.tmem_stride_words >> 2));
			up.width = (int32_t)pixel_count;
			up.height = 1;
			up.tmem_stride_words >>= 2;
		} else {
			up.width = (int32_t)pixel_count;
			up.height = 1;
			up.tmem_stride_words = 0;
			mode = UP_TILE;
		}
	} else {
		ux = info->slo >> 2;
		uy = info->tlo >> 2;
		up.width = (int32_t)((((info->shi >> 2) - (info->slo >> 2)) + 1) & 0xfff);
		up.height = (int32_t)(((info->thi >> 2) - (info->tlo >> 2)) + 1);
	}

	if (!up.width)
		return;

	switch (info->size) {
	case 1: up.vram_effective_width = (up.width + 7) & ~7; break;
	case 2: up.vram_effective_width = mode == UP_TLUT ? up.width : ((up.width + 3) & ~3); break;
	case 3: up.vram_effective_width = (up.width + 1) & ~1; break;
	default: break;
	}

	switch (t->size) {
	case 0: up.width = ((up.width + 15) & ~15) >> 2; break;
	case 1: up.width = ((up.width + 7) & ~7) >> 1; break;
	case 2:
		up.width = (up.width + 3) & ~3;
		if (t->fmt == TF_YUV)
			up.width >>= 1;
		break;
	default: up.width = (up.width + 1) & ~1; break;
	}

	if (up.height > 1 && mode == UP_TLUT)
		return;

	up.vram_addr = (int32_t)(info->tex_addr + ((info->tex_width * uy + ux) << (info->size - 1)));
	up.vram_width = mode == UP_BLOCK ? up.vram_effective_width : (int32_t)info->tex_width;
	up.vram_size = (int32_t)info->size;
	up.tmem_offset = (int32_t)((t->offset + tmem_offset) & 0xfff);
	up.tmem_size = (int32_t)t->size;
	up.tmem_fmt = (int32_t)t->fmt;
	up.mode = (int32_t)mode;
	apply_upload(o, &up);
	/* One TMEM instance per upload; rdp_renderer.cpp:3482-3485. */
	if (++o->pass.num_uploads + 1 >= 256)
		pass_flush(o);
}

static void load_tile(rdpo *o, uint32_t tile, const load_info_t *info)
{
	if (tmem_upload_needs_flush(o, info->tex_addr))
		pass_flush(o);
	if (info->mode != UP_BLOCK) {
		if ((info->thi >> 2) < (info->tlo >> 2))
			return;
		unsigned pc = (((info->shi >> 2) - (info->slo >> 2)) + 1) & 0xfff;
		if (!pc)
			return;
	} else {
		unsigned pc = ((info->shi - info->slo) + 1) & 0xfff;
		if (!pc || pc > 2048)
			return;
	}

	if (info->mode == UP_TILE) {
		tile_t *t = &o->tiles[tile];
		unsigned px_per_line = (((info->shi >> 2) - (info->slo >> 2)) + 1) & 0xfff;
		unsigned qw_per_line = ((px_per_line << t->size) + 15) >> 4;
		if (t->size > info->size)
			qw_per_line <<= t->size - info->size;
		else if (t->size < info->size)
			qw_per_line >>= info->size - t->size;
		unsigned bytes_per_line = qw_per_line * 8 > t->stride ? qw_per_line * 8 : t->stride;
		unsigned max_bytes = 0x1000;
		if (t->fmt == TF_YUV)
			max_bytes /= 2;
		unsigned num_lines = ((info->thi >> 2) - (info->tlo >> 2)) + 1;
		unsigned total = bytes_per_line * num_lines;
		if (total > max_bytes) {
			unsigned max_lines = max_bytes / bytes_per_line;
			max_lines &= ~1u;
			if (max_lines == 0)
				return;
			for (unsigned line = 0; line < num_lines; line += max_lines) {
				unsigned n = num_lines - line < max_lines ? num_lines - line : max_lines;
				load_info_t tmp = *info;
				tmp.tlo = info->tlo + (line << 2);
				tmp.thi = tmp.tlo + ((n - 1) << 2);
				load_tile_iteration(o, tile, &tmp, line * t->stride);
			}
			t->slo = info->slo;
			t->shi = info->shi;
			t->tlo = info->tlo;
			t->thi = info->thi;
		} else
			load_tile_iteration(o, tile, info, 0);
	} else
		load_tile_iteration(o, tile, info, 0);
}

/* ------------------------------------------------------------------------ */
/* Command decode: rdp_device.cpp:268-959.                                  */
/* ------------------------------------------------------------------------ */
unsigned rdpo_command_length(uint32_t first_word)
{
	static const uint8_t lengths[64] = {
		2, 0, 0, 0, 2, 0, 0, 0, 8, 12, 24, 28, 24, 28, 40, 44,
		0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
		0, 0, 0, 0, 4, 4, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
		2, 0, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
	};
	return lengths[(first_word >> 24) & 63];
}

static void decode_triangle(const rdpo *o, tri_t *s, const uint32_t *w)
{
	memset(s, 0, sizeof(*s));
	int copy_cycle = (o->ras_flags & RAS_COPY) != 0;
	int flip = (w[0] & 0x800000u) != 0;
	int sign_dxhdy = (w[5] & 0x80000000u) != 0;
	int do_offset = flip == sign_dxhdy;
	if (flip)
		s->flags |= TRI_FLIP;
	if (do_offset)
		s->flags |= TRI_DO_OFFSET;
	if (copy_cycle)
		s->flags |= TRI_SKIP_XFRAC;
	if (o->quirks & 0x100)
		s->flags |= TRI_NATIVE_LOD;
	s->tile = (w[0] >> 16) & 63;
	s->yl = sext((int32_t)w[0], 14);
	s->ym = sext((int32_t)(w[1] >> 16), 14);
	s->yh = sext((int32_t)w[1], 14);
	s->xl = sext((int32_t)w[2], 28) >> 1;
	s->xh = sext((int32_t)w[4], 28) >> 1;
	s->xm = sext((int32_t)w[6], 28) >> 1;
	s->dxldy = sext((int32_t)(w[3] >> 2), 28) >> 1;
	s->dxhdy = sext((int32_t)(w[5] >> 2), 28) >> 1;
	s->dxmdy = sext((int32_t)(w[7] >> 2), 28) >> 1;
}

static void decode_rgba(attr_t *a, const uint32_t *w)
{
	for (int i = 0; i < 2; i++) {
		a->rgba[2 * i + 0] = (int32_t)((w[i] & 0xffff0000u) | ((w[4 + i] >> 16) & 0xffff));
		a->rgba[2 * i + 1] = (int32_t)((w[i] << 16) | (w[4 + i] & 0xffff));
		a->drgba_dx[2 * i + 0] = (int32_t)((w[2 + i] & 0xffff0000u) | ((w[6 + i] >> 16) & 0xffff));
		a->drgba_dx[2 * i + 1] = (int32_t)((w[2 + i] << 16) | (w[6 + i] & 0xffff));
		a->drgba_de[2 * i + 0] = (int32_t)((w[8 + i] & 0xffff0000u) | ((w[12 + i] >> 16) & 0xffff));
		a->drgba_de[2 * i + 1] = (int32_t)((w[8 + i] << 16) | (w[12 + i] & 0xffff));
		a->drgba_dy[2 * i + 0] = (int32_t)((w[10 + i] & 0xffff0000u) | ((w[14 + i] >> 16) & 0xffff));
		a->drgba_dy[2 * i + 1] = (int32_t)((w[10 + i] << 16) | (w[14 + i] & 0xffff));
	}
}

static void decode_tex(attr_t *a, const uint32_t *w)
{
	a->stzw[0] = (int32_t)((w[0] & 0xffff0000u) | ((w[4] >> 16) & 0xffff));
	a->stzw[1] = (int32_t)(((w[0] << 16) & 0xffff0000u) | (w[4] & 0xffff));
	a->stzw[3] = (int32_t)((w[1] & 0xffff0000u) | ((w[5] >> 16) & 0xffff));
	a->dstzw_dx[0] = (int32_t)((w[2] & 0xffff0000u) | ((w[6] >> 16) & 0xffff));
	a->dstzw_dx[1] = (int32_t)(((w[2] << 16) & 0xffff0000u) | (w[6] & 0xffff));
	a->dstzw_dx[3] = (int32_t)((w[3] & 0xffff0000u) | ((w[7] >> 16) & 0xffff));
	a->dstzw_de[0] = (int32_t)((w[8] & 0xffff0000u) | ((w[12] >> 16) & 0xffff));
	a->dstzw_de[1] = (int32_t)(((w[8] << 16) & 0xffff0000u) | (w[12] & 0xffff));
	a->dstzw_de[3] = (int32_t)((w[9] & 0xffff0000u) | ((w[13] >> 16) & 0xffff));
	a->dstzw_dy[0] = (int32_t)((w[10] & 0xffff0000u) | ((w[14] >> 16) & 0xffff));
	a->dstzw_dy[1] = (int32_t)(((w[10] << 16) & 0xffff0000u) | (w[14] & 0xffff));
	a->dstzw_dy[3] = (int32_t)((w[11] & 0xffff0000u) | ((w[15] >> 16) & 0xffff));
}

static void decode_z(attr_t *a, const uint32_t *w)
{
	a->stzw[2] = (int32_t)w[0];
	a->dstzw_dx[2] = (int32_t)w[1];
	a->dstzw_de[2] = (int32_t)w[2];
	a->dstzw_dy[2] = (int32_t)w[3];
}

static void op_triangle(rdpo *o, unsigned op, const uint32_t *w)
{
	tri_t s;
	attr_t a;
	memset(&a, 0, sizeof(a));
	decode_triangle(o, &s, w);
	const uint32_t *p = w + 8;
	int shade = (op & 4) != 0, tex = (op & 2) != 0, zbuf = (op & 1) != 0;
	if (shade) {
		decode_rgba(&a, p);
		p += 16;
	}
	if (tex) {
		decode_tex(&a, p);
		p += 16;
	}
	if (zbuf)
		decode_z(&a, p);
	draw_primitive(o, &s, &a);
}

static void op_rectangle(rdpo *o, unsigned op, const uint32_t *w)
{
	uint32_t xl = (w[0] >> 12) & 0xfff, yl = w[0] & 0xfff;
	uint32_t xh = (w[1] >> 12) & 0xfff, yh = w[1] & 0xfff;
	if (o->ras_flags & (RAS_COPY | RAS_FILL))
		yl |= 3;
	tri_t s;
	attr_t a;
	memset(&s, 0, sizeof(s));
	memset(&a, 0, sizeof(a));
	s.xh = (int32_t)(xh << 13);
	s.xl = (int32_t)(xl << 13);
	s.xm = (int32_t)(xl << 13);
	s.ym = (int32_t)yl;
	s.yl = (int32_t)yl;
	s.yh = (int32_t)yh;
	if (op == 0x36) {
		s.flags = TRI_FLIP | TRI_DISABLE_UPSCALING;
	} else {
		uint32_t tile = (w[1] >> 24) & 7;
		int32_t ss = (int32_t)((w[2] >> 16) & 0xffff), tt = (int32_t)(w[2] & 0xffff);
		int32_t dsdx = sext((int32_t)((w[3] >> 16) & 0xffff), 16), dtdy = sext((int32_t)(w[3] & 0xffff), 16);
		s.flags = TRI_FLIP;
		if (op == 0x25 || (o->quirks & 1))
			s.flags |= TRI_DISABLE_UPSCALING;
		if (o->quirks & 0x100)
			s.flags |= TRI_NATIVE_LOD;
		s.tile = tile;
		a.stzw[0] = ss << 16;
		a.stzw[1] = tt << 16;
		if (op == 0x24) {
			a.dstzw_dx[0] = dsdx << 11;
			a.dstzw_de[1] = dtdy << 11;
			a.dstzw_dy[1] = dtdy << 11;
		} else {
			a.dstzw_dx[1] = dtdy << 11;
			a.dstzw_de[0] = dsdx << 11;
			a.dstzw_dy[0] = dsdx << 11;
		}
		if (o->ras_flags & RAS_COPY)
			s.flags |= TRI_SKIP_XFRAC;
	}
	draw_primitive(o, &s, &a);
}

#define SET_FLAG(var, cond, bit) do { if (cond) (var) |= (bit); else (var) &= ~(uint32_t)(bit); } while (0)

static void op_set_other_modes(rdpo *o, const uint32_t *w)
{
	uint32_t w0 = w[0], w1 = w[1];
	SET_FLAG(o->ras_flags, w0 & (1 << 19), RAS_PERSPECTIVE);
	SET_FLAG(o->ras_flags, w0 & (1 << 18), RAS_DETAIL_LOD);
	SET_FLAG(o->ras_flags, w0 & (1 << 17), RAS_SHARPEN_LOD);
	SET_FLAG(o->ras_flags, w0 & (1 << 16), RAS_TEX_LOD);
	SET_FLAG(o->ras_flags, w0 & (1 << 15), RAS_TLUT);
	SET_FLAG(o->ras_flags, w0 & (1 << 14), RAS_TLUT_TYPE);
	SET_FLAG(o->ras_flags, w0 & (1 << 13), RAS_SAMPLE_QUAD);
	SET_FLAG(o->ras_flags, w0 & (1 << 12), RAS_MID_TEXEL);
	SET_FLAG(o->ras_flags, w0 & (1 << 11), RAS_BILERP0);
	SET_FLAG(o->ras_flags, w0 & (1 << 10), RAS_BILERP1);
	SET_FLAG(o->ras_flags, w0 & (1 << 9), RAS_CONVERT_ONE);
	SET_FLAG(o->db_flags, w1 & (1 << 14), DB_FORCE_BLEND);
	SET_FLAG(o->ras_flags, w1 & (1 << 13), RAS_ALPHA_CVG_SELECT);
	SET_FLAG(o->ras_flags, w1 & (1 << 12), RAS_CVG_TIMES_ALPHA);
	SET_FLAG(o->db_flags, w1 & (1 << 7), DB_COLOR_ON_CVG);
	SET_FLAG(o->db_flags, w1 & (1 << 6), DB_IMAGE_READ);
	SET_FLAG(o->db_flags, w1 & (1 << 5), DB_DEPTH_UPDATE);
	SET_FLAG(o->db_flags, w1 & (1 << 4), DB_DEPTH_TEST);
	SET_FLAG(o->ras_flags, w1 & (1 << 3), RAS_AA);
	SET_FLAG(o->db_flags, w1 & (1 << 3), DB_AA);
	SET_FLAG(o->ras_flags, w1 & (1 << 1), RAS_ALPHA_TEST_DITHER);
	SET_FLAG(o->ras_flags, w1 & (1 << 0), RAS_ALPHA_TEST);
	o->dither = (w0 >> 4) & 0xf;
	SET_FLAG(o->db_flags, (o->dither >> 2) != 3, DB_DITHER);
	o->coverage_mode = (w1 >> 8) & 3;
	o->z_mode = (w1 >> 10) & 3;
	o->ras_flags &= ~(uint32_t)(RAS_MULTI_CYCLE | RAS_FILL | RAS_COPY);
	o->db_flags &= ~(uint32_t)DB_MULTI_CYCLE;
	switch ((w0 >> 20) & 3) {
	case 1: o->ras_flags |= RAS_MULTI_CYCLE; o->db_flags |= DB_MULTI_CYCLE; break;
	case 2: o->ras_flags |= RAS_COPY; break;
	case 3: o->ras_flags |= RAS_FILL; break;
	default: break;
	}
	o->blend[0][0] = (w1 >> 30) & 3;
	o->blend[1][0] = (w1 >> 28) & 3;
	o->blend[0][1] = (w1 >> 26) & 3;
	o->blend[1][1] = (w1 >> 24) & 3;
	o->blend[0][2] = (w1 >> 22) & 3;
	o->blend[1][2] = (w1 >> 20) & 3;
	o->blend[0][3] = (w1 >> 18) & 3;
	o->blend[1][3] = (w1 >> 16) & 3;
	o->use_prim_depth = (w1 >> 2) & 1;
}

static void op_set_combine(rdpo *o, const uint32_t *w)
{
	uint32_t w0 = w[0], w1 = w[1];
	o->comb[0][0] = (w0 >> 20) & 0xf;
	o->comb[0][2] = (w0 >> 15) & 0x1f;
	o->comb[0][1] = (w1 >> 28) & 0xf;
	o->comb[0][3] = (w1 >> 15) & 0x7;
	o->comb[0][4] = (w0 >> 12) & 0x7;
	o->comb[0][5] = (w1 >> 12) & 0x7;
	o->comb[0][6] = (w0 >> 9) & 0x7;
	o->comb[0][7] = (w1 >> 9) & 0x7;
	o->comb[1][0] = (w0 >> 5) & 0xf;
	o->comb[1][2] = (w0 >> 0) & 0x1f;
	o->comb[1][1] = (w1 >> 24) & 0xf;
	o->comb[1][3] = (w1 >> 6) & 0x7;
	o->comb[1][4] = (w1 >> 21) & 0x7;
	o->comb[1][5] = (w1 >> 3) & 0x7;
	o->comb[1][6] = (w1 >> 18) & 0x7;
	o->comb[1][7] = (w1 >> 0) & 0x7;
}

static void op_set_tile(rdpo *o, const uint32_t *w)
{
	uint32_t tile = (w[1] >> 24) & 7;
	tile_t *t = &o->tiles[tile];
	t->offset = ((w[0] >> 0) & 511) << 3;
	t->stride = ((w[0] >> 9) & 511) << 3;
	t->size = (w[0] >> 19) & 3;
	t->fmt = (w[0] >> 21) & 7;
	t->palette = (w[1] >> 20) & 15;
	t->shift_s = (w[1] >> 0) & 15;
	t->mask_s = (w[1] >> 4) & 15;
	t->shift_t = (w[1] >> 10) & 15;
	t->mask_t = (w[1] >> 14) & 15;
	t->flags = 0;
	if (w[1] & (1 << 8)) t->flags |= TILE_MIRROR_S;
	if (w[1] & (1 << 9)) t->flags |= TILE_CLAMP_S;
	if (w[1] & (1 << 18)) t->flags |= TILE_MIRROR_T;
	if (w[1] & (1 << 19)) t->flags |= TILE_CLAMP_T;
	if (t->mask_s > 10) t->mask_s = 10;
	else if (t->mask_s == 0) t->flags |= TILE_CLAMP_S;
	if (t->mask_t > 10) t->mask_t = 10;
	else if (t->mask_t == 0) t->flags |= TILE_CLAMP_T;
}

static void op_load(rdpo *o, uint32_t mode, const uint32_t *w)
{
	load_info_t info;
	info.tex_addr = o->teximg.addr;
	info.tex_width = o->teximg.width;
	info.fmt = o->teximg.fmt;
	info.size = o->teximg.size;
	info.slo = (w[0] >> 12) & 0xfff;
	info.shi = (w[1] >> 12) & 0xfff;
	info.tlo = w[0] & 0xfff;
	info.thi = w[1] & 0xfff;
	info.mode = mode;
	load_tile(o, (w[1] >> 24) & 7, &info);
}

static void execute(rdpo *o, const uint32_t *w)
{
	unsigned op = (w[0] >> 24) & 63;
	switch (op) {
	case 0x08: case 0x09: case 0x0a: case 0x0b: case 0x0c: case 0x0d: case 0x0e: case 0x0f:
		op_triangle(o, op, w);
		break;
	case 0x24: case 0x25: case 0x36:
		op_rectangle(o, op, w);
		break;
	case 0x2a: /* SetKeyGB */
		o->key_width[1] = (w[0] >> 12) & 0xfff;
		o->key_width[2] = w[0] & 0xfff;
		o->key_center[1] = (w[1] >> 24) & 0xff;
		o->key_scale[1] = (w[1] >> 16) & 0xff;
		o->key_center[2] = (w[1] >> 8) & 0xff;
		o->key_scale[2] = w[1] & 0xff;
		break;
	case 0x2b: /* SetKeyR */
		o->key_width[0] = (w[1] >> 16) & 0xfff;
		o->key_center[0] = (w[1] >> 8) & 0xff;
		o->key_scale[0] = w[1] & 0xff;
		break;
	case 0x2c: { /* SetConvert */
		uint64_t m = ((uint64_t)w[0] << 32) | w[1];
		int32_t k[6];
		for (int i = 0; i < 6; i++)
			k[i] = (int32_t)((m >> (45 - 9 * i)) & 0x1ff);
		for (int i = 0; i < 4; i++)
			o->convert[i] = 2 * sext(k[i], 9) + 1;
		o->convert[4] = k[4];
		o->convert[5] = k[5];
		break;
	}
	case 0x2d: /* SetScissor */
		o->scissor.xlo = (w[0] >> 12) & 0xfff;
		o->scissor.xhi = (w[1] >> 12) & 0xfff;
		o->scissor.ylo = w[0] & 0xfff;
		o->scissor.yhi = w[1] & 0xfff;
		SET_FLAG(o->ras_flags, w[1] & (1 << 25), RAS_INTERLACE_FIELD);
		SET_FLAG(o->ras_flags, w[1] & (1 << 24), RAS_INTERLACE_KEEP_ODD);
		break;
	case 0x2e: /* SetPrimDepth */
		o->prim_depth = (int32_t)((w[1] >> 16) & 0x7fff) << 16;
		o->prim_dz = w[1] & 0xffff;
		break;
	case 0x2f:
		op_set_other_modes(o, w);
		break;
	case 0x30:
		op_load(o, UP_TLUT, w);
		break;
	case 0x32: { /* SetTileSize */
		tile_t *t = &o->tiles[(w[1] >> 24) & 7];
		t->slo = (w[0] >> 12) & 0xfff;
		t->shi = (w[1] >> 12) & 0xfff;
		t->tlo = w[0] & 0xfff;
		t->thi = w[1] & 0xfff;
		break;
	}
	case 0x33:
		op_load(o, UP_BLOCK, w);
		break;
	case 0x34:
		op_load(o, UP_TILE, w);
		break;
	case 0x35:
		op_set_tile(o, w);
		break;
	case 0x37: o->fill_color = w[1]; break;
	case 0x38: o->fog_color = w[1]; break;
	case 0x39: o->blend_color = w[1]; break;
	case 0x3a:
		o->min_level = (w[0] >> 8) & 31;
		o->prim_lod_frac = w[0] & 0xff;
		o->prim_color = w[1];
		break;
	case 0x3b: o->env_color = w[1]; break;
	case 0x3c:
		op_set_combine(o, w);
		break;
	case 0x3d:
		o->teximg.fmt = (w[0] >> 21) & 7;
		o->teximg.size = (w[0] >> 19) & 3;
		o->teximg.width = (w[0] & 0x3ff) + 1;
		o->teximg.addr = w[1] & 0x00ffffffu;
		break;
	case 0x29: /* SyncFull: flush_and_signal, rdp_device.cpp:263-266 */
		pass_flush(o);
		break;
	case 0x3e: /* set_depth_framebuffer, rdp_renderer.cpp:688-694 */
		if ((w[1] & 0xffffff) != o->depth_addr)
			pass_flush(o);
		o->depth_addr = w[1] & 0xffffff;
		break;
	case 0x3f: { /* set_color_framebuffer, rdp_renderer.cpp:679-686 */
		unsigned fmt = (w[0] >> 21) & 7, size = (w[0] >> 19) & 3;
		uint32_t width = (w[0] & 1023) + 1, addr = w[1] & 0xffffff, ff;
		switch (size) {
		case 0: ff = FB_I4; break;
		case 1: ff = FB_I8; break;
		case 2: ff = fmt ? FB_IA88 : FB_RGBA5551; break;
		default: ff = FB_RGBA8888; break;
		}
		if (addr != o->fb_addr || width != o->fb_width || ff != o->fb_fmt)
			pass_flush(o);
		o->fb_width = width;
		o->fb_addr = addr;
		o->fb_fmt = ff;
		break;
	}
	case 0x04: /* MetaSetQuirks */
		o->quirks = w[1];
		break;
	default:
		break;
	}
}

size_t rdpo_process(rdpo *o, const uint32_t *words, size_t num_words)
{
	size_t i = 0;
	while (i < num_words) {
		unsigned len = rdpo_command_length(words[i]);
		if (len == 0) {
			i++;
			continue;
		}
		if (i + len > num_words)
			break;
		execute(o, words + i);
		i += len;
	}
	return i;
}

void rdpo_flush(rdpo *o) { pass_flush(o); }

size_t rdpo_pass_ends(const rdpo *o, uint64_t *out, size_t cap)
{
	for (size_t i = 0; i < o->pass_log_len && i < cap; i++)
		out[i] = o->pass_log[i];
	return o->pass_log_len;
}

/* submit_update_upscaled_domain_external (rdp_renderer.cpp:1936-1941): the
 * pre update over the range the VI reads, colour only, length counted in
 * pixels of 1 << size_log2 bytes. */
static void up_vi_domain_pre(void *ctx, uint32_t offset, uint32_t length, unsigned size_log2)
{
	up_domain_update((rdpo *)ctx, 0, offset, offset, length, size_log2);
}

int rdpo_scanout_ex(rdpo *o, const rdpo_scanout_options *opts, uint32_t *out, unsigned *width, unsigned *height)
{
	/* CommandProcessor::scanout flushes the renderer first (rdp_device.cpp:1095-1106). */
	pass_flush(o);
	vi_source_t src;
	memset(&src, 0, sizeof(src));
	src.rdram_size = o->rdram_size;
	if (o->up_log2) {
		src.rdram = o->up_rdram;
		src.hidden = o->up_hidden;
		src.scale_log2 = o->up_log2;
		src.domain_pre = up_vi_domain_pre;
		src.ctx = o;
	} else {
		src.rdram = o->rdram;
		src.hidden = o->hidden;
	}
	return vi_scanout(o->vi, o->vi_regs, &src, opts, out, width, height);
}

int rdpo_scanout(rdpo *o, uint32_t *out, unsigned *width, unsigned *height)
{
	rdpo_scanout_options opts;
	rdpo_scanout_options_default(&opts);
	return rdpo_scanout_ex(o, &opts, out, width, height);
}

void rdpo_begin_vi_register_per_scanline(rdpo *o, uint32_t flags) { vi_begin_per_scanline(o->vi, o->vi_regs, flags); }
void rdpo_set_vi_register_for_scanline(rdpo *o, uint32_t reg_bit, uint32_t value)
{
	vi_set_for_scanline(o->vi, reg_bit, value);
}
void rdpo_latch_vi_register_for_scanline(rdpo *o, unsigned vi_line) { vi_latch_for_scanline(o->vi, vi_line); }
void rdpo_end_vi_register_per_scanline(rdpo *o) { vi_end_per_scanline(o->vi); }

void rdpo_set_vi_register(rdpo *o, unsigned reg, uint32_t value)
{
	if (reg < 14)
		o->vi_regs[reg] = value;
}

/* ------------------------------------------------------------------------ */
/* Hooks that let the test-suite pin individual functions.                  */
/* ------------------------------------------------------------------------ */
uint32_t rdpo_test_z_compress(int32_t z) { return z_compress(z); }
int32_t rdpo_test_z_decompress(uint32_t z) { return z_decompress(z); }
int32_t rdpo_test_clamp_9bit(int32_t c) { return clamp_9bit(c); }
int32_t rdpo_test_clamp_z(int32_t z) { return clamp_z(z); }
int32_t rdpo_test_dz_compress_host(int32_t dz) { return dz_compress_host(dz); }
int32_t rdpo_test_normalize_dzpix(int32_t dz) { return normalize_dzpix(dz); }
int32_t rdpo_test_combiner_equation(int32_t a, int32_t b, int32_t c, int32_t d)
{
	int32_t A[4] = { a, a, a, a }, B[4] = { b, b, b, b }, C[4] = { c, c, c, c }, D[4] = { d, d, d, d }, out[4];
	combiner_equation(A, B, C, D, out);
	return out[0];
}
uint32_t rdpo_test_noise(uint32_t x, uint32_t y, uint32_t prim) { return reseed_noise(x, y, prim); }
void rdpo_test_perspective_divide(int32_t s, int32_t t, int32_t w, int32_t out[3])
{
	int32_t stw[3] = { s, t, w }, st[2];
	int ovf = 0;
	perspective_divide(stw, st, &ovf);
	out[0] = st[0];
	out[1] = st[1];
	out[2] = ovf;
}
int rdpo_test_coverage(const uint16_t xleft[4], const uint16_t xright[4], int x)
{
	span_t s;
	memset(&s, 0, sizeof(s));
	for (int i = 0; i < 4; i++) {
		s.xleft[i] = xleft[i];
		s.xright[i] = xright[i];
	}
	return compute_coverage(&s, x);
}
void rdpo_test_rgb_dither(int32_t rgb[3], int dith) { rgb_dither(rgb, dith); }

/* Hooks for the function pins of tests/test_oracle_pins.py.  Records are
 * passed as flat integer arrays:
 *   triangle: xh, xm, xl, yh, ym, yl, dxhdy, dxmdy, dxldy, flags;
 *   attributes: rgba, drgba_dx, drgba_de, drgba_dy, stzw, dstzw_dx,
 *               dstzw_de, dstzw_dy (4 each);
 *   tile: slo, shi, tlo, thi, offset, stride, fmt, size, palette, mask_s,
 *         shift_s, mask_t, shift_t, flags. */
static void unpack_tri(const int32_t v[10], tri_t *t)
{
	memset(t, 0, sizeof(*t));
	t->xh = v[0];
	t->xm = v[1];
	t->xl = v[2];
	t->yh = v[3];
	t->ym = v[4];
	t->yl = v[5];
	t->dxhdy = v[6];
	t->dxmdy = v[7];
	t->dxldy = v[8];
	t->flags = (uint32_t)v[9];
}

static void unpack_attr(const int32_t v[32], attr_t *a)
{
	for (int i = 0; i < 4; i++) {
		a->rgba[i] = v[i];
		a->drgba_dx[i] = v[4 + i];
		a->drgba_de[i] = v[8 + i];
		a->drgba_dy[i] = v[12 + i];
		a->stzw[i] = v[16 + i];
		a->dstzw_dx[i] = v[20 + i];
		a->dstzw_de[i] = v[24 + i];
		a->dstzw_dy[i] = v[28 + i];
	}
}

static void unpack_tile(const uint32_t v[14], tile_t *t)
{
	t->slo = v[0];
	t->shi = v[1];
	t->tlo = v[2];
	t->thi = v[3];
	t->offset = v[4];
	t->stride = v[5];
	t->fmt = v[6];
	t->size = v[7];
	t->palette = v[8];
	t->mask_s = v[9];
	t->shift_s = v[10];
	t->mask_t = v[11];
	t->shift_t = v[12];
	t->flags = v[13];
}

/* out: rgba[4], stzw[4], xleft[4], xright[4], base_x, start_x, end_x,
 * lodlength, valid. */
void rdpo_test_span_setup(const int32_t tri[10], const int32_t attr[32], const uint32_t scissor[4], int y, int32_t out[21])
{
	tri_t t;
	attr_t a;
	span_t s;
	unpack_tri(tri, &t);
	unpack_attr(attr, &a);
	memset(&s, 0, sizeof(s));
	span_setup(&t, &a, scissor, y, &s, 0);
	for (int i = 0; i < 4; i++) {
		out[i] = s.rgba[i];
		out[4 + i] = s.stzw[i];
		out[8 + i] = s.xleft[i];
		out[12 + i] = s.xright[i];
	}
	out[16] = s.base_x;
	out[17] = s.start_x;
	out[18] = s.end_x;
	out[19] = s.lodlength;
	out[20] = s.valid;
}

void rdpo_test_interpolate_rgba(const int32_t span_rgba[4], const int32_t attr[32], int dx, int coverage, int32_t out[4])
{
	span_t s;
	attr_t a;
	memset(&s, 0, sizeof(s));
	unpack_attr(attr, &a);
	memcpy(s.rgba, span_rgba, sizeof(s.rgba));
	interpolate_rgba(&s, &a, dx, coverage, out, 0);
}

/* out: st[2], st_dx[2], st_dy[2], z, overflow. */
void rdpo_test_interpolate_stz(const int32_t span_stzw[4], const int32_t attr[32], int dx, int coverage, int perspective,
                               int uses_lod, int flip_dir, int32_t out[8])
{
	span_t s;
	attr_t a;
	memset(&s, 0, sizeof(s));
	unpack_attr(attr, &a);
	memcpy(s.stzw, span_stzw, sizeof(s.stzw));
	int32_t st[2], st_dx[2] = { 0, 0 }, st_dy[2] = { 0, 0 }, z;
	int ovf = 0;
	interpolate_stz(&s, &a, dx, coverage, perspective, uses_lod, flip_dir, st, st_dx, st_dy, &z, &ovf, 0);
	out[0] = st[0];
	out[1] = st[1];
	out[2] = st_dx[0];
	out[3] = st_dx[1];
	out[4] = st_dy[0];
	out[5] = st_dy[1];
	out[6] = z;
	out[7] = ovf;
}

void rdpo_test_fetch_texel(const rdpo *o, const uint32_t tile[14], uint32_t s, uint32_t t, int tlut, int tlut_type,
                           uint32_t lut_offset, uint32_t addr_xor, int32_t out[4])
{
	tile_t tl;
	unpack_tile(tile, &tl);
	fetch_texel(o, &tl, s, t, tlut, tlut_type, lut_offset, addr_xor, out);
}

void rdpo_test_fetch_yuv(const rdpo *o, const uint32_t tile[14], uint32_t s, uint32_t t, uint32_t chroma_x, int32_t out[4])
{
	tile_t tl;
	unpack_tile(tile, &tl);
	fetch_yuv16(o, &tl, s, t, chroma_x, out);
}

int rdpo_test_texel_mask(const uint32_t tile[14], int coord, int axis_t)
{
	tile_t tl;
	unpack_tile(tile, &tl);
	return axis_t ? texel_mask_t(&tl, coord) : texel_mask_s(&tl, coord);
}

int rdpo_test_clamp_and_shift(int clamp_bit, int coord, int lo, int hi, int shift)
{
	return clamp_and_shift_coord(clamp_bit, coord, lo, hi, shift);
}

void rdpo_test_sample_texture(const rdpo *o, const uint32_t tile[14], const int32_t st[2], int tlut, int tlut_type,
                              int sample_quad, int mid_texel, int convert_one, int bilerp, const int32_t factors[4],
                              const int32_t prev[4], int32_t out[4])
{
	tile_t tl;
	unpack_tile(tile, &tl);
	sample_texture(o, &tl, st, tlut, tlut_type, sample_quad, mid_texel, convert_one, bilerp, factors, prev, out);
}

/* stq: st, st_dx, st_dy; out: tile0, tile1, lod_frac. */
void rdpo_test_compute_lod(uint32_t tile, uint32_t max_level, int min_lod, const int32_t stq[6], int ovf, int tex_lod_en,
                           int sharpen, int detail, int32_t out[3])
{
	uint32_t t0 = tile & 7, t1 = (tile + 1) & 7;
	int32_t frac = 0;
	compute_lod_2cycle(&t0, &t1, &frac, max_level, min_lod, stq, stq + 2, stq + 4, ovf, tex_lod_en, sharpen, detail);
	out[0] = (int32_t)t0;
	out[1] = (int32_t)t1;
	out[2] = frac;
}

/* in: shade[4], combined[4], texel0[4], texel1[4], lod_frac, noise;
 * k: the muladd, mulsub, mul and add constant slots (4 each). */
void rdpo_test_select(const int32_t in[18], const int32_t k[16], int which, int sel_rgb, int sel_a, int32_t out[4])
{
	comb_in_t ci;
	memcpy(ci.shade, in, 16);
	memcpy(ci.combined, in + 4, 16);
	memcpy(ci.texel0, in + 8, 16);
	memcpy(ci.texel1, in + 12, 16);
	ci.lod_frac = in[16];
	ci.noise = in[17];
	ci.const_muladd = k;
	ci.const_mulsub = k + 4;
	ci.const_mul = k + 8;
	ci.const_add = k + 12;
	select4(&ci, which, sel_rgb, sel_a, out);
}

/* Returns pass; out: coverage after the test, blend_en, cvg_wrap, shift0, shift1. */
int rdpo_test_depth_test(int32_t z, int32_t dz, int32_t dz_compressed, uint32_t cur_depth, uint32_t cur_dz, int coverage,
                         int cur_cvg, int z_compare, int z_mode, int force_blend, int aa, int32_t out[5])
{
	int cvg = coverage, blend_en = 0, wrap = 0, shift[2] = { 0, 0 };
	int pass = depth_test(z, dz, dz_compressed, cur_depth, cur_dz, &cvg, cur_cvg, z_compare, z_mode, force_blend, aa,
	                      &blend_en, &wrap, shift);
	out[0] = cvg;
	out[1] = blend_en;
	out[2] = wrap;
	out[3] = shift[0];
	out[4] = shift[1];
	return pass;
}

/* in: pixel[4], memory[4], fog[4], blendc[4], shade_alpha; modes: 1a, 1b, 2a, 2b. */
void rdpo_test_blender(const rdpo *o, const int32_t in[17], const int32_t modes[4], int force_blend, int blend_en,
                       int color_on_cvg, int cvg_wrap, int shift0, int shift1, int final_cycle, int32_t out[3])
{
	blend_in_t b;
	memcpy(b.pixel, in, 16);
	memcpy(b.memory, in + 4, 16);
	memcpy(b.fog, in + 8, 16);
	memcpy(b.blendc, in + 12, 16);
	b.shade_alpha = in[16];
	const uint8_t m[4] = { (uint8_t)modes[0], (uint8_t)modes[1], (uint8_t)modes[2], (uint8_t)modes[3] };
	const int shift[2] = { shift0, shift1 };
	blender(o, &b, m, force_blend, blend_en, color_on_cvg, cvg_wrap, shift, final_cycle, out);
}

int rdpo_test_blend_coverage(int coverage, int memory_coverage, int blend_en, int mode)
{
	return blend_coverage(coverage, memory_coverage, blend_en, mode);
}

void rdpo_test_load_color(const rdpo *o, uint32_t fmt, uint32_t index, int32_t out[4]) { load_color(o, fmt, index, out); }
void rdpo_test_store_color(rdpo *o, uint32_t fmt, uint32_t index, const int32_t c[4]) { store_color(o, fmt, index, c); }

void rdpo_test_load_depth(const rdpo *o, uint32_t index, uint32_t out[2]) { load_depth(o, index, &out[0], &out[1]); }
void rdpo_test_store_depth(rdpo *o, uint32_t index, uint32_t depth, uint32_t dz) { store_depth(o, index, depth, dz); }

/* A pixel through the fill or copy pipe from the given colour and depth
 * state; io: color[4], depth, dz (in and out). */
void rdpo_test_fill_color(uint32_t fmt, uint32_t color_index, int alias, uint32_t col, int32_t io[6])
{
	pixel_state_t p;
	memset(&p, 0, sizeof(p));
	memcpy(p.color, io, 16);
	p.depth = (uint32_t)io[4];
	p.dz = (uint32_t)io[5];
	p.color_index = color_index;
	fill_color(&p, fmt, alias, col);
	memcpy(io, p.color, 16);
	io[4] = (int32_t)p.depth;
	io[5] = (int32_t)p.dz;
}

void rdpo_test_copy_pipeline(uint32_t fmt, int alias, uint32_t word, int32_t io[6])
{
	pixel_state_t p;
	memset(&p, 0, sizeof(p));
	memcpy(p.color, io, 16);
	p.depth = (uint32_t)io[4];
	p.dz = (uint32_t)io[5];
	copy_pipeline(&p, fmt, alias, word);
	memcpy(io, p.color, 16);
	io[4] = (int32_t)p.depth;
	io[5] = (int32_t)p.dz;
}

/* State digest with the definition of the device digest (rdp_mi_state_digest
 * of the library under test): the sum modulo 2^64 of SplitMix64(key ^ (i << 32
 * | w_i)) over RDRAM words, hidden RDRAM bytes packed four to a little-endian
 * word and TMEM halfwords packed two to a little-endian word. */
static uint64_t digest_mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

static uint64_t digest_region(const uint8_t *p, size_t words, uint64_t key)
{
	uint64_t acc = 0;
	for (size_t i = 0; i < words; i++) {
		uint32_t w;
		memcpy(&w, p + 4 * i, 4);
		acc += digest_mix64(key ^ (((uint64_t)i << 32) | w));
	}
	return acc;
}

uint64_t rdpo_state_digest(const rdpo *o)
{
	return digest_region(o->rdram, o->rdram_size / 4, 0x243f6a8885a308d3ull) +
	       digest_region(o->hidden, o->rdram_size / 8, 0x13198a2e03707344ull) +
	       digest_region((const uint8_t *)o->tmem, 1024, 0xa4093822299f31d0ull);
}
