/*
 * vi_oracle.c - CPU restatement of the paraLLEl-RDP video interface scanout.
 *
 * TEST INFRASTRUCTURE ONLY (see rdp_oracle.h): the correctness oracle for the
 * device VI path.  Follows, stage by stage:
 *   register decode     video_interface.cpp:176-367 (decode_vi_registers)
 *   per-scanline regs   video_interface.cpp:1126-1213
 *   frame control       video_interface.cpp:1215-1426 (scanout)
 *   VRAM extraction     shaders/extract_vram.comp:61-106
 *   AA / dither filter  shaders/vi_fetch.frag:58-163
 *   divot filter        shaders/vi_divot.frag:36-93
 *   scale, gamma        shaders/vi_scale.frag:32-153, video_interface.cpp:796-998
 * Every image the reference renders is materialised here as a plain array;
 * no upscaling (scaling factor 1).
 */
#include "vi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Planted-mistake builds, as in rdp_oracle.c (-DRDPO_PLANT=n). */
#ifndef RDPO_PLANT
#define RDPO_PLANT 0
#endif
#define PLANTED(n) (RDPO_PLANT == (n))

enum {
	VI_TYPE_MASK = 3,
	VI_TYPE_RGBA5551 = 2,
	VI_TYPE_RGBA8888 = 3,
	VI_GAMMA_DITHER = 1 << 2,
	VI_GAMMA = 1 << 3,
	VI_DIVOT = 1 << 4,
	VI_SERRATE = 1 << 6,
	VI_AA_MASK = 3 << 8,
	VI_AA_RESAMP_ONLY = 2 << 8,
	VI_AA_RESAMP_REPLICATE = 3 << 8,
	VI_DITHER_FILTER = 1 << 16,
	VI_META_AA = 1 << 17,
	VI_META_SCALE = 1 << 18,
};

/* rdp_common.hpp:375-391 */
#define V_SYNC_NTSC 525
#define H_OFFSET_NTSC 108
#define H_OFFSET_PAL 128
#define V_OFFSET_NTSC 34
#define V_OFFSET_PAL 44
#define V_RES_NTSC 480
#define V_RES_PAL 576
#define SCANOUT_WIDTH 640
#define V_END_PAL ((V_OFFSET_PAL + V_RES_PAL) | 1)
#define V_END_NTSC ((V_OFFSET_NTSC + V_RES_NTSC) | 1)
#define MAX_OUTPUT_SCANLINES (V_RES_PAL >> 1)

enum { REG_CONTROL, REG_ORIGIN, REG_WIDTH, REG_INTR, REG_V_CURRENT_LINE, REG_TIMING, REG_V_SYNC, REG_H_SYNC,
       REG_LEAP, REG_H_START, REG_V_START, REG_V_BURST, REG_X_SCALE, REG_Y_SCALE };

typedef struct {
	int32_t h_start, h_start_clamp, h_end_clamp, x_start, x_add, y_start, y_add, y_base;
} vi_line_t;

typedef struct {
	uint32_t status;
	int vi_width, vi_offset, v_current_line, v_start, init_y_add, v_res, is_pal;
	int max_x, max_y, h_start, h_res, h_start_clamp, h_res_clamp;
} vi_regs_t;

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

void vi_state_init(vi_state_t *s) { memset(s, 0, sizeof(*s)); }

void vi_state_free(vi_state_t *s)
{
	free(s->prev);
	s->prev = NULL;
	s->prev_cap = 0;
}

/* ------------------------------------------------------------------------ */
/* Per-scanline registers, video_interface.cpp:1126-1213.                   */
/* ------------------------------------------------------------------------ */
void vi_begin_per_scanline(vi_state_t *s, const uint32_t regs[14], uint32_t flags)
{
	s->per_line_flags = flags;
	s->h_start_latched = regs[REG_H_START];
	s->x_scale_latched = regs[REG_X_SCALE];
	s->h_start_lines[0] = regs[REG_H_START];
	s->x_scale_lines[0] = regs[REG_X_SCALE];
	s->per_line_line = 0;
	s->per_line_ended = 0;
}

void vi_set_for_scanline(vi_state_t *s, uint32_t reg_bit, uint32_t value)
{
	if ((s->per_line_flags & reg_bit) == 0)
		return;
	if (reg_bit == 1)
		s->h_start_latched = value;
	else if (reg_bit == 2)
		s->x_scale_latched = value;
}

void vi_latch_for_scanline(vi_state_t *s, unsigned vi_line)
{
	if (vi_line > VI_ORACLE_V_END_MAX - 1)
		vi_line = VI_ORACLE_V_END_MAX - 1;
	if (vi_line <= s->per_line_line)
		return;
	unsigned n = s->per_line_line;
	while (++n < vi_line) {
		s->h_start_lines[n] = s->h_start_lines[s->per_line_line];
		s->x_scale_lines[n] = s->x_scale_lines[s->per_line_line];
	}
	s->h_start_lines[n] = s->h_start_latched;
	s->x_scale_lines[n] = s->x_scale_latched;
	s->per_line_line = n;
}

void vi_end_per_scanline(vi_state_t *s)
{
	if (s->per_line_flags == 0 || s->per_line_ended)
		return;
	unsigned n = s->per_line_line;
	while (++n < VI_ORACLE_V_END_MAX) {
		s->h_start_lines[n] = s->h_start_lines[s->per_line_line];
		s->x_scale_lines[n] = s->x_scale_lines[s->per_line_line];
	}
	s->per_line_ended = 1;
}

/* ------------------------------------------------------------------------ */
/* decode_vi_registers, video_interface.cpp:176-367.                        */
/* ------------------------------------------------------------------------ */
typedef struct {
	vi_regs_t *reg;
	int h_start_clamp_lo, h_end_clamp_hi, h_start_lo, h_end_hi;
	int y_start, y_add;
} analyze_ctx_t;

static void analyze_line(analyze_ctx_t *c, int x_start, int x_add, int h_start, int h_end, vi_line_t *line)
{
	int left_clamp = 0, right_clamp = 0;
	h_start -= c->reg->is_pal ? H_OFFSET_PAL : H_OFFSET_NTSC;
	h_end -= c->reg->is_pal ? H_OFFSET_PAL : H_OFFSET_NTSC;
	if (h_start < 0) {
		x_start -= x_add * h_start;
		h_start = 0;
		left_clamp = 1;
	}
	if (h_end > SCANOUT_WIDTH) {
		h_end = SCANOUT_WIDTH;
		right_clamp = 1;
	}
	int h_start_clamp = h_start + (left_clamp ? 0 : 8);
	int h_end_clamp = h_end - (right_clamp ? 0 : 7);
	int h_res = h_end - h_start;
	int max_x = (x_start + h_res * x_add) >> 10;
	c->reg->max_x = imax(c->reg->max_x, max_x);
	c->h_start_lo = imin(c->h_start_lo, h_start);
	c->h_end_hi = imax(c->h_end_hi, h_end);
	c->h_start_clamp_lo = imin(c->h_start_clamp_lo, h_start_clamp);
	c->h_end_clamp_hi = imax(c->h_end_clamp_hi, h_end_clamp);
	if (line) {
		line->h_start = h_start;
		line->h_start_clamp = h_start_clamp;
		line->h_end_clamp = h_end_clamp;
		line->x_start = x_start;
		line->x_add = x_add;
		line->y_start = c->y_start;
		line->y_add = c->y_add;
		line->y_base = 0;
	}
}

static void decode_vi_registers(const vi_state_t *s, const uint32_t regs[14], vi_regs_t *reg, vi_line_t *lines)
{
	memset(reg, 0, sizeof(*reg));
	reg->status = regs[REG_CONTROL];
	reg->vi_width = (int)(regs[REG_WIDTH] & 0xfff);
	reg->vi_offset = (int)(regs[REG_ORIGIN] & 0xffffff);
	reg->v_current_line = (int)(regs[REG_V_CURRENT_LINE] & 1);
	reg->v_start = (int)((regs[REG_V_START] >> 16) & 0x3ff);

	int y_start = (int)((regs[REG_Y_SCALE] >> 16) & 0xfff);
	int y_add = (int)(regs[REG_Y_SCALE] & 0xfff);
	reg->init_y_add = y_add;

	int v_end = (int)(regs[REG_V_START] & 0x3ff);
	int v_sync = (int)(regs[REG_V_SYNC] & 0x3ff);
	reg->is_pal = (unsigned)v_sync > (V_SYNC_NTSC + 25);

	int v_end_max = reg->is_pal ? V_END_PAL : V_END_NTSC;
	if (v_end > v_end_max)
		v_end = v_end_max;
	if (reg->v_start > v_end_max)
		reg->v_start = v_end_max;

	int v_start_offset = reg->is_pal ? V_OFFSET_PAL : V_OFFSET_NTSC;
	reg->v_res = (v_end - reg->v_start) >> 1;
	reg->v_start = (reg->v_start - v_start_offset) / 2; /* truncating division */
	if (reg->v_start < 0) {
		y_start -= y_add * reg->v_start;
		reg->v_start = 0;
	}
	reg->v_res = imin(reg->v_res, MAX_OUTPUT_SCANLINES - reg->v_start);

	analyze_ctx_t c;
	c.reg = reg;
	c.h_start_clamp_lo = 0x7fffffff;
	c.h_end_clamp_hi = 0;
	c.h_start_lo = 0x7fffffff;
	c.h_end_hi = 0;
	c.y_start = y_start;
	c.y_add = y_add;
	reg->max_x = 0;
	int degenerate_y = reg->v_res <= 0;

	if (lines) {
		if (degenerate_y)
			memset(lines, 0, sizeof(vi_line_t) * MAX_OUTPUT_SCANLINES);
		else {
			for (int l = 0; l < reg->v_start; l++)
				memset(&lines[l], 0, sizeof(vi_line_t));
			for (int l = reg->v_start + reg->v_res; l < MAX_OUTPUT_SCANLINES; l++)
				memset(&lines[l], 0, sizeof(vi_line_t));
		}
	}

	if (degenerate_y || !s->per_line_ended) {
		int x_start = (int)((regs[REG_X_SCALE] >> 16) & 0xfff);
		int x_add = (int)(regs[REG_X_SCALE] & 0xfff);
		int h_start = (int)((regs[REG_H_START] >> 16) & 0x3ff);
		int h_end = (int)(regs[REG_H_START] & 0x3ff);
		vi_line_t info;
		analyze_line(&c, x_start, x_add, h_start, h_end, &info);
		if (lines)
			for (int l = reg->v_start; l < reg->v_start + reg->v_res; l++)
				lines[l] = info;
	} else {
		for (int l = reg->v_start; l < reg->v_start + reg->v_res; l++) {
			int eff = 2 * l + v_start_offset + (reg->v_current_line == 0);
			uint32_t xs = s->x_scale_lines[eff], hs = s->h_start_lines[eff];
			analyze_line(&c, (int)((xs >> 16) & 0xfff), (int)(xs & 0xfff), (int)((hs >> 16) & 0x3ff),
			             (int)(hs & 0x3ff), lines ? &lines[l] : NULL);
		}
	}

	reg->max_y = (y_start + reg->v_res * y_add) >> 10;
	reg->h_start = c.h_start_lo;
	reg->h_res = c.h_end_hi - c.h_start_lo;
	reg->h_start_clamp = c.h_start_clamp_lo;
	reg->h_res_clamp = c.h_end_clamp_hi - c.h_start_clamp_lo;
}

/* ------------------------------------------------------------------------ */
/* Images: RGBA as 4 x uint32 per texel (rgb + coverage in a).              */
/* ------------------------------------------------------------------------ */
typedef struct {
	uint32_t c[4];
} texel_t;

typedef struct {
	int w, h;
	texel_t *t;
} image_t;

static texel_t img_get(const image_t *im, int x, int y)
{
	texel_t z = { { 0, 0, 0, 0 } };
	if (x < 0 || y < 0 || x >= im->w || y >= im->h)
		return z; /* out-of-range texel fetch reads zero */
	return im->t[(size_t)y * (size_t)im->w + (size_t)x];
}

/* extract_vram.comp:61-106; (x, y) is a pixel of the 2^L-times image,
 * sample (x mod 2^L, y mod 2^L) of native pixel (x >> L, y >> L). */
static texel_t fetch_color(const vi_source_t *src, uint32_t status, int fb_offset, int fb_width, int x, int y)
{
	texel_t t;
	const int L = (int)src->scale_log2, S = 1 << L;
	const uint32_t rdram_size = src->rdram_size;
	int slice = PLANTED(36) ? (x & (S - 1)) * S + (y & (S - 1)) : (y & (S - 1)) * S + (x & (S - 1));
	x >>= L;
	y >>= L;
	const uint8_t *rdram = src->rdram + (size_t)slice * rdram_size;
	const uint8_t *hidden = src->hidden + (size_t)slice * (rdram_size >> 1);
	int type = (int)(status & VI_TYPE_MASK);
	int linear = y * fb_width + x + fb_offset;
	if (type == VI_TYPE_RGBA8888) {
		linear &= (int)(rdram_size >> 2) - 1;
		uint32_t word;
		memcpy(&word, rdram + 4 * (size_t)linear, 4);
		t.c[0] = word >> 24;
		t.c[1] = (word >> 16) & 0xff;
		t.c[2] = (word >> 8) & 0xff;
		t.c[3] = (word >> 5) & 7;
	} else if (type == VI_TYPE_RGBA5551) {
		linear &= (int)(rdram_size >> 1) - 1;
		uint16_t word;
		memcpy(&word, rdram + 2 * (size_t)(linear ^ 1), 2);
		uint32_t hid = hidden[linear];
		t.c[0] = (word >> 8) & 0xf8;
		t.c[1] = (word >> 3) & 0xf8;
		t.c[2] = ((uint32_t)word << 2) & 0xf8;
		t.c[3] = ((word & 1u) << 2) | hid;
	} else
		memset(&t, 0, sizeof(t));
	if (!(status & VI_META_AA))
		t.c[3] = 7;
	return t;
}

static void check_neighbor(texel_t cand, uint32_t lo[3], uint32_t hi[3], uint32_t slo[3], uint32_t shi[3])
{
	if (cand.c[3] != 7)
		return;
	for (int i = 0; i < 3; i++) {
		uint32_t mx = cand.c[i] > lo[i] ? cand.c[i] : lo[i];
		uint32_t mn = cand.c[i] < hi[i] ? cand.c[i] : hi[i];
		if (mx < slo[i]) slo[i] = mx;
		if (mn > shi[i]) shi[i] = mn;
		if (cand.c[i] < lo[i]) lo[i] = cand.c[i];
		if (cand.c[i] > hi[i]) hi[i] = cand.c[i];
	}
}

static int clamp1(int v) { return v < -1 ? -1 : (v > 1 ? 1 : v); }

/* vi_fetch.frag:58-163 for one output texel (layer 0 and, with the fetch bug,
 * layer 1).  E is the extracted VRAM image; (px, py) is already offset by
 * (+2, +2). */
static void vi_fetch_texel(const image_t *E, uint32_t status, int fetch_bug, int px, int py, texel_t *out0,
                           texel_t *out1)
{
	texel_t mid = img_get(E, px, py);
	uint32_t color[3], color_bug[3];
	if (mid.c[3] != 7) {
		uint32_t lo[3], hi[3], slo[3], shi[3];
		for (int i = 0; i < 3; i++)
			lo[i] = hi[i] = slo[i] = shi[i] = mid.c[i];
		texel_t lu = img_get(E, px - 1, py - 1), ru = img_get(E, px + 1, py - 1);
		texel_t tl = img_get(E, px - 2, py), tr = img_get(E, px + 2, py);
		texel_t ld = img_get(E, px - 1, py + 1), rd = img_get(E, px + 1, py + 1);
		check_neighbor(lu, lo, hi, slo, shi);
		check_neighbor(ru, lo, hi, slo, shi);
		check_neighbor(tl, lo, hi, slo, shi);
		check_neighbor(tr, lo, hi, slo, shi);
		uint32_t lo_b[3], hi_b[3], slo_b[3], shi_b[3];
		memcpy(lo_b, lo, sizeof(lo));
		memcpy(hi_b, hi, sizeof(hi));
		memcpy(slo_b, slo, sizeof(slo));
		memcpy(shi_b, shi, sizeof(shi));
		check_neighbor(ld, lo, hi, slo, shi);
		check_neighbor(rd, lo, hi, slo, shi);
		if (fetch_bug) {
			check_neighbor(tl, lo_b, hi_b, slo_b, shi_b);
			check_neighbor(tr, lo_b, hi_b, slo_b, shi_b);
			for (int i = 0; i < 3; i++) {
				if (mid.c[i] == lo[i]) slo[i] = lo[i];
				if (mid.c[i] == hi[i]) shi[i] = hi[i];
				if (mid.c[i] == lo_b[i]) slo_b[i] = lo_b[i];
				if (mid.c[i] == hi_b[i]) shi_b[i] = hi_b[i];
			}
		}
		uint32_t coeff = 7u - mid.c[3];
		for (int i = 0; i < 3; i++) {
			uint32_t off = slo[i] + shi[i] - (mid.c[i] << 1);
			color[i] = (mid.c[i] + ((off * coeff + 4u) >> 3)) & 0xffu;
			uint32_t off_b = slo_b[i] + shi_b[i] - (mid.c[i] << 1);
			color_bug[i] = (mid.c[i] + ((off_b * coeff + 4u) >> 3)) & 0xffu;
		}
	} else if (status & VI_DITHER_FILTER) {
		int tmp[3], acc[3] = { 0, 0, 0 }, acc_b[3];
		for (int i = 0; i < 3; i++)
			tmp[i] = (int)(mid.c[i] >> 3);
		for (int y = -1; y <= 0; y++)
			for (int x = -1; x <= 1; x++) {
				texel_t t = img_get(E, px + x, py + y);
				for (int i = 0; i < 3; i++)
					acc[i] += clamp1((int)(t.c[i] >> 3) - tmp[i]);
			}
		memcpy(acc_b, acc, sizeof(acc));
		texel_t a = img_get(E, px - 1, py + 1), b = img_get(E, px + 1, py + 1), c = img_get(E, px, py + 1);
		texel_t l = img_get(E, px - 1, py), r = img_get(E, px + 1, py);
		for (int i = 0; i < 3; i++) {
			acc[i] += clamp1((int)(a.c[i] >> 3) - tmp[i]);
			acc[i] += clamp1((int)(b.c[i] >> 3) - tmp[i]);
			acc[i] += clamp1((int)(c.c[i] >> 3) - tmp[i]);
			acc_b[i] += clamp1((int)(l.c[i] >> 3) - tmp[i]);
			acc_b[i] += clamp1((int)(r.c[i] >> 3) - tmp[i]);
			/* An out-of-range uint written to an 8-bit UINT attachment keeps
			 * its low 8 bits. */
			color[i] = ((mid.c[i] & 0xf8u) + (uint32_t)acc[i]) & 0xffu;
			color_bug[i] = ((mid.c[i] & 0xf8u) + (uint32_t)acc_b[i]) & 0xffu;
		}
	} else {
		for (int i = 0; i < 3; i++)
			color[i] = color_bug[i] = mid.c[i];
	}
	for (int i = 0; i < 3; i++) {
		out0->c[i] = color[i];
		out1->c[i] = color_bug[i];
	}
	out0->c[3] = out1->c[3] = mid.c[3];
}

static uint32_t median3(uint32_t l, uint32_t c, uint32_t r)
{
	uint32_t t;
	if (l < c) { t = l; l = c; c = t; }
	if (c < r) { t = c; c = r; r = t; }
	if (l < c) { t = l; l = c; c = t; }
	return c;
}

/* vi_divot.frag:54-92 */
static texel_t divot_texel(const image_t *A, int x, int y)
{
	texel_t l = img_get(A, x, y), m = img_get(A, x + 1, y), r = img_get(A, x + 2, y);
	if ((l.c[3] & m.c[3] & r.c[3]) == 7)
		return m;
	texel_t o;
	for (int i = 0; i < 3; i++)
		o.c[i] = median3(l.c[i], m.c[i], r.c[i]);
	o.c[3] = m.c[3];
	return o;
}

/* noise.h:29-69 */
static uint32_t reseed_noise(uint32_t x, uint32_t y, uint32_t z)
{
	const uint32_t P = 1103515245u;
	uint32_t s0 = x, s1 = y, s2 = z;
	for (int i = 0; i < 3; i++) {
		uint32_t n0 = ((s0 >> 8) ^ s1) * P;
		uint32_t n1 = ((s1 >> 8) ^ s2) * P;
		uint32_t n2 = ((s2 >> 8) ^ s0) * P;
		s0 = n0;
		s1 = n1;
		s2 = n2;
	}
	return (s0 >> 16) & 0xffff;
}

static uint32_t vi_lerp(uint32_t a, uint32_t b, uint32_t l) { return (a + (((b - a) * l + 16u) >> 5)) & 0xffu; }

static unsigned isqrt_u(unsigned v)
{
	unsigned r = 0;
	while ((r + 1) * (r + 1) <= v)
		r++;
	return r;
}

static uint8_t gamma_lookup(uint32_t idx)
{
	/* luts.hpp:1034-1555 in closed form: [0, 256) is isqrt(256 c) & ~1,
	 * [256, 256 + 16384) is isqrt(4 j) & ~1. */
	if (idx < 256)
		return (uint8_t)(isqrt_u(idx * 256) & ~1u);
	return (uint8_t)(isqrt_u((idx - 256) * 4) & ~1u);
}

void rdpo_scanout_options_default(rdpo_scanout_options *o)
{
	memset(o, 0, sizeof(*o));
	o->upscale_deinterlacing = 1;
	o->vi_aa = o->vi_scale = o->vi_serrate = o->vi_dither_filter = o->vi_divot_filter = o->vi_gamma_dither = 1;
}

static void shift_rect(int *ox, int *oy, int *w, int *h, int dx, int dy)
{
	*ox += dx;
	*oy += dy;
	if (*ox < 0) {
		*w += *ox;
		*ox = 0;
	}
	if (*oy < 0) {
		*h += *oy;
		*oy = 0;
	}
	if (*w < 0)
		*w = 0;
	if (*h < 0)
		*h = 0;
}

/* VideoInterface::scanout_memory_range, video_interface.cpp:369-396, on
 * the registers as they decode after the scanout has cleared the
 * per-scanline state. */
static void scanout_memory_range(const vi_state_t *s, const uint32_t regs[14], uint32_t *offset, uint32_t *length)
{
	vi_regs_t r;
	decode_vi_registers(s, regs, &r, NULL);
	int divot = (r.status & VI_DIVOT) != 0;
	int aa_width = r.max_x + 2 + 4 + divot * 2;
	int aa_height = r.max_y + 1 + 4;
	int x_off = divot ? -3 : -2, y_off = -2;
	if (r.vi_offset == 0 || r.h_res <= 0 || r.h_start >= SCANOUT_WIDTH) {
		*offset = *length = 0;
		return;
	}
	int pixel_size = (((r.status & VI_TYPE_MASK) | VI_TYPE_RGBA5551) == VI_TYPE_RGBA8888) ? 4 : 2;
	uint32_t vi_offset = (uint32_t)r.vi_offset & ~(uint32_t)(pixel_size - 1);
	vi_offset += (uint32_t)((x_off + y_off * r.vi_width) * pixel_size);
	*offset = vi_offset;
	*length = (uint32_t)((aa_height * r.vi_width + aa_width) * pixel_size);
}

/* VideoInterface::scanout, video_interface.cpp:1215-1426 (no downscale
 * steps), with vram_fetch_stage (:405-510) and scale_stage (:719-998);
 * src->scale_log2 is the renderer's upscaling factor. */
int vi_scanout(vi_state_t *s, const uint32_t regs[14], const vi_source_t *src, const rdpo_scanout_options *opts,
               uint32_t *out, unsigned *width, unsigned *height)
{
	const int L = (int)src->scale_log2, SF = 1 << L;
	vi_line_t lines[MAX_OUTPUT_SCANLINES];
	vi_regs_t r;
	decode_vi_registers(s, regs, &r, lines);
	s->per_line_flags = 0; /* clear_per_scanline_state */
	s->per_line_ended = 0;
	*width = *height = 0;

	if (r.vi_offset == 0) {
		s->prev_valid = 0;
		return 0;
	}
	if (!opts->vi_serrate)
		r.status &= ~(uint32_t)VI_SERRATE;
	int is_aa = (r.status & VI_AA_MASK) < VI_AA_RESAMP_ONLY && opts->vi_aa;
	int is_bilinear = (r.status & VI_AA_MASK) < VI_AA_RESAMP_REPLICATE && opts->vi_scale;
	r.status &= ~(uint32_t)(VI_AA_MASK | VI_META_AA | VI_META_SCALE);
	if (is_aa)
		r.status |= VI_META_AA;
	if (is_bilinear)
		r.status |= VI_META_SCALE;
	if (!opts->vi_gamma_dither)
		r.status &= ~(uint32_t)VI_GAMMA_DITHER;
	if (!opts->vi_divot_filter)
		r.status &= ~(uint32_t)VI_DIVOT;
	if (!opts->vi_dither_filter)
		r.status &= ~(uint32_t)VI_DITHER_FILTER;

	int is_blank = (r.status & VI_TYPE_RGBA5551) == 0;
	if (is_blank && s->previous_frame_blank) {
		s->frame_count++;
		s->prev_valid = 0;
		return 0;
	}
	if (is_blank)
		s->prev_valid = 0;
	r.status |= VI_TYPE_RGBA5551;
	s->previous_frame_blank = is_blank;
	int divot = (r.status & VI_DIVOT) != 0;

	if (r.h_res <= 0 || r.h_start >= SCANOUT_WIDTH) {
		s->frame_count++;
		if (opts->persist_frame_on_invalid_input && (s->frame_count - s->last_valid_frame_count < 4)) {
			if (s->prev_valid) {
				*width = s->prev_width;
				*height = s->prev_height;
				memcpy(out, s->prev, sizeof(uint32_t) * s->prev_width * s->prev_height);
			}
		} else
			s->prev_valid = 0;
		return 0;
	}
	s->last_valid_frame_count = s->frame_count;
	int degenerate = r.h_res <= 0 || r.v_res <= 0;
	int fetch_bug = r.init_y_add < 1024 && SF == 1;

	/* video_interface.cpp:1299-1318 */
	r.h_start *= SF;
	r.h_start_clamp *= SF;
	r.v_start *= SF;
	r.h_res *= SF;
	r.h_res_clamp *= SF;
	r.v_res *= SF;
	r.max_x = r.max_x * SF + (SF - 1);
	r.max_y = r.max_y * SF + (SF - 1);
	for (int l = 0; l < MAX_OUTPUT_SCANLINES; l++) {
		lines[l].h_start *= SF;
		lines[l].h_start_clamp *= SF;
		lines[l].h_end_clamp *= SF;
		lines[l].x_start *= SF;
		lines[l].y_start *= SF;
		lines[l].y_base *= SF;
	}

	/* Stages 1-3: extract, AA/dither fetch filter, divot. */
	image_t E = { 0, 0, NULL }, A[2] = { { 0, 0, NULL }, { 0, 0, NULL } }, D[2] = { { 0, 0, NULL }, { 0, 0, NULL } };
	const image_t *S[2] = { NULL, NULL };
	if (!degenerate) {
		if (L && src->domain_pre) {
			uint32_t offset, length;
			scanout_memory_range(s, regs, &offset, &length);
			src->domain_pre(src->ctx, offset, length, (r.status & VI_TYPE_MASK) == VI_TYPE_RGBA8888 ? 2 : 1);
		}
		E.w = r.max_x + 2 + 4 + divot * 2;
		E.h = r.max_y + 1 + 4;
		E.t = (texel_t *)malloc(sizeof(texel_t) * (size_t)E.w * (size_t)E.h);
		int fb_offset = (r.status & VI_TYPE_MASK) == VI_TYPE_RGBA8888 ? r.vi_offset >> 2 : r.vi_offset >> 1;
		int x_off = divot ? -3 : -2, y_off = -2;
		for (int y = 0; y < E.h; y++)
			for (int x = 0; x < E.w; x++)
				E.t[(size_t)y * E.w + x] = fetch_color(src, r.status, fb_offset, r.vi_width, x + x_off, y + y_off);

		for (int l = 0; l < 2; l++) {
			A[l].w = r.max_x + 3 + divot * 2;
			A[l].h = r.max_y + 2;
			A[l].t = (texel_t *)malloc(sizeof(texel_t) * (size_t)A[l].w * (size_t)A[l].h);
		}
		for (int y = 0; y < A[0].h; y++)
			for (int x = 0; x < A[0].w; x++)
				vi_fetch_texel(&E, r.status, fetch_bug, x + 2, y + 2, &A[0].t[(size_t)y * A[0].w + x],
				               &A[1].t[(size_t)y * A[1].w + x]);
		if (divot) {
			for (int l = 0; l < 2; l++) {
				D[l].w = r.max_x + 2;
				D[l].h = r.max_y + 2;
				D[l].t = (texel_t *)malloc(sizeof(texel_t) * (size_t)D[l].w * (size_t)D[l].h);
				for (int y = 0; y < D[l].h; y++)
					for (int x = 0; x < D[l].w; x++)
						D[l].t[(size_t)y * D[l].w + x] = divot_texel(&A[l], x, y);
			}
			S[0] = &D[0];
			S[1] = &D[1];
		} else {
			S[0] = &A[0];
			S[1] = &A[1];
		}
	}

	/* Stage 4: scale_stage. */
	int serrate = (r.status & VI_SERRATE) != 0 && !opts->upscale_deinterlacing;
	int W = SCANOUT_WIDTH * SF;
	int H = ((r.is_pal ? V_RES_PAL : V_RES_NTSC) >> (serrate ? 0 : 1)) * SF;
	int crop_l, crop_r, crop_t, crop_b;
	if (opts->crop_enable) {
		crop_l = (int)opts->crop_left;
		crop_r = (int)opts->crop_right;
		crop_t = (int)opts->crop_top;
		crop_b = (int)opts->crop_bottom;
		if (serrate) {
			crop_t *= 2;
			crop_b *= 2;
		}
	} else {
		int cy = (int)opts->crop_overscan_pixels * (serrate ? 2 : 1);
		int cx = (int)lroundf((float)cy * ((float)W / (float)H));
		crop_l = crop_r = cx;
		crop_t = crop_b = cy;
	}
	crop_l *= SF;
	crop_r *= SF;
	crop_t *= SF;
	crop_b *= SF;
	int OW = W, OH = H;
	if (crop_l + crop_r < W && crop_t + crop_b < H) {
		OW = W - (crop_l + crop_r);
		OH = H - (crop_t + crop_b);
	}

	int v_start = r.v_start, v_res = r.v_res;
	int serrate_shift = 0, serrate_mask = 0, serrate_select = 0, info_y_shift = L;
	if (serrate) {
		v_start *= 2;
		v_res *= 2;
		serrate_shift = 1;
		serrate_mask = 1;
		serrate_select = r.v_current_line == 0;
		info_y_shift++;
	}

	memset(out, 0, sizeof(uint32_t) * (size_t)OW * (size_t)OH);
	int draw_x = r.h_start, draw_y = v_start, draw_w = r.h_res, draw_h = v_res;
	shift_rect(&draw_x, &draw_y, &draw_w, &draw_h, -crop_l, -crop_t);
	if (!degenerate && r.h_res > 0 && v_res > 0 && draw_w > 0 && draw_h > 0) {
		for (int py = draw_y; py < draw_y + draw_h && py < OH; py++) {
			for (int px = draw_x; px < draw_x + draw_w && px < OW; px++) {
				int cx = px + crop_l, cy = py + crop_t;
				const vi_line_t *li = &lines[cy >> info_y_shift];
				cy -= v_start;
				if (cx < li->h_start_clamp || cx >= li->h_end_clamp || ((cy & serrate_mask) != serrate_select))
					continue;
				cx -= li->h_start;
				cy >>= serrate_shift;
				uint32_t noise = 0;
				if (r.status & VI_GAMMA_DITHER)
					noise = reseed_noise((uint32_t)cx, (uint32_t)cy, s->frame_count);
				int x = cx * li->x_add + li->x_start;
				int y = (cy - li->y_base) * li->y_add + li->y_start;
				int bx = x >> 10, by = y >> 10;
				texel_t c00t = img_get(S[0], bx, by);
				uint32_t c00[3] = { c00t.c[0], c00t.c[1], c00t.c[2] };
				int bug = 0;
				if (fetch_bug) {
					int prev_y = (y - r.init_y_add) >> 10;
					int next_y = (y + r.init_y_add) >> 10;
					if (cy != 0 && by == prev_y && by != next_y)
						bug = 1;
				}
				if (r.status & VI_META_SCALE) {
					uint32_t xf = (uint32_t)((x >> 5) & 31), yf = (uint32_t)((y >> 5) & 31);
					texel_t c10 = img_get(S[0], bx + 1, by);
					texel_t c01 = img_get(S[bug], bx, by + 1);
					texel_t c11 = img_get(S[bug], bx + 1, by + 1);
					for (int i = 0; i < 3; i++) {
						uint32_t a = vi_lerp(c00[i], c01.c[i], yf);
						uint32_t b = vi_lerp(c10.c[i], c11.c[i], yf);
						c00[i] = vi_lerp(a, b, xf);
					}
				}
				if (r.status & VI_GAMMA) {
					if (r.status & VI_GAMMA_DITHER) {
						uint32_t n[3] = { noise & 0x3f, (noise >> 6) & 0x3f, ((noise >> 9) & 0x38) | (noise & 7) };
						for (int i = 0; i < 3; i++)
							c00[i] = gamma_lookup((c00[i] << 6) + n[i] + 256u);
					} else
						for (int i = 0; i < 3; i++)
							c00[i] = gamma_lookup(c00[i]);
				} else if (r.status & VI_GAMMA_DITHER) {
					for (int i = 0; i < 3; i++) {
						uint32_t v = c00[i] + ((noise >> i) & 1u);
						c00[i] = v > 0xff ? 0xff : v;
					}
				}
				out[(size_t)py * OW + px] = c00[0] | (c00[1] << 8) | (c00[2] << 16) | 0xff000000u;
			}
		}
	}

	/* Blend in the previous frame where this one drew nothing
	 * (vi_blend_fields.frag; video_interface.cpp:919-981). */
	if (s->prev_valid && opts->blend_previous_frame) {
		int rects[3][4];
		int nr = 0;
		int ph = (int)s->prev_height;
		if (degenerate) {
			if (r.h_res > 0) {
				rects[nr][0] = r.h_start; rects[nr][1] = 0; rects[nr][2] = r.h_res; rects[nr][3] = ph; nr++;
			}
		} else {
			if (r.h_res > 0 && v_start > 0) {
				rects[nr][0] = r.h_start; rects[nr][1] = 0; rects[nr][2] = r.h_res; rects[nr][3] = v_start; nr++;
			}
			if (r.h_res_clamp > 0 && v_res > 0) {
				rects[nr][0] = r.h_start_clamp; rects[nr][1] = v_start; rects[nr][2] = r.h_res_clamp;
				rects[nr][3] = v_res; nr++;
			}
			if (r.h_res > 0 && ph > v_start + v_res) {
				rects[nr][0] = r.h_start; rects[nr][1] = v_start + v_res; rects[nr][2] = r.h_res;
				rects[nr][3] = ph - (v_start + v_res); nr++;
			}
		}
		for (int k = 0; k < nr; k++) {
			int rx = rects[k][0], ry = rects[k][1], rw = rects[k][2], rh = rects[k][3];
			shift_rect(&rx, &ry, &rw, &rh, -crop_l, -crop_t);
			for (int py = ry; py < ry + rh && py < OH; py++)
				for (int px = rx; px < rx + rw && px < OW; px++) {
					uint32_t *d = &out[(size_t)py * OW + px];
					if ((*d >> 24) != 0)
						continue; /* destination alpha 1: keep */
					uint32_t p = 0;
					if ((unsigned)px < s->prev_width && (unsigned)py < s->prev_height)
						p = s->prev[(size_t)py * s->prev_width + px];
					/* rgb * alpha, alpha channel not written */
					*d = (p >> 24) ? (p & 0x00ffffffu) : 0;
				}
		}
	}

	*width = (unsigned)OW;
	*height = (unsigned)OH;
	if ((size_t)OW * (size_t)OH > s->prev_cap) {
		free(s->prev);
		s->prev_cap = (size_t)OW * (size_t)OH;
		s->prev = (uint32_t *)malloc(sizeof(uint32_t) * s->prev_cap);
		if (!s->prev)
			abort();
	}
	memcpy(s->prev, out, sizeof(uint32_t) * (size_t)OW * (size_t)OH);
	s->prev_width = (unsigned)OW;
	s->prev_height = (unsigned)OH;
	s->prev_valid = 1;
	s->frame_count++;

	free(E.t);
	free(A[0].t);
	free(A[1].t);
	free(D[0].t);
	free(D[1].t);
	return 0;
}
